// wamp_capi.cpp -- C-ABI of libwamp.so (declared in include/wamp.h).
//
// Host side of the drop-in boundary: argument checks, scratch ownership and
// the stage sequence of one SSP-RK3 step (SC:2028-2059).  No entry point
// synchronises the device or falls back to a host computation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "wamp.h"

#include "wamp_internal.h"
#include "wamp_math.h"
#include "wamp_types.h"

using namespace wamp;

static Eos make_eos(const wamp_eos& e) { return make_eos_values(e.gamma1, e.pinf1, e.gamma2, e.pinf2); }

extern "C" {
hipError_t wamp_launch_faces(const Geo*, const Eos*, const State*, const Faces*, int base,
                             int shear, int face_pp, int vf_always, double t1, int precision,
                             int low, hipStream_t);
hipError_t wamp_launch_limit_update(const Geo*, const Eos*, const State* Un, const State* U,
                                    const Faces*, const State* Uo, double alpha, double beta,
                                    const double* dt_dev, double* dt_acc, int repair,
                                    hipStream_t);
hipError_t wamp_launch_lu_march(const Geo*, const Eos*, const State* Un, const State* U,
                               const Faces*, const State* Uo, double alpha, double beta,
                               const double* dt_dev, double* dt_acc, int repair, int seg,
                               hipStream_t);
hipError_t wamp_launch_lu_lo(const Geo*, const Eos*, const State* Un, const State* U,
                             const Faces*, const State* Uo, double alpha, double beta,
                             const double* dt_dev, double* dt_acc, int repair, int seg,
                             hipStream_t);
hipError_t wamp_launch_dt(const Geo*, const Eos*, const State*, double* dt_acc, double* dt_out,
                          double cfl, hipStream_t);
hipError_t wamp_launch_dt_finalize(double* dt_acc, double* dt_out, double cfl, hipStream_t);
hipError_t wamp_launch_selftest_arith(const double* a, const double* b, double* out, long long n,
                                      hipStream_t);
hipError_t wamp_launch_selftest_elementary(const double* x, const double* y, double* out,
                                           long long n, hipStream_t);
hipError_t wamp_launch_prepare(const Geo*, const Eos*, const State*, int repair, hipStream_t);
hipError_t wamp_launch_enforce(const Geo*, const Eos*, const State*, hipStream_t);
hipError_t wamp_launch_cons_to_prim(const Geo*, const Eos*, const State*, double*, double*,
                                    double*, hipStream_t);
hipError_t wamp_launch_bc(const Geo*, const State*, double*, double*, double*, hipStream_t);
hipError_t wamp_launch_ic_droplets(const Geo*, const Eos*, const State*, const DropletIC*,
                                  hipStream_t);
hipError_t wamp_launch_ic_triple(const Geo*, const Eos*, const State*, const TripleIC*,
                                 hipStream_t);
hipError_t wamp_launch_faces_batch(const CaseDesc* cs, const int* tile_off, int ncase,
                                   const BatchScal* sc, int stage, int base, int shear,
                                   int face_pp, int vf_always, double t1, int precision, int tile,
                                   int blocks, hipStream_t st);
hipError_t wamp_launch_lu_batch(const CaseDesc* cs, const int* item_off, int ncase,
                                const BatchScal* sc, int stage, int repair, int seg, int blocks,
                                int f32, hipStream_t st);
hipError_t wamp_launch_batch_finalize(const CaseDesc* cs, int ncase, const BatchScal* sc,
                                      hipStream_t st);
hipError_t wamp_launch_batch_dt0(const Geo* G, const Eos* e, const State* U, double* acc,
                                 const BatchScal* sc, int k, double cfl, hipStream_t st);
int wamp3_scheme_available(int base, int shear, int face_pp);
hipError_t wamp3_launch_cells(const Geo3*, const Eos*, const State3*, double* P, double* S,
                              double* D, int shear, hipStream_t);
hipError_t wamp3_launch_faces(const Geo3*, const Eos*, const State3*, const double* P,
                              const double* S, const double* D, const Faces3*, int base, int shear,
                              int vf_always, int face_pp, double t1, hipStream_t);
hipError_t wamp3_launch_faces_mixed(const Geo3*, const Eos*, const State3*, const double* P,
                                    const double* S, const double* D, const Faces3*, int base,
                                    int shear, int vf_always, int face_pp, double t1, hipStream_t);
hipError_t wamp3_launch_limit_update(const Geo3*, const Eos*, const State3* Un, const State3* U,
                                     const Faces3*, double* const th[3], const State3* Uo,
                                     double alpha, double beta, const double* dt_dev,
                                     double* dt_acc, int repair, hipStream_t);
hipError_t wamp3_launch_dt(const Geo3*, const Eos*, const State3*, double* dt_acc, hipStream_t);
hipError_t wamp3_launch_prepare(const Geo3*, const Eos*, const State3*, int repair, hipStream_t);
hipError_t wamp3_launch_batch_stage(const Case3Desc* cs, const int* off, int ncase,
                                    const BatchScal* sc, int stage, int base, int shear,
                                    int vf_always, int face_pp, int repair, double t1,
                                    int precision, const int* blocks, hipStream_t st);
hipError_t wamp3_launch_batch_finalize(const Case3Desc* cs, int ncase, const BatchScal* sc,
                                       hipStream_t st);
hipError_t wamp3_launch_batch_prepare(const Geo3* G, const Eos* e, const State3* U, int repair,
                                      double* acc, const BatchScal* sc, int k, double cfl,
                                      hipStream_t st);
}

// A context is a step plan plus the scratch it owns.
struct wamp_ctx : StepPlan {
  double* block;      // one allocation: faces | stage buffers | dt accumulator
  int64_t bytes;
  StepScratch scr;
  // measurement hook: events around the kernels of up to timing_cap steps
  int timing_cap;
  int timing_n;
  hipEvent_t* ev;  // 7 per step: faces/limit_update boundaries of 3 stages
};

static constexpr int EV_PER_STEP = 7;

static thread_local std::string g_err;

namespace wamp {

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return 0;
  // reported here, so clear the thread's last-error slot: the launch wrappers
  // read it after each launch and would otherwise report this error again
  // (e.g. a refused hipMalloc) against a later, unrelated launch
  (void)hipGetLastError();
  return fail((int)e, "%s: %s", what, hipGetErrorString(e));
}

int state_of(double* const U[6], State& s, const char* name) {
  if (!U) return fail(-2, "%s: NULL array list", name);
  for (int m = 0; m < 6; ++m) {
    if (!U[m]) return fail(-2, "%s[%d] is NULL", name, m);
    s.f[m] = U[m];
  }
  return 0;
}

// Bytes of the face planes of one case: 28 FP64 planes, or, in mixed
// precision, the four FP64 energy planes and 24 FP32 planes (the mixed face
// kernels' other fluxes are FP32 values; see Faces).
size_t face_bytes(size_t plane, int precision) {
  return precision == 0 ? 28 * plane * sizeof(double)
                        : 4 * plane * sizeof(double) + 24 * plane * sizeof(float);
}
// Faces over `bytes = face_bytes(plane, precision)` at p (8-byte aligned).
Faces carve_faces(char* p, size_t plane, int precision) {
  Faces f{(double*)p, (long long)plane};
  if (precision != 0) {
    f.h = (float*)(p + 4 * plane * sizeof(double));
    f.hs = (long long)plane;
  }
  return f;
}
size_t step_scratch_bytes(size_t plane, int precision) {
  return face_bytes(plane, precision) + (12 * plane + 32) * sizeof(double);
}
StepScratch carve_step_scratch(void* base, size_t plane, int precision) {
  StepScratch s;
  s.faces = carve_faces((char*)base, plane, precision);
  double* p = (double*)((char*)base + face_bytes(plane, precision));  // a multiple of 8
  for (int m = 0; m < 6; ++m, p += plane) s.u1.f[m] = p;
  for (int m = 0; m < 6; ++m, p += plane) s.u2.f[m] = p;
  s.dt_acc = p;
  return s;
}

int check_case(const wamp_grid* grid, const wamp_eos* eos, const wamp_scheme* scheme,
               int32_t precision) {
  if (!grid || !eos || !scheme) return fail(-1, "grid, eos and scheme must be non-NULL");
  if (grid->nx < 1 || grid->ny < 1) return fail(-3, "nx, ny must be >= 1 (got %d, %d)", grid->nx, grid->ny);
  if (grid->g < 4) return fail(-3, "g must be >= 4 (the face stencil reaches 4 cells), got %d", grid->g);
  if (grid->ld < grid->ny + 2 * grid->g)
    return fail(-3, "ld (%d) must be >= ny + 2g (%d)", grid->ld, grid->ny + 2 * grid->g);
  if (!(grid->dx > 0.0) || !(grid->dy > 0.0)) return fail(-3, "dx, dy must be > 0");
  if ((int64_t)grid->nx * grid->ny > (int64_t)1 << 30) return fail(-3, "grid too large");
  for (int k = 0; k < 4; ++k)
    if (grid->bc[k] != 0 && grid->bc[k] != 1) return fail(-3, "bc[%d] must be 0 or 1", k);
  if (!(eos->gamma1 > 1.0) || !(eos->gamma2 > 1.0)) return fail(-3, "gamma must be > 1");
  if (scheme->base < 0 || scheme->base > 1) return fail(-3, "scheme.base must be 0 or 1");
  if (scheme->shear < 0 || scheme->shear > 2) return fail(-3, "scheme.shear must be 0..2");
  if (scheme->vf_thinc < 0 || scheme->vf_thinc > 1) return fail(-3, "scheme.vf_thinc must be 0 or 1");
  if (scheme->face_pp < 0 || scheme->face_pp > 1) return fail(-3, "scheme.face_pp must be 0 or 1");
  if (!(scheme->cfl > 0.0)) return fail(-3, "scheme.cfl must be > 0");
  if (precision != WAMP_PREC_FP64 && precision != WAMP_PREC_MIXED)
    return fail(-4, "precision %d is not available (0 = FP64, 1 = mixed)", precision);
  return 0;
}

void plan_fill(StepPlan& p, const wamp_grid& grid, const wamp_eos& eos, const wamp_scheme& scheme,
               int precision, int device, int cu_count) {
  p.geo.nx = grid.nx;
  p.geo.ny = grid.ny;
  p.geo.g = grid.g;
  p.geo.ld = grid.ld;
  p.geo.dx = grid.dx;
  p.geo.dy = grid.dy;
  for (int k = 0; k < 4; ++k) p.geo.bc[k] = grid.bc[k];
  p.eos = make_eos(eos);
  p.scheme = scheme;
  p.precision = precision;
  p.device = device;
  p.t1 = wm_tanh(1.8 * 0.5);
  p.lu_kind = -1;
  p.lu_seg = 0;
  if (const char* v = std::getenv("WAMP_LU_KIND")) p.lu_kind = std::atoi(v);
  if (const char* v = std::getenv("WAMP_LU_SEG")) p.lu_seg = std::atoi(v);
  p.cu_count = cu_count;
}

}  // namespace wamp

static int check_ctx(const wamp_ctx* c) {
  if (!c) return fail(-1, "NULL context");
  return 0;
}

extern "C" {

const char* wamp_last_error(void) { return g_err.c_str(); }


int wamp_ctx_create(const wamp_grid* grid, const wamp_eos* eos, const wamp_scheme* scheme,
                    int32_t precision, wamp_ctx** out) {
  if (!out) return fail(-1, "out is NULL");
  *out = nullptr;
  if (int rc = check_case(grid, eos, scheme, precision)) return rc;
  int device = 0, cu_count = 0;
  int rc = hip_check(hipGetDevice(&device), "hipGetDevice");
  if (!rc)
    rc = hip_check(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, device),
                   "hipDeviceGetAttribute");
  if (rc) return rc;
  wamp_ctx* c = new (std::nothrow) wamp_ctx();
  if (!c) return fail(-5, "host allocation failed");
  plan_fill(*c, *grid, *eos, *scheme, precision, device, cu_count);

  const size_t plane = (size_t)(grid->nx + 2 * grid->g) * (size_t)grid->ld;
  const size_t bytes = step_scratch_bytes(plane, precision);
  rc = hip_check(hipMalloc((void**)&c->block, bytes), "hipMalloc(scratch)");
  if (rc) { delete c; return rc; }
  c->bytes = (int64_t)bytes;
  c->scr = carve_step_scratch(c->block, plane, precision);
  rc = hip_check(wamp_launch_fill(c->scr.dt_acc, 1.0e30, nullptr), "fill dt accumulator");
  if (!rc) rc = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (rc) {
    (void)hipFree(c->block);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

static void timing_free(wamp_ctx* c) {
  if (c->ev) {
    for (int k = 0; k < c->timing_cap * EV_PER_STEP; ++k) (void)hipEventDestroy(c->ev[k]);
    delete[] c->ev;
  }
  c->ev = nullptr;
  c->timing_cap = 0;
  c->timing_n = 0;
}

int wamp_ctx_destroy(wamp_ctx* c) {
  if (!c) return 0;
  timing_free(c);
  int rc = 0;
  if (c->block) rc = hip_check(hipFree(c->block), "hipFree(scratch)");
  delete c;
  return rc;
}

int64_t wamp_ctx_scratch_bytes(const wamp_ctx* c) { return c ? c->bytes : -1; }

int wamp_apply_bc(wamp_ctx* c, double* const U[6], double* const P[3], void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  double *up = nullptr, *vp = nullptr, *pp = nullptr;
  if (P) {
    if (!P[0] || !P[1] || !P[2]) return fail(-2, "P has a NULL entry");
    up = P[0]; vp = P[1]; pp = P[2];
  }
  return hip_check(wamp_launch_bc(&c->geo, &s, up, vp, pp, (hipStream_t)stream), "apply_bc");
}

int wamp_cons_to_prim(wamp_ctx* c, double* const U[6], double* const P[3], void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  if (!P || !P[0] || !P[1] || !P[2]) return fail(-2, "P must hold three non-NULL arrays");
  return hip_check(wamp_launch_cons_to_prim(&c->geo, &c->eos, &s, P[0], P[1], P[2],
                                            (hipStream_t)stream), "cons_to_prim");
}

int wamp_enforce_admissible(wamp_ctx* c, double* const U[6], void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  if (!c->scheme.repair) return 0;
  return hip_check(wamp_launch_enforce(&c->geo, &c->eos, &s, (hipStream_t)stream),
                   "enforce_admissible");
}

int wamp_prepare(wamp_ctx* c, double* const U[6], void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  return hip_check(wamp_launch_prepare(&c->geo, &c->eos, &s, c->scheme.repair, (hipStream_t)stream),
                   "prepare");
}

static int faces_of(const wamp_ctx* c, double* const F[28], Faces& f) {
  if (!F) return fail(-2, "F is NULL");
  for (int m = 0; m < 28; ++m)
    if (!F[m]) return fail(-2, "F[%d] is NULL", m);
  const long long st = (long long)(F[1] - F[0]);
  const long long plane = (long long)(c->geo.nx + 2 * c->geo.g) * c->geo.ld;
  if (st < plane && -st < plane)
    return fail(-2, "face planes overlap (spacing %lld < plane %lld elements)", st, plane);
  for (int m = 2; m < 28; ++m)
    if ((long long)(F[m] - F[0]) != (long long)m * st)
      return fail(-2, "the 28 face planes must be equally spaced in one allocation (F[%d])", m);
  f = Faces{F[0], st};
  return 0;
}

// Rows per wave of the marching limiter (k_lu_march).  Its waves, one per
// (58-column strip, segment), are packed 4 to a block and run two blocks per
// CU, so the device holds 8 x CUs of them at a time; a wave of `seg` rows
// runs seg + 6 fronts.  Cost: rounds x (seg + 6), rounds = waves / (8 x CUs)
// rounded up; for each round count the shortest segment whose waves fit is
// the candidate, as for k_lu_lo below.
constexpr int MJ_COLS = 58;  // output columns per limiter wave (64 lanes - 2 x 3 halo)
static int rows_for_rounds(long long nx, long long strips, long long slots, int fronts_extra,
                           bool paired) {
  const long long max_seg = std::max(4LL, nx);
  long long best = max_seg, best_cost = -1;
  for (long long r = 1;; ++r) {
    const long long nseg = std::max(1LL, r * slots / strips);  // segments per strip in r rounds
    const long long seg = std::max(4LL, std::min(max_seg, (nx + nseg - 1) / nseg));
    const long long waves = strips * ((nx + seg - 1) / seg);
    const long long rounds = (waves + slots - 1) / slots;
    long long fronts = seg + fronts_extra;
    if (paired) fronts = (fronts + 1) & ~1LL;
    const long long cost = rounds * fronts;
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = seg;
    }
    if (seg == 4 || r > 64) break;
  }
  return (int)best;
}
static int lu_rows(const StepPlan* c) {
  if (c->lu_seg) return c->lu_seg;
  const long long strips = (c->geo.ny + MJ_COLS - 1) / MJ_COLS;
  return rows_for_rounds(c->geo.nx, strips, 8LL * std::max(1, c->cu_count), 6, false);
}

// Rows per wave of the march that forms the first-order fluxes itself
// (k_lu_lo).  Its waves, one per (56-column strip, segment), are packed 4 to
// a block and run one block per CU, so the device holds 4 x CUs of them at a
// time; a wave of `seg` rows runs seg + 6 fronts (seg + 7 when that is odd:
// the fronts run in pairs).  Cost: rounds x fronts, where rounds = waves /
// (4 x CUs) rounded up.  For each round count R the shortest segment whose
// waves fit in R rounds is the candidate, and the cheapest candidate wins:
// at 4096x2048 on 256 CUs one round of 27 segments of 152 rows (158 fronts)
// instead of five rounds of 32-row segments (190 fronts).  Segments stay at
// 4 rows or more.
constexpr int LO_COLS = 56;
static int lu_rows_lo(const StepPlan* c) {
  if (c->lu_seg) return c->lu_seg;
  const long long strips = (c->geo.ny + LO_COLS - 1) / LO_COLS;
  return rows_for_rounds(c->geo.nx, strips, 4LL * std::max(1, c->cu_count), 6, true);
}

// Limiter kernel: the LDS-tile kernel (every tile in parallel, four short
// phases) has the lower latency on small grids, whose march would leave most
// SIMDs idle; the marching kernel moves fewer bytes and wins from about
// 2^18 cells (measured: 512x256 0.12 vs 0.19 ms per step, 1024x512 0.37 vs
// 0.33, 2048x1024 1.45 vs 1.14; with the one-round rows per wave the
// marches win from 1024x512 on and the LDS-tile kernel up to 512x256).
// Kind 2 is the march that forms the first-order fluxes itself (FP64 face
// storage only), so the face kernel skips them; it is the FP64 step's march
// (4096x2048 M10, per launch: faces 1.02 instead of 1.27 ms with the
// first-order planes, limiter 0.665 ms).  Every kind gives bitwise the same
// result.  `planes`: the
// caller hands over all 28 face planes (the split entry point), which kind 2
// reads only when WAMP_LU_KIND asks for it.
static int lu_kind(const StepPlan* c, bool planes = false) {
  int k = c->lu_kind >= 0 ? c->lu_kind : (long long)c->geo.nx * c->geo.ny <= (1LL << 18) ? 0 : 2;
  if (k == 2 && (c->precision != 0 || (planes && c->lu_kind != 2))) k = 1;
  return k;
}

static int launch_lu(const StepPlan* c, int kind, const State& un, const State& u, const Faces& f,
                     const State& uo, double alpha, double beta, const double* dt_dev,
                     double* dt_acc, hipStream_t st) {
  hipError_t e;
  if (kind == 2)
    e = wamp_launch_lu_lo(&c->geo, &c->eos, &un, &u, &f, &uo, alpha, beta, dt_dev, dt_acc,
                          c->scheme.repair, lu_rows_lo(c), st);
  else if (kind == 1)
    e = wamp_launch_lu_march(&c->geo, &c->eos, &un, &u, &f, &uo, alpha, beta, dt_dev, dt_acc,
                             c->scheme.repair, lu_rows(c), st);
  else
    e = wamp_launch_limit_update(&c->geo, &c->eos, &un, &u, &f, &uo, alpha, beta, dt_dev, dt_acc,
                                 c->scheme.repair, st);
  return hip_check(e, "limit_update");
}

// low = 0: the face kernel skips the first-order fluxes (limiter kind 2).
static int launch_faces(const StepPlan* c, const State& s, const Faces& f, int low,
                        hipStream_t st) {
  return hip_check(wamp_launch_faces(&c->geo, &c->eos, &s, &f, c->scheme.base, c->scheme.shear,
                                     c->scheme.face_pp, c->scheme.vf_thinc, c->t1, c->precision,
                                     low, st),
                   "faces");
}

int wamp_compute_fluxes(wamp_ctx* c, double* const U[6], double* const F[28], void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  Faces f;
  if (int rc = faces_of(c, F, f)) return rc;
  return launch_faces(c, s, f, 1, (hipStream_t)stream);
}

int wamp_limit_update(wamp_ctx* c, double* const Un[6], double* const U[6], double* const F[28],
                      double* const Uout[6], double alpha, double beta, const double* dt_dev,
                      double* dt_acc, void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State sn, s, so;
  if (int rc = state_of(Un, sn, "Un")) return rc;
  if (int rc = state_of(U, s, "U")) return rc;
  if (int rc = state_of(Uout, so, "Uout")) return rc;
  for (int m = 0; m < 6; ++m)
    if (so.f[m] == s.f[m]) return fail(-2, "Uout[%d] aliases U[%d]", m, m);
  if (!dt_dev) return fail(-2, "dt_dev must be non-NULL");
  Faces f;
  if (int rc = faces_of(c, F, f)) return rc;
  return launch_lu(c, lu_kind(c, true), sn, s, f, so, alpha, beta, dt_dev, dt_acc,
                   (hipStream_t)stream);
}

int wamp_get_dt(wamp_ctx* c, double* const U[6], double* dt_dev, void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  if (!dt_dev) return fail(-2, "dt_dev is NULL");
  return hip_check(wamp_launch_dt(&c->geo, &c->eos, &s, c->scr.dt_acc, dt_dev, c->scheme.cfl,
                                  (hipStream_t)stream), "get_dt");
}

int wamp_step(wamp_ctx* c, double* const U[6], const double* dt_dev, double* dt_next_dev,
              void* stream) {
  if (int rc = check_ctx(c)) return rc;
  State a;
  if (int rc = state_of(U, a, "U")) return rc;
  if (!dt_dev || !dt_next_dev) return fail(-2, "dt_dev and dt_next_dev must be non-NULL");
  hipEvent_t* ev = nullptr;
  if (c->timing_n < c->timing_cap) ev = c->ev + (size_t)c->timing_n++ * EV_PER_STEP;
  return step_enqueue(*c, a, c->scr, dt_dev, dt_next_dev, (hipStream_t)stream, ev);
}

}  // extern "C"

int wamp::step_enqueue(const StepPlan& p, const State& a, const StepScratch& scr,
                       const double* dt_dev, double* dt_next_dev, hipStream_t st, hipEvent_t* ev) {
  auto mark = [&](int k) -> int {
    return ev ? hip_check(hipEventRecord(ev[k], st), "hipEventRecord") : 0;
  };
  const State* cur[3] = {&a, &scr.u1, &scr.u2};
  const State* dst[3] = {&scr.u1, &scr.u2, &a};
  const double coef[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
  // stage 1: U1 = Un + dt R(Un)                               (SC:2031-2037)
  // stage 2: U2 = 3/4 Un + 1/4 (U1 + dt R(U1))                (SC:2040-2046)
  // stage 3: Un+1 = 1/3 Un + 2/3 (U2 + dt R(U2)), written over Un, with the
  //          CFL minimum of the new state folded in           (SC:2049-2059)
  const int kind = lu_kind(&p);
  if (int rc = mark(0)) return rc;
  for (int s = 0; s < 3; ++s) {
    if (int rc = launch_faces(&p, *cur[s], scr.faces, kind != 2, st)) return rc;
    if (int rc = mark(2 * s + 1)) return rc;
    if (int rc = launch_lu(&p, kind, a, *cur[s], scr.faces, *dst[s], coef[s][0], coef[s][1],
                           dt_dev, s == 2 ? scr.dt_acc : nullptr, st))
      return rc;
    if (int rc = mark(2 * s + 2)) return rc;
  }
  return hip_check(wamp_launch_dt_finalize(scr.dt_acc, dt_next_dev, p.scheme.cfl, st), "dt");
}

extern "C" {

// ---- initial conditions on the device ------------------------------------
static int geo_of(const wamp_grid* grid, Geo& G) {
  if (!grid) return fail(-1, "grid is NULL");
  if (grid->nx < 1 || grid->ny < 1 || grid->g < 4 || grid->ld < grid->ny + 2 * grid->g ||
      !(grid->dx > 0.0) || !(grid->dy > 0.0))
    return fail(-3, "invalid grid");
  G = Geo{grid->nx, grid->ny, grid->g, grid->ld, grid->dx, grid->dy,
          {grid->bc[0], grid->bc[1], grid->bc[2], grid->bc[3]}};
  return 0;
}

int wamp_init_droplets(const wamp_grid* grid, const wamp_eos* eos, const wamp_droplet_ic* ic,
                       double* const U[6], void* stream) {
  Geo G;
  if (int rc = geo_of(grid, G)) return rc;
  if (!eos || !ic) return fail(-1, "eos and ic must be non-NULL");
  if (!(eos->gamma1 > 1.0) || !(eos->gamma2 > 1.0)) return fail(-3, "gamma must be > 1");
  if (ic->n_drops < 0 || (ic->n_drops > 0 && !ic->drops)) return fail(-2, "drops is NULL");
  if (!(ic->width > 0.0)) return fail(-3, "width must be > 0");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  DropletIC d{ic->xmin, ic->ymin, ic->shock_x, ic->rho_post, ic->p_post, ic->u_post,
              ic->rho_water, ic->rho_air, ic->p0, ic->width, ic->drops, ic->n_drops};
  const Eos e = make_eos(*eos);
  return hip_check(wamp_launch_ic_droplets(&G, &e, &s, &d, (hipStream_t)stream), "ic_droplets");
}

int wamp_init_triple_point(const wamp_grid* grid, const wamp_eos* eos, const wamp_triple_ic* ic,
                           double* const U[6], void* stream) {
  Geo G;
  if (int rc = geo_of(grid, G)) return rc;
  if (!eos || !ic) return fail(-1, "eos and ic must be non-NULL");
  if (!(eos->gamma1 > 1.0) || !(eos->gamma2 > 1.0)) return fail(-3, "gamma must be > 1");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  TripleIC t{ic->xmin, ic->ymin, ic->x_split, ic->y_split, ic->rho_left, ic->p_left,
             ic->rho_top, ic->p_top, ic->rho_bottom, ic->p_bottom};
  const Eos e = make_eos(*eos);
  return hip_check(wamp_launch_ic_triple(&G, &e, &s, &t, (hipStream_t)stream), "ic_triple");
}

int wamp_selftest_arith(const double* a, const double* b, double* out, int64_t n, void* stream) {
  if (n < 0 || n > (int64_t)1 << 30) return fail(-3, "n must be in [0, 2^30]");
  if (n == 0) return 0;
  if (!a || !b || !out) return fail(-2, "a, b and out must be non-NULL");
  return hip_check(wamp_launch_selftest_arith(a, b, out, (long long)n, (hipStream_t)stream),
                   "selftest_arith");
}

int wamp_selftest_elementary(const double* x, const double* y, double* out, int64_t n,
                             void* stream) {
  if (n < 0 || n > (int64_t)1 << 30) return fail(-3, "n must be in [0, 2^30]");
  if (n == 0) return 0;
  if (!x || !y || !out) return fail(-2, "x, y and out must be non-NULL");
  return hip_check(wamp_launch_selftest_elementary(x, y, out, (long long)n, (hipStream_t)stream),
                   "selftest_elementary");
}

int wamp_timing_enable(wamp_ctx* c, int32_t max_steps) {
  if (int rc = check_ctx(c)) return rc;
  if (max_steps < 0 || max_steps > 100000) return fail(-3, "max_steps must be in [0, 100000]");
  timing_free(c);
  if (max_steps == 0) return 0;
  c->ev = new (std::nothrow) hipEvent_t[(size_t)max_steps * EV_PER_STEP];
  if (!c->ev) return fail(-5, "host allocation failed");
  for (int k = 0; k < max_steps * EV_PER_STEP; ++k) {
    if (int rc = hip_check(hipEventCreate(&c->ev[k]), "hipEventCreate")) {
      for (int q = 0; q < k; ++q) (void)hipEventDestroy(c->ev[q]);
      delete[] c->ev;
      c->ev = nullptr;
      return rc;
    }
  }
  c->timing_cap = max_steps;
  c->timing_n = 0;
  return 0;
}

int wamp_timing_read(wamp_ctx* c, double ms[2], int64_t launches[2]) {
  if (int rc = check_ctx(c)) return rc;
  if (!ms || !launches) return fail(-2, "ms and launches must be non-NULL");
  ms[0] = ms[1] = 0.0;
  launches[0] = launches[1] = 0;
  for (int n = 0; n < c->timing_n; ++n) {
    hipEvent_t* ev = c->ev + (size_t)n * EV_PER_STEP;
    if (int rc = hip_check(hipEventSynchronize(ev[EV_PER_STEP - 1]), "hipEventSynchronize")) return rc;
    for (int s = 0; s < 3; ++s) {
      float a = 0.f, b = 0.f;
      if (int rc = hip_check(hipEventElapsedTime(&a, ev[2 * s], ev[2 * s + 1]), "elapsed")) return rc;
      if (int rc = hip_check(hipEventElapsedTime(&b, ev[2 * s + 1], ev[2 * s + 2]), "elapsed")) return rc;
      ms[0] += a;
      ms[1] += b;
      launches[0] += 1;
      launches[1] += 1;
    }
  }
  c->timing_n = 0;
  return 0;
}

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------
}  // extern "C"

struct wamp_batch {
  wamp_scheme scheme;
  int precision;
  int max_cases;
  int device;
  int cu_count;
  double t1;
  // tuning
  int tile;            // 0: 8x32 face tiles, 1: 16x16
  int faces_bpcu;      // persistent face blocks per CU
  int lu_bpcu;         // persistent limiter blocks (4 waves) per CU; 0: one wave per item
  int lu_seg;          // limiter rows per wave (in effect)
  int lu_seg_req;      // rows per wave asked for; 0: chosen at each commit
  // host mirror of the case table
  std::vector<CaseDesc> desc;
  std::vector<int> used;         // slot occupied
  std::vector<double*> scratch;  // per-slot scratch allocation
  std::vector<double> t0, tend;
  // device side
  CaseDesc* desc_dev;
  int* off_dev;   // tile offsets [max+1] | item offsets [max+1]
  double* scal;   // dt | acc | t | tend, max_cases each
  int* done;
  BatchScal sc;
  int committed;
  int total_tiles, total_items;
  // graph
  hipStream_t cap_stream;
  hipGraphExec_t graph;
  int graph_steps;
};

static const int TILE_TI[2] = {8, 16};
static const int TILE_TJ[2] = {32, 16};
static const int MJ_LANES = 58;

static void case_work(wamp_batch* b, CaseDesc& d) {
  const int ti = TILE_TI[b->tile], tj = TILE_TJ[b->tile];
  d.tiles_j = (d.geo.ny + 1 + tj - 1) / tj;
  d.tiles = d.tiles_j * ((d.geo.nx + 1 + ti - 1) / ti);
  d.strips = (d.geo.ny + MJ_LANES - 1) / MJ_LANES;
  d.items = d.strips * ((d.geo.nx + b->lu_seg - 1) / b->lu_seg);
}

// Limiter rows per wave for the cases resident at a commit: k_lu_batch runs
// two 4-wave blocks per CU (one when the persistent form has one per CU), a
// wave of `rows` rows marches rows + 6 fronts, and all cases' items share
// the device, so the rows with the fewest rounds x fronts over the whole
// table win (as rows_for_rounds does for one case).  A few large cases then
// get short segments that fill the device in one round; many cases longer
// ones with less warm-up.
static int batch_rows(const wamp_batch* b) {
  static const int cand[] = {4,   5,   6,   7,   8,   10,  12,   14,   16,   19,   22,
                             26,  30,  34,  40,  48,  56,  64,   76,   90,   106,  128,
                             152, 180, 216, 256, 300, 360, 430,  512,  600,  720,  860,
                             1024, 1200, 1450, 1700, 2048, 2400, 2900, 3400, 4096};
  const long long slots = 4LL * (b->lu_bpcu == 1 ? 1 : 2) * std::max(1, b->cu_count);
  int max_nx = 4;
  for (int k = 0; k < b->max_cases; ++k)
    if (b->used[k]) max_nx = std::max(max_nx, b->desc[k].geo.nx);
  int best = 26;
  long long best_cost = -1;
  for (int r : cand) {
    if (r > max_nx && best_cost >= 0) break;
    long long waves = 0;
    for (int k = 0; k < b->max_cases; ++k) {
      if (!b->used[k]) continue;
      const Geo& g = b->desc[k].geo;
      waves += (long long)((g.ny + MJ_LANES - 1) / MJ_LANES) * ((g.nx + r - 1) / r);
    }
    const long long cost = ((waves + slots - 1) / slots) * (r + 6);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = r;
    }
  }
  return best;
}

static void graph_drop(wamp_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
}

extern "C" {

int wamp_batch_create(const wamp_scheme* scheme, int32_t precision, int32_t max_cases,
                      wamp_batch** out) {
  if (!out) return fail(-1, "out is NULL");
  *out = nullptr;
  if (!scheme) return fail(-1, "scheme is NULL");
  if (max_cases < 1 || max_cases > 4096) return fail(-3, "max_cases must be in [1, 4096]");
  if (precision != WAMP_PREC_FP64 && precision != WAMP_PREC_MIXED)
    return fail(-4, "precision %d is not available (0 = FP64, 1 = mixed)", precision);
  const int b_ = scheme->base, s_ = scheme->shear, p_ = scheme->face_pp;
  if (!((b_ == 0 && s_ == 1 && p_ == 0) || (b_ == 0 && s_ == 0 && p_ == 1) ||
        (b_ == 1 && s_ == 2 && p_ == 0)))
    return fail(-3, "batches support the SC (0,1,0), M10 (0,0,1) and MP5 (1,2,0) "
                    "(base, shear, face_pp) schemes");
  if (!(scheme->cfl > 0.0)) return fail(-3, "scheme.cfl must be > 0");
  wamp_batch* b = new (std::nothrow) wamp_batch();
  if (!b) return fail(-5, "host allocation failed");
  b->scheme = *scheme;
  b->precision = precision;
  b->max_cases = max_cases;
  b->t1 = wm_tanh(1.8 * 0.5);
  b->tile = 0;
  b->faces_bpcu = 3;
  b->lu_bpcu = 0;
  b->lu_seg = 16;
  b->lu_seg_req = 16;
  b->desc.assign(max_cases, CaseDesc());
  b->used.assign(max_cases, 0);
  b->scratch.assign(max_cases, nullptr);
  b->t0.assign(max_cases, 0.0);
  b->tend.assign(max_cases, 0.0);
  int rc = hip_check(hipGetDevice(&b->device), "hipGetDevice");
  if (!rc) rc = hip_check(hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount,
                                                b->device), "hipDeviceGetAttribute");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->desc_dev, sizeof(CaseDesc) * max_cases), "hipMalloc");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->off_dev, sizeof(int) * 2 * (max_cases + 1)), "hipMalloc");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->scal, sizeof(double) * 4 * max_cases), "hipMalloc");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->done, sizeof(int) * max_cases), "hipMalloc");
  if (!rc) rc = hip_check(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
  if (rc) {
    delete b;
    return rc;
  }
  b->sc.dt = b->scal;
  b->sc.acc = b->scal + max_cases;
  b->sc.t = b->scal + 2 * max_cases;
  b->sc.tend = b->scal + 3 * max_cases;
  b->sc.done = b->done;
  std::vector<double> init(4 * max_cases, 0.0);
  for (int k = 0; k < max_cases; ++k) init[max_cases + k] = 1.0e30;
  std::vector<int> ones(max_cases, 1);
  rc = hip_check(hipMemcpy(b->scal, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice), "hipMemcpy");
  if (!rc) rc = hip_check(hipMemcpy(b->done, ones.data(), sizeof(int) * max_cases, hipMemcpyHostToDevice), "hipMemcpy");
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int wamp_batch_destroy(wamp_batch* b) {
  if (!b) return 0;
  graph_drop(b);
  for (double* p : b->scratch)
    if (p) (void)hipFree(p);
  (void)hipFree(b->desc_dev);
  (void)hipFree(b->off_dev);
  (void)hipFree(b->scal);
  (void)hipFree(b->done);
  (void)hipStreamDestroy(b->cap_stream);
  delete b;
  return 0;
}

int wamp_batch_set_case(wamp_batch* b, int32_t slot, const wamp_grid* grid, const wamp_eos* eos,
                        double* const U[6], double t0, double t_end) {
  if (!b) return fail(-1, "NULL batch");
  if (slot < 0 || slot >= b->max_cases) return fail(-3, "slot %d out of range", slot);
  if (!grid || !eos) return fail(-1, "grid and eos must be non-NULL");
  if (grid->nx < 1 || grid->ny < 1 || grid->g < 4 || grid->ld < grid->ny + 2 * grid->g ||
      !(grid->dx > 0.0) || !(grid->dy > 0.0))
    return fail(-3, "invalid grid for slot %d", slot);
  if (!(eos->gamma1 > 1.0) || !(eos->gamma2 > 1.0)) return fail(-3, "gamma must be > 1");
  State st;
  if (int rc = state_of(U, st, "U")) return rc;
  const size_t plane = (size_t)(grid->nx + 2 * grid->g) * (size_t)grid->ld;
  const size_t fbytes = face_bytes(plane, b->precision);
  // the new scratch first: if it cannot be had, the slot keeps its case and
  // the batch stays as it was (still committed, graph still valid)
  char* base = nullptr;
  if (int rc = hip_check(hipMalloc((void**)&base, fbytes + 12 * plane * sizeof(double)),
                         "hipMalloc(case scratch)"))
    return rc;
  if (b->scratch[slot]) {
    // the old scratch may still be in use by enqueued work of the batch
    (void)hipDeviceSynchronize();
    (void)hipFree(b->scratch[slot]);
  }
  b->scratch[slot] = (double*)base;
  graph_drop(b);  // a captured graph holds the old case table
  CaseDesc& d = b->desc[slot];
  d = CaseDesc();
  d.geo = Geo{grid->nx, grid->ny, grid->g, grid->ld, grid->dx, grid->dy,
              {grid->bc[0], grid->bc[1], grid->bc[2], grid->bc[3]}};
  d.eos = make_eos(*eos);
  for (int m = 0; m < 6; ++m) d.u[0][m] = U[m];
  d.f = carve_faces(base, plane, b->precision);
  double* p = (double*)(base + fbytes);
  for (int m = 0; m < 6; ++m, p += plane) d.u[1][m] = p;
  for (int m = 0; m < 6; ++m, p += plane) d.u[2][m] = p;
  d.cfl = b->scheme.cfl;
  b->used[slot] = 1;
  b->t0[slot] = t0;
  b->tend[slot] = t_end;
  b->committed = 0;
  return 0;
}

int wamp_batch_clear_case(wamp_batch* b, int32_t slot) {
  if (!b) return fail(-1, "NULL batch");
  if (slot < 0 || slot >= b->max_cases) return fail(-3, "slot %d out of range", slot);
  if (b->scratch[slot]) {
    (void)hipDeviceSynchronize();  // enqueued work of the batch may still use it
    (void)hipFree(b->scratch[slot]);
  }
  b->scratch[slot] = nullptr;
  b->used[slot] = 0;
  b->desc[slot] = CaseDesc();
  b->committed = 0;
  return 0;
}

int wamp_batch_set_tuning(wamp_batch* b, int32_t tile, int32_t faces_bpcu, int32_t lu_bpcu,
                          int32_t lu_seg) {
  if (!b) return fail(-1, "NULL batch");
  if (tile < 0 || tile > 1) return fail(-3, "tile must be 0 (8x32) or 1 (16x16)");
  if (faces_bpcu < 1 || faces_bpcu > 16) return fail(-3, "face blocks per CU must be in [1, 16]");
  if (lu_bpcu < 0 || lu_bpcu > 16) return fail(-3, "limiter blocks per CU must be in [0, 16]");
  if (lu_seg != 0 && (lu_seg < 4 || lu_seg > 4096))
    return fail(-3, "lu_seg must be 0 (chosen at each commit) or in [4, 4096]");
  b->tile = tile;
  b->faces_bpcu = faces_bpcu;
  b->lu_bpcu = lu_bpcu;
  b->lu_seg_req = lu_seg;
  if (lu_seg) b->lu_seg = lu_seg;
  b->committed = 0;
  return 0;
}

// Upload the case table and work-list offsets; a slot's time and end time
// are (re)initialised only when its case was (re)set since the last commit.
int wamp_batch_commit(wamp_batch* b, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  hipStream_t st = (hipStream_t)stream;
  graph_drop(b);
  b->committed = 0;  // until the new table is on the device
  const int n = b->max_cases;
  if (b->lu_seg_req == 0) b->lu_seg = batch_rows(b);
  std::vector<int> off(2 * (n + 1), 0);
  for (int k = 0; k < n; ++k) {
    int tiles = 0, items = 0;
    if (b->used[k]) {
      case_work(b, b->desc[k]);
      tiles = b->desc[k].tiles;
      items = b->desc[k].items;
    }
    off[k + 1] = off[k] + tiles;
    off[n + 1 + k + 1] = off[n + 1 + k] + items;
  }
  b->total_tiles = off[n];
  b->total_items = off[2 * n + 1];
  if (int rc = hip_check(hipMemcpyAsync(b->desc_dev, b->desc.data(), sizeof(CaseDesc) * n,
                                        hipMemcpyHostToDevice, st), "hipMemcpyAsync(desc)")) return rc;
  if (int rc = hip_check(hipMemcpyAsync(b->off_dev, off.data(), sizeof(int) * off.size(),
                                        hipMemcpyHostToDevice, st), "hipMemcpyAsync(offsets)")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;
  b->committed = 1;
  return 0;
}

// Initial sequence (repair + clamp, SC:1998-2001) and first dt of the slots
// listed in `slots` (all occupied slots when n_slots < 0); sets t = t0.
int wamp_batch_prepare(wamp_batch* b, const int32_t* slots, int32_t n_slots, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  if (!b->committed) return fail(-6, "batch not committed");
  hipStream_t st = (hipStream_t)stream;
  std::vector<int> list;
  if (n_slots < 0) {
    for (int k = 0; k < b->max_cases; ++k)
      if (b->used[k]) list.push_back(k);
  } else {
    if (!slots && n_slots > 0) return fail(-2, "slots is NULL");
    for (int q = 0; q < n_slots; ++q) {
      if (slots[q] < 0 || slots[q] >= b->max_cases || !b->used[slots[q]])
        return fail(-3, "slot %d is not occupied", slots[q]);
      list.push_back(slots[q]);
    }
  }
  for (int k : list) {
    const CaseDesc& d = b->desc[k];
    State s;
    for (int m = 0; m < 6; ++m) s.f[m] = d.u[0][m];
    double vals[2] = {b->t0[k], b->tend[k]};
    if (int rc = hip_check(hipMemcpyAsync(b->sc.t + k, &vals[0], sizeof(double), hipMemcpyHostToDevice, st), "t0")) return rc;
    if (int rc = hip_check(hipMemcpyAsync(b->sc.tend + k, &vals[1], sizeof(double), hipMemcpyHostToDevice, st), "tend")) return rc;
    if (int rc = hip_check(wamp_launch_prepare(&d.geo, &d.eos, &s, b->scheme.repair, st), "prepare")) return rc;
    if (int rc = hip_check(wamp_launch_batch_dt0(&d.geo, &d.eos, &s, b->sc.acc + k, &b->sc, k, d.cfl, st), "dt0")) return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;  // vals lifetime
  }
  return 0;
}

// Mark slots finished (frozen) or not.
int wamp_batch_set_done(wamp_batch* b, int32_t slot, int32_t done, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  if (slot < 0 || slot >= b->max_cases) return fail(-3, "slot %d out of range", slot);
  int v = done ? 1 : 0;
  if (int rc = hip_check(hipMemcpyAsync(b->done + slot, &v, sizeof(int), hipMemcpyHostToDevice,
                                        (hipStream_t)stream), "done")) return rc;
  return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

static int batch_enqueue_step(wamp_batch* b, hipStream_t st) {
  const int n = b->max_cases;
  const int fb = std::max(1, std::min(b->total_tiles, b->faces_bpcu * b->cu_count));
  // limiter: one wave per item (lu_bpcu 0), or lu_bpcu persistent blocks per CU
  const int per_item = std::max(1, (b->total_items + 3) / 4);
  const int lb = b->lu_bpcu == 0 ? per_item : std::min(per_item, b->lu_bpcu * b->cu_count);
  for (int s = 0; s < 3; ++s) {
    if (int rc = hip_check(wamp_launch_faces_batch(b->desc_dev, b->off_dev, n, &b->sc, s,
                                                   b->scheme.base, b->scheme.shear,
                                                   b->scheme.face_pp, b->scheme.vf_thinc, b->t1,
                                                   b->precision, b->tile, fb, st), "faces_batch"))
      return rc;
    if (int rc = hip_check(wamp_launch_lu_batch(b->desc_dev, b->off_dev + n + 1, n, &b->sc, s,
                                                b->scheme.repair, b->lu_seg, lb,
                                                b->precision != 0, st), "lu_batch"))
      return rc;
  }
  return hip_check(wamp_launch_batch_finalize(b->desc_dev, n, &b->sc, st), "finalize");
}

// Capture `steps` consecutive steps in a hipGraph (on the batch's own
// capture stream); wamp_batch_step then replays it.  0 drops the graph.
int wamp_batch_graph(wamp_batch* b, int32_t steps) {
  if (!b) return fail(-1, "NULL batch");
  if (!b->committed) return fail(-6, "batch not committed");
  graph_drop(b);
  if (steps <= 0) return 0;
  hipGraph_t g = nullptr;
  if (int rc = hip_check(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal),
                         "hipStreamBeginCapture")) return rc;
  int rc = 0;
  for (int k = 0; k < steps && !rc; ++k) rc = batch_enqueue_step(b, b->cap_stream);
  int rc2 = hip_check(hipStreamEndCapture(b->cap_stream, &g), "hipStreamEndCapture");
  if (rc || rc2) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : rc2;
  }
  rc = hip_check(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (rc) {
    b->graph = nullptr;
    return rc;
  }
  b->graph_steps = steps;
  return 0;
}

// Advance every occupied, unfinished case: with a captured graph, one replay
// (graph_steps steps); otherwise one step.  Returns the steps enqueued (> 0).
int wamp_batch_step(wamp_batch* b, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  if (!b->committed) return fail(-6, "batch not committed");
  hipStream_t st = (hipStream_t)stream;
  if (b->graph) {
    if (int rc = hip_check(hipGraphLaunch(b->graph, st), "hipGraphLaunch")) return rc;
    return b->graph_steps;
  }
  if (int rc = batch_enqueue_step(b, st)) return rc;
  return 1;
}

// Copy t, dt and done of every slot to the host (synchronises the stream).
int wamp_batch_status(wamp_batch* b, double* t, double* dt, int32_t* done, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  hipStream_t st = (hipStream_t)stream;
  const int n = b->max_cases;
  if (t && hip_check(hipMemcpyAsync(t, b->sc.t, sizeof(double) * n, hipMemcpyDeviceToHost, st), "t")) return -7;
  if (dt && hip_check(hipMemcpyAsync(dt, b->sc.dt, sizeof(double) * n, hipMemcpyDeviceToHost, st), "dt")) return -7;
  if (done && hip_check(hipMemcpyAsync(done, b->done, sizeof(int) * n, hipMemcpyDeviceToHost, st), "done")) return -7;
  return hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

int64_t wamp_batch_work(const wamp_batch* b, int32_t which) {
  if (!b) return -1;
  return which == 0 ? b->total_tiles : which == 1 ? b->total_items : b->lu_seg;
}

// ---- three-dimensional extension -------------------------------------------
// One 3D case: 48 face planes, two 7-plane stage buffers, pressure, entropy
// and Ducros planes, three limiter theta planes and the dt accumulator, in
// one allocation.
struct wamp3_ctx {
  Geo3 geo;
  Eos eos;
  wamp_scheme scheme;
  int precision;
  double t1;
  double* block;
  int64_t bytes;
  Faces3 faces;
  State3 u1, u2;
  double *P, *S, *D;
  double* th[3];
  double* dt_acc;
};

static int state3_of(double* const U[7], State3& s, const char* name) {
  if (!U) return fail(-2, "%s: NULL array list", name);
  for (int m = 0; m < 7; ++m) {
    if (!U[m]) return fail(-2, "%s[%d] is NULL", name, m);
    s.f[m] = U[m];
  }
  return 0;
}

}  // extern "C"

// Checks of a 3D grid and EOS shared by contexts and batches; 0 when valid.
static int check_grid3(const wamp_grid3* grid, const wamp_eos* eos) {
  if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1)
    return fail(-3, "nx, ny, nz must be >= 1 (got %d, %d, %d)", grid->nx, grid->ny, grid->nz);
  if (grid->g < 4) return fail(-3, "g must be >= 4 (the face stencil reaches 4 cells), got %d", grid->g);
  if (grid->ldz < grid->nz + 2 * grid->g)
    return fail(-3, "ldz (%d) must be >= nz + 2g (%d)", grid->ldz, grid->nz + 2 * grid->g);
  if (!(grid->dx > 0.0) || !(grid->dy > 0.0) || !(grid->dz > 0.0)) return fail(-3, "dx, dy, dz must be > 0");
  const int64_t plane64 = (int64_t)(grid->nx + 2 * grid->g) * (grid->ny + 2 * grid->g) * grid->ldz;
  if (plane64 > (int64_t)1 << 31) return fail(-3, "grid too large (padded plane > 2^31 cells)");
  if (grid->nx + 2 * grid->g > 65535)
    return fail(-3, "nx + 2g must be <= 65535 (one launch row per x index)");
  for (int k = 0; k < 6; ++k)
    if (grid->bc[k] != 0 && grid->bc[k] != 1) return fail(-3, "bc[%d] must be 0 or 1", k);
  if (!(eos->gamma1 > 1.0) || !(eos->gamma2 > 1.0)) return fail(-3, "gamma must be > 1");
  return 0;
}
static int check_scheme3(const wamp_scheme* scheme) {
  if (scheme->vf_thinc < 0 || scheme->vf_thinc > 1) return fail(-3, "scheme.vf_thinc must be 0 or 1");
  if (!wamp3_scheme_available(scheme->base, scheme->shear, scheme->face_pp))
    return fail(-4, "3D: scheme (base %d, shear %d, face_pp %d) is not built; available: SC (0,1,0), "
                "M10 (0,0,1), MP5 (1,2,0)", scheme->base, scheme->shear, scheme->face_pp);
  if (!(scheme->cfl > 0.0)) return fail(-3, "scheme.cfl must be > 0");
  return 0;
}
static Geo3 make_geo3(const wamp_grid3* grid) {
  Geo3 G;
  G.nx = grid->nx;
  G.ny = grid->ny;
  G.nz = grid->nz;
  G.g = grid->g;
  G.sy = grid->ldz;
  G.sx = (long long)(grid->ny + 2 * grid->g) * grid->ldz;
  G.dx = grid->dx;
  G.dy = grid->dy;
  G.dz = grid->dz;
  for (int k = 0; k < 6; ++k) G.bc[k] = grid->bc[k];
  return G;
}
static size_t plane3(const wamp_grid3* grid) {
  return (size_t)(grid->nx + 2 * grid->g) * (grid->ny + 2 * grid->g) * grid->ldz;
}
// face planes: 48 FP64, or in mixed precision 6 FP64 energy planes and 42
// FP32 planes (a multiple of 8 bytes: 42 * 4 * plane)
static size_t faces3_bytes(size_t plane, int precision) {
  return precision == 0 ? 48 * plane * sizeof(double)
                        : 6 * plane * sizeof(double) + 42 * plane * sizeof(float);
}
static Faces3 carve_faces3(char* base, size_t plane, int precision) {
  Faces3 f{(double*)base, (long long)plane};
  if (precision != 0) {
    f.h = (float*)(base + 6 * plane * sizeof(double));
    f.hs = (long long)plane;
  }
  return f;
}

extern "C" {

int wamp3_ctx_create(const wamp_grid3* grid, const wamp_eos* eos, const wamp_scheme* scheme,
                     int32_t precision, wamp3_ctx** out) {
  if (!out) return fail(-1, "out is NULL");
  *out = nullptr;
  if (!grid || !eos || !scheme) return fail(-1, "grid, eos and scheme must be non-NULL");
  if (int rc = check_grid3(grid, eos)) return rc;
  if (int rc = check_scheme3(scheme)) return rc;
  if (precision != WAMP_PREC_FP64 && precision != WAMP_PREC_MIXED)
    return fail(-4, "precision %d is not available (0 = FP64, 1 = mixed)", precision);

  wamp3_ctx* c = new (std::nothrow) wamp3_ctx();
  if (!c) return fail(-5, "host allocation failed");
  c->geo = make_geo3(grid);
  c->eos = make_eos(*eos);
  c->scheme = *scheme;
  c->precision = precision;
  c->t1 = wm_tanh(1.8 * 0.5);
  const size_t plane = plane3(grid);
  const size_t fbytes = faces3_bytes(plane, precision);
  const size_t bytes = fbytes + ((14 + 3 + 3) * plane + 32) * sizeof(double);
  int rc = hip_check(hipMalloc((void**)&c->block, bytes), "hipMalloc(3D scratch)");
  if (rc) { delete c; return rc; }
  c->bytes = (int64_t)bytes;
  c->faces = carve_faces3((char*)c->block, plane, precision);
  double* p = (double*)((char*)c->block + fbytes);
  for (int m = 0; m < 7; ++m, p += plane) c->u1.f[m] = p;
  for (int m = 0; m < 7; ++m, p += plane) c->u2.f[m] = p;
  c->P = p; p += plane;
  c->S = p; p += plane;
  c->D = p; p += plane;
  for (int s = 0; s < 3; ++s, p += plane) c->th[s] = p;
  c->dt_acc = p;
  rc = hip_check(wamp_launch_fill(c->dt_acc, 1.0e30, nullptr), "fill dt accumulator");
  if (!rc) rc = hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (rc) {
    (void)hipFree(c->block);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int wamp3_ctx_destroy(wamp3_ctx* c) {
  if (!c) return 0;
  int rc = 0;
  if (c->block) rc = hip_check(hipFree(c->block), "hipFree(3D scratch)");
  delete c;
  return rc;
}

int64_t wamp3_ctx_scratch_bytes(const wamp3_ctx* c) { return c ? c->bytes : -1; }

int wamp3_prepare(wamp3_ctx* c, double* const U[7], void* stream) {
  if (!c) return fail(-1, "NULL context");
  State3 s;
  if (int rc = state3_of(U, s, "U")) return rc;
  return hip_check(wamp3_launch_prepare(&c->geo, &c->eos, &s, c->scheme.repair, (hipStream_t)stream),
                   "3D prepare");
}

int wamp3_get_dt(wamp3_ctx* c, double* const U[7], double* dt_dev, void* stream) {
  if (!c) return fail(-1, "NULL context");
  State3 s;
  if (int rc = state3_of(U, s, "U")) return rc;
  if (!dt_dev) return fail(-2, "dt_dev is NULL");
  hipStream_t st = (hipStream_t)stream;
  if (int rc = hip_check(wamp3_launch_dt(&c->geo, &c->eos, &s, c->dt_acc, st), "3D dt")) return rc;
  return hip_check(wamp_launch_dt_finalize(c->dt_acc, dt_dev, c->scheme.cfl, st), "3D dt");
}

static int faces3_of(const wamp3_ctx* c, double* const F[48], Faces3& f) {
  if (!F) return fail(-2, "F is NULL");
  for (int m = 0; m < 48; ++m)
    if (!F[m]) return fail(-2, "F[%d] is NULL", m);
  const long long st = (long long)(F[1] - F[0]);
  const long long plane = (long long)(c->geo.nx + 2 * c->geo.g) * c->geo.sx;
  if (st < plane && -st < plane)
    return fail(-2, "face planes overlap (spacing %lld < plane %lld elements)", st, plane);
  for (int m = 2; m < 48; ++m)
    if ((long long)(F[m] - F[0]) != (long long)m * st)
      return fail(-2, "the 48 face planes must be equally spaced in one allocation (F[%d])", m);
  f = Faces3{F[0], st};
  return 0;
}

int wamp3_compute_fluxes(wamp3_ctx* c, double* const U[7], double* const F[48], void* stream) {
  if (!c) return fail(-1, "NULL context");
  State3 s;
  if (int rc = state3_of(U, s, "U")) return rc;
  Faces3 f;
  if (int rc = faces3_of(c, F, f)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const wamp_scheme& sc = c->scheme;
  if (int rc = hip_check(wamp3_launch_cells(&c->geo, &c->eos, &s, c->P, c->S, c->D, sc.shear, st),
                         "3D cells"))
    return rc;
  auto faces = c->precision == 0 ? wamp3_launch_faces : wamp3_launch_faces_mixed;
  return hip_check(faces(&c->geo, &c->eos, &s, c->P, c->S, c->D, &f, sc.base, sc.shear,
                         sc.vf_thinc, sc.face_pp, c->t1, st),
                   "3D faces");
}

int wamp3_limit_update(wamp3_ctx* c, double* const Un[7], double* const U[7], double* const F[48],
                       double* const Uout[7], double alpha, double beta, const double* dt_dev,
                       double* dt_acc, void* stream) {
  if (!c) return fail(-1, "NULL context");
  State3 sn, s, so;
  if (int rc = state3_of(Un, sn, "Un")) return rc;
  if (int rc = state3_of(U, s, "U")) return rc;
  if (int rc = state3_of(Uout, so, "Uout")) return rc;
  for (int m = 0; m < 7; ++m)
    if (so.f[m] == s.f[m]) return fail(-2, "Uout[%d] aliases U[%d]", m, m);
  if (!dt_dev) return fail(-2, "dt_dev must be non-NULL");
  Faces3 f;
  if (int rc = faces3_of(c, F, f)) return rc;
  return hip_check(wamp3_launch_limit_update(&c->geo, &c->eos, &sn, &s, &f, c->th, &so, alpha,
                                             beta, dt_dev, dt_acc, c->scheme.repair,
                                             (hipStream_t)stream),
                   "3D limit/update");
}

int wamp3_step(wamp3_ctx* c, double* const U[7], const double* dt_dev, double* dt_next_dev,
               void* stream) {
  if (!c) return fail(-1, "NULL context");
  State3 a;
  if (int rc = state3_of(U, a, "U")) return rc;
  if (!dt_dev || !dt_next_dev) return fail(-2, "dt_dev and dt_next_dev must be non-NULL");
  hipStream_t st = (hipStream_t)stream;
  const State3* cur[3] = {&a, &c->u1, &c->u2};
  const State3* dst[3] = {&c->u1, &c->u2, &a};
  const double coef[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
  const wamp_scheme& sc = c->scheme;
  for (int s = 0; s < 3; ++s) {
    if (int rc = hip_check(wamp3_launch_cells(&c->geo, &c->eos, cur[s], c->P, c->S, c->D, sc.shear, st),
                           "3D cells"))
      return rc;
    auto faces = c->precision == 0 ? wamp3_launch_faces : wamp3_launch_faces_mixed;
    if (int rc = hip_check(faces(&c->geo, &c->eos, cur[s], c->P, c->S, c->D, &c->faces, sc.base,
                                 sc.shear, sc.vf_thinc, sc.face_pp, c->t1, st),
                           "3D faces"))
      return rc;
    if (int rc = hip_check(wamp3_launch_limit_update(&c->geo, &c->eos, &a, cur[s], &c->faces, c->th,
                                                     dst[s], coef[s][0], coef[s][1], dt_dev,
                                                     s == 2 ? c->dt_acc : nullptr, sc.repair, st),
                           "3D limit/update"))
      return rc;
  }
  return hip_check(wamp_launch_dt_finalize(c->dt_acc, dt_next_dev, sc.cfl, st), "3D dt");
}

// ---- batches of 3D cases ---------------------------------------------------
// Slots of resident 3D cases sharing one scheme and precision.  Each step is
// one launch per kernel and stage over every active case's blocks, with
// per-case dt, time, end time and done flag on the device (as wamp_batch).
struct wamp3_batch {
  wamp_scheme scheme;
  int precision;
  int max_cases;
  double t1;
  std::vector<Case3Desc> desc;
  std::vector<int> used;
  std::vector<char*> scratch;
  std::vector<double> t0, tend;
  Case3Desc* desc_dev;
  int* off_dev;  // B3_KINDS prefix arrays of max_cases + 1 block offsets
  double* scal;  // dt | acc | t | tend
  int* done;
  BatchScal sc;
  int committed;
  int blocks[B3_KINDS];
  hipStream_t cap_stream;
  hipGraphExec_t graph;
  int graph_steps;
};

static void graph3_drop(wamp3_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
}

int wamp3_batch_create(const wamp_scheme* scheme, int32_t precision, int32_t max_cases,
                       wamp3_batch** out) {
  if (!out) return fail(-1, "out is NULL");
  *out = nullptr;
  if (!scheme) return fail(-1, "scheme is NULL");
  if (max_cases < 1 || max_cases > 4096) return fail(-3, "max_cases must be in [1, 4096]");
  if (precision != WAMP_PREC_FP64 && precision != WAMP_PREC_MIXED)
    return fail(-4, "precision %d is not available (0 = FP64, 1 = mixed)", precision);
  if (int rc = check_scheme3(scheme)) return rc;
  wamp3_batch* b = new (std::nothrow) wamp3_batch();
  if (!b) return fail(-5, "host allocation failed");
  b->scheme = *scheme;
  b->precision = precision;
  b->max_cases = max_cases;
  b->t1 = wm_tanh(1.8 * 0.5);
  b->desc.assign(max_cases, Case3Desc());
  b->used.assign(max_cases, 0);
  b->scratch.assign(max_cases, nullptr);
  b->t0.assign(max_cases, 0.0);
  b->tend.assign(max_cases, 0.0);
  int rc = hip_check(hipMalloc((void**)&b->desc_dev, sizeof(Case3Desc) * max_cases), "hipMalloc");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->off_dev, sizeof(int) * B3_KINDS * (max_cases + 1)), "hipMalloc");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->scal, sizeof(double) * 4 * max_cases), "hipMalloc");
  if (!rc) rc = hip_check(hipMalloc((void**)&b->done, sizeof(int) * max_cases), "hipMalloc");
  if (!rc) rc = hip_check(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
  if (!rc) {
    b->sc.dt = b->scal;
    b->sc.acc = b->scal + max_cases;
    b->sc.t = b->scal + 2 * max_cases;
    b->sc.tend = b->scal + 3 * max_cases;
    b->sc.done = b->done;
    std::vector<double> init(4 * max_cases, 0.0);
    for (int k = 0; k < max_cases; ++k) init[max_cases + k] = 1.0e30;
    std::vector<int> ones(max_cases, 1);
    rc = hip_check(hipMemcpy(b->scal, init.data(), sizeof(double) * init.size(), hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = hip_check(hipMemcpy(b->done, ones.data(), sizeof(int) * max_cases, hipMemcpyHostToDevice), "hipMemcpy");
  }
  if (rc) {
    (void)hipFree(b->desc_dev);
    (void)hipFree(b->off_dev);
    (void)hipFree(b->scal);
    (void)hipFree(b->done);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

int wamp3_batch_destroy(wamp3_batch* b) {
  if (!b) return 0;
  graph3_drop(b);
  for (char* p : b->scratch)
    if (p) (void)hipFree(p);
  (void)hipFree(b->desc_dev);
  (void)hipFree(b->off_dev);
  (void)hipFree(b->scal);
  (void)hipFree(b->done);
  (void)hipStreamDestroy(b->cap_stream);
  delete b;
  return 0;
}

// Put a 3D case in a slot (allocates its scratch: the face planes, two stage
// buffers, pressure / entropy / Ducros and the three theta planes).
int wamp3_batch_set_case(wamp3_batch* b, int32_t slot, const wamp_grid3* grid,
                         const wamp_eos* eos, double* const U[7], double t0, double t_end) {
  if (!b) return fail(-1, "NULL batch");
  if (slot < 0 || slot >= b->max_cases) return fail(-3, "slot %d out of range", slot);
  if (!grid || !eos) return fail(-1, "grid and eos must be non-NULL");
  if (int rc = check_grid3(grid, eos)) return rc;
  State3 s;
  if (int rc = state3_of(U, s, "U")) return rc;
  const size_t plane = plane3(grid);
  const size_t fbytes = faces3_bytes(plane, b->precision);
  char* base = nullptr;
  if (int rc = hip_check(hipMalloc((void**)&base, fbytes + (14 + 3 + 3) * plane * sizeof(double)),
                         "hipMalloc(3D case scratch)"))
    return rc;
  if (b->scratch[slot]) {
    (void)hipDeviceSynchronize();  // enqueued work of the batch may still use it
    (void)hipFree(b->scratch[slot]);
  }
  b->scratch[slot] = base;
  graph3_drop(b);
  Case3Desc& d = b->desc[slot];
  d = Case3Desc();
  d.geo = make_geo3(grid);
  d.eos = make_eos(*eos);
  for (int m = 0; m < 7; ++m) d.u[0][m] = U[m];
  d.f = carve_faces3(base, plane, b->precision);
  double* p = (double*)(base + fbytes);
  for (int m = 0; m < 7; ++m, p += plane) d.u[1][m] = p;
  for (int m = 0; m < 7; ++m, p += plane) d.u[2][m] = p;
  d.P = p; p += plane;
  d.S = p; p += plane;
  d.D = p; p += plane;
  for (int q = 0; q < 3; ++q, p += plane) d.th[q] = p;
  d.cfl = b->scheme.cfl;
  b->used[slot] = 1;
  b->t0[slot] = t0;
  b->tend[slot] = t_end;
  b->committed = 0;
  return 0;
}

int wamp3_batch_clear_case(wamp3_batch* b, int32_t slot) {
  if (!b) return fail(-1, "NULL batch");
  if (slot < 0 || slot >= b->max_cases) return fail(-3, "slot %d out of range", slot);
  if (b->scratch[slot]) {
    (void)hipDeviceSynchronize();
    (void)hipFree(b->scratch[slot]);
  }
  b->scratch[slot] = nullptr;
  b->used[slot] = 0;
  b->desc[slot] = Case3Desc();
  b->committed = 0;
  return 0;
}

// Upload the case table and the per-kind block offsets.
int wamp3_batch_commit(wamp3_batch* b, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  hipStream_t st = (hipStream_t)stream;
  graph3_drop(b);
  b->committed = 0;  // until the new table and its block counts are on the device
  const int n = b->max_cases;
  std::vector<int> off(B3_KINDS * (n + 1), 0);
  int blocks[B3_KINDS];
  for (int kind = 0; kind < B3_KINDS; ++kind) {
    int* o = off.data() + kind * (n + 1);
    long long acc = 0;
    for (int k = 0; k < n; ++k) {
      if (b->used[k]) acc += kind_blocks3(b->desc[k].geo, kind);
      if (acc > (1LL << 31) - 1) return fail(-3, "batch too large (more than 2^31 blocks)");
      o[k + 1] = (int)acc;
    }
    blocks[kind] = o[n];
  }
  if (int rc = hip_check(hipMemcpyAsync(b->desc_dev, b->desc.data(), sizeof(Case3Desc) * n,
                                        hipMemcpyHostToDevice, st), "hipMemcpyAsync(desc)")) return rc;
  if (int rc = hip_check(hipMemcpyAsync(b->off_dev, off.data(), sizeof(int) * off.size(),
                                        hipMemcpyHostToDevice, st), "hipMemcpyAsync(offsets)")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;
  for (int kind = 0; kind < B3_KINDS; ++kind) b->blocks[kind] = blocks[kind];
  b->committed = 1;
  return 0;
}

// Initial sequence (repair + clamp) and first dt of the listed slots (all
// occupied slots when n_slots < 0); sets t = t0.
int wamp3_batch_prepare(wamp3_batch* b, const int32_t* slots, int32_t n_slots, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  if (!b->committed) return fail(-6, "batch not committed");
  hipStream_t st = (hipStream_t)stream;
  std::vector<int> list;
  if (n_slots < 0) {
    for (int k = 0; k < b->max_cases; ++k)
      if (b->used[k]) list.push_back(k);
  } else {
    if (!slots && n_slots > 0) return fail(-2, "slots is NULL");
    for (int q = 0; q < n_slots; ++q) {
      if (slots[q] < 0 || slots[q] >= b->max_cases || !b->used[slots[q]])
        return fail(-3, "slot %d is not occupied", slots[q]);
      list.push_back(slots[q]);
    }
  }
  for (int k : list) {
    const Case3Desc& d = b->desc[k];
    State3 s;
    for (int m = 0; m < 7; ++m) s.f[m] = d.u[0][m];
    double vals[2] = {b->t0[k], b->tend[k]};
    if (int rc = hip_check(hipMemcpyAsync(b->sc.t + k, &vals[0], sizeof(double), hipMemcpyHostToDevice, st), "t0")) return rc;
    if (int rc = hip_check(hipMemcpyAsync(b->sc.tend + k, &vals[1], sizeof(double), hipMemcpyHostToDevice, st), "tend")) return rc;
    if (int rc = hip_check(wamp3_launch_batch_prepare(&d.geo, &d.eos, &s, b->scheme.repair,
                                                      b->sc.acc + k, &b->sc, k, d.cfl, st),
                           "3D batch prepare"))
      return rc;
    if (int rc = hip_check(hipStreamSynchronize(st), "hipStreamSynchronize")) return rc;  // vals lifetime
  }
  return 0;
}

static int batch3_enqueue_step(wamp3_batch* b, hipStream_t st) {
  const wamp_scheme& sc = b->scheme;
  for (int s = 0; s < 3; ++s)
    if (int rc = hip_check(wamp3_launch_batch_stage(b->desc_dev, b->off_dev, b->max_cases, &b->sc, s,
                                                    sc.base, sc.shear, sc.vf_thinc, sc.face_pp,
                                                    sc.repair, b->t1, b->precision, b->blocks, st),
                           "3D batch stage"))
      return rc;
  return hip_check(wamp3_launch_batch_finalize(b->desc_dev, b->max_cases, &b->sc, st),
                   "3D batch finalize");
}

// Capture `steps` consecutive steps in a hipGraph; 0 drops the graph.
int wamp3_batch_graph(wamp3_batch* b, int32_t steps) {
  if (!b) return fail(-1, "NULL batch");
  if (!b->committed) return fail(-6, "batch not committed");
  graph3_drop(b);
  if (steps <= 0) return 0;
  hipGraph_t g = nullptr;
  if (int rc = hip_check(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal),
                         "hipStreamBeginCapture")) return rc;
  int rc = 0;
  for (int k = 0; k < steps && !rc; ++k) rc = batch3_enqueue_step(b, b->cap_stream);
  int rc2 = hip_check(hipStreamEndCapture(b->cap_stream, &g), "hipStreamEndCapture");
  if (rc || rc2) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : rc2;
  }
  rc = hip_check(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(g);
  if (rc) {
    b->graph = nullptr;
    return rc;
  }
  b->graph_steps = steps;
  return 0;
}

// Advance every occupied, unfinished case: with a captured graph, one replay
// (graph_steps steps); otherwise one step.  Returns the steps enqueued (> 0).
int wamp3_batch_step(wamp3_batch* b, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  if (!b->committed) return fail(-6, "batch not committed");
  hipStream_t st = (hipStream_t)stream;
  if (b->graph) {
    if (int rc = hip_check(hipGraphLaunch(b->graph, st), "hipGraphLaunch")) return rc;
    return b->graph_steps;
  }
  if (int rc = batch3_enqueue_step(b, st)) return rc;
  return 1;
}

// Copy t, dt and done of every slot to the host (synchronises the stream).
int wamp3_batch_status(wamp3_batch* b, double* t, double* dt, int32_t* done, void* stream) {
  if (!b) return fail(-1, "NULL batch");
  hipStream_t st = (hipStream_t)stream;
  const int n = b->max_cases;
  if (t && hip_check(hipMemcpyAsync(t, b->sc.t, sizeof(double) * n, hipMemcpyDeviceToHost, st), "t")) return -7;
  if (dt && hip_check(hipMemcpyAsync(dt, b->sc.dt, sizeof(double) * n, hipMemcpyDeviceToHost, st), "dt")) return -7;
  if (done && hip_check(hipMemcpyAsync(done, b->done, sizeof(int) * n, hipMemcpyDeviceToHost, st), "done")) return -7;
  return hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

}  // extern "C"
