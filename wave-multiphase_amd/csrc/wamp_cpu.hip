// wamp_cpu.hip -- the CPU device: the FP64 step on the host (the reference's
// --cpu, SC:15-21 / M10:14-20, which runs the same Warp kernels on the CPU).
//
// The per-cell code is the device code itself: wamp_faces.h's tile phases
// (staging, indicators and sensors, the x- and y-face of one position) and
// wamp_limit.h's per-cell limiter stages and update are host-and-device
// functions.  Host builds of their FP64 arithmetic use the IEEE divide and
// square root that the device sequences reproduce bit for bit, and the same
// pow and tanh (wamp_math.h), so a host step gives the device step's bits
// (but for the last bits of subnormal values, below 2^-1022, where the
// kernels' shortened division is not the IEEE quotient).
// Only the drivers differ: a face tile is one loop per phase instead of 256
// threads and barriers, the limiter runs each theta stage over whole planes,
// and the work of a sweep is split over std::threads by rows or tiles.
//
// Compiled with -ffp-contract=off like the FP64 kernels.  No HIP runtime call
// is made here: the entry points take host pointers and return when done.
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "wamp.h"
#include "wamp_device.h"
#include "wamp_faces.h"
#include "wamp_internal.h"
#include "wamp_limit.h"
#include "wamp_math.h"
#include "wamp_types.h"

namespace wamp {
namespace {

using CTile = Tile8x32;

// fn(lo, hi, worker) over [0, n) in contiguous chunks, one per worker.
template <class Fn>
void parallel_for(int workers, int n, const Fn& fn) {
  const int t = std::max(1, std::min(workers, n));
  if (t == 1) {
    fn(0, n, 0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(t - 1);
  for (int w = 1; w < t; ++w)
    pool.emplace_back([&fn, n, t, w] { fn((int)((int64_t)n * w / t), (int)((int64_t)n * (w + 1) / t), w); });
  fn(0, (int)((int64_t)n / t), 0);
  for (auto& th : pool) th.join();
}

// Worker threads for threads = 0: OMP_NUM_THREADS when set (the usual cap
// of a process's CPU share), else the CPUs of the process's affinity mask.
int default_threads() {
  if (const char* v = std::getenv("OMP_NUM_THREADS")) {
    const long n = std::strtol(v, nullptr, 10);
    if (n > 0) return (int)std::min<long>(n, 1024);
  }
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) {
    const int n = CPU_COUNT(&set);
    if (n > 0) return std::min(n, 1024);
  }
  return (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 1024u));
}

struct CpuCase {
  Geo geo;
  Eos eos;
  wamp_scheme scheme;
  double t1;
  int threads;
  size_t plane;
  std::vector<double> faces, u1, u2, theta;
  std::unique_ptr<FaceSmem<CTile>[]> smem;  // one tile buffer per worker
};

State planes_of(std::vector<double>& v, size_t plane) {
  State s;
  for (int m = 0; m < 6; ++m) s.f[m] = v.data() + (size_t)m * plane;
  return s;
}

// ---- face fluxes ----------------------------------------------------------
template <int BASE, int SHEAR, int FACEPP>
void faces_t(CpuCase& c, const State& U, const Faces& F) {
  const Geo& G = c.geo;
  const int tj = (G.ny + 1 + CTile::TJ - 1) / CTile::TJ;
  const int ti = (G.nx + 1 + CTile::TI - 1) / CTile::TI;
  const int vf = c.scheme.vf_thinc;
  parallel_for(c.threads, ti * tj, [&](int lo, int hi, int w) {
    FaceSmem<CTile>& S = c.smem[w];
    for (int t = lo; t < hi; ++t) {
      const int i0 = G.g - 1 + (t / tj) * CTile::TI;
      const int j0 = G.g - 1 + (t % tj) * CTile::TJ;
      using SS = StageSet<CTile, SHEAR>;
      for (int k = 0; k < SS::N; ++k) tile_stage_cell<CTile>(G, c.eos, U, i0, j0, SS::cell(k), S);
      for (int k = 0; k < CTile::CXN; ++k) tile_sensor_x<CTile, SHEAR>(S, k);
      for (int k = 0; k < CTile::CYN; ++k) tile_sensor_y<CTile, SHEAR>(S, k);
      for (int p = 0; p < CTile::THREADS; ++p)
        tile_faces_at<CTile, BASE, SHEAR, FACEPP, double>(G, c.eos, F, vf, c.t1, 1, i0, j0, p, S);
    }
  });
}

void faces(CpuCase& c, const State& U, const Faces& F) {
  const int key = c.scheme.base * 6 + c.scheme.shear * 2 + c.scheme.face_pp;
  switch (key) {
    case 0: faces_t<0, 0, 0>(c, U, F); break;
    case 1: faces_t<0, 0, 1>(c, U, F); break;
    case 2: faces_t<0, 1, 0>(c, U, F); break;
    case 3: faces_t<0, 1, 1>(c, U, F); break;
    case 4: faces_t<0, 2, 0>(c, U, F); break;
    case 5: faces_t<0, 2, 1>(c, U, F); break;
    case 6: faces_t<1, 0, 0>(c, U, F); break;
    case 7: faces_t<1, 0, 1>(c, U, F); break;
    case 8: faces_t<1, 1, 0>(c, U, F); break;
    case 9: faces_t<1, 1, 1>(c, U, F); break;
    case 10: faces_t<1, 2, 0>(c, U, F); break;
    default: faces_t<1, 2, 1>(c, U, F); break;
  }
}

// ---- limiter, SSP-RK3 combine, repair, clamp, CFL bound ---------------------
// Face thetas of interior cell c from the cell thetas of stages < nb; cells
// outside the interior hold 1 (pp_blend_x/y SC:1500-1504, 1536-1540).
inline void faces_of(const CpuCase& c, size_t cell, int nb, CellFaces& cf) {
  const size_t ld = (size_t)c.geo.ld;
  for (int k = 0; k < nb; ++k) {
    const double* t = c.theta.data() + (size_t)k * c.plane;
    cf.t[0][k] = face_theta(t[cell - ld], t[cell]);
    cf.t[1][k] = face_theta(t[cell], t[cell + ld]);
    cf.t[2][k] = face_theta(t[cell - 1], t[cell]);
    cf.t[3][k] = face_theta(t[cell], t[cell + 1]);
  }
}

// Returns the CFL minimum of the new state (1e30 unless want_dt).
double limit_update(CpuCase& c, const State& Un, const State& U, const Faces& F, const State& Uo,
                    double alpha, double beta, double dt, bool want_dt) {
  const Geo& G = c.geo;
  const size_t ld = (size_t)G.ld;
  const double inv_dx = 1.0 / G.dx;
  const double inv_dy = 1.0 / G.dy;
  for (int s = 0; s < 3; ++s) {  // pp_cell_theta SC:1436-1476, one plane per stage
    double* th = c.theta.data() + (size_t)s * c.plane;
    parallel_for(c.threads, G.nx, [&](int lo, int hi, int) {
      for (int i = G.g + lo; i < G.g + hi; ++i)
        for (int j = G.g; j < G.g + G.ny; ++j) {
          const size_t cell = (size_t)i * ld + (size_t)j;
          CellFaces cf;
          faces_of(c, cell, s, cf);
          th[cell] = s == 0   ? cell_theta<0, false>(c.eos, U, F, cell, ld, cf, inv_dx, inv_dy, dt)
                     : s == 1 ? cell_theta<1, false>(c.eos, U, F, cell, ld, cf, inv_dx, inv_dy, dt)
                              : cell_theta<2, false>(c.eos, U, F, cell, ld, cf, inv_dx, inv_dy, dt);
        }
    });
  }
  std::vector<double> mins((size_t)std::max(1, c.threads), 1.0e30);
  parallel_for(c.threads, G.nx, [&](int lo, int hi, int w) {
    double m = 1.0e30;
    for (int i = G.g + lo; i < G.g + hi; ++i)
      for (int j = G.g; j < G.g + G.ny; ++j) {
        const size_t cell = (size_t)i * ld + (size_t)j;
        CellFaces cf;
        faces_of(c, cell, 3, cf);
        double un[6];
        cell_update<false>(c.eos, Un, U, F, cell, ld, cf, inv_dx, inv_dy, dt, alpha, beta,
                           c.scheme.repair, un);
        for (int q = 0; q < 6; ++q) Uo.f[q][cell] = un[q];
        if (want_dt) m = wmin(m, cell_dt(c.eos, un, G.dx, G.dy));
      }
    mins[(size_t)w] = m;
  });
  double m = 1.0e30;
  for (double v : mins) m = wmin(m, v);
  return m;
}

// min over interior cells of the CFL bound (timestep_kernel SC:1686-1705)
double dt_min(const CpuCase& c, const State& U) {
  const Geo& G = c.geo;
  std::vector<double> mins((size_t)std::max(1, c.threads), 1.0e30);
  parallel_for(c.threads, G.nx, [&](int lo, int hi, int w) {
    double m = 1.0e30;
    for (int i = G.g + lo; i < G.g + hi; ++i)
      for (int j = G.g; j < G.g + G.ny; ++j) {
        const size_t cell = (size_t)i * (size_t)G.ld + (size_t)j;
        double u[6];
        for (int q = 0; q < 6; ++q) u[q] = U.f[q][cell];
        m = wmin(m, cell_dt(c.eos, u, G.dx, G.dy));
      }
    mins[(size_t)w] = m;
  });
  double m = 1.0e30;
  for (double v : mins) m = wmin(m, v);
  return m;
}

}  // namespace
}  // namespace wamp

using namespace wamp;

struct wamp_cpu_ctx : CpuCase {};

extern "C" {

int wamp_cpu_ctx_create(const wamp_grid* grid, const wamp_eos* eos, const wamp_scheme* scheme,
                        int32_t threads, wamp_cpu_ctx** out) {
  if (!out) return fail(-1, "out is NULL");
  *out = nullptr;
  if (int rc = check_case(grid, eos, scheme, WAMP_PREC_FP64)) return rc;
  if (threads < 0 || threads > 1024) return fail(-3, "threads must be 0 (all cores) .. 1024, got %d", threads);
  wamp_cpu_ctx* c = new (std::nothrow) wamp_cpu_ctx();
  if (!c) return fail(-5, "host allocation failed");
  c->geo = Geo{grid->nx, grid->ny, grid->g, grid->ld, grid->dx, grid->dy,
               {grid->bc[0], grid->bc[1], grid->bc[2], grid->bc[3]}};
  c->eos = make_eos_values(eos->gamma1, eos->pinf1, eos->gamma2, eos->pinf2);
  c->scheme = *scheme;
  c->t1 = wm_tanh(1.8 * 0.5);
  c->threads = threads > 0 ? threads : default_threads();
  c->plane = (size_t)(grid->nx + 2 * grid->g) * (size_t)grid->ld;
  try {
    c->faces.assign(28 * c->plane, 0.0);
    c->u1.assign(6 * c->plane, 0.0);
    c->u2.assign(6 * c->plane, 0.0);
    c->theta.assign(3 * c->plane, 1.0);  // cells outside the interior keep theta = 1
    c->smem.reset(new FaceSmem<CTile>[(size_t)c->threads]);
  } catch (const std::bad_alloc&) {
    const size_t c_plane = c->plane;
    delete c;
    return fail(-5, "host allocation failed (40 planes of %zu cells)", c_plane);
  }
  *out = c;
  return 0;
}

int wamp_cpu_ctx_destroy(wamp_cpu_ctx* c) {
  delete c;
  return 0;
}

int wamp_cpu_prepare(wamp_cpu_ctx* c, double* const U[6]) {
  if (!c) return fail(-1, "ctx is NULL");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  const Geo& G = c->geo;
  parallel_for(c->threads, G.nx, [&](int lo, int hi, int) {
    for (int i = G.g + lo; i < G.g + hi; ++i)
      for (int j = G.g; j < G.g + G.ny; ++j) {
        const size_t cell = (size_t)i * (size_t)G.ld + (size_t)j;
        double u[6];
        for (int q = 0; q < 6; ++q) u[q] = s.f[q][cell];
        if (c->scheme.repair) repair_cell(c->eos, u);
        u[5] = clamp_al(u[5]);
        for (int q = 0; q < 6; ++q) s.f[q][cell] = u[q];
      }
  });
  return 0;
}

int wamp_cpu_apply_bc(wamp_cpu_ctx* c, double* const U[6], double* const P[3]) {
  if (!c) return fail(-1, "ctx is NULL");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  if (P && (!P[0] || !P[1] || !P[2])) return fail(-2, "P[0..2] must be non-NULL");
  const Geo& G = c->geo;
  const int nyg = G.ny + 2 * G.g, nxg = G.nx + 2 * G.g;
  // ghost cells read only interior cells, which no iteration writes
  parallel_for(c->threads, nxg, [&](int lo, int hi, int) {
    for (int i = lo; i < hi; ++i)
      for (int j = 0; j < nyg; ++j) {
        if (i >= G.g && i < G.g + G.nx && j >= G.g && j < G.g + G.ny) continue;
        double sru, srv;
        const size_t src = map_cell(G, i, j, sru, srv);
        const size_t cell = (size_t)i * (size_t)G.ld + (size_t)j;
        for (int q = 0; q < 6; ++q) {
          double v = s.f[q][src];
          if (q == 2 && sru < 0.0) v = -1.0 * v;
          if (q == 3 && srv < 0.0) v = -1.0 * v;
          s.f[q][cell] = v;
        }
        if (P) {
          double uu = P[0][src], vv = P[1][src];
          if (sru < 0.0) uu = -1.0 * uu;
          if (srv < 0.0) vv = -1.0 * vv;
          P[0][cell] = uu;
          P[1][cell] = vv;
          P[2][cell] = P[2][src];
        }
      }
  });
  return 0;
}

int wamp_cpu_cons_to_prim(wamp_cpu_ctx* c, double* const U[6], double* const P[3]) {
  if (!c) return fail(-1, "ctx is NULL");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  if (!P || !P[0] || !P[1] || !P[2]) return fail(-2, "P[0..2] must be non-NULL");
  const Geo& G = c->geo;
  const int nyg = G.ny + 2 * G.g, nxg = G.nx + 2 * G.g;
  parallel_for(c->threads, nxg, [&](int lo, int hi, int) {
    for (int i = lo; i < hi; ++i)
      for (int j = 0; j < nyg; ++j) {
        const size_t cell = (size_t)i * (size_t)G.ld + (size_t)j;
        const double a1 = clamp_al(s.f[5][cell]);
        s.f[5][cell] = a1;
        double uu, vv;
        const double p = pressure(c->eos, s.f[0][cell], s.f[1][cell], s.f[2][cell], s.f[3][cell],
                                  s.f[4][cell], a1, uu, vv);
        P[0][cell] = uu;
        P[1][cell] = vv;
        P[2][cell] = p;
      }
  });
  return 0;
}

int wamp_cpu_get_dt(wamp_cpu_ctx* c, double* const U[6], double* dt) {
  if (!c || !dt) return fail(-1, "ctx and dt must be non-NULL");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  *dt = c->scheme.cfl * dt_min(*c, s);
  return 0;
}

int wamp_cpu_compute_fluxes(wamp_cpu_ctx* c, double* const U[6], double* const F[28]) {
  if (!c) return fail(-1, "ctx is NULL");
  State s;
  if (int rc = state_of(U, s, "U")) return rc;
  if (!F) return fail(-2, "F is NULL");
  const long long stride = (long long)(F[1] - F[0]);
  for (int m = 0; m < 28; ++m)
    if (!F[m] || F[m] != F[0] + (long long)m * stride)
      return fail(-2, "the 28 flux planes must be equally spaced in one allocation");
  if (stride < (long long)c->plane && stride > -(long long)c->plane)
    return fail(-2, "flux planes overlap (spacing %lld < %zu cells)", stride, c->plane);
  Faces f{F[0], stride};
  faces(*c, s, f);
  return 0;
}

int wamp_cpu_step(wamp_cpu_ctx* c, double* const U[6], double dt, double* dt_next) {
  if (!c) return fail(-1, "ctx is NULL");
  State a;
  if (int rc = state_of(U, a, "U")) return rc;
  if (!(dt > 0.0)) return fail(-3, "dt must be > 0");
  const State u1 = planes_of(c->u1, c->plane), u2 = planes_of(c->u2, c->plane);
  const Faces f{c->faces.data(), (long long)c->plane};
  const State* cur[3] = {&a, &u1, &u2};
  const State* dst[3] = {&u1, &u2, &a};
  const double coef[3][2] = {{0.0, 1.0}, {0.75, 0.25}, {1.0 / 3.0, 2.0 / 3.0}};
  double m = 1.0e30;
  for (int s = 0; s < 3; ++s) {  // the stage sequence of SC:2028-2059
    faces(*c, *cur[s], f);
    m = limit_update(*c, a, *cur[s], f, *dst[s], coef[s][0], coef[s][1], dt, s == 2);
  }
  if (dt_next) *dt_next = c->scheme.cfl * m;
  return 0;
}

int32_t wamp_cpu_threads(const wamp_cpu_ctx* c) { return c ? c->threads : 0; }

}  // extern "C"
