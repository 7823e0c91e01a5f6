// wamp3_faces.h -- the 3D face-flux kernel (k3_faces) and the cell-access
// helpers of the 3D step, shared by two translation units: wamp3_kernels.hip
// instantiates the FP64 kernels (contraction off) and wamp3_faces_mixed.hip
// the FP32-mixed ones (contraction on, namespace wamp::mixed), as
// wamp_faces.h does for the 2D face kernel.
//
// Every 3D formula is the 2D one (SC: citations) with the third velocity
// component added in a fixed place, chosen so that a z-invariant state with
// w = 0 takes exactly the 2D step of its (x, y) slice and a y-invariant state
// with v = 0 exactly the 2D step of its (x, z) slice (z in the role of y),
// for the same dt: sums gain their third term last, (a + b) + c; tangential
// momentum products keep the 2D order rho * (u before v before w); each
// sweep's two shear waves use the 2D sign conventions (x-sweep: v, w as the
// 2D x-sweep's v; y-sweep: u as the 2D y-sweep's u, w as the x-sweep's v;
// z-sweep: u as the 2D y-sweep's u, v as the x-sweep's v).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wamp_device.h"
#include "wamp_types.h"

WAMP_DEV_BEGIN

__device__ __forceinline__ size_t at3(const Geo3& G, int i, int j, int k) {
  return (size_t)i * (size_t)G.sx + (size_t)j * (size_t)G.sy + (size_t)k;
}
__device__ __forceinline__ bool interior3(const Geo3& G, int i, int j, int k) {
  return i >= G.g && i < G.g + G.nx && j >= G.g && j < G.g + G.ny && k >= G.g && k < G.g + G.nz;
}

// Boundary fill by index mapping, axis by axis as map_cell (bc_x then bc_y,
// SC:1616-1682, extended by z): a ghost index maps to the edge interior
// cell; slip walls negate the wall-normal momentum.
__device__ __forceinline__ size_t map3(const Geo3& G, int i, int j, int k, double* sg) {
  sg[0] = sg[1] = sg[2] = 1.0;
  if (i < G.g) { i = G.g; if (G.bc[0]) sg[0] = -1.0; }
  else if (i >= G.g + G.nx) { i = G.g + G.nx - 1; if (G.bc[1]) sg[0] = -1.0; }
  if (j < G.g) { j = G.g; if (G.bc[2]) sg[1] = -1.0; }
  else if (j >= G.g + G.ny) { j = G.g + G.ny - 1; if (G.bc[3]) sg[1] = -1.0; }
  if (k < G.g) { k = G.g; if (G.bc[4]) sg[2] = -1.0; }
  else if (k >= G.g + G.nz) { k = G.g + G.nz - 1; if (G.bc[5]) sg[2] = -1.0; }
  return at3(G, i, j, k);
}
// The conserved state of cell (i, j, k) after the boundary fill.
__device__ __forceinline__ void load3(const Geo3& G, const State3& U, int i, int j, int k, double* q) {
  double sg[3];
  const size_t c = map3(G, i, j, k, sg);
#pragma unroll
  for (int m = 0; m < 7; ++m) q[m] = U.f[m][c];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (sg[d] < 0.0) q[2 + d] = -1.0 * q[2 + d];
}

// ------------------------------------------------------------- EOS --------
// cons_to_prim pressure (SC:856-860) with w (always FP64).
__device__ __forceinline__ double pressure3(const Eos& e, const double* q, double& uu, double& vv,
                                            double& ww) {
  double rho = q[0] + q[1];
  uu = ddiv(q[2], rho);
  vv = ddiv(q[3], rho);
  ww = ddiv(q[4], rho);
  double gb, pb;
  mix_gb_pb(e, q[6], gb, pb);
  return ddiv(q[5] - pb - 0.5 * rho * ((uu * uu + vv * vv) + ww * ww), gb);
}
// prim_to_energy SC:105-110 with w.
template <typename T>
__device__ __forceinline__ T prim_energy3(const Eos& e, T ar1, T ar2, T u, T v, T w, T p, T a1) {
  T rho = ar1 + ar2;
  T gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  return gb * p + pb + T(0.5) * rho * ((u * u + v * v) + w * w);
}
// rho_ctilde2_sc SC:174-178 on a (ar1, ar2, ru, rv, rw, p, a1) state.
__device__ __forceinline__ double rho_ct2_sc3(const Eos& e, const double* s) {
  double a1 = clamp01_pp(s[6]);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  return ddiv(s[5] + pi, g - 1.0);
}

// ------------------------------------------------------------ faces -------
// Face-constant part of the eigenvectors (SC:462-478) for sweep DIR.  n is
// the normal velocity; t1, t2 the two shear waves' tangential velocities as
// the projections use them (t1 = -u on the y- and z-sweeps, the 2D y-sweep's
// convention); vel the three face-averaged velocities.
template <typename T>
struct Basis3 {
  T c, k, inv2c, inv2c2, y1c2, y2c2, Y1, Y2, c2, n, t1, t2;
  T vel[3];
};
template <int DIR, typename T>
__device__ __forceinline__ Basis3<T> make_basis3(T c, T rho, T ar1, T ar2, const T* vel) {
  Basis3<T> b;
  b.c = c;
  b.inv2c = dv(T(0.5), c);
  const T c2 = c * c;
  b.inv2c2 = dv(T(0.5), c2);
  b.Y1 = dv(ar1, rho);
  b.Y2 = dv(ar2, rho);
  b.y1c2 = dv(b.Y1, c2);
  b.y2c2 = dv(b.Y2, c2);
  b.n = vel[DIR];
  b.k = b.n * b.inv2c;
  b.t1 = DIR == 0 ? vel[1] : -vel[0];
  b.t2 = DIR == 2 ? vel[1] : vel[2];
  b.c2 = c * c;
#pragma unroll
  for (int d = 0; d < 3; ++d) b.vel[d] = vel[d];
  return b;
}
// L . u on u = (ar1, ar2, ru, rv, rw, p, a1); w = (acoustic-, entropy 1,
// entropy 2, shear 1, shear 2, a1, acoustic+)  (to_char_x SC:480-499,
// to_char_y SC:537-556).
template <int DIR, typename T>
__device__ __forceinline__ void to_char3(const Basis3<T>& b, const T* u, T* w) {
  const T un_m = u[2 + DIR];
  w[0] = b.k * u[0] + b.k * u[1] - b.inv2c * un_m + b.inv2c2 * u[5];
  w[1] = u[0] - b.y1c2 * u[5];
  w[2] = u[1] - b.y2c2 * u[5];
  if (DIR == 0) {
    w[3] = -b.t1 * u[0] - b.t1 * u[1] + u[3];
    w[4] = -b.t2 * u[0] - b.t2 * u[1] + u[4];
  } else if (DIR == 1) {
    w[3] = -b.t1 * u[0] - b.t1 * u[1] - u[2];
    w[4] = -b.t2 * u[0] - b.t2 * u[1] + u[4];
  } else {
    w[3] = -b.t1 * u[0] - b.t1 * u[1] - u[2];
    w[4] = -b.t2 * u[0] - b.t2 * u[1] + u[3];
  }
  w[5] = u[6];
  w[6] = -b.k * u[0] - b.k * u[1] + b.inv2c * un_m + b.inv2c2 * u[5];
}
// R . w  (to_phys_x SC:504-534, to_phys_y SC:559-589)
template <int DIR, typename T>
__device__ __forceinline__ void to_phys3(const Basis3<T>& b, const T* w, T* s) {
  const T n = b.n, c = b.c;
  s[0] = b.Y1 * (w[0] + w[6]) + w[1];
  s[1] = b.Y2 * (w[0] + w[6]) + w[2];
  s[2 + DIR] = (n - c) * w[0] + n * w[1] + n * w[2] + (n + c) * w[6];
  if (DIR == 0) {
    s[3] = b.vel[1] * (w[0] + w[1] + w[2] + w[6]) + w[3];
    s[4] = b.vel[2] * (w[0] + w[1] + w[2] + w[6]) + w[4];
  } else if (DIR == 1) {
    s[2] = b.vel[0] * (w[0] + w[1] + w[2] + w[6]) - w[3];
    s[4] = b.vel[2] * (w[0] + w[1] + w[2] + w[6]) + w[4];
  } else {
    s[2] = b.vel[0] * (w[0] + w[1] + w[2] + w[6]) - w[3];
    s[3] = b.vel[1] * (w[0] + w[1] + w[2] + w[6]) + w[4];
  }
  s[5] = b.c2 * (w[0] + w[6]);
  s[6] = w[5];
}

// wong_stage_lower_s6 SC:189-201 on 7 components.
__device__ __forceinline__ void wong_lower3(const double* b, double* s, double bv, double hv,
                                            double floor_) {
  double theta = 1.0;
  if (isnan_(bv) || isnan_(hv) || bv < floor_) {
    theta = 0.0;
  } else if (hv < floor_) {
    double denom = bv - hv;
    theta = denom > TINY ? ddiv(bv - floor_, denom) : 0.0;
  }
#pragma unroll
  for (int m = 0; m < 7; ++m) s[m] = b[m] + theta * (s[m] - b[m]);
}
// The M10 face positivity sequence for one side (M10:1007-1021) on
// (ar1, ar2, ru, rv, rw, p, a1), as face_state_pp (always FP64).
__device__ __forceinline__ void face_state_pp3(const Eos& e, double* hi, const double* base) {
  double b[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) b[m] = base[m];
  b[6] = clamp01_pp(b[6]);
  b[0] = floor_pp(b[0], PP_RHO);
  b[1] = floor_pp(b[1], PP_RHO);
  wong_lower3(b, hi, b[0], hi[0], PP_RHO);
  wong_lower3(b, hi, b[1], hi[1], PP_RHO);
  wong_lower3(b, hi, b[6], hi[6], PP_ALPHA);
  wong_lower3(b, hi, 1.0 - b[6], 1.0 - hi[6], PP_ALPHA);
  wong_lower3(b, hi, rho_ct2_sc3(e, b), rho_ct2_sc3(e, hi), PP_CTILDE);
  bool back = hi[0] < PP_HS_RHO || hi[1] < PP_HS_RHO || hi[6] < PP_HS_ALPHA ||
              hi[6] > 1.0 - PP_HS_ALPHA;
  if (back) {
#pragma unroll
    for (int m = 0; m < 7; ++m) hi[m] = b[m];
  }
  double v = rho_ct2_sc3(e, hi);
  if (v < PP_HS_CTILDE || isnan_(hi[5])) {
#pragma unroll
    for (int m = 0; m < 7; ++m) hi[m] = b[m];
    v = rho_ct2_sc3(e, hi);
  }
  if (hi[0] < 0.0 || isnan_(hi[0])) hi[0] = base[0];
  if (hi[1] < 0.0 || isnan_(hi[1])) hi[1] = base[1];
  if (hi[6] < 0.0 || hi[6] > 1.0) {
    hi[6] = base[6];
    v = rho_ct2_sc3(e, hi);
  }
  if (v < PP_CTILDE || isnan_(hi[5])) {
    hi[5] = base[5];
    v = rho_ct2_sc3(e, hi);
  }
  const double a1 = clamp01_pp(hi[6]);
  if (isnan_(hi[5]) || v < PP_CTILDE) {
    double g, pi;
    mix_g_pi(e, a1, g, pi);
    hi[5] = PP_CTILDE * (g - 1.0) - pi;
  }
  hi[6] = a1;
}

// Physical momentum flux of component d through a DIR face: rho n n + p for
// the normal component, rho a b with (a, b) in u, v, w order otherwise.
template <int DIR, typename T>
__device__ __forceinline__ T mflux3(const T* v, T rho, T p, int d) {
  if (d == DIR) return rho * v[d] * v[d] + p;
  return d < DIR ? rho * v[d] * v[DIR] : rho * v[DIR] * v[d];
}

// HLLC flux + contact velocity (hllc_x/_src SC:594-687, hllc_y/_src
// SC:690-771) on primitives (ar1, ar2, u, v, w, p, a1).
template <int DIR>
__device__ __forceinline__ void hllc3(const Eos& e, const double* L, const double* R, double* f,
                                      double& ustar) {
  const double* vL = L + 2;
  const double* vR = R + 2;
  const double nL = vL[DIR], nR = vR[DIR];
  const double pL = L[5], pR = R[5];
  double rhoL = L[0] + L[1];
  double rhoR = R[0] + R[1];
  double cL = sound_speed(e, rhoL, pL, L[6]);
  double cR = sound_speed(e, rhoR, pR, R[6]);
  double EL = prim_energy3(e, L[0], L[1], vL[0], vL[1], vL[2], pL, L[6]);
  double ER = prim_energy3(e, R[0], R[1], vR[0], vR[1], vR[2], pR, R[6]);
  double n_avg = 0.5 * (nL + nR);
  double c_avg = 0.5 * (cL + cR);
  double SL = wmin(nL - cL, n_avg - c_avg);
  double SR = wmax(nR + cR, n_avg + c_avg);
  double denom = rhoL * (SL - nL) - rhoR * (SR - nR);
  if (fabs(denom) < TINY) denom = TINY;
  double SP = ddiv(pR - pL + rhoL * nL * (SL - nL) - rhoR * nR * (SR - nR), denom);
  double ratL = (SL - nL) / (SL - SP);
  double ratR = (SR - nR) / (SR - SP);
  if (SL > 0.0) {
    f[0] = L[0] * nL;
    f[1] = L[1] * nL;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[2 + d] = mflux3<DIR>(vL, rhoL, pL, d);
    f[5] = (EL + pL) * nL;
    f[6] = L[6] * nL;
  } else if (SL <= 0.0 && 0.0 < SP) {
    double ratio = ratL;
    double EL_star = EL + (SP - nL) * (SP * rhoL + ddiv(pL, SL - nL));
    f[0] = L[0] * nL + SL * (L[0] * ratio - L[0]);
    f[1] = L[1] * nL + SL * (L[1] * ratio - L[1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double star = d == DIR ? rhoL * SP * ratio - rhoL * vL[d] : rhoL * vL[d] * ratio - rhoL * vL[d];
      f[2 + d] = mflux3<DIR>(vL, rhoL, pL, d) + SL * star;
    }
    f[5] = (EL + pL) * nL + SL * (EL_star * ratio - EL);
    f[6] = L[6] * nL + SL * (L[6] * ratio - L[6]);
  } else if (SP <= 0.0 && 0.0 <= SR) {
    double ratio = ratR;
    double ER_star = ER + (SP - nR) * (SP * rhoR + ddiv(pR, SR - nR));
    f[0] = R[0] * nR + SR * (R[0] * ratio - R[0]);
    f[1] = R[1] * nR + SR * (R[1] * ratio - R[1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double star = d == DIR ? rhoR * SP * ratio - rhoR * vR[d] : rhoR * vR[d] * ratio - rhoR * vR[d];
      f[2 + d] = mflux3<DIR>(vR, rhoR, pR, d) + SR * star;
    }
    f[5] = (ER + pR) * nR + SR * (ER_star * ratio - ER);
    f[6] = R[6] * nR + SR * (R[6] * ratio - R[6]);
  } else {
    f[0] = R[0] * nR;
    f[1] = R[1] * nR;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[2 + d] = mflux3<DIR>(vR, rhoR, pR, d);
    f[5] = (ER + pR) * nR;
    f[6] = R[6] * nR;
  }
  double left_part = nL + wmin(0.0, SL) * (ratL - 1.0);
  double right_part = nR + wmax(0.0, SR) * (ratR - 1.0);
  double sp_sign = wsign(SP);
  ustar = 0.5 * (1.0 + sp_sign) * left_part + 0.5 * (1.0 - sp_sign) * right_part;
}

// HLLC for the mixed-precision path, as hllc_mixed: wave speeds, contact
// speed and the mass / momentum / volume-fraction fluxes in FP32, the energy
// flux in FP64 from the FP32 primitives (E carries the p_inf term).
template <int DIR>
__device__ __forceinline__ void hllc3_mixed(const Eos& e, const float* L, const float* R, double* f,
                                            double& ustar) {
  const float* vL = L + 2;
  const float* vR = R + 2;
  const float nL = vL[DIR], nR = vR[DIR];
  const float pL = L[5], pR = R[5];
  float rhoL = L[0] + L[1];
  float rhoR = R[0] + R[1];
  float cL = sound_speed<float>(e, rhoL, pL, L[6]);
  float cR = sound_speed<float>(e, rhoR, pR, R[6]);
  float n_avg = 0.5f * (nL + nR);
  float c_avg = 0.5f * (cL + cR);
  float SL = wmin(nL - cL, n_avg - c_avg);
  float SR = wmax(nR + cR, n_avg + c_avg);
  float denom = rhoL * (SL - nL) - rhoR * (SR - nR);
  if (fabsf(denom) < float(TINY)) denom = float(TINY);
  float SP = dv(pR - pL + rhoL * nL * (SL - nL) - rhoR * nR * (SR - nR), denom);
  float ratL = dv(SL - nL, SL - SP);
  float ratR = dv(SR - nR, SR - SP);
  const double EL = prim_energy3<double>(e, L[0], L[1], vL[0], vL[1], vL[2], pL, L[6]);
  const double ER = prim_energy3<double>(e, R[0], R[1], vR[0], vR[1], vR[2], pR, R[6]);
  if (SL > 0.f) {
    f[0] = L[0] * nL;
    f[1] = L[1] * nL;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[2 + d] = mflux3<DIR>(vL, rhoL, pL, d);
    f[5] = (EL + (double)pL) * (double)nL;
    f[6] = L[6] * nL;
  } else if (SL <= 0.f && 0.f < SP) {
    float ratio = ratL;
    f[0] = L[0] * nL + SL * (L[0] * ratio - L[0]);
    f[1] = L[1] * nL + SL * (L[1] * ratio - L[1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float star = d == DIR ? rhoL * SP * ratio - rhoL * vL[d] : rhoL * vL[d] * ratio - rhoL * vL[d];
      f[2 + d] = mflux3<DIR>(vL, rhoL, pL, d) + SL * star;
    }
    const double sl = SL, sp = SP, nl = nL, rl = rhoL, pl = pL;
    const double EL_star = EL + (sp - nl) * (sp * rl + ddiv(pl, sl - nl));
    f[5] = (EL + pl) * nl + sl * (EL_star * (double)ratio - EL);
    f[6] = L[6] * nL + SL * (L[6] * ratio - L[6]);
  } else if (SP <= 0.f && 0.f <= SR) {
    float ratio = ratR;
    f[0] = R[0] * nR + SR * (R[0] * ratio - R[0]);
    f[1] = R[1] * nR + SR * (R[1] * ratio - R[1]);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float star = d == DIR ? rhoR * SP * ratio - rhoR * vR[d] : rhoR * vR[d] * ratio - rhoR * vR[d];
      f[2 + d] = mflux3<DIR>(vR, rhoR, pR, d) + SR * star;
    }
    const double sr = SR, sp = SP, nr = nR, rr = rhoR, pr = pR;
    const double ER_star = ER + (sp - nr) * (sp * rr + ddiv(pr, sr - nr));
    f[5] = (ER + pr) * nr + sr * (ER_star * (double)ratio - ER);
    f[6] = R[6] * nR + SR * (R[6] * ratio - R[6]);
  } else {
    f[0] = R[0] * nR;
    f[1] = R[1] * nR;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[2 + d] = mflux3<DIR>(vR, rhoR, pR, d);
    f[5] = (ER + (double)pR) * (double)nR;
    f[6] = R[6] * nR;
  }
  float left_part = nL + wmin(0.f, SL) * (ratL - 1.f);
  float right_part = nR + wmax(0.f, SR) * (ratR - 1.f);
  float sp_sign = wsign(SP);
  ustar = 0.5f * (1.f + sp_sign) * left_part + 0.5f * (1.f - sp_sign) * right_part;
}

template <int DIR>
__device__ __forceinline__ long long stride3(const Geo3& G) {
  return DIR == 0 ? G.sx : DIR == 1 ? G.sy : 1;
}

// Store of face plane m: the FP64 kernels write FP64 storage; the mixed ones
// write whichever storage F describes.
template <typename T>
__device__ __forceinline__ void fstore3(const Faces3& F, int m, size_t c, double v) {
  if constexpr (sizeof(T) == sizeof(float)) {
    if (F.h) {
      if (face3_is_energy(m)) F.b[(long long)face3_slot64(m) * F.s + (long long)c] = v;
      else F.h[(long long)face3_slot32(m) * F.hs + (long long)c] = (float)v;
      return;
    }
    // FP64 storage of the mixed kernels' planes (the split entry points):
    // the FP32 planes' values widened, as the limiter reads them.
    if (!face3_is_energy(m)) v = (double)(float)v;
  }
  F.b[(long long)m * F.s + (long long)c] = v;
}

// One face of sweep DIR between cells (i, j, k) and its +DIR neighbour:
// high-order fluxes + u* into out[0..7], first-order ones into out[8..15]
// (compute_fluxes_x/y SC:1030-1128 and fx/gy_low SC:1303-1394 with w).
// T: the arithmetic of projection, reconstruction and HLLC (double, or float
// in the mixed unit); pressure, indicators, the M10 face limiter and the
// energy flux stay FP64.
template <int DIR, int BASE, int SHEAR, int FACEPP, typename T>
__device__ __forceinline__ void face3(const Geo3& G, const Eos& e, const State3& U,
                                      const double* __restrict__ P, const double* __restrict__ S,
                                      const double* __restrict__ D, int vf_always, double t1d, int i,
                                      int j, int k, double* out) {
  const long long s = stride3<DIR>(G);
  const size_t c0 = at3(G, i, j, k);
  constexpr bool wide = BASE == 1 || SHEAR == 2;
  constexpr int klo = wide ? 0 : 1, khi = wide ? 6 : 5;
  const T t1 = (T)t1d;
  // indicators at cells -1..2 (cator2_val SC:318-325, ducros SC:875-916)
  double ci[4], sg[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const long long r = (long long)c0 + (long long)(q - 1) * s;
    ci[q] = cator2(S[r - 2 * s], S[r - s], S[r], S[r + s], S[r + 2 * s]);
    sg[q] = 0.0;
    if (SHEAR != 0) {
      double pm2 = P[r - 2 * s], pm1 = P[r - s], p0 = P[r], pp1 = P[r + s], pp2 = P[r + 2 * s];
      double aa = fabs(-pm2 + 16.0 * pm1 - 30.0 * p0 + 16.0 * pp1 - pp2);
      double bb = fabs(pm2 + 16.0 * pm1 + 30.0 * p0 + 16.0 * pp1 + pp2);
      sg[q] = ddiv(aa, bb + 1.0e-30) * D[r];
    }
  }
  // stencil cells -2..3 (boundary-filled) with their pressure.  On the x- and
  // y-sweeps, where every stencil of the block is interior, the boundary
  // mapping is the identity and the cells are addressed directly; the test
  // covers the block's index range along the sweep (k3_faces' 64 x 4 blocks:
  // one i, 4 j) so that it is uniform over each wave.  The z-sweep keeps the
  // mapping: with 64 k per block most of its blocks touch a z boundary, and
  // carrying both paths made it slower (0.86 vs 0.82 ms at 320x160x160).
  const int ip = DIR == 0 ? i : j;
  const int np = DIR == 0 ? G.nx : G.ny;
  const int lo = DIR == 0 ? ip : ip - (int)threadIdx.y;
  const int hi = DIR == 0 ? ip : lo + 3;
  const bool inner = DIR != 2 && lo + (klo - 2) >= G.g && hi + (khi - 3) < G.g + np;
  double u[6][7];
#pragma unroll
  for (int q = klo; q < khi; ++q) {
    const int o = q - 2;
    double cq[7];
    if (DIR != 2 && inner) {
      const size_t cc = (size_t)((long long)c0 + (long long)o * s);
#pragma unroll
      for (int m = 0; m < 7; ++m) cq[m] = U.f[m][cc];
    } else {
      load3(G, U, i + (DIR == 0 ? o : 0), j + (DIR == 1 ? o : 0), k + (DIR == 2 ? o : 0), cq);
    }
    u[q][0] = cq[0];
    u[q][1] = cq[1];
    u[q][2] = cq[2];
    u[q][3] = cq[3];
    u[q][4] = cq[4];
    u[q][5] = P[(long long)c0 + (long long)o * s];
    u[q][6] = cq[6];
  }
  {  // high order (SC:1030-1128, M10:960-1040)
    bool use_thinc = wmin(wmin(ci[0], ci[1]), wmin(ci[2], ci[3])) < THINC_THRESH;
    bool use_smooth = false;
    if (SHEAR != 0) use_smooth = wmax(wmax(sg[0], sg[1]), wmax(sg[2], sg[3])) < SIGMA_SMOOTH;
    const double* A = u[2];
    const double* B = u[3];
    T ar1a = T(0.5) * ((T)A[0] + (T)B[0]);
    T ar2a = T(0.5) * ((T)A[1] + (T)B[1]);
    T a1a = T(0.5) * ((T)A[6] + (T)B[6]);
    T pa = T(0.5) * ((T)A[5] + (T)B[5]);
    T rhoa = ar1a + ar2a;
    T vel[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[d] = dv(T(0.5) * ((T)A[2 + d] + (T)B[2 + d]), rhoa);
    T ca = sound_speed<T>(e, rhoa, pa, a1a);
    Basis3<T> Bs = make_basis3<DIR, T>(ca, rhoa, ar1a, ar2a, vel);
    T w[6][7];
#pragma unroll
    for (int q = klo; q < khi; ++q) {
      T uq[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) uq[m] = (T)u[q][m];
      to_char3<DIR>(Bs, uq, w[q]);
    }
    T wl[7], wr[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      if (BASE == 0) {
        wl[m] = muscl_l(w[1][m], w[2][m], w[3][m]);
        wr[m] = muscl_r(w[2][m], w[3][m], w[4][m]);
      } else {
        wl[m] = mp5_l(w[0][m], w[1][m], w[2][m], w[3][m], w[4][m]);
        wr[m] = mp5_r(w[1][m], w[2][m], w[3][m], w[4][m], w[5][m]);
      }
    }
    if (SHEAR != 0 && use_smooth) {
#pragma unroll
      for (int m = 3; m < 5; ++m) {
        T iph = SHEAR == 1 ? central4(w[1][m], w[2][m], w[3][m], w[4][m])
                           : central6(w[0][m], w[1][m], w[2][m], w[3][m], w[4][m], w[5][m]);
        wl[m] = iph;
        wr[m] = iph;
      }
    }
    if (use_thinc) {
      wl[1] = thinc_l(w[1][1], w[2][1], w[3][1], t1);
      wl[2] = thinc_l(w[1][2], w[2][2], w[3][2], t1);
      wr[1] = thinc_r(w[2][1], w[3][1], w[4][1], t1);
      wr[2] = thinc_r(w[2][2], w[3][2], w[4][2], t1);
    }
    if (use_thinc || vf_always) {
      wl[5] = thinc_l(w[1][5], w[2][5], w[3][5], t1);
      wr[5] = thinc_r(w[2][5], w[3][5], w[4][5], t1);
    }
    T phL[7], phR[7], phfL[7], phfR[7];
    to_phys3<DIR>(Bs, wl, phL);
    to_phys3<DIR>(Bs, wr, phR);
    to_phys3<DIR>(Bs, w[2], phfL);
    to_phys3<DIR>(Bs, w[3], phfR);
    const bool cbad = ca < T(0) || isnan_(ca);
    if (FACEPP == 0) {  // SC:1103-1116
      if (phL[0] < T(0) || isnan_(phL[0])) phL[0] = phfL[0];
      if (phL[1] < T(0) || isnan_(phL[1])) phL[1] = phfL[1];
      if (phL[6] < T(0) || phL[6] > T(1)) phL[6] = phfL[6];
      if (isnan_(phL[5])) phL[5] = phfL[5];
      if (phR[0] < T(0) || isnan_(phR[0])) phR[0] = phfR[0];
      if (phR[1] < T(0) || isnan_(phR[1])) phR[1] = phfR[1];
      if (phR[6] < T(0) || phR[6] > T(1)) phR[6] = phfR[6];
      if (isnan_(phR[5])) phR[5] = phfR[5];
      if (cbad) {
#pragma unroll
        for (int m = 0; m < 7; ++m) {
          phL[m] = phfL[m];
          phR[m] = phfR[m];
        }
      }
    } else if constexpr (sizeof(T) == sizeof(double)) {  // M10:1007-1025
      face_state_pp3(e, phL, phfL);
      if (cbad) {
#pragma unroll
        for (int m = 0; m < 7; ++m) phL[m] = phfL[m];
      }
      face_state_pp3(e, phR, phfR);
      if (cbad) {
#pragma unroll
        for (int m = 0; m < 7; ++m) phR[m] = phfR[m];
      }
    } else {  // the same on FP64 copies: the thresholds are always applied in FP64
      double d[7], df[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        d[m] = phL[m];
        df[m] = phfL[m];
      }
      face_state_pp3(e, d, df);
#pragma unroll
      for (int m = 0; m < 7; ++m) phL[m] = cbad ? phfL[m] : (T)d[m];
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        d[m] = phR[m];
        df[m] = phfR[m];
      }
      face_state_pp3(e, d, df);
#pragma unroll
      for (int m = 0; m < 7; ++m) phR[m] = cbad ? phfR[m] : (T)d[m];
    }
    T rhoL = phL[0] + phL[1], rhoR = phR[0] + phR[1];
    T PL[7] = {phL[0], phL[1], dv(phL[2], rhoL), dv(phL[3], rhoL), dv(phL[4], rhoL), phL[5], phL[6]};
    T PR[7] = {phR[0], phR[1], dv(phR[2], rhoR), dv(phR[3], rhoR), dv(phR[4], rhoR), phR[5], phR[6]};
    if constexpr (sizeof(T) == sizeof(double)) {
      hllc3<DIR>(e, PL, PR, out, out[7]);
    } else {
      hllc3_mixed<DIR>(e, PL, PR, out, out[7]);
    }
  }
  {  // first order (fx_low_kernel SC:1303-1347), primitives formed in FP64
    const double* A = u[2];
    const double* B = u[3];
    double rhoL = A[0] + A[1];
    double rhoR = B[0] + B[1];
    double aL = clamp01_pp(A[6]);
    double aR = clamp01_pp(B[6]);
    double pL = A[5];
    double pR = B[5];
    double gL, piL, gR, piR;
    mix_g_pi(e, aL, gL, piL);
    mix_g_pi(e, aR, gR, piR);
    if (isnan_(pL) || ddiv(pL + piL, gL - 1.0) < PP_CTILDE) pL = PP_CTILDE * (gL - 1.0) - piL;
    if (isnan_(pR) || ddiv(pR + piR, gR - 1.0) < PP_CTILDE) pR = PP_CTILDE * (gR - 1.0) - piR;
    T PL[7] = {(T)A[0], (T)A[1], (T)ddiv(A[2], rhoL), (T)ddiv(A[3], rhoL), (T)ddiv(A[4], rhoL), (T)pL,
               (T)aL};
    T PR[7] = {(T)B[0], (T)B[1], (T)ddiv(B[2], rhoR), (T)ddiv(B[3], rhoR), (T)ddiv(B[4], rhoR), (T)pR,
               (T)aR};
    if constexpr (sizeof(T) == sizeof(double)) {
      hllc3<DIR>(e, PL, PR, out + 8, out[15]);
    } else {
      hllc3_mixed<DIR>(e, PL, PR, out + 8, out[15]);
    }
  }
}

// Faces of sweep DIR: (i, j, k) with the DIR index in [g-1, g+n) and the
// other two interior; one thread per face, k fastest.  (bx, by, bz): the
// block's coordinates in the grid3 block grid of the sweep.
template <int DIR, int BASE, int SHEAR, int FACEPP, typename T>
__device__ __forceinline__ void faces3_block(const Geo3& G, const Eos& e, const State3& U,
                                             const double* __restrict__ P,
                                             const double* __restrict__ S,
                                             const double* __restrict__ D, const Faces3& F,
                                             int vf_always, double t1, int bx, int by, int bz) {
  const int k = G.g - (DIR == 2 ? 1 : 0) + (int)(bx * 64 + threadIdx.x);
  const int j = G.g - (DIR == 1 ? 1 : 0) + (int)(by * 4 + threadIdx.y);
  const int i = G.g - (DIR == 0 ? 1 : 0) + bz;
  if (k >= G.g + G.nz || j >= G.g + G.ny) return;
  double out[16];
  face3<DIR, BASE, SHEAR, FACEPP, T>(G, e, U, P, S, D, vf_always, t1, i, j, k, out);
  const size_t c = at3(G, i, j, k);
#pragma unroll
  for (int m = 0; m < 16; ++m) fstore3<T>(F, DIR * 16 + m, c, out[m]);
}
template <int DIR, int BASE, int SHEAR, int FACEPP, typename T>
__global__ void __launch_bounds__(256) k3_faces(Geo3 G, Eos e, State3 U, const double* __restrict__ P,
                                                const double* __restrict__ S,
                                                const double* __restrict__ D, Faces3 F, int vf_always,
                                                double t1) {
  faces3_block<DIR, BASE, SHEAR, FACEPP, T>(G, e, U, P, S, D, F, vf_always, t1, (int)blockIdx.x,
                                            (int)blockIdx.y, (int)blockIdx.z);
}

inline dim3 grid3(int nk, int nj, int ni) {
  return dim3((unsigned)((nk + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)ni);
}

// ---------------------------------------------------------- batches -------
// A batched launch runs the blocks of every active case of the table: its
// flat block index b falls in case k where off[k] <= b < off[k + 1] (prefix
// sums of the cases' block counts of one work kind), and the local index
// b - off[k] is the case's block in its own grid3 block grid.  Every thread
// of a block belongs to one case, so the block-wide reductions and barriers
// of the single-case kernels hold.
__device__ __forceinline__ int find_case3(const int* __restrict__ off, int n, int b) {
  int lo = 0, hi = n;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid;
  }
  return lo;
}
// Case and block coordinates of this block for work kind `kind`; false when
// the case has finished (the whole block then returns).
__device__ __forceinline__ bool batch_block3(const Case3Desc* __restrict__ cs,
                                             const int* __restrict__ off, int ncase,
                                             const int* __restrict__ done, int kind, int& k,
                                             int& bx, int& by, int& bz) {
  const int* o = off + kind * (ncase + 1);
  const int b = (int)blockIdx.x;
  k = __builtin_amdgcn_readfirstlane(find_case3(o, ncase, b));
  if (done[k]) return false;
  int nk, nj, ni;
  kind_extent3(cs[k].geo, kind, nk, nj, ni);
  const int gx = (nk + 63) / 64, gy = (nj + 3) / 4;
  const int local = b - o[k];
  bx = local % gx;
  by = (local / gx) % gy;
  bz = local / (gx * gy);
  return true;
}
// The faces of sweep DIR of every active case on stage `stage`'s input.
template <int DIR, int BASE, int SHEAR, int FACEPP, typename T>
__global__ void __launch_bounds__(256) k3b_faces(const Case3Desc* __restrict__ cs,
                                                 const int* __restrict__ off, int ncase,
                                                 BatchScal sc, int stage, int vf_always, double t1) {
  int k, bx, by, bz;
  if (!batch_block3(cs, off, ncase, sc.done, B3_FACE0 + DIR, k, bx, by, bz)) return;
  const Case3Desc& c = cs[k];
  State3 U;
#pragma unroll
  for (int m = 0; m < 7; ++m) U.f[m] = c.u[stage][m];
  faces3_block<DIR, BASE, SHEAR, FACEPP, T>(c.geo, c.eos, U, c.P, c.S, c.D, c.f, vf_always, t1, bx,
                                            by, bz);
}

// The three sweeps of one stage for one scheme.
template <int BASE, int SHEAR, int FACEPP, typename T>
inline hipError_t launch_faces3_t(const Geo3& G, const Eos& e, const State3& U, const double* P,
                                  const double* S, const double* D, const Faces3& F, int vf_always,
                                  double t1, hipStream_t st) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL((k3_faces<0, BASE, SHEAR, FACEPP, T>), grid3(G.nz, G.ny, G.nx + 1), b, 0, st,
                     G, e, U, P, S, D, F, vf_always, t1);
  hipLaunchKernelGGL((k3_faces<1, BASE, SHEAR, FACEPP, T>), grid3(G.nz, G.ny + 1, G.nx), b, 0, st,
                     G, e, U, P, S, D, F, vf_always, t1);
  hipLaunchKernelGGL((k3_faces<2, BASE, SHEAR, FACEPP, T>), grid3(G.nz + 1, G.ny, G.nx), b, 0, st,
                     G, e, U, P, S, D, F, vf_always, t1);
  return hipGetLastError();
}
// The three sweeps of one stage of a batch; blocks[d]: the table's total
// block count of sweep d.
template <int BASE, int SHEAR, int FACEPP, typename T>
inline hipError_t launch_faces3_batch_t(const Case3Desc* cs, const int* off, int ncase,
                                        const BatchScal& sc, int stage, int vf_always, double t1,
                                        const int* blocks, hipStream_t st) {
  const dim3 b(64, 4);
  if (blocks[0] > 0)
    hipLaunchKernelGGL((k3b_faces<0, BASE, SHEAR, FACEPP, T>), dim3(blocks[0]), b, 0, st, cs, off,
                       ncase, sc, stage, vf_always, t1);
  if (blocks[1] > 0)
    hipLaunchKernelGGL((k3b_faces<1, BASE, SHEAR, FACEPP, T>), dim3(blocks[1]), b, 0, st, cs, off,
                       ncase, sc, stage, vf_always, t1);
  if (blocks[2] > 0)
    hipLaunchKernelGGL((k3b_faces<2, BASE, SHEAR, FACEPP, T>), dim3(blocks[2]), b, 0, st, cs, off,
                       ncase, sc, stage, vf_always, t1);
  return hipGetLastError();
}
template <typename T>
inline hipError_t launch_faces3_batch(const Case3Desc* cs, const int* off, int ncase,
                                      const BatchScal& sc, int stage, int base, int shear,
                                      int vf_always, int face_pp, double t1, const int* blocks,
                                      hipStream_t st) {
  if (base == 0 && shear == 1 && face_pp == 0)
    return launch_faces3_batch_t<0, 1, 0, T>(cs, off, ncase, sc, stage, vf_always, t1, blocks, st);
  if (base == 0 && shear == 0 && face_pp == 1)
    return launch_faces3_batch_t<0, 0, 1, T>(cs, off, ncase, sc, stage, vf_always, t1, blocks, st);
  if (base == 1 && shear == 2 && face_pp == 0)
    return launch_faces3_batch_t<1, 2, 0, T>(cs, off, ncase, sc, stage, vf_always, t1, blocks, st);
  return hipErrorInvalidValue;
}
template <typename T>
inline hipError_t launch_faces3(const Geo3& G, const Eos& e, const State3& U, const double* P,
                                const double* S, const double* D, const Faces3& F, int base,
                                int shear, int vf_always, int face_pp, double t1, hipStream_t st) {
  if (base == 0 && shear == 1 && face_pp == 0)
    return launch_faces3_t<0, 1, 0, T>(G, e, U, P, S, D, F, vf_always, t1, st);
  if (base == 0 && shear == 0 && face_pp == 1)
    return launch_faces3_t<0, 0, 1, T>(G, e, U, P, S, D, F, vf_always, t1, st);
  if (base == 1 && shear == 2 && face_pp == 0)
    return launch_faces3_t<1, 2, 0, T>(G, e, U, P, S, D, F, vf_always, t1, st);
  return hipErrorInvalidValue;
}

WAMP_DEV_END
