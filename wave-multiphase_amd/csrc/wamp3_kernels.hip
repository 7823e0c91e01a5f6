// wamp3_kernels.hip -- the three-dimensional extension of the step: the
// same SSP-RK3 stage (boundary fill, characteristic MUSCL/MP5 + THINC +
// central reconstruction, HLLC with the alpha1 source, the three-stage Wong
// limiter, repair, alpha clamp, CFL dt) with a z-sweep and the 7-variable
// state [a1 rho1, a2 rho2, rho u, rho v, rho w, E, a1].
//
// The reference solves 2D only (its README mentions 3D runs whose code it
// does not hold).  Every 3D formula here is the 2D one (SC: citations) with
// the third velocity component added in a fixed place, chosen so that
//  * a z-invariant state with w = 0 takes exactly the 2D step of its (x, y)
//    slice, and
//  * a y-invariant state with v = 0 takes exactly the 2D step of its (x, z)
//    slice with z in the role of y,
// for the same dt.  Sums gain their third term last, (a + b) + c, tangential
// momentum products keep the 2D order rho * (u before v before w), and each
// sweep's two shear waves use the 2D sign conventions: x-sweep (v, w) as the
// 2D x-sweep's v; y-sweep u as the 2D y-sweep's u and w as the x-sweep's v;
// z-sweep u as the 2D y-sweep's u and v as the x-sweep's v.
//
// Kernels per stage: k3_cells (pressure, entropy and the Ducros factor once
// per cell of the padded grid), k3_faces<DIR> (one thread per face: high-
// and first-order fluxes of one sweep), k3_theta<S> (one limiter theta per
// cell and stage), k3_update (blended fluxes, RK combine, repair, clamp, dt).
// FP64 throughout, FMA contraction off, built with wamp_kernels.hip's flags.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wamp_device.h"
#include "wamp_types.h"

namespace wamp {

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
// cons_to_prim pressure (SC:856-860) with w.
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
__device__ __forceinline__ double prim_energy3(const Eos& e, double ar1, double ar2, double u,
                                               double v, double w, double p, double a1) {
  double rho = ar1 + ar2;
  double gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  return gb * p + pb + 0.5 * rho * ((u * u + v * v) + w * w);
}
// rho_ctilde2_cons SC:242-258 with w.
__device__ __forceinline__ double rho_ct2_cons3(const Eos& e, const double* q) {
  double a1 = clamp01_pp(q[6]);
  double m1 = q[0], m2 = q[1];
  if (m1 < PP_RHO || isnan_(m1)) m1 = PP_RHO;
  if (m2 < PP_RHO || isnan_(m2)) m2 = PP_RHO;
  double rho = m1 + m2;
  double uu = ddiv(q[2], rho);
  double vv = ddiv(q[3], rho);
  double ww = ddiv(q[4], rho);
  double kin = 0.5 * rho * ((uu * uu + vv * vv) + ww * ww);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  return q[5] - kin - pi;
}
// rho_ctilde2_sc SC:174-178 on a (ar1, ar2, ru, rv, rw, p, a1) state.
__device__ __forceinline__ double rho_ct2_sc3(const Eos& e, const double* s) {
  double a1 = clamp01_pp(s[6]);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  return ddiv(s[5] + pi, g - 1.0);
}

// ------------------------------------------------------- per cell ---------
// Ducros factor div^2 / (div^2 + |curl|^2 + 1e-30) of cell (i, j, k) from
// the boundary-filled neighbours (ducros_kernel SC:875-916, with w).
__device__ __forceinline__ void vel3(const Geo3& G, const State3& U, int i, int j, int k, double* v) {
  double q[7];
  load3(G, U, i, j, k, q);
  const double rho = q[0] + q[1];
  v[0] = ddiv(q[2], rho);
  v[1] = ddiv(q[3], rho);
  v[2] = ddiv(q[4], rho);
}
__device__ __forceinline__ double ducros3(const Geo3& G, const State3& U, int i, int j, int k) {
  double xm[3], xp[3], ym[3], yp[3], zm[3], zp[3];
  vel3(G, U, i - 1, j, k, xm);
  vel3(G, U, i + 1, j, k, xp);
  vel3(G, U, i, j - 1, k, ym);
  vel3(G, U, i, j + 1, k, yp);
  vel3(G, U, i, j, k - 1, zm);
  vel3(G, U, i, j, k + 1, zp);
  const double der_ux = 0.5 * (-xm[0] + xp[0]), der_vx = 0.5 * (-xm[1] + xp[1]);
  const double der_wx = 0.5 * (-xm[2] + xp[2]);
  const double der_uy = 0.5 * (-ym[0] + yp[0]), der_vy = 0.5 * (-ym[1] + yp[1]);
  const double der_wy = 0.5 * (-ym[2] + yp[2]);
  const double der_uz = 0.5 * (-zm[0] + zp[0]), der_vz = 0.5 * (-zm[1] + zp[1]);
  const double der_wz = 0.5 * (-zm[2] + zp[2]);
  const double div_u = (der_ux + der_vy) + der_wz;
  const double cz = der_vx - der_uy, cx = der_wy - der_vz, cy = der_uz - der_wx;
  return ddiv(div_u * div_u, div_u * div_u + (cz * cz + (cx * cx + cy * cy)) + 1.0e-30);
}

// Pressure, entropy (cell_entropy SC:308-314) and, for the shear schemes,
// the Ducros factor of every cell of the padded grid, ghosts included (a
// ghost holds the values of its boundary-filled state).
template <int SHEAR>
__global__ void __launch_bounds__(256) k3_cells(Geo3 G, Eos e, State3 U, double* __restrict__ P,
                                                double* __restrict__ S, double* __restrict__ D) {
  const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
  if (k >= G.nz + 2 * G.g || j >= G.ny + 2 * G.g) return;
  double q[7];
  load3(G, U, i, j, k, q);
  double uu, vv, ww;
  const double p = pressure3(e, q, uu, vv, ww);
  const size_t c = at3(G, i, j, k);
  P[c] = p;
  S[c] = entropy(e, q[0], q[1], p, q[6]);
  if (SHEAR != 0) D[c] = ducros3(G, U, i, j, k);
}

// ------------------------------------------------------------ faces -------
// Face-constant part of the eigenvectors (SC:462-478) for sweep DIR.  n is
// the normal velocity; t1, t2 the two shear waves' tangential velocities as
// the projections use them (t1 = -u on the y- and z-sweeps, the 2D y-sweep's
// convention); vel the three face-averaged velocities.
struct Basis3 {
  double c, k, inv2c, inv2c2, y1c2, y2c2, Y1, Y2, c2, n, t1, t2;
  double vel[3];
};
template <int DIR>
__device__ __forceinline__ Basis3 make_basis3(double c, double rho, double ar1, double ar2,
                                              const double* vel) {
  Basis3 b;
  b.c = c;
  b.inv2c = ddiv(0.5, c);
  const double c2 = c * c;
  b.inv2c2 = ddiv(0.5, c2);
  b.Y1 = ddiv(ar1, rho);
  b.Y2 = ddiv(ar2, rho);
  b.y1c2 = ddiv(b.Y1, c2);
  b.y2c2 = ddiv(b.Y2, c2);
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
template <int DIR>
__device__ __forceinline__ void to_char3(const Basis3& b, const double* u, double* w) {
  const double un_m = u[2 + DIR];
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
template <int DIR>
__device__ __forceinline__ void to_phys3(const Basis3& b, const double* w, double* s) {
  const double n = b.n, c = b.c;
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
// (ar1, ar2, ru, rv, rw, p, a1), as face_state_pp.
__device__ __forceinline__ void face_state_pp3(const Eos& e, double* hi, const double* base) {
  double b[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) b[m] = base[m];
  b[6] = clamp01_pp(b[6]);
  if (b[0] < PP_RHO || isnan_(b[0])) b[0] = PP_RHO;
  if (b[1] < PP_RHO || isnan_(b[1])) b[1] = PP_RHO;
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

// HLLC flux + contact velocity (hllc_x/_src SC:594-687, hllc_y/_src
// SC:690-771) on primitives (ar1, ar2, u, v, w, p, a1).  Tangential momentum
// fluxes are rho * a * b with a, b the two velocities in u, v, w order.
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
  // physical momentum flux of component d on one side
  auto mflux = [&](const double* v, double rho, double p, int d) -> double {
    if (d == DIR) return rho * v[d] * v[d] + p;
    return d < DIR ? rho * v[d] * v[DIR] : rho * v[DIR] * v[d];
  };
  if (SL > 0.0) {
    f[0] = L[0] * nL;
    f[1] = L[1] * nL;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[2 + d] = mflux(vL, rhoL, pL, d);
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
      f[2 + d] = mflux(vL, rhoL, pL, d) + SL * star;
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
      f[2 + d] = mflux(vR, rhoR, pR, d) + SR * star;
    }
    f[5] = (ER + pR) * nR + SR * (ER_star * ratio - ER);
    f[6] = R[6] * nR + SR * (R[6] * ratio - R[6]);
  } else {
    f[0] = R[0] * nR;
    f[1] = R[1] * nR;
#pragma unroll
    for (int d = 0; d < 3; ++d) f[2 + d] = mflux(vR, rhoR, pR, d);
    f[5] = (ER + pR) * nR;
    f[6] = R[6] * nR;
  }
  double left_part = nL + wmin(0.0, SL) * (ratL - 1.0);
  double right_part = nR + wmax(0.0, SR) * (ratR - 1.0);
  double sp_sign = wsign(SP);
  ustar = 0.5 * (1.0 + sp_sign) * left_part + 0.5 * (1.0 - sp_sign) * right_part;
}

template <int DIR>
__device__ __forceinline__ long long stride3(const Geo3& G) {
  return DIR == 0 ? G.sx : DIR == 1 ? G.sy : 1;
}

// One face of sweep DIR between cells (i, j, k) and its +DIR neighbour:
// high-order fluxes + u* into out[0..7], first-order ones into out[8..15]
// (compute_fluxes_x/y SC:1030-1128 and fx/gy_low SC:1303-1394 with w).
template <int DIR, int BASE, int SHEAR, int FACEPP>
__device__ __forceinline__ void face3(const Geo3& G, const Eos& e, const State3& U,
                                      const double* __restrict__ P, const double* __restrict__ S,
                                      const double* __restrict__ D, int vf_always, double t1, int i,
                                      int j, int k, double* out) {
  const long long s = stride3<DIR>(G);
  const size_t c0 = at3(G, i, j, k);
  constexpr bool wide = BASE == 1 || SHEAR == 2;
  constexpr int klo = wide ? 0 : 1, khi = wide ? 6 : 5;
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
  // stencil cells -2..3 (boundary-filled) with their pressure
  double u[6][7];
#pragma unroll
  for (int q = klo; q < khi; ++q) {
    const int o = q - 2;
    double cq[7];
    load3(G, U, i + (DIR == 0 ? o : 0), j + (DIR == 1 ? o : 0), k + (DIR == 2 ? o : 0), cq);
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
    double ar1a = 0.5 * (A[0] + B[0]);
    double ar2a = 0.5 * (A[1] + B[1]);
    double a1a = 0.5 * (A[6] + B[6]);
    double pa = 0.5 * (A[5] + B[5]);
    double rhoa = ar1a + ar2a;
    double vel[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[d] = ddiv(0.5 * (A[2 + d] + B[2 + d]), rhoa);
    double ca = sound_speed(e, rhoa, pa, a1a);
    Basis3 Bs = make_basis3<DIR>(ca, rhoa, ar1a, ar2a, vel);
    double w[6][7];
#pragma unroll
    for (int q = klo; q < khi; ++q) to_char3<DIR>(Bs, u[q], w[q]);
    double wl[7], wr[7];
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
        double iph = SHEAR == 1 ? central4(w[1][m], w[2][m], w[3][m], w[4][m])
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
    double phL[7], phR[7], phfL[7], phfR[7];
    to_phys3<DIR>(Bs, wl, phL);
    to_phys3<DIR>(Bs, wr, phR);
    to_phys3<DIR>(Bs, w[2], phfL);
    to_phys3<DIR>(Bs, w[3], phfR);
    const bool cbad = ca < 0.0 || isnan_(ca);
    if (FACEPP == 0) {  // SC:1103-1116
      if (phL[0] < 0.0 || isnan_(phL[0])) phL[0] = phfL[0];
      if (phL[1] < 0.0 || isnan_(phL[1])) phL[1] = phfL[1];
      if (phL[6] < 0.0 || phL[6] > 1.0) phL[6] = phfL[6];
      if (isnan_(phL[5])) phL[5] = phfL[5];
      if (phR[0] < 0.0 || isnan_(phR[0])) phR[0] = phfR[0];
      if (phR[1] < 0.0 || isnan_(phR[1])) phR[1] = phfR[1];
      if (phR[6] < 0.0 || phR[6] > 1.0) phR[6] = phfR[6];
      if (isnan_(phR[5])) phR[5] = phfR[5];
      if (cbad) {
#pragma unroll
        for (int m = 0; m < 7; ++m) {
          phL[m] = phfL[m];
          phR[m] = phfR[m];
        }
      }
    } else {  // M10:1007-1025
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
    }
    double rhoL = phL[0] + phL[1], rhoR = phR[0] + phR[1];
    double PL[7] = {phL[0], phL[1], ddiv(phL[2], rhoL), ddiv(phL[3], rhoL), ddiv(phL[4], rhoL),
                    phL[5], phL[6]};
    double PR[7] = {phR[0], phR[1], ddiv(phR[2], rhoR), ddiv(phR[3], rhoR), ddiv(phR[4], rhoR),
                    phR[5], phR[6]};
    hllc3<DIR>(e, PL, PR, out, out[7]);
  }
  {  // first order (fx_low_kernel SC:1303-1347)
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
    double PL[7] = {A[0], A[1], ddiv(A[2], rhoL), ddiv(A[3], rhoL), ddiv(A[4], rhoL), pL, aL};
    double PR[7] = {B[0], B[1], ddiv(B[2], rhoR), ddiv(B[3], rhoR), ddiv(B[4], rhoR), pR, aR};
    hllc3<DIR>(e, PL, PR, out + 8, out[15]);
  }
}

// Faces of sweep DIR: (i, j, k) with the DIR index in [g-1, g+n) and the
// other two interior; one thread per face, k fastest.
template <int DIR, int BASE, int SHEAR, int FACEPP>
__global__ void __launch_bounds__(256) k3_faces(Geo3 G, Eos e, State3 U, const double* __restrict__ P,
                                                const double* __restrict__ S,
                                                const double* __restrict__ D, Faces3 F, int vf_always,
                                                double t1) {
  const int k = G.g - (DIR == 2 ? 1 : 0) + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.g - (DIR == 1 ? 1 : 0) + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.g - (DIR == 0 ? 1 : 0) + (int)blockIdx.z;
  if (k >= G.g + G.nz || j >= G.g + G.ny) return;
  double out[16];
  face3<DIR, BASE, SHEAR, FACEPP>(G, e, U, P, S, D, vf_always, t1, i, j, k, out);
  const size_t c = at3(G, i, j, k);
#pragma unroll
  for (int m = 0; m < 16; ++m) F[DIR * 16 + m][c] = out[m];
}

// ---------------------------------------------------------- limiter -------
// Face theta of stage s between the cell and its neighbour (non-interior
// cells hold 1), pp_blend_x/y SC:1500-1504.
__device__ __forceinline__ double th3(const Geo3& G, const double* T, int i, int j, int k) {
  return interior3(G, i, j, k) ? T[at3(G, i, j, k)] : 1.0;
}
struct CellFaces3 {
  double t[6][3];  // x-, x+, y-, y+, z-, z+ for each applied stage
};
__device__ __forceinline__ void faces_of3(const Geo3& G, double* const* T, int nb, int i, int j,
                                          int k, CellFaces3& cf) {
  for (int s = 0; s < nb; ++s) {
    const double* t = T[s];
    const double tc = t[at3(G, i, j, k)];
    double th = 1.0;
    cf.t[0][s] = wmin(wmin(th, th3(G, t, i - 1, j, k)), tc);
    cf.t[1][s] = wmin(wmin(th, tc), th3(G, t, i + 1, j, k));
    cf.t[2][s] = wmin(wmin(th, th3(G, t, i, j - 1, k)), tc);
    cf.t[3][s] = wmin(wmin(th, tc), th3(G, t, i, j + 1, k));
    cf.t[4][s] = wmin(wmin(th, th3(G, t, i, j, k - 1)), tc);
    cf.t[5][s] = wmin(wmin(th, tc), th3(G, t, i, j, k + 1));
  }
}
__device__ __forceinline__ double blend3(double H, double L, const double* th, int nb) {
  double f = H;
  for (int s = 0; s < nb; ++s) f = L + th[s] * (f - L);
  return f;
}
// Residual of component m at cell c from blended (rh) and first-order (rl)
// face fluxes; m == 6 adds the alpha source with alpha = a (SC:1444-1460).
__device__ __forceinline__ void residual3(const Geo3& G, const Faces3& F, int m, size_t c,
                                          const CellFaces3& cf, int nb, const double* inv, double a,
                                          double& rh, double& rl) {
  const size_t cm[3] = {c - (size_t)G.sx, c - (size_t)G.sy, c - 1};
  double dh[3], dl[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double Lp = F[d * 16 + 8 + m][c], Lm = F[d * 16 + 8 + m][cm[d]];
    const double Hp = blend3(F[d * 16 + m][c], Lp, cf.t[2 * d + 1], nb);
    const double Hm = blend3(F[d * 16 + m][cm[d]], Lm, cf.t[2 * d], nb);
    dh[d] = Hp - Hm;
    dl[d] = Lp - Lm;
  }
  rh = (dh[0] * inv[0] + dh[1] * inv[1]) + dh[2] * inv[2];
  rl = (dl[0] * inv[0] + dl[1] * inv[1]) + dl[2] * inv[2];
  if (m == 6) {
    double sh[3], sl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double Lp = F[d * 16 + 15][c], Lm = F[d * 16 + 15][cm[d]];
      sh[d] = blend3(F[d * 16 + 7][c], Lp, cf.t[2 * d + 1], nb) -
              blend3(F[d * 16 + 7][cm[d]], Lm, cf.t[2 * d], nb);
      sl[d] = Lp - Lm;
    }
    rh = rh - a * sh[0] * inv[0] - a * sh[1] * inv[1] - a * sh[2] * inv[2];
    rl = rl - a * sl[0] * inv[0] - a * sl[1] * inv[1] - a * sl[2] * inv[2];
  }
}

// theta_lower SC:123-138 / theta_cons_stage SC:268-292 on 7 components.
template <int STAGE>
__device__ __forceinline__ double theta_stage3(const Eos& e, const double* l, const double* h) {
  double theta = 1.0;
  if (STAGE == 0) {
    theta = wmin(theta, theta_lower(l[0], h[0], PP_RHO));
    theta = wmin(theta, theta_lower(l[1], h[1], PP_RHO));
  } else if (STAGE == 1) {
    theta = wmin(theta, theta_lower(l[6], h[6], PP_ALPHA));
    theta = wmin(theta, theta_lower(1.0 - l[6], 1.0 - h[6], PP_ALPHA));
  } else {
    theta = theta_lower(rho_ct2_cons3(e, l), rho_ct2_cons3(e, h), PP_CTILDE);
  }
  if (theta < 0.0 || isnan_(theta)) theta = 0.0;
  if (theta > 1.0) theta = 1.0;
  return theta;
}

struct Theta3 {
  double* t[3];
};

// One limiter stage's theta of every interior cell (pp_cell_theta
// SC:1436-1476), from the fluxes blended by the earlier stages' thetas.
template <int STAGE>
__global__ void __launch_bounds__(256) k3_theta(Geo3 G, Eos e, State3 U, Faces3 F, Theta3 T,
                                                const double* __restrict__ dt_dev) {
  const int k = G.g + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.g + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.g + (int)blockIdx.z;
  if (k >= G.g + G.nz || j >= G.g + G.ny) return;
  const double dt = *dt_dev;
  const double inv[3] = {1.0 / G.dx, 1.0 / G.dy, 1.0 / G.dz};
  const size_t c = at3(G, i, j, k);
  CellFaces3 cf;
  faces_of3(G, T.t, STAGE, i, j, k, cf);
  const double al = U.f[6][c];
  double l[7], h[7];
  if (STAGE == 0) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      double rh, rl;
      residual3(G, F, m, c, cf, 0, inv, al, rh, rl);
      const double u = U.f[m][c];
      l[m] = u - dt * rl;
      h[m] = u - dt * rh;
    }
  } else if (STAGE == 1) {
    double rh, rl;
    residual3(G, F, 6, c, cf, 1, inv, al, rh, rl);
    l[6] = al - dt * rl;
    h[6] = al - dt * rh;
  } else {
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      double rh, rl;
      residual3(G, F, m, c, cf, 2, inv, al, rh, rl);
      const double u = U.f[m][c];
      l[m] = u - dt * rl;
      h[m] = u - dt * rh;
    }
  }
  T.t[STAGE][c] = theta_stage3<STAGE>(e, l, h);
}

// enforce_admissible_kernel SC:918-958 with w.
__device__ __forceinline__ void repair_cell3(const Eos& e, double* u) {
  double a1 = clamp01_pp(u[6]);
  double m1 = u[0], m2 = u[1], mx = u[2], my = u[3], mz = u[4], E = u[5];
  if (isnan_(m1) || m1 < PP_RHO) m1 = PP_RHO;
  if (isnan_(m2) || m2 < PP_RHO) m2 = PP_RHO;
  if (isnan_(mx)) mx = 0.0;
  if (isnan_(my)) my = 0.0;
  if (isnan_(mz)) mz = 0.0;
  double rho = m1 + m2;
  double uu = ddiv(mx, rho);
  double vv = ddiv(my, rho);
  double ww = ddiv(mz, rho);
  double kin = 0.5 * rho * ((uu * uu + vv * vv) + ww * ww);
  double gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  double p = ddiv(E - pb - kin, gb);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  double p_min = PP_CTILDE * (g - 1.0) - pi;
  const double q[7] = {m1, m2, mx, my, mz, E, a1};
  if (isnan_(E) || isnan_(p) || rho_ct2_cons3(e, q) < PP_CTILDE) E = gb * p_min + pb + kin;
  u[0] = m1;
  u[1] = m2;
  u[2] = mx;
  u[3] = my;
  u[4] = mz;
  u[5] = E;
  u[6] = a1;
}
__device__ __forceinline__ double clamp_al3(double a1) {  // SC:852-853
  if (a1 < AL_CLIP) a1 = AL_CLIP;
  if (a1 > 1.0 - AL_CLIP) a1 = 1.0 - AL_CLIP;
  return a1;
}
// CFL time step of one cell: the 2D sx sy / (sx + sy) of timestep_kernel
// (SC:1700-1704) generalised to 1 / (1/sx + 1/sy + 1/sz).
__device__ __forceinline__ double cell_dt3(const Eos& e, const double* u, const Geo3& G) {
  double uu, vv, ww;
  double p = pressure3(e, u, uu, vv, ww);
  double rho = u[0] + u[1];
  double c = sound_speed(e, rho, p, u[6]);
  double sx = ddiv(G.dx, fabs(uu) + c);
  double sy = ddiv(G.dy, fabs(vv) + c);
  double sz = ddiv(G.dz, fabs(ww) + c);
  return ddiv(sx * sy * sz, (sx * sy + sy * sz) + sx * sz);
}
__device__ __forceinline__ void block_min3(double v, double* acc) {
  __shared__ double wmins[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = wmin(v, __shfl_xor(v, off, 64));
  const int t = threadIdx.y * 64 + threadIdx.x;
  if ((t & 63) == 0) wmins[t >> 6] = v;
  __syncthreads();
  if (t == 0) {
    double m = wmin(wmin(wmins[0], wmins[1]), wmin(wmins[2], wmins[3]));
    atomicMin((unsigned long long*)acc, (unsigned long long)__double_as_longlong(m));
  }
}

// Final fluxes, SSP-RK3 combine (rk3_kernel SC:1592-1611), repair, clamp;
// with dt_acc the CFL minimum of the new state is folded in.
__global__ void __launch_bounds__(256) k3_update(Geo3 G, Eos e, State3 Un, State3 U, Faces3 F,
                                                 Theta3 T, State3 Uo, double alpha, double beta,
                                                 const double* __restrict__ dt_dev, double* dt_acc,
                                                 int repair) {
  const int k = G.g + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.g + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.g + (int)blockIdx.z;
  double local_dt = 1.0e30;
  if (k < G.g + G.nz && j < G.g + G.ny) {
    const double dt = *dt_dev;
    const double inv[3] = {1.0 / G.dx, 1.0 / G.dy, 1.0 / G.dz};
    const size_t c = at3(G, i, j, k);
    CellFaces3 cf;
    faces_of3(G, T.t, 3, i, j, k, cf);
    const double ap = U.f[6][c];
    double un[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      double rh, rl;
      residual3(G, F, m, c, cf, 3, inv, ap, rh, rl);
      un[m] = alpha * Un.f[m][c] + beta * (U.f[m][c] - dt * rh);
    }
    if (repair) repair_cell3(e, un);
    un[6] = clamp_al3(un[6]);
#pragma unroll
    for (int m = 0; m < 7; ++m) Uo.f[m][c] = un[m];
    if (dt_acc) local_dt = cell_dt3(e, un, G);
  }
  if (dt_acc) block_min3(local_dt, dt_acc);
}

// CFL minimum over the interior (timestep_kernel SC:1686-1705).
__global__ void __launch_bounds__(256) k3_dt(Geo3 G, Eos e, State3 U, double* dt_acc) {
  const int k = G.g + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.g + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.g + (int)blockIdx.z;
  double v = 1.0e30;
  if (k < G.g + G.nz && j < G.g + G.ny) {
    const size_t c = at3(G, i, j, k);
    double u[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) u[m] = U.f[m][c];
    v = cell_dt3(e, u, G);
  }
  block_min3(v, dt_acc);
}

// Interior repair + clamp (the initial sequence SC:1998-2001).
__global__ void __launch_bounds__(256) k3_prepare(Geo3 G, Eos e, State3 U, int repair) {
  const int k = G.g + (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.g + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.g + (int)blockIdx.z;
  if (k >= G.g + G.nz || j >= G.g + G.ny) return;
  const size_t c = at3(G, i, j, k);
  double u[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) u[m] = U.f[m][c];
  if (repair) repair_cell3(e, u);
  u[6] = clamp_al3(u[6]);
#pragma unroll
  for (int m = 0; m < 7; ++m) U.f[m][c] = u[m];
}

static dim3 grid3(int nk, int nj, int ni) {
  return dim3((unsigned)((nk + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)ni);
}

template <int BASE, int SHEAR, int FACEPP>
static hipError_t launch_faces3_t(const Geo3& G, const Eos& e, const State3& U, const double* P,
                                  const double* S, const double* D, const Faces3& F, int vf_always,
                                  double t1, hipStream_t st) {
  const dim3 b(64, 4);
  hipLaunchKernelGGL((k3_faces<0, BASE, SHEAR, FACEPP>), grid3(G.nz, G.ny, G.nx + 1), b, 0, st, G,
                     e, U, P, S, D, F, vf_always, t1);
  hipLaunchKernelGGL((k3_faces<1, BASE, SHEAR, FACEPP>), grid3(G.nz, G.ny + 1, G.nx), b, 0, st, G,
                     e, U, P, S, D, F, vf_always, t1);
  hipLaunchKernelGGL((k3_faces<2, BASE, SHEAR, FACEPP>), grid3(G.nz + 1, G.ny, G.nx), b, 0, st, G,
                     e, U, P, S, D, F, vf_always, t1);
  return hipGetLastError();
}

}  // namespace wamp

extern "C" {
using namespace wamp;

// Schemes with a 3D instantiation: the reference's three (SC, M10, MP5).
int wamp3_scheme_available(int base, int shear, int face_pp) {
  return (base == 0 && shear == 1 && face_pp == 0) || (base == 0 && shear == 0 && face_pp == 1) ||
         (base == 1 && shear == 2 && face_pp == 0);
}

hipError_t wamp3_launch_cells(const Geo3* G, const Eos* e, const State3* U, double* P, double* S,
                              double* D, int shear, hipStream_t st) {
  const dim3 b(64, 4);
  const dim3 g = grid3(G->nz + 2 * G->g, G->ny + 2 * G->g, G->nx + 2 * G->g);
  if (shear)
    hipLaunchKernelGGL(k3_cells<1>, g, b, 0, st, *G, *e, *U, P, S, D);
  else
    hipLaunchKernelGGL(k3_cells<0>, g, b, 0, st, *G, *e, *U, P, S, D);
  return hipGetLastError();
}

hipError_t wamp3_launch_faces(const Geo3* G, const Eos* e, const State3* U, const double* P,
                              const double* S, const double* D, const Faces3* F, int base, int shear,
                              int vf_always, int face_pp, double t1, hipStream_t st) {
  if (base == 0 && shear == 1 && face_pp == 0)
    return launch_faces3_t<0, 1, 0>(*G, *e, *U, P, S, D, *F, vf_always, t1, st);
  if (base == 0 && shear == 0 && face_pp == 1)
    return launch_faces3_t<0, 0, 1>(*G, *e, *U, P, S, D, *F, vf_always, t1, st);
  if (base == 1 && shear == 2 && face_pp == 0)
    return launch_faces3_t<1, 2, 0>(*G, *e, *U, P, S, D, *F, vf_always, t1, st);
  return hipErrorInvalidValue;
}

hipError_t wamp3_launch_limit_update(const Geo3* G, const Eos* e, const State3* Un, const State3* U,
                                     const Faces3* F, double* const th[3], const State3* Uo,
                                     double alpha, double beta, const double* dt_dev,
                                     double* dt_acc, int repair, hipStream_t st) {
  const dim3 b(64, 4);
  const dim3 g = grid3(G->nz, G->ny, G->nx);
  Theta3 T{{th[0], th[1], th[2]}};
  hipLaunchKernelGGL(k3_theta<0>, g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
  hipLaunchKernelGGL(k3_theta<1>, g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
  hipLaunchKernelGGL(k3_theta<2>, g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
  hipLaunchKernelGGL(k3_update, g, b, 0, st, *G, *e, *Un, *U, *F, T, *Uo, alpha, beta, dt_dev,
                     dt_acc, repair);
  return hipGetLastError();
}

hipError_t wamp3_launch_dt(const Geo3* G, const Eos* e, const State3* U, double* dt_acc,
                           hipStream_t st) {
  hipLaunchKernelGGL(k3_dt, grid3(G->nz, G->ny, G->nx), dim3(64, 4), 0, st, *G, *e, *U, dt_acc);
  return hipGetLastError();
}

hipError_t wamp3_launch_prepare(const Geo3* G, const Eos* e, const State3* U, int repair,
                                hipStream_t st) {
  hipLaunchKernelGGL(k3_prepare, grid3(G->nz, G->ny, G->nx), dim3(64, 4), 0, st, *G, *e, *U,
                     repair);
  return hipGetLastError();
}

}  // extern "C"
