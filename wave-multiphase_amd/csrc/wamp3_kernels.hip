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
// and first-order fluxes of one sweep; source in wamp3_faces.h, FP32-mixed
// instantiation in wamp3_faces_mixed.hip), k3_theta<S> (one limiter theta
// per cell and stage), k3_update (blended fluxes, RK combine, repair, clamp,
// dt).  FP64, FMA contraction off, built with wamp_kernels.hip's flags.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wamp3_faces.h"
#include "wamp_device.h"
#include "wamp_types.h"

namespace wamp {

// rho_ctilde2_cons SC:242-258 with w.
__device__ __forceinline__ double rho_ct2_cons3(const Eos& e, const double* q) {
  double a1 = clamp01_pp(q[6]);
  double m1 = q[0], m2 = q[1];
  m1 = floor_pp(m1, PP_RHO);
  m2 = floor_pp(m2, PP_RHO);
  double rho = m1 + m2;
  double uu = ddiv(q[2], rho);
  double vv = ddiv(q[3], rho);
  double ww = ddiv(q[4], rho);
  double kin = 0.5 * rho * ((uu * uu + vv * vv) + ww * ww);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  return q[5] - kin - pi;
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
// (bx, by, bz): the block's coordinates in the grid3 block grid of the padded grid.
template <int SHEAR>
__device__ __forceinline__ void cells3_block(const Geo3& G, const Eos& e, const State3& U,
                                             double* __restrict__ P, double* __restrict__ S,
                                             double* __restrict__ D, int bx, int by, int bz) {
  const int k = bx * 64 + (int)threadIdx.x, j = by * 4 + (int)threadIdx.y, i = bz;
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
template <int SHEAR>
__global__ void __launch_bounds__(256) k3_cells(Geo3 G, Eos e, State3 U, double* __restrict__ P,
                                                double* __restrict__ S, double* __restrict__ D) {
  cells3_block<SHEAR>(G, e, U, P, S, D, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
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
// Face plane m at cell offset c in either storage of Faces3 (F32: mixed).
template <bool F32>
__device__ __forceinline__ double fload3(const Faces3& F, int m, size_t c) {
  if constexpr (F32) {
    if (face3_is_energy(m)) return F.b[(long long)face3_slot64(m) * F.s + (long long)c];
    return (double)F.h[(long long)face3_slot32(m) * F.hs + (long long)c];
  } else {
    return F.b[(long long)m * F.s + (long long)c];
  }
}
// Residual of component m at cell c from blended (rh) and first-order (rl)
// face fluxes; m == 6 adds the alpha source with alpha = a (SC:1444-1460).
template <bool F32>
__device__ __forceinline__ void residual3(const Geo3& G, const Faces3& F, int m, size_t c,
                                          const CellFaces3& cf, int nb, const double* inv, double a,
                                          double& rh, double& rl) {
  const size_t cm[3] = {c - (size_t)G.sx, c - (size_t)G.sy, c - 1};
  double dh[3], dl[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double Lp = fload3<F32>(F, d * 16 + 8 + m, c), Lm = fload3<F32>(F, d * 16 + 8 + m, cm[d]);
    const double Hp = blend3(fload3<F32>(F, d * 16 + m, c), Lp, cf.t[2 * d + 1], nb);
    const double Hm = blend3(fload3<F32>(F, d * 16 + m, cm[d]), Lm, cf.t[2 * d], nb);
    dh[d] = Hp - Hm;
    dl[d] = Lp - Lm;
  }
  rh = (dh[0] * inv[0] + dh[1] * inv[1]) + dh[2] * inv[2];
  rl = (dl[0] * inv[0] + dl[1] * inv[1]) + dl[2] * inv[2];
  if (m == 6) {
    double sh[3], sl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double Lp = fload3<F32>(F, d * 16 + 15, c), Lm = fload3<F32>(F, d * 16 + 15, cm[d]);
      sh[d] = blend3(fload3<F32>(F, d * 16 + 7, c), Lp, cf.t[2 * d + 1], nb) -
              blend3(fload3<F32>(F, d * 16 + 7, cm[d]), Lm, cf.t[2 * d], nb);
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
template <int STAGE, bool F32>
__device__ __forceinline__ void theta3_block(const Geo3& G, const Eos& e, const State3& U,
                                             const Faces3& F, const Theta3& T, double dt, int bx,
                                             int by, int bz) {
  const int k = G.g + bx * 64 + (int)threadIdx.x;
  const int j = G.g + by * 4 + (int)threadIdx.y;
  const int i = G.g + bz;
  if (k >= G.g + G.nz || j >= G.g + G.ny) return;
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
      residual3<F32>(G, F, m, c, cf, 0, inv, al, rh, rl);
      const double u = U.f[m][c];
      l[m] = u - dt * rl;
      h[m] = u - dt * rh;
    }
  } else if (STAGE == 1) {
    double rh, rl;
    residual3<F32>(G, F, 6, c, cf, 1, inv, al, rh, rl);
    l[6] = al - dt * rl;
    h[6] = al - dt * rh;
  } else {
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      double rh, rl;
      residual3<F32>(G, F, m, c, cf, 2, inv, al, rh, rl);
      const double u = U.f[m][c];
      l[m] = u - dt * rl;
      h[m] = u - dt * rh;
    }
  }
  T.t[STAGE][c] = theta_stage3<STAGE>(e, l, h);
}
template <int STAGE, bool F32>
__global__ void __launch_bounds__(256) k3_theta(Geo3 G, Eos e, State3 U, Faces3 F, Theta3 T,
                                                const double* __restrict__ dt_dev) {
  theta3_block<STAGE, F32>(G, e, U, F, T, *dt_dev, (int)blockIdx.x, (int)blockIdx.y,
                           (int)blockIdx.z);
}

// enforce_admissible_kernel SC:918-958 with w.
__device__ __forceinline__ void repair_cell3(const Eos& e, double* u) {
  double a1 = clamp01_pp(u[6]);
  double m1 = u[0], m2 = u[1], mx = u[2], my = u[3], mz = u[4], E = u[5];
  m1 = floor_pp(m1, PP_RHO);
  m2 = floor_pp(m2, PP_RHO);
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
template <bool F32>
__device__ __forceinline__ void update3_block(const Geo3& G, const Eos& e, const State3& Un,
                                              const State3& U, const Faces3& F, const Theta3& T,
                                              const State3& Uo, double alpha, double beta,
                                              double dt, double* dt_acc, int repair, int bx,
                                              int by, int bz) {
  const int k = G.g + bx * 64 + (int)threadIdx.x;
  const int j = G.g + by * 4 + (int)threadIdx.y;
  const int i = G.g + bz;
  double local_dt = 1.0e30;
  if (k < G.g + G.nz && j < G.g + G.ny) {
    const double inv[3] = {1.0 / G.dx, 1.0 / G.dy, 1.0 / G.dz};
    const size_t c = at3(G, i, j, k);
    CellFaces3 cf;
    faces_of3(G, T.t, 3, i, j, k, cf);
    const double ap = U.f[6][c];
    double un[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      double rh, rl;
      residual3<F32>(G, F, m, c, cf, 3, inv, ap, rh, rl);
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
template <bool F32>
__global__ void __launch_bounds__(256) k3_update(Geo3 G, Eos e, State3 Un, State3 U, Faces3 F,
                                                 Theta3 T, State3 Uo, double alpha, double beta,
                                                 const double* __restrict__ dt_dev, double* dt_acc,
                                                 int repair) {
  update3_block<F32>(G, e, Un, U, F, T, Uo, alpha, beta, *dt_dev, dt_acc, repair, (int)blockIdx.x,
                     (int)blockIdx.y, (int)blockIdx.z);
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

// ---------------------------------------------------------- batches -------
// One launch per kernel and stage over every active case of a table
// (batch_block3, wamp3_faces.h): the single-case kernels' bodies on each
// case's own grid, planes, EOS and dt.  Stage s reads U^s (u[s]) and writes
// u[s + 1] (stage 2: u[0]); the RK coefficients are the step's (SC:2035-2053).
__device__ __forceinline__ void stage_states3(const Case3Desc& c, int stage, State3& Un,
                                              State3& U, State3& Uo) {
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    Un.f[m] = c.u[0][m];
    U.f[m] = c.u[stage][m];
    Uo.f[m] = c.u[stage == 2 ? 0 : stage + 1][m];
  }
}
template <int SHEAR>
__global__ void __launch_bounds__(256) k3b_cells(const Case3Desc* __restrict__ cs,
                                                 const int* __restrict__ off, int ncase,
                                                 BatchScal sc, int stage) {
  int k, bx, by, bz;
  if (!batch_block3(cs, off, ncase, sc.done, B3_CELLS, k, bx, by, bz)) return;
  const Case3Desc& c = cs[k];
  State3 Un, U, Uo;
  stage_states3(c, stage, Un, U, Uo);
  cells3_block<SHEAR>(c.geo, c.eos, U, c.P, c.S, c.D, bx, by, bz);
}
template <int STAGE, bool F32>
__global__ void __launch_bounds__(256) k3b_theta(const Case3Desc* __restrict__ cs,
                                                 const int* __restrict__ off, int ncase,
                                                 BatchScal sc, int stage) {
  int k, bx, by, bz;
  if (!batch_block3(cs, off, ncase, sc.done, B3_INT, k, bx, by, bz)) return;
  const Case3Desc& c = cs[k];
  State3 Un, U, Uo;
  stage_states3(c, stage, Un, U, Uo);
  const Theta3 T{{c.th[0], c.th[1], c.th[2]}};
  theta3_block<STAGE, F32>(c.geo, c.eos, U, c.f, T, sc.dt[k], bx, by, bz);
}
template <bool F32>
__global__ void __launch_bounds__(256) k3b_update(const Case3Desc* __restrict__ cs,
                                                  const int* __restrict__ off, int ncase,
                                                  BatchScal sc, int stage, int repair) {
  int k, bx, by, bz;
  if (!batch_block3(cs, off, ncase, sc.done, B3_INT, k, bx, by, bz)) return;
  const Case3Desc& c = cs[k];
  State3 Un, U, Uo;
  stage_states3(c, stage, Un, U, Uo);
  const Theta3 T{{c.th[0], c.th[1], c.th[2]}};
  const double alpha = stage == 0 ? 0.0 : stage == 1 ? 0.75 : 1.0 / 3.0;
  const double beta = stage == 0 ? 1.0 : stage == 1 ? 0.25 : 2.0 / 3.0;
  update3_block<F32>(c.geo, c.eos, Un, U, c.f, T, Uo, alpha, beta, sc.dt[k],
                     stage == 2 ? sc.acc + k : nullptr, repair, bx, by, bz);
}
// After a step: t += dt; next dt = cfl * min, clipped to the end time
// (SC:2025-2026); a case that reached its end time is frozen.
__global__ void k3b_finalize(const Case3Desc* __restrict__ cs, int ncase, BatchScal sc) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncase || sc.done[k]) return;
  const double t = sc.t[k] + sc.dt[k];
  sc.t[k] = t;
  double dt = cs[k].cfl * sc.acc[k];
  sc.acc[k] = 1.0e30;
  const double te = sc.tend[k];
  if (te > 0.0) {
    if (t >= te - 1.0e-14) {
      sc.done[k] = 1;
      dt = 0.0;
    } else if (t + dt > te) {
      dt = te - t;
    }
  }
  sc.dt[k] = dt;
}
// Initial dt of slot k from its CFL minimum (clipped to the end time).
__global__ void k3b_dt0(double* acc, BatchScal sc, int k, double cfl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double dt = cfl * *acc;
    *acc = 1.0e30;
    const double te = sc.tend[k];
    if (te > 0.0 && sc.t[k] + dt > te) dt = te - sc.t[k];
    sc.dt[k] = dt;
    sc.done[k] = (te > 0.0 && sc.t[k] >= te - 1.0e-14) ? 1 : 0;
  }
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
  return launch_faces3<double>(*G, *e, *U, P, S, D, *F, base, shear, vf_always, face_pp, t1, st);
}

hipError_t wamp3_launch_limit_update(const Geo3* G, const Eos* e, const State3* Un, const State3* U,
                                     const Faces3* F, double* const th[3], const State3* Uo,
                                     double alpha, double beta, const double* dt_dev,
                                     double* dt_acc, int repair, hipStream_t st) {
  const dim3 b(64, 4);
  const dim3 g = grid3(G->nz, G->ny, G->nx);
  Theta3 T{{th[0], th[1], th[2]}};
  if (F->h) {
    hipLaunchKernelGGL((k3_theta<0, true>), g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
    hipLaunchKernelGGL((k3_theta<1, true>), g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
    hipLaunchKernelGGL((k3_theta<2, true>), g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
    hipLaunchKernelGGL(k3_update<true>, g, b, 0, st, *G, *e, *Un, *U, *F, T, *Uo, alpha, beta,
                       dt_dev, dt_acc, repair);
  } else {
    hipLaunchKernelGGL((k3_theta<0, false>), g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
    hipLaunchKernelGGL((k3_theta<1, false>), g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
    hipLaunchKernelGGL((k3_theta<2, false>), g, b, 0, st, *G, *e, *U, *F, T, dt_dev);
    hipLaunchKernelGGL(k3_update<false>, g, b, 0, st, *G, *e, *Un, *U, *F, T, *Uo, alpha, beta,
                       dt_dev, dt_acc, repair);
  }
  return hipGetLastError();
}

hipError_t wamp3_launch_dt(const Geo3* G, const Eos* e, const State3* U, double* dt_acc,
                           hipStream_t st) {
  hipLaunchKernelGGL(k3_dt, grid3(G->nz, G->ny, G->nx), dim3(64, 4), 0, st, *G, *e, *U, dt_acc);
  return hipGetLastError();
}

hipError_t wamp3_launch_faces_batch_mixed(const Case3Desc* cs, const int* off, int ncase,
                                         const BatchScal* sc, int stage, int base, int shear,
                                         int vf_always, int face_pp, double t1, const int* blocks,
                                         hipStream_t st);

// One RK stage of a 3D batch: cells, the three sweeps, the three theta
// stages and the update, each one launch over the table.  blocks[kind]: the
// table's total block count of each work kind.
hipError_t wamp3_launch_batch_stage(const Case3Desc* cs, const int* off, int ncase,
                                    const BatchScal* sc, int stage, int base, int shear,
                                    int vf_always, int face_pp, int repair, double t1,
                                    int precision, const int* blocks, hipStream_t st) {
  const dim3 b(64, 4);
  if (blocks[B3_CELLS] > 0) {
    if (shear)
      hipLaunchKernelGGL(k3b_cells<1>, dim3(blocks[B3_CELLS]), b, 0, st, cs, off, ncase, *sc, stage);
    else
      hipLaunchKernelGGL(k3b_cells<0>, dim3(blocks[B3_CELLS]), b, 0, st, cs, off, ncase, *sc, stage);
  }
  hipError_t e = precision == 0
                     ? launch_faces3_batch<double>(cs, off, ncase, *sc, stage, base, shear,
                                                   vf_always, face_pp, t1, blocks + B3_FACE0, st)
                     : wamp3_launch_faces_batch_mixed(cs, off, ncase, sc, stage, base, shear,
                                                      vf_always, face_pp, t1, blocks + B3_FACE0, st);
  if (e != hipSuccess) return e;
  const int nb = blocks[B3_INT];
  if (nb <= 0) return hipGetLastError();
  if (precision) {
    hipLaunchKernelGGL((k3b_theta<0, true>), dim3(nb), b, 0, st, cs, off, ncase, *sc, stage);
    hipLaunchKernelGGL((k3b_theta<1, true>), dim3(nb), b, 0, st, cs, off, ncase, *sc, stage);
    hipLaunchKernelGGL((k3b_theta<2, true>), dim3(nb), b, 0, st, cs, off, ncase, *sc, stage);
    hipLaunchKernelGGL(k3b_update<true>, dim3(nb), b, 0, st, cs, off, ncase, *sc, stage, repair);
  } else {
    hipLaunchKernelGGL((k3b_theta<0, false>), dim3(nb), b, 0, st, cs, off, ncase, *sc, stage);
    hipLaunchKernelGGL((k3b_theta<1, false>), dim3(nb), b, 0, st, cs, off, ncase, *sc, stage);
    hipLaunchKernelGGL((k3b_theta<2, false>), dim3(nb), b, 0, st, cs, off, ncase, *sc, stage);
    hipLaunchKernelGGL(k3b_update<false>, dim3(nb), b, 0, st, cs, off, ncase, *sc, stage, repair);
  }
  return hipGetLastError();
}

hipError_t wamp3_launch_batch_finalize(const Case3Desc* cs, int ncase, const BatchScal* sc,
                                       hipStream_t st) {
  hipLaunchKernelGGL(k3b_finalize, dim3((ncase + 63) / 64), dim3(64), 0, st, cs, ncase, *sc);
  return hipGetLastError();
}

// Initial sequence of slot k: repair and clamp, then its first dt.
hipError_t wamp3_launch_batch_prepare(const Geo3* G, const Eos* e, const State3* U, int repair,
                                      double* acc, const BatchScal* sc, int k, double cfl,
                                      hipStream_t st) {
  const dim3 g = grid3(G->nz, G->ny, G->nx);
  hipLaunchKernelGGL(k3_prepare, g, dim3(64, 4), 0, st, *G, *e, *U, repair);
  hipLaunchKernelGGL(k3_dt, g, dim3(64, 4), 0, st, *G, *e, *U, acc);
  hipLaunchKernelGGL(k3b_dt0, dim3(1), dim3(64), 0, st, acc, *sc, k, cfl);
  return hipGetLastError();
}

hipError_t wamp3_launch_prepare(const Geo3* G, const Eos* e, const State3* U, int repair,
                                hipStream_t st) {

  hipLaunchKernelGGL(k3_prepare, grid3(G->nz, G->ny, G->nx), dim3(64, 4), 0, st, *G, *e, *U,
                     repair);
  return hipGetLastError();
}

}  // extern "C"
