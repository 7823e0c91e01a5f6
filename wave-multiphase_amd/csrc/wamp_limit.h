// wamp_limit.h -- per-cell pieces of the limiter / update stage (Wong
// thetas, blended-flux residuals, SSP-RK3 combine inputs, repair, alpha
// clamp, CFL bound), shared by the limiter kernels (wamp_kernels.hip) and
// the host step of the CPU device (wamp_cpu.hip).  FP64 only.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "wamp_device.h"
#include "wamp_types.h"

namespace wamp {

// Blended flux after `nb` limiter stages: f <- L + theta_k (f - L), k < nb
// (pp_blend_x_flux_kernel SC:1506-1512 applied once per stage).
WAMP_FN double blend(double H, double L, const double* th, int nb) {
  double f = H;
  for (int k = 0; k < nb; ++k) f = L + th[k] * (f - L);
  return f;
}

struct CellFaces {
  // theta of the four faces for each applied stage: x-, x+, y-, y+
  double t[4][3];
};

// Face theta from the two adjacent cells' theta (non-interior cells hold 1),
// pp_blend_x/y SC:1500-1504, 1536-1540.
WAMP_FN double face_theta(double ta, double tb) {
  double th = 1.0;
  th = wmin(th, ta);
  th = wmin(th, tb);
  return th;
}

// Residual of component m (0..5) at cell c from blended (rh) and low (rl)
// face fluxes; m == 5 adds the alpha source with alpha = a (SC:1444-1460).
template <bool F32>
WAMP_FN void residual(const Faces& F, int m, size_t c, size_t ld,
                                         const CellFaces& cf, int nb, double inv_dx,
                                         double inv_dy, double a, double& rh, double& rl) {
  const size_t cx = c - ld, cy = c - 1;
  double Lxp = fload<F32>(F, 14 + m, c), Lxm = fload<F32>(F, 14 + m, cx);
  double Lyp = fload<F32>(F, 21 + m, c), Lym = fload<F32>(F, 21 + m, cy);
  double Hxp = blend(fload<F32>(F, m, c), Lxp, cf.t[1], nb);
  double Hxm = blend(fload<F32>(F, m, cx), Lxm, cf.t[0], nb);
  double Hyp = blend(fload<F32>(F, 7 + m, c), Lyp, cf.t[3], nb);
  double Hym = blend(fload<F32>(F, 7 + m, cy), Lym, cf.t[2], nb);
  rh = (Hxp - Hxm) * inv_dx + (Hyp - Hym) * inv_dy;
  rl = (Lxp - Lxm) * inv_dx + (Lyp - Lym) * inv_dy;
  if (m == 5) {
    double sLxp = fload<F32>(F, 20, c), sLxm = fload<F32>(F, 20, cx);
    double sLyp = fload<F32>(F, 27, c), sLym = fload<F32>(F, 27, cy);
    double sHxp = blend(fload<F32>(F, 6, c), sLxp, cf.t[1], nb);
    double sHxm = blend(fload<F32>(F, 6, cx), sLxm, cf.t[0], nb);
    double sHyp = blend(fload<F32>(F, 13, c), sLyp, cf.t[3], nb);
    double sHym = blend(fload<F32>(F, 13, cy), sLym, cf.t[2], nb);
    rh = rh - a * (sHxp - sHxm) * inv_dx - a * (sHyp - sHym) * inv_dy;
    rl = rl - a * (sLxp - sLxm) * inv_dx - a * (sLyp - sLym) * inv_dy;
  }
}

// enforce_admissible_kernel SC:918-958 on one cell (in registers).
WAMP_FN void repair_cell(const Eos& e, double* u) {
  double a1 = clamp01_pp(u[5]);
  double m1 = u[0], m2 = u[1], mx = u[2], my = u[3], E = u[4];
  m1 = floor_pp(m1, PP_RHO);
  m2 = floor_pp(m2, PP_RHO);
  if (isnan_(mx)) mx = 0.0;
  if (isnan_(my)) my = 0.0;
  double rho = m1 + m2;  // >= 2e-10: both partial densities are floored
  double uu = ddiv(mx, rho);
  double vv = ddiv(my, rho);
  double kin = 0.5 * rho * (uu * uu + vv * vv);
  double gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  double p = ddiv(E - pb - kin, gb);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  double p_min = PP_CTILDE * (g - 1.0) - pi;
  if (isnan_(E) || isnan_(p) || rho_ct2_cons(e, m1, m2, mx, my, E, a1) < PP_CTILDE)
    E = gb * p_min + pb + kin;
  u[0] = m1;
  u[1] = m2;
  u[2] = mx;
  u[3] = my;
  u[4] = E;
  u[5] = a1;
}

WAMP_FN double clamp_al(double a1) {  // SC:852-853
  if (a1 < AL_CLIP) a1 = AL_CLIP;
  if (a1 > 1.0 - AL_CLIP) a1 = 1.0 - AL_CLIP;
  return a1;
}

// timestep_kernel SC:1700-1704 on one cell whose a1 is already clamped.
WAMP_FN double cell_dt(const Eos& e, const double* u, double dx, double dy) {
  double uu, vv;
  double p = pressure(e, u[0], u[1], u[2], u[3], u[4], u[5], uu, vv);
  double rho = u[0] + u[1];
  double c = sound_speed(e, rho, p, u[5]);
  double sx = ddiv(dx, fabs(uu) + c);
  double sy = ddiv(dy, fabs(vv) + c);
  return ddiv(sx * sy, sx + sy);
}

// Theta of limiter stage S at interior cell c (pp_cell_theta SC:1436-1476):
// the first-order and the blended high-order updates of the stage's
// variables -- S = 0 the partial densities, 1 alpha, 2 the full state for
// rho c~^2 -- with the face thetas of the stages before S in cf.
template <int S, bool F32>
WAMP_FN double cell_theta(const Eos& e, const State& U, const Faces& F, size_t c, size_t ld,
                          const CellFaces& cf, double inv_dx, double inv_dy, double dt) {
  const double al = U.f[5][c];
  double l[6], h[6];
  if (S == 0) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      double rh, rl;
      residual<F32>(F, m, c, ld, cf, 0, inv_dx, inv_dy, al, rh, rl);
      double u = U.f[m][c];
      l[m] = u - dt * rl;
      h[m] = u - dt * rh;
    }
  } else if (S == 1) {
    double rh, rl;
    residual<F32>(F, 5, c, ld, cf, 1, inv_dx, inv_dy, al, rh, rl);
    l[5] = al - dt * rl;
    h[5] = al - dt * rh;
  } else {
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      double rh, rl;
      residual<F32>(F, m, c, ld, cf, 2, inv_dx, inv_dy, al, rh, rl);
      double u = U.f[m][c];
      l[m] = u - dt * rl;
      h[m] = u - dt * rh;
    }
  }
  return theta_stage<S>(e, l, h);
}

// New state of interior cell c: the SSP-RK3 combine alpha Un + beta (U -
// dt R) with the fluxes blended by all three stages (rk3_kernel
// SC:1592-1611), the emergency repair (SC:918-958) and the alpha clamp.
template <bool F32>
WAMP_FN void cell_update(const Eos& e, const State& Un, const State& U, const Faces& F, size_t c,
                         size_t ld, const CellFaces& cf, double inv_dx, double inv_dy, double dt,
                         double alpha, double beta, int repair, double* un) {
  const double ap = U.f[5][c];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    double rh, rl;
    residual<F32>(F, m, c, ld, cf, 3, inv_dx, inv_dy, ap, rh, rl);
    un[m] = alpha * Un.f[m][c] + beta * (U.f[m][c] - dt * rh);
  }
  if (repair) repair_cell(e, un);
  un[5] = clamp_al(un[5]);
}

}  // namespace wamp
