// wamp_device.h -- device math of the five-equation stiffened-gas step (gfx950).
//
// Scalar building blocks of the fused face-flux and limiter kernels.  Every
// arithmetic expression keeps the reference's evaluation order (Python's
// left-to-right association) so the FP64 path reproduces it to round-off;
// the FP64 build compiles with contraction off.  Citations are to the
// reference's Warp functions: SC = sainath_SC_it_works.py,
// M10 = Sainath_M10.py.
//
// Design points that differ from the reference (exact, order-preserving):
//  * the HLLC flux and its contact velocity u* share one evaluation of the
//    wave speeds (hllc_x SC:594-662 and hllc_x_src SC:664-687 recompute the
//    same cL, cR, SL, SR, SP from the same arguments);
//  * face constants (u_n/2c, Y_k/c^2, ...) of the characteristic projection
//    are formed once per face instead of once per stencil cell;
//  * per-cell quantities (pressure, entropy, indicator, Ducros sensor) are
//    formed once per cell by the callers, not once per face.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "wamp_math.h"
#include "wamp_types.h"

// The scalar helpers are host-and-device functions: the FP64 step also runs
// on the host (wamp_cpu.hip, the CPU device) from this same source.  Device
// builds use the hardware reciprocal / square-root sequences below, host
// builds the IEEE operations those sequences reproduce bit for bit in their
// operand range (below it -- subnormal quotients -- the last bits may
// differ), so a host step gives the device step's bits.
#define WAMP_FN __host__ __device__ __forceinline__

// The FP64 kernels (wamp_kernels.hip) and the FP32-mixed face kernels
// (wamp_faces_mixed.hip, defining WAMP_MIXED_TU) are separate translation
// units whose device helpers differ in FP64 division; the mixed unit's
// helpers live in the inline namespace wamp::mixed so the two sets never
// share a symbol.
#ifdef WAMP_MIXED_TU
#define WAMP_DEV_BEGIN namespace wamp { inline namespace mixed {
#define WAMP_DEV_END } }
#else
#define WAMP_DEV_BEGIN namespace wamp {
#define WAMP_DEV_END }
#endif

WAMP_DEV_BEGIN

// Warp builtin semantics: comparison selects, sign(0) = +1.
template <typename T> WAMP_FN T wmin(T a, T b) { return a < b ? a : b; }
template <typename T> WAMP_FN T wmax(T a, T b) { return a > b ? a : b; }
template <typename T> WAMP_FN T wsign(T x) { return x < T(0) ? T(-1) : T(1); }
template <typename T> WAMP_FN bool isnan_(T x) { return __builtin_isnan(x); }
// FP32 (the mixed mode's face arithmetic only): the hardware v_min_f32 /
// v_max_f32, one instruction instead of a compare and a select.  For
// operands that are not NaN they return the same value as the comparison
// selects (a pair of zeros may come back with the other sign).
template <> WAMP_FN float wmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <> WAMP_FN float wmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }
// sign inside the minmod limiters: there every term is multiplied by the
// smallest magnitude of its arguments, which is zero whenever an argument is
// a zero of either sign, so copysign (one v_bfi) gives the same value as
// wsign for FP32.
template <typename T> WAMP_FN T msign(T x) { return wsign(x); }
template <> WAMP_FN float msign<float>(float x) { return __builtin_copysignf(1.0f, x); }

// Division, square root and tanh by arithmetic type.  The FP64 kernels keep
// the correctly rounded results (the parity path), through the shortened
// sequences below where the operands allow.  The FP32 face arithmetic of the
// mixed-precision mode uses the hardware approximations (v_rcp_f32,
// v_sqrt_f32 and an exp-based tanh, about 1-2 ulp): on CDNA4 non-packed FP32
// issues at the FP64 rate, so the mixed mode gains only by dropping the
// multi-instruction IEEE sequences.

// a / b from y = RN(1/b) with one correction step (Markstein 1990; Muller et
// al., Handbook of Floating-Point Arithmetic, the FMA-based division
// theorem): q = RN(a y), r = a - b q (exact with an FMA), RN(q + r y) is the
// correctly rounded quotient for finite a, b with no overflow or underflow,
// i.e. bitwise a / b.  Three instructions where y is known (a host constant
// or a reciprocal shared by several numerators).
WAMP_FN double quotient(double a, double y, double b) {
  const double q = a * y;
  const double r = fma(-b, q, a);
  return fma(r, y, q);
}

// 1/b from v_rcp_f64 and two Newton steps: the reciprocal steps of the
// compiler's IEEE divide sequence (v_div_scale, v_rcp, 4 x v_fma, v_mul,
// v_fma, v_div_fmas, v_div_fixup) without its operand scaling.
WAMP_FN double recip_nr(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  return fma(y, e, y);
#else
  return 1.0 / b;
#endif
}

// FP64 division.
//  * FP64 kernels: quotient(a, recip_nr(b), b) is the IEEE divide sequence
//    minus v_div_scale / v_div_fmas / v_div_fixup, which change nothing when
//    no operand needs scaling or special-casing, so it is a / b bit for bit
//    for finite b with 2^-1000 <= |b| <= 2^1000 and finite a with a = 0 or
//    |a| >= 2^-960 and |a / b| in [2^-1000, 2^700] (a zero quotient may come
//    back with the other sign).  For b = 0 or a non-finite b it gives NaN
//    where the divide gives +-inf, so it is used where the divisor cannot
//    vanish in an admissible state (densities, sound speeds, mixture gammas,
//    limiter denominators kept above TINY); the HLLC ratios over SL - SP and
//    SR - SP keep the IEEE divide.  Eight instructions instead of eleven,
//    and the compiler keeps one reciprocal per divisor, so each further
//    numerator of a shared divisor costs three.
//  * The mixed unit's remaining FP64 arithmetic (pressure recovery,
//    indicators, face positivity limiter, energy flux): one Newton step and
//    the correction, within one ulp of the quotient, six instructions.
WAMP_FN double ddiv(double a, double b) {
#if !defined(__HIP_DEVICE_COMPILE__)
  return a / b;  // host: the IEEE quotient
#elif defined(WAMP_MIXED_TU)
  double y = __builtin_amdgcn_rcp(b);
  const double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  const double q = a * y;
  const double r = fma(-b, q, a);
  return fma(r, y, q);
#else
  return quotient(a, recip_nr(b), b);
#endif
}
WAMP_FN double dv(double a, double b) { return ddiv(a, b); }
WAMP_FN float dv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a / b;  // FP32 arithmetic runs on the device only
#endif
}
// FP64 square root of the sound speed (argument clamped to >= TINY = 1e-20).
// FP64 kernels: the compiler's IEEE sqrt sequence (v_rsq_f64, two products,
// seven FMAs) without its scaling of arguments below 2^-767 and its select
// for zero and +inf, which act only outside this argument range: bitwise
// sqrt(x) for finite x >= 2^-767; NaN stays NaN (+inf would give NaN, not
// +inf, but the argument is finite in any admissible state).  The mixed unit
// keeps the library sqrt for its few FP64 square roots.
WAMP_FN double sq(double x) {
#if !defined(__HIP_DEVICE_COMPILE__) || defined(WAMP_MIXED_TU)
  return sqrt(x);
#else
  const double g = __builtin_amdgcn_rsq(x);
  double s = x * g;
  double h = g * 0.5;
  const double r = fma(-h, s, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  double d = fma(-s, s, x);
  s = fma(d, h, s);
  d = fma(-s, s, x);
  return fma(d, h, s);
#endif
}
WAMP_FN float sq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
// FP64 tanh (THINC) out of line: the face kernel has up to 12 THINC sites;
// one shared body keeps the FP64 face kernel's code within the 64 KB
// instruction cache of a CU pair.  It runs only where THINC fires.  The FP64
// kernels use wm_tanh (wamp_math.h: IEEE operations only, so host builds of
// the same source give the same bits); the mixed unit keeps the library tanh
// for its few FP64 THINC values.
#ifdef WAMP_MIXED_TU
inline __host__ __device__ __noinline__ double th(double x) { return tanh(x); }
#else
inline __host__ __device__ __noinline__ double th(double x) { return wm_tanh(x); }
#endif
WAMP_FN float th(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  float t = __expf(-2.0f * fabsf(x));
  float r = (1.0f - t) * __builtin_amdgcn_rcpf(1.0f + t);
  return x < 0.0f ? -r : r;
#else
  return tanhf(x);
#endif
}

// Face plane m at cell offset c in either storage of Faces.  The limiter
// kernels take the storage as a template argument (F32: mixed storage).
template <bool F32>
WAMP_FN double fload(const Faces& F, int m, size_t c) {
  if constexpr (F32) {
    if (face_is_energy(m)) return F.b[(long long)(m / 7) * F.s + (long long)c];
    return (double)F.h[(long long)face_slot32(m) * F.hs + (long long)c];
  } else {
    return F.b[(long long)m * F.s + (long long)c];
  }
}
// Store of the face kernels: the FP64 kernels always write FP64 storage; the
// mixed ones write whichever storage F describes (the split entry point
// wamp_compute_fluxes hands them 28 FP64 planes).
template <typename T>
WAMP_FN void fstore(const Faces& F, int m, size_t c, double v) {
  if constexpr (sizeof(T) == sizeof(float)) {
    if (F.h) {
      if (face_is_energy(m)) F.b[(long long)(m / 7) * F.s + (long long)c] = v;
      else F.h[(long long)face_slot32(m) * F.hs + (long long)c] = (float)v;
      return;
    }
  }
  F.b[(long long)m * F.s + (long long)c] = v;
}

// Reference constants, SC:44-70.
constexpr double TINY = 1.0e-20;
constexpr double EPSM = 1.0e-20;
constexpr double MP5C = 4.0;
constexpr double B2 = 4.0 / 3.0;
constexpr double KAP = 1.0 / 3.0;
constexpr double ONE_60 = 1.0 / 60.0;
constexpr double YUXIN = 0.9 * 0.35 * 1.0e-2 / (1.0 - 0.9 * 0.35);
constexpr double THINC_THRESH = 0.35;
constexpr double PP_RHO = 1.0e-10;
constexpr double PP_ALPHA = 1.0e-10;
constexpr double PP_CTILDE = 1.0e-8;
constexpr double PP_HS_RHO = 1.0e-11;
constexpr double PP_HS_ALPHA = 1.0e-11;
constexpr double PP_HS_CTILDE = 1.0e-9;
constexpr double SIGMA_SMOOTH = 0.01;
constexpr double AL_CLIP = 1.0e-8;  // cons_to_prim clamp, SC:852-853

// ---------------------------------------------------------------- EOS ------
// mix_gamma_pibig SC:83-88
template <typename T>
WAMP_FN void mix_gb_pb(const Eos& e, T a1, T& gb, T& pb) {
  T a2 = T(1) - a1;
  gb = dv(a1, T(e.g1) - T(1)) + dv(a2, T(e.g2) - T(1));
  pb = dv(a1 * T(e.g1) * T(e.p1), T(e.g1) - T(1)) + dv(a2 * T(e.g2) * T(e.p2), T(e.g2) - T(1));
}
// FP64: the four divisions by the constants g-1 use the host reciprocals
// (bitwise the same quotients; gamma > 1 is checked at context creation).
template <>
WAMP_FN void mix_gb_pb<double>(const Eos& e, double a1, double& gb, double& pb) {
  const double a2 = 1.0 - a1;
  const double d1 = e.g1 - 1.0, d2 = e.g2 - 1.0;
  gb = quotient(a1, e.r1, d1) + quotient(a2, e.r2, d2);
  pb = quotient(a1 * e.g1 * e.p1, e.r1, d1) + quotient(a2 * e.g2 * e.p2, e.r2, d2);
}
// mix_gamma_pinf SC:90-95
template <typename T>
WAMP_FN void mix_g_pi(const Eos& e, T a1, T& g, T& pi) {
  T gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  g = dv(T(1), gb) + T(1);
  pi = dv((g - T(1)) * pb, g);
}
// sound_speed SC:97-103
template <typename T>
WAMP_FN T sound_speed(const Eos& e, T rho, T p, T a1) {
  T g, pi;
  mix_g_pi(e, a1, g, pi);
  T val = dv(g * (p + pi), rho);
  if (val < T(TINY)) val = T(TINY);
  return sq(val);
}
// prim_to_energy SC:105-110
template <typename T>
WAMP_FN T prim_energy(const Eos& e, T ar1, T ar2, T u, T v, T p, T a1) {
  T rho = ar1 + ar2;
  T gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  return gb * p + pb + T(0.5) * rho * (u * u + v * v);
}
// clamp01_pp SC:112-121: NaN -> PP_ALPHA, then clamped to [PP_ALPHA,
// 1 - PP_ALPHA].  fmax returns its other operand for a NaN argument (IEEE
// maxNum), so fmin(fmax(a, lo), hi) is the reference's branch sequence value
// for value: one v_max and one v_min instead of compares, selects and
// constant moves.
WAMP_FN double clamp01_pp(double a) {
  return __builtin_fmin(__builtin_fmax(a, PP_ALPHA), 1.0 - PP_ALPHA);
}
WAMP_FN float clamp01_pp(float a) {
  return __builtin_fminf(__builtin_fmaxf(a, (float)PP_ALPHA), 1.0f - (float)PP_ALPHA);
}
// x < lo || isnan(x) ? lo : x (the partial-density floors SC:223-226,
// SC:249-252 and SC:933-936), as one fmax for the same reason.
WAMP_FN double floor_pp(double x, double lo) { return __builtin_fmax(x, lo); }
// pressure recovery from the conserved state, cons_to_prim_kernel SC:856-860
// (always FP64: p is the small difference of E and the p_inf term).
WAMP_FN double pressure(const Eos& e, double ar1, double ar2, double ru,
                                           double rv, double E, double a1, double& uu,
                                           double& vv) {
  double rho = ar1 + ar2;
  uu = dv(ru, rho);
  vv = dv(rv, rho);
  double gb, pb;
  mix_gb_pb(e, a1, gb, pb);
  return ddiv(E - pb - 0.5 * rho * (uu * uu + vv * vv), gb);
}

// Primitive state of one cell as the first-order flux sees it (fx_low_kernel
// SC:1303-1347): partial densities, velocities u = ru / rho, v = rv / rho,
// the pressure floored where rho c~^2 < PP_CTILDE (SC:1335-1338) and the
// clamped alpha.  The face kernels and the limiter march that forms the
// first-order fluxes itself both go through this function.
WAMP_FN void low_prims(const Eos& e, double ar1, double ar2, double uu, double vv, double p,
                       double al, double* P) {
  const double a = clamp01_pp(al);
  double g, pi;
  mix_g_pi(e, a, g, pi);
  if (isnan_(p) || ddiv(p + pi, g - 1.0) < PP_CTILDE) p = PP_CTILDE * (g - 1.0) - pi;
  P[0] = ar1;
  P[1] = ar2;
  P[2] = uu;
  P[3] = vv;
  P[4] = p;
  P[5] = a;
}

// ------------------------------------------------------- positivity --------
// pp_theta_lower SC:123-138
WAMP_FN double theta_lower(double base, double high, double floor_) {
  double theta = 1.0;
  if (isnan_(base) || isnan_(high)) {
    theta = 0.0;
  } else if (high < floor_) {
    double denom = base - high;
    theta = denom > TINY ? ddiv(base - floor_, denom) : 0.0;
  }
  if (theta < 0.0) theta = 0.0;
  if (theta > 1.0) theta = 1.0;
  return theta;
}
// rho_ctilde2_cons SC:242-258
WAMP_FN double rho_ct2_cons(const Eos& e, double m1, double m2, double mx,
                                               double my, double E, double a1_in) {
  double a1 = clamp01_pp(a1_in);
  m1 = floor_pp(m1, PP_RHO);
  m2 = floor_pp(m2, PP_RHO);
  double rho = m1 + m2;  // >= 2e-10: both partial densities are floored
  double uu = dv(mx, rho);
  double vv = dv(my, rho);
  double kin = 0.5 * rho * (uu * uu + vv * vv);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  return E - kin - pi;
}
// theta_cons_stage SC:268-292, specialised per stage
template <int STAGE>
WAMP_FN double theta_stage(const Eos& e, const double* l, const double* h) {
  double theta = 1.0;
  if (STAGE == 0) {
    theta = wmin(theta, theta_lower(l[0], h[0], PP_RHO));
    theta = wmin(theta, theta_lower(l[1], h[1], PP_RHO));
  } else if (STAGE == 1) {
    theta = wmin(theta, theta_lower(l[5], h[5], PP_ALPHA));
    theta = wmin(theta, theta_lower(1.0 - l[5], 1.0 - h[5], PP_ALPHA));
  } else {
    double lc = rho_ct2_cons(e, l[0], l[1], l[2], l[3], l[4], l[5]);
    double hc = rho_ct2_cons(e, h[0], h[1], h[2], h[3], h[4], h[5]);
    theta = theta_lower(lc, hc, PP_CTILDE);
  }
  if (theta < 0.0 || isnan_(theta)) theta = 0.0;
  if (theta > 1.0) theta = 1.0;
  return theta;
}

// ------------------------------------------------- reconstruction ----------
template <typename T>
WAMP_FN T mm2(T x, T y) {  // SC:295-297
  return T(0.5) * (msign(x) + msign(y)) * wmin(fabs(x), fabs(y));
}
template <typename T>
WAMP_FN T mm4(T a, T b, T c, T d) {  // SC:299-304
  T s = msign(a) + msign(b);
  return T(0.125) * s * fabs((msign(a) + msign(c)) * (msign(a) + msign(d))) *
         wmin(wmin(fabs(a), fabs(b)), wmin(fabs(c), fabs(d)));
}
// cell_entropy SC:308-314
WAMP_FN double entropy(const Eos& e, double ar1, double ar2, double p, double a1) {
  double rho = ar1 + ar2;
  if (rho < TINY) rho = TINY;
  double g, pi;
  mix_g_pi(e, a1, g, pi);
#ifdef WAMP_MIXED_TU
  // mixed unit: rho^g = exp(g log rho) for rho >= TINY.  Its relative error
  // is about |g log rho| ulps (< 1e-14 for the densities and gammas here)
  // against the library pow, which also handles negative, zero and infinite
  // bases; the indicator built on it only gates the reconstruction.
  return ddiv(p, exp(g * log(rho)));
#else
  // FP64 kernels: wm_pow (wamp_math.h), within one ulp of rho^g
  return ddiv(p, wm_pow(rho, g));
#endif
}
// cator2_val SC:318-325
WAMP_FN double cator2(double em2, double em1, double e0, double ep1, double ep2) {
  double aa = 0.25 * fabs(em2 - 4.0 * em1 + 3.0 * e0) + (13.0 / 12.0) * fabs(e0 - 2.0 * em1 + em2);
  double bb = 0.25 * fabs(3.0 * e0 - 4.0 * ep1 + ep2) + (13.0 / 12.0) * fabs(e0 - 2.0 * ep1 + ep2);
  return ddiv(2.0 * aa * bb + YUXIN, aa * aa + bb * bb + YUXIN);
}
// thinc_ul SC:327-337; t1 = tanh(beta/2) is a per-launch constant
template <typename T>
WAMP_FN T thinc_l(T vm1, T v0, T v1, T t1) {
  T result = v0;
  if ((v1 - v0) * (v0 - vm1) > T(0)) {
    T qa = T(0.5) * (v1 + vm1);
    T qd = T(0.5) * (v1 - vm1);
    T feng = dv(v0 - qa, qd);
    T t2 = th(feng * T(1.8) * T(0.5));
    result = qa + dv(qd * (t1 + dv(t2, t1)), T(1) + t2);
  }
  return result;
}
// thinc_ur SC:339-349
template <typename T>
WAMP_FN T thinc_r(T v0, T v1, T v2, T t1) {
  T result = v1;
  if ((v2 - v1) * (v1 - v0) > T(0)) {
    T qa = T(0.5) * (v2 + v0);
    T qd = T(0.5) * (v2 - v0);
    T feng = dv(v1 - qa, qd);
    T t2 = th(feng * T(1.8) * T(0.5));
    result = qa - dv(qd * (t1 - dv(t2, t1)), T(1) - t2);
  }
  return result;
}
// mp5_ul SC:353-378 (the blend of SC:371-375 is overwritten by SC:376)
template <typename T>
WAMP_FN T mp5_l(T vm2, T vm1, T v0, T vp1, T vp2) {
  T VOR = T(ONE_60) * (T(2) * vm2 - T(13) * vm1 + T(47) * v0 + T(27) * vp1 - T(3) * vp2);
  T VMP = v0 + mm2(vp1 - v0, T(MP5C) * (v0 - vm1));
  T result = VOR;
  if ((VOR - v0) * (VOR - VMP) >= T(EPSM)) {
    T DJM1 = vm2 - T(2) * vm1 + v0;
    T DJ = vm1 - T(2) * v0 + vp1;
    T DJP1 = v0 - T(2) * vp1 + vp2;
    T DM4JPH = mm4(T(4) * DJ - DJP1, T(4) * DJP1 - DJ, DJ, DJP1);
    T DM4JMH = mm4(T(4) * DJ - DJM1, T(4) * DJM1 - DJ, DJ, DJM1);
    T VUL = v0 + T(MP5C) * (v0 - vm1);
    T VAV = T(0.5) * (v0 + vp1);
    T VMD = VAV - T(0.5) * DM4JPH;
    T VLC = v0 + T(0.5) * (v0 - vm1) + T(B2) * DM4JMH;
    T VMIN = wmax(wmin(wmin(v0, vp1), VMD), wmin(wmin(v0, VUL), VLC));
    T VMAX = wmin(wmax(wmax(v0, vp1), VMD), wmax(wmax(v0, VUL), VLC));
    result = VOR + mm2(VMIN - VOR, VMAX - VOR);
  }
  return result;
}
// mp5_ur SC:380-405
template <typename T>
WAMP_FN T mp5_r(T vm1, T v0, T vp1, T vp2, T vp3) {
  T VOR = T(ONE_60) * (T(-3) * vm1 + T(27) * v0 + T(47) * vp1 - T(13) * vp2 + T(2) * vp3);
  T VMP = vp1 + mm2(v0 - vp1, T(MP5C) * (vp1 - vp2));
  T result = VOR;
  if ((VOR - vp1) * (VOR - VMP) >= T(EPSM)) {
    T DJM1 = vm1 - T(2) * v0 + vp1;
    T DJ = v0 - T(2) * vp1 + vp2;
    T DJP1 = vp1 - T(2) * vp2 + vp3;
    T DM4JPH = mm4(T(4) * DJ - DJP1, T(4) * DJP1 - DJ, DJ, DJP1);
    T DM4JMH = mm4(T(4) * DJ - DJM1, T(4) * DJM1 - DJ, DJ, DJM1);
    T VUL = vp1 + T(MP5C) * (vp1 - vp2);
    T VAV = T(0.5) * (vp1 + v0);
    T VMD = VAV - T(0.5) * DM4JMH;
    T VLC = vp1 + T(0.5) * (vp1 - vp2) + T(B2) * DM4JPH;
    T VMIN = wmax(wmin(wmin(vp1, v0), VMD), wmin(wmin(vp1, VUL), VLC));
    T VMAX = wmin(wmax(wmax(vp1, v0), VMD), wmax(wmax(vp1, VUL), VLC));
    result = VOR + mm2(VMIN - VOR, VMAX - VOR);
  }
  return result;
}
// central6_avg SC:407-412
template <typename T>
WAMP_FN T central6(T vm2, T vm1, T v0, T vp1, T vp2, T vp3) {
  T left = T(ONE_60) * (T(2) * vm2 - T(13) * vm1 + T(47) * v0 + T(27) * vp1 - T(3) * vp2);
  T right = T(ONE_60) * (T(-3) * vm1 + T(27) * v0 + T(47) * vp1 - T(13) * vp2 + T(2) * vp3);
  return T(0.5) * (left + right);
}
// 4th-order central interface value, SC:1077-1080
template <typename T>
WAMP_FN T central4(T vm1, T v0, T vp1, T vp2) {
  return (-T(1.0 / 12.0) * vm1 + T(7.0 / 12.0) * v0 + T(7.0 / 12.0) * vp1 - T(1.0 / 12.0) * vp2);
}
// muscl_char_ul / muscl_char_ur SC:430-446:
// v +- 1/4 ((1 - KAP) mm2(d, 2 e) + (1 + KAP) mm2(e, 2 d)).  Both minmods share
// their sign factor: sign(2 e) = sign(e) and sign(2 d) = sign(d) (also for
// zeros), so 0.5 (sign(d) + sign(e)) is formed once; every value is the one
// mm2 gives.
template <typename T>
WAMP_FN T muscl_pair(T d, T e) {
  const T s = T(0.5) * (msign(d) + msign(e));
  const T a = s * wmin(fabs(d), fabs(T(2) * e));  // mm2(d, 2 e)
  const T b = s * wmin(fabs(e), fabs(T(2) * d));  // mm2(e, 2 d)
  return T(0.25) * ((T(1) - T(KAP)) * a + (T(1) + T(KAP)) * b);
}
template <typename T>
WAMP_FN T muscl_l(T vm1, T v0, T vp1) {
  T dm = v0 - vm1;
  T d0 = vp1 - v0;
  return v0 + muscl_pair(dm, d0);
}
template <typename T>
WAMP_FN T muscl_r(T v0, T vp1, T vp2) {
  T dp = vp2 - vp1;
  T d0 = vp1 - v0;
  return vp1 - muscl_pair(dp, d0);
}

// ---------------------------------------------- characteristic basis -------
// Face-constant part of L and R (SC:462-478), formed once per face.
template <typename T>
struct Basis {
  T c, k, inv2c, inv2c2, y1c2, y2c2, ut, Y1, Y2, un_, ut_, c2;
};
// DIR 0: x-face (normal u, SC:480-534); DIR 1: y-face (normal v, SC:537-589)
template <int DIR, typename T>
WAMP_FN Basis<T> make_basis(T c, T rho, T ar1, T ar2, T u_avg, T v_avg) {
  Basis<T> b;
  b.c = c;
  b.inv2c = dv(T(0.5), c);
  const T c2 = c * c;
  b.inv2c2 = dv(T(0.5), c2);
  T Y1 = dv(ar1, rho);
  T Y2 = dv(ar2, rho);
  b.Y1 = Y1;
  b.Y2 = Y2;
  b.y1c2 = dv(Y1, c2);
  b.y2c2 = dv(Y2, c2);
  T un = DIR == 0 ? u_avg : v_avg;
  b.k = un * b.inv2c;
  b.ut = DIR == 0 ? v_avg : -u_avg;  // to_char_x: ut = v;  to_char_y: ut_f = -u
  b.un_ = u_avg;
  b.ut_ = v_avg;
  b.c2 = c * c;
  return b;
}
// L . u  (to_char_x SC:480-499 / to_char_y SC:537-556)
template <int DIR, typename T>
WAMP_FN void to_char(const Basis<T>& b, const T* u, T* w) {
  const T un_m = DIR == 0 ? u[2] : u[3];  // normal momentum
  const T ut_m = DIR == 0 ? u[3] : u[2];  // tangential momentum
  w[0] = b.k * u[0] + b.k * u[1] - b.inv2c * un_m + b.inv2c2 * u[4];
  w[1] = u[0] - b.y1c2 * u[4];
  w[2] = u[1] - b.y2c2 * u[4];
  if (DIR == 0)
    w[3] = -b.ut * u[0] - b.ut * u[1] + ut_m;
  else
    w[3] = -b.ut * u[0] - b.ut * u[1] - ut_m;
  w[4] = u[5];
  w[5] = -b.k * u[0] - b.k * u[1] + b.inv2c * un_m + b.inv2c2 * u[4];
}
// R . w  (to_phys_x SC:504-534 / to_phys_y SC:559-589)
template <int DIR, typename T>
WAMP_FN void to_phys(const Basis<T>& b, const T* w, T* s) {
  const T u_avg = b.un_, v_avg = b.ut_, c = b.c;
  s[0] = b.Y1 * (w[0] + w[5]) + w[1];
  s[1] = b.Y2 * (w[0] + w[5]) + w[2];
  if (DIR == 0) {
    s[2] = (u_avg - c) * w[0] + u_avg * w[1] + u_avg * w[2] + (u_avg + c) * w[5];
    s[3] = v_avg * (w[0] + w[1] + w[2] + w[5]) + w[3];
  } else {
    s[2] = u_avg * (w[0] + w[1] + w[2] + w[5]) - w[3];
    s[3] = (v_avg - c) * w[0] + v_avg * w[1] + v_avg * w[2] + (v_avg + c) * w[5];
  }
  s[4] = b.c2 * (w[0] + w[5]);
  s[5] = w[4];
}

// ----------------------------------------------- Wong face limiter ---------
// clamp01_pp as the reference's branches (compares and selects).  The FP64
// face kernels with the face positivity limiter run at the 168 VGPRs that
// three waves per SIMD allow; with rho_ct2_sc's clamp in this form their
// registers fit (the v_max / v_min form there made the M10 kernel spill 8
// values to scratch and run 7 % slower).
WAMP_FN double clamp01_sel(double a1) {
  double a = a1;
  if (isnan_(a)) a = PP_ALPHA;
  if (a < PP_ALPHA) a = PP_ALPHA;
  if (a > 1.0 - PP_ALPHA) a = 1.0 - PP_ALPHA;
  return a;
}
// rho_ctilde2_sc SC:174-178
WAMP_FN double rho_ct2_sc(const Eos& e, const double* s) {
  double a1 = clamp01_sel(s[5]);
  double g, pi;
  mix_g_pi(e, a1, g, pi);
  return ddiv(s[4] + pi, g - 1.0);
}
// wong_stage_lower_s6 SC:189-201 (in place on s)
WAMP_FN void wong_lower(const double* b, double* s, double bv, double hv,
                                           double floor_) {
  double theta = 1.0;
  if (isnan_(bv) || isnan_(hv) || bv < floor_) {
    theta = 0.0;
  } else if (hv < floor_) {
    double denom = bv - hv;
    theta = denom > TINY ? ddiv(bv - floor_, denom) : 0.0;
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) s[m] = b[m] + theta * (s[m] - b[m]);
}
// The M10 face positivity sequence for one side (M10:1007-1021), in place on
// the reconstructed state hi, with base the first-order state of that side:
// pp_limit_reconstructed_state (SC:217-240), the fallbacks of the partial
// densities, alpha1 and p to base, and pp_floor_pressure_state (SC:180-187).
// rho*c~^2 of the state leaving the limiter's hard-stop check is carried into
// the later checks and recomputed only in the (rare) branches that change p
// or alpha1, so every check sees the value a fresh rho_ctilde2_sc of the
// current state gives.
WAMP_FN void face_state_pp(const Eos& e, double* hi, const double* base) {
  double b[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) b[m] = base[m];
  b[5] = clamp01_pp(b[5]);
  b[0] = floor_pp(b[0], PP_RHO);
  b[1] = floor_pp(b[1], PP_RHO);
  wong_lower(b, hi, b[0], hi[0], PP_RHO);
  wong_lower(b, hi, b[1], hi[1], PP_RHO);
  wong_lower(b, hi, b[5], hi[5], PP_ALPHA);
  wong_lower(b, hi, 1.0 - b[5], 1.0 - hi[5], PP_ALPHA);
  wong_lower(b, hi, rho_ct2_sc(e, b), rho_ct2_sc(e, hi), PP_CTILDE);
  bool back = hi[0] < PP_HS_RHO || hi[1] < PP_HS_RHO || hi[5] < PP_HS_ALPHA ||
              hi[5] > 1.0 - PP_HS_ALPHA;
  if (back) {
#pragma unroll
    for (int m = 0; m < 6; ++m) hi[m] = b[m];
  }
  double v = rho_ct2_sc(e, hi);
  if (v < PP_HS_CTILDE || isnan_(hi[4])) {
#pragma unroll
    for (int m = 0; m < 6; ++m) hi[m] = b[m];
    v = rho_ct2_sc(e, hi);
  }
  // fallbacks to the first-order state (M10:1011-1015)
  if (hi[0] < 0.0 || isnan_(hi[0])) hi[0] = base[0];
  if (hi[1] < 0.0 || isnan_(hi[1])) hi[1] = base[1];
  if (hi[5] < 0.0 || hi[5] > 1.0) {
    hi[5] = base[5];
    v = rho_ct2_sc(e, hi);
  }
  if (v < PP_CTILDE || isnan_(hi[4])) {
    hi[4] = base[4];
    v = rho_ct2_sc(e, hi);
  }
  // pp_floor_pressure_state
  const double a1 = clamp01_pp(hi[5]);
  if (isnan_(hi[4]) || v < PP_CTILDE) {
    double g, pi;
    mix_g_pi(e, a1, g, pi);
    hi[4] = PP_CTILDE * (g - 1.0) - pi;
  }
  hi[5] = a1;
}

// ---------------------------------------------------------------- HLLC -----
// hllc_x + hllc_x_src (SC:594-687) / hllc_y + hllc_y_src (SC:690-771) fused:
// one evaluation of cL, cR, SL, SR, SP feeds the flux and the contact
// velocity u*.  Inputs are primitives (ar1, ar2, u, v, p, a1) per side and
// each side's sound speed c and total energy E (hllc below forms them; the
// limiter march that forms first-order fluxes carries them per cell).
//
// The reference's four cases (SL > 0; SL <= 0 < SP; SP <= 0 <= SR; else)
// take the left state in the first two and the right state in the last two,
// and add the star-state correction S (U*_K - U_K) in the middle two.  Here
// the side is selected first and one flux expression is evaluated: each
// component is the reference's expression of its case on the selected
// operands (the correction is selected in, not added as zero, so every
// result equals the case's expression bit for bit), and lanes that take
// different cases no longer run four bodies.
template <int DIR, typename T>
WAMP_FN void hllc_ce(const Eos& e, const T* L, const T* R, T cL, T cR, T EL, T ER, T* f,
                     T& ustar) {
  const T ar1L = L[0], ar2L = L[1], uL = L[2], vL = L[3], pL = L[4], a1L = L[5];
  const T ar1R = R[0], ar2R = R[1], uR = R[2], vR = R[3], pR = R[4], a1R = R[5];
  const T nL = DIR == 0 ? uL : vL, nR = DIR == 0 ? uR : vR;
  T rhoL = ar1L + ar2L;
  T rhoR = ar1R + ar2R;
  T n_avg = T(0.5) * (nL + nR);
  T c_avg = T(0.5) * (cL + cR);
  T SL = wmin(nL - cL, n_avg - c_avg);
  T SR = wmax(nR + cR, n_avg + c_avg);
  T denom = rhoL * (SL - nL) - rhoR * (SR - nR);
  if (fabs(denom) < T(TINY)) denom = T(TINY);
  T SP = dv(pR - pL + rhoL * nL * (SL - nL) - rhoR * nR * (SR - nR), denom);  // |denom| >= TINY
  T ratL = (SL - nL) / (SL - SP);
  T ratR = (SR - nR) / (SR - SP);

  const bool c1 = SL > T(0);
  const bool c2 = !c1 && SL <= T(0) && T(0) < SP;
  const bool c3 = !c1 && !c2 && SP <= T(0) && T(0) <= SR;
  const bool left = c1 || c2, star = c2 || c3;
  const T ar1 = left ? ar1L : ar1R, ar2 = left ? ar2L : ar2R;
  const T u = left ? uL : uR, v = left ? vL : vR, p = left ? pL : pR, a1 = left ? a1L : a1R;
  const T rho = left ? rhoL : rhoR, n = left ? nL : nR, E = left ? EL : ER;
  const T S = left ? SL : SR, ratio = left ? ratL : ratR;
  // SL - nL <= -cL < 0 and SR - nR >= cR > 0 wherever the star state is taken
  const T E_star = E + (SP - n) * (SP * rho + dv(p, S - n));
  T b[6], c[6];  // the case's flux without and with the star correction
  b[0] = ar1 * n;
  b[1] = ar2 * n;
  if (DIR == 0) {
    b[2] = rho * u * u + p;
    b[3] = rho * u * v;
  } else {
    b[2] = rho * u * v;
    b[3] = rho * v * v + p;
  }
  b[4] = (E + p) * n;
  b[5] = a1 * n;
  c[0] = b[0] + S * (ar1 * ratio - ar1);
  c[1] = b[1] + S * (ar2 * ratio - ar2);
  if (DIR == 0) {
    c[2] = b[2] + S * (rho * SP * ratio - rho * u);
    c[3] = b[3] + S * (rho * v * ratio - rho * v);
  } else {
    c[2] = b[2] + S * (rho * u * ratio - rho * u);
    c[3] = b[3] + S * (rho * SP * ratio - rho * v);
  }
  c[4] = b[4] + S * (E_star * ratio - E);
  c[5] = b[5] + S * (a1 * ratio - a1);
#pragma unroll
  for (int m = 0; m < 6; ++m) f[m] = star ? c[m] : b[m];
  T left_part = nL + wmin(T(0), SL) * (ratL - T(1));
  T right_part = nR + wmax(T(0), SR) * (ratR - T(1));
  T sp_sign = wsign(SP);
  ustar = T(0.5) * (T(1) + sp_sign) * left_part + T(0.5) * (T(1) - sp_sign) * right_part;
}
template <int DIR, typename T>
WAMP_FN void hllc(const Eos& e, const T* L, const T* R, T* f, T& ustar) {
  const T cL = sound_speed(e, L[0] + L[1], L[4], L[5]);
  const T cR = sound_speed(e, R[0] + R[1], R[4], R[5]);
  const T EL = prim_energy(e, L[0], L[1], L[2], L[3], L[4], L[5]);
  const T ER = prim_energy(e, R[0], R[1], R[2], R[3], R[4], R[5]);
  hllc_ce<DIR>(e, L, R, cL, cR, EL, ER, f, ustar);
}

// HLLC for the mixed-precision path: wave speeds, contact speed and the
// mass / momentum / volume-fraction fluxes in FP32; the energy flux in FP64.
// E = gb p + pb + kin carries pb = gamma pi / (gamma - 1) (7765 for the
// nondimensional water against p ~ 1), so an FP32 energy flux would put
// ~eps32 * pb of noise into E and hence into the recovered pressure.
template <int DIR>
WAMP_FN void hllc_mixed(const Eos& e, const float* L, const float* R, double* f,
                                           double& ustar) {
  const float ar1L = L[0], ar2L = L[1], uL = L[2], vL = L[3], pL = L[4], a1L = L[5];
  const float ar1R = R[0], ar2R = R[1], uR = R[2], vR = R[3], pR = R[4], a1R = R[5];
  const float nL = DIR == 0 ? uL : vL, nR = DIR == 0 ? uR : vR;
  float rhoL = ar1L + ar2L;
  float rhoR = ar1R + ar2R;
  float cL = sound_speed<float>(e, rhoL, pL, a1L);
  float cR = sound_speed<float>(e, rhoR, pR, a1R);
  float n_avg = 0.5f * (nL + nR);
  float c_avg = 0.5f * (cL + cR);
  float SL = wmin(nL - cL, n_avg - c_avg);
  float SR = wmax(nR + cR, n_avg + c_avg);
  float denom = rhoL * (SL - nL) - rhoR * (SR - nR);
  if (fabsf(denom) < float(TINY)) denom = float(TINY);
  float SP = dv(pR - pL + rhoL * nL * (SL - nL) - rhoR * nR * (SR - nR), denom);
  float ratL = dv(SL - nL, SL - SP);
  float ratR = dv(SR - nR, SR - SP);
  // FP64 energies from the FP32 primitives
  const double EL = prim_energy<double>(e, ar1L, ar2L, uL, vL, pL, a1L);
  const double ER = prim_energy<double>(e, ar1R, ar2R, uR, vR, pR, a1R);
  if (SL > 0.f) {
    f[0] = ar1L * nL;
    f[1] = ar2L * nL;
    if (DIR == 0) { f[2] = rhoL * uL * uL + pL; f[3] = rhoL * uL * vL; }
    else          { f[2] = rhoL * uL * vL;      f[3] = rhoL * vL * vL + pL; }
    f[4] = (EL + (double)pL) * (double)nL;
    f[5] = a1L * nL;
  } else if (SL <= 0.f && 0.f < SP) {
    float ratio = ratL;
    f[0] = ar1L * nL + SL * (ar1L * ratio - ar1L);
    f[1] = ar2L * nL + SL * (ar2L * ratio - ar2L);
    if (DIR == 0) {
      f[2] = (rhoL * uL * uL + pL) + SL * (rhoL * SP * ratio - rhoL * uL);
      f[3] = (rhoL * uL * vL) + SL * (rhoL * vL * ratio - rhoL * vL);
    } else {
      f[2] = (rhoL * uL * vL) + SL * (rhoL * uL * ratio - rhoL * uL);
      f[3] = (rhoL * vL * vL + pL) + SL * (rhoL * SP * ratio - rhoL * vL);
    }
    const double sl = SL, sp = SP, nl = nL, rl = rhoL, pl = pL;
    const double EL_star = EL + (sp - nl) * (sp * rl + ddiv(pl, sl - nl));
    f[4] = (EL + pl) * nl + sl * (EL_star * (double)ratio - EL);
    f[5] = a1L * nL + SL * (a1L * ratio - a1L);
  } else if (SP <= 0.f && 0.f <= SR) {
    float ratio = ratR;
    f[0] = ar1R * nR + SR * (ar1R * ratio - ar1R);
    f[1] = ar2R * nR + SR * (ar2R * ratio - ar2R);
    if (DIR == 0) {
      f[2] = (rhoR * uR * uR + pR) + SR * (rhoR * SP * ratio - rhoR * uR);
      f[3] = (rhoR * uR * vR) + SR * (rhoR * vR * ratio - rhoR * vR);
    } else {
      f[2] = (rhoR * uR * vR) + SR * (rhoR * uR * ratio - rhoR * uR);
      f[3] = (rhoR * vR * vR + pR) + SR * (rhoR * SP * ratio - rhoR * vR);
    }
    const double sr = SR, sp = SP, nr = nR, rr = rhoR, pr = pR;
    const double ER_star = ER + (sp - nr) * (sp * rr + ddiv(pr, sr - nr));
    f[4] = (ER + pr) * nr + sr * (ER_star * (double)ratio - ER);
    f[5] = a1R * nR + SR * (a1R * ratio - a1R);
  } else {
    f[0] = ar1R * nR;
    f[1] = ar2R * nR;
    if (DIR == 0) { f[2] = rhoR * uR * uR + pR; f[3] = rhoR * uR * vR; }
    else          { f[2] = rhoR * uR * vR;      f[3] = rhoR * vR * vR + pR; }
    f[4] = (ER + (double)pR) * (double)nR;
    f[5] = a1R * nR;
  }
  float left_part = nL + wmin(0.f, SL) * (ratL - 1.f);
  float right_part = nR + wmax(0.f, SR) * (ratR - 1.f);
  float sp_sign = wsign(SP);
  ustar = 0.5f * (1.f + sp_sign) * left_part + 0.5f * (1.f - sp_sign) * right_part;
}

WAMP_DEV_END
