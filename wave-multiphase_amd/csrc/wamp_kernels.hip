// wamp_kernels.hip -- fused HIP kernels of the SSP-RK3 step for gfx950 (FP64).
//
// Per RK stage the reference launches ~20 kernels over 47 arrays
// (SC:2031-2037).  Here a stage is two launches, a face kernel and a limiter:
//
//   k_faces        (wamp_faces.h) one LDS tile of 8x32 face positions: loads the 6
//                  conserved fields of the tile plus a 3/4-cell halo once
//                  (the boundary fill SC:1616-1682 is applied by index
//                  mapping while loading), forms pressure, entropy, the
//                  entropy indicator and the Ducros sensor once per cell,
//                  then every thread evaluates one x-face and one y-face,
//                  high order (fx_kernel SC:966-1133 / M10:868-1045) and,
//                  unless the limiter forms them itself, first order
//                  (fx_low_kernel SC:1303-1347), and stores the face arrays.
//   limiter        the three Wong theta stages (pp_cell_theta / pp_blend,
//                  SC:1397-1548), the final blended fluxes, the SSP-RK3
//                  combine with the alpha-source term (rk3_kernel
//                  SC:1553-1611), the emergency repair (SC:918-958), the alpha
//                  clamp (SC:851-854) and, on the last stage, the CFL minimum
//                  (SC:1686-1705), in one of three kernels with bitwise the
//                  same results:
//     k_limit_update  a 16x16 cell tile with a 3-cell halo frame in LDS
//                     (grids up to 2^18 cells);
//     k_lu_march      a wave per (58-column strip, segment of rows) marching
//                     along i, neighbours by DPP lane shifts, carries in
//                     registers and an LDS ring (larger grids, mixed storage);
//     k_lu_lo         the march that also forms the first-order fluxes from
//                     the stage input it streams (the FP64 step on larger
//                     grids; the face kernel then writes only the 14
//                     high-order planes).
//   k_faces_batch / k_lu_batch run the same code over the case table of a
//   batch.
//
// Compiled with -ffp-contract=off: every expression keeps the reference's
// association.  pow (the entropy) and tanh (THINC) come from wamp_math.h, so
// the FP64 results equal the oracle built with those functions bit for bit.
// This file instantiates the FP64 face kernels; the FP32-mixed ones live in
// wamp_faces_mixed.hip and are reached through wamp_launch_faces_mixed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wamp_device.h"
#include "wamp_faces.h"
#include "wamp_limit.h"
#include "wamp_types.h"

namespace wamp {

// ========================================================================
// k_limit_update
// ========================================================================
constexpr int LT = 16;             // cells per tile side
constexpr int LH = 3;              // theta halo
constexpr int LF = LT + 2 * LH;    // frame side (22)
constexpr int LTHREADS = 256;

// Block-wide min of a positive double, one atomic per block (bit pattern
// order equals numeric order for non-negative doubles).
__device__ __forceinline__ void block_min_atomic(double v, double* acc) {
  __shared__ double wmins[LTHREADS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = wmin(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wmins[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmins[0];
    for (int w = 1; w < LTHREADS / 64; ++w) m = wmin(m, wmins[w]);
    atomicMin((unsigned long long*)acc, (unsigned long long)__double_as_longlong(m));
  }
}

template <bool F32>
__global__ void __launch_bounds__(LTHREADS)
k_limit_update(Geo G, Eos e, State Un, State U, Faces F, State Uo, double alpha, double beta,
               const double* __restrict__ dt_dev, double* dt_acc, int repair) {
  __shared__ double th[3][LF * LF];
  const int tid = threadIdx.x;
  const int i0 = G.g + (int)blockIdx.y * LT;
  const int j0 = G.g + (int)blockIdx.x * LT;
  const size_t ld = (size_t)G.ld;
  const double dt = *dt_dev;
  const double inv_dx = 1.0 / G.dx;
  const double inv_dy = 1.0 / G.dy;

  auto interior = [&](int i, int j) {
    return i >= G.g && i < G.g + G.nx && j >= G.g && j < G.g + G.ny;
  };
  auto faces_of = [&](int a, int b, int nb, CellFaces& cf) {
    for (int k = 0; k < nb; ++k) {
      const double* t = th[k];
      const int r = a * LF + b;
      cf.t[0][k] = face_theta(t[r - LF], t[r]);
      cf.t[1][k] = face_theta(t[r], t[r + LF]);
      cf.t[2][k] = face_theta(t[r - 1], t[r]);
      cf.t[3][k] = face_theta(t[r], t[r + 1]);
    }
  };

  // Three theta stages on frames of halo 3, 2, 1 (pp_cell_theta SC:1436-1476).
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int lo = s, n = LF - 2 * s;
    for (int k = tid; k < n * n; k += LTHREADS) {
      const int a = lo + k / n, b = lo + k % n;
      const int i = i0 - LH + a, j = j0 - LH + b;
      double theta = 1.0;
      if (interior(i, j)) {
        const size_t c = (size_t)i * ld + (size_t)j;
        CellFaces cf;
        faces_of(a, b, s, cf);
        if (s == 0) theta = cell_theta<0, F32>(e, U, F, c, ld, cf, inv_dx, inv_dy, dt);
        else if (s == 1) theta = cell_theta<1, F32>(e, U, F, c, ld, cf, inv_dx, inv_dy, dt);
        else theta = cell_theta<2, F32>(e, U, F, c, ld, cf, inv_dx, inv_dy, dt);
      }
      th[s][a * LF + b] = theta;
    }
    __syncthreads();
  }

  // Final fluxes, SSP-RK3 combine (rk3_kernel SC:1592-1611), repair, clamp.
  const int a = LH + tid / LT, b = LH + tid % LT;
  const int i = i0 - LH + a, j = j0 - LH + b;
  double local_dt = 1.0e30;
  if (interior(i, j)) {
    const size_t c = (size_t)i * ld + (size_t)j;
    CellFaces cf;
    faces_of(a, b, 3, cf);
    double un[6];
    cell_update<F32>(e, Un, U, F, c, ld, cf, inv_dx, inv_dy, dt, alpha, beta, repair, un);
#pragma unroll
    for (int m = 0; m < 6; ++m) Uo.f[m][c] = un[m];
    if (dt_acc) local_dt = cell_dt(e, un, G.dx, G.dy);
  }
  if (dt_acc) block_min_atomic(local_dt, dt_acc);
}

// ========================================================================
// k_lu_march -- the limiter + update of k_limit_update as a register
// pipeline.  One wave owns 64 consecutive j (58 output columns plus a 3-lane
// halo each side: the update of a cell needs theta2 at +-1, theta1 at +-2,
// theta0 at +-3 cells) and marches along i.  At front row f it forms
// theta0(f), theta1(f-1), theta2(f-2) and the update of row f-3.
//
// Every face and state value is read from memory once per wave: a value
// needed again at a later front is carried, in registers or in a wave-private
// LDS ring indexed by row, instead of being re-fetched (a re-fetch one to
// three fronts later misses the L2 once all resident waves' windows are
// counted):
//   x-faces  row f: components {0,1} (theta0); row f-1: {5,6} (theta1); row
//            f-2: {2,3,4} (theta2).  Rows f-2..f of {0,1} and f-2..f-1 of
//            {5,6} stay in registers; rows f-3..f-2 of all components too.
//   y-faces  row f: {0,1} (theta0); row f-1: {5,6} (theta1); row f-2:
//            {2,3,4} (theta2).  {0,1} and {5,6} wait in LDS until theta2.
//            The face below a cell is the neighbouring lane's (shuffle),
//            except for theta0, whose lower faces are loaded so theta0 is
//            exact on every lane.
//   state    U0, U1 at theta0, U5 at theta1, U2..U4 at theta2; all wait in
//            LDS until the update one front after theta2.
// Values of rows the wave has not visited yet read as zero (the ring is
// cleared first); they only feed thetas of rows that no output depends on.
// ========================================================================
constexpr int MW = 64;
constexpr int MH = 3;
constexpr int MJ = MW - 2 * MH;

// LDS carry ring, doubles per lane: [slot][value] blocks of 64 lanes.
constexpr int LC_Y01 = 0;    // y-faces H0 H1 L0 L1, 3 row slots
constexpr int LC_Y56 = 12;   // y-faces H5 H6 L5 L6, 2 row slots
constexpr int LC_U01 = 20;   // U0 U1, 4 row slots
constexpr int LC_U5 = 28;    // U5, 3 row slots
constexpr int LC_U234 = 31;  // U2 U3 U4, 2 row slots
constexpr int LU_CARRY = 37;
constexpr int LU_WAVES = 4;  // waves per limiter block

// Neighbouring lane's value (lane_up: lane i gets lane i-1; lane_down: lane i
// gets lane i+1) by a DPP wave shift of each 32-bit half: one VALU move per
// half instead of an LDS permute.  Lane 0 (lane_up) and lane 63 (lane_down)
// get 0 (the move's bound control writes zero where the source lane is out
// of range, so no zero has to be placed in the destination first); they are
// halo lanes, whose thetas never reach an output lane.
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double lane_shift(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double lane_up(double v) { return lane_shift<DPP_WAVE_SHR1>(v); }
__device__ __forceinline__ double lane_down(double v) { return lane_shift<DPP_WAVE_SHL1>(v); }

// f <- L + theta_k (f - L) for k < nb (SC:1506-1512 once per limiter stage)
__device__ __forceinline__ double blend3(double H, double L, const double* th, int nb) {
  double f = H;
  for (int k = 0; k < nb; ++k) f = L + th[k] * (f - L);
  return f;
}

// One wave's work item: columns strip*58-3 .. +64, rows seg*rows .. +rows.
// `cw` is this wave's LU_CARRY x 64 doubles of LDS.  Returns this lane's CFL
// minimum of the updated cells (1e30 if none or if want_dt is false).
template <bool F32>
__device__ __forceinline__ double lu_wave(const Geo& G, const Eos& e, const State& Un,
                                          const State& U, const Faces& F, const State& Uo,
                                          double alpha, double beta, double dt, int repair,
                                          bool want_dt, int strip, int seg, int rows, double* cw) {
  const int lane = threadIdx.x & (MW - 1);
  const int j = G.g + strip * MJ - MH + lane;
  const int i0 = G.g + seg * rows;
  if (i0 >= G.g + G.nx) return 1.0e30;  // wave-uniform
  const int i_end = min(i0 + rows, G.g + G.nx);
  const size_t ld = (size_t)G.ld;
  const double inv_dx = 1.0 / G.dx;
  const double inv_dy = 1.0 / G.dy;
  const bool col_in = j >= G.g && j < G.g + G.ny;          // interior column
  const bool yface_in = j >= G.g - 1 && j < G.g + G.ny;   // y-face (r, j+1/2) exists
  const bool out_lane = lane >= MH && lane < MH + MJ && col_in;
  auto row_in = [&](int r) { return r >= G.g && r < G.g + G.nx; };
  auto xrow_in = [&](int r) { return r >= G.g - 1 && r < G.g + G.nx; };  // x-face r+1/2 exists
  auto at = [&](int r) { return (size_t)r * ld + (size_t)j; };
  auto lds = [&](int q) -> double& { return cw[q * MW + lane]; };
  // face component m: high F[base + m], low F[base + 14 + m]; base 0 = x, 7 = y
  // Face value at row r, column j + dj (dj = 0 or -1), zero where !ok.  On the
  // mixed storage the load is issued on every lane at a clamped in-allocation
  // address (rows 0..nx+2g-1, columns 0..ld-1; j >= g-3 >= 1) and masked
  // afterwards: a masked FP32 load would be converted to FP64 inside its
  // branch, and that conversion waits for the load.
  const int nrows = G.nx + 2 * G.g;
  const size_t jc = (size_t)min(j, G.ld - 1);
  auto face = [&](bool ok, int m, size_t c, int r, int dj) -> double {  // c = at(r) + dj
    if constexpr (F32) {
      const size_t cs = (size_t)min(max(r, 0), nrows - 1) * ld + jc + dj;
      const double v = fload<true>(F, m, cs);
      return ok ? v : 0.0;
    } else {
      return ok ? fload<false>(F, m, c) : 0.0;
    }
  };

#pragma unroll
  for (int q = 0; q < LU_CARRY; ++q) lds(q) = 0.0;

  double x01[3][4] = {};   // x-face rows f-2, f-1, f: H0 H1 L0 L1
  double x56[2][4] = {};   // x-face rows f-2, f-1: H5 H6 L5 L6
  double xa[2][14] = {};   // x-face rows f-3, f-2: H0..6 L0..6
  double F2xm[7] = {};     // final flux at x-face row f-4 (below the update row)
  // A face's blended flux is the same for the cells on either side of it:
  // the x-face below theta1's / theta2's row is the face above the previous
  // front's row (the same face thetas), and a y-face below a lane is the face
  // above the lane before (a lane shift).  Those are carried or shifted, not
  // blended twice.
  double B1x[2] = {};      // x-face row f-2, H5 H6 blended once
  double B2x[7] = {};      // x-face row f-3, H0..H6 blended twice
  double F2yu[7] = {};     // final flux at the y-faces above row f-3
  double t0[5] = {1.0, 1.0, 1.0, 1.0, 1.0};  // theta0 rows f-4 .. f
  double t1[4] = {1.0, 1.0, 1.0, 1.0};       // theta1 rows f-4 .. f-1
  double t2[3] = {1.0, 1.0, 1.0};            // theta2 rows f-4 .. f-2
  double local_dt = 1.0e30;

  // The first theta0 row, i0-3, is part of the first output row's dependency
  // cone (update(i0) <- theta2(i0-1) <- theta1(i0-2) <- theta0(i0-3)), so its
  // lower x-face, row i0-4, is loaded before the march starts.
  {
    const int r = i0 - 4;
    const bool ok = xrow_in(r) && col_in;
    const size_t c = at(r);
    x01[2][0] = face(ok, 0, c, r, 0);
    x01[2][1] = face(ok, 1, c, r, 0);
    x01[2][2] = face(ok, 14, c, r, 0);
    x01[2][3] = face(ok, 15, c, r, 0);
  }

  for (int f = i0 - 3; f < i_end + 3; ++f) {
    const int s3f = f % 3, s3m1 = (f + 2) % 3, s3m2 = (f + 1) % 3;  // slots of rows f, f-1, f-2
    const int s2m1 = (f + 1) & 1, s2m2 = f & 1;                      // rows f-1, f-2 (f-3 = f-1)
    const int s4f = f & 3, s4m2 = (f + 2) & 3, s4m3 = (f + 1) & 3;  // rows f, f-2, f-3
    // ---------------- x-face loads of this front ----------------
    {
      const bool ok = xrow_in(f) && col_in;
      const size_t c = at(f);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        x01[0][q] = x01[1][q];
        x01[1][q] = x01[2][q];
      }
      x01[2][0] = face(ok, 0, c, f, 0);
      x01[2][1] = face(ok, 1, c, f, 0);
      x01[2][2] = face(ok, 14, c, f, 0);
      x01[2][3] = face(ok, 15, c, f, 0);
    }
    {
      const bool ok = xrow_in(f - 1) && col_in;
      const size_t c = at(f - 1);
#pragma unroll
      for (int q = 0; q < 4; ++q) x56[0][q] = x56[1][q];
      x56[1][0] = face(ok, 5, c, f - 1, 0);
      x56[1][1] = face(ok, 6, c, f - 1, 0);
      x56[1][2] = face(ok, 19, c, f - 1, 0);
      x56[1][3] = face(ok, 20, c, f - 1, 0);
    }
    {
      // row f-2: components 0, 1 and 5, 6 are already held (same predicate)
      const bool ok = xrow_in(f - 2) && col_in;
      const size_t c = at(f - 2);
#pragma unroll
      for (int m = 0; m < 14; ++m) xa[0][m] = xa[1][m];
      xa[1][0] = x01[0][0];
      xa[1][1] = x01[0][1];
      xa[1][7] = x01[0][2];
      xa[1][8] = x01[0][3];
      xa[1][5] = x56[0][0];
      xa[1][6] = x56[0][1];
      xa[1][12] = x56[0][2];
      xa[1][13] = x56[0][3];
#pragma unroll
      for (int m = 2; m < 5; ++m) {
        xa[1][m] = face(ok, m, c, f - 2, 0);
        xa[1][7 + m] = face(ok, 14 + m, c, f - 2, 0);
      }
    }

    // ---- the front's remaining global loads, all issued before its arithmetic so
    //      that they are in flight together (each stage below only waits for its own)
    const bool upd = f - 3 >= i0 && f - 3 < i_end && out_lane;  // row f-3 is output
    double yc[4], yd0[4], u01[2];  // theta0, row f
    {
      const size_t c = at(f);
      const bool ok = row_in(f) && col_in;
      const bool yok = row_in(f) && yface_in;
      yc[0] = face(yok, 7, c, f, 0);
      yc[1] = face(yok, 8, c, f, 0);
      yc[2] = face(yok, 21, c, f, 0);
      yc[3] = face(yok, 22, c, f, 0);
      // lower y-faces are loaded (not shuffled) so theta0 is exact on every lane
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        yd0[m] = face(ok, 7 + m, c - 1, f, -1);
        yd0[2 + m] = face(ok, 21 + m, c - 1, f, -1);
      }
      u01[0] = ok ? U.f[0][c] : 0.0;
      u01[1] = ok ? U.f[1][c] : 0.0;
    }
    double yu[4], al1;  // theta1, row f-1
    {
      const size_t c = at(f - 1);
      const bool ok = row_in(f - 1) && col_in;
      const bool yok = row_in(f - 1) && yface_in;
      yu[0] = face(yok, 12, c, f - 1, 0);
      yu[1] = face(yok, 13, c, f - 1, 0);
      yu[2] = face(yok, 26, c, f - 1, 0);
      yu[3] = face(yok, 27, c, f - 1, 0);
      al1 = ok ? U.f[5][c] : 0.0;
    }
    double yn24[6], u234[3];  // theta2, row f-2: y-face components 2..4 (H, L), U2..U4
    {
      const size_t c = at(f - 2);
      const bool ok = row_in(f - 2) && col_in;
      const bool yok = row_in(f - 2) && yface_in;
#pragma unroll
      for (int m = 2; m < 5; ++m) {
        yn24[m - 2] = face(yok, 7 + m, c, f - 2, 0);
        yn24[m + 1] = face(yok, 21 + m, c, f - 2, 0);
      }
      u234[0] = ok ? U.f[2][c] : 0.0;
      u234[1] = ok ? U.f[3][c] : 0.0;
      u234[2] = ok ? U.f[4][c] : 0.0;
    }

    // ---------------- theta0 at row f (raw high-order fluxes) ----------------
    double th = 1.0;
    {
      const bool ok = row_in(f) && col_in;
      if (ok) {
        double l[6], h[6];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          double rh = (x01[2][m] - x01[1][m]) * inv_dx + (yc[m] - yd0[m]) * inv_dy;
          double rl = (x01[2][2 + m] - x01[1][2 + m]) * inv_dx + (yc[2 + m] - yd0[2 + m]) * inv_dy;
          l[m] = u01[m] - dt * rl;
          h[m] = u01[m] - dt * rh;
        }
        th = theta_stage<0>(e, l, h);
      }
      // carried values are stored once consumed: a store issued right after its
      // load would wait for every load issued before it
#pragma unroll
      for (int q = 0; q < 4; ++q) lds(LC_Y01 + 4 * s3f + q) = yc[q];
      lds(LC_U01 + 2 * s4f) = u01[0];
      lds(LC_U01 + 2 * s4f + 1) = u01[1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) t0[k] = t0[k + 1];
    t0[4] = th;

    // ---------------- theta1 at row f-1 (fluxes blended once) ----------------
    {
      const int r = f - 1;
      const bool ok = row_in(r) && col_in;
      const double al = al1;
      double tf[4];
      tf[0] = face_theta(t0[2], t0[3]);
      tf[1] = face_theta(t0[3], t0[4]);
      tf[2] = face_theta(lane_up(t0[3]), t0[3]);
      tf[3] = face_theta(t0[3], lane_down(t0[3]));
      double yd[4];  // [2], [3]: L5 L6 of the face below
#pragma unroll
      for (int q = 2; q < 4; ++q) yd[q] = lane_up(yu[q]);
      // the faces above and to the side of the row, blended once on every lane
      const double Hxp5 = blend3(x56[1][0], x56[1][2], &tf[1], 1);
      const double Hxp6 = blend3(x56[1][1], x56[1][3], &tf[1], 1);
      const double Hyp5 = blend3(yu[0], yu[2], &tf[3], 1);
      const double Hyp6 = blend3(yu[1], yu[3], &tf[3], 1);
      const double Hxm5 = B1x[0], Hxm6 = B1x[1];
      const double Hym5 = lane_up(Hyp5), Hym6 = lane_up(Hyp6);
      B1x[0] = Hxp5;
      B1x[1] = Hxp6;
      double th1 = 1.0;
      if (ok) {
        double rh = (Hxp5 - Hxm5) * inv_dx + (Hyp5 - Hym5) * inv_dy;
        double rl = (x56[1][2] - x56[0][2]) * inv_dx + (yu[2] - yd[2]) * inv_dy;
        rh = rh - al * (Hxp6 - Hxm6) * inv_dx - al * (Hyp6 - Hym6) * inv_dy;
        rl = rl - al * (x56[1][3] - x56[0][3]) * inv_dx - al * (yu[3] - yd[3]) * inv_dy;
        double l[6], h[6];
        l[5] = al - dt * rl;
        h[5] = al - dt * rh;
        th1 = theta_stage<1>(e, l, h);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) lds(LC_Y56 + 4 * s2m1 + q) = yu[q];
      lds(LC_U5 + s3m1) = al;
#pragma unroll
      for (int k = 0; k < 3; ++k) t1[k] = t1[k + 1];
      t1[3] = th1;
    }

    // ---------------- theta2 at row f-2 (fluxes blended twice) ----------------
    double yn[14];  // y-faces above row f-2 (H0..6, L0..6)
    double hm[7], hu[7];  // x-face row f-3, y-faces above row f-2: blended twice
    {
      const int r = f - 2;
      const bool ok = row_in(r) && col_in;
      yn[0] = lds(LC_Y01 + 4 * s3m2 + 0);
      yn[1] = lds(LC_Y01 + 4 * s3m2 + 1);
      yn[7] = lds(LC_Y01 + 4 * s3m2 + 2);
      yn[8] = lds(LC_Y01 + 4 * s3m2 + 3);
      yn[5] = lds(LC_Y56 + 4 * s2m2 + 0);
      yn[6] = lds(LC_Y56 + 4 * s2m2 + 1);
      yn[12] = lds(LC_Y56 + 4 * s2m2 + 2);
      yn[13] = lds(LC_Y56 + 4 * s2m2 + 3);
#pragma unroll
      for (int m = 2; m < 5; ++m) {
        yn[m] = yn24[m - 2];
        yn[7 + m] = yn24[m + 1];
      }
      double tf[4][2];
      tf[0][0] = face_theta(t0[1], t0[2]);
      tf[1][0] = face_theta(t0[2], t0[3]);
      tf[2][0] = face_theta(lane_up(t0[2]), t0[2]);
      tf[3][0] = face_theta(t0[2], lane_down(t0[2]));
      tf[0][1] = face_theta(t1[1], t1[2]);
      tf[1][1] = face_theta(t1[2], t1[3]);
      tf[2][1] = face_theta(lane_up(t1[2]), t1[2]);
      tf[3][1] = face_theta(t1[2], lane_down(t1[2]));
      // the faces above and to the side of the row, blended twice on every lane
      double hp[7], hd[7], ydl[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        hp[m] = blend3(xa[1][m], xa[1][7 + m], tf[1], 2);
        hu[m] = blend3(yn[m], yn[7 + m], tf[3], 2);
        hm[m] = B2x[m];
        hd[m] = lane_up(hu[m]);
        ydl[m] = lane_up(yn[7 + m]);
        B2x[m] = hp[m];
      }
      double th2 = 1.0;
      if (ok) {
        double uc[6];
        uc[0] = lds(LC_U01 + 2 * s4m2);
        uc[1] = lds(LC_U01 + 2 * s4m2 + 1);
        uc[2] = u234[0];
        uc[3] = u234[1];
        uc[4] = u234[2];
        uc[5] = lds(LC_U5 + s3m2);
        const double al = uc[5];
        double l[6], h[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          double rh = (hp[m] - hm[m]) * inv_dx + (hu[m] - hd[m]) * inv_dy;
          double rl = (xa[1][7 + m] - xa[0][7 + m]) * inv_dx + (yn[7 + m] - ydl[m]) * inv_dy;
          if (m == 5) {
            rh = rh - al * (hp[6] - hm[6]) * inv_dx - al * (hu[6] - hd[6]) * inv_dy;
            rl = rl - al * (xa[1][13] - xa[0][13]) * inv_dx - al * (yn[13] - ydl[6]) * inv_dy;
          }
          const double u = uc[m];
          l[m] = u - dt * rl;
          h[m] = u - dt * rh;
        }
        th2 = theta_stage<2>(e, l, h);
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) lds(LC_U234 + 3 * s2m2 + q) = u234[q];
#pragma unroll
      for (int k = 0; k < 2; ++k) t2[k] = t2[k + 1];
      t2[2] = th2;
    }

    // ---- final fluxes (three blends) of the faces whose thetas are now complete:
    //      x-face row f-3 (between rows f-3, f-2) and the y-faces above row f-2
    //      (their first two blends are theta2's: one more with the stage-2 face theta)
    double F2xp[7], F2yn[7];
    {
      const double tx = face_theta(t2[1], t2[2]);             // rows f-3, f-2
      const double ty = face_theta(t2[2], lane_down(t2[2]));  // row f-2 and the lane after
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        F2xp[m] = blend3(hm[m], xa[0][7 + m], &tx, 1);
        F2yn[m] = blend3(hu[m], yn[7 + m], &ty, 1);
      }
    }

    // ---------------- update of row f-3 ----------------
    {
      const int r = f - 3;
      const size_t c = at(r);
      double unv[6];
#pragma unroll
      for (int m = 0; m < 6; ++m) unv[m] = upd ? Un.f[m][c] : 0.0;
      double yd[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) yd[m] = lane_up(F2yu[m]);
      if (upd) {
        // row f-3 is the previous front's theta2 row (s2m1 is its U2..U4 slot)
        double uc[6];
        uc[0] = lds(LC_U01 + 2 * s4m3);
        uc[1] = lds(LC_U01 + 2 * s4m3 + 1);
        uc[2] = lds(LC_U234 + 3 * s2m1);
        uc[3] = lds(LC_U234 + 3 * s2m1 + 1);
        uc[4] = lds(LC_U234 + 3 * s2m1 + 2);
        uc[5] = lds(LC_U5 + s3f);  // (f-3) % 3 == f % 3
        const double ap = uc[5];
        double un[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          double res = (F2xp[m] - F2xm[m]) * inv_dx + (F2yu[m] - yd[m]) * inv_dy;
          if (m == 5)
            res = res - ap * (F2xp[6] - F2xm[6]) * inv_dx - ap * (F2yu[6] - yd[6]) * inv_dy;
          un[m] = alpha * unv[m] + beta * (uc[m] - dt * res);
        }
        if (repair) repair_cell(e, un);
        un[5] = clamp_al(un[5]);
#pragma unroll
        for (int m = 0; m < 6; ++m) Uo.f[m][c] = un[m];
        if (want_dt) local_dt = wmin(local_dt, cell_dt(e, un, G.dx, G.dy));
      }
    }
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      F2xm[m] = F2xp[m];
      F2yu[m] = F2yn[m];
    }
  }
  return local_dt;
}

__device__ __forceinline__ void wave_min_atomic(double v, double* acc) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = wmin(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & (MW - 1)) == 0)
    atomicMin((unsigned long long*)acc, (unsigned long long)__double_as_longlong(v));
}

// One wave per (strip, segment of `rows` rows), numbered strip-fastest and
// packed 4 to a block, so a strip's segment count need not be a multiple of
// 4: the host picks `rows` so that the waves fill the device's two-blocks-
// per-CU slots in as few rounds as it can.  (Strip-fastest measured faster
// than segment-fastest for this kernel, two waves per SIMD, whose loads wait
// on memory: at 4096x2048 on the mixed storage 0.518 against 0.565 ms per
// launch at 74 rows, 0.559 against 0.698 ms at 26 rows.)
template <bool F32>
__global__ void __launch_bounds__(256, 2)
k_lu_march(Geo G, Eos e, State Un, State U, Faces F, State Uo, double alpha, double beta,
           const double* __restrict__ dt_dev, double* dt_acc, int repair, int rows, int nseg,
           int nitems) {
  __shared__ double carry[LU_WAVES * LU_CARRY * MW];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / MW);  // wave-uniform
  // Workgroups go to the 8 XCDs round-robin by linear id.  Give each XCD a
  // contiguous run of blocks, i.e. of whole rows of neighbouring strips, so
  // that neighbouring strips, which share their halo columns' cache lines,
  // and neighbouring segments, which share warm-up rows, run on the same
  // XCD's L2.  The map is a bijection on [0, N); which block does which
  // strip and segment does not change any result.
  const int nblk = (int)gridDim.x;
  const int lin = (int)blockIdx.x;
  const int xcd = lin & 7, base = nblk >> 3, extra = nblk & 7;
  const int q = xcd * base + (xcd < extra ? xcd : extra) + (lin >> 3);
  const int item = q * LU_WAVES + wave;
  if (item >= nitems) return;  // wave-uniform
  const int nstrip = nitems / nseg;  // nitems = strips x nseg
  const int seg = item / nstrip;
  const int strip = item - seg * nstrip;
  double v = lu_wave<F32>(G, e, Un, U, F, Uo, alpha, beta, *dt_dev, repair, dt_acc != nullptr,
                          strip, seg, rows, carry + wave * (LU_CARRY * MW));
  if (dt_acc) wave_min_atomic(v, dt_acc);
}

// ========================================================================
// k_lu_lo -- the march of k_lu_march with the first-order fluxes formed in
// the march (FP64 face storage).  The face kernel then writes only the 14
// high-order planes: the first-order flux of a face depends on its two cells
// alone (fx_low_kernel SC:1303-1347, gy_low_kernel SC:1350-1394), and the
// march streams those cells' state anyway.  Per launch that removes 112 B per
// cell of face writes and 112 B per cell of face reads, and the first-order
// HLLC moves out of the face kernel, which is bound by VALU issue, into this
// one, which runs at one wave per SIMD and is bound by that wave's
// instruction issue.
//
// One wave owns 64 consecutive j: 56 output columns and a 4-lane halo each
// side.  The y-face above lane l is formed from the cells of lanes l and l+1
// (a lane shift), so lane 63's is not formed; with the output columns at
// lanes 4..59 nothing an output depends on reaches lanes 0 or 63 (update(l)
// <- theta2(l+-1) <- theta1(l+-2) <- theta0(l+-3) <- y-faces of lanes
// l-4 .. l+2).
//
// At front f the wave takes row f+1 of the stage input U (at the boundary-
// mapped cell, slip signs applied, as the face kernel stages it), forms its
// first-order primitives (low_prims, the face kernel's function) and from
// them and row f's the first-order fluxes of x-face row f and of the y-faces
// of row f.  Then, as k_lu_march: theta0(f), theta1(f-1), theta2(f-2), the
// final fluxes and the update of row f-3.  Carried between fronts:
//   registers  row f's primitives; first-order x-face rows f-3..f (all
//              components); high-order x-face rows as in k_lu_march;
//   LDS ring   first-order y-faces rows f-2..f, high-order y-face {0,1}
//              rows f-2..f and {5,6} rows f-2..f-1, U rows f-3..f+1.
// The wave runs alone on its SIMD (one 4-wave block per CU: the ring takes
// 31 KB of LDS per wave and the carries most of the 512 registers), so the
// next front's global loads are issued before the current front's
// arithmetic (fronts in pairs over two load buffers), at in-bounds addresses
// (rows and columns clamped).  Loaded and formed values of faces outside the
// domain are used as they come: they feed only thetas of cells outside the
// interior (which stay 1) and fluxes that no output row or column reads.
// ========================================================================
constexpr int LOH = 4;                // halo lanes each side
constexpr int LOJ = MW - 2 * LOH;     // output columns per wave (56)
constexpr int LO_YH01 = 0;            // y-face H0 H1, 3 row slots
constexpr int LO_YH56 = 6;            // y-face H5 H6, 2 row slots
constexpr int LO_YL = 10;             // y-face first-order L0..L6, 3 row slots
constexpr int LO_U = 31;              // U0..U5, 5 row slots
constexpr int LO_CARRY = 61;
constexpr int LO_WAVES = 4;

// One front's global loads.
struct LoLoads {
  double q[6];        // U row f+1 at its boundary-mapped cell
  double sgu, sgv;    // slip signs of that cell
  double xh01[2];     // x-face H0 H1 row f
  double xh56[2];     // x-face H5 H6 row f-1
  double xh24[3];     // x-face H2 H3 H4 row f-2
  double yh01[2];     // y-face H0 H1 row f
  double yh56[2];     // y-face H5 H6 row f-1
  double yh24[3];     // y-face H2 H3 H4 row f-2
  double un[6];       // Un row f-3
};

__device__ __forceinline__ void lo_load(const Geo& G, const State& Un, const State& U,
                                        const Faces& F, int j, int jc, int f, LoLoads& L) {
  const size_t ld = (size_t)G.ld;
  const int last = G.nx + 2 * G.g - 1;
  auto at = [&](int r) { return (size_t)min(max(r, 0), last) * ld + (size_t)jc; };
  double sru, srv;
  const size_t cq = map_cell(G, f + 1, j, sru, srv);
#pragma unroll
  for (int m = 0; m < 6; ++m) L.q[m] = U.f[m][cq];
  L.sgu = sru;
  L.sgv = srv;
  const size_t c0 = at(f), c1 = at(f - 1), c2 = at(f - 2), c3 = at(f - 3);
  L.xh01[0] = F[0][c0];
  L.xh01[1] = F[1][c0];
  L.xh56[0] = F[5][c1];
  L.xh56[1] = F[6][c1];
  L.yh01[0] = F[7][c0];
  L.yh01[1] = F[8][c0];
  L.yh56[0] = F[12][c1];
  L.yh56[1] = F[13][c1];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    L.xh24[k] = F[2 + k][c2];
    L.yh24[k] = F[9 + k][c2];
  }
#pragma unroll
  for (int m = 0; m < 6; ++m) L.un[m] = Un.f[m][c3];
}

// First-order primitives of a staged cell (the conserved values of its
// boundary-mapped cell, slip signs sgu / sgv) -- pressure recovery as the face
// kernel's staging (tile_stage_cell), then low_prims -- with the sound speed
// and total energy that hllc forms from them.  Every face of the cell reuses
// these: the march forms them once per cell instead of once per face side.
__device__ __forceinline__ void lo_cell(const Eos& e, const double* q, double sgu, double sgv,
                                        double* P, double& c, double& E) {
  const double ru = sgu < 0.0 ? -1.0 * q[2] : q[2];
  const double rv = sgv < 0.0 ? -1.0 * q[3] : q[3];
  double uu, vv;
  const double p = pressure(e, q[0], q[1], ru, rv, q[4], q[5], uu, vv);
  low_prims(e, q[0], q[1], uu, vv, p, q[5], P);
  c = sound_speed(e, P[0] + P[1], P[4], P[5]);
  E = prim_energy(e, P[0], P[1], P[2], P[3], P[4], P[5]);
}

__device__ __forceinline__ double lu_wave_lo(const Geo& G, const Eos& e, const State& Un,
                                             const State& U, const Faces& F, const State& Uo,
                                             double alpha, double beta, double dt, int repair,
                                             bool want_dt, int strip, int seg, int rows,
                                             double* cw) {
  const int lane = threadIdx.x & (MW - 1);
  const int j = G.g + strip * LOJ - LOH + lane;  // >= g - 4 >= 0
  const int i0 = G.g + seg * rows;
  if (i0 >= G.g + G.nx) return 1.0e30;  // wave-uniform
  const int i_end = min(i0 + rows, G.g + G.nx);
  const double inv_dx = 1.0 / G.dx;
  const double inv_dy = 1.0 / G.dy;
  const int jc = min(j, G.ld - 1);
  const bool col_in = j >= G.g && j < G.g + G.ny;
  const bool yface_in = j >= G.g - 1 && j < G.g + G.ny;
  const bool out_lane = lane >= LOH && lane < LOH + LOJ && col_in;
  auto row_in = [&](int r) { return r >= G.g && r < G.g + G.nx; };
  auto xrow_in = [&](int r) { return r >= G.g - 1 && r < G.g + G.nx; };
  auto lds = [&](int q) -> double& { return cw[q * MW + lane]; };

#pragma unroll
  for (int q = 0; q < LO_CARRY; ++q) lds(q) = 0.0;

  double P0[6];               // first-order primitives of row f
  double Lx[4][7];            // first-order x-face fluxes rows f-3 .. f
  double xh01[3][2] = {};     // high-order x-face H0 H1 rows f-2 .. f
  double xh56[2][2] = {};     // H5 H6 rows f-2, f-1
  double xa[2][7] = {};       // H0..H6 rows f-3, f-2
  double F2xm[7] = {};        // final flux at x-face row f-4 (below the update row)
  // A face's blended flux is the same for the cells on either side of it:
  // the x-face below theta1's / theta2's row is the face above the previous
  // front's row (the same face thetas), and a y-face below a lane is the face
  // above the lane before (a lane shift).  Those are carried or shifted, not
  // blended twice.
  double B1x[2] = {};         // x-face row f-2, H5 H6 blended once
  double B2x[7] = {};         // x-face row f-3, H0..H6 blended twice
  double F2yu[7] = {};        // final flux at the y-faces above row f-3
  double t0[5] = {1.0, 1.0, 1.0, 1.0, 1.0};  // theta0 rows f-4 .. f
  double t1[4] = {1.0, 1.0, 1.0, 1.0};       // theta1 rows f-4 .. f-1
  double t2[3] = {1.0, 1.0, 1.0};            // theta2 rows f-4 .. f-2
  double local_dt = 1.0e30;

  // Before the first front (f = i0-3): row i0-3's primitives and U (theta0 of
  // row i0-3 reads its U0, U1 from the ring), and x-face row i0-4 below it --
  // theta0(i0-3) is in the first output row's dependency cone.
  double c0, E0;  // row f's sound speed and total energy
  {
    double q[6], Pm[6], cm, Em;
    LoLoads a;
    lo_load(G, Un, U, F, j, jc, i0 - 5, a);  // U row i0-4 (faces unused)
    lo_cell(e, a.q, a.sgu, a.sgv, Pm, cm, Em);
    double sru, srv;
    const size_t cq = map_cell(G, i0 - 3, j, sru, srv);
#pragma unroll
    for (int m = 0; m < 6; ++m) q[m] = U.f[m][cq];
    lo_cell(e, q, sru, srv, P0, c0, E0);
#pragma unroll
    for (int m = 0; m < 6; ++m) lds(LO_U + 6 * ((i0 - 3) % 5) + m) = q[m];
    double lx[7];
    hllc_ce<0>(e, Pm, P0, cm, c0, Em, E0, lx, lx[6]);
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      Lx[0][m] = 0.0;
      Lx[1][m] = 0.0;
      Lx[2][m] = 0.0;
      Lx[3][m] = lx[m];
    }
    const size_t c = (size_t)(i0 - 4) * (size_t)G.ld + (size_t)jc;  // i0 - 4 >= 0
    xh01[2][0] = F[0][c];
    xh01[2][1] = F[1][c];
  }

  // ring slots of rows f (U ring: row mod 5; y rings: row mod 3), advanced per
  // front instead of divided (f >= i0 - 3 >= 1)
  int s5f = (i0 - 3) % 5, s3f = (i0 - 3) % 3;
  // One front; `cur` holds its global loads (issued one front earlier).
  auto front = [&](const int f, const LoLoads& cur) __attribute__((always_inline)) {
    const int s5p1 = s5f == 4 ? 0 : s5f + 1, s5m1 = s5f == 0 ? 4 : s5f - 1;
    const int s5m2 = s5m1 == 0 ? 4 : s5m1 - 1, s5m3 = s5m2 == 0 ? 4 : s5m2 - 1;
    const int s3m1 = s3f == 0 ? 2 : s3f - 1, s3m2 = s3m1 == 0 ? 2 : s3m1 - 1;
    const int s2m1 = (f + 1) & 1, s2m2 = f & 1;

    // ---- row f+1's primitives; first-order fluxes of x-face row f and of
    //      the y-faces of row f
    double ly[7];
    {
      double P1[6], lx[7], Pd[6], c1, E1;
      lo_cell(e, cur.q, cur.sgu, cur.sgv, P1, c1, E1);
#pragma unroll
      for (int m = 0; m < 6; ++m) lds(LO_U + 6 * s5p1 + m) = cur.q[m];
      hllc_ce<0>(e, P0, P1, c0, c1, E0, E1, lx, lx[6]);
#pragma unroll
      for (int m = 0; m < 6; ++m) Pd[m] = lane_down(P0[m]);
      hllc_ce<1>(e, P0, Pd, c0, lane_down(c0), E0, lane_down(E0), ly, ly[6]);
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        Lx[0][m] = Lx[1][m];
        Lx[1][m] = Lx[2][m];
        Lx[2][m] = Lx[3][m];
        Lx[3][m] = lx[m];
        lds(LO_YL + 7 * s3f + m) = ly[m];
      }
#pragma unroll
      for (int m = 0; m < 6; ++m) P0[m] = P1[m];
      c0 = c1;
      E0 = E1;
    }

    // ---- high-order x-faces of this front
    {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        xh01[0][q] = xh01[1][q];
        xh01[1][q] = xh01[2][q];
        xh01[2][q] = cur.xh01[q];
        xh56[0][q] = xh56[1][q];
        xh56[1][q] = cur.xh56[q];
      }
#pragma unroll
      for (int m = 0; m < 7; ++m) xa[0][m] = xa[1][m];
      xa[1][0] = xh01[0][0];
      xa[1][1] = xh01[0][1];
      xa[1][5] = xh56[0][0];
      xa[1][6] = xh56[0][1];
#pragma unroll
      for (int k = 0; k < 3; ++k) xa[1][2 + k] = cur.xh24[k];
    }

    // ---------------- theta0 at row f (raw high-order fluxes) ----------------
    {
      const bool ok = row_in(f) && col_in;
      double yc[2], ycd[2], lyd[2];
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        yc[m] = cur.yh01[m];
        ycd[m] = lane_up(yc[m]);
        lyd[m] = lane_up(ly[m]);
      }
      double th = 1.0;
      if (ok) {
        double l[6], h[6];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          const double u = lds(LO_U + 6 * s5f + m);
          double rh = (xh01[2][m] - xh01[1][m]) * inv_dx + (yc[m] - ycd[m]) * inv_dy;
          double rl = (Lx[3][m] - Lx[2][m]) * inv_dx + (ly[m] - lyd[m]) * inv_dy;
          l[m] = u - dt * rl;
          h[m] = u - dt * rh;
        }
        th = theta_stage<0>(e, l, h);
      }
      lds(LO_YH01 + 2 * s3f) = yc[0];
      lds(LO_YH01 + 2 * s3f + 1) = yc[1];
#pragma unroll
      for (int k = 0; k < 4; ++k) t0[k] = t0[k + 1];
      t0[4] = th;
    }

    // ---------------- theta1 at row f-1 (fluxes blended once) ----------------
    {
      const int r = f - 1;
      const bool ok = row_in(r) && col_in;
      double tf[4];
      tf[0] = face_theta(t0[2], t0[3]);
      tf[1] = face_theta(t0[3], t0[4]);
      tf[2] = face_theta(lane_up(t0[3]), t0[3]);
      tf[3] = face_theta(t0[3], lane_down(t0[3]));
      double yu[2], yul[2], ydl[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        yu[q] = cur.yh56[q];
        yul[q] = lds(LO_YL + 7 * s3m1 + 5 + q);
        ydl[q] = lane_up(yul[q]);
      }
      // the faces above and to the side of the row, blended once on every lane
      const double Hxp5 = blend3(xh56[1][0], Lx[2][5], &tf[1], 1);
      const double Hxp6 = blend3(xh56[1][1], Lx[2][6], &tf[1], 1);
      const double Hyp5 = blend3(yu[0], yul[0], &tf[3], 1);
      const double Hyp6 = blend3(yu[1], yul[1], &tf[3], 1);
      const double Hxm5 = B1x[0], Hxm6 = B1x[1];
      const double Hym5 = lane_up(Hyp5), Hym6 = lane_up(Hyp6);
      B1x[0] = Hxp5;
      B1x[1] = Hxp6;
      double th1 = 1.0;
      if (ok) {
        const double al = lds(LO_U + 6 * s5m1 + 5);
        double rh = (Hxp5 - Hxm5) * inv_dx + (Hyp5 - Hym5) * inv_dy;
        double rl = (Lx[2][5] - Lx[1][5]) * inv_dx + (yul[0] - ydl[0]) * inv_dy;
        rh = rh - al * (Hxp6 - Hxm6) * inv_dx - al * (Hyp6 - Hym6) * inv_dy;
        rl = rl - al * (Lx[2][6] - Lx[1][6]) * inv_dx - al * (yul[1] - ydl[1]) * inv_dy;
        double l[6], h[6];
        l[5] = al - dt * rl;
        h[5] = al - dt * rh;
        th1 = theta_stage<1>(e, l, h);
      }
      lds(LO_YH56 + 2 * s2m1) = yu[0];
      lds(LO_YH56 + 2 * s2m1 + 1) = yu[1];
#pragma unroll
      for (int k = 0; k < 3; ++k) t1[k] = t1[k + 1];
      t1[3] = th1;
    }

    // ---------------- theta2 at row f-2 (fluxes blended twice) ----------------
    double yn[14];  // y-faces above row f-2 (H0..6, L0..6)
    double hm[7], hu[7];  // x-face row f-3, y-faces above row f-2: blended twice
    {
      const int r = f - 2;
      const bool ok = row_in(r) && col_in;
      yn[0] = lds(LO_YH01 + 2 * s3m2);
      yn[1] = lds(LO_YH01 + 2 * s3m2 + 1);
      yn[5] = lds(LO_YH56 + 2 * s2m2);
      yn[6] = lds(LO_YH56 + 2 * s2m2 + 1);
#pragma unroll
      for (int k = 0; k < 3; ++k) yn[2 + k] = cur.yh24[k];
#pragma unroll
      for (int m = 0; m < 7; ++m) yn[7 + m] = lds(LO_YL + 7 * s3m2 + m);
      double tf[4][2];
      tf[0][0] = face_theta(t0[1], t0[2]);
      tf[1][0] = face_theta(t0[2], t0[3]);
      tf[2][0] = face_theta(lane_up(t0[2]), t0[2]);
      tf[3][0] = face_theta(t0[2], lane_down(t0[2]));
      tf[0][1] = face_theta(t1[1], t1[2]);
      tf[1][1] = face_theta(t1[2], t1[3]);
      tf[2][1] = face_theta(lane_up(t1[2]), t1[2]);
      tf[3][1] = face_theta(t1[2], lane_down(t1[2]));
      // the faces above and to the side of the row, blended twice on every lane
      double hp[7], hd[7], ydl[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        hp[m] = blend3(xa[1][m], Lx[1][m], tf[1], 2);
        hu[m] = blend3(yn[m], yn[7 + m], tf[3], 2);
        hm[m] = B2x[m];
        hd[m] = lane_up(hu[m]);
        ydl[m] = lane_up(yn[7 + m]);
        B2x[m] = hp[m];
      }
      double th2 = 1.0;
      if (ok) {
        double uc[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) uc[m] = lds(LO_U + 6 * s5m2 + m);
        const double al = uc[5];
        double l[6], h[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          double rh = (hp[m] - hm[m]) * inv_dx + (hu[m] - hd[m]) * inv_dy;
          double rl = (Lx[1][m] - Lx[0][m]) * inv_dx + (yn[7 + m] - ydl[m]) * inv_dy;
          if (m == 5) {
            rh = rh - al * (hp[6] - hm[6]) * inv_dx - al * (hu[6] - hd[6]) * inv_dy;
            rl = rl - al * (Lx[1][6] - Lx[0][6]) * inv_dx - al * (yn[13] - ydl[6]) * inv_dy;
          }
          const double u = uc[m];
          l[m] = u - dt * rl;
          h[m] = u - dt * rh;
        }
        th2 = theta_stage<2>(e, l, h);
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) t2[k] = t2[k + 1];
      t2[2] = th2;
    }

    // ---- final fluxes (three blends): x-face row f-3, y-faces above row f-2.
    //      Their first two blends are theta2's (the same face thetas), so one
    //      more blend with the stage-2 face theta finishes them.
    double F2xp[7], F2yn[7];
    {
      const double tx = face_theta(t2[1], t2[2]);             // rows f-3, f-2
      const double ty = face_theta(t2[2], lane_down(t2[2]));  // row f-2 and the lane after
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        F2xp[m] = blend3(hm[m], Lx[0][m], &tx, 1);
        F2yn[m] = blend3(hu[m], yn[7 + m], &ty, 1);
      }
    }

    // ---------------- update of row f-3 ----------------
    {
      const int r = f - 3;
      const bool upd = r >= i0 && r < i_end && out_lane;
      double yd[7];
#pragma unroll
      for (int m = 0; m < 7; ++m) yd[m] = lane_up(F2yu[m]);
      if (upd) {
        const size_t c = (size_t)r * (size_t)G.ld + (size_t)j;
        double uc[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) uc[m] = lds(LO_U + 6 * s5m3 + m);
        const double ap = uc[5];
        double un[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
          double res = (F2xp[m] - F2xm[m]) * inv_dx + (F2yu[m] - yd[m]) * inv_dy;
          if (m == 5)
            res = res - ap * (F2xp[6] - F2xm[6]) * inv_dx - ap * (F2yu[6] - yd[6]) * inv_dy;
          un[m] = alpha * cur.un[m] + beta * (uc[m] - dt * res);
        }
        if (repair) repair_cell(e, un);
        un[5] = clamp_al(un[5]);
#pragma unroll
        for (int m = 0; m < 6; ++m) Uo.f[m][c] = un[m];
        if (want_dt) local_dt = wmin(local_dt, cell_dt(e, un, G.dx, G.dy));
      }
    }
#pragma unroll
    for (int m = 0; m < 7; ++m) {
      F2xm[m] = F2xp[m];
      F2yu[m] = F2yn[m];
    }
    s5f = s5p1;
    s3f = s3f == 2 ? 0 : s3f + 1;
  };

  // Fronts i0-3 .. i_end+2, run in pairs over two load buffers that swap
  // roles, so a front's loads are in flight during the previous front's
  // arithmetic and no register copy waits for them.  An odd count gets one
  // more front after the last: it updates no row (row i_end is past the
  // segment) and its loads stay in bounds.
  LoLoads A, B;
  lo_load(G, Un, U, F, j, jc, i0 - 3, A);
  const int f_end = i0 - 3 + (((i_end - i0 + 6) + 1) & ~1);
  for (int f = i0 - 3; f < f_end; f += 2) {
    lo_load(G, Un, U, F, j, jc, f + 1, B);
    front(f, A);
    lo_load(G, Un, U, F, j, jc, f + 2, A);
    front(f + 1, B);
  }
  return local_dt;
}

// One wave per (strip, segment of `rows` rows), numbered segment-fastest and
// packed LO_WAVES to a block, so a strip's segment count need not be a
// multiple of LO_WAVES: the host picks `rows` so that the waves fill the
// device's one-block-per-CU slots in as few rounds as it can.  (Segment-
// fastest measured faster than strip-fastest for this kernel: at 4096x2048,
// 152 rows, 0.692 against 0.722 ms per launch.)
__global__ void __launch_bounds__(64 * LO_WAVES, 1)
k_lu_lo(Geo G, Eos e, State Un, State U, Faces F, State Uo, double alpha, double beta,
        const double* __restrict__ dt_dev, double* dt_acc, int repair, int rows, int nseg,
        int nitems) {
  __shared__ double carry[LO_WAVES * LO_CARRY * MW];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / MW);  // wave-uniform
  // the XCD-aware block map of k_lu_march (a bijection on [0, N)): each XCD
  // takes a contiguous run of blocks, i.e. of whole neighbouring strips
  const int nblk = (int)gridDim.x;
  const int lin = (int)blockIdx.x;
  const int xcd = lin & 7, base = nblk >> 3, extra = nblk & 7;
  const int q = xcd * base + (xcd < extra ? xcd : extra) + (lin >> 3);
  const int item = q * LO_WAVES + wave;
  if (item >= nitems) return;  // wave-uniform
  const int strip = item / nseg;
  const int seg = item - strip * nseg;
  double v = lu_wave_lo(G, e, Un, U, F, Uo, alpha, beta, *dt_dev, repair, dt_acc != nullptr, strip,
                        seg, rows, carry + wave * (LO_CARRY * MW));
  if (dt_acc) wave_min_atomic(v, dt_acc);
}

// ========================================================================
// arithmetic self-test: the FP64 kernels' division and square-root
// sequences beside the compiler's IEEE operations, element by element
// ========================================================================
__global__ void __launch_bounds__(256) k_selftest_arith(const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  out[i] = ddiv(x, y);
  out[n + i] = x / y;
  out[2 * n + i] = sq(y);
  out[3 * n + i] = sqrt(y);
}

// the FP64 kernels' pow and tanh (wamp_math.h) element by element
__global__ void __launch_bounds__(256) k_selftest_elementary(const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             double* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = wm_pow(x[i], y[i]);
  out[n + i] = th(x[i]);
}

// ========================================================================
// small kernels
// ========================================================================
constexpr int STHREADS = 256;

__global__ void __launch_bounds__(STHREADS) k_dt(Geo G, Eos e, State U, double* dt_acc) {
  const int n = G.nx * G.ny;
  const int k = blockIdx.x * STHREADS + threadIdx.x;
  double v = 1.0e30;
  if (k < n) {
    const size_t c = (size_t)(G.g + k / G.ny) * (size_t)G.ld + (size_t)(G.g + k % G.ny);
    double u[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) u[m] = U.f[m][c];
    v = cell_dt(e, u, G.dx, G.dy);
  }
  // same block reduction as the fused epilogue
  __shared__ double wmins[STHREADS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = wmin(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) wmins[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmins[0];
    for (int w = 1; w < STHREADS / 64; ++w) m = wmin(m, wmins[w]);
    atomicMin((unsigned long long*)dt_acc, (unsigned long long)__double_as_longlong(m));
  }
}

// dt = cfl * min (get_dt SC:1995); reset the accumulator to 1e30 (SC:1987).
__global__ void k_dt_finalize(double* dt_acc, double* dt_out, double cfl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double m = *dt_acc;
    *dt_out = cfl * m;
    *dt_acc = 1.0e30;
  }
}

__global__ void k_fill(double* p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Interior repair + clamp (the initial sequence SC:1998-2001).
__global__ void __launch_bounds__(STHREADS) k_prepare(Geo G, Eos e, State U, int repair) {
  const int n = G.nx * G.ny;
  const int k = blockIdx.x * STHREADS + threadIdx.x;
  if (k >= n) return;
  const size_t c = (size_t)(G.g + k / G.ny) * (size_t)G.ld + (size_t)(G.g + k % G.ny);
  double u[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) u[m] = U.f[m][c];
  if (repair) repair_cell(e, u);
  u[5] = clamp_al(u[5]);
#pragma unroll
  for (int m = 0; m < 6; ++m) U.f[m][c] = u[m];
}

// enforce_admissible_kernel over the padded grid (SC:918-958).
__global__ void __launch_bounds__(STHREADS) k_enforce(Geo G, Eos e, State U) {
  const int nyg = G.ny + 2 * G.g, nxg = G.nx + 2 * G.g;
  const int k = blockIdx.x * STHREADS + threadIdx.x;
  if (k >= nxg * nyg) return;
  const size_t c = (size_t)(k / nyg) * (size_t)G.ld + (size_t)(k % nyg);
  double u[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) u[m] = U.f[m][c];
  repair_cell(e, u);
#pragma unroll
  for (int m = 0; m < 6; ++m) U.f[m][c] = u[m];
}

// cons_to_prim_kernel over the padded grid (SC:840-864).
__global__ void __launch_bounds__(STHREADS) k_cons_to_prim(Geo G, Eos e, State U, double* up,
                                                           double* vp, double* pp) {
  const int nyg = G.ny + 2 * G.g, nxg = G.nx + 2 * G.g;
  const int k = blockIdx.x * STHREADS + threadIdx.x;
  if (k >= nxg * nyg) return;
  const size_t c = (size_t)(k / nyg) * (size_t)G.ld + (size_t)(k % nyg);
  double a1 = clamp_al(U.f[5][c]);
  U.f[5][c] = a1;
  double uu, vv;
  double p = pressure(e, U.f[0][c], U.f[1][c], U.f[2][c], U.f[3][c], U.f[4][c], a1, uu, vv);
  up[c] = uu;
  vp[c] = vv;
  pp[c] = p;
}

// bc_x then bc_y over all ghost cells (SC:1616-1682); each ghost cell reads
// its mapped interior cell, which no thread writes.
__global__ void __launch_bounds__(STHREADS) k_bc(Geo G, State U, double* up, double* vp,
                                                 double* pp) {
  const int nyg = G.ny + 2 * G.g, nxg = G.nx + 2 * G.g;
  const int k = blockIdx.x * STHREADS + threadIdx.x;
  if (k >= nxg * nyg) return;
  const int i = k / nyg, j = k % nyg;
  if (i >= G.g && i < G.g + G.nx && j >= G.g && j < G.g + G.ny) return;
  double sru, srv;
  const size_t src = map_cell(G, i, j, sru, srv);
  const size_t c = (size_t)i * (size_t)G.ld + (size_t)j;
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    double v = U.f[m][src];
    if (m == 2 && sru < 0.0) v = -1.0 * v;
    if (m == 3 && srv < 0.0) v = -1.0 * v;
    U.f[m][c] = v;
  }
  if (up) {
    double uu = up[src], vv = vp[src];
    if (sru < 0.0) uu = -1.0 * uu;
    if (srv < 0.0) vv = -1.0 * vv;
    up[c] = uu;
    vp[c] = vv;
    pp[c] = pp[src];
  }
}

// ========================================================================
// Batched mode: many independent cases (any grid and EOS, one scheme and
// precision) advanced by one persistent launch per kernel kind.  The work
// list is the concatenation of every active case's face tiles (faces) or
// wave items (limiter); a block/wave finds its case by binary search in the
// prefix offsets.  Per-case dt, time, end time and a done flag live on the
// device, so a captured step graph needs no host interaction.
// ========================================================================

// One wave per work item (case, strip, segment of `rows` rows) of the
// concatenated list.
template <bool F32>
__global__ void __launch_bounds__(256, 2)
k_lu_batch(const CaseDesc* __restrict__ cs, const int* __restrict__ item_off, int ncase,
           BatchScal sc, int stage, int repair, int rows) {
  // the wave index is uniform across the wave: keep it (and everything derived
  // from it: case, descriptor, rows) in scalar registers
  __shared__ double carry[LU_WAVES * LU_CARRY * MW];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / MW);
  const int total = item_off[ncase];
  const double alpha = stage == 0 ? 0.0 : stage == 1 ? 0.75 : 1.0 / 3.0;
  const double beta = stage == 0 ? 1.0 : stage == 1 ? 0.25 : 2.0 / 3.0;
  // items wave, wave + waves in the grid, ...: one item per wave when the
  // grid has a wave per item, a persistent stride otherwise
  const int stride = (int)gridDim.x * (blockDim.x / MW);
  for (int it = blockIdx.x * (blockDim.x / MW) + wave; it < total; it += stride) {
    const int k = find_case(item_off, ncase, it);
    if (sc.done[k]) continue;
    const CaseDesc& c = cs[k];
    const int li = it - item_off[k];
    State Un, U, Uo;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      Un.f[m] = c.u[0][m];
      U.f[m] = c.u[stage][m];
      Uo.f[m] = c.u[stage == 2 ? 0 : stage + 1][m];
    }
    double v = lu_wave<F32>(c.geo, c.eos, Un, U, c.f, Uo, alpha, beta, sc.dt[k], repair,
                            stage == 2, li % c.strips, li / c.strips, rows,
                            carry + wave * (LU_CARRY * MW));
    if (stage == 2) wave_min_atomic(v, sc.acc + k);
  }
}

// After a step: t += dt; next dt = cfl * min, clipped to the end time
// (SC:2025-2026); a case that reached its end time is frozen.
__global__ void k_batch_finalize(const CaseDesc* __restrict__ cs, int ncase, BatchScal sc) {
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

// Initial dt of one slot from get_dt's accumulator (clipped to the end time).
__global__ void k_batch_dt0(double* acc, BatchScal sc, int k, double cfl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double dt = cfl * *acc;
    *acc = 1.0e30;
    const double te = sc.tend[k];
    if (te > 0.0 && sc.t[k] + dt > te) dt = te - sc.t[k];
    sc.dt[k] = dt;
    sc.acc[k] = 1.0e30;
    sc.done[k] = (te > 0.0 && sc.t[k] >= te - 1.0e-14) ? 1 : 0;
  }
}

// ========================================================================
// initial conditions on the device
// ========================================================================
// Conserved state from primitives with two phase densities (the stiffened-gas
// energy gb p + pb + kin); divisions by g-1 are IEEE divides as in the host
// restatement of the same formula.
__device__ __forceinline__ void ic_store(const Geo& G, const Eos& e, const State& U, size_t c,
                                         double a1, double rho1, double rho2, double u, double v,
                                         double p) {
  const double ar1 = a1 * rho1;
  const double ar2 = (1.0 - a1) * rho2;
  const double rho = ar1 + ar2;
  const double a2 = 1.0 - a1;
  const double gb = a1 / (e.g1 - 1.0) + a2 / (e.g2 - 1.0);
  const double pb = a1 * e.g1 * e.p1 / (e.g1 - 1.0) + a2 * e.g2 * e.p2 / (e.g2 - 1.0);
  const double E = gb * p + pb + 0.5 * rho * (u * u + v * v);
  U.f[0][c] = ar1;
  U.f[1][c] = ar2;
  U.f[2][c] = rho * u;
  U.f[3][c] = rho * v;
  U.f[4][c] = E;
  U.f[5][c] = a1;
}

// One padded cell per thread over (nx+2g) x ld; cells outside the interior are zero.
__global__ void __launch_bounds__(256) k_ic_droplets(Geo G, Eos e, State U, DropletIC ic) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = (long long)(G.nx + 2 * G.g) * G.ld;
  if (k >= n) return;
  const int i = (int)(k / G.ld), j = (int)(k % G.ld);
  const size_t c = (size_t)k;
  if (i < G.g || i >= G.g + G.nx || j < G.g || j >= G.g + G.ny) {
#pragma unroll
    for (int m = 0; m < 6; ++m) U.f[m][c] = 0.0;
    return;
  }
  const double x = ic.xmin + ((double)(i - G.g) + 0.5) * G.dx;
  const double y = ic.ymin + ((double)(j - G.g) + 0.5) * G.dy;
  double water = 0.0;
  for (int q = 0; q < ic.n_drops; ++q) {
    const double xc = ic.drops[4 * q], yc = ic.drops[4 * q + 1];
    const double r = ic.drops[4 * q + 2], rc = ic.drops[4 * q + 3];
    const double dx = x - xc, dy = y - yc;
    const double d = sqrt(dx * dx + dy * dy);
    double a = 0.5 * (1.0 - tanh((d - r) / ic.width));
    if (rc > 0.0) a = a * (1.0 - 0.5 * (1.0 - tanh((d - rc) / ic.width)));
    water = water > a ? water : a;  // maximum (a is never NaN)
  }
  double a1 = water;
  if (a1 < 1.0e-6) a1 = 1.0e-6;
  if (a1 > 1.0 - 1.0e-6) a1 = 1.0 - 1.0e-6;
  const bool post = x < ic.shock_x;
  ic_store(G, e, U, c, a1, ic.rho_water, post ? ic.rho_post : ic.rho_air,
           post ? ic.u_post : 0.0, 0.0, post ? ic.p_post : ic.p0);
}

__global__ void __launch_bounds__(256) k_ic_triple(Geo G, Eos e, State U, TripleIC ic) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long n = (long long)(G.nx + 2 * G.g) * G.ld;
  if (k >= n) return;
  const int i = (int)(k / G.ld), j = (int)(k % G.ld);
  const size_t c = (size_t)k;
  if (i < G.g || i >= G.g + G.nx || j < G.g || j >= G.g + G.ny) {
#pragma unroll
    for (int m = 0; m < 6; ++m) U.f[m][c] = 0.0;
    return;
  }
  const double x = ic.xmin + ((double)(i - G.g) + 0.5) * G.dx;
  const double y = ic.ymin + ((double)(j - G.g) + 0.5) * G.dy;
  const bool left = x < ic.x_split;
  const bool top = !left && y >= ic.y_split;
  const double a1 = (left || top) ? 1.0 - 1.0e-6 : 1.0e-6;
  const double rho = left ? ic.rho_left : (top ? ic.rho_top : ic.rho_bottom);
  const double p = left ? ic.p_left : (top ? ic.p_top : ic.p_bottom);
  ic_store(G, e, U, c, a1, rho, rho, 0.0, 0.0, p);
}

// ========================================================================
// launch helpers (called by the C-ABI layer)
// ========================================================================
template <int BASE, int SHEAR, int FACEPP, typename T>
static void launch_faces_t(const Geo& G, const Eos& e, const State& U, const Faces& F,
                           int vf_always, double t1, int low, hipStream_t st) {
  // 16x16 tiles stage 2.07 cells per face position instead of 8x32's 2.29, and
  // every FP64 scheme gains from that (4096x2048 per launch: M10 1.008 instead
  // of 1.025 ms, SC 0.804 instead of 0.821, MP5 1.098 instead of 1.124).  The
  // FP32-mixed kernels (wamp_faces_mixed.hip) measured faster with 8x32.
  static_assert(sizeof(T) == 8, "the FP64 translation unit launches FP64 face kernels only");
  using TL = Tile16x16;
  dim3 grid((G.ny + 1 + TL::TJ - 1) / TL::TJ, (G.nx + 1 + TL::TI - 1) / TL::TI);
  hipLaunchKernelGGL((k_faces<TL, BASE, SHEAR, FACEPP, T>), grid, dim3(FTHREADS), 0, st, G, e, U,
                     F, vf_always, t1, low);
}

}  // namespace wamp

// Exposed to wamp_capi.cpp
extern "C" {
using namespace wamp;

// FP32-mixed face kernels (wamp_faces_mixed.hip)
hipError_t wamp_launch_faces_mixed(const Geo* G, const Eos* e, const State* U, const Faces* F,
                                   int base, int shear, int face_pp, int vf_always, double t1,
                                   hipStream_t st);
hipError_t wamp_launch_faces_batch_mixed(const CaseDesc* cs, const int* tile_off, int ncase,
                                         const BatchScal* sc, int stage, int base, int shear,
                                         int face_pp, int vf_always, double t1, int tile,
                                         int blocks, hipStream_t st);

// low = 0: the first-order planes are not written (the limiter march that
// forms the first-order fluxes itself reads only the 14 high-order planes).
hipError_t wamp_launch_faces(const Geo* G, const Eos* e, const State* U, const Faces* F,
                             int base, int shear, int face_pp, int vf_always, double t1,
                             int precision, int low, hipStream_t st) {
  if (precision != 0)
    return wamp_launch_faces_mixed(G, e, U, F, base, shear, face_pp, vf_always, t1, st);
#define WAMP_F(B, S, P)                                                         \
  if (base == B && shear == S && face_pp == P) {                               \
    launch_faces_t<B, S, P, double>(*G, *e, *U, *F, vf_always, t1, low, st);   \
    return hipGetLastError();                                                  \
  }
  WAMP_F(0, 0, 0) WAMP_F(0, 0, 1) WAMP_F(0, 1, 0) WAMP_F(0, 1, 1) WAMP_F(0, 2, 0) WAMP_F(0, 2, 1)
  WAMP_F(1, 0, 0) WAMP_F(1, 0, 1) WAMP_F(1, 1, 0) WAMP_F(1, 1, 1) WAMP_F(1, 2, 0) WAMP_F(1, 2, 1)
#undef WAMP_F
  return hipErrorInvalidValue;
}

hipError_t wamp_launch_limit_update(const Geo* G, const Eos* e, const State* Un, const State* U,
                                    const Faces* F, const State* Uo, double alpha, double beta,
                                    const double* dt_dev, double* dt_acc, int repair,
                                    hipStream_t st) {
  dim3 grid((G->ny + LT - 1) / LT, (G->nx + LT - 1) / LT);
  if (F->h)
    hipLaunchKernelGGL(k_limit_update<true>, grid, dim3(LTHREADS), 0, st, *G, *e, *Un, *U, *F, *Uo,
                       alpha, beta, dt_dev, dt_acc, repair);
  else
    hipLaunchKernelGGL(k_limit_update<false>, grid, dim3(LTHREADS), 0, st, *G, *e, *Un, *U, *F,
                       *Uo, alpha, beta, dt_dev, dt_acc, repair);
  return hipGetLastError();
}

hipError_t wamp_launch_lu_march(const Geo* G, const Eos* e, const State* Un, const State* U,
                               const Faces* F, const State* Uo, double alpha, double beta,
                               const double* dt_dev, double* dt_acc, int repair, int seg,
                               hipStream_t st) {
  if (seg < 1) return hipErrorInvalidValue;
  const long long nseg = (G->nx + seg - 1) / seg;
  const long long items = (long long)((G->ny + MJ - 1) / MJ) * nseg;
  if (items > (1LL << 30)) return hipErrorInvalidValue;
  const int blocks = (int)((items + LU_WAVES - 1) / LU_WAVES);
  if (F->h)
    hipLaunchKernelGGL(k_lu_march<true>, dim3(blocks), dim3(64 * LU_WAVES), 0, st, *G, *e, *Un, *U,
                       *F, *Uo, alpha, beta, dt_dev, dt_acc, repair, seg, (int)nseg, (int)items);
  else
    hipLaunchKernelGGL(k_lu_march<false>, dim3(blocks), dim3(64 * LU_WAVES), 0, st, *G, *e, *Un,
                       *U, *F, *Uo, alpha, beta, dt_dev, dt_acc, repair, seg, (int)nseg,
                       (int)items);
  return hipGetLastError();
}

// The march that forms the first-order fluxes itself (FP64 face storage; the
// first-order planes of F are not read), `seg` rows per wave (any seg >= 1).
hipError_t wamp_launch_lu_lo(const Geo* G, const Eos* e, const State* Un, const State* U,
                             const Faces* F, const State* Uo, double alpha, double beta,
                             const double* dt_dev, double* dt_acc, int repair, int seg,
                             hipStream_t st) {
  if (F->h || seg < 1) return hipErrorInvalidValue;
  const long long nseg = (G->nx + seg - 1) / seg;
  const long long items = (long long)((G->ny + LOJ - 1) / LOJ) * nseg;
  if (items > (1LL << 30)) return hipErrorInvalidValue;
  const int blocks = (int)((items + LO_WAVES - 1) / LO_WAVES);
  hipLaunchKernelGGL(k_lu_lo, dim3(blocks), dim3(64 * LO_WAVES), 0, st, *G, *e, *Un, *U, *F, *Uo,
                     alpha, beta, dt_dev, dt_acc, repair, seg, (int)nseg, (int)items);
  return hipGetLastError();
}

hipError_t wamp_launch_dt(const Geo* G, const Eos* e, const State* U, double* dt_acc,
                          double* dt_out, double cfl, hipStream_t st) {
  const int n = G->nx * G->ny;
  hipLaunchKernelGGL(k_dt, dim3((n + STHREADS - 1) / STHREADS), dim3(STHREADS), 0, st, *G, *e,
                     *U, dt_acc);
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, dt_acc, dt_out, cfl);
  return hipGetLastError();
}

hipError_t wamp_launch_dt_finalize(double* dt_acc, double* dt_out, double cfl, hipStream_t st) {
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(64), 0, st, dt_acc, dt_out, cfl);
  return hipGetLastError();
}

hipError_t wamp_launch_selftest_elementary(const double* x, const double* y, double* out,
                                           long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_selftest_elementary, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x,
                     y, out, n);
  return hipGetLastError();
}

hipError_t wamp_launch_selftest_arith(const double* a, const double* b, double* out, long long n,
                                      hipStream_t st) {
  hipLaunchKernelGGL(k_selftest_arith, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b,
                     out, n);
  return hipGetLastError();
}

hipError_t wamp_launch_fill(double* p, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

hipError_t wamp_launch_prepare(const Geo* G, const Eos* e, const State* U, int repair,
                               hipStream_t st) {
  const int n = G->nx * G->ny;
  hipLaunchKernelGGL(k_prepare, dim3((n + STHREADS - 1) / STHREADS), dim3(STHREADS), 0, st, *G,
                     *e, *U, repair);
  return hipGetLastError();
}

hipError_t wamp_launch_enforce(const Geo* G, const Eos* e, const State* U, hipStream_t st) {
  const int n = (G->nx + 2 * G->g) * (G->ny + 2 * G->g);
  hipLaunchKernelGGL(k_enforce, dim3((n + STHREADS - 1) / STHREADS), dim3(STHREADS), 0, st, *G,
                     *e, *U);
  return hipGetLastError();
}

hipError_t wamp_launch_cons_to_prim(const Geo* G, const Eos* e, const State* U, double* up,
                                    double* vp, double* pp, hipStream_t st) {
  const int n = (G->nx + 2 * G->g) * (G->ny + 2 * G->g);
  hipLaunchKernelGGL(k_cons_to_prim, dim3((n + STHREADS - 1) / STHREADS), dim3(STHREADS), 0, st,
                     *G, *e, *U, up, vp, pp);
  return hipGetLastError();
}

hipError_t wamp_launch_bc(const Geo* G, const State* U, double* up, double* vp, double* pp,
                          hipStream_t st) {
  const int n = (G->nx + 2 * G->g) * (G->ny + 2 * G->g);
  hipLaunchKernelGGL(k_bc, dim3((n + STHREADS - 1) / STHREADS), dim3(STHREADS), 0, st, *G, *U, up,
                     vp, pp);
  return hipGetLastError();
}

hipError_t wamp_launch_ic_droplets(const Geo* G, const Eos* e, const State* U, const DropletIC* ic,
                                  hipStream_t st) {
  const long long n = (long long)(G->nx + 2 * G->g) * G->ld;
  hipLaunchKernelGGL(k_ic_droplets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, *G, *e,
                     *U, *ic);
  return hipGetLastError();
}

hipError_t wamp_launch_ic_triple(const Geo* G, const Eos* e, const State* U, const TripleIC* ic,
                                 hipStream_t st) {
  const long long n = (long long)(G->nx + 2 * G->g) * G->ld;
  hipLaunchKernelGGL(k_ic_triple, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, *G, *e,
                     *U, *ic);
  return hipGetLastError();
}

hipError_t wamp_launch_faces_batch(const CaseDesc* cs, const int* tile_off, int ncase,
                                   const BatchScal* sc, int stage, int base, int shear,
                                   int face_pp, int vf_always, double t1, int precision, int tile,
                                   int blocks, hipStream_t st) {
  if (precision != 0)
    return wamp_launch_faces_batch_mixed(cs, tile_off, ncase, sc, stage, base, shear, face_pp,
                                         vf_always, t1, tile, blocks, st);
#define WAMP_FB(TL, B, S, P, T)                                                              \
  hipLaunchKernelGGL((k_faces_batch<TL, B, S, P, T>), dim3(blocks), dim3(FTHREADS), 0, st, cs, \
                     tile_off, ncase, *sc, stage, vf_always, t1)
#define WAMP_FBS(B, S, P)                                                          \
  if (base == B && shear == S && face_pp == P) {                                  \
    if (tile == 0) WAMP_FB(Tile8x32, B, S, P, double);                             \
    else WAMP_FB(Tile16x16, B, S, P, double);                                      \
    return hipGetLastError();                                                      \
  }
  // the three named schemes: SC (0,1,0), M10 (0,0,1), MUSCL/MP5 + central6 (1,2,0)
  WAMP_FBS(0, 1, 0) WAMP_FBS(0, 0, 1) WAMP_FBS(1, 2, 0)
#undef WAMP_FBS
#undef WAMP_FB
  return hipErrorInvalidValue;
}

hipError_t wamp_launch_lu_batch(const CaseDesc* cs, const int* item_off, int ncase,
                                const BatchScal* sc, int stage, int repair, int seg, int blocks,
                                int f32, hipStream_t st) {
  if (seg < 1) return hipErrorInvalidValue;
  if (f32)
    hipLaunchKernelGGL(k_lu_batch<true>, dim3(blocks), dim3(256), 0, st, cs, item_off, ncase, *sc,
                       stage, repair, seg);
  else
    hipLaunchKernelGGL(k_lu_batch<false>, dim3(blocks), dim3(256), 0, st, cs, item_off, ncase,
                       *sc, stage, repair, seg);
  return hipGetLastError();
}

hipError_t wamp_launch_batch_finalize(const CaseDesc* cs, int ncase, const BatchScal* sc,
                                      hipStream_t st) {
  hipLaunchKernelGGL(k_batch_finalize, dim3((ncase + 63) / 64), dim3(64), 0, st, cs, ncase, *sc);
  return hipGetLastError();
}

hipError_t wamp_launch_batch_dt0(const Geo* G, const Eos* e, const State* U, double* acc,
                                 const BatchScal* sc, int k, double cfl, hipStream_t st) {
  const int n = G->nx * G->ny;
  hipLaunchKernelGGL(k_dt, dim3((n + STHREADS - 1) / STHREADS), dim3(STHREADS), 0, st, *G, *e, *U,
                     acc);
  hipLaunchKernelGGL(k_batch_dt0, dim3(1), dim3(64), 0, st, acc, *sc, k, cfl);
  return hipGetLastError();
}

size_t wamp_case_desc_size(void) { return sizeof(CaseDesc); }
}
