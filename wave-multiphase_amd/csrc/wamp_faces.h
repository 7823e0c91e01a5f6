// wamp_faces.h -- the fused face-flux kernel (k_faces, k_faces_batch) and its
// tile routine, shared by two translation units: wamp_kernels.hip
// instantiates the FP64 kernels (contraction off, the parity path) and
// wamp_faces_mixed.hip the FP32-mixed ones (contraction on, fast FP64
// division; namespace wamp::mixed).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wamp_device.h"
#include "wamp_types.h"

WAMP_DEV_BEGIN

// ------------------------------------------------------------------------
// Boundary fill by index mapping.  Ghost cell (i, j) holds the edge
// interior cell: bc_x copies cell g / g+nx-1 into every x-ghost layer
// (SC:1630-1646), then bc_y copies row g / g+ny-1 of the x-filled array
// into every y-ghost layer (SC:1664-1682), so corner cells hold the corner
// interior cell.  Slip walls negate the wall-normal momentum.
// ------------------------------------------------------------------------
WAMP_FN size_t map_cell(const Geo& G, int i, int j, double& sru, double& srv) {
  sru = 1.0;
  srv = 1.0;
  if (i < G.g) {
    i = G.g;
    if (G.bc[0]) sru = -1.0;
  } else if (i >= G.g + G.nx) {
    i = G.g + G.nx - 1;
    if (G.bc[1]) sru = -1.0;
  }
  if (j < G.g) {
    j = G.g;
    if (G.bc[2]) srv = -1.0;
  } else if (j >= G.g + G.ny) {
    j = G.g + G.ny - 1;
    if (G.bc[3]) srv = -1.0;
  }
  return (size_t)i * (size_t)G.ld + (size_t)j;
}

// ========================================================================
// k_faces
// ========================================================================
// Tile of TI x TJ face positions (j contiguous), 256 threads.
template <int TI_, int TJ_>
struct FTile {
  static constexpr int TI = TI_, TJ = TJ_;
  static constexpr int RI = TI + 7, RJ = TJ + 7;  // LDS rectangle: cells i0-3 .. i0+TI+3
  static constexpr int RN = RI * RJ;
  static constexpr int CXN = (TI + 3) * TJ;       // x-indicator cells i0-1..i0+TI+1, j0..j0+TJ-1
  static constexpr int CYN = TI * (TJ + 3);       // y-indicator cells i0..i0+TI-1, j0-1..j0+TJ+1
  static constexpr int THREADS = TI * TJ;
};
using Tile8x32 = FTile<8, 32>;
using Tile16x16 = FTile<16, 16>;
constexpr int FTHREADS = 256;

template <class TL>
struct FaceSmem {
  double q[6][TL::RN];  // ar1, ar2, ru, rv, p, a1
  double ent[TL::RN];
  double cx[TL::CXN], cy[TL::CYN];
  double sx[TL::CXN], sy[TL::CYN];
};

// Ducros x pressure sensor (ducros_kernel SC:875-916) at LDS cell r.
template <class TL>
WAMP_FN double lds_vel(const FaceSmem<TL>& S, int r, int comp) {
  return dv(S.q[comp][r], S.q[0][r] + S.q[1][r]);
}
template <class TL, int DIR>
WAMP_FN double ducros_sigma(const FaceSmem<TL>& S, int r) {
  constexpr int FRJ = TL::RJ;
  double der_ux = 0.5 * (-lds_vel(S, r - FRJ, 2) + lds_vel(S, r + FRJ, 2));
  double der_vx = 0.5 * (-lds_vel(S, r - FRJ, 3) + lds_vel(S, r + FRJ, 3));
  double der_uy = 0.5 * (-lds_vel(S, r - 1, 2) + lds_vel(S, r + 1, 2));
  double der_vy = 0.5 * (-lds_vel(S, r - 1, 3) + lds_vel(S, r + 1, 3));
  double div_u = der_ux + der_vy;
  double curl_u = der_vx - der_uy;
  double ducros = ddiv(div_u * div_u, div_u * div_u + curl_u * curl_u + 1.0e-30);
  const int s = DIR == 0 ? FRJ : 1;
  const double* p = S.q[4];
  double pm2 = p[r - 2 * s], pm1 = p[r - s], p0 = p[r], pp1 = p[r + s], pp2 = p[r + 2 * s];
  double aa = fabs(-pm2 + 16.0 * pm1 - 30.0 * p0 + 16.0 * pp1 - pp2);
  double bb = fabs(pm2 + 16.0 * pm1 + 30.0 * p0 + 16.0 * pp1 + pp2);
  return ddiv(aa, bb + 1.0e-30) * ducros;
}

// High-order flux through one face whose left cell is LDS index r0; cells
// along the normal are r0 + k*s.  ci / sg: indicator and Ducros sensor at
// the cells -1..2.  Writes 6 fluxes + u* to out.
template <class TL, int DIR, int BASE, int SHEAR, int FACEPP, typename T>
WAMP_FN void face_high(const Eos& e, const FaceSmem<TL>& S, int r0, const double* ci,
                                          const double* sg, int vf_always, double t1d, double* out) {
  constexpr int s = DIR == 0 ? TL::RJ : 1;
  constexpr bool wide = BASE == 1 || SHEAR == 2;  // needs cells -2 and +3
  constexpr int klo = wide ? 0 : 1, khi = wide ? 6 : 5;
  const T t1 = (T)t1d;

  bool use_thinc = wmin(wmin(ci[0], ci[1]), wmin(ci[2], ci[3])) < THINC_THRESH;
  bool use_smooth = false;
  if (SHEAR != 0) use_smooth = wmax(wmax(sg[0], sg[1]), wmax(sg[2], sg[3])) < SIGMA_SMOOTH;

  // face-averaged state, SC:1039-1048
  const int rA = r0, rB = r0 + s;
  T ar1a = T(0.5) * ((T)S.q[0][rA] + (T)S.q[0][rB]);
  T ar2a = T(0.5) * ((T)S.q[1][rA] + (T)S.q[1][rB]);
  T a1a = T(0.5) * ((T)S.q[5][rA] + (T)S.q[5][rB]);
  T pa = T(0.5) * ((T)S.q[4][rA] + (T)S.q[4][rB]);
  T rhoa = ar1a + ar2a;
  T u_avg = dv(T(0.5) * ((T)S.q[2][rA] + (T)S.q[2][rB]), rhoa);
  T v_avg = dv(T(0.5) * ((T)S.q[3][rA] + (T)S.q[3][rB]), rhoa);
  T ca = sound_speed<T>(e, rhoa, pa, a1a);
  Basis<T> B = make_basis<DIR>(ca, rhoa, ar1a, ar2a, u_avg, v_avg);

  // characteristic projection of the stencil cells -2..3 (SC:1051-1056)
  T w[6][6];
#pragma unroll
  for (int k = klo; k < khi; ++k) {
    T u[6];
    const int r = r0 + (k - 2) * s;
#pragma unroll
    for (int m = 0; m < 6; ++m) u[m] = (T)S.q[m][r];
    to_char<DIR>(B, u, w[k]);
  }
  T wl[6], wr[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    if (BASE == 0) {
      wl[m] = muscl_l(w[1][m], w[2][m], w[3][m]);
      wr[m] = muscl_r(w[2][m], w[3][m], w[4][m]);
    } else {
      wl[m] = mp5_l(w[0][m], w[1][m], w[2][m], w[3][m], w[4][m]);
      wr[m] = mp5_r(w[1][m], w[2][m], w[3][m], w[4][m], w[5][m]);
    }
  }
  if (SHEAR != 0 && use_smooth) {
    T iph = SHEAR == 1 ? central4(w[1][3], w[2][3], w[3][3], w[4][3])
                       : central6(w[0][3], w[1][3], w[2][3], w[3][3], w[4][3], w[5][3]);
    wl[3] = iph;
    wr[3] = iph;
  }
  if (use_thinc) {
    wl[1] = thinc_l(w[1][1], w[2][1], w[3][1], t1);
    wl[2] = thinc_l(w[1][2], w[2][2], w[3][2], t1);
    wr[1] = thinc_r(w[2][1], w[3][1], w[4][1], t1);
    wr[2] = thinc_r(w[2][2], w[3][2], w[4][2], t1);
  }
  if (use_thinc || vf_always) {
    wl[4] = thinc_l(w[1][4], w[2][4], w[3][4], t1);
    wr[4] = thinc_r(w[2][4], w[3][4], w[4][4], t1);
  }
  T phL[6], phR[6], phfL[6], phfR[6];
  to_phys<DIR>(B, wl, phL);
  to_phys<DIR>(B, wr, phR);
  to_phys<DIR>(B, w[2], phfL);
  to_phys<DIR>(B, w[3], phfR);

  const bool cbad = ca < T(0) || isnan_(ca);
  if (FACEPP == 0) {  // SC:1103-1111
    if (phL[0] < T(0) || isnan_(phL[0])) phL[0] = phfL[0];
    if (phL[1] < T(0) || isnan_(phL[1])) phL[1] = phfL[1];
    if (phL[5] < T(0) || phL[5] > T(1)) phL[5] = phfL[5];
    if (isnan_(phL[4])) phL[4] = phfL[4];
    if (phR[0] < T(0) || isnan_(phR[0])) phR[0] = phfR[0];
    if (phR[1] < T(0) || isnan_(phR[1])) phR[1] = phfR[1];
    if (phR[5] < T(0) || phR[5] > T(1)) phR[5] = phfR[5];
    if (isnan_(phR[4])) phR[4] = phfR[4];
    if (cbad) {  // SC:1113-1116
#pragma unroll
      for (int m = 0; m < 6; ++m) {
        phL[m] = phfL[m];
        phR[m] = phfR[m];
      }
    }
  } else if constexpr (sizeof(T) == sizeof(double)) {  // M10:1007-1021
    // each side's sound-speed fallback (M10:1023-1025) right after its own
    // sequence, so the first-order state of L is dead while R is limited
    face_state_pp(e, phL, phfL);
    if (cbad) {
#pragma unroll
      for (int m = 0; m < 6; ++m) phL[m] = phfL[m];
    }
    face_state_pp(e, phR, phfR);
    if (cbad) {
#pragma unroll
      for (int m = 0; m < 6; ++m) phR[m] = phfR[m];
    }
  } else {  // the same on FP64 copies: the thresholds are always applied in FP64
    double d[6], df[6];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      d[m] = phL[m];
      df[m] = phfL[m];
    }
    face_state_pp(e, d, df);
#pragma unroll
    for (int m = 0; m < 6; ++m) phL[m] = cbad ? phfL[m] : (T)d[m];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      d[m] = phR[m];
      df[m] = phfR[m];
    }
    face_state_pp(e, d, df);
#pragma unroll
    for (int m = 0; m < 6; ++m) phR[m] = cbad ? phfR[m] : (T)d[m];
  }
  // primitives for HLLC, SC:1118-1125
  T rhoL = phL[0] + phL[1], rhoR = phR[0] + phR[1];
  T PL[6] = {phL[0], phL[1], dv(phL[2], rhoL), dv(phL[3], rhoL), phL[4], phL[5]};
  T PR[6] = {phR[0], phR[1], dv(phR[2], rhoR), dv(phR[3], rhoR), phR[4], phR[5]};
  if constexpr (sizeof(T) == sizeof(double)) {
    hllc<DIR>(e, PL, PR, out, out[6]);
  } else {
    hllc_mixed<DIR>(e, PL, PR, out, out[6]);
  }
}

// First-order flux (fx_low_kernel SC:1303-1347 / gy_low_kernel SC:1350-1394).
template <class TL, int DIR, typename T>
WAMP_FN void face_low(const Eos& e, const FaceSmem<TL>& S, int r0, double* out) {
  constexpr int s = DIR == 0 ? TL::RJ : 1;
  double pd[2][6];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int r = r0 + k * s;
    const double rho = S.q[0][r] + S.q[1][r];
    low_prims(e, S.q[0][r], S.q[1][r], dv(S.q[2][r], rho), dv(S.q[3][r], rho), S.q[4][r],
              S.q[5][r], pd[k]);
  }
  T PL[6], PR[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) {
    PL[m] = (T)pd[0][m];
    PR[m] = (T)pd[1][m];
  }
  if constexpr (sizeof(T) == sizeof(double)) {
    hllc<DIR>(e, PL, PR, out, out[6]);
  } else {
    hllc_mixed<DIR>(e, PL, PR, out, out[6]);
  }
}

// One tile of face positions (bi, bj) of one case in three phases, each a
// function of one work index so that the device runs a phase with one
// thread per index (a barrier between phases) and the host with a loop:
//   A  stage cell r of the rectangle: conserved values, pressure and
//      entropy once per cell (the cells StageSet lists);
//   B  entropy indicator (cator2_val SC:318-325) and Ducros sensor once per
//      cell along each sweep direction (k < TL::CXN, k < TL::CYN);
//   C  the x-face and the y-face at tile position t (t < TL::THREADS).
// Phase A in two halves, so that a thread can issue the loads of all its
// cells before it computes any of them: the conserved values of rectangle
// cell r (its boundary-mapped cell) and the slip signs ...
struct StagedCell {
  double v[6];
  double sru, srv;
};
template <class TL>
WAMP_FN void tile_load_cell(const Geo& G, const State& U, int i0, int j0, int r, StagedCell& c) {
  constexpr int FRJ = TL::RJ;
  const int i = i0 - 3 + r / FRJ;
  const int j = j0 - 3 + r % FRJ;
  const size_t o = map_cell(G, i, j, c.sru, c.srv);
#pragma unroll
  for (int m = 0; m < 6; ++m) c.v[m] = U.f[m][o];
}
// ... then the slip signs applied, pressure and entropy, and the LDS stores.
template <class TL>
WAMP_FN void tile_store_cell(const Eos& e, const StagedCell& c, int r, FaceSmem<TL>& S) {
  const double ar1 = c.v[0], ar2 = c.v[1], E = c.v[4], al = c.v[5];
  double ru = c.v[2], rv = c.v[3];
  if (c.sru < 0.0) ru = -1.0 * ru;
  if (c.srv < 0.0) rv = -1.0 * rv;
  double uu, vv;
  double p = pressure(e, ar1, ar2, ru, rv, E, al, uu, vv);
  S.q[0][r] = ar1;
  S.q[1][r] = ar2;
  S.q[2][r] = ru;
  S.q[3][r] = rv;
  S.q[4][r] = p;
  S.q[5][r] = al;
  S.ent[r] = entropy(e, ar1, ar2, p, al);
}
template <class TL>
WAMP_FN void tile_stage_cell(const Geo& G, const Eos& e, const State& U, int i0, int j0, int r,
                             FaceSmem<TL>& S) {
  StagedCell c;
  tile_load_cell<TL>(G, U, i0, j0, r, c);
  tile_store_cell<TL>(e, c, r, S);
}
// The rectangle cells phase A stages, by work index k < N: only those that
// phases B and C read.  The x-faces and their indicators read LDS columns
// 3..TJ+2 over all rows, the y-faces and theirs LDS rows 3..TI+2 over all
// columns; the corner blocks outside that cross are never read, except, with
// the Ducros sensor (SHEAR != 0), the 8 cells its velocity differences reach
// at the sensor cells next to the corners: (2, 2), (2, TJ+3), (2, TJ+4),
// (TI+3, 2), (TI+3, TJ+3), (TI+3, TJ+4), (TI+4, 2), (TI+4, TJ+3).  At 16x16
// that is 480 (488) of the 529 cells: two passes of the block's 256 threads
// instead of three.
template <class TL, int SHEAR>
struct StageSet {
  static constexpr int BAND = TL::TI * TL::RJ;        // LDS rows 3..TI+2, every column
  static constexpr int CROSS = BAND + 7 * TL::TJ;     // + rows 0-2, TI+3..TI+6 at columns 3..TJ+2
  static constexpr int N = CROSS + (SHEAR != 0 ? 8 : 0);
  WAMP_FN static int cell(int k) {
    if (k < BAND) return 3 * TL::RJ + k;
    if (k < CROSS) {
      const int q = k - BAND, rr = q / TL::TJ;  // rr 0..6: rows 0-2, TI+3..TI+6
      return (rr < 3 ? rr : TL::TI + rr) * TL::RJ + 3 + q % TL::TJ;
    }
    const int q = k - CROSS;  // 0..7, the Ducros cells in the order listed above
    const int row = q < 3 ? 2 : (q < 6 ? TL::TI + 3 : TL::TI + 4);
    const int col = q % 3 == 0 ? 2 : TL::TJ + 2 + q % 3;
    return row * TL::RJ + col;
  }
};
template <class TL, int SHEAR>
WAMP_FN void tile_sensor_x(FaceSmem<TL>& S, int k) {
  constexpr int FTJ = TL::TJ, FRJ = TL::RJ;
  const int r = (k / FTJ + 2) * FRJ + (k % FTJ + 3);
  S.cx[k] = cator2(S.ent[r - 2 * FRJ], S.ent[r - FRJ], S.ent[r], S.ent[r + FRJ], S.ent[r + 2 * FRJ]);
  if (SHEAR != 0) S.sx[k] = ducros_sigma<TL, 0>(S, r);
}
template <class TL, int SHEAR>
WAMP_FN void tile_sensor_y(FaceSmem<TL>& S, int k) {
  constexpr int FTJ = TL::TJ, FRJ = TL::RJ;
  const int r = (k / (FTJ + 3) + 3) * FRJ + (k % (FTJ + 3) + 2);
  S.cy[k] = cator2(S.ent[r - 2], S.ent[r - 1], S.ent[r], S.ent[r + 1], S.ent[r + 2]);
  if (SHEAR != 0) S.sy[k] = ducros_sigma<TL, 1>(S, r);
}
template <class TL, int BASE, int SHEAR, int FACEPP, typename T>
WAMP_FN void tile_faces_at(const Geo& G, const Eos& e, const Faces& F, int vf_always, double t1,
                           int low, int i0, int j0, int t, const FaceSmem<TL>& S) {
  constexpr int FTJ = TL::TJ, FRJ = TL::RJ;
  const int ti = t / FTJ, tj = t % FTJ;
  const int i = i0 + ti, j = j0 + tj;
  const int r0 = (ti + 3) * FRJ + (tj + 3);
  const size_t o = (size_t)i * (size_t)G.ld + (size_t)j;
  if (i >= G.g - 1 && i <= G.g + G.nx - 1 && j >= G.g && j <= G.g + G.ny - 1) {
    double ci[4], sg[4], out[7];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ci[k] = S.cx[(ti + k) * FTJ + tj];
      sg[k] = SHEAR != 0 ? S.sx[(ti + k) * FTJ + tj] : 0.0;
    }
    face_high<TL, 0, BASE, SHEAR, FACEPP, T>(e, S, r0, ci, sg, vf_always, t1, out);
#pragma unroll
    for (int m = 0; m < 7; ++m) fstore<T>(F, m, o, out[m]);
    if (low) {
      face_low<TL, 0, T>(e, S, r0, out);
#pragma unroll
      for (int m = 0; m < 7; ++m) fstore<T>(F, 14 + m, o, out[m]);
    }
  }
  if (i >= G.g && i <= G.g + G.nx - 1 && j >= G.g - 1 && j <= G.g + G.ny - 1) {
    double ci[4], sg[4], out[7];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ci[k] = S.cy[ti * (FTJ + 3) + tj + k];
      sg[k] = SHEAR != 0 ? S.sy[ti * (FTJ + 3) + tj + k] : 0.0;
    }
    face_high<TL, 1, BASE, SHEAR, FACEPP, T>(e, S, r0, ci, sg, vf_always, t1, out);
#pragma unroll
    for (int m = 0; m < 7; ++m) fstore<T>(F, 7 + m, o, out[m]);
    if (low) {
      face_low<TL, 1, T>(e, S, r0, out);
#pragma unroll
      for (int m = 0; m < 7; ++m) fstore<T>(F, 21 + m, o, out[m]);
    }
  }
}

// Tile (bi, bj) by the 256 threads of a block.
template <class TL, int BASE, int SHEAR, int FACEPP, typename T>
__device__ __forceinline__ void faces_tile(const Geo& G, const Eos& e, const State& U, const Faces& F,
                                           int vf_always, double t1, int low, int bi, int bj,
                                           FaceSmem<TL>& S) {
  const int tid = threadIdx.x;
  const int i0 = G.g - 1 + bi * TL::TI;
  const int j0 = G.g - 1 + bj * TL::TJ;
  // Phase A: every load of the thread's cells first (a thread without a cell
  // in the last pass loads its first cell again), then their arithmetic, so
  // the wave waits on memory once instead of once per pass.
  using SS = StageSet<TL, SHEAR>;
  constexpr int NP = (SS::N + FTHREADS - 1) / FTHREADS;
  static_assert(SS::N >= FTHREADS, "every thread stages a first cell");
  StagedCell cells[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int k = tid + q * FTHREADS;
    tile_load_cell<TL>(G, U, i0, j0, SS::cell(k < SS::N ? k : tid), cells[q]);
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const int k = tid + q * FTHREADS;
    if (k < SS::N) tile_store_cell<TL>(e, cells[q], SS::cell(k), S);
  }
  __syncthreads();
  // Phase B: one x and one y indicator per thread, then the remaining ones
  // in a third pass whose x and y parts start on separate waves, so that no
  // wave runs more than three of them (a strided loop over each kind in turn
  // would give wave 0 four).
  constexpr int LX = TL::CXN - FTHREADS, LY = TL::CYN - FTHREADS, OY = (LX + 63) / 64 * 64;
  static_assert(LX > 0 && LY > 0 && OY + LY <= FTHREADS, "tile shape outside phase B's three passes");
  tile_sensor_x<TL, SHEAR>(S, tid);
  tile_sensor_y<TL, SHEAR>(S, tid);
  if (tid < LX)
    tile_sensor_x<TL, SHEAR>(S, FTHREADS + tid);
  else if (tid >= OY && tid < OY + LY)
    tile_sensor_y<TL, SHEAR>(S, FTHREADS + tid - OY);
  __syncthreads();
  tile_faces_at<TL, BASE, SHEAR, FACEPP, T>(G, e, F, vf_always, t1, low, i0, j0, tid, S);
}

template <class TL, int BASE, int SHEAR, int FACEPP, typename T>
__global__ void __launch_bounds__(FTHREADS, (BASE == 0 || sizeof(T) == 4) ? 3 : 2)
k_faces(Geo G, Eos e, State U, Faces F, int vf_always, double t1, int low) {
  __shared__ FaceSmem<TL> S;
  faces_tile<TL, BASE, SHEAR, FACEPP, T>(G, e, U, F, vf_always, t1, low, (int)blockIdx.y,
                                         (int)blockIdx.x, S);
}

__device__ __forceinline__ int find_case(const int* __restrict__ off, int n, int t) {
  int lo = 0, hi = n;  // off[lo] <= t < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <class TL, int BASE, int SHEAR, int FACEPP, typename T>
__global__ void __launch_bounds__(FTHREADS, (BASE == 0 || sizeof(T) == 4) ? 3 : 2)
k_faces_batch(const CaseDesc* __restrict__ cs, const int* __restrict__ tile_off, int ncase,
              BatchScal sc, int stage, int vf_always, double t1) {
  __shared__ FaceSmem<TL> S;
  const int total = tile_off[ncase];
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    const int k = find_case(tile_off, ncase, t);
    if (sc.done[k]) continue;  // block-uniform
    const CaseDesc& c = cs[k];
    const int lt = t - tile_off[k];
    State U;
#pragma unroll
    for (int m = 0; m < 6; ++m) U.f[m] = c.u[stage][m];
    const Faces F = c.f;
    faces_tile<TL, BASE, SHEAR, FACEPP, T>(c.geo, c.eos, U, F, vf_always, t1, 1, lt / c.tiles_j,
                                           lt % c.tiles_j, S);
    __syncthreads();  // the next tile overwrites the LDS
  }
}

WAMP_DEV_END
