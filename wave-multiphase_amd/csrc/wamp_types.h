// wamp_types.h -- argument structures shared by the kernels and the C-ABI layer.
#pragma once

namespace wamp {

// Two stiffened-gas phases.  r1, r2 are the correctly rounded reciprocals
// 1/(g1-1), 1/(g2-1), formed on the host (make_eos); the FP64 device EOS
// divides by g-1 through them.
struct Eos {
  double g1, p1, g2, p2;
  double r1, r2;
};
inline Eos make_eos_values(double g1, double p1, double g2, double p2) {  // host
  return Eos{g1, p1, g2, p2, 1.0 / (g1 - 1.0), 1.0 / (g2 - 1.0)};
}

struct Geo {  // one case's grid: interior nx x ny, g ghost layers, row stride ld
  int nx, ny, g, ld;
  double dx, dy;
  int bc[4];
};
struct State {  // the six conserved planes [a1 rho1, a2 rho2, rho u, rho v, E, a1]
  double* f[6];
};
// The 28 face planes fx0..5, srcx, gy0..5, srcy, fxl0..5, srcxl, gyl0..5,
// srcyl (groups of 7: high x, high y, low x, low y).
//  * FP64 storage (h == nullptr): equally spaced in one allocation, plane m
//    starts at b + m * s.
//  * Mixed storage (h != nullptr; the FP32-mixed face kernels, whose fluxes
//    other than the energy flux are FP32 values): the four energy planes
//    (m % 7 == 4) in FP64 at b + (m / 7) * s, the other 24 in FP32 at
//    h + face_slot32(m) * hs.  Storing those FP32 values in FP32 loses
//    nothing, so both storages give the limiter the same numbers.
struct Faces {
  double* b;
  long long s;
  float* h = nullptr;
  long long hs = 0;
  __host__ __device__ double* operator[](int m) const { return b + (long long)m * s; }  // FP64 storage
};
__host__ __device__ constexpr int face_slot32(int m) { return m - m / 7 - (m % 7 > 4 ? 1 : 0); }
__host__ __device__ constexpr bool face_is_energy(int m) { return m % 7 == 4; }

// One case of a batch (device-resident table).
struct CaseDesc {
  Geo geo;
  Eos eos;
  double* u[3][6];  // [0] state (U^n, overwritten by stage 3), [1] U^1, [2] U^2
  Faces f;          // face planes
  int tiles_j, tiles;        // face tiles along j and in total
  int strips, items;         // limiter work items: strips x segments
  double cfl;
};

// Initial conditions built on the device (cases of the measured workloads).
// Planar shock at shock_x moving +x into air at rest holding water droplets
// (xc, yc, r, r_cavity) smeared over `width` with a tanh profile.
struct DropletIC {
  double xmin, ymin;
  double shock_x, rho_post, p_post, u_post;
  double rho_water, rho_air, p0;
  double width;
  const double* drops;  // n_drops x 4, device memory
  int n_drops;
};
// Three constant states split at x_split and (for x >= x_split) y_split:
// left and top in phase 1, bottom in phase 2 (the gas-gas triple point).
struct TripleIC {
  double xmin, ymin, x_split, y_split;
  double rho_left, p_left, rho_top, p_top, rho_bottom, p_bottom;
};

// Three-dimensional grid: interior nx x ny x nz, g ghost layers, cell
// (i, j, k) at i * sx + j * sy + k (k contiguous; sy >= nz + 2g,
// sx = (ny + 2g) * sy).  bc: x-lo, x-hi, y-lo, y-hi, z-lo, z-hi.
struct Geo3 {
  int nx, ny, nz, g;
  long long sx, sy;
  double dx, dy, dz;
  int bc[6];
};
struct State3 {  // [a1 rho1, a2 rho2, rho u, rho v, rho w, E, a1]
  double* f[7];
};
// The 48 face planes of a 3D step: direction d (x, y, z) holds 16 planes at
// d * 16: the high-order fluxes of the 7 conserved variables and u*, then
// the first-order ones.  The face d + 1/2 of cell (i, j, k) is stored at
// the cell's index.
//  * FP64 storage (h == nullptr): plane m starts at b + m * s.
//  * Mixed storage (the FP32-mixed face kernels, whose fluxes other than the
//    energy flux are FP32 values): the six energy planes (m % 8 == 5) in
//    FP64 at b + face3_slot64(m) * s, the other 42 in FP32 at
//    h + face3_slot32(m) * hs.
struct Faces3 {
  double* b;
  long long s;
  float* h = nullptr;
  long long hs = 0;
  __host__ __device__ double* operator[](int m) const { return b + (long long)m * s; }  // FP64 storage
};
__host__ __device__ constexpr bool face3_is_energy(int m) { return m % 8 == 5; }
__host__ __device__ constexpr int face3_slot64(int m) { return m / 8; }
__host__ __device__ constexpr int face3_slot32(int m) { return (m / 8) * 7 + (m % 8 < 5 ? m % 8 : m % 8 - 1); }

// Per-slot device scalars of a batch.
struct BatchScal {
  double* dt;
  double* acc;
  double* t;
  double* tend;  // <= 0: no end time
  int* done;
};

// One case of a 3D batch (device-resident table).  Its work is counted in
// blocks of the 3D kernels' (64, 4) thread shape, per kind: the padded grid
// (cells), the faces of each sweep, the interior (limiter and update).
enum { B3_CELLS = 0, B3_FACE0 = 1, B3_INT = 4, B3_KINDS = 5 };
struct Case3Desc {
  Geo3 geo;
  Eos eos;
  double* u[3][7];  // [0] state (U^n, overwritten by stage 3), [1] U^1, [2] U^2
  double *P, *S, *D;  // pressure, entropy, Ducros planes
  Faces3 f;
  double* th[3];    // limiter theta planes
  double cfl;
};
// (nk, nj, ni) extents of work kind `kind` on grid G; its blocks are
// ceil(nk / 64) x ceil(nj / 4) x ni.
__host__ __device__ inline void kind_extent3(const Geo3& G, int kind, int& nk, int& nj, int& ni) {
  if (kind == B3_CELLS) {
    nk = G.nz + 2 * G.g;
    nj = G.ny + 2 * G.g;
    ni = G.nx + 2 * G.g;
  } else {
    nk = G.nz + (kind == B3_FACE0 + 2 ? 1 : 0);
    nj = G.ny + (kind == B3_FACE0 + 1 ? 1 : 0);
    ni = G.nx + (kind == B3_FACE0 ? 1 : 0);
  }
}
__host__ __device__ inline long long kind_blocks3(const Geo3& G, int kind) {
  int nk, nj, ni;
  kind_extent3(G, kind, nk, nj, ni);
  return (long long)((nk + 63) / 64) * ((nj + 3) / 4) * ni;
}

}  // namespace wamp
