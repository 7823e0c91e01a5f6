// wamp_faces_mixed.hip -- the FP32-mixed face kernels (precision = 1).
//
// The same kernel source as the FP64 path (wamp_faces.h), instantiated with
// FP32 face arithmetic in its own translation unit so that it can be built
// differently from the parity path:
//  * FMA contraction on (-ffp-contract=on: a*b+c within one expression
//    becomes one v_fma), which the FP64 unit keeps off to follow the
//    reference's rounding;
//  * the FP64 divisions left in these kernels (pressure recovery, cell
//    indicators, face positivity limiter, energy flux) use the reciprocal
//    with one Newton step and a residual correction (ddiv, within one ulp);
//  * FP32 min/max are the hardware instructions.
// None of this changes what is FP64: the conserved state, the pressure
// recovery from E, the cell indicators, the positivity thresholds and the
// energy flux stay in FP64 arithmetic.
#define WAMP_MIXED_TU 1
#include "wamp_faces.h"

namespace wamp {
inline namespace mixed {

template <int BASE, int SHEAR, int FACEPP>
static void launch_faces_m(const Geo& G, const Eos& e, const State& U, const Faces& F,
                           int vf_always, double t1, hipStream_t st) {
  using TL = Tile8x32;
  dim3 grid((G.ny + 1 + TL::TJ - 1) / TL::TJ, (G.nx + 1 + TL::TI - 1) / TL::TI);
  hipLaunchKernelGGL((k_faces<TL, BASE, SHEAR, FACEPP, float>), grid, dim3(FTHREADS), 0, st, G,
                     e, U, F, vf_always, t1, 1);
}

}  // namespace mixed
}  // namespace wamp

extern "C" {
using namespace wamp;

hipError_t wamp_launch_faces_mixed(const Geo* G, const Eos* e, const State* U, const Faces* F,
                                   int base, int shear, int face_pp, int vf_always, double t1,
                                   hipStream_t st) {
#define WAMP_F(B, S, P)                                                   \
  if (base == B && shear == S && face_pp == P) {                         \
    launch_faces_m<B, S, P>(*G, *e, *U, *F, vf_always, t1, st);          \
    return hipGetLastError();                                            \
  }
  WAMP_F(0, 0, 0) WAMP_F(0, 0, 1) WAMP_F(0, 1, 0) WAMP_F(0, 1, 1) WAMP_F(0, 2, 0) WAMP_F(0, 2, 1)
  WAMP_F(1, 0, 0) WAMP_F(1, 0, 1) WAMP_F(1, 1, 0) WAMP_F(1, 1, 1) WAMP_F(1, 2, 0) WAMP_F(1, 2, 1)
#undef WAMP_F
  return hipErrorInvalidValue;
}

hipError_t wamp_launch_faces_batch_mixed(const CaseDesc* cs, const int* tile_off, int ncase,
                                         const BatchScal* sc, int stage, int base, int shear,
                                         int face_pp, int vf_always, double t1, int tile,
                                         int blocks, hipStream_t st) {
#define WAMP_FB(TL, B, S, P)                                                                    \
  hipLaunchKernelGGL((k_faces_batch<TL, B, S, P, float>), dim3(blocks), dim3(FTHREADS), 0, st, \
                     cs, tile_off, ncase, *sc, stage, vf_always, t1)
#define WAMP_FBS(B, S, P)                                   \
  if (base == B && shear == S && face_pp == P) {           \
    if (tile == 0) WAMP_FB(Tile8x32, B, S, P);              \
    else WAMP_FB(Tile16x16, B, S, P);                       \
    return hipGetLastError();                               \
  }
  // the three named schemes: SC (0,1,0), M10 (0,0,1), MUSCL/MP5 + central6 (1,2,0)
  WAMP_FBS(0, 1, 0) WAMP_FBS(0, 0, 1) WAMP_FBS(1, 2, 0)
#undef WAMP_FBS
#undef WAMP_FB
  return hipErrorInvalidValue;
}
}
