// wamp3_faces_mixed.hip -- the FP32-mixed 3D face kernels (precision = 1).
//
// The same kernel source as the FP64 3D path (wamp3_faces.h), instantiated
// with FP32 projection, reconstruction and HLLC in its own translation unit,
// built as wamp_faces_mixed.hip is: FMA contraction on, the one-ulp FP64
// division, hardware FP32 min/max.  Pressure, the cell indicators (read from
// the FP64 planes k3_cells writes), the M10 face limiter and the energy flux
// stay FP64.  The kernels write the mixed face storage of Faces3.
#define WAMP_MIXED_TU 1
#include "wamp3_faces.h"

extern "C" hipError_t wamp3_launch_faces_mixed(const wamp::Geo3* G, const wamp::Eos* e,
                                               const wamp::State3* U, const double* P,
                                               const double* S, const double* D,
                                               const wamp::Faces3* F, int base, int shear,
                                               int vf_always, int face_pp, double t1,
                                               hipStream_t st) {
  return wamp::mixed::launch_faces3<float>(*G, *e, *U, P, S, D, *F, base, shear, vf_always, face_pp,
                                           t1, st);
}

extern "C" hipError_t wamp3_launch_faces_batch_mixed(const wamp::Case3Desc* cs, const int* off,
                                                     int ncase, const wamp::BatchScal* sc,
                                                     int stage, int base, int shear,
                                                     int vf_always, int face_pp, double t1,
                                                     const int* blocks, hipStream_t st) {
  return wamp::mixed::launch_faces3_batch<float>(cs, off, ncase, *sc, stage, base, shear,
                                                 vf_always, face_pp, t1, blocks, st);
}
