// wamp_internal.h -- pieces of the C-ABI layer shared by wamp_capi.cpp and
// wamp_noalloc.hip: error reporting, argument checks, the step's scratch
// layout and the stage sequence of one SSP-RK3 step (SC:2028-2059).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "wamp.h"
#include "wamp_types.h"

namespace wamp {

// Record the message of a failing call (wamp_last_error) and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// 0 for hipSuccess, else the hipError_t with its message recorded.
int hip_check(hipError_t e, const char* what);
// The six state planes of a caller's pointer list.
int state_of(double* const U[6], State& s, const char* name);

// Everything one step of one case needs besides its buffers.
struct StepPlan {
  Geo geo;
  Eos eos;
  wamp_scheme scheme;
  int precision;
  int device;
  double t1;     // tanh(beta_thinc / 2), SC:334
  int lu_kind;   // 0: LDS-tile limiter/update kernel, 1: marching kernel, 2: march forming the
                 // first-order fluxes itself (FP64), -1: by grid size
  int lu_seg;    // rows per wave of the marching kernel (0: by grid size)
  int cu_count;  // compute units of the device (limiter segment choice)
};

// Check a case's grid, EOS, scheme and precision (0 or a negative code).
int check_case(const wamp_grid* grid, const wamp_eos* eos, const wamp_scheme* scheme,
               int32_t precision);
// Fill a plan from checked arguments; reads WAMP_LU_KIND / WAMP_LU_SEG.
// Device and compute-unit count are the caller's (no runtime call here).
void plan_fill(StepPlan& p, const wamp_grid& grid, const wamp_eos& eos,
               const wamp_scheme& scheme, int precision, int device, int cu_count);

// Scratch of one step: the 28 face planes (mixed precision: 4 FP64 energy
// planes + 24 FP32 planes), the stage buffers U1 and U2 and the dt
// accumulator, carved from one block of step_scratch_bytes(plane, precision)
// bytes (8-byte aligned) for `plane` = (nx + 2g) * ld cells per plane.
struct StepScratch {
  Faces faces;
  State u1, u2;
  double* dt_acc;
};
size_t face_bytes(size_t plane, int precision);
Faces carve_faces(char* p, size_t plane, int precision);
size_t step_scratch_bytes(size_t plane, int precision);
StepScratch carve_step_scratch(void* base, size_t plane, int precision);

// Enqueue one SSP-RK3 step in place on `a` (dt read from dt_dev, cfl * min
// of the new state written to dt_next; the accumulator must hold 1e30 and
// is left at 1e30).  ev: NULL, or 7 events recorded around the kernels.
int step_enqueue(const StepPlan& p, const State& a, const StepScratch& s, const double* dt_dev,
                 double* dt_next, hipStream_t st, hipEvent_t* ev);

}  // namespace wamp

extern "C" hipError_t wamp_launch_fill(double* p, double v, hipStream_t);
