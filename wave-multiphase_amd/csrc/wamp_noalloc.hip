// wamp_noalloc.hip -- wamp_step_noalloc: one SSP-RK3 step (as wamp_step)
// that allocates nothing, for callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops, harnesses that forbid it).
//
// The step's scratch -- the 28 face planes, the stage buffers U1 and U2 and
// the dt accumulator -- lives in `scratch`, a static device array of
// WAMP_NOALLOC_SCRATCH_BYTES that every device the library runs on holds
// until the process ends.  Its address on a device comes from
// hipGetSymbolAddress, once per device, with that device current; kernels
// get that address, never the host-side name of the array.  Every call on a
// device shares that device's array, so calls that overlap on one device
// (different streams or threads) race; the array keeps the last call's
// contents, and each call first resets the one value that must start fresh
// (the dt accumulator, with a fill kernel, no memset).
#include <hip/hip_runtime.h>

#include <mutex>

#include "wamp.h"
#include "wamp_internal.h"

using namespace wamp;

static __device__ __align__(256) double scratch[WAMP_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // per device: the device address of `scratch`
int g_cus[kMaxDevices];        // per device: compute units
}  // namespace

// The current device's copy of `scratch` and its compute-unit count, looked
// up on the first call on that device.
static int device_scratch(int dev, void** ptr, int* cus) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(scratch)"))
      return rc;
    if (!p) return fail(-9, "hipGetSymbolAddress(scratch) gave a NULL address");
    int n = 0;
    if (int rc = hip_check(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev),
                           "hipDeviceGetAttribute"))
      return rc;
    g_cus[dev] = n;
    g_scratch[dev] = p;
  }
  *ptr = g_scratch[dev];
  *cus = g_cus[dev];
  return 0;
}

extern "C" int wamp_step_noalloc(const wamp_grid* grid, const wamp_eos* eos,
                                 const wamp_scheme* scheme, int32_t precision, double* const U[6],
                                 const double* dt_dev, double* dt_next_dev, void* stream) {
  if (int rc = check_case(grid, eos, scheme, precision)) return rc;
  State a;
  if (int rc = state_of(U, a, "U")) return rc;
  if (!dt_dev || !dt_next_dev) return fail(-2, "dt_dev and dt_next_dev must be non-NULL");
  const size_t plane = (size_t)(grid->nx + 2 * grid->g) * (size_t)grid->ld;
  const size_t need = step_scratch_bytes(plane, precision);
  if (need > (size_t)WAMP_NOALLOC_SCRATCH_BYTES)
    return fail(-8,
                "wamp_step_noalloc: a %d x %d grid (ld %d, %s) needs %zu bytes of step scratch; "
                "the static scratch holds %zu (use wamp_ctx_create + wamp_step for this grid)",
                grid->nx, grid->ny, grid->ld, precision == 0 ? "FP64" : "mixed", need,
                (size_t)WAMP_NOALLOC_SCRATCH_BYTES);
  int dev = 0;
  if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice")) return rc;
  if (dev < 0 || dev >= kMaxDevices) return fail(-9, "device %d beyond the %d supported", dev, kMaxDevices);
  void* base = nullptr;
  int cus = 0;
  if (int rc = device_scratch(dev, &base, &cus)) return rc;
  StepPlan p;
  plan_fill(p, *grid, *eos, *scheme, precision, dev, cus);
  const StepScratch scr = carve_step_scratch(base, plane, precision);
  hipStream_t st = (hipStream_t)stream;
  if (int rc = hip_check(wamp_launch_fill(scr.dt_acc, 1.0e30, st), "fill dt accumulator")) return rc;
  return step_enqueue(p, a, scr, dt_dev, dt_next_dev, st, nullptr);
}
