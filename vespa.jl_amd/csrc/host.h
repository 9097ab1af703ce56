// Host-side error plumbing shared by the translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

namespace vh {

// Records a formatted message for vh_last_error() and returns code.
int fail(int code, const char* fmt, const char* a = "", long long b = 0);

}  // namespace vh

#define VH_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return ::vh::fail(-2, "HIP error %s (code %lld) at " #call,    \
                                            hipGetErrorString(e_), (long long)e_);       \
  } while (0)

namespace vh {

// Clock kernels of the uniform-grid context, shared with the block context.
// clock: [0] t, [1] dt, [2] max CFL term, [3] steps.
void enqueue_dt_finalize(hipStream_t s, const double* partials, int np, double cfl, double tfinal, int clip,
                         double* clock);
void enqueue_set_dt(hipStream_t s, double* clock, double dt);
void enqueue_advance_clock(hipStream_t s, double* clock);

}  // namespace vh
