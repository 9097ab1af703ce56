// Shared pieces of the MUSCL stage kernels (both arithmetic profiles) and the host side.
#pragma once

#include <hip/hip_runtime.h>

namespace vh {

struct MusclDev {
  int n[3];
  int ng;
  int tiles_x, tiles_y, chunks_z, zchunk;
  int xcd_remap, nf;                  // nf: 5 fields, or 6 with the gas energy ρe
  long long ncell;
  double gamma, theta, small, dtdx;   // dtdx = T(dt)/T(dx), exact in T
  double eta1;                        // dual-energy selection threshold
};

// Julia's max / min on floats are IEEE 754-2019 maximum / minimum (NaN wins, −0 < +0).
template <typename T>
__device__ __forceinline__ T jmax(T x, T y) {
  return __builtin_elementwise_maximum(x, y);
}
template <typename T>
__device__ __forceinline__ T jmin(T x, T y) {
  return __builtin_elementwise_minimum(x, y);
}
__device__ __forceinline__ double absT(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float absT(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtT(float x) { return __builtin_sqrtf(x); }


namespace exact {
// One RK stage of the unsplit MUSCL step (profile-specific arithmetic). variant:
// 0 = 16×16 tiles, 1 = 32×8; nf = 5, or 6 with the gas energy.
void muscl_launch_stage(int dtype, int variant, int nf, bool stage2, bool rec, const MusclDev& G, hipStream_t s,
                        const void* Us, const void* U0, void* Uout, void* frec);
}
namespace fast {
void muscl_launch_stage(int dtype, int variant, int nf, bool stage2, bool rec, const MusclDev& G, hipStream_t s,
                        const void* Us, const void* U0, void* Uout, void* frec);
}

}  // namespace vh
