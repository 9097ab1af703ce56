// Types and helpers shared by both arithmetic profiles of the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef VH_FAST_PROFILE
#define VH_FAST_PROFILE 0
#endif

// Each arithmetic profile lives in its own namespace so that the two builds of
// the same templates never collide at link time.
#if VH_FAST_PROFILE
#define VH_PROFILE_NS fast
#else
#define VH_PROFILE_NS exact
#endif

namespace vh {

enum : int { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTING = 2 };

template <typename T>
struct Prim {
  T r, vx, vy, vz, p;
};

template <typename T>
struct Cons {
  T d, mx, my, mz, e;
};

// Uniform-grid description passed by value to every kernel.
struct GridDev {
  int n[3];        // cells per axis (axes beyond rank: 1)
  int bc[6];       // [2*axis + side], side 0 = lo, 1 = hi
  int rank;
  int zchunk;      // planes per workgroup along z
  int tiles_x;     // tiles along x
  int tiles_y;     // tiles along y
  int chunks_z;    // z chunks
  int xcd_remap;   // 1: contiguous tile ranges per XCD
  long long ncell;
  double gamma;
  double igm1;     // 1/(γ−1)
  double V;        // cell volume (geometry precision)
  double w[3];     // cell widths
  double area[3];  // face areas
};

__device__ __forceinline__ double vmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }

// Map a (possibly out-of-range) coordinate on one axis to the cell that
// supplies its state, following neighbor() (lib/RefMesh/src/RefMesh.jl:147-161).
// Returns the source coordinate and sets `gh` to the boundary side whose ghost
// applies (0: none, 1: lo, 2: hi).
__device__ __forceinline__ int resolve(int i, int n, int bclo, int& gh) {
  gh = 0;
  if (i >= 0 && i < n) return i;
  if (bclo == BC_PERIODIC) {
    int r = i % n;
    return r < 0 ? r + n : r;
  }
  if (i < 0) {
    gh = 1;
    return 0;
  }
  gh = 2;
  return n - 1;
}

template <typename T>
__device__ __forceinline__ Cons<T> load_cons(const T* __restrict__ U, long long n, long long c) {
  Cons<T> u;
  u.d = U[c];
  u.mx = U[n + c];
  u.my = U[2 * n + c];
  u.mz = U[3 * n + c];
  u.e = U[4 * n + c];
  return u;
}

template <typename T>
__device__ __forceinline__ void store_cons(T* __restrict__ U, long long n, long long c, const Cons<T>& u) {
  U[c] = u.d;
  U[n + c] = u.mx;
  U[2 * n + c] = u.my;
  U[3 * n + c] = u.mz;
  U[4 * n + c] = u.e;
}

// Workgroup maximum of a double (wave64 butterfly, then LDS across waves).
template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = vmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r = vmax(r, red[i]);
  return r;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r += red[i];
  return r;
}

}  // namespace vh
