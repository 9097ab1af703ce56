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

enum : int { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTING = 2, BC_PARENT = 3 };

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
  void* bflux;     // boundary-flux register (5 Tf per boundary face slot) or null
  long long bbase[6];  // first slot of each boundary plane (axis 0 lo, axis 0 hi, axis 1 lo, ...)
  const void* phi; // gravitational potential (ncell Tf) when self-gravity is on, else null
  double gravG;    // gravitational constant
  int gper;        // bit d: φ periodic along axis d
  int pad_;
  const void* pg;  // ParentGhost states (5 Tf primitives per boundary face slot) or null
};

// ---- batched blocks (block-structured AMR) ----

// One block of the mesh: cells per axis, boundary conditions where no other
// block touches a face, offset of its first cell in the SoA arrays, geometry.
// Face slot s = 2*axis + side (side 0 = lo, 1 = hi).
struct BlockDev {
  int n[3];
  int bc[6];
  int level;
  int haslinks;    // 1 if any cell of the block touches a link face
  int repoff[6];   // offset of the slot's representative map in repmap, -1 if no block touches it
  int lcoff[6];    // offset of the slot's link-cell map in lcmap, -1 if no link face is on it
  long long off;   // first cell of the block in the leaf order
  double w[3];
  double area[3];
  double V;
};

// A tile of work: block, x/y origin and shape of the tile, z range.
struct TileDev {
  int block, x0, y0, z0, z1, tx, ty, pad;
};

// A face between two different blocks (emitted once: same level from the hi
// side, coarse-fine from the fine side, with the emitting cell's area). nL/nR:
// lo and hi representative neighbours of iL and iR along the axis (>= 0: cell;
// < 0: ghost of the cell itself with boundary code -1-value).
struct LinkFaceDev {
  int iL, iR;
  int nL[2], nR[2];
  int axis;
  int pad;
  double area;
};

// Gather record of a cell that touches link faces (16 bytes, packed; read it
// through the accessors below).
// type(s): 0 flux computed by the tile (or none), 1 one link face emitted by
// this cell at its own walk position, 2 a group of link faces emitted by cells
// of another block. start(s)/count(s): the slot's faces in slot_faces (walk
// order); the slots' lists follow each other from `start`, so start(s) is
// start + the counts of the slots before s. grp(0..ngrp): the type-2 slots in
// increasing emitting-block order; the first npre come from blocks that precede
// this one. finer: bit s set when slot s is a group of finer cells (the coarse
// side of a flux register).
//   tc: bits 2s..2s+1 type(s), bits 12+3s..14+3s count(s) (at most 4)
//   gr: bits 3q..3q+2 grp(q) (7: none), bits 18-20 npre, 21-23 ngrp, 24-29 finer
struct LinkCellDev {
  int start;
  unsigned tc, gr;
  int pad;
};
__host__ __device__ inline int lc_type(const LinkCellDev& R, int s) { return (int)((R.tc >> (2 * s)) & 3u); }
__host__ __device__ inline int lc_count(const LinkCellDev& R, int s) { return (int)((R.tc >> (12 + 3 * s)) & 7u); }
__host__ __device__ inline int lc_start(const LinkCellDev& R, int s) {
  int o = R.start;
  for (int q = 0; q < s; ++q) o += lc_count(R, q);
  return o;
}
__host__ __device__ inline int lc_grp(const LinkCellDev& R, int q) {
  const int v = (int)((R.gr >> (3 * q)) & 7u);
  return v == 7 ? -1 : v;
}
__host__ __device__ inline int lc_npre(const LinkCellDev& R) { return (int)((R.gr >> 18) & 7u); }
__host__ __device__ inline int lc_ngrp(const LinkCellDev& R) { return (int)((R.gr >> 21) & 7u); }
__host__ __device__ inline int lc_finer(const LinkCellDev& R) { return (int)((R.gr >> 24) & 63u); }

// Everything a block kernel needs, passed by value.
struct BlocksDev {
  const BlockDev* blocks;
  const TileDev* tiles;
  const int* repmap;
  const int* lcmap;
  const LinkCellDev* linkcells;
  const int* slot_faces;
  const LinkFaceDev* faces;
  const void* fluxbuf;     // 5 values per link face (AoS)
  const void* gravg;       // g = −∇φ of the step (3 SoA fields of the state type) with self-gravity, else null
  long long ncell;
  double gamma, igm1;
  double gravG;            // gravitational constant (self-gravity)
  int ntiles;
  int nfaces;
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

// Streaming forms for values read or written once per launch (the RK base state
// and the stage output): non-temporal, so that they do not push out of the L2 the
// lines that neighbouring tiles re-read as their halo (+2-3% for the uniform and
// block stage kernels at 256³ / 3000 blocks).
#ifndef VH_NT_STREAM
#define VH_NT_STREAM 1
#endif
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#if VH_NT_STREAM
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
template <typename T>
__device__ __forceinline__ void st_stream(T v, T* p) {
#if VH_NT_STREAM
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

template <typename T>
__device__ __forceinline__ Cons<T> load_cons_stream(const T* __restrict__ U, long long n, long long c) {
#if VH_NT_STREAM
  Cons<T> u;
  u.d = __builtin_nontemporal_load(U + c);
  u.mx = __builtin_nontemporal_load(U + n + c);
  u.my = __builtin_nontemporal_load(U + 2 * n + c);
  u.mz = __builtin_nontemporal_load(U + 3 * n + c);
  u.e = __builtin_nontemporal_load(U + 4 * n + c);
  return u;
#else
  return load_cons(U, n, c);
#endif
}

template <typename T>
__device__ __forceinline__ void store_cons_stream(T* __restrict__ U, long long n, long long c, const Cons<T>& u) {
#if VH_NT_STREAM
  __builtin_nontemporal_store(u.d, U + c);
  __builtin_nontemporal_store(u.mx, U + n + c);
  __builtin_nontemporal_store(u.my, U + 2 * n + c);
  __builtin_nontemporal_store(u.mz, U + 3 * n + c);
  __builtin_nontemporal_store(u.e, U + 4 * n + c);
#else
  store_cons(U, n, c, u);
#endif
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
