// Kernel launchers of one arithmetic profile. This file is compiled twice:
// VH_FAST_PROFILE=0 with -ffp-contract=off (exact, bit-identical to the reference's
// evaluation order) and VH_FAST_PROFILE=1 with FMA contraction (fast).
#include "profiles.h"
#include "stencil.h"
#include "blocks.h"

namespace vh {
namespace VH_PROFILE_NS {

// Tier-1 face batch: one lane per face (ext/VespaKAFluxExt.jl:23-35 contract).
template <typename T>
__global__ void __launch_bounds__(256) face_flux_kernel(const T* __restrict__ WL, const T* __restrict__ WR,
                                                        const int32_t* __restrict__ axis, long long nf, T g,
                                                        T igm1, T* __restrict__ F) {
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  Prim<T> L, R;
  const T* l = WL + 5 * f;
  const T* r = WR + 5 * f;
  L.r = l[0]; L.vx = l[1]; L.vy = l[2]; L.vz = l[3]; L.p = l[4];
  R.r = r[0]; R.vx = r[1]; R.vy = r[2]; R.vz = r[3]; R.p = r[4];
  const int d = axis[f];
  // constant axis per branch keeps the normal-component selection in registers
  const Cons<T> o = d <= 1 ? hllc(L, R, g, igm1, 1) : (d == 2 ? hllc(L, R, g, igm1, 2) : hllc(L, R, g, igm1, 3));
  T* out = F + 5 * f;
  out[0] = o.d; out[1] = o.mx; out[2] = o.my; out[3] = o.mz; out[4] = o.e;
}

template <typename T, int RANK, int TX, int TY, bool FLOOR64, bool PER, bool GRAV, bool PG = false>
static void stage_tp(bool stage2, int blocks, hipStream_t s, const void* Us, const void* U0, void* Uout,
                     const double* dtp, double* partials, const GridDev& G) {
  if (stage2)
    hipLaunchKernelGGL((rk_stage_kernel<T, RANK, TX, TY, true, FLOOR64, PER, GRAV, PG>), dim3(blocks), dim3(TX * TY),
                       0, s, (const T*)Us, (const T*)U0, (T*)Uout, dtp, partials, G);
  else
    hipLaunchKernelGGL((rk_stage_kernel<T, RANK, TX, TY, false, FLOOR64, PER, GRAV, PG>), dim3(blocks), dim3(TX * TY),
                       0, s, (const T*)Us, (const T*)U0, (T*)Uout, dtp, partials, G);
}

// Grids periodic on every axis run the kernel with the boundary logic compiled
// out; with self-gravity the gravity form runs (boundary logic kept); a grid
// with a ParentGhost side runs the form that reads their states (with or
// without gravity); the rest the general form.
template <typename T, int RANK, int TX, int TY, bool FLOOR64>
static void stage_t(bool stage2, int blocks, hipStream_t s, const void* Us, const void* U0, void* Uout,
                    const double* dtp, double* partials, const GridDev& G) {
  bool per = true, parent = false;
  for (int d = 0; d < RANK; ++d) {
    per = per && G.bc[2 * d] == BC_PERIODIC && G.bc[2 * d + 1] == BC_PERIODIC;
    parent = parent || G.bc[2 * d] == BC_PARENT || G.bc[2 * d + 1] == BC_PARENT;
  }
  if (G.phi && parent)
    stage_tp<T, RANK, TX, TY, FLOOR64, false, true, true>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
  else if (G.phi) stage_tp<T, RANK, TX, TY, FLOOR64, false, true>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
  else if (per) stage_tp<T, RANK, TX, TY, FLOOR64, true, false>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
  else if (parent)
    stage_tp<T, RANK, TX, TY, FLOOR64, false, false, true>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
  else stage_tp<T, RANK, TX, TY, FLOOR64, false, false>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
}

template <typename T, bool FLOOR64>
static void stage_dispatch(int rank, int variant, bool stage2, int blocks, hipStream_t s, const void* Us,
                           const void* U0, void* Uout, const double* dtp, double* partials, const GridDev& G) {
  if (rank == 3) {
    switch (variant) {
      case 0: stage_t<T, 3, 32, 8, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
      case 1: stage_t<T, 3, 64, 4, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
      case 2: stage_t<T, 3, 16, 16, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
      default: stage_t<T, 3, 64, 8, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
    }
  } else if (rank == 2) {
    switch (variant) {
      case 0: stage_t<T, 2, 32, 8, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
      case 1: stage_t<T, 2, 64, 4, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
      default: stage_t<T, 2, 16, 16, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
    }
  } else {
    switch (variant) {
      case 0: stage_t<T, 1, 256, 1, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
      default: stage_t<T, 1, 128, 1, FLOOR64>(stage2, blocks, s, Us, U0, Uout, dtp, partials, G); break;
    }
  }
}

void launch_stage(int dtype, int rank, int variant, bool stage2, int blocks, hipStream_t s, const void* Us,
                  const void* U0, void* Uout, const double* dtp, double* partials, const GridDev& G) {
  if (dtype == 0)
    stage_dispatch<double, false>(rank, variant, stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
  else
    stage_dispatch<float, true>(rank, variant, stage2, blocks, s, Us, U0, Uout, dtp, partials, G);
}

void launch_cfl(int dtype, int rank, int blocks, hipStream_t s, const void* U, const GridDev& G,
                double* partials) {
  if (dtype == 0) {
    if (rank == 3) hipLaunchKernelGGL((cfl_kernel<double, 3, false>), dim3(blocks), dim3(256), 0, s, (const double*)U, G, partials);
    else if (rank == 2) hipLaunchKernelGGL((cfl_kernel<double, 2, false>), dim3(blocks), dim3(256), 0, s, (const double*)U, G, partials);
    else hipLaunchKernelGGL((cfl_kernel<double, 1, false>), dim3(blocks), dim3(256), 0, s, (const double*)U, G, partials);
  } else {
    if (rank == 3) hipLaunchKernelGGL((cfl_kernel<float, 3, true>), dim3(blocks), dim3(256), 0, s, (const float*)U, G, partials);
    else if (rank == 2) hipLaunchKernelGGL((cfl_kernel<float, 2, true>), dim3(blocks), dim3(256), 0, s, (const float*)U, G, partials);
    else hipLaunchKernelGGL((cfl_kernel<float, 1, true>), dim3(blocks), dim3(256), 0, s, (const float*)U, G, partials);
  }
}

void launch_face_flux(int dtype, long long nf, hipStream_t s, const void* WL, const void* WR, const int32_t* axis,
                      double gamma, void* F) {
  const unsigned blocks = (unsigned)((nf + 255) / 256);
  const double igm1 = 1.0 / (gamma - 1.0);
  if (dtype == 0)
    hipLaunchKernelGGL((face_flux_kernel<double>), dim3(blocks), dim3(256), 0, s, (const double*)WL,
                       (const double*)WR, axis, nf, gamma, igm1, (double*)F);
  else
    hipLaunchKernelGGL((face_flux_kernel<float>), dim3(blocks), dim3(256), 0, s, (const float*)WL,
                       (const float*)WR, axis, nf, float(gamma), float(igm1), (float*)F);
}

// ---- batched blocks ----

void launch_link_flux(int dtype, int nfaces, hipStream_t s, const void* U, const BlocksDev& B, void* F) {
  if (nfaces <= 0) return;
  const unsigned blocks = (unsigned)((nfaces + 255) / 256);
  if (dtype == 0)
    hipLaunchKernelGGL((link_flux_kernel<double, false>), dim3(blocks), dim3(256), 0, s, (const double*)U, B,
                       (double*)F);
  else
    hipLaunchKernelGGL((link_flux_kernel<float, true>), dim3(blocks), dim3(256), 0, s, (const float*)U, B,
                       (float*)F);
}

template <typename T, int RANK, bool FLOOR64, bool GRAV, int CTX = 0, int CTY = 0>
static void block_stage_tg(bool stage2, int wgs, hipStream_t s, const void* Us, const void* U0, void* Uout,
                           const double* dtp, double* partials, const BlocksDev& B) {
  constexpr int NT = CTX ? CTX * CTY : BNT;
  if (stage2)
    hipLaunchKernelGGL((block_stage_kernel<T, RANK, true, FLOOR64, GRAV, CTX, CTY>), dim3(wgs), dim3(NT), 0, s,
                       (const T*)Us, (const T*)U0, (T*)Uout, dtp, partials, B);
  else
    hipLaunchKernelGGL((block_stage_kernel<T, RANK, false, FLOOR64, GRAV, CTX, CTY>), dim3(wgs), dim3(NT), 0, s,
                       (const T*)Us, (const T*)U0, (T*)Uout, dtp, partials, B);
}

// With self-gravity (B.gravg set) the stage adds the gravity source of the
// step's g field to every cell's flux sum (runtime-shape kernel). Tile class cls
// (block_shape_class) picks the compile-time shape of rank-3 tiles.
template <typename T, int RANK, bool FLOOR64>
static void block_stage_t(int cls, bool stage2, int wgs, hipStream_t s, const void* Us, const void* U0, void* Uout,
                          const double* dtp, double* partials, const BlocksDev& B) {
  if (B.gravg) {
    block_stage_tg<T, RANK, FLOOR64, true>(stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
    return;
  }
  if constexpr (RANK == 3) {
    switch (cls) {
      case 1: block_stage_tg<T, 3, FLOOR64, false, 32, 8>(stage2, wgs, s, Us, U0, Uout, dtp, partials, B); return;
      case 2: block_stage_tg<T, 3, FLOOR64, false, 16, 16>(stage2, wgs, s, Us, U0, Uout, dtp, partials, B); return;
      case 3: block_stage_tg<T, 3, FLOOR64, false, 24, 8>(stage2, wgs, s, Us, U0, Uout, dtp, partials, B); return;
      case 4: block_stage_tg<T, 3, FLOOR64, false, 8, 8>(stage2, wgs, s, Us, U0, Uout, dtp, partials, B); return;
      default: break;
    }
  }
  block_stage_tg<T, RANK, FLOOR64, false>(stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
}

void launch_block_stage(int dtype, int rank, int cls, bool stage2, int wgs, hipStream_t s, const void* Us,
                        const void* U0, void* Uout, const double* dtp, double* partials, const BlocksDev& B) {
  if (wgs <= 0) return;
  if (dtype == 0) {
    if (rank == 3) block_stage_t<double, 3, false>(cls, stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
    else if (rank == 2) block_stage_t<double, 2, false>(cls, stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
    else block_stage_t<double, 1, false>(cls, stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
  } else {
    if (rank == 3) block_stage_t<float, 3, true>(cls, stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
    else if (rank == 2) block_stage_t<float, 2, true>(cls, stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
    else block_stage_t<float, 1, true>(cls, stage2, wgs, s, Us, U0, Uout, dtp, partials, B);
  }
}

template <typename T, bool FLOOR64>
static void block_cfl_t(int rank, int wgs, hipStream_t s, const void* U, const BlocksDev& B, const int* blist,
                        const long long* prefix, int nb, double* partials) {
  if (rank == 3)
    hipLaunchKernelGGL((block_cfl_kernel<T, 3, FLOOR64>), dim3(wgs), dim3(256), 0, s, (const T*)U, B, blist, prefix,
                       nb, partials);
  else if (rank == 2)
    hipLaunchKernelGGL((block_cfl_kernel<T, 2, FLOOR64>), dim3(wgs), dim3(256), 0, s, (const T*)U, B, blist, prefix,
                       nb, partials);
  else
    hipLaunchKernelGGL((block_cfl_kernel<T, 1, FLOOR64>), dim3(wgs), dim3(256), 0, s, (const T*)U, B, blist, prefix,
                       nb, partials);
}

void launch_block_cfl(int dtype, int rank, int wgs, hipStream_t s, const void* U, const BlocksDev& B,
                      const int* blist, const long long* prefix, int nb, double* partials) {
  if (wgs <= 0 || nb <= 0) return;
  if (dtype == 0) block_cfl_t<double, false>(rank, wgs, s, U, B, blist, prefix, nb, partials);
  else block_cfl_t<float, true>(rank, wgs, s, U, B, blist, prefix, nb, partials);
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
