// Host-side declarations of the kernel launchers built once per arithmetic
// profile (stages.hip compiled with VH_FAST = 0 and 1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace vh {

struct TileShape {
  int tx, ty;
};

// Tile variants per rank (index = the variant id of vh_grid_set_launch).
static const TileShape kTiles3[] = {{32, 8}, {64, 4}, {16, 16}, {64, 8}};
static const TileShape kTiles2[] = {{32, 8}, {64, 4}, {16, 16}};
static const TileShape kTiles1[] = {{256, 1}, {128, 1}};

inline int tile_count_for_rank(int rank) { return rank == 3 ? 4 : (rank == 2 ? 3 : 2); }
inline TileShape tile_for(int rank, int variant) {
  if (rank == 3) return kTiles3[variant];
  if (rank == 2) return kTiles2[variant];
  return kTiles1[variant];
}

// Tile shapes the rank-3 block stage kernel is compiled for (class 1..4; one
// lane per cell). Tiles of any other shape, of rank 1 and 2 meshes and of meshes
// with self-gravity run the runtime-shape kernel (class 0, 256 lanes).
static const TileShape kBlockShapes[] = {{32, 8}, {16, 16}, {24, 8}, {8, 8}};
constexpr int kBlockClasses = 5;
inline int block_shape_class(int rank, int tx, int ty) {
  if (rank != 3) return 0;
  for (int c = 0; c < 4; ++c)
    if (kBlockShapes[c].tx == tx && kBlockShapes[c].ty == ty) return c + 1;
  return 0;
}
inline int block_class_lanes(int cls) { return cls ? kBlockShapes[cls - 1].tx * kBlockShapes[cls - 1].ty : 256; }

#define VH_DECLARE_PROFILE(ns)                                                                      \
  namespace ns {                                                                                    \
  /* dtype 0: double state; 1: float state (pressure floor in double). */                          \
  void launch_stage(int dtype, int rank, int variant, bool stage2, int blocks, hipStream_t s,       \
                    const void* Us, const void* U0, void* Uout, const double* dtp, double* partials, \
                    const GridDev& G);                                                              \
  void launch_cfl(int dtype, int rank, int blocks, hipStream_t s, const void* U, const GridDev& G,  \
                  double* partials);                                                                \
  void launch_face_flux(int dtype, long long nf, hipStream_t s, const void* WL, const void* WR,     \
                        const int32_t* axis, double gamma, void* F);                                \
  void launch_link_flux(int dtype, int nfaces, hipStream_t s, const void* U, const BlocksDev& B,    \
                        void* F);                                                                   \
  void launch_block_stage(int dtype, int rank, int cls, bool stage2, int wgs, hipStream_t s,         \
                          const void* Us, const void* U0, void* Uout, const double* dtp,            \
                          double* partials, const BlocksDev& B);                                    \
  void launch_block_cfl(int dtype, int rank, int wgs, hipStream_t s, const void* U,                 \
                        const BlocksDev& B, const int* blist, const long long* prefix, int nb,      \
                        double* partials);                                                          \
  }

VH_DECLARE_PROFILE(exact)
VH_DECLARE_PROFILE(fast)

}  // namespace vh
