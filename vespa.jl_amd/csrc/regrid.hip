// Device side of the block-mesh regrid (gfx950).
//
//   indicator_kernel  density_gradient_indicator (src/driver.jl:400-414) for every
//                     leaf: the largest |ρn − ρc| / max(ρc, ρn, eps) over the
//                     representative neighbours across the 2·rank faces that are
//                     interior (a domain boundary is skipped; a periodic side wraps
//                     inside the block), with the per-block maximum, which is all
//                     the host needs to pick the blocks to refine;
//   remap_kernel      the conservative transfer of the state onto a new block
//                     mesh (prolong! by injection, lib/RefMesh/src/RefMesh.jl:284-302):
//                     every cell of a new block takes the conserved values of the
//                     old cell that contains it, so Σ U·V is unchanged (nf
//                     fields: the five conserved ones, or φ alone).
//
// One wavefront per block; each lane walks the block's cells with a stride of 64.
// Both kernels read the state once and write their output once: HBM-bound
// integer/gather work, launched once per regrid.
#include <float.h>

#include "../../include/vespahip.h"
#include "host.h"

namespace vh {
namespace {

constexpr int kRegridLanes = 64;

// Representative neighbour of cell (i, j, k) of block bk across face (d, side),
// following the block topology: the next cell in the block, the wrapped cell on
// a periodic side, the neighbour recorded in the face's representative map, or
// -1 at a domain boundary.
__device__ __forceinline__ long long neighbor_cell(const BlockDev& bk, const int* __restrict__ repmap, int d, int side,
                                                   int i, int j, int k) {
  const int n = bk.n[d];
  const int p = (d == 0 ? i : (d == 1 ? j : k)) + (side ? 1 : -1);
  int q = p;
  if (p < 0 || p >= n) {
    if (bk.bc[2 * d + side] != BC_PERIODIC) {
      const int ro = bk.repoff[2 * d + side];
      if (ro < 0) return -1;
      const int pi = d == 0 ? k * bk.n[1] + j : (d == 1 ? k * bk.n[0] + i : j * bk.n[0] + i);
      return repmap[ro + pi];
    }
    q = p < 0 ? n - 1 : 0;
  }
  const int ii = d == 0 ? q : i, jj = d == 1 ? q : j, kk = d == 2 ? q : k;
  return bk.off + ((long long)kk * bk.n[1] + jj) * bk.n[0] + ii;
}

__device__ __forceinline__ double abs_t(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float abs_t(float x) { return __builtin_fabsf(x); }

// Julia's max: NaN wins, then the larger value, +0.0 over -0.0.
__device__ __forceinline__ double jmax(double x, double y) {
  if (x != x) return x;
  if (y != y) return y;
  if (y > x) return y;
  if (x > y) return x;
  return __builtin_signbit(y) ? x : y;
}

// The difference is taken at the state's precision and the ratio in Float64,
// as Julia promotes abs(ρn − ρc) / max(ρc, ρn, eps()).
template <typename T>
__device__ __forceinline__ double jump(T rc, T rn) {
  const double num = (double)abs_t(rn - rc);
  const double den = jmax(jmax((double)rc, (double)rn), DBL_EPSILON);
  return num / den;
}

template <typename T>
__global__ __launch_bounds__(kRegridLanes) void indicator_kernel(const BlockDev* __restrict__ blocks,
                                                                 const int* __restrict__ repmap, int rank,
                                                                 const T* __restrict__ rho, double* __restrict__ cell_g,
                                                                 double* __restrict__ block_max) {
  const BlockDev& bk = blocks[blockIdx.x];
  const int nx = bk.n[0], ny = bk.n[1];
  const long long nc = (long long)nx * ny * bk.n[2];
  double gb = 0.0;
  for (long long c = threadIdx.x; c < nc; c += kRegridLanes) {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long long)nx * ny));
    const T rc = rho[bk.off + c];
    double g = 0.0;
    for (int d = 0; d < rank; ++d)
      for (int side = 0; side < 2; ++side) {
        const long long nb = neighbor_cell(bk, repmap, d, side, i, j, k);
        if (nb >= 0) g = jmax(g, jump(rc, rho[nb]));
      }
    if (cell_g) cell_g[bk.off + c] = g;
    gb = jmax(gb, g);
  }
  for (int s = kRegridLanes / 2; s > 0; s >>= 1) gb = jmax(gb, __shfl_xor(gb, s, kRegridLanes));
  if (threadIdx.x == 0) block_max[blockIdx.x] = gb;
}

template <typename T>
__global__ __launch_bounds__(kRegridLanes) void remap_kernel(const BlockDev* __restrict__ dst_blocks,
                                                             const BlockDev* __restrict__ src_blocks,
                                                             const RemapDev* __restrict__ map, const T* __restrict__ Us,
                                                             long long ns, T* __restrict__ Ud, long long nd, int nf) {
  const BlockDev& db = dst_blocks[blockIdx.x];
  const RemapDev r = map[blockIdx.x];
  const BlockDev& sb = src_blocks[r.src];
  const int nx = db.n[0], ny = db.n[1];
  const long long nc = (long long)nx * ny * db.n[2];
  const int snx = sb.n[0], sny = sb.n[1];
  for (long long c = threadIdx.x; c < nc; c += kRegridLanes) {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long long)nx * ny));
    const long long s = sb.off + ((long long)(r.o[2] + (k >> r.shift)) * sny + (r.o[1] + (j >> r.shift))) * snx +
                        (r.o[0] + (i >> r.shift));
    for (int f = 0; f < nf; ++f) Ud[f * nd + db.off + c] = Us[f * ns + s];
  }
}

}  // namespace

void regrid_indicator(hipStream_t s, int dtype, int nblocks, const BlockDev* blocks, const int* repmap, int rank,
                      const void* rho, double* cell_g, double* block_max) {
  if (nblocks <= 0) return;
  if (dtype == VH_F64)
    indicator_kernel<double><<<nblocks, kRegridLanes, 0, s>>>(blocks, repmap, rank, (const double*)rho, cell_g,
                                                              block_max);
  else
    indicator_kernel<float><<<nblocks, kRegridLanes, 0, s>>>(blocks, repmap, rank, (const float*)rho, cell_g,
                                                             block_max);
}

void regrid_remap(hipStream_t s, int dtype, int ndst, const BlockDev* dst_blocks, const BlockDev* src_blocks,
                  const RemapDev* map, const void* Us, long long ns, void* Ud, long long nd, int nf) {
  if (ndst <= 0) return;
  if (dtype == VH_F64)
    remap_kernel<double><<<ndst, kRegridLanes, 0, s>>>(dst_blocks, src_blocks, map, (const double*)Us, ns,
                                                       (double*)Ud, nd, nf);
  else
    remap_kernel<float><<<ndst, kRegridLanes, 0, s>>>(dst_blocks, src_blocks, map, (const float*)Us, ns, (float*)Ud,
                                                      nd, nf);
}

}  // namespace vh
