// Self-gravity on a uniform grid (gfx950): the Poisson solve ∇²φ = 4πGρ by
// conjugate gradients on A = −∇²·V, and g = −∇φ.
//
// The reference solves on the composite leaf mesh with a CSR adjacency built on
// the host (src/gravity.jl:248-336, ext/VespaKAGravityExt.jl:21-63). On a uniform
// grid that adjacency is the 2·rank-point stencil, so the matrix-vector product
// here is a stencil gather (one lane per cell, no index arrays), summed in
// Float64 in the row order the CSR build produces — lo neighbours that precede
// the cell (z, y, x), the cell's hi neighbours (x, y, z), then lo neighbours
// wrapped by a periodic φ boundary (x, y, z) — so each product matches the
// reference's bit for bit. Non-periodic φ faces contribute nothing (homogeneous
// Neumann, src/gravity.jl:124-128).
//
// One CG iteration is five stream-ordered launches (matvec + p·Ap partials, α,
// update + r·r partials, β, direction) that read a device state block; once the
// residual drops below the tolerance (or p·Ap ≤ 0, or maxiter) the state is
// marked done and every later launch returns at once, so the host enqueues
// iterations in chunks and synchronises once per chunk. Dot products are
// Float64 (src/gravity.jl:338-343); reductions go through per-workgroup partials
// summed in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "gravity.h"
#include "host.h"

namespace vh {

constexpr int GNT = 256;
constexpr double kFourPi = 4.0 * 3.141592653589793;

__device__ __forceinline__ double gsum_block(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < GNT / 64; ++i) r += red[i];
  return r;
}

// Coordinates of cell c and the neighbour along axis d on side s (0 lo, 1 hi):
// returns false for a non-periodic φ boundary.
__device__ __forceinline__ bool gnb(const GridDev& G, int gper, long long c, const int* co, int d, int s,
                                    long long& nb) {
  const long long st = d == 0 ? 1 : (d == 1 ? (long long)G.n[0] : (long long)G.n[0] * G.n[1]);
  const int n = G.n[d];
  if (s == 0) {
    if (co[d] > 0) { nb = c - st; return true; }
    if (gper >> d & 1) { nb = c + (long long)(n - 1) * st; return true; }
    return false;
  }
  if (co[d] + 1 < n) { nb = c + st; return true; }
  if (gper >> d & 1) { nb = c - (long long)(n - 1) * st; return true; }
  return false;
}

__device__ __forceinline__ void gcoords(const GridDev& G, long long c, int* co) {
  co[0] = (int)(c % G.n[0]);
  const long long q = c / G.n[0];
  co[1] = (int)(q % G.n[1]);
  co[2] = (int)(q / G.n[1]);
}

struct CgState {
  double rsold, alpha, beta, relres, bnorm;
  double iters;      // iterations taken (as double)
  int done;          // 1: converged, stalled or out of iterations
  int pad;
};
static_assert(sizeof(CgState) <= kGravStateBytes, "CG state block");

// out = A·x (Float64 sum, stored at T); partial (x·out) per workgroup when
// partials is non-null. Skipped when the state is done.
template <typename T>
__global__ void __launch_bounds__(GNT) lap_kernel(const T* __restrict__ x, T* __restrict__ out, GridDev G, int gper,
                                                  double c0, double c1, double c2, double* __restrict__ partials,
                                                  const CgState* __restrict__ st) {
  __shared__ double red[GNT / 64];
  if (st && st->done) return;
  const double coef[3] = {c0, c1, c2};
  double dot = 0.0;
  for (long long c = (long long)blockIdx.x * GNT + threadIdx.x; c < G.ncell; c += (long long)gridDim.x * GNT) {
    int co[3];
    gcoords(G, c, co);
    const double xc = (double)x[c];
    double s = 0.0;
    long long nb;
    for (int d = G.rank - 1; d >= 0; --d)
      if (co[d] > 0 && gnb(G, gper, c, co, d, 0, nb)) s += coef[d] * (xc - (double)x[nb]);
    for (int d = 0; d < G.rank; ++d)
      if (gnb(G, gper, c, co, d, 1, nb)) s += coef[d] * (xc - (double)x[nb]);
    for (int d = 0; d < G.rank; ++d)
      if (co[d] == 0 && gnb(G, gper, c, co, d, 0, nb)) s += coef[d] * (xc - (double)x[nb]);
    const T o = (T)s;
    out[c] = o;
    dot += xc * (double)o;
  }
  if (partials) {
    dot = gsum_block(dot, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = dot;
  }
}

// out = A·x on the CSR operator: per row Σ coef·(x_c − x_nb) in Float64 in row
// order (_lap_csr_cpu!, src/gravity.jl:326-336), stored at T; partial (x·out)
// per workgroup when partials is non-null. Skipped when the state is done.
template <typename T>
__global__ void __launch_bounds__(GNT) csr_lap_kernel(const T* __restrict__ x, T* __restrict__ out,
                                                      const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                                      const double* __restrict__ coef, long long n,
                                                      double* __restrict__ partials, const CgState* st) {
  __shared__ double red[GNT / 64];
  if (st && st->done) return;
  double dot = 0.0;
  for (long long c = (long long)blockIdx.x * GNT + threadIdx.x; c < n; c += (long long)gridDim.x * GNT) {
    const double xc = (double)x[c];
    double s = 0.0;
    for (int k = rowptr[c]; k < rowptr[c + 1]; ++k) s += coef[k] * (xc - (double)x[colidx[k]]);
    const T o = (T)s;
    out[c] = o;
    dot += xc * (double)o;
  }
  if (partials) {
    dot = gsum_block(dot, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = dot;
  }
}

// Partials of Σ a·b (b null: Σ a·w with a constant weight w).
template <typename T>
__global__ void __launch_bounds__(GNT) dot_kernel(const T* __restrict__ a, const T* __restrict__ b, double w,
                                                  long long n, double* __restrict__ partials) {
  __shared__ double red[GNT / 64];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT)
    s += b ? (double)a[i] * (double)b[i] : w * (double)a[i];
  s = gsum_block(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// b = T(−4πG(ρ − ρ̄)·V) with ρ̄ = sums[0]/sums[1] (0 without the projection);
// V per cell from vols when given, else the constant vol.
template <typename T>
__global__ void __launch_bounds__(GNT) rhs_kernel(const T* __restrict__ rho, T* __restrict__ b, long long n, double G,
                                                  double vol, const double* __restrict__ vols,
                                                  const double* __restrict__ rbar) {
  const double rb = rbar ? *rbar : 0.0;
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT)
    b[i] = (T)(-kFourPi * G * ((double)rho[i] - rb) * (vols ? vols[i] : vol) * 1.0);
}

// Partials of Σ V_i·a_i with per-cell volumes.
template <typename T>
__global__ void __launch_bounds__(GNT) vdot_kernel(const T* __restrict__ a, const double* __restrict__ vols,
                                                   long long n, double* __restrict__ partials) {
  __shared__ double red[GNT / 64];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT)
    s += vols[i] * (double)a[i];
  s = gsum_block(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// g_d = −(φ_hi − φ_lo)/T(d_lo + d_hi) per cell from the CSR stencil (_gax,
// src/gravity.jl:403-410), 3 SoA fields; axes beyond the rank are 0.
template <typename T>
__global__ void __launch_bounds__(GNT) csr_accel_kernel(const T* __restrict__ phi, T* __restrict__ g,
                                                        const int* __restrict__ glo, const int* __restrict__ ghi,
                                                        const double* __restrict__ gsum, long long n, int rank) {
  for (long long c = (long long)blockIdx.x * GNT + threadIdx.x; c < n; c += (long long)gridDim.x * GNT)
    for (int d = 0; d < 3; ++d) {
      if (d >= rank) {
        g[d * n + c] = T(0);
        continue;
      }
      const T wlo = phi[glo[d * n + c]], whi = phi[ghi[d * n + c]];
      g[d * n + c] = -(whi - wlo) / T(gsum[d * n + c]);
    }
}

// Single-workgroup sum of np partials into *out (fixed order per lane, then the tree).
__global__ void __launch_bounds__(GNT) psum_kernel(const double* __restrict__ partials, int np, double* out) {
  __shared__ double red[GNT / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += GNT) s += partials[i];
  s = gsum_block(s, red);
  if (threadIdx.x == 0) *out = s;
}

// scal[0] = Σ ρV, scal[1] = Σ V (summed cell by cell on the host, as the
// reference does)  →  scal[2] = ρ̄ = scal[0]/scal[1].
__global__ void ratio_kernel(double* scal) { scal[2] = scal[0] / scal[1]; }

// r = b − Ap; p = r.
template <typename T>
__global__ void __launch_bounds__(GNT) cg_init_kernel(const T* __restrict__ b, const T* __restrict__ Ap, T* r, T* p,
                                                      long long n) {
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT) {
    const T v = b[i] - Ap[i];
    r[i] = v;
    p[i] = v;
  }
}

// scal[3] = r·r, scal[4] = b·b  →  state.
__global__ void cg_start_kernel(const double* scal, CgState* st, long long maxiter) {
  st->rsold = scal[3];
  const double bn = sqrt(scal[4]);
  st->bnorm = bn == 0.0 ? 1.0 : bn;
  st->relres = sqrt(st->rsold) / st->bnorm;
  st->iters = 0.0;
  st->done = maxiter <= 0 ? 1 : 0;
}

// Iteration k begins: α = rsold / p·Ap, or stop when p·Ap ≤ 0.
__global__ void __launch_bounds__(GNT) cg_alpha_kernel(const double* __restrict__ partials, int np, CgState* st) {
  __shared__ double red[GNT / 64];
  if (st->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += GNT) s += partials[i];
  s = gsum_block(s, red);
  if (threadIdx.x == 0) {
    st->iters += 1.0;
    if (s <= 0.0) st->done = 1;
    else st->alpha = st->rsold / s;
  }
}

// φ += α p; r −= α Ap (at T); partial r·r.
template <typename T>
__global__ void __launch_bounds__(GNT) cg_update_kernel(T* __restrict__ phi, T* __restrict__ r, const T* __restrict__ p,
                                                        const T* __restrict__ Ap, long long n,
                                                        double* __restrict__ partials, const CgState* __restrict__ st) {
  __shared__ double red[GNT / 64];
  if (st->done) return;
  const T a = (T)st->alpha;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT) {
    phi[i] += a * p[i];
    const T ri = r[i] - a * Ap[i];
    r[i] = ri;
    s += (double)ri * (double)ri;
  }
  s = gsum_block(s, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// relres = √(r·r)/‖b‖; stop at the tolerance or at maxiter; else β = rsnew/rsold.
__global__ void __launch_bounds__(GNT) cg_beta_kernel(const double* __restrict__ partials, int np, CgState* st,
                                                      double tol, long long maxiter) {
  __shared__ double red[GNT / 64];
  if (st->done) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += GNT) s += partials[i];
  s = gsum_block(s, red);
  if (threadIdx.x == 0) {
    st->relres = sqrt(s) / st->bnorm;
    if (st->relres <= tol || st->iters >= (double)maxiter) {
      st->done = 1;
    } else {
      st->beta = s / st->rsold;
      st->rsold = s;
    }
  }
}

// p = r + β p (at T).
template <typename T>
__global__ void __launch_bounds__(GNT) cg_dir_kernel(const T* __restrict__ r, T* __restrict__ p, long long n,
                                                     const CgState* __restrict__ st) {
  if (st->done) return;
  const T b = (T)st->beta;
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT)
    p[i] = r[i] + b * p[i];
}

// φ −= T(Σ V·φ / Σ V) (scal[5] = Σ Vφ, scal[1] = Σ V).
template <typename T>
__global__ void __launch_bounds__(GNT) sub_mean_kernel(T* __restrict__ phi, long long n, const double* scal) {
  const T mu = (T)(scal[5] / scal[1]);
  for (long long i = (long long)blockIdx.x * GNT + threadIdx.x; i < n; i += (long long)gridDim.x * GNT)
    phi[i] -= mu;
}

// g_d = −(φ_hi − φ_lo)/T(d_lo + d_hi) per cell (src/gravity.jl:403-410), 3 SoA fields.
template <typename T>
__global__ void __launch_bounds__(GNT) accel_kernel(const T* __restrict__ phi, T* __restrict__ g, GridDev G, int gper) {
  for (long long c = (long long)blockIdx.x * GNT + threadIdx.x; c < G.ncell; c += (long long)gridDim.x * GNT) {
    int co[3];
    gcoords(G, c, co);
    for (int d = 0; d < 3; ++d)
      g[d * G.ncell + c] = d < G.rank ? VH_PROFILE_NS::grav_axis<T>(phi, G, gper, c, co, d) : T(0);
  }
}

// ---- host side ----

static int grid_blocks(long long n) {
  long long b = (n + GNT - 1) / GNT;
  if (b > 4096) b = 4096;
  return (int)(b < 1 ? 1 : b);
}

long long grav_work_values(long long ncell) { return 5 * ncell; }

void grav_apply(hipStream_t s, int dtype, const GridDev& G, int gper, const void* x, void* out) {
  double coef[3] = {0, 0, 0};
  for (int d = 0; d < G.rank; ++d) coef[d] = G.area[d] / (0.5 * (G.w[d] + G.w[d]));
  const int nb = grid_blocks(G.ncell);
  if (dtype == 0)
    hipLaunchKernelGGL(lap_kernel<double>, dim3(nb), dim3(GNT), 0, s, (const double*)x, (double*)out, G, gper,
                       coef[0], coef[1], coef[2], (double*)nullptr, (const CgState*)nullptr);
  else
    hipLaunchKernelGGL(lap_kernel<float>, dim3(nb), dim3(GNT), 0, s, (const float*)x, (float*)out, G, gper, coef[0],
                       coef[1], coef[2], (double*)nullptr, (const CgState*)nullptr);
}

void grav_accel(hipStream_t s, int dtype, const GridDev& G, int gper, const void* phi, void* g) {
  const int nb = grid_blocks(G.ncell);
  if (dtype == 0)
    hipLaunchKernelGGL(accel_kernel<double>, dim3(nb), dim3(GNT), 0, s, (const double*)phi, (double*)g, G, gper);
  else
    hipLaunchKernelGGL(accel_kernel<float>, dim3(nb), dim3(GNT), 0, s, (const float*)phi, (float*)g, G, gper);
}

// The two operator forms of the CG driver: the uniform stencil (constant cell
// volume) and the CSR composite operator of a block mesh (per-cell volumes).
struct UniformOp {
  const GridDev* G;
  int gper;
  double coef[3];
  long long n() const { return G->ncell; }
  const double* vols() const { return nullptr; }
  double vol() const { return G->V; }
  template <typename T>
  void lap(hipStream_t s, int nb, const T* x, T* out, double* partials, const CgState* st) const {
    hipLaunchKernelGGL(lap_kernel<T>, dim3(nb), dim3(GNT), 0, s, x, out, *G, gper, coef[0], coef[1], coef[2], partials,
                       st);
  }
};
struct CsrOp {
  const GravCsr* A;
  long long n() const { return A->n; }
  const double* vols() const { return A->vol; }
  double vol() const { return 0.0; }
  template <typename T>
  void lap(hipStream_t s, int nb, const T* x, T* out, double* partials, const CgState* st) const {
    hipLaunchKernelGGL(csr_lap_kernel<T>, dim3(nb), dim3(GNT), 0, s, x, out, A->rowptr, A->colidx, A->coef, A->n,
                       partials, st);
  }
};

// Σ V·a into *out (constant or per-cell volume).
template <typename T, typename Op>
static void vsum(hipStream_t s, int nb, const Op& op, const T* a, double* partials, double* out) {
  if (op.vols())
    hipLaunchKernelGGL(vdot_kernel<T>, dim3(nb), dim3(GNT), 0, s, a, op.vols(), op.n(), partials);
  else
    hipLaunchKernelGGL(dot_kernel<T>, dim3(nb), dim3(GNT), 0, s, a, (const T*)nullptr, op.vol(), op.n(), partials);
  hipLaunchKernelGGL(psum_kernel, dim3(1), dim3(GNT), 0, s, partials, nb, out);
}

template <typename T, typename Op>
static int solve_t(hipStream_t s, const Op& op, const GravParams& P, GravBuffers& B, long long* iters,
                   double* relres) {
  const long long n = op.n();
  const int nb = grid_blocks(n);
  T* phi = (T*)B.phi;
  T* r = (T*)B.work;
  T* p = r + n;
  T* Ap = r + 2 * n;
  T* b = r + 3 * n;
  const T* rho = (const T*)B.rho;
  CgState* st = (CgState*)B.state;
  double* scal = B.scal;
  // right-hand side (src/gravity.jl:230-246)
  if (P.project_mean) {
    vsum<T>(s, nb, op, rho, B.partials, scal + 0);
    hipLaunchKernelGGL(ratio_kernel, dim3(1), dim3(1), 0, s, scal);
  }
  hipLaunchKernelGGL(rhs_kernel<T>, dim3(nb), dim3(GNT), 0, s, rho, b, n, P.G, op.vol(), op.vols(),
                     P.project_mean ? (const double*)(scal + 2) : (const double*)nullptr);
  // r = b − Aφ, p = r (src/gravity.jl:280-289)
  op.template lap<T>(s, nb, (const T*)phi, Ap, (double*)nullptr, (const CgState*)nullptr);
  hipLaunchKernelGGL(cg_init_kernel<T>, dim3(nb), dim3(GNT), 0, s, (const T*)b, (const T*)Ap, r, p, n);
  hipLaunchKernelGGL(dot_kernel<T>, dim3(nb), dim3(GNT), 0, s, (const T*)r, (const T*)r, 0.0, n, B.partials);
  hipLaunchKernelGGL(psum_kernel, dim3(1), dim3(GNT), 0, s, B.partials, nb, scal + 3);
  hipLaunchKernelGGL(dot_kernel<T>, dim3(nb), dim3(GNT), 0, s, (const T*)b, (const T*)b, 0.0, n, B.partials);
  hipLaunchKernelGGL(psum_kernel, dim3(1), dim3(GNT), 0, s, B.partials, nb, scal + 4);
  hipLaunchKernelGGL(cg_start_kernel, dim3(1), dim3(1), 0, s, (const double*)scal, st, P.maxiter);
  // iterations, in chunks between host checks
  CgState h{};
  for (long long done_it = 0; done_it < P.maxiter;) {
    const long long chunk = std::min<long long>(P.chunk, P.maxiter - done_it);
    for (long long k = 0; k < chunk; ++k) {
      op.template lap<T>(s, nb, (const T*)p, Ap, B.partials, (const CgState*)st);
      hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(GNT), 0, s, (const double*)B.partials, nb, st);
      hipLaunchKernelGGL(cg_update_kernel<T>, dim3(nb), dim3(GNT), 0, s, phi, r, (const T*)p, (const T*)Ap, n,
                         B.partials, (const CgState*)st);
      hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(GNT), 0, s, (const double*)B.partials, nb, st, P.tol,
                         P.maxiter);
      hipLaunchKernelGGL(cg_dir_kernel<T>, dim3(nb), dim3(GNT), 0, s, (const T*)r, p, n, (const CgState*)st);
    }
    done_it += chunk;
    VH_HIP(hipMemcpyAsync(B.host_state, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    VH_HIP(hipStreamSynchronize(s));
    h = *(const CgState*)B.host_state;
    if (h.done) break;
  }
  // volume-weighted mean projection (src/gravity.jl:310-320)
  if (P.project_mean) {
    vsum<T>(s, nb, op, (const T*)phi, B.partials, scal + 5);
    hipLaunchKernelGGL(sub_mean_kernel<T>, dim3(nb), dim3(GNT), 0, s, phi, n, (const double*)scal);
  }
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(s));
  *iters = (long long)h.iters;
  *relres = h.relres;
  return 0;
}

int grav_solve(hipStream_t s, int dtype, const GridDev& G, const GravParams& P, GravBuffers& B, long long* iters,
               double* relres) {
  UniformOp op{&G, P.gper, {0, 0, 0}};
  for (int d = 0; d < G.rank; ++d) op.coef[d] = G.area[d] / (0.5 * (G.w[d] + G.w[d]));
  if (dtype == 0) return solve_t<double>(s, op, P, B, iters, relres);
  return solve_t<float>(s, op, P, B, iters, relres);
}

void grav_csr_apply(hipStream_t s, int dtype, const GravCsr& A, const void* x, void* out) {
  const CsrOp op{&A};
  const int nb = grid_blocks(A.n);
  if (dtype == 0) op.lap<double>(s, nb, (const double*)x, (double*)out, nullptr, nullptr);
  else op.lap<float>(s, nb, (const float*)x, (float*)out, nullptr, nullptr);
}

void grav_csr_accel(hipStream_t s, int dtype, const GravCsr& A, const void* phi, void* g) {
  const int nb = grid_blocks(A.n);
  if (dtype == 0)
    hipLaunchKernelGGL(csr_accel_kernel<double>, dim3(nb), dim3(GNT), 0, s, (const double*)phi, (double*)g, A.glo,
                       A.ghi, A.gsum, A.n, A.rank);
  else
    hipLaunchKernelGGL(csr_accel_kernel<float>, dim3(nb), dim3(GNT), 0, s, (const float*)phi, (float*)g, A.glo,
                       A.ghi, A.gsum, A.n, A.rank);
}

int grav_csr_solve(hipStream_t s, int dtype, const GravCsr& A, const GravParams& P, GravBuffers& B, long long* iters,
                   double* relres) {
  const CsrOp op{&A};
  if (dtype == 0) return solve_t<double>(s, op, P, B, iters, relres);
  return solve_t<float>(s, op, P, B, iters, relres);
}

}  // namespace vh
