// Host-side error plumbing shared by the translation units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

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

namespace vh {

// Reverse-mode launchers (adjoint.hip), Float64 uniform grids. Face buffers
// gL/gR per axis hold 5 SoA fields over adj_face_slots(G, axis) face slots.
long long adj_face_slots(const GridDev& G, int axis);
void adj_launch_prim(hipStream_t s, const double* U, double* W, const GridDev& G);
void adj_launch_faces(hipStream_t s, const double* W, const double* lam, const GridDev& G, double* const gL[3],
                      double* const gR[3]);
void adj_launch_cells(hipStream_t s, const double* U, const double* W, double* const gL[3], double* const gR[3],
                      const double* base1, double a1, const double* base2, double a2, double beta,
                      const double* dtp, const GridDev& G, double* out);
// Fused single-kernel form for 2D grids periodic on both axes whose sizes are
// multiples of the 32×8 tile; returns false (nothing launched) otherwise.
bool adj_launch_jat_fused(hipStream_t s, const double* U, const double* lam, const double* base1, double a1,
                          const double* base2, double a2, double beta, const double* dtp, const GridDev& G,
                          double* out);

}  // namespace vh

namespace vh {

// Self-gravity on uniform grids (gravity.hip). φ and the CG vectors are stored
// at the state's precision; dot products and the matrix-vector sums in Float64.
struct GravParams {
  double G = 1.0;
  int gper = 0;            // bit d: φ periodic along axis d
  int project_mean = 1;
  long long maxiter = 500;
  double tol = 1e-8;
  int chunk = 16;          // iterations enqueued between host checks
};
struct GravBuffers {
  void* phi = nullptr;     // ncell values
  void* work = nullptr;    // 4·ncell values: r, p, Ap, b
  const void* rho = nullptr;
  void* state = nullptr;   // CG state block (device)
  double* scal = nullptr;  // 8 scalars (device)
  double* partials = nullptr;  // 4096 per-workgroup partial sums (device)
  void* host_state = nullptr;  // pinned copy of the state block
};
constexpr int kGravStateBytes = 64;
void grav_apply(hipStream_t s, int dtype, const GridDev& G, int gper, const void* x, void* out);
void grav_accel(hipStream_t s, int dtype, const GridDev& G, int gper, const void* phi, void* g);
int grav_solve(hipStream_t s, int dtype, const GridDev& G, const GravParams& P, GravBuffers& B, long long* iters,
               double* relres);

// The composite operator of a block mesh in CSR form (the reference's
// PoissonKAStruct, src/gravity.jl:455-600), built on the host per mesh: row c
// lists its face neighbours in the face walk's order with coef = area/d, and
// the gravity stencil gives per cell and axis the lo/hi neighbour (the cell
// itself at a Neumann side) and d_lo + d_hi. Device pointers.
struct GravCsr {
  long long n = 0;
  const int* rowptr = nullptr;      // n + 1
  const int* colidx = nullptr;
  const double* coef = nullptr;
  const double* vol = nullptr;      // per-cell volume
  const int* glo = nullptr;         // 3·n, [axis][cell]
  const int* ghi = nullptr;
  const double* gsum = nullptr;     // 3·n, [axis][cell]
  int rank = 1;
};
void grav_csr_apply(hipStream_t s, int dtype, const GravCsr& A, const void* x, void* out);
void grav_csr_accel(hipStream_t s, int dtype, const GravCsr& A, const void* phi, void* g);
int grav_csr_solve(hipStream_t s, int dtype, const GravCsr& A, const GravParams& P, GravBuffers& B, long long* iters,
                   double* relres);

}  // namespace vh

namespace vh {

// Block-mesh regrid (regrid.hip). RemapDev: per destination block, the source
// block that contains it, the level difference and the position of its first
// cell in the source block's cells.
struct RemapDev {
  int src, shift;
  int o[3];
};
void regrid_indicator(hipStream_t s, int dtype, int nblocks, const BlockDev* blocks, const int* repmap, int rank,
                      const void* rho, double* cell_g, double* block_max);
void regrid_remap(hipStream_t s, int dtype, int ndst, const BlockDev* dst_blocks, const BlockDev* src_blocks,
                  const RemapDev* map, const void* Us, long long ns, void* Ud, long long nd, int nf);

}  // namespace vh
