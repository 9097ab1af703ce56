// Launchers of the reverse-mode kernels (adjoint.h), Float64 uniform grids.
// Built once, in the fast arithmetic profile (FMA contraction, refined hardware
// reciprocals): the reverse mode needs no bit parity with the reference.
#include "adjoint.h"
#include "host.h"

namespace vh {

static unsigned grid_for(long long n) { return (unsigned)((n + 255) / 256); }

long long adj_face_slots(const GridDev& G, int axis) {
  long long s = 1;
  for (int d = 0; d < 3; ++d) s *= (d == axis ? G.n[d] + 1 : G.n[d]);
  return s;
}

void adj_launch_prim(hipStream_t s, const double* U, double* W, const GridDev& G) {
  hipLaunchKernelGGL(VH_PROFILE_NS::adj_prim_kernel, dim3(grid_for(G.ncell)), dim3(256), 0, s, U, W, G);
}

void adj_launch_faces(hipStream_t s, const double* W, const double* lam, const GridDev& G, double* const gL[3],
                      double* const gR[3]) {
  hipLaunchKernelGGL(VH_PROFILE_NS::adj_face_kernel<0>, dim3(grid_for(adj_face_slots(G, 0))), dim3(256), 0, s, W, lam, G,
                     gL[0], gR[0]);
  if (G.rank > 1)
    hipLaunchKernelGGL(VH_PROFILE_NS::adj_face_kernel<1>, dim3(grid_for(adj_face_slots(G, 1))), dim3(256), 0, s, W, lam, G,
                       gL[1], gR[1]);
  if (G.rank > 2)
    hipLaunchKernelGGL(VH_PROFILE_NS::adj_face_kernel<2>, dim3(grid_for(adj_face_slots(G, 2))), dim3(256), 0, s, W, lam, G,
                       gL[2], gR[2]);
}

void adj_launch_cells(hipStream_t s, const double* U, const double* W, double* const gL[3], double* const gR[3],
                      const double* base1, double a1, const double* base2, double a2, double beta,
                      const double* dtp, const GridDev& G, double* out) {
  VH_PROFILE_NS::AdjFaces F;
  for (int d = 0; d < 3; ++d) {
    F.gL[d] = gL[d];
    F.gR[d] = gR[d];
  }
  const dim3 grid(grid_for(G.ncell));
  if (G.rank == 3)
    hipLaunchKernelGGL(VH_PROFILE_NS::adj_cell_kernel<3>, grid, dim3(256), 0, s, U, W, F, base1, a1, base2, a2, beta, dtp, G,
                       out);
  else if (G.rank == 2)
    hipLaunchKernelGGL(VH_PROFILE_NS::adj_cell_kernel<2>, grid, dim3(256), 0, s, U, W, F, base1, a1, base2, a2, beta, dtp, G,
                       out);
  else
    hipLaunchKernelGGL(VH_PROFILE_NS::adj_cell_kernel<1>, grid, dim3(256), 0, s, U, W, F, base1, a1, base2, a2, beta, dtp, G,
                       out);
}

bool adj_launch_jat_fused(hipStream_t s, const double* U, const double* lam, const double* base1, double a1,
                          const double* base2, double a2, double beta, const double* dtp, const GridDev& G,
                          double* out) {
  const bool periodic = G.bc[0] == BC_PERIODIC && G.bc[2] == BC_PERIODIC;
  if (G.rank != 2 || !periodic || G.n[0] % VH_PROFILE_NS::JTX || G.n[1] % VH_PROFILE_NS::JTY) return false;
  const dim3 grid(G.n[0] / VH_PROFILE_NS::JTX, G.n[1] / VH_PROFILE_NS::JTY);
  hipLaunchKernelGGL(VH_PROFILE_NS::adj_jat2d_periodic_kernel, grid, dim3(VH_PROFILE_NS::JNT), 0, s, U, lam, base1, a1,
                     base2, a2, beta, dtp, G, out);
  return true;
}

}  // namespace vh
