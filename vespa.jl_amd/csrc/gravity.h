// Gravitational acceleration and the gas source term (uniform grids), shared by
// the stage kernels of both arithmetic profiles and the gravity kernels.
#pragma once

#include "common.h"

namespace vh {
namespace VH_PROFILE_NS {

// g_d = −(φ_hi − φ_lo)/Tf(d_lo + d_hi) at cell c with coordinates co
// (src/gravity.jl:403-410): centre distances ½(w + w) to the face neighbours;
// a side on a non-periodic φ boundary (bit d of gper clear) takes φ_c and the
// half width (homogeneous Neumann).
template <typename T>
__device__ __forceinline__ T grav_axis(const T* __restrict__ phi, const GridDev& G, int gper, long long c,
                                       const int* co, int d) {
  const long long st = d == 0 ? 1 : (d == 1 ? (long long)G.n[0] : (long long)G.n[0] * G.n[1]);
  const int n = G.n[d];
  const double w = G.w[d];
  const T pc = phi[c];
  T wlo, whi;
  double dlo, dhi;
  if (co[d] > 0) {
    wlo = phi[c - st];
    dlo = 0.5 * (w + w);
  } else if (gper >> d & 1) {
    wlo = phi[c + (long long)(n - 1) * st];
    dlo = 0.5 * (w + w);
  } else {
    wlo = pc;
    dlo = 0.5 * w;
  }
  if (co[d] + 1 < n) {
    whi = phi[c + st];
    dhi = 0.5 * (w + w);
  } else if (gper >> d & 1) {
    whi = phi[c - (long long)(n - 1) * st];
    dhi = 0.5 * (w + w);
  } else {
    whi = pc;
    dhi = 0.5 * w;
  }
  return -(whi - wlo) / T(dlo + dhi);
}

// apply_gravity_source! (src/gravity.jl:388-447): the flux sum acc of the
// stage's input state U gets (0, −ρg_x·V, −ρg_y·V, −ρg_z·V, −(ρv·g)·V) added,
// component by component, after the fluxes.
template <typename T, int RANK>
__device__ __forceinline__ void grav_source_add(Cons<T>& acc, const Cons<T>& U, const T* __restrict__ phi,
                                                const GridDev& G, long long c, int gx, int gy, int gz) {
  const int co[3] = {gx, gy, gz};
  T g[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) g[d] = d < RANK ? grav_axis<T>(phi, G, G.gper, c, co, d) : T(0);
  const T V = T(G.V);
  const T rho = U.d;
  T es = T(0);
  es += U.mx * g[0];
  es += U.my * g[1];
  es += U.mz * g[2];
  const T esV = es * V;
  acc.d = acc.d + T(0);
  acc.mx = acc.mx + (-(rho * g[0] * V));
  acc.my = acc.my + (-(rho * g[1] * V));
  acc.mz = acc.mz + (-(rho * g[2] * V));
  acc.e = acc.e + (-esV);
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
