// Device-side ideal-gas hydro kernels for the flux-divergence path (gfx950).
//
// These are the per-face / per-cell building blocks shared by the batched
// face-flux kernel and the fused stencil stage kernels.
//
// Two arithmetic profiles are compiled from this one source (VH_FAST_PROFILE = 0 / 1):
//
// * exact (VH_FAST_PROFILE = 0, built with -ffp-contract=off): every expression keeps
//   the reference's evaluation order, so results are bit-identical to the CPU
//   restatement:
//     cons2prim      src/eos.jl:8-25
//     prim2cons / E  src/eos.jl:28-32 (E = p/(γ-1) + (½ρ)((vx²+vy²)+vz²), a division)
//     minmod         src/reconstruct.jl:6-16 (a·b ≤ 0 → 0; |a| < |b| ? a : b)
//     euler_flux     src/riemann.jl:9-20
//     hllc_flux      src/riemann.jl:33-66, _hllc_star src/riemann.jl:69-88
//   The HLLC evaluation selects the upwind side K first (SL ≥ 0 → L; SR ≤ 0 → R;
//   else the sign of S*) and forms only that side's E, physical flux and star
//   state; every value reaching the output is produced by the same operations
//   as in the reference.
//
// * fast (VH_FAST_PROFILE = 1, built with FMA contraction): the same formulas with
//   a*b+c fused, E = p·(1/(γ−1)), sound speeds c = (γp)·rsqrt(γpρ) refined by
//   Newton steps, divisions through a refined reciprocal, and the star flux
//   written through the star state itself, so that one reciprocal 1/(S_K − S*)
//   replaces the three divisions by ρ, by ρ(S_K − u_n) and by (S_K − S*) and the
//   physical flux F_K is formed only on supersonic faces. With q = ρ(S_K − u_n),
//   ρ* = q/(S_K − S*), p* = p + q(S* − u_n) and
//     ρ_*E_* = [(S_K − u_n)·E + (S* − u_n)·(q·S* + p)] / (S_K − S*),
//   the Rankine-Hugoniot form F*_K = F_K + S_K·(U*_K − U_K) equals
//     F*_K = (ρ*S*,  ρ*S*·S* + p* (normal),  ρ*S*·v_t (tangential),  (ρ_*E_* + p*)·S*),
//   which costs about ten fewer operations per face.
//   Results agree with the exact profile to a few ulps per face.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace vh {
namespace VH_PROFILE_NS {

__device__ __forceinline__ double vabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float vabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double vsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float vsqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double vfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float vfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Reciprocal from the hardware estimate (relative error below 2^-24 on gfx950)
// refined by one third-order step: e = 1 − b·r, r ← r + r·(e + e²), error
// O(e³); then a/b as a·r with one residual correction.
__device__ __forceinline__ double fast_rcp(double b) {
  const double r = __builtin_amdgcn_rcp(b);
  const double e = vfma(-b, r, 1.0);
  return vfma(r, vfma(e, e, e), r);
}
__device__ __forceinline__ float fast_rcp(float b) {
  float r = __builtin_amdgcn_rcpf(b);
  return vfma(vfma(-b, r, 1.0f), r, r);
}
template <typename T>
__device__ __forceinline__ T fast_div(T a, T b) {
  const T r = fast_rcp(b);
  const T q = a * r;
  return vfma(vfma(-b, q, a), r, q);
}
// sqrt(x/y) for x >= 0, y > 0: x·rsqrt(x·y), the rsqrt estimate (relative error
// below 2^-24) refined by one third-order step: e = 1 − a·s², s ← s + s·e·(½ + ⅜e)
// (x·y is kept at or above the smallest normal so that x = 0 gives 0).
__device__ __forceinline__ double fast_sqrt_ratio(double x, double y) {
  const double a = vmax(x * y, 2.2250738585072014e-308);
  const double s = __builtin_amdgcn_rsq(a);
  const double e = vfma(-(a * s), s, 1.0);
  return x * vfma(s, e * vfma(e, 0.375, 0.5), s);
}
__device__ __forceinline__ float fast_sqrt_ratio(float x, float y) {
  const float a = vmax(x * y, 1.17549435e-38f);
  float s = __builtin_amdgcn_rsqf(a);
  const float h = 0.5f * a;
  s = s * vfma(-h * s, s, 1.5f);
  return x * s;
}

// Reciprocal used by cons2prim (1/ρ).
template <typename T>
__device__ __forceinline__ T recip(T x) {
#if VH_FAST_PROFILE
  return fast_rcp(x);
#else
  return T(1) / x;
#endif
}

// cons2prim (src/eos.jl:8-25). With FLOOR64 (the FP32 form) the pressure and its
// positivity floor are evaluated in double from the Tf energy and kinetic terms.
template <typename T, bool FLOOR64>
__device__ __forceinline__ Prim<T> cons2prim(const Cons<T>& U, T g, double gd) {
  Prim<T> W;
  const T rho = U.d;
  const T inv = recip(rho);
  W.r = rho;
  W.vx = U.mx * inv;
  W.vy = U.my * inv;
  W.vz = U.mz * inv;
  const T kinetic = (T(0.5) * rho) * ((W.vx * W.vx + W.vy * W.vy) + W.vz * W.vz);
  if constexpr (FLOOR64) {
    const double pd = (gd - 1.0) * (double(U.e) - double(kinetic));
    W.p = T(vmax(pd, 1e-12 * double(kinetic)));
  } else {
    const T p = (g - T(1)) * (U.e - kinetic);
    W.p = vmax(p, T(1e-12) * kinetic);
  }
  return W;
}

// Total energy from primitives (the E of prim2cons and euler_flux). `igm1` is
// 1/(γ−1) (used by the fast profile only).
template <typename T>
__device__ __forceinline__ T energy(const Prim<T>& W, T g, T igm1) {
#if VH_FAST_PROFILE
  return W.p * igm1 + (T(0.5) * W.r) * ((W.vx * W.vx + W.vy * W.vy) + W.vz * W.vz);
#else
  (void)igm1;
  return W.p / (g - T(1)) + (T(0.5) * W.r) * ((W.vx * W.vx + W.vy * W.vy) + W.vz * W.vz);
#endif
}

#if VH_FAST_PROFILE
// minmod(a, b) is the median of (a, b, 0): the smaller-magnitude argument when
// the signs agree, else 0. Float64: max(min(a,b), min(max(a,b), 0)), four
// min/max; Float32: one v_med3_f32. Differs from the reference form only when
// a·b underflows to zero with both signs equal (a subnormal-scale slope).
__device__ __forceinline__ double minmod(double a, double b) { return vmax(vmin(a, b), vmin(vmax(a, b), 0.0)); }
__device__ __forceinline__ float minmod(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.0f); }
#else
template <typename T>
__device__ __forceinline__ T minmod(T a, T b) {
  if (a * b <= T(0)) return T(0);
  return vabs(a) < vabs(b) ? a : b;
}
#endif

// limited_slope(l, c, r) = minmod(c - l, r - c) per primitive component.
template <typename T>
__device__ __forceinline__ Prim<T> slope(const Prim<T>& l, const Prim<T>& c, const Prim<T>& r) {
  Prim<T> s;
  s.r = minmod(c.r - l.r, r.r - c.r);
  s.vx = minmod(c.vx - l.vx, r.vx - c.vx);
  s.vy = minmod(c.vy - l.vy, r.vy - c.vy);
  s.vz = minmod(c.vz - l.vz, r.vz - c.vz);
  s.p = minmod(c.p - l.p, r.p - c.p);
  return s;
}

// _face_value: W + h·slope with h = ±½ (src/driver.jl:227-233). The fast
// profile states the fused form so that the two faces of a cell cost one FMA
// each rather than a shared product and two adds.
template <typename T>
__device__ __forceinline__ Prim<T> face_value(const Prim<T>& w, const Prim<T>& s, T h) {
  Prim<T> o;
#if VH_FAST_PROFILE
  o.r = vfma(h, s.r, w.r);
  o.vx = vfma(h, s.vx, w.vx);
  o.vy = vfma(h, s.vy, w.vy);
  o.vz = vfma(h, s.vz, w.vz);
  o.p = vfma(h, s.p, w.p);
#else
  o.r = w.r + h * s.r;
  o.vx = w.vx + h * s.vx;
  o.vy = w.vy + h * s.vy;
  o.vz = w.vz + h * s.vz;
  o.p = w.p + h * s.p;
#endif
  return o;
}

// ghost_state (src/driver.jl:117-122). axis is 1-based.
template <typename T>
__device__ __forceinline__ Prim<T> ghost(const Prim<T>& w, int bc, int axis) {
  Prim<T> o = w;
  if (bc == BC_REFLECTING) {
    if (axis == 1) o.vx = -w.vx;
    else if (axis == 2) o.vy = -w.vy;
    else o.vz = -w.vz;
  }
  return o;
}

template <typename T>
__device__ __forceinline__ T normal_v(const Prim<T>& w, int dim) {
  return dim == 1 ? w.vx : (dim == 2 ? w.vy : w.vz);
}

template <typename T>
__device__ __forceinline__ T sound(T g, T p, T rho) {
#if VH_FAST_PROFILE
  return fast_sqrt_ratio(g * p, rho);
#else
  return vsqrt(g * p / rho);
#endif
}

// HLLC flux across a face normal to `dim` (1-based), src/riemann.jl:33-88.
template <typename T>
__device__ __forceinline__ Cons<T> hllc(const Prim<T>& WL, const Prim<T>& WR, T g, T igm1, int dim) {
  const T rL = WL.r, rR = WR.r, pL = WL.p, pR = WR.p;
  const T unL = normal_v(WL, dim), unR = normal_v(WR, dim);
  const T cL = sound(g, pL, rL);
  const T cR = sound(g, pR, rR);
  const T SL = vmin(unL - cL, unR - cR);
  const T SR = vmax(unL + cL, unR + cR);
  const bool supL = SL >= T(0);
  const bool supR = !supL && (SR <= T(0));
#if VH_FAST_PROFILE
  const T aL = rL * (SL - unL), aR = rR * (SR - unR);
  const T Ss = fast_div(vfma(aL, unL, pR - pL) - aR * unR, aL - aR);
#else
  const T Ss = (pR - pL + rL * unL * (SL - unL) - rR * unR * (SR - unR)) /
               (rL * (SL - unL) - rR * (SR - unR));
#endif
  const bool useL = supL || (!supR && Ss >= T(0));

  const Prim<T> W = useL ? WL : WR;
  const T SK = useL ? SL : SR;
  const T un = useL ? unL : unR;
  const T E = energy(W, g, igm1);
#if VH_FAST_PROFILE
  Cons<T> o;
  if (supL || supR) {
    o.d = W.r * un;
    o.mx = o.d * W.vx + (dim == 1 ? W.p : T(0));
    o.my = o.d * W.vy + (dim == 2 ? W.p : T(0));
    o.mz = o.d * W.vz + (dim == 3 ? W.p : T(0));
    o.e = (E + W.p) * un;
    return o;
  }
  const T sku = SK - un, su = Ss - un;
  const T q = W.r * sku;
  const T inv = fast_rcp(SK - Ss);
  const T ps = vfma(q, su, W.p);
  const T es = vfma(sku, E, su * vfma(q, Ss, W.p)) * inv;
  o.d = (q * inv) * Ss;
  o.mx = dim == 1 ? vfma(o.d, Ss, ps) : o.d * W.vx;
  o.my = dim == 2 ? vfma(o.d, Ss, ps) : o.d * W.vy;
  o.mz = dim == 3 ? vfma(o.d, Ss, ps) : o.d * W.vz;
  o.e = (es + ps) * Ss;
  return o;
#else
  const T m1 = W.r * W.vx, m2 = W.r * W.vy, m3 = W.r * W.vz;
  Cons<T> F;
  F.d = W.r * un;
  F.mx = m1 * un + (dim == 1 ? W.p : T(0));
  F.my = m2 * un + (dim == 2 ? W.p : T(0));
  F.mz = m3 * un + (dim == 3 ? W.p : T(0));
  F.e = (E + W.p) * un;
  if (supL || supR) return F;

  // F*K = FK + SK (U*K − UK), src/riemann.jl:69-88.
  const T q = W.r * (SK - un);
  const T vsx = dim == 1 ? Ss : W.vx;
  const T vsy = dim == 2 ? Ss : W.vy;
  const T vsz = dim == 3 ? Ss : W.vz;
  const T factor = q / (SK - Ss);
  const T Es = E / W.r + (Ss - un) * (Ss + W.p / q);
  const T fE = factor * Es;
  Cons<T> o;
  o.d = F.d + SK * (factor - W.r);
  o.mx = F.mx + SK * (factor * vsx - m1);
  o.my = F.my + SK * (factor * vsy - m2);
  o.mz = F.mz + SK * (factor * vsz - m3);
  o.e = F.e + SK * (fE - E);
  return o;
#endif
}

template <typename T>
__device__ __forceinline__ void acc_add(Cons<T>& a, const Cons<T>& F, T s) {
  a.d += s * F.d;
  a.mx += s * F.mx;
  a.my += s * F.my;
  a.mz += s * F.mz;
  a.e += s * F.e;
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
