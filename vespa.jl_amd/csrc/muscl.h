// Stage kernel of PPMKernels' unsplit SSP-RK2 MUSCL step, in the arithmetic profile
// selected by VH_FAST_PROFILE (compiled once per profile by muscl_kernels.hip):
//
// * exact (VH_FAST_PROFILE = 0, -ffp-contract=off): the reference's operations in
//   its order, divisions and square roots correctly rounded — bit-identical to the
//   reference's arithmetic;
// * fast (VH_FAST_PROFILE = 1, contraction on): a*b+c fused, one refined reciprocal
//   per cell for cons→prim (1/ρ) and per face for the HLL fluxes (1/(a⁺+a⁻)), sound
//   speeds as (γp)·rsqrt(γpρ) refined; results within a few ulps per face of the
//   exact profile.
#pragma once

#include "hydro.h"
#include "muscl_common.h"

namespace vh {
namespace VH_PROFILE_NS {

template <typename T>
struct MuW {
  T r, e, vx, vy, vz;
};

// _plm_pt (lib/PPMKernels/src/muscl.jl:13-27): v + ½·minmod3((v−vm1)θ, (vp1−v)θ,
// ½(vp1−vm1)) with Enzo's sign(0) = −1.
template <typename T>
__device__ __forceinline__ T plm_pt(T vm1, T v, T vp1, T th) {
  const T a = (v - vm1) * th, b = (vp1 - v) * th, c = T(0.5) * (vp1 - vm1);
  const T sa = a > T(0) ? T(1) : T(-1), sb = b > T(0) ? T(1) : T(-1), sc = c > T(0) ? T(1) : T(-1);
  const T s = T(0.25) * (sa + sb) * absT(sa + sc) * jmin(absT(a), jmin(absT(b), absT(c)));
  return v + T(0.5) * s;
}
// Both face values of the middle cell at once. The lo-face value is _plm_pt with
// the neighbours exchanged (the reference's R state); its minmod arguments are the
// exact negatives of the hi face's in swapped order, so it reuses their magnitudes
// and minimum and only recomputes the signs: bit-identical to evaluating it anew,
// zero signs included.
// The leading factor of minmod3, 0.25·(sa+sb)·|sa+sd| with s = +1 for a positive
// argument and −1 otherwise, is exactly ±1 when the three signs agree and a zero
// otherwise: −0 when sa = sb = −1 ≠ sd, +0 in the other cases (the signs the
// products give). Selected from the comparisons instead of multiplied out.
template <typename T>
__device__ __forceinline__ T minmod_factor(bool pa, bool pb, bool pd) {
  const T full = pa ? T(1) : T(-1);
  const T zero = pa ? T(0) : T(-0.0);
  return pa == pb ? (pa == pd ? full : zero) : T(0);
}
template <typename T>
__device__ __forceinline__ void plm_pair(T m, T c, T p, T th, T& hi, T& lo) {
  const T a = (c - m) * th, b = (p - c) * th, d = T(0.5) * (p - m);
  const T mn = jmin(absT(a), jmin(absT(b), absT(d)));
  hi = c + T(0.5) * (minmod_factor<T>(a > T(0), b > T(0), d > T(0)) * mn);
  lo = c + T(0.5) * (minmod_factor<T>(b < T(0), a < T(0), d < T(0)) * mn);
}
template <typename T>
__device__ __forceinline__ void face_pair(const MuW<T>& m, const MuW<T>& c, const MuW<T>& p, T th, T sr, MuW<T>& hi,
                                          MuW<T>& lo) {
  plm_pair(m.r, c.r, p.r, th, hi.r, lo.r);
  hi.r = jmax(hi.r, sr);
  lo.r = jmax(lo.r, sr);
  plm_pair(m.e, c.e, p.e, th, hi.e, lo.e);
  plm_pair(m.vx, c.vx, p.vx, th, hi.vx, lo.vx);
  plm_pair(m.vy, c.vy, p.vy, th, hi.vy, lo.vy);
  plm_pair(m.vz, c.vz, p.vz, th, hi.vz, lo.vz);
}
// Hi-face state of cell c from its neighbours m (lo side) and p (hi side); the
// lo-face state is the same with m and p exchanged (the reference's R state).
template <typename T>
__device__ __forceinline__ MuW<T> face_state(const MuW<T>& m, const MuW<T>& c, const MuW<T>& p, T th, T sr) {
  MuW<T> o;
  o.r = jmax(plm_pt(m.r, c.r, p.r, th), sr);
  o.e = plm_pt(m.e, c.e, p.e, th);
  o.vx = plm_pt(m.vx, c.vx, p.vx, th);
  o.vy = plm_pt(m.vy, c.vy, p.vy, th);
  o.vz = plm_pt(m.vz, c.vz, p.vz, th);
  return o;
}

// (a⁺F_L + a⁻F_R − a⁺a⁻(U_R − U_L)) / (a⁺ + a⁻); rinv = 1/(a⁺ + a⁻) in the fast profile.
template <typename T>
__device__ __forceinline__ T hll1(T ap, T am, T fl, T fr, T ul, T ur, T rinv) {
#if VH_FAST_PROFILE
  return (ap * fl + am * fr - ap * am * (ur - ul)) * rinv;
#else
  (void)rinv;
  return (ap * fl + am * fr - ap * am * (ur - ul)) / (ap + am);
#endif
}
// sqrt(γp/ρ)
template <typename T>
__device__ __forceinline__ T sound(T gp, T rho) {
#if VH_FAST_PROFILE
  return fast_sqrt_ratio(gp, rho);
#else
  return sqrtT(gp / rho);
#endif
}

// _hll6's hydro fluxes (lib/PPMKernels/src/muscl.jl:37-58) for states whose
// velocities are already in (normal, t1, t2) order in (vx, vy, vz). F: D, normal,
// t1, t2 momentum, energy.
template <typename T, int NF>
__device__ __forceinline__ void hll_rot(const MuW<T>& L, const MuW<T>& R, T g, T gm1, T F[NF]) {
  const T h = T(0.5);
  const T rl = L.r, el = L.e, ul = L.vx, vl = L.vy, wl = L.vz;
  const T rr = R.r, er = R.e, ur = R.vx, vr = R.vy, wr = R.vz;
  const T v2l = ul * ul + vl * vl + wl * wl, pl = gm1 * rl * el, csl = sound(g * pl, rl), etl = el + h * v2l;
  const T UlS1 = rl * ul, UlS2 = rl * vl, UlS3 = rl * wl, UlE = rl * etl;
  const T FlD = rl * ul, FlS1 = UlS1 * ul + pl, FlS2 = UlS2 * ul, FlS3 = UlS3 * ul, FlE = (UlE + pl) * ul;
  const T lpl = ul + csl, lml = ul - csl;
  const T v2r = ur * ur + vr * vr + wr * wr, pr = gm1 * rr * er, csr = sound(g * pr, rr), etr = er + h * v2r;
  const T UrS1 = rr * ur, UrS2 = rr * vr, UrS3 = rr * wr, UrE = rr * etr;
  const T FrD = rr * ur, FrS1 = UrS1 * ur + pr, FrS2 = UrS2 * ur, FrS3 = UrS3 * ur, FrE = (UrE + pr) * ur;
  const T lpr = ur + csr, lmr = ur - csr;
  const T ap = jmax(T(0), jmax(lpl, lpr));
  const T am = jmax(T(0), jmax(-lml, -lmr));
#if VH_FAST_PROFILE
  const T rinv = fast_rcp(ap + am);
#else
  const T rinv = T(0);
#endif
  F[0] = hll1(ap, am, FlD, FrD, rl, rr, rinv);
  F[1] = hll1(ap, am, FlS1, FrS1, UlS1, UrS1, rinv);
  F[2] = hll1(ap, am, FlS2, FrS2, UlS2, UrS2, rinv);
  F[3] = hll1(ap, am, FlS3, FrS3, UlS3, UrS3, rinv);
  F[4] = hll1(ap, am, FlE, FrE, UlE, UrE, rinv);
  if constexpr (NF == 6) {   // _hll6's gas-energy flux: U = ρe, F = ρe·u
    const T UlG = rl * el, UrG = rr * er;
    F[5] = hll1(ap, am, UlG * ul, UrG * ur, UlG, UrG, rinv);
  }
}

// Velocities of a state in the cyclic roles of axis a: (normal, t1, t2) =
// x: (vx, vy, vz), y: (vy, vz, vx), z: (vz, vx, vy).
template <typename T>
__device__ __forceinline__ MuW<T> roles(const MuW<T>& w, int a) {
  MuW<T> o = w;
  if (a == 1) {
    o.vx = w.vy; o.vy = w.vz; o.vz = w.vx;
  } else if (a == 2) {
    o.vx = w.vz; o.vy = w.vx; o.vz = w.vy;
  }
  return o;
}
// HLL flux through a face normal to axis a, returned in grid order (D, Sx, Sy, Sz, E[, ρe]).
template <typename T, int NF>
__device__ __forceinline__ void hll_axis(const MuW<T>& L, const MuW<T>& R, int a, T g, T gm1, T out[NF]) {
  T F[NF];
  hll_rot<T, NF>(roles(L, a), roles(R, a), g, gm1, F);
  out[0] = F[0];
  out[4] = F[4];
  if constexpr (NF == 6) out[5] = F[5];
  if (a == 0) {
    out[1] = F[1]; out[2] = F[2]; out[3] = F[3];
  } else if (a == 1) {
    out[2] = F[1]; out[3] = F[2]; out[1] = F[3];
  } else {
    out[3] = F[1]; out[1] = F[2]; out[2] = F[3];
  }
}

// _dual_eint (lib/PPMKernels/src/muscl_grid.jl:41-47): the total-energy eint where the
// pressure is dynamically significant, else the advected one; floored.
// tod = τ/ρ and god = ρe/ρ (quotients formed by the caller).
template <typename T>
__device__ __forceinline__ T dual_eint(T tod, T god, T v2, T g, T gm1, T eta1, T small) {
  const T eint_tot = tod - T(0.5) * v2;
  const T eint_dual = god;
  const T cs2 = eint_tot > T(0) ? g * gm1 * eint_tot : T(0);
  const bool use_tot = (cs2 > eta1 * v2) && (eint_tot > T(0.5) * eint_dual);
  return jmax(use_tot ? eint_tot : eint_dual, small);
}

// cons → prim (_cons2prim_k!, lib/PPMKernels/src/muscl_grid.jl:23-33, eint floored; with
// the gas energy, _cons2prim_dual_k!, :49-60).
template <typename T, int NF>
__device__ __forceinline__ MuW<T> to_prim(const T* __restrict__ U, long long n, long long c, T sr, T g, T gm1,
                                          T eta1) {
  MuW<T> w;
  const T d = U[c];
#if VH_FAST_PROFILE
  const T inv = fast_rcp(d);
  const T u = U[n + c] * inv, v = U[2 * n + c] * inv, ww = U[3 * n + c] * inv, tod = U[4 * n + c] * inv;
#else
  const T u = U[n + c] / d, v = U[2 * n + c] / d, ww = U[3 * n + c] / d, tod = U[4 * n + c] / d;
#endif
  w.r = d;
  w.vx = u;
  w.vy = v;
  w.vz = ww;
  if constexpr (NF == 6) {
#if VH_FAST_PROFILE
    const T god = U[5 * n + c] * inv;
#else
    const T god = U[5 * n + c] / d;
#endif
    w.e = dual_eint(tod, god, u * u + v * v + ww * ww, g, gm1, eta1, sr);
  } else {
    w.e = jmax(tod - T(0.5) * (u * u + v * v + ww * ww), sr);
  }
  return w;
}

template <typename T, int TX, int TY, int NF>
struct MusLDS {
  static constexpr int WX = TX + 4;
  static constexpr int WS = (TY + 4) * WX;                 // per-field stride of w
  static constexpr int NXS = TY * (TX + 2);
  static constexpr int FS = NXS + (TY + 2) * TX;           // per-field stride of f
  T w[5][TY + 4][TX + 4];
  T f[NF][FS];
  __device__ static int xs(int row, int i) { return row * (TX + 2) + i; }   // x face i of row (0..TX+1)
  __device__ static int ys(int j, int col) { return NXS + j * TX + col; }   // y face j of column (0..TY+1)
};

template <typename T>
__device__ __forceinline__ void put_w(T* base, int stride, int i, const MuW<T>& w) {
  base[i] = w.r; base[stride + i] = w.e; base[2 * stride + i] = w.vx; base[3 * stride + i] = w.vy;
  base[4 * stride + i] = w.vz;
}
template <typename T>
__device__ __forceinline__ MuW<T> get_w(const T* base, int stride, int i) {
  MuW<T> w;
  w.r = base[i]; w.e = base[stride + i]; w.vx = base[2 * stride + i]; w.vy = base[3 * stride + i];
  w.vz = base[4 * stride + i];
  return w;
}
template <typename T, int NF>
__device__ __forceinline__ void put_f(T* base, int stride, int i, const T F[NF]) {
  for (int q = 0; q < NF; ++q) base[q * stride + i] = F[q];
}

__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Waves per SIMD the stage kernel is compiled for (Float64 / Float32 state).
#ifndef VH_MUSCL_WAVES
#define VH_MUSCL_WAVES 2
#endif
#ifndef VH_MUSCL_WAVES_F32
#define VH_MUSCL_WAVES_F32 4
#endif

template <typename T, int TX, int TY, bool STAGE2, bool FREC, int NF>
__global__ void __launch_bounds__(TX* TY, sizeof(T) == 4 ? VH_MUSCL_WAVES_F32 : VH_MUSCL_WAVES) muscl_stage_kernel(const T* __restrict__ Us, const T* U0, T* Uout,
                                                               T* __restrict__ frec, MusclDev G) {
  constexpr int NT = TX * TY;
  __shared__ MusLDS<T, TX, TY, NF> S;
  using L_ = MusLDS<T, TX, TY, NF>;
  const T g = T(G.gamma), gm1 = g - T(1), th = T(G.theta), sr = T(G.small), dtdx = T(G.dtdx), e1 = T(G.eta1);
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  const int nblk = G.tiles_x * G.tiles_y * G.chunks_z;
  // Workgroup i runs on XCD i % 8; give each XCD a contiguous range of tiles so
  // tiles that share halos share an L2 (any count of workgroups).
  int b = blockIdx.x;
  if (G.xcd_remap) {
    const int q = nblk / 8, r = nblk % 8, x = b % 8;
    b = x * q + min(x, r) + b / 8;
  }
  const int tile = b % (G.tiles_x * G.tiles_y);
  const int chunk = b / (G.tiles_x * G.tiles_y);
  const int x0 = (tile % G.tiles_x) * TX, y0 = (tile / G.tiles_x) * TY;
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2], ng = G.ng;
  const int kz0 = chunk * G.zchunk;
  const int kz1 = min(nz, kz0 + G.zchunk);
  const long long n = G.ncell;
  const long long plane = (long long)nx * ny;
  const int gx = x0 + tx, gy = y0 + ty;
  const bool inside = gx < nx && gy < ny;
  const long long col = (long long)clampi(gy, ny) * nx + clampi(gx, nx);
  auto zsrc = [&](int kk) { return col + (long long)clampi(kk, nz) * plane; };

  // ---- z pipeline prologue: planes kz0-2 .. kz0+2, the face below kz0 ----
  MuW<T> w0, w1, w2, zhi;   // planes k, k+1, k+2 of the column; hi z state of plane k
  T fzlo[NF];
  {
    const MuW<T> wm2 = to_prim<T, NF>(Us, n, zsrc(kz0 - 2), sr, g, gm1, e1);
    const MuW<T> wm1 = to_prim<T, NF>(Us, n, zsrc(kz0 - 1), sr, g, gm1, e1);
    w0 = to_prim<T, NF>(Us, n, zsrc(kz0), sr, g, gm1, e1);
    w1 = to_prim<T, NF>(Us, n, zsrc(kz0 + 1), sr, g, gm1, e1);
    w2 = to_prim<T, NF>(Us, n, zsrc(kz0 + 2), sr, g, gm1, e1);
    MuW<T> zlo;
    face_pair(wm1, w0, w1, th, sr, zhi, zlo);
    if (kz0 >= ng && kz0 <= nz - ng) hll_axis<T, NF>(face_state(wm2, wm1, w0, th, sr), zlo, 2, g, gm1, fzlo);
  }

  // Halo cells of a plane (two deep, cross-shaped): NHX in x, NHY in y.
  constexpr int NHX = 4 * TY, NHY = 4 * TX;
  auto halo_pos = [&](int h, int& lx, int& ly) {
    if (h < NHX) {
      const int r = h / 4, q = h % 4;
      ly = r;
      lx = q < 2 ? q - 2 : TX + (q - 2);
    } else {
      const int hh = h - NHX;
      const int r = hh / TX, q = hh % TX;
      lx = q;
      ly = r < 2 ? r - 2 : TY + (r - 2);
    }
  };
  T* wf = &S.w[0][0][0];
  T* ff = &S.f[0][0];

  for (int k = kz0; k < kz1; ++k) {
    const long long c = (long long)k * plane + (long long)gy * nx + gx;
    const int rot = (k + (int)blockIdx.x) % (NT / 64);
    const int lt = (tid + 64 * rot) % NT;   // rotating lane index for the side work

    // ---- P0: the plane's primitives (tile + halo) in LDS ----
    put_w(wf, L_::WS, (ty + 2) * L_::WX + tx + 2, w0);
    for (int h = lt; h < NHX + NHY; h += NT) {
      int lx, ly;
      halo_pos(h, lx, ly);
      const long long src = (long long)k * plane + (long long)clampi(y0 + ly, ny) * nx + clampi(x0 + lx, nx);
      put_w(wf, L_::WS, (ly + 2) * L_::WX + lx + 2, to_prim<T, NF>(Us, n, src, sr, g, gm1, e1));
    }
    __syncthreads();

    // ---- PA: face states; hi states to LDS, lo states kept in registers ----
    MuW<T> xlo, ylo;
    {
      const int ic = (ty + 2) * L_::WX + tx + 2;
      MuW<T> hi;
      face_pair(get_w(wf, L_::WS, ic - 1), w0, get_w(wf, L_::WS, ic + 1), th, sr, hi, xlo);
      put_w(ff, L_::FS, L_::xs(ty, tx + 1), hi);
      face_pair(get_w(wf, L_::WS, ic - L_::WX), w0, get_w(wf, L_::WS, ic + L_::WX), th, sr, hi, ylo);
      put_w(ff, L_::FS, L_::ys(ty + 1, tx), hi);
    }
    // Rim: hi state of the cell before the tile and lo state of the cell after
    // it, per row (x) and per column (y).
    constexpr int NRX = 2 * TY, NRY = 2 * TX;
    for (int h = lt; h < NRX + NRY; h += NT) {
      const bool isy = h >= NRX;
      const int q = isy ? h - NRX : h;
      const int line = q >> 1;
      const bool hi = q & 1;
      int ic, dst;
      if (!isy) {
        ic = (line + 2) * L_::WX + (hi ? TX + 2 : 1);
        dst = L_::xs(line, hi ? TX + 1 : 0);
      } else {
        ic = (hi ? TY + 2 : 1) * L_::WX + line + 2;
        dst = L_::ys(hi ? TY + 1 : 0, line);
      }
      const int d = isy ? L_::WX : 1;
      const MuW<T> wm = get_w(wf, L_::WS, ic - d), wc = get_w(wf, L_::WS, ic), wp = get_w(wf, L_::WS, ic + d);
      put_w(ff, L_::FS, dst, hi ? face_state(wp, wc, wm, th, sr) : face_state(wm, wc, wp, th, sr));
    }
    __syncthreads();

    // ---- PB: HLL fluxes (each face once), written over the L state ----
    if (gy < ny && gx >= ng && gx <= nx - ng) {
      T F[NF];
      const int sl = L_::xs(ty, tx);
      hll_axis<T, NF>(get_w(ff, L_::FS, sl), xlo, 0, g, gm1, F);
      put_f<T, NF>(ff, L_::FS, sl, F);
    }
    if (gx < nx && gy >= ng && gy <= ny - ng) {
      T F[NF];
      const int sl = L_::ys(ty, tx);
      hll_axis<T, NF>(get_w(ff, L_::FS, sl), ylo, 1, g, gm1, F);
      put_f<T, NF>(ff, L_::FS, sl, F);
    }
    // tile hi-edge faces: x face TX of each row, y face TY of each column
    for (int h = lt; h < TY + TX; h += NT) {
      const bool isy = h >= TY;
      const int line = isy ? h - TY : h;
      const int gp = isy ? y0 + TY : x0 + TX;
      const int na = isy ? ny : nx;
      const bool ok = (isy ? (x0 + line < nx) : (y0 + line < ny)) && gp >= ng && gp <= na - ng;
      if (ok) {
        const int sl = isy ? L_::ys(TY, line) : L_::xs(line, TX);
        const int srr = isy ? L_::ys(TY + 1, line) : L_::xs(line, TX + 1);
        T F[NF];
        hll_axis<T, NF>(get_w(ff, L_::FS, sl), get_w(ff, L_::FS, srr), isy ? 1 : 0, g, gm1, F);
        put_f<T, NF>(ff, L_::FS, sl, F);
      }
    }
    // z face k+½ of the lane's column: hi state of plane k, lo state of plane k+1
    T fzhi[NF];
    MuW<T> zhi1, zlo1;
    face_pair(w0, w1, w2, th, sr, zhi1, zlo1);
    if (k + 1 >= ng && k + 1 <= nz - ng) hll_axis<T, NF>(zhi, zlo1, 2, g, gm1, fzhi);
    __syncthreads();

    // ---- PC: divergence in the reference's order, RK combine, flux record ----
    if (inside) {
      const bool ax = gx >= ng && gx < nx - ng, ay = gy >= ng && gy < ny - ng, az = k >= ng && k < nz - ng;
      T acc[NF];
      for (int q = 0; q < NF; ++q) acc[q] = T(0);
      if (ax) {
        for (int q = 0; q < NF; ++q) {
          const T lo = ff[q * L_::FS + L_::xs(ty, tx)], hi = ff[q * L_::FS + L_::xs(ty, tx + 1)];
          acc[q] = acc[q] + (-(hi - lo) * dtdx);
        }
      }
      for (int q = 0; q < NF; ++q) {
        T t = T(0);
        if (ay) {
          const T lo = ff[q * L_::FS + L_::ys(ty, tx)], hi = ff[q * L_::FS + L_::ys(ty + 1, tx)];
          t = -(hi - lo) * dtdx;
        }
        acc[q] = acc[q] + t;
      }
      for (int q = 0; q < NF; ++q) {
        T t = T(0);
        if (az) t = -(fzhi[q] - fzlo[q]) * dtdx;
        acc[q] = acc[q] + t;
      }
      T o[NF];
      for (int q = 0; q < NF; ++q) {
        const T us = Us[q * n + c];
        if constexpr (STAGE2) o[q] = T(0.5) * (U0[q * n + c] + us + acc[q]);
        else o[q] = U0[q * n + c] + acc[q];   // U0: the stage's base (Us itself unless bc! refilled a copy)
      }
      if constexpr (STAGE2 && NF == 6) {
        // dual_energy_sync! (lib/PPMKernels/src/muscl_grid.jl:97-107) on the new state
#if VH_FAST_PROFILE
        const T d = o[0], inv = fast_rcp(d), u = o[1] * inv, v = o[2] * inv, w = o[3] * inv;
        const T tod = o[4] * inv, god = o[5] * inv;
#else
        const T d = o[0], u = o[1] / d, v = o[2] / d, w = o[3] / d;
        const T tod = o[4] / d, god = o[5] / d;
#endif
        const T v2 = u * u + v * v + w * w;
        const T e = dual_eint(tod, god, v2, g, gm1, e1, sr);
        o[5] = d * e;
        o[4] = d * (e + T(0.5) * v2);
      }
      // plain stores: non-temporal ones measured slower for this kernel in Float32
      for (int q = 0; q < NF; ++q) Uout[q * n + c] = o[q];
      if constexpr (FREC) {
        // frec[a][q][c]: ½F through the lo face of c along a, where that face is
        // one of the axis' active faces; stage 1 stores 0 + s (every cell), stage 2 adds s.
        const bool fx = gx >= ng && gx <= nx - ng, fy = gy >= ng && gy <= ny - ng, fz = k >= ng && k <= nz - ng;
        const T w = T(0.5);
        for (int q = 0; q < NF; ++q) {
          const T sxv = fx ? w * ff[q * L_::FS + L_::xs(ty, tx)] : T(0);
          const T syv = fy ? w * ff[q * L_::FS + L_::ys(ty, tx)] : T(0);
          const T szv = fz ? w * fzlo[q] : T(0);
          T* px = frec + (0 * NF + q) * n + c;
          T* py = frec + (1 * NF + q) * n + c;
          T* pz = frec + (2 * NF + q) * n + c;
          if constexpr (!STAGE2) {
            *px = T(0) + sxv;
            *py = T(0) + syv;
            *pz = T(0) + szv;
          } else {
            if (fx) *px = *px + sxv;
            if (fy) *py = *py + syv;
            if (fz) *pz = *pz + szv;
          }
        }
      }
    }

    // advance the z pipeline
    for (int q = 0; q < NF; ++q) fzlo[q] = fzhi[q];
    zhi = zhi1;
    w0 = w1;
    w1 = w2;
    if (k + 1 < kz1) w2 = to_prim<T, NF>(Us, n, zsrc(k + 3), sr, g, gm1, e1);
    __syncthreads();
  }
}


// One RK stage: template dispatch over dtype, tile shape, nf, stage and recording.
template <typename T, int TX, int TY, int NF>
inline void muscl_launch_nf(bool stage2, bool rec, const MusclDev& G, hipStream_t s, const void* Us, const void* U0,
                            void* Uout, void* frec) {
  const unsigned nb = (unsigned)(G.tiles_x * G.tiles_y * G.chunks_z);
  T* fr = (T*)frec;
  if (!stage2) {
    // stage 1 reads its base (U0, Us unless bc! refilled a copy) for every cell
    if (rec) muscl_stage_kernel<T, TX, TY, false, true, NF><<<nb, TX * TY, 0, s>>>((const T*)Us, (const T*)U0, (T*)Uout, fr, G);
    else muscl_stage_kernel<T, TX, TY, false, false, NF><<<nb, TX * TY, 0, s>>>((const T*)Us, (const T*)U0, (T*)Uout, fr, G);
  } else {
    if (rec) muscl_stage_kernel<T, TX, TY, true, true, NF><<<nb, TX * TY, 0, s>>>((const T*)Us, (const T*)U0, (T*)Uout, fr, G);
    else muscl_stage_kernel<T, TX, TY, true, false, NF><<<nb, TX * TY, 0, s>>>((const T*)Us, (const T*)U0, (T*)Uout, fr, G);
  }
}
template <typename T, int TX, int TY>
inline void muscl_launch_t(int nf, bool stage2, bool rec, const MusclDev& G, hipStream_t s, const void* Us,
                           const void* U0, void* Uout, void* frec) {
  if (nf == 6) muscl_launch_nf<T, TX, TY, 6>(stage2, rec, G, s, Us, U0, Uout, frec);
  else muscl_launch_nf<T, TX, TY, 5>(stage2, rec, G, s, Us, U0, Uout, frec);
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
