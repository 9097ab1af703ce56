// Unsplit SSP-RK2 MUSCL step of PPMKernels (lib/PPMKernels/src/muscl_grid.jl:212-370,
// lib/PPMKernels/src/muscl.jl) on a ghosted 3D grid, fused per RK stage (gfx950).
//
// The reference runs, per stage, a cons→prim kernel, then per axis a transpose, the
// PLM+HLL flux line over every pencil, a divergence kernel into a slab and an
// untranspose-add into dU, then the RK combine: ~20 kernels and ~60 full-grid
// passes per step. Here one launch per stage does it all for a tile of 16×16 (or
// 32×8) cells marching along z: the plane's primitives (plus a two-deep halo in x
// and y) are staged in LDS, each cell reconstructs its hi- and lo-face states per
// axis, each face's HLL flux is evaluated once, and each cell sums its divergence
// in the reference's order and writes the combined state; z faces are pipelined
// in registers down the column. Optional flux recording (½F per stage at the lo
// face of each cell, grid-order momenta) is written in the same pass.
//
// Arithmetic follows the reference operation for operation at the state's
// precision (compiled with -ffp-contract=off), so results are reproducible bit for bit.
#include <float.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/vespahip.h"
#include "host.h"

namespace vh {
namespace {

struct MusclDev {
  int n[3];
  int ng;
  int tiles_x, tiles_y, chunks_z, zchunk;
  int xcd_remap, nf;                  // nf: 5 fields, or 6 with the gas energy ρe
  long long ncell;
  double gamma, theta, small, dtdx;   // dtdx = T(dt)/T(dx), exact in T
  double eta1;                        // dual-energy selection threshold
};

template <typename T>
struct MuW {
  T r, e, vx, vy, vz;
};

// Julia's max / min on floats are IEEE 754-2019 maximum / minimum (NaN wins, −0 < +0).
template <typename T>
__device__ __forceinline__ T jmax(T x, T y) {
  return __builtin_elementwise_maximum(x, y);
}
template <typename T>
__device__ __forceinline__ T jmin(T x, T y) {
  return __builtin_elementwise_minimum(x, y);
}
__device__ __forceinline__ double absT(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float absT(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtT(float x) { return __builtin_sqrtf(x); }

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

template <typename T>
__device__ __forceinline__ T hll1(T ap, T am, T fl, T fr, T ul, T ur) {
  return (ap * fl + am * fr - ap * am * (ur - ul)) / (ap + am);
}

// _hll6's hydro fluxes (lib/PPMKernels/src/muscl.jl:37-58) for states whose
// velocities are already in (normal, t1, t2) order in (vx, vy, vz). F: D, normal,
// t1, t2 momentum, energy.
template <typename T, int NF>
__device__ __forceinline__ void hll_rot(const MuW<T>& L, const MuW<T>& R, T g, T gm1, T F[NF]) {
  const T h = T(0.5);
  const T rl = L.r, el = L.e, ul = L.vx, vl = L.vy, wl = L.vz;
  const T rr = R.r, er = R.e, ur = R.vx, vr = R.vy, wr = R.vz;
  const T v2l = ul * ul + vl * vl + wl * wl, pl = gm1 * rl * el, csl = sqrtT(g * pl / rl), etl = el + h * v2l;
  const T UlS1 = rl * ul, UlS2 = rl * vl, UlS3 = rl * wl, UlE = rl * etl;
  const T FlD = rl * ul, FlS1 = UlS1 * ul + pl, FlS2 = UlS2 * ul, FlS3 = UlS3 * ul, FlE = (UlE + pl) * ul;
  const T lpl = ul + csl, lml = ul - csl;
  const T v2r = ur * ur + vr * vr + wr * wr, pr = gm1 * rr * er, csr = sqrtT(g * pr / rr), etr = er + h * v2r;
  const T UrS1 = rr * ur, UrS2 = rr * vr, UrS3 = rr * wr, UrE = rr * etr;
  const T FrD = rr * ur, FrS1 = UrS1 * ur + pr, FrS2 = UrS2 * ur, FrS3 = UrS3 * ur, FrE = (UrE + pr) * ur;
  const T lpr = ur + csr, lmr = ur - csr;
  const T ap = jmax(T(0), jmax(lpl, lpr));
  const T am = jmax(T(0), jmax(-lml, -lmr));
  F[0] = hll1(ap, am, FlD, FrD, rl, rr);
  F[1] = hll1(ap, am, FlS1, FrS1, UlS1, UrS1);
  F[2] = hll1(ap, am, FlS2, FrS2, UlS2, UrS2);
  F[3] = hll1(ap, am, FlS3, FrS3, UlS3, UrS3);
  F[4] = hll1(ap, am, FlE, FrE, UlE, UrE);
  if constexpr (NF == 6) {   // _hll6's gas-energy flux: U = ρe, F = ρe·u
    const T UlG = rl * el, UrG = rr * er;
    F[5] = hll1(ap, am, UlG * ul, UrG * ur, UlG, UrG);
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
template <typename T>
__device__ __forceinline__ T dual_eint(T tau, T d, T ge, T v2, T g, T gm1, T eta1, T small) {
  const T eint_tot = tau / d - T(0.5) * v2;
  const T eint_dual = ge / d;
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
  const T u = U[n + c] / d, v = U[2 * n + c] / d, ww = U[3 * n + c] / d;
  w.r = d;
  w.vx = u;
  w.vy = v;
  w.vz = ww;
  if constexpr (NF == 6) w.e = dual_eint(U[4 * n + c], d, U[5 * n + c], u * u + v * v + ww * ww, g, gm1, eta1, sr);
  else w.e = jmax(U[4 * n + c] / d - T(0.5) * (u * u + v * v + ww * ww), sr);
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
        else o[q] = us + acc[q];
      }
      if constexpr (STAGE2 && NF == 6) {
        // dual_energy_sync! (lib/PPMKernels/src/muscl_grid.jl:97-107) on the new state
        const T d = o[0], u = o[1] / d, v = o[2] / d, w = o[3] / d;
        const T v2 = u * u + v * v + w * w;
        const T e = dual_eint(o[4], d, o[5], v2, g, gm1, e1, sr);
        o[5] = d * e;
        o[4] = d * (e + T(0.5) * v2);
      }
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

// fill_periodic! (lib/PPMKernels/src/muscl_grid.jl:644-670) for the five fields:
// every ghost cell copies the periodic image inside the interior.
template <typename T>
__global__ void __launch_bounds__(256) fill_periodic_kernel(T* __restrict__ U, MusclDev G) {
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2], ng = G.ng;
  const long long n = G.ncell;
  const int nix = nx - 2 * ng, niy = ny - 2 * ng, niz = nz - 2 * ng;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    const int i = (int)(c % nx);
    const long long t = c / nx;
    const int j = (int)(t % ny), k = (int)(t / ny);
    const bool gh = i < ng || i >= nx - ng || j < ng || j >= ny - ng || k < ng || k >= nz - ng;
    if (!gh) continue;
    const int si = i < ng ? i + nix : (i >= nx - ng ? i - nix : i);
    const int sj = j < ng ? j + niy : (j >= ny - ng ? j - niy : j);
    const int sk = k < ng ? k + niz : (k >= nz - ng ? k - niz : k);
    const long long s = ((long long)sk * ny + sj) * nx + si;
    for (int q = 0; q < G.nf; ++q) U[q * n + c] = U[q * n + s];
  }
}

// max_wavespeed (lib/PPMKernels/src/muscl_grid.jl:672-695): per-workgroup maxima.
template <typename T>
__global__ void __launch_bounds__(256) wavespeed_kernel(const T* __restrict__ U, MusclDev G,
                                                        double* __restrict__ partials) {
  __shared__ double red[4];
  const T g = T(G.gamma);
  const long long n = G.ncell;
  bool any = false;
  T m = T(0);
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
    const T d = U[c], u = U[n + c] / d, v = U[2 * n + c] / d, w = U[3 * n + c] / d;
    const T eint = U[4 * n + c] / d - T(0.5) * (u * u + v * v + w * w);
    const T cs = sqrtT(jmax(g * (g - T(1)) * eint, T(0)));
    const T s = cs + jmax(absT(u), jmax(absT(v), absT(w)));
    m = any ? jmax(m, s) : s;
    any = true;
  }
  double md = any ? (double)m : -HUGE_VAL;
  for (int o = 32; o > 0; o >>= 1) md = jmax(md, __shfl_xor(md, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = md;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = jmax(jmax(red[0], red[1]), jmax(red[2], red[3]));
}

// boundary_flux_register's gathers (lib/PPMKernels/src/reflux_capture.jl:52-77,86-131) for
// one axis a: the lo and hi boundary planes ([side][field][u2][u1]) and the interior
// hi faces of active cells 1..act-1 ([field][u2][u1][c]), each value Float64(F)·sc.
// b1 < b2 are the other two axes; u1, u2 run over their active cells.
template <typename T>
__global__ void __launch_bounds__(256) flux_register_kernel(const T* __restrict__ frec, MusclDev G, int a, double sc,
                                                            double* __restrict__ planes, double* __restrict__ inner) {
  const int b1 = a == 0 ? 1 : 0, b2 = a == 2 ? 1 : 2;
  const int ng = G.ng;
  const int act[3] = {G.n[0] - 2 * ng, G.n[1] - 2 * ng, G.n[2] - 2 * ng};
  const long long n = G.ncell;
  const long long str[3] = {1, G.n[0], (long long)G.n[0] * G.n[1]};
  const long long m1 = act[b1], m2 = act[b2];
  const long long nplane = 2 * 5 * m1 * m2;
  const long long ninner = inner && act[a] > 1 ? 5 * (long long)(act[a] - 1) * m1 * m2 : 0;
  for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < nplane + ninner; t += (long long)gridDim.x * 256) {
    long long q = t < nplane ? t : t - nplane;
    int p, comp;
    long long u1, u2;
    if (t < nplane) {
      u1 = q % m1;
      q /= m1;
      u2 = q % m2;
      q /= m2;
      comp = (int)(q % 5);
      p = q / 5 ? ng + act[a] : ng;
    } else {
      const long long c = q % (act[a] - 1);
      q /= act[a] - 1;
      u1 = q % m1;
      q /= m1;
      u2 = q % m2;
      comp = (int)(q / m2);
      p = ng + (int)c + 1;
    }
    const long long cell = p * str[a] + (ng + u1) * str[b1] + (ng + u2) * str[b2];
    const double v = (double)frec[((long long)a * G.nf + comp) * n + cell] * sc;
    if (t < nplane) planes[t] = v;
    else inner[t - nplane] = v;
  }
}

constexpr int kWsWgs = 1024;

}  // namespace
}  // namespace vh

using namespace vh;

struct vh_muscl {
  MusclDev G{};
  int dtype = VH_F64, device = 0;
  size_t esize = 8;
  double dx = 1.0;
  int variant = 0;   // 0: 16×16 tiles, 1: 32×8
  void* U = nullptr;
  void* U1 = nullptr;
  void* frec = nullptr;
  double* partials = nullptr;
  hipStream_t stream = nullptr;
};

namespace {

int tile_x(int variant) { return variant == 1 ? 32 : 16; }
int tile_y(int variant) { return variant == 1 ? 8 : 16; }

void set_tiles(vh_muscl* m) {
  const int TX = tile_x(m->variant), TY = tile_y(m->variant);
  m->G.tiles_x = (m->G.n[0] + TX - 1) / TX;
  m->G.tiles_y = (m->G.n[1] + TY - 1) / TY;
  m->G.chunks_z = (m->G.n[2] + m->G.zchunk - 1) / m->G.zchunk;
}

template <typename T, int TX, int TY, int NF>
void launch_stage_nf(vh_muscl* m, bool stage2, bool rec, const void* Us, const void* U0, void* Uout) {
  const unsigned nb = (unsigned)(m->G.tiles_x * m->G.tiles_y * m->G.chunks_z);
  T* fr = (T*)m->frec;
  if (!stage2) {
    if (rec) muscl_stage_kernel<T, TX, TY, false, true, NF><<<nb, TX * TY, 0, m->stream>>>((const T*)Us, nullptr, (T*)Uout, fr, m->G);
    else muscl_stage_kernel<T, TX, TY, false, false, NF><<<nb, TX * TY, 0, m->stream>>>((const T*)Us, nullptr, (T*)Uout, fr, m->G);
  } else {
    if (rec) muscl_stage_kernel<T, TX, TY, true, true, NF><<<nb, TX * TY, 0, m->stream>>>((const T*)Us, (const T*)U0, (T*)Uout, fr, m->G);
    else muscl_stage_kernel<T, TX, TY, true, false, NF><<<nb, TX * TY, 0, m->stream>>>((const T*)Us, (const T*)U0, (T*)Uout, fr, m->G);
  }
}

template <typename T, int TX, int TY>
void launch_stage_t(vh_muscl* m, bool stage2, bool rec, const void* Us, const void* U0, void* Uout) {
  if (m->G.nf == 6) launch_stage_nf<T, TX, TY, 6>(m, stage2, rec, Us, U0, Uout);
  else launch_stage_nf<T, TX, TY, 5>(m, stage2, rec, Us, U0, Uout);
}

void launch_stage(vh_muscl* m, bool stage2, bool rec, const void* Us, const void* U0, void* Uout) {
  if (m->dtype == VH_F64) {
    if (m->variant == 1) launch_stage_t<double, 32, 8>(m, stage2, rec, Us, U0, Uout);
    else launch_stage_t<double, 16, 16>(m, stage2, rec, Us, U0, Uout);
  } else {
    if (m->variant == 1) launch_stage_t<float, 32, 8>(m, stage2, rec, Us, U0, Uout);
    else launch_stage_t<float, 16, 16>(m, stage2, rec, Us, U0, Uout);
  }
}

// One step: U1 = U + dU(U), then U = ½((U + U1) + dU(U1)) written over U.
void enqueue_step(vh_muscl* m, double dt, bool rec) {
  const double dtdx = m->dtype == VH_F64 ? dt / m->dx : (double)((float)dt / (float)m->dx);
  m->G.dtdx = dtdx;
  launch_stage(m, false, rec, m->U, nullptr, m->U1);
  launch_stage(m, true, rec, m->U1, m->U, m->U);
}

}  // namespace

extern "C" {

int vh_muscl_create(vh_muscl** out, const int32_t* dims, int32_t ng, double gamma, double theta, double small_rho,
                    double dx, int32_t dtype) {
  if (!out || !dims) return fail(-1, "null argument%s", "");
  *out = nullptr;
  if (dtype != VH_F64 && dtype != VH_F32) return fail(-1, "dtype must be VH_F64 or VH_F32%s (got %lld)", "", dtype);
  if (ng < 2) return fail(-1, "the MUSCL stencil needs at least 2 ghost zones%s (got %lld)", "", ng);
  for (int d = 0; d < 3; ++d)
    if (dims[d] < 2 * ng + 1 || dims[d] > (1 << 20))
      return fail(-1, "each axis needs 2*ng+1 .. 2^20 cells%s (axis %lld)", "", d);
  if (!(gamma > 1.0) || !(dx > 0.0)) return fail(-1, "gamma must exceed 1 and dx be positive%s", "");
  const long long n = (long long)dims[0] * dims[1] * dims[2];
  if (n > 0x7fffffffLL) return fail(-1, "grid too large%s (%lld cells)", "", n);
  vh_muscl* m = new vh_muscl();
  for (int d = 0; d < 3; ++d) m->G.n[d] = dims[d];
  m->G.ng = ng;
  m->G.zchunk = 8;
  m->G.xcd_remap = 1;
  m->G.ncell = n;
  m->G.gamma = gamma;
  m->G.theta = theta;
  m->G.small = small_rho;
  m->G.nf = 5;
  m->G.eta1 = 1e-3;
  m->dx = dx;
  m->dtype = dtype;
  m->esize = dtype == VH_F64 ? 8 : 4;
  m->variant = 0;
  set_tiles(m);
  if (hipGetDevice(&m->device) != hipSuccess) {
    delete m;
    return fail(-2, "no HIP device%s", "");
  }
  const size_t bytes = (size_t)5 * n * m->esize;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess || hipMalloc(&m->U, bytes) != hipSuccess ||
      hipMalloc(&m->U1, bytes) != hipSuccess || hipMalloc(&m->partials, sizeof(double) * kWsWgs) != hipSuccess) {
    vh_muscl_destroy(m);
    return fail(-2, "device allocation of %s%lld bytes failed", "", (long long)(2 * bytes));
  }
  if (hipMemsetAsync(m->U, 0, bytes, m->stream) != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess) {
    vh_muscl_destroy(m);
    return fail(-2, "device initialisation failed%s", "");
  }
  *out = m;
  return 0;
}

int vh_muscl_destroy(vh_muscl* m) {
  if (!m) return 0;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->U);
  (void)hipFree(m->U1);
  (void)hipFree(m->frec);
  (void)hipFree(m->partials);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

int64_t vh_muscl_ncell(const vh_muscl* m) { return m ? m->G.ncell : -1; }

int vh_muscl_upload(vh_muscl* m, const void* U) {
  if (!m || !U) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipMemcpyAsync(m->U, U, (size_t)m->G.nf * m->G.ncell * m->esize, hipMemcpyHostToDevice, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_muscl_download(vh_muscl* m, void* U) {
  if (!m || !U) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipMemcpyAsync(U, m->U, (size_t)m->G.nf * m->G.ncell * m->esize, hipMemcpyDeviceToHost, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_muscl_state_device_ptr(vh_muscl* m, void** ptr) {
  if (!m || !ptr) return fail(-1, "null argument%s", "");
  *ptr = m->U;
  return 0;
}

int vh_muscl_enable_dual_energy(vh_muscl* m, double eta1) {
  if (!m) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  VH_HIP(hipStreamSynchronize(m->stream));
  m->G.eta1 = eta1;
  if (m->G.nf == 6) return 0;
  const size_t b5 = (size_t)5 * m->G.ncell * m->esize, b6 = (size_t)6 * m->G.ncell * m->esize;
  void *U = nullptr, *U1 = nullptr;
  if (hipMalloc(&U, b6) != hipSuccess || hipMalloc(&U1, b6) != hipSuccess) {
    (void)hipFree(U);
    return fail(-2, "device allocation of %s%lld bytes failed", "", (long long)(2 * b6));
  }
  VH_HIP(hipMemcpyAsync(U, m->U, b5, hipMemcpyDeviceToDevice, m->stream));
  VH_HIP(hipMemsetAsync((char*)U + b5, 0, b6 - b5, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  (void)hipFree(m->U);
  (void)hipFree(m->U1);
  (void)hipFree(m->frec);
  m->U = U;
  m->U1 = U1;
  m->frec = nullptr;
  m->G.nf = 6;
  return 0;
}

int vh_muscl_set_tiles(vh_muscl* m, int32_t variant, int32_t zchunk) {
  if (!m) return fail(-1, "null argument%s", "");
  if (variant < 0 || variant > 1) return fail(-1, "tile variant must be 0 (16x16) or 1 (32x8)%s", "");
  if (zchunk < 1 || zchunk > 1024) return fail(-1, "zchunk out of range%s (%lld)", "", zchunk);
  VH_HIP(hipStreamSynchronize(m->stream));
  m->variant = variant;
  m->G.zchunk = zchunk;
  set_tiles(m);
  return 0;
}

int vh_muscl_step(vh_muscl* m, double dt, int32_t record) {
  if (!m) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  if (record && !m->frec) VH_HIP(hipMalloc(&m->frec, (size_t)3 * m->G.nf * m->G.ncell * m->esize));
  enqueue_step(m, dt, record != 0);
  VH_HIP(hipGetLastError());
  return 0;
}

int vh_muscl_fluxrec_download(vh_muscl* m, void* frec) {
  if (!m || !frec) return fail(-1, "null argument%s", "");
  if (!m->frec) return fail(-1, "no step has recorded fluxes yet%s", "");
  VH_HIP(hipMemcpyAsync(frec, m->frec, (size_t)3 * m->G.nf * m->G.ncell * m->esize, hipMemcpyDeviceToHost,
                        m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_muscl_fill_periodic(vh_muscl* m) {
  if (!m) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  const unsigned nb = (unsigned)std::min<long long>(4096, (m->G.ncell + 255) / 256);
  if (m->dtype == VH_F64) fill_periodic_kernel<double><<<nb, 256, 0, m->stream>>>((double*)m->U, m->G);
  else fill_periodic_kernel<float><<<nb, 256, 0, m->stream>>>((float*)m->U, m->G);
  VH_HIP(hipGetLastError());
  return 0;
}

int vh_muscl_max_wavespeed(vh_muscl* m, double* out) {
  if (!m || !out) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(m->device));
  const int nb = (int)std::min<long long>(kWsWgs, (m->G.ncell + 255) / 256);
  if (m->dtype == VH_F64) wavespeed_kernel<double><<<nb, 256, 0, m->stream>>>((const double*)m->U, m->G, m->partials);
  else wavespeed_kernel<float><<<nb, 256, 0, m->stream>>>((const float*)m->U, m->G, m->partials);
  VH_HIP(hipGetLastError());
  std::vector<double> p(nb);
  VH_HIP(hipMemcpyAsync(p.data(), m->partials, sizeof(double) * nb, hipMemcpyDeviceToHost, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  double r = p[0];
  for (int i = 1; i < nb; ++i) r = std::isnan(r) ? r : (std::isnan(p[i]) || p[i] > r ? p[i] : r);
  *out = r;
  return 0;
}

int vh_muscl_flux_register(vh_muscl* m, double dt, double* planes, double* interior) {
  if (!m || !planes) return fail(-1, "null argument%s", "");
  if (!m->frec) return fail(-1, "no step has recorded fluxes yet%s", "");
  VH_HIP(hipSetDevice(m->device));
  const int ng = m->G.ng;
  const long long act[3] = {m->G.n[0] - 2LL * ng, m->G.n[1] - 2LL * ng, m->G.n[2] - 2LL * ng};
  long long np = 0, ni = 0;
  for (int a = 0; a < 3; ++a) {
    const long long t = act[(a + 1) % 3] * act[(a + 2) % 3];
    np += 10 * t;
    ni += 5 * (act[a] - 1) * t;
  }
  double* d = nullptr;
  VH_HIP(hipMalloc(&d, sizeof(double) * (np + (interior ? ni : 0))));
  const double sc = dt * (m->dx * m->dx);   // ∫F·area·dt = F·dt·dx²
  long long op = 0, oi = 0;
  hipError_t e = hipSuccess;
  for (int a = 0; a < 3 && e == hipSuccess; ++a) {
    const long long t = act[(a + 1) % 3] * act[(a + 2) % 3];
    const long long work = 10 * t + (interior ? 5 * (act[a] - 1) * t : 0);
    const unsigned nb = (unsigned)std::min<long long>(4096, (work + 255) / 256);
    double* pl = d + op;
    double* in = interior ? d + np + oi : nullptr;
    if (m->dtype == VH_F64)
      flux_register_kernel<double><<<nb, 256, 0, m->stream>>>((const double*)m->frec, m->G, a, sc, pl, in);
    else
      flux_register_kernel<float><<<nb, 256, 0, m->stream>>>((const float*)m->frec, m->G, a, sc, pl, in);
    e = hipGetLastError();
    op += 10 * t;
    oi += 5 * (act[a] - 1) * t;
  }
  if (e == hipSuccess) e = hipMemcpyAsync(planes, d, sizeof(double) * np, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess && interior)
    e = hipMemcpyAsync(interior, d + np, sizeof(double) * ni, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(-2, "HIP error %s in the flux register gather%s", hipGetErrorString(e));
  return 0;
}

int vh_muscl_sync(vh_muscl* m) {
  if (!m) return fail(-1, "null argument%s", "");
  VH_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int vh_muscl_time_steps(vh_muscl* m, double dt, int64_t nsteps, int32_t record, double* ms_step, double* ms_stage2) {
  if (!m || nsteps <= 0 || nsteps > 100000) return fail(-1, "bad argument%s (nsteps %lld)", "", nsteps);
  VH_HIP(hipSetDevice(m->device));
  if (record && !m->frec) VH_HIP(hipMalloc(&m->frec, (size_t)3 * m->G.nf * m->G.ncell * m->esize));
  enqueue_step(m, dt, record != 0);   // untimed warm-up
  hipEvent_t e0, e1, s0, s1;
  VH_HIP(hipEventCreate(&e0));
  VH_HIP(hipEventCreate(&e1));
  VH_HIP(hipEventCreate(&s0));
  VH_HIP(hipEventCreate(&s1));
  VH_HIP(hipEventRecord(e0, m->stream));
  for (int64_t k = 0; k < nsteps; ++k) enqueue_step(m, dt, record != 0);
  VH_HIP(hipEventRecord(e1, m->stream));
  VH_HIP(hipStreamSynchronize(m->stream));
  float tot = 0;
  VH_HIP(hipEventElapsedTime(&tot, e0, e1));
  // stage-2 launches timed one by one
  double st2 = 0;
  const int k2 = (int)std::min<int64_t>(nsteps, 20);
  for (int k = 0; k < k2; ++k) {
    launch_stage(m, false, record != 0, m->U, nullptr, m->U1);
    VH_HIP(hipEventRecord(s0, m->stream));
    launch_stage(m, true, record != 0, m->U1, m->U, m->U);
    VH_HIP(hipEventRecord(s1, m->stream));
    VH_HIP(hipStreamSynchronize(m->stream));
    float a = 0;
    VH_HIP(hipEventElapsedTime(&a, s0, s1));
    st2 += a;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(s0);
  (void)hipEventDestroy(s1);
  VH_HIP(hipGetLastError());
  if (ms_step) *ms_step = (double)tot / nsteps;
  if (ms_stage2) *ms_stage2 = st2 / k2;
  return 0;
}

}  // extern "C"
