// Fused SSP-RK2 stage kernels for a uniform ghost-free grid (gfx950).
//
// One kernel launch is one RK stage over the whole grid: PLM/minmod
// reconstruction, HLLC at every face, the flux-divergence accumulation and the
// stage update, all from a single read of the stencil field. The second stage
// also fuses the SSP combine and the FP64 CFL reduction of the new state.
//
// Work decomposition (2.5-D march): a workgroup owns a TX×TY column of cells and
// marches through a chunk of z planes. For the current plane it stages the
// primitive state with a two-cell halo in x and y in LDS (a cross-shaped
// footprint; corners are never read), computes every x- and y-face flux of the
// tile once into LDS, while each lane carries its own column's z pipeline
// (W(k), W(k+1), W(k+2), the z slope and the z-lo face flux) in registers, so
// every face flux in the grid is evaluated exactly once per chunk plane.
//
// Accumulation order. The reference walks faces cell by cell, column-major,
// axis by axis (lib/RefMesh/src/RefMesh.jl:187-197) and scatters ±F·area into
// the two cells (src/driver.jl:236-281). For one cell that order is:
//   1. lo faces whose partner precedes the cell (non-wrapped interior), z, y, x;
//   2. per axis x, y, z: its own hi face, then the lo face if it is a domain
//      boundary (emitted right after the hi face of the same axis);
//   3. lo faces wrapped by a periodic boundary (partner visited later), x, y, z.
// The gather below adds the six face fluxes in exactly that order, so the
// floating-point sum matches the reference scatter bit for bit.
#pragma once

#include "hydro.h"

namespace vh {

struct GridDev {
  int n[3];        // cells per axis (axes beyond rank: 1)
  int bc[6];       // [2*axis + side], side 0 = lo, 1 = hi
  int rank;
  int zchunk;      // planes per workgroup along z
  int tiles_x;     // tiles along x
  int tiles_y;     // tiles along y
  int chunks_z;    // z chunks
  int xcd_remap;   // 1: contiguous tile ranges per XCD
  long long ncell;
  double gamma;
  double V;        // cell volume (geometry precision)
  double w[3];     // cell widths
  double area[3];  // face areas
};

// Map a (possibly out-of-range) coordinate on one axis to the cell that
// supplies its state, following neighbor() (lib/RefMesh/src/RefMesh.jl:147-161).
// Returns the source coordinate and sets `gh` to the boundary side whose ghost
// applies (0: none, 1: lo, 2: hi).
__device__ __forceinline__ int resolve(int i, int n, int bclo, int &gh) {
  gh = 0;
  if (i >= 0 && i < n) return i;
  if (bclo == BC_PERIODIC) {
    int r = i % n;
    return r < 0 ? r + n : r;
  }
  if (i < 0) { gh = 1; return 0; }
  gh = 2;
  return n - 1;
}

template <typename T>
__device__ __forceinline__ Cons<T> load_cons(const T* __restrict__ U, long long n, long long c) {
  Cons<T> u;
  u.d = U[c];
  u.mx = U[n + c];
  u.my = U[2 * n + c];
  u.mz = U[3 * n + c];
  u.e = U[4 * n + c];
  return u;
}

template <typename T>
__device__ __forceinline__ void store_cons(T* __restrict__ U, long long n, long long c, const Cons<T>& u) {
  U[c] = u.d;
  U[n + c] = u.mx;
  U[2 * n + c] = u.my;
  U[3 * n + c] = u.mz;
  U[4 * n + c] = u.e;
}

// Lane-level block maximum of a double (wave64 butterfly + LDS across waves).
template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = vmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r = vmax(r, red[i]);
  return r;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r += red[i];
  return r;
}

// Per-cell CFL term of compute_dt (src/driver.jl:178-185), Tr = double.
template <typename T, int RANK>
__device__ __forceinline__ double cfl_term(const Prim<T>& W, T g, const GridDev& G) {
  const T c = vsqrt(g * W.p / W.r);
  double m = double(vabs(W.vx) + c) / G.w[0];
  if (RANK > 1) m = vmax(m, double(vabs(W.vy) + c) / G.w[1]);
  if (RANK > 2) m = vmax(m, double(vabs(W.vz) + c) / G.w[2]);
  return m;
}

// LDS footprint of one plane (primitive state with halo 2 in x and y) plus the
// x- and y-face flux tiles.
template <typename T, int TX, int TY>
struct PlaneLDS {
  T w[5][TY + 4][TX + 4];
  T fx[5][TY][TX + 1];
  T fy[5][TY + 1][TX];
  double red[16];
};

template <typename T, int TX, int TY>
__device__ __forceinline__ void lds_put_w(PlaneLDS<T, TX, TY>& S, int sy, int sx, const Prim<T>& w) {
  S.w[0][sy][sx] = w.r;
  S.w[1][sy][sx] = w.vx;
  S.w[2][sy][sx] = w.vy;
  S.w[3][sy][sx] = w.vz;
  S.w[4][sy][sx] = w.p;
}
template <typename T, int TX, int TY>
__device__ __forceinline__ Prim<T> lds_get_w(const PlaneLDS<T, TX, TY>& S, int sy, int sx) {
  Prim<T> w;
  w.r = S.w[0][sy][sx];
  w.vx = S.w[1][sy][sx];
  w.vy = S.w[2][sy][sx];
  w.vz = S.w[3][sy][sx];
  w.p = S.w[4][sy][sx];
  return w;
}

// State of a column cell at plane index kk (any integer) with the z boundary
// rule applied: periodic wrap, or the ghost of the nearest boundary cell.
template <typename T, bool FLOOR64>
__device__ __forceinline__ Prim<T> column_w(const T* __restrict__ U, const GridDev& G, long long col,
                                            int kk, T g) {
  int gh;
  const int src = resolve(kk, G.n[2], G.bc[4], gh);
  const long long plane = (long long)G.n[0] * G.n[1];
  Prim<T> w = cons2prim<T, FLOOR64>(load_cons(U, G.ncell, col + (long long)src * plane), g, G.gamma);
  if (gh) w = ghost(w, G.bc[4 + gh - 1], 3);
  return w;
}

// z-face flux between column cells k and k+1 given W and z slopes of both.
// Handles the domain-boundary faces at k = -1 (lo) and k = nz-1 (hi).
template <typename T>
__device__ __forceinline__ Cons<T> zface(int k, const GridDev& G, const Prim<T>& wa, const Prim<T>& sa,
                                         const Prim<T>& wb, const Prim<T>& sb, T g) {
  const int nz = G.n[2];
  const bool per = G.bc[4] == BC_PERIODIC;
  Prim<T> L, R;
  if (!per && k < 0) {
    R = face_value(wb, sb, T(-0.5));
    L = ghost(R, G.bc[4], 3);
  } else if (!per && k >= nz - 1) {
    L = face_value(wa, sa, T(0.5));
    R = ghost(L, G.bc[5], 3);
  } else {
    L = face_value(wa, sa, T(0.5));
    R = face_value(wb, sb, T(-0.5));
  }
  return hllc(L, R, g, 3);
}

// One RK stage. STAGE2 = false: Uout = Us − (dt/V)·acc(Us).
// STAGE2 = true: Uout = ½·U0 + ½·(Us − (dt/V)·acc(Us)) written over U0 (point
// reads only), plus the per-workgroup max of the CFL term of the new state.
template <typename T, int RANK, int TX, int TY, bool STAGE2, bool FLOOR64>
__global__ void __launch_bounds__(TX* TY) rk_stage_kernel(const T* __restrict__ Us, const T* U0,
                                                          T* Uout, const double* __restrict__ dtp,
                                                          double* __restrict__ partials, GridDev G) {
  constexpr int NT = TX * TY;
  __shared__ PlaneLDS<T, TX, TY> S;
  const T g = T(G.gamma);
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;

  // Workgroup → (tile, z chunk). Consecutive hardware dispatch slots go round
  // robin over the 8 XCDs; remapping gives each XCD a contiguous tile range so
  // tiles that share halos share an L2.
  const int nblk = G.tiles_x * G.tiles_y * G.chunks_z;
  int b = blockIdx.x;
  if (G.xcd_remap && (nblk % 8) == 0) b = (b % 8) * (nblk / 8) + b / 8;
  const int tile = b % (G.tiles_x * G.tiles_y);
  const int chunk = b / (G.tiles_x * G.tiles_y);
  const int x0 = (tile % G.tiles_x) * TX, y0 = (tile / G.tiles_x) * TY;
  const int nx = G.n[0], ny = G.n[1], nz = G.n[2];
  const int kz0 = chunk * G.zchunk;
  const int kz1 = min(nz, kz0 + G.zchunk);
  const long long n = G.ncell;
  const long long plane = (long long)nx * ny;

  const double dt = *dtp;
  const T invV = T(dt) / (T(1.0) * T(G.V));
  const T a0 = T(G.area[0]), a1 = T(G.area[1]), a2 = T(G.area[2]);

  const int gx = x0 + tx, gy = y0 + ty;
  const bool active = gx < nx && gy < ny;
  // The lane's own column (clamped so inactive lanes read valid memory).
  int ghx, ghy;
  const int sxc = resolve(gx, nx, G.bc[0], ghx);
  const int syc = resolve(gy, ny, G.bc[2], ghy);
  const long long col = (long long)syc * nx + sxc;

  // ---- z pipeline prologue ----
  Prim<T> wz0, wz1, wz2, sz0;
  Cons<T> fzlo;
  if constexpr (RANK == 3) {
    const Prim<T> wm2 = column_w<T, FLOOR64>(Us, G, col, kz0 - 2, g);
    const Prim<T> wm1 = column_w<T, FLOOR64>(Us, G, col, kz0 - 1, g);
    wz0 = column_w<T, FLOOR64>(Us, G, col, kz0, g);
    wz1 = column_w<T, FLOOR64>(Us, G, col, kz0 + 1, g);
    wz2 = column_w<T, FLOOR64>(Us, G, col, kz0 + 2, g);
    // slope of cell kz0-1 needs its own ghost rule at the lo boundary: if kz0-1
    // is the ghost plane its slope is never used (the lo face is a boundary face).
    const Prim<T> sm1 = slope(wm2, wm1, wz0);
    sz0 = slope(wm1, wz0, wz1);
    fzlo = zface(kz0 - 1, G, wm1, sm1, wz0, sz0, g);
  }

  double cfl_max = 0.0;

  for (int k = kz0; k < kz1; ++k) {
    // ---- stage the plane's primitive state in LDS ----
    Prim<T> wc;
    if constexpr (RANK == 3) {
      wc = wz0;
    } else {
      wc = cons2prim<T, FLOOR64>(load_cons(Us, n, col), g, G.gamma);
    }
    {
      // Own slot: apply the x/y ghost when the lane sits past a non-periodic edge.
      Prim<T> w = wc;
      if (ghx) w = ghost(w, G.bc[ghx - 1], 1);
      if (ghy) w = ghost(w, G.bc[2 + ghy - 1], 2);
      lds_put_w(S, ty + 2, tx + 2, w);
    }
    constexpr int NHX = 4 * TY;
    constexpr int NHY = (RANK > 1) ? 4 * TX : 0;
    for (int h = tid; h < NHX + NHY; h += NT) {
      int lx, ly;
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
      int hgx, hgy;
      const int sx = resolve(x0 + lx, nx, G.bc[0], hgx);
      const int sy = resolve(y0 + ly, ny, G.bc[2], hgy);
      Prim<T> w = cons2prim<T, FLOOR64>(load_cons(Us, n, (long long)k * plane + (long long)sy * nx + sx), g,
                                        G.gamma);
      if (hgx) w = ghost(w, G.bc[hgx - 1], 1);
      if (hgy) w = ghost(w, G.bc[2 + hgy - 1], 2);
      lds_put_w(S, ly + 2, lx + 2, w);
    }
    __syncthreads();

    // ---- x- and y-face fluxes of the tile (each face once) ----
    constexpr int NFX = (TX + 1) * TY;
    constexpr int NFY = (RANK > 1) ? TX * (TY + 1) : 0;
    for (int f = tid; f < NFX + NFY; f += NT) {
      int ax, fi, fj;  // ax 0: x-face at local x = fi (between fi-1 and fi), row fj
      if (f < NFX) {
        ax = 0; fi = f % (TX + 1); fj = f / (TX + 1);
      } else {
        ax = 1; fi = (f - NFX) % TX; fj = (f - NFX) / TX;
      }
      // Global index of the cell on the +side of the face along ax, and row.
      const int gpos = ax == 0 ? x0 + fi : y0 + fj;
      const int grow = ax == 0 ? y0 + fj : x0 + fi;
      const int nax = ax == 0 ? nx : ny;
      const int nrow = ax == 0 ? ny : nx;
      if (grow >= nrow || gpos > nax) continue;
      // LDS slots of the four cells gpos-2 .. gpos+1 along ax.
      const int sy0 = ax == 0 ? fj + 2 : fj;        // slot of gpos-2 (y)
      const int sx0 = ax == 0 ? fi : fi + 2;        // slot of gpos-2 (x)
      const int dy = ax == 0 ? 0 : 1, dx = ax == 0 ? 1 : 0;
      const Prim<T> wa2 = lds_get_w(S, sy0, sx0);
      const Prim<T> wa = lds_get_w(S, sy0 + dy, sx0 + dx);
      const Prim<T> wb = lds_get_w(S, sy0 + 2 * dy, sx0 + 2 * dx);
      const Prim<T> wb2 = lds_get_w(S, sy0 + 3 * dy, sx0 + 3 * dx);
      const bool per = G.bc[2 * ax] == BC_PERIODIC;
      const int dim = ax + 1;
      Prim<T> L, R;
      if (!per && gpos == 0) {
        R = face_value(wb, slope(wa, wb, wb2), T(-0.5));
        L = ghost(R, G.bc[2 * ax], dim);
      } else if (!per && gpos == nax) {
        L = face_value(wa, slope(wa2, wa, wb), T(0.5));
        R = ghost(L, G.bc[2 * ax + 1], dim);
      } else {
        L = face_value(wa, slope(wa2, wa, wb), T(0.5));
        R = face_value(wb, slope(wa, wb, wb2), T(-0.5));
      }
      const Cons<T> F = hllc(L, R, g, dim);
      if (ax == 0) {
        S.fx[0][fj][fi] = F.d; S.fx[1][fj][fi] = F.mx; S.fx[2][fj][fi] = F.my;
        S.fx[3][fj][fi] = F.mz; S.fx[4][fj][fi] = F.e;
      } else {
        S.fy[0][fj][fi] = F.d; S.fy[1][fj][fi] = F.mx; S.fy[2][fj][fi] = F.my;
        S.fy[3][fj][fi] = F.mz; S.fy[4][fj][fi] = F.e;
      }
    }

    // ---- z-face at k+½ for the lane's column ----
    Cons<T> fzhi;
    Prim<T> sz1;
    if constexpr (RANK == 3) {
      sz1 = slope(wz0, wz1, wz2);
      fzhi = zface(k, G, wz0, sz0, wz1, sz1, g);
    }
    __syncthreads();

    // ---- gather the six face fluxes in the reference order and update ----
    if (active) {
      const long long c = (long long)k * plane + (long long)gy * nx + gx;
      Cons<T> fxlo, fxhi, fylo, fyhi;
      fxlo.d = S.fx[0][ty][tx]; fxlo.mx = S.fx[1][ty][tx]; fxlo.my = S.fx[2][ty][tx];
      fxlo.mz = S.fx[3][ty][tx]; fxlo.e = S.fx[4][ty][tx];
      fxhi.d = S.fx[0][ty][tx + 1]; fxhi.mx = S.fx[1][ty][tx + 1]; fxhi.my = S.fx[2][ty][tx + 1];
      fxhi.mz = S.fx[3][ty][tx + 1]; fxhi.e = S.fx[4][ty][tx + 1];
      if constexpr (RANK > 1) {
        fylo.d = S.fy[0][ty][tx]; fylo.mx = S.fy[1][ty][tx]; fylo.my = S.fy[2][ty][tx];
        fylo.mz = S.fy[3][ty][tx]; fylo.e = S.fy[4][ty][tx];
        fyhi.d = S.fy[0][ty + 1][tx]; fyhi.mx = S.fy[1][ty + 1][tx]; fyhi.my = S.fy[2][ty + 1][tx];
        fyhi.mz = S.fy[3][ty + 1][tx]; fyhi.e = S.fy[4][ty + 1][tx];
      }
      // Lo-face kinds: 0 interior (partner earlier), 1 domain boundary, 2 periodic wrap.
      const int kx = gx > 0 ? 0 : (G.bc[0] == BC_PERIODIC ? 2 : 1);
      const int ky = gy > 0 ? 0 : (G.bc[2] == BC_PERIODIC ? 2 : 1);
      const int kz = k > 0 ? 0 : (G.bc[4] == BC_PERIODIC ? 2 : 1);
      Cons<T> acc;
      acc.d = acc.mx = acc.my = acc.mz = acc.e = T(0);
      if (RANK > 2 && kz == 0) acc_add(acc, fzlo, -a2);
      if (RANK > 1 && ky == 0) acc_add(acc, fylo, -a1);
      if (kx == 0) acc_add(acc, fxlo, -a0);
      acc_add(acc, fxhi, a0);
      if (kx == 1) acc_add(acc, fxlo, -a0);
      if constexpr (RANK > 1) {
        acc_add(acc, fyhi, a1);
        if (ky == 1) acc_add(acc, fylo, -a1);
      }
      if constexpr (RANK > 2) {
        acc_add(acc, fzhi, a2);
        if (kz == 1) acc_add(acc, fzlo, -a2);
      }
      if (kx == 2) acc_add(acc, fxlo, -a0);
      if (RANK > 1 && ky == 2) acc_add(acc, fylo, -a1);
      if (RANK > 2 && kz == 2) acc_add(acc, fzlo, -a2);

      const Cons<T> us = load_cons(Us, n, c);
      Cons<T> o;
      if constexpr (!STAGE2) {
        // _euler_apply!: dst = base − invV·acc (src/driver.jl:297-307).
        o.d = us.d - invV * acc.d;
        o.mx = us.mx - invV * acc.mx;
        o.my = us.my - invV * acc.my;
        o.mz = us.mz - invV * acc.mz;
        o.e = us.e - invV * acc.e;
      } else {
        // _rk2_combine!: ½·u0 + ½·(u1 − invV·acc) (src/driver.jl:310-321).
        const Cons<T> u0 = load_cons(U0, n, c);
        const T h = T(0.5);
        o.d = h * u0.d + h * (us.d - invV * acc.d);
        o.mx = h * u0.mx + h * (us.mx - invV * acc.mx);
        o.my = h * u0.my + h * (us.my - invV * acc.my);
        o.mz = h * u0.mz + h * (us.mz - invV * acc.mz);
        o.e = h * u0.e + h * (us.e - invV * acc.e);
        cfl_max = vmax(cfl_max, cfl_term<T, RANK>(cons2prim<T, FLOOR64>(o, g, G.gamma), g, G));
      }
      store_cons(Uout, n, c, o);
    }

    if constexpr (RANK == 3) {
      fzlo = fzhi;
      sz0 = sz1;
      wz0 = wz1;
      wz1 = wz2;
      if (k + 1 < kz1) wz2 = column_w<T, FLOOR64>(Us, G, col, k + 3, g);
    }
    __syncthreads();
  }

  if constexpr (STAGE2) {
    const double m = block_max<NT>(cfl_max, S.red);
    if (tid == 0) partials[blockIdx.x] = m;
  }
}

}  // namespace vh
