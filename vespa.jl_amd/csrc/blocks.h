// Batched block-structured AMR stage kernels (gfx950).
//
// A mesh is a set of blocks of uniform cells at a refinement level (2:1
// balanced), leaves ordered block by block and column-major inside a block.
// Faces follow HGBackend's enumeration (lib/HGBackend/src/HGBackend.jl:202-276):
// a conforming face is emitted once from its hi side, a coarse↔fine face once
// from the fine side as a sub-face with the fine area; reconstruction across a
// block face uses one representative neighbour per face (the neighbour cell, the
// coarse cell containing the position, or the fine cell at the lo corner of the
// coarse face).
//
// Per RK stage:
//   link_flux_kernel   one lane per face between two different blocks: the
//                      HLLC flux from representative-neighbour reconstruction,
//                      written to a face buffer;
//   block_stage_kernel one persistent launch over a device tile list; each
//                      workgroup runs the 2.5-D plane march of stencil.h on its
//                      tiles (runtime tile shape, block geometry per tile) and,
//                      for cells that touch link faces, takes those fluxes from
//                      the buffer and adds everything in the order of the face
//                      walk: foreign groups of lower-index blocks, the cell's own
//                      block contributions (the uniform rule, with own link faces
//                      at their walk positions), then foreign groups of
//                      higher-index blocks.
#pragma once

#include "stencil.h"

namespace vh {
namespace VH_PROFILE_NS {

// Waves per SIMD the block stage kernel is compiled for (Float64 / Float32 state).
#ifndef VH_BLOCK_MIN_WAVES
#define VH_BLOCK_MIN_WAVES 3
#endif
#ifndef VH_BLOCK_MIN_WAVES_F32
#define VH_BLOCK_MIN_WAVES_F32 3
#endif

constexpr int BNT = 256;     // lanes per workgroup
constexpr int BW = 432;      // max (TY+4)*(TX+4) over the tile shapes used
constexpr int BFS = 592;     // max TY*(TX+2) + (TY+2)*TX (TX, TY <= 32, TX*TY <= 256)

// Face slots of a TX×TY tile, one array per field (as PlaneLDS in stencil.h):
//   x slot (r, s) = r*(TX+2) + s, s = 0..TX: hi face state of cell s-1 of row r,
//     then the flux of x face s; s = TX+1: lo face state of halo cell TX;
//   y slot (r, c) = TY*(TX+2) + r*TX + c, r = 0..TY: hi face state of cell r-1
//     of column c, then the flux of y face r; r = TY+1: lo state of halo cell TY.
// The final workgroup reduction reuses the start of w (after a barrier).
// Compile-time tile shapes (TX×TY, one lane per cell) have their exact LDS
// footprint; the runtime shape (CTX = 0) the maximum above.
template <typename T, int CTX = 0, int CTY = 0>
struct alignas(16) BlockLDS {
  static constexpr int W = CTX ? (CTY + 4) * (CTX + 4) : BW;
  static constexpr int F = CTX ? CTY * (CTX + 2) + (CTY + 2) * CTX : BFS;
  T w[5][W];
  T f[5][F];
};

template <typename L, typename T>
__device__ __forceinline__ void bl_putw(L& S, int i, const Prim<T>& w) {
  S.w[0][i] = w.r; S.w[1][i] = w.vx; S.w[2][i] = w.vy; S.w[3][i] = w.vz; S.w[4][i] = w.p;
}

// Per-tile block description kept in registers (the rest of BlockDev is read
// from memory where it is needed).
struct BlockGeo {
  int nx, ny, nz;
  long long off;
};

// Face-plane index of a cell on the face normal to axis d (column-major over the
// other two axes), used to look up the per-face maps of a block.
__device__ __forceinline__ int plane_index(const BlockGeo& bg, int d, int i, int j, int k) {
  return d == 0 ? k * bg.ny + j : (d == 1 ? k * bg.nx + i : j * bg.nx + i);
}

// Source cell (global index) of position (i, j, k) of a block, out of range on
// axis D only: periodic wrap inside the block, the representative neighbour
// across a face that touches another block, else the nearest boundary cell
// with gh = 1 (lo) or 2 (hi) telling the caller to take its ghost.
template <int D>
__device__ __forceinline__ long long block_src(const BlockGeo& bg, const BlockDev* __restrict__ bk,
                                               const int* __restrict__ repmap, int bclo, int i, int j, int k,
                                               int& gh) {
  gh = 0;
  const int n = D == 0 ? bg.nx : (D == 1 ? bg.ny : bg.nz);
  int p = D == 0 ? i : (D == 1 ? j : k);
  if (p < 0 || p >= n) {
    const int side = p < 0 ? 0 : 1;
    if (bclo == BC_PERIODIC) {
      const int r = p % n;
      p = r < 0 ? r + n : r;
    } else {
      p = side ? n - 1 : 0;
      const int ro = bk->repoff[2 * D + side];
      const int ii = D == 0 ? p : i, jj = D == 1 ? p : j, kk = D == 2 ? p : k;
      if (ro >= 0) {
        const int rep = repmap[ro + plane_index(bg, D, ii, jj, kk)];
        if (rep >= 0) return rep;
      }
      gh = side + 1;
    }
  }
  const int ii = D == 0 ? p : i, jj = D == 1 ? p : j, kk = D == 2 ? p : k;
  return bg.off + ((long long)kk * bg.ny + jj) * bg.nx + ii;
}

template <typename T, bool FLOOR64>
__device__ __forceinline__ Prim<T> link_face_state(const T* __restrict__ U, long long n, int c, int lo, int hi,
                                                   int axis, T g, double gd, T h) {
  const Prim<T> wc = cons2prim<T, FLOOR64>(load_cons(U, n, c), g, gd);
  const Prim<T> wl = lo >= 0 ? cons2prim<T, FLOOR64>(load_cons(U, n, lo), g, gd) : ghost(wc, -1 - lo, axis + 1);
  const Prim<T> wr = hi >= 0 ? cons2prim<T, FLOOR64>(load_cons(U, n, hi), g, gd) : ghost(wc, -1 - hi, axis + 1);
  return face_value(wc, slope(wl, wc, wr), h);
}

// One lane per link face: _face_value of both cells with their representative
// neighbours (src/driver.jl:227-233), then the Riemann flux (src/riemann.jl:33-88).
template <typename T, bool FLOOR64>
__global__ void __launch_bounds__(256) link_flux_kernel(const T* __restrict__ U, BlocksDev B, T* __restrict__ F) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= B.nfaces) return;
  const LinkFaceDev* L = B.faces + f;
  const int iL = L->iL, iR = L->iR, axis = L->axis;
  const int l0 = L->nL[0], l1 = L->nL[1], r0 = L->nR[0], r1 = L->nR[1];
  const T g = T(B.gamma), igm1 = T(B.igm1);
  const Prim<T> WL = link_face_state<T, FLOOR64>(U, B.ncell, iL, l0, l1, axis, g, B.gamma, T(0.5));
  const Prim<T> WR = link_face_state<T, FLOOR64>(U, B.ncell, iR, r0, r1, axis, g, B.gamma, T(-0.5));
  // constant axis per branch keeps the normal-component selection in registers
  const Cons<T> o = axis == 0 ? hllc(WL, WR, g, igm1, 1) : (axis == 1 ? hllc(WL, WR, g, igm1, 2) : hllc(WL, WR, g, igm1, 3));
  T* out = F + 5 * (long long)f;
  out[0] = o.d; out[1] = o.mx; out[2] = o.my; out[3] = o.mz; out[4] = o.e;
}

// Contribution of a group of link faces to cell c: ±area·F in stored order
// (_acc_add! with ±area, src/driver.jl:98-99, 236-251).
template <typename T>
__device__ __forceinline__ void add_link_group(Cons<T>& acc, const BlocksDev& B, const T* __restrict__ F, int start,
                                               int count, long long c) {
  for (int q = 0; q < count; ++q) {
    const int f = B.slot_faces[start + q];
    const LinkFaceDev* L = B.faces + f;
    const T a = T(L->area);
    const T s = (c == (long long)L->iL) ? a : -a;
    Cons<T> fl;
    const T* p = F + 5 * (long long)f;
    fl.d = p[0]; fl.mx = p[1]; fl.my = p[2]; fl.mz = p[3]; fl.e = p[4];
    acc_add(acc, fl, s);
  }
}


// One RK stage over a list of tiles (persistent: workgroup w takes tiles w,
// w + gridDim.x, ...). STAGE2 as in rk_stage_kernel; the CFL partial of the
// new state covers the workgroup's tiles.
// apply_gravity_source! (src/gravity.jl:412-447) for cell c of a block with
// volume V: (0, −ρg_x·V, −ρg_y·V, −ρg_z·V, −(ρv·g)·V) added to the flux sum, g
// read from the step's field (axes beyond the rank hold 0).
template <typename T>
__device__ __forceinline__ void block_grav_add(Cons<T>& acc, const Cons<T>& U, const T* __restrict__ gf, long long n,
                                               long long c, T V) {
  const T g0 = gf[c], g1 = gf[n + c], g2 = gf[2 * n + c];
  const T rho = U.d;
  T es = T(0);
  es += U.mx * g0;
  es += U.my * g1;
  es += U.mz * g2;
  const T esV = es * V;
  acc.d = acc.d + T(0);
  acc.mx = acc.mx + (-(rho * g0 * V));
  acc.my = acc.my + (-(rho * g1 * V));
  acc.mz = acc.mz + (-(rho * g2 * V));
  acc.e = acc.e + (-esV);
}

// CTX×CTY > 0: every tile of the launch has that shape (one lane per cell; the
// index arithmetic folds at compile time); 0: the tile shape is read per tile
// and a workgroup of BNT lanes covers it.
template <typename T, int RANK, bool STAGE2, bool FLOOR64, bool GRAV = false, int CTX = 0, int CTY = 0>
__global__ void __launch_bounds__(CTX ? CTX * CTY : BNT, sizeof(T) == 4 ? VH_BLOCK_MIN_WAVES_F32 : VH_BLOCK_MIN_WAVES) block_stage_kernel(const T* __restrict__ Us, const T* U0, T* Uout,
                                                                       const double* __restrict__ dtp,
                                                                       double* __restrict__ partials, BlocksDev B) {
  constexpr int NT = CTX ? CTX * CTY : BNT;   // lanes per workgroup
  constexpr int NW = NT / 64;                 // waves per workgroup
  static_assert(NT % 64 == 0 && NT <= BNT, "tile shape");
  using LDS = BlockLDS<T, CTX, CTY>;
  constexpr int LWP = LDS::W, LFP = LDS::F;   // field pitches of w and f
  __shared__ LDS S;
  const T g = T(B.gamma);
  const T igm1 = T(B.igm1);
  const int tid = threadIdx.x;
  const long long n = B.ncell;
  const double dt = *dtp;
  const T* __restrict__ LF = (const T*)B.fluxbuf;
  double cfl_max = 0.0;

  for (int t = blockIdx.x; t < B.ntiles; t += gridDim.x) {
    CflMax<T, RANK> cfl;   // this tile's CFL numerators (one block: one set of widths)
    const TileDev tl = B.tiles[t];
    const BlockDev* __restrict__ bk = B.blocks + tl.block;
    BlockGeo bg;
    bg.nx = bk->n[0];
    bg.ny = bk->n[1];
    bg.nz = bk->n[2];
    bg.off = bk->off;
    const int nx = bg.nx, ny = bg.ny, nz = bg.nz;
    const int bxl = bk->bc[0], bxh = bk->bc[1], byl = bk->bc[2], byh = bk->bc[3], bzl = bk->bc[4], bzh = bk->bc[5];
    const int TX = CTX ? CTX : tl.tx, TY = CTX ? CTY : tl.ty;
    const int WX = TX + 4;
    const bool lane = CTX ? true : tid < TX * TY;
    const int tx = lane ? tid % TX : 0, ty = lane ? tid / TX : 0;
    const int x0 = tl.x0, y0 = tl.y0, kz0 = tl.z0, kz1 = tl.z1;
    const T invV = T(dt) / (T(1.0) * T(bk->V));
    const T a0 = T(bk->area[0]), a1 = T(bk->area[1]), a2 = T(bk->area[2]);
    const int gx = x0 + tx, gy = y0 + ty;
    const bool active = lane && gx < nx && gy < ny;
    const int sxc = gx < nx ? gx : nx - 1;
    const int syc = gy < ny ? gy : ny - 1;

    // ---- z pipeline prologue ----
    Prim<T> wz0, wz1, wz2, sz0;
    Cons<T> fzlo;
    if constexpr (RANK == 3) {
      auto zcol = [&](int kk) -> Prim<T> {
        int gh;
        const long long cc = block_src<2>(bg, bk, B.repmap, bzl, sxc, syc, kk, gh);
        Prim<T> w = cons2prim<T, FLOOR64>(load_cons(Us, n, cc), g, B.gamma);
        if (gh) w = ghost(w, gh == 1 ? bzl : bzh, 3);
        return w;
      };
      const Prim<T> wm2 = zcol(kz0 - 2), wm1 = zcol(kz0 - 1);
      wz0 = zcol(kz0);
      wz1 = zcol(kz0 + 1);
      wz2 = zcol(kz0 + 2);
      const Prim<T> sm1 = slope(wm2, wm1, wz0);
      sz0 = slope(wm1, wz0, wz1);
      if (bzl != BC_PERIODIC && kz0 == 0) {
        const Prim<T> R = face_value(wz0, sz0, T(-0.5));
        fzlo = hllc(ghost(R, bzl, 3), R, g, igm1, 3);
      } else {
        fzlo = hllc(face_value(wm1, sm1, T(0.5)), face_value(wz0, sz0, T(-0.5)), g, igm1, 3);
      }
    }

    for (int k = kz0; k < kz1; ++k) {
      const long long c = bg.off + ((long long)k * ny + gy) * nx + gx;
      Cons<T> unext;
      int ghn = 0;
      if constexpr (RANK == 3) {
        if (k + 1 < kz1) unext = load_cons(Us, n, block_src<2>(bg, bk, B.repmap, bzl, sxc, syc, k + 3, ghn));
      }

      // ---- P0: the plane's primitive state with a cross-shaped halo ----
      Prim<T> wc;
      if (lane) {
        if constexpr (RANK == 3) {
          wc = wz0;
        } else {
          wc = cons2prim<T, FLOOR64>(load_cons(Us, n, bg.off + ((long long)k * ny + syc) * nx + sxc), g, B.gamma);
        }
        // Lanes past the block edge hold the neighbour position (periodic wrap,
        // representative cell or ghost), read only by slopes of edge cells.
        if (gx >= nx) {
          int gh;
          const long long sidx = block_src<0>(bg, bk, B.repmap, bxl, gx, syc, k, gh);
          const Prim<T> w = cons2prim<T, FLOOR64>(load_cons(Us, n, sidx), g, B.gamma);
          wc = gh ? ghost(w, gh == 1 ? bxl : bxh, 1) : w;
        } else if (gy >= ny) {
          int gh;
          const long long sidx = block_src<1>(bg, bk, B.repmap, byl, gx, gy, k, gh);
          const Prim<T> w = cons2prim<T, FLOOR64>(load_cons(Us, n, sidx), g, B.gamma);
          wc = gh ? ghost(w, gh == 1 ? byl : byh, 2) : w;
        }
        bl_putw(S, (ty + 2) * WX + tx + 2, wc);
      }
      // Halo, rim and edge work goes to lanes counted from a wave that moves
      // by one per plane, so every SIMD gets its share.
      const int lt = NW == 1 ? tid : (tid + 64 * ((k + t) % NW)) % NT;
      const int NHX = 4 * TY;
      const int NHY = (RANK > 1) ? 4 * TX : 0;
      for (int h = lt; h < NHX + NHY; h += NT) {
        int lx, ly;
        Prim<T> w;
        if (h < NHX) {
          const int r = h >> 2, q = h & 3;
          ly = r;
          lx = q < 2 ? q - 2 : TX + (q - 2);
          const int cy = y0 + ly < ny ? y0 + ly : ny - 1;
          int gh;
          const long long sidx = block_src<0>(bg, bk, B.repmap, bxl, x0 + lx, cy, k, gh);
          w = cons2prim<T, FLOOR64>(load_cons(Us, n, sidx), g, B.gamma);
          if (gh) w = ghost(w, gh == 1 ? bxl : bxh, 1);
        } else {
          const int hh = h - NHX;
          const int r = hh / TX, q = hh - r * TX;
          lx = q;
          ly = r < 2 ? r - 2 : TY + (r - 2);
          const int cx = x0 + lx < nx ? x0 + lx : nx - 1;
          int gh;
          const long long sidx = block_src<1>(bg, bk, B.repmap, byl, cx, y0 + ly, k, gh);
          w = cons2prim<T, FLOOR64>(load_cons(Us, n, sidx), g, B.gamma);
          if (gh) w = ghost(w, gh == 1 ? byl : byh, 2);
        }
        bl_putw(S, (ly + 2) * WX + lx + 2, w);
      }
      __syncthreads();

      // ---- PA: slopes once per cell; publish hi face states ----
      const T* wf = &S.w[0][0];
      T* ff = &S.f[0][0];
      const int NXS = TY * (TX + 2);
      Prim<T> xlo, ylo;
      if (lane) {
        const int o = (ty + 2) * WX + tx + 2;
        const Prim<T> sx = slope(lds_get5(wf, LWP, o - 1), wc, lds_get5(wf, LWP, o + 1));
        xlo = face_value(wc, sx, T(-0.5));
        lds_put5(ff, LFP, ty * (TX + 2) + tx + 1, face_value(wc, sx, T(0.5)));
        if constexpr (RANK > 1) {
          const Prim<T> sy = slope(lds_get5(wf, LWP, o - WX), wc, lds_get5(wf, LWP, o + WX));
          ylo = face_value(wc, sy, T(-0.5));
          lds_put5(ff, LFP, NXS + (ty + 1) * TX + tx, face_value(wc, sy, T(0.5)));
        }
      }
      // Halo rim through one code path: x hi state of cell -1 and lo state of
      // cell TX per row, y hi state of cell -1 and lo state of cell TY per column.
      const int NRX = 2 * TY;
      const int NRY = (RANK > 1) ? 2 * TX : 0;
      for (int h = lt; h < NRX + NRY; h += NT) {
        const bool isy = h >= NRX;
        const int q = isy ? h - NRX : h;
        const int line = q >> 1;
        const bool hi = q & 1;
        int ic, dst;
        if (!isy) {
          ic = (line + 2) * WX + (hi ? TX + 2 : 1);
          dst = line * (TX + 2) + (hi ? TX + 1 : 0);
        } else {
          ic = (hi ? TY + 2 : 1) * WX + line + 2;
          dst = NXS + (hi ? TY + 1 : 0) * TX + line;
        }
        const int d = isy ? WX : 1;
        const Prim<T> wm = lds_get5(wf, LWP, ic);
        const Prim<T> sl = slope(lds_get5(wf, LWP, ic - d), wm, lds_get5(wf, LWP, ic + d));
        lds_put5(ff, LFP, dst, face_value(wm, sl, hi ? T(-0.5) : T(0.5)));
      }
      __syncthreads();

      // ---- PB: face fluxes (each face of the tile once) ----
      if (lane && gy < ny && gx <= nx) {
        const int ix = ty * (TX + 2) + tx;
        lds_putf(ff, LFP, ix, axis_face<false>(gx, nx, bxl, bxh, lds_get5(ff, LFP, ix), xlo, g, igm1, 1));
      }
      if constexpr (RANK > 1) {
        if (lane && gx < nx && gy <= ny) {
          const int iy = NXS + ty * TX + tx;
          lds_putf(ff, LFP, iy, axis_face<false>(gy, ny, byl, byh, lds_get5(ff, LFP, iy), ylo, g, igm1, 2));
        }
      }
      // Tile hi-edge faces (x face TX of each row, y face TY of each column)
      // through one code path, a y face in the frame of an x face.
      const int NEX = TY;
      const int NEY = (RANK > 1) ? TX : 0;
      for (int h = lt; h < NEX + NEY; h += NT) {
        const bool isy = h >= NEX;
        const int line = isy ? h - NEX : h;
        const int gp = isy ? y0 + TY : x0 + TX;
        const int na = isy ? ny : nx;
        const bool inside = isy ? (x0 + line < nx) : (y0 + line < ny);
        if (inside && gp <= na) {
          const int sl = isy ? NXS + TY * TX + line : line * (TX + 2) + TX;
          const int sr = isy ? sl + TX : sl + 1;
          const Prim<T> L = swap_xy(lds_get5(ff, LFP, sl), isy);
          const Prim<T> R = swap_xy(lds_get5(ff, LFP, sr), isy);
          lds_putf(ff, LFP, sl,
                   swap_xy(axis_face<false>(gp, na, isy ? byl : bxl, isy ? byh : bxh, L, R, g, igm1, 1), isy));
        }
      }

      Cons<T> fzhi;
      Prim<T> sz1;
      if constexpr (RANK == 3) {
        sz1 = slope(wz0, wz1, wz2);
        const Prim<T> L = face_value(wz0, sz0, T(0.5));
        if (bzl != BC_PERIODIC && k == nz - 1)
          fzhi = hllc(L, ghost(L, bzh, 3), g, igm1, 3);
        else
          fzhi = hllc(L, face_value(wz1, sz1, T(-0.5)), g, igm1, 3);
      }
      __syncthreads();

      // ---- PC: gather in the face-walk order and update ----
      if (active) {
        const Cons<T> uown = load_cons(Us, n, c);
        Cons<T> u0own;
        if constexpr (STAGE2) u0own = load_cons_stream(U0, n, c);
        const int ix = ty * (TX + 2) + tx;
        const int iy = NXS + ty * TX + tx;
        auto flux = [&](int s) -> Cons<T> {
          if (s == 0) return lds_getf(ff, LFP, ix);
          if (s == 1) return lds_getf(ff, LFP, ix + 1);
          if (s == 2) return lds_getf(ff, LFP, iy);
          if (s == 3) return lds_getf(ff, LFP, iy + TX);
          if (s == 4) return fzlo;
          return fzhi;
        };
        const T ar[3] = {a0, a1, a2};
        // Lo-face kinds: 0 interior (partner earlier), 1 domain boundary, 2 periodic wrap.
        const int kind[3] = {gx > 0 ? 0 : (bxl == BC_PERIODIC ? 2 : 1), gy > 0 ? 0 : (byl == BC_PERIODIC ? 2 : 1),
                             k > 0 ? 0 : (bzl == BC_PERIODIC ? 2 : 1)};
        int lc = -1;
        if (bk->haslinks) {
          const int pos[3] = {gx, gy, k};
          const int nn[3] = {nx, ny, nz};
#pragma unroll
          for (int d = 0; d < RANK; ++d) {
            const int pi = plane_index(bg, d, gx, gy, k);
            const int o0 = bk->lcoff[2 * d], o1 = bk->lcoff[2 * d + 1];
            if (pos[d] == 0 && o0 >= 0) {
              const int v = B.lcmap[o0 + pi];
              if (v >= 0) lc = v;
            }
            if (pos[d] == nn[d] - 1 && o1 >= 0) {
              const int v = B.lcmap[o1 + pi];
              if (v >= 0) lc = v;
            }
          }
        }
        Cons<T> acc;
        acc.d = acc.mx = acc.my = acc.mz = acc.e = T(0);
        if (lc < 0 && kind[0] == 0 && (RANK < 2 || kind[1] == 0) && (RANK < 3 || kind[2] == 0)) {
          // interior cell: lo faces z, y, x, then hi faces x, y, z
#pragma unroll
          for (int d = RANK - 1; d >= 0; --d) acc_add(acc, flux(2 * d), -ar[d]);
#pragma unroll
          for (int d = 0; d < RANK; ++d) acc_add(acc, flux(2 * d + 1), ar[d]);
        } else if (lc < 0) {
          // the uniform walk order of rk_stage_kernel
#pragma unroll
          for (int d = RANK - 1; d >= 0; --d)
            if (kind[d] == 0) acc_add(acc, flux(2 * d), -ar[d]);
#pragma unroll
          for (int d = 0; d < RANK; ++d) {
            acc_add(acc, flux(2 * d + 1), ar[d]);
            if (kind[d] == 1) acc_add(acc, flux(2 * d), -ar[d]);
          }
#pragma unroll
          for (int d = 0; d < RANK; ++d)
            if (kind[d] == 2) acc_add(acc, flux(2 * d), -ar[d]);
        } else {
          const LinkCellDev R = B.linkcells[lc];
          const int npre = lc_npre(R), ngrp = lc_ngrp(R);
          // faces emitted by cells of preceding blocks
          for (int q = 0; q < npre; ++q) {
            const int sl = lc_grp(R, q);
            add_link_group(acc, B, LF, lc_start(R, sl), lc_count(R, sl), c);
          }
          // interior lo faces (partner earlier in the block), z, y, x
#pragma unroll
          for (int d = RANK - 1; d >= 0; --d)
            if (lc_type(R, 2 * d) == 0 && kind[d] == 0) acc_add(acc, flux(2 * d), -ar[d]);
          // per axis: the hi face, then a boundary lo face or an own coarse-side lo link
#pragma unroll
          for (int d = 0; d < RANK; ++d) {
            const int th = lc_type(R, 2 * d + 1), tlo = lc_type(R, 2 * d);
            if (th == 0) acc_add(acc, flux(2 * d + 1), ar[d]);
            else if (th == 1) add_link_group(acc, B, LF, lc_start(R, 2 * d + 1), 1, c);
            if (tlo == 0 && kind[d] == 1) acc_add(acc, flux(2 * d), -ar[d]);
            else if (tlo == 1) add_link_group(acc, B, LF, lc_start(R, 2 * d), 1, c);
          }
          // lo faces wrapped by a periodic boundary (partner later in the block)
#pragma unroll
          for (int d = 0; d < RANK; ++d)
            if (lc_type(R, 2 * d) == 0 && kind[d] == 2) acc_add(acc, flux(2 * d), -ar[d]);
          // faces emitted by cells of following blocks
          for (int q = npre; q < ngrp; ++q) {
            const int sl = lc_grp(R, q);
            add_link_group(acc, B, LF, lc_start(R, sl), lc_count(R, sl), c);
          }
        }

        if constexpr (GRAV) block_grav_add<T>(acc, uown, (const T*)B.gravg, n, c, T(bk->V));

        Cons<T> o;
        if constexpr (!STAGE2) {
          o.d = uown.d - invV * acc.d;
          o.mx = uown.mx - invV * acc.mx;
          o.my = uown.my - invV * acc.my;
          o.mz = uown.mz - invV * acc.mz;
          o.e = uown.e - invV * acc.e;
        } else {
          const T hh = T(0.5);
          o.d = hh * u0own.d + hh * (uown.d - invV * acc.d);
          o.mx = hh * u0own.mx + hh * (uown.mx - invV * acc.mx);
          o.my = hh * u0own.my + hh * (uown.my - invV * acc.my);
          o.mz = hh * u0own.mz + hh * (uown.mz - invV * acc.mz);
          o.e = hh * u0own.e + hh * (uown.e - invV * acc.e);
          cfl.add(cons2prim<T, FLOOR64>(o, g, B.gamma), g);
        }
        store_cons_stream(Uout, n, c, o);
      }

      if constexpr (RANK == 3) {
        fzlo = fzhi;
        sz0 = sz1;
        wz0 = wz1;
        wz1 = wz2;
        if (k + 1 < kz1) {
          Prim<T> w = cons2prim<T, FLOOR64>(unext, g, B.gamma);
          if (ghn) w = ghost(w, ghn == 1 ? bzl : bzh, 3);
          wz2 = w;
        }
      }
    }
    if constexpr (STAGE2) cfl_max = vmax(cfl_max, cfl.value(bk->w, GRAV, B.gravG));
  }

  if constexpr (STAGE2) {
    const double m = block_max<NT>(cfl_max, reinterpret_cast<double*>(&S.w[0][0]));
    if (tid == 0) partials[blockIdx.x] = m;
  }
}

// Per-workgroup CFL maximum (src/driver.jl:178-185) over the cells of a list of
// blocks; prefix[q] = cells of the first q listed blocks.
template <typename T, int RANK, bool FLOOR64>
__global__ void __launch_bounds__(256) block_cfl_kernel(const T* __restrict__ U, BlocksDev B,
                                                        const int* __restrict__ blist,
                                                        const long long* __restrict__ prefix, int nb,
                                                        double* __restrict__ partials) {
  __shared__ double red[16];
  const T g = T(B.gamma);
  const long long total = prefix[nb];
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const BlockDev& bk = B.blocks[blist[lo]];
    const Prim<T> W = cons2prim<T, FLOOR64>(load_cons(U, B.ncell, bk.off + (i - prefix[lo])), g, B.gamma);
    const T cs = vsqrt(g * W.p / W.r);
    double v = double(vabs(W.vx) + cs) / bk.w[0];
    if (RANK > 1) v = vmax(v, double(vabs(W.vy) + cs) / bk.w[1]);
    if (RANK > 2) v = vmax(v, double(vabs(W.vz) + cs) / bk.w[2]);
    // free-fall limiter √(4πGρ)/½ with self-gravity (src/gravity.jl:449-452)
    if (B.gravg) v = vmax(v, double(W.r) > 0.0 ? __builtin_sqrt(4.0 * 3.141592653589793 * B.gravG * double(W.r)) / 0.5 : 0.0);
    m = vmax(m, v);
  }
  m = block_max<256>(m, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
