// Fused SSP-RK2 stage kernels for a uniform ghost-free grid (gfx950).
//
// One kernel launch is one RK stage over the whole grid: PLM/minmod
// reconstruction, HLLC at every face, the flux-divergence accumulation and the
// stage update, all from a single read of the stencil field. The second stage
// also fuses the SSP combine and the FP64 CFL reduction of the new state.
//
// Work decomposition (2.5-D march): a workgroup owns a TX×TY column of cells and
// marches through a chunk of z planes. Per plane:
//   P0  stage the primitive state W of the plane with a two-cell halo in x and y
//       in LDS (cross-shaped footprint; corners are never read);
//   PA  each lane computes its cell's x and y slopes once, keeps the lo-side
//       face states in registers and publishes the hi-side ones to LDS; a few
//       lanes do the same for the one-cell halo rim;
//   PB  each lane evaluates the HLLC flux of its cell's x-lo and y-lo faces and
//       writes it over the face state it just consumed (every published face
//       state is read by exactly one face); the tile's hi-edge faces are spread
//       over the lanes; meanwhile each lane advances its own z pipeline
//       (W(k+1), W(k+2), the z slope and the z-lo flux live in registers);
//   PC  each lane gathers its six face fluxes in the reference order and
//       writes the updated state.
// Every face flux and every slope of the tile is evaluated once per plane.
//
// Accumulation order. The reference walks faces cell by cell, column-major,
// axis by axis (lib/RefMesh/src/RefMesh.jl:187-197) and scatters ±F·area into
// the two cells (src/driver.jl:236-281). For one cell that order is:
//   1. lo faces whose partner precedes the cell (non-wrapped interior), z, y, x;
//   2. per axis x, y, z: its own hi face, then the lo face if it is a domain
//      boundary (emitted right after the hi face of the same axis);
//   3. lo faces wrapped by a periodic boundary (partner visited later), x, y, z.
// The gather adds the six fluxes in exactly that order, so in the exact profile
// the floating-point sum matches the reference scatter bit for bit.
#pragma once

#include "hydro.h"

#ifndef VH_MIN_WAVES
#define VH_MIN_WAVES 3
#endif
// Float32 state: fewer registers per value, so more waves can be resident.
#ifndef VH_MIN_WAVES_F32
#define VH_MIN_WAVES_F32 4
#endif
// 1: issue the loads of the cell's own state at the start of the plane (more
// latency hidden, more registers live across the plane); 0: load it in the
// update phase.
#ifndef VH_EARLY_OWN
#define VH_EARLY_OWN 0
#endif

namespace vh {
namespace VH_PROFILE_NS {

// Per-cell CFL term of compute_dt (src/driver.jl:178-185), Tr = double.
template <typename T, int RANK>
__device__ __forceinline__ double cfl_term(const Prim<T>& W, T g, const GridDev& G) {
  const T c = vsqrt(g * W.p / W.r);
  double m = double(vabs(W.vx) + c) / G.w[0];
  if (RANK > 1) m = vmax(m, double(vabs(W.vy) + c) / G.w[1]);
  if (RANK > 2) m = vmax(m, double(vabs(W.vz) + c) / G.w[2]);
  return m;
}

template <typename T, int TX, int TY>
struct PlaneLDS {
  T w[5][TY + 4][TX + 4];   // primitive state, halo 2 in x and y
  T xh[5][TY][TX + 1];      // x: hi face state of cell s-1 → then flux of x face s
  T yh[5][TY + 1][TX];      // y: hi face state of cell s-1 → then flux of y face s
  T xl[5][TY];              // x: lo face state of halo cell TX
  T yl[5][TX];              // y: lo face state of halo cell TY
  double red[16];
};

#define VH_PUT5(arr, idx, v) \
  do {                       \
    arr[0] idx = (v).r;      \
    arr[1] idx = (v).vx;     \
    arr[2] idx = (v).vy;     \
    arr[3] idx = (v).vz;     \
    arr[4] idx = (v).p;      \
  } while (0)
#define VH_GET5(v, arr, idx) \
  do {                       \
    (v).r = arr[0] idx;      \
    (v).vx = arr[1] idx;     \
    (v).vy = arr[2] idx;     \
    (v).vz = arr[3] idx;     \
    (v).p = arr[4] idx;      \
  } while (0)
#define VH_PUTF(arr, idx, f) \
  do {                       \
    arr[0] idx = (f).d;      \
    arr[1] idx = (f).mx;     \
    arr[2] idx = (f).my;     \
    arr[3] idx = (f).mz;     \
    arr[4] idx = (f).e;      \
  } while (0)
#define VH_GETF(f, arr, idx) \
  do {                       \
    (f).d = arr[0] idx;      \
    (f).mx = arr[1] idx;     \
    (f).my = arr[2] idx;     \
    (f).mz = arr[3] idx;     \
    (f).e = arr[4] idx;      \
  } while (0)

template <typename T, int TX, int TY>
__device__ __forceinline__ Prim<T> lds_w(const PlaneLDS<T, TX, TY>& S, int sy, int sx) {
  Prim<T> w;
  VH_GET5(w, S.w, [sy][sx]);
  return w;
}

// State of a column cell at plane index kk (any integer) with the z boundary
// rule applied: periodic wrap, or the ghost of the nearest boundary cell.
template <typename T, bool FLOOR64>
__device__ __forceinline__ Prim<T> column_w(const Cons<T>& u, const GridDev& G, int kk, T g) {
  int gh;
  (void)resolve(kk, G.n[2], G.bc[4], gh);
  Prim<T> w = cons2prim<T, FLOOR64>(u, g, G.gamma);
  if (gh) w = ghost(w, G.bc[4 + gh - 1], 3);
  return w;
}
__device__ __forceinline__ long long column_src(const GridDev& G, long long col, int kk) {
  int gh;
  const int src = resolve(kk, G.n[2], G.bc[4], gh);
  return col + (long long)src * ((long long)G.n[0] * G.n[1]);
}

// Flux of the face between cells gpos-1 and gpos along an axis with n cells,
// given the hi state of gpos-1 (L) and the lo state of gpos (R); applies the
// domain-boundary rule of _flux_face! (src/driver.jl:255-281).
template <typename T>
__device__ __forceinline__ Cons<T> axis_face(int gpos, int n, int bclo, int bchi, Prim<T> L, Prim<T> R, T g,
                                             T igm1, int dim) {
  if (bclo != BC_PERIODIC) {
    if (gpos == 0) L = ghost(R, bclo, dim);
    else if (gpos == n) R = ghost(L, bchi, dim);
  }
  return hllc(L, R, g, igm1, dim);
}

// Boundary-flux register (src/reflux.jl:105-132): b[slot] += (s·F)·area for a
// domain-boundary face, s = ½dt per RK stage. Slots: planes ordered (axis 0 lo,
// axis 0 hi, axis 1 lo, ...), each column-major over the other axes; idx is the
// face's index in its plane. Each boundary face is evaluated by one lane per
// stage, so the read-modify-write needs no atomics.
template <typename T>
__device__ __forceinline__ void bflux_add(const GridDev& G, int axis, int side, long long idx, const Cons<T>& F,
                                          T s, T a) {
  long long base = 0;
  for (int e = 0; e < axis; ++e) base += 2 * (G.ncell / G.n[e]);
  if (side) base += G.ncell / G.n[axis];
  T* b = (T*)G.bflux + 5 * (base + idx);
  b[0] = b[0] + (s * F.d) * a;
  b[1] = b[1] + (s * F.mx) * a;
  b[2] = b[2] + (s * F.my) * a;
  b[3] = b[3] + (s * F.mz) * a;
  b[4] = b[4] + (s * F.e) * a;
}

// One RK stage. STAGE2 = false: Uout = Us − (dt/V)·acc(Us).
// STAGE2 = true: Uout = ½·U0 + ½·(Us − (dt/V)·acc(Us)) written over U0 (point
// reads only), plus the per-workgroup max of the CFL term of the new state.
template <typename T, int RANK, int TX, int TY, bool STAGE2, bool FLOOR64>
__global__ void __launch_bounds__(TX* TY, sizeof(T) == 4 ? VH_MIN_WAVES_F32 : VH_MIN_WAVES) rk_stage_kernel(const T* __restrict__ Us, const T* U0,
                                                          T* Uout, const double* __restrict__ dtp,
                                                          double* __restrict__ partials, GridDev G) {
  constexpr int NT = TX * TY;
  __shared__ PlaneLDS<T, TX, TY> S;
  const T g = T(G.gamma);
  const T igm1 = T(G.igm1);
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;

  // Workgroup → (tile, z chunk). Consecutive dispatch slots go round robin over
  // the 8 XCDs; the remap gives each XCD a contiguous tile range so tiles that
  // share halos share an L2.
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
  // The lane's own column (clamped or wrapped so every lane reads valid memory).
  int ghx, ghy;
  const int sxc = resolve(gx, nx, G.bc[0], ghx);
  const int syc = resolve(gy, ny, G.bc[2], ghy);
  const long long col = (long long)syc * nx + sxc;

  // ---- z pipeline prologue ----
  Prim<T> wz0, wz1, wz2, sz0;
  Cons<T> fzlo;
  if constexpr (RANK == 3) {
    const Cons<T> um2 = load_cons(Us, n, column_src(G, col, kz0 - 2));
    const Cons<T> um1 = load_cons(Us, n, column_src(G, col, kz0 - 1));
    const Cons<T> u0 = load_cons(Us, n, column_src(G, col, kz0));
    const Cons<T> u1 = load_cons(Us, n, column_src(G, col, kz0 + 1));
    const Cons<T> u2 = load_cons(Us, n, column_src(G, col, kz0 + 2));
    const Prim<T> wm2 = column_w<T, FLOOR64>(um2, G, kz0 - 2, g);
    const Prim<T> wm1 = column_w<T, FLOOR64>(um1, G, kz0 - 1, g);
    wz0 = column_w<T, FLOOR64>(u0, G, kz0, g);
    wz1 = column_w<T, FLOOR64>(u1, G, kz0 + 1, g);
    wz2 = column_w<T, FLOOR64>(u2, G, kz0 + 2, g);
    // The slope of the ghost plane kz0-1 is never used when kz0-1 is outside a
    // non-periodic domain: that face is a boundary face (lo state = ghost).
    const Prim<T> sm1 = slope(wm2, wm1, wz0);
    sz0 = slope(wm1, wz0, wz1);
    const int bz = G.bc[4];
    if (bz != BC_PERIODIC && kz0 == 0) {
      const Prim<T> R = face_value(wz0, sz0, T(-0.5));
      fzlo = hllc(ghost(R, bz, 3), R, g, igm1, 3);
      if (G.bflux && active) bflux_add(G, 2, 0, (long long)gy * nx + gx, fzlo, T(0.5 * dt), a2);
    } else {
      fzlo = hllc(face_value(wm1, sm1, T(0.5)), face_value(wz0, sz0, T(-0.5)), g, igm1, 3);
    }
  }

  double cfl_max = 0.0;

  for (int k = kz0; k < kz1; ++k) {
    const long long c = (long long)k * plane + (long long)gy * nx + gx;
    // Early loads: the next z plane of the column and this cell's own state.
    Cons<T> unext, uown, u0own;
    if constexpr (RANK == 3) {
      if (k + 1 < kz1) unext = load_cons(Us, n, column_src(G, col, k + 3));
    }
#if VH_EARLY_OWN
    if (active) {
      uown = load_cons(Us, n, c);
      if constexpr (STAGE2) u0own = load_cons(U0, n, c);
    }
#endif

    // ---- P0: stage the plane's primitive state in LDS ----
    Prim<T> wc;
    if constexpr (RANK == 3) {
      wc = wz0;
    } else {
      wc = cons2prim<T, FLOOR64>(load_cons(Us, n, col), g, G.gamma);
    }
    {
      Prim<T> w = wc;
      if (ghx) w = ghost(w, G.bc[ghx - 1], 1);
      if (ghy) w = ghost(w, G.bc[2 + ghy - 1], 2);
      VH_PUT5(S.w, [ty + 2][tx + 2], w);
      wc = w;
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
      VH_PUT5(S.w, [ly + 2][lx + 2], w);
    }
    __syncthreads();

    // ---- PA: slopes once per cell; publish hi face states ----
    Prim<T> xlo, ylo;
    {
      const Prim<T> sx = slope(lds_w(S, ty + 2, tx + 1), wc, lds_w(S, ty + 2, tx + 3));
      xlo = face_value(wc, sx, T(-0.5));
      const Prim<T> xhi = face_value(wc, sx, T(0.5));
      VH_PUT5(S.xh, [ty][tx + 1], xhi);
      if constexpr (RANK > 1) {
        const Prim<T> sy = slope(lds_w(S, ty + 1, tx + 2), wc, lds_w(S, ty + 3, tx + 2));
        ylo = face_value(wc, sy, T(-0.5));
        const Prim<T> yhi = face_value(wc, sy, T(0.5));
        VH_PUT5(S.yh, [ty + 1][tx], yhi);
      }
    }
    // Halo rim: x hi state of cell -1 and lo state of cell TX per row; y hi
    // state of cell -1 and lo state of cell TY per column.
    constexpr int NRX = 2 * TY;
    constexpr int NRY = (RANK > 1) ? 2 * TX : 0;
    for (int h = tid; h < NRX + NRY; h += NT) {
      if (h < NRX) {
        const int r = h >> 1;
        if ((h & 1) == 0) {   // cell -1: slots 0, 1, 2
          const Prim<T> wm = lds_w(S, r + 2, 1);
          const Prim<T> s = slope(lds_w(S, r + 2, 0), wm, lds_w(S, r + 2, 2));
          const Prim<T> v = face_value(wm, s, T(0.5));
          VH_PUT5(S.xh, [r][0], v);
        } else {              // cell TX: slots TX+1, TX+2, TX+3
          const Prim<T> wp = lds_w(S, r + 2, TX + 2);
          const Prim<T> s = slope(lds_w(S, r + 2, TX + 1), wp, lds_w(S, r + 2, TX + 3));
          const Prim<T> v = face_value(wp, s, T(-0.5));
          VH_PUT5(S.xl, [r], v);
        }
      } else {
        const int hh = h - NRX;
        const int cc = hh >> 1;
        if ((hh & 1) == 0) {
          const Prim<T> wm = lds_w(S, 1, cc + 2);
          const Prim<T> s = slope(lds_w(S, 0, cc + 2), wm, lds_w(S, 2, cc + 2));
          const Prim<T> v = face_value(wm, s, T(0.5));
          VH_PUT5(S.yh, [0][cc], v);
        } else {
          const Prim<T> wp = lds_w(S, TY + 2, cc + 2);
          const Prim<T> s = slope(lds_w(S, TY + 1, cc + 2), wp, lds_w(S, TY + 3, cc + 2));
          const Prim<T> v = face_value(wp, s, T(-0.5));
          VH_PUT5(S.yl, [cc], v);
        }
      }
    }
    __syncthreads();

    // ---- PB: face fluxes (each face once), written over the consumed state ----
    const int bx = G.bc[0], by = G.bc[2];
    if (gy < ny && gx <= nx) {
      Prim<T> L;
      VH_GET5(L, S.xh, [ty][tx]);
      const Cons<T> F = axis_face(gx, nx, bx, G.bc[1], L, xlo, g, igm1, 1);
      VH_PUTF(S.xh, [ty][tx], F);
      if (G.bflux && bx != BC_PERIODIC && (gx == 0 || gx == nx))
        bflux_add(G, 0, gx == nx, (long long)k * ny + gy, F, T(0.5 * dt), a0);
    }
    if constexpr (RANK > 1) {
      if (gx < nx && gy <= ny) {
        Prim<T> L;
        VH_GET5(L, S.yh, [ty][tx]);
        const Cons<T> F = axis_face(gy, ny, by, G.bc[3], L, ylo, g, igm1, 2);
        VH_PUTF(S.yh, [ty][tx], F);
        if (G.bflux && by != BC_PERIODIC && (gy == 0 || gy == ny))
          bflux_add(G, 1, gy == ny, (long long)k * nx + gx, F, T(0.5 * dt), a1);
      }
    }
    // Tile hi-edge faces: x face TX of each row, y face TY of each column.
    constexpr int NEX = TY;
    constexpr int NEY = (RANK > 1) ? TX : 0;
    for (int h = tid; h < NEX + NEY; h += NT) {
      if (h < NEX) {
        const int r = h;
        const int gp = x0 + TX;
        if (y0 + r < ny && gp <= nx) {
          Prim<T> L, R;
          VH_GET5(L, S.xh, [r][TX]);
          VH_GET5(R, S.xl, [r]);
          const Cons<T> F = axis_face(gp, nx, bx, G.bc[1], L, R, g, igm1, 1);
          VH_PUTF(S.xh, [r][TX], F);
          if (G.bflux && bx != BC_PERIODIC && gp == nx)
            bflux_add(G, 0, 1, (long long)k * ny + y0 + r, F, T(0.5 * dt), a0);
        }
      } else {
        const int cc = h - NEX;
        const int gp = y0 + TY;
        if (x0 + cc < nx && gp <= ny) {
          Prim<T> L, R;
          VH_GET5(L, S.yh, [TY][cc]);
          VH_GET5(R, S.yl, [cc]);
          const Cons<T> F = axis_face(gp, ny, by, G.bc[3], L, R, g, igm1, 2);
          VH_PUTF(S.yh, [TY][cc], F);
          if (G.bflux && by != BC_PERIODIC && gp == ny)
            bflux_add(G, 1, 1, (long long)k * nx + x0 + cc, F, T(0.5 * dt), a1);
        }
      }
    }

    // z face at k+½ for the lane's column.
    Cons<T> fzhi;
    Prim<T> sz1;
    if constexpr (RANK == 3) {
      sz1 = slope(wz0, wz1, wz2);
      const int bz = G.bc[4];
      const Prim<T> L = face_value(wz0, sz0, T(0.5));
      if (bz != BC_PERIODIC && k == nz - 1) {
        fzhi = hllc(L, ghost(L, G.bc[5], 3), g, igm1, 3);
        if (G.bflux && active) bflux_add(G, 2, 1, (long long)gy * nx + gx, fzhi, T(0.5 * dt), a2);
      } else
        fzhi = hllc(L, face_value(wz1, sz1, T(-0.5)), g, igm1, 3);
    }
    __syncthreads();

    // ---- PC: gather the six face fluxes in the reference order and update ----
    if (active) {
#if !VH_EARLY_OWN
      uown = load_cons(Us, n, c);
      if constexpr (STAGE2) u0own = load_cons(U0, n, c);
#endif
      Cons<T> fxlo, fxhi, fylo, fyhi;
      VH_GETF(fxlo, S.xh, [ty][tx]);
      VH_GETF(fxhi, S.xh, [ty][tx + 1]);
      if constexpr (RANK > 1) {
        VH_GETF(fylo, S.yh, [ty][tx]);
        VH_GETF(fyhi, S.yh, [ty + 1][tx]);
      }
      // Lo-face kinds: 0 interior (partner earlier), 1 domain boundary, 2 periodic wrap.
      const int kx = gx > 0 ? 0 : (bx == BC_PERIODIC ? 2 : 1);
      const int ky = gy > 0 ? 0 : (by == BC_PERIODIC ? 2 : 1);
      const int kz = k > 0 ? 0 : (G.bc[4] == BC_PERIODIC ? 2 : 1);
      Cons<T> acc;
      acc.d = acc.mx = acc.my = acc.mz = acc.e = T(0);
      if (kx == 0 && (RANK < 2 || ky == 0) && (RANK < 3 || kz == 0)) {
        // interior cell (every wave away from the lo domain faces): the same
        // order as below without per-lane selects
        if constexpr (RANK > 2) acc_add(acc, fzlo, -a2);
        if constexpr (RANK > 1) acc_add(acc, fylo, -a1);
        acc_add(acc, fxlo, -a0);
        acc_add(acc, fxhi, a0);
        if constexpr (RANK > 1) acc_add(acc, fyhi, a1);
        if constexpr (RANK > 2) acc_add(acc, fzhi, a2);
      } else {
        if constexpr (RANK > 2) {
          if (kz == 0) acc_add(acc, fzlo, -a2);
        }
        if constexpr (RANK > 1) {
          if (ky == 0) acc_add(acc, fylo, -a1);
        }
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
        if constexpr (RANK > 1) {
          if (ky == 2) acc_add(acc, fylo, -a1);
        }
        if constexpr (RANK > 2) {
          if (kz == 2) acc_add(acc, fzlo, -a2);
        }
      }

      Cons<T> o;
      if constexpr (!STAGE2) {
        // _euler_apply!: dst = base − invV·acc (src/driver.jl:297-307).
        o.d = uown.d - invV * acc.d;
        o.mx = uown.mx - invV * acc.mx;
        o.my = uown.my - invV * acc.my;
        o.mz = uown.mz - invV * acc.mz;
        o.e = uown.e - invV * acc.e;
      } else {
        // _rk2_combine!: ½·u0 + ½·(u1 − invV·acc) (src/driver.jl:310-321).
        const T h = T(0.5);
        o.d = h * u0own.d + h * (uown.d - invV * acc.d);
        o.mx = h * u0own.mx + h * (uown.mx - invV * acc.mx);
        o.my = h * u0own.my + h * (uown.my - invV * acc.my);
        o.mz = h * u0own.mz + h * (uown.mz - invV * acc.mz);
        o.e = h * u0own.e + h * (uown.e - invV * acc.e);
        cfl_max = vmax(cfl_max, cfl_term<T, RANK>(cons2prim<T, FLOOR64>(o, g, G.gamma), g, G));
      }
      store_cons(Uout, n, c, o);
    }

    if constexpr (RANK == 3) {
      fzlo = fzhi;
      sz0 = sz1;
      wz0 = wz1;
      wz1 = wz2;
      if (k + 1 < kz1) wz2 = column_w<T, FLOOR64>(unext, G, k + 3, g);
    }
  }

  if constexpr (STAGE2) {
    const double m = block_max<NT>(cfl_max, S.red);
    if (tid == 0) partials[blockIdx.x] = m;
  }
}

// Per-workgroup max of the CFL term over a grid-stride range of cells.
template <typename T, int RANK, bool FLOOR64>
__global__ void __launch_bounds__(256) cfl_kernel(const T* __restrict__ U, GridDev G,
                                                  double* __restrict__ partials) {
  __shared__ double red[16];
  const T g = T(G.gamma);
  double m = 0.0;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < G.ncell; c += (long long)gridDim.x * 256)
    m = vmax(m, cfl_term<T, RANK>(cons2prim<T, FLOOR64>(load_cons(U, G.ncell, c), g, G.gamma), g, G));
  m = block_max<256>(m, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
