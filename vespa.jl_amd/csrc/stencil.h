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

#include "gravity.h"
#include "hydro.h"

#ifndef VH_MIN_WAVES
#define VH_MIN_WAVES 3
#endif
// Float32 state: fewer registers per value, so more waves can be resident.
#ifndef VH_MIN_WAVES_F32
#define VH_MIN_WAVES_F32 4
#endif
// 1: in stage 1 of the periodic Float64 form, issue the loads of the cell's own
// state at the start of the plane (more latency hidden, 10 more registers live
// across the plane, which the boundary forms and stage 2 need elsewhere; the
// Float32 form measured 3% slower with it); 0: load it in the update phase.
#ifndef VH_EARLY_OWN
#define VH_EARLY_OWN 1
#endif
// 1: load the halo cells of the next plane while the current one is processed
// (tiles whose halo fits one cell per lane); 0: load them when they are staged.
#ifndef VH_HALO_PREFETCH
#define VH_HALO_PREFETCH 0
#endif

namespace vh {
namespace VH_PROFILE_NS {

// Running maximum of the CFL term of compute_dt (src/driver.jl:178-185):
// max over cells and axes of Tr(|v_d| + c)/w_d with Tr = double. The per-axis
// numerators are reduced first and divided by the width once: for w > 0 a
// correctly rounded x ↦ x/w is monotone, so max_i(a_i/w) = (max_i a_i)/w
// exactly, and the result equals the per-cell form bit for bit. The sound
// speed is the profile's (exact: sqrt((γp)/ρ); fast: the refined rsqrt form).
// With self-gravity the free-fall limiter √(4πGρ)/η, η = ½ (src/gravity.jl:449-452),
// joins the maximum; it is monotone in ρ, so it is evaluated once at the largest ρ.
template <typename T, int RANK>
struct CflMax {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, rmax = 0.0;
  __device__ __forceinline__ void add(const Prim<T>& W, T g) {
    const T c = sound(g, W.p, W.r);
    a0 = vmax(a0, double(vabs(W.vx) + c));
    if (RANK > 1) a1 = vmax(a1, double(vabs(W.vy) + c));
    if (RANK > 2) a2 = vmax(a2, double(vabs(W.vz) + c));
    rmax = vmax(rmax, double(W.r));
  }
  __device__ __forceinline__ double value(const double* w, bool grav = false, double Gc = 0.0) const {
    double m = a0 / w[0];
    if (RANK > 1) m = vmax(m, a1 / w[1]);
    if (RANK > 2) m = vmax(m, a2 / w[2]);
    if (grav) m = vmax(m, rmax > 0.0 ? __builtin_sqrt(4.0 * 3.141592653589793 * Gc * rmax) / 0.5 : 0.0);
    return m;
  }
};

// Cell index of position i along an axis of n cells. PER: the axis is periodic
// (wrap, no ghost) and i is a stencil position, i >= -n; positions from 2n on
// lie past every cell a face of the grid reads (partial tiles of small grids)
// and are clamped to a valid cell. Else the generic boundary rule of resolve().
template <bool PER>
__device__ __forceinline__ int resolve_axis(int i, int n, int bclo, int& gh) {
  if constexpr (PER) {
    gh = 0;
    i = i < 0 ? i + n : (i >= n ? i - n : i);
    return i < n ? i : n - 1;
  } else {
    return resolve(i, n, bclo, gh);
  }
}

// Face slots share one array per field so that the x and y halo-rim work runs
// through a single code path:
//   x slot (r, s), s = 0..TX:  hi face state of cell s-1 of row r, then the flux
//                              of x face s; s = TX+1: lo face state of halo cell TX;
//   y slot (r, c), r = 0..TY:  hi face state of cell r-1 of column c, then the
//                              flux of y face r; r = TY+1: lo face state of halo cell TY.
// The workgroup reduction at the end of stage 2 reuses the start of w (it runs
// after a barrier that follows the last read of w), which keeps the 32×8
// Float64 tile at 40 KiB: four workgroups per compute unit.
template <typename T, int TX, int TY, int RANK>
struct alignas(16) PlaneLDS {
  static constexpr int HY = RANK > 1 ? 2 : 0;    // halo rows in y
  static constexpr int WX = TX + 4;              // row pitch of w
  static constexpr int WS = (TY + 2 * HY) * WX;  // field pitch of w
  static constexpr int NXS = TY * (TX + 2);
  static constexpr int NYS = RANK > 1 ? (TY + 2) * TX : 0;
  static constexpr int FS = NXS + NYS;           // field pitch of f
  __device__ static constexpr int xs(int r, int sl) { return r * (TX + 2) + sl; }
  __device__ static constexpr int ys(int r, int c) { return NXS + r * TX + c; }
  T w[5][TY + 2 * HY][TX + 4];   // primitive state, halo 2 in x (and y)
  T f[5][FS];                    // face slots
};

template <typename T>
__device__ __forceinline__ Prim<T> lds_get5(const T* a, int pitch, int i) {
  Prim<T> v;
  v.r = a[i];
  v.vx = a[pitch + i];
  v.vy = a[2 * pitch + i];
  v.vz = a[3 * pitch + i];
  v.p = a[4 * pitch + i];
  return v;
}
template <typename T>
__device__ __forceinline__ void lds_put5(T* a, int pitch, int i, const Prim<T>& v) {
  a[i] = v.r;
  a[pitch + i] = v.vx;
  a[2 * pitch + i] = v.vy;
  a[3 * pitch + i] = v.vz;
  a[4 * pitch + i] = v.p;
}
template <typename T>
__device__ __forceinline__ Cons<T> lds_getf(const T* a, int pitch, int i) {
  Cons<T> v;
  v.d = a[i];
  v.mx = a[pitch + i];
  v.my = a[2 * pitch + i];
  v.mz = a[3 * pitch + i];
  v.e = a[4 * pitch + i];
  return v;
}
template <typename T>
__device__ __forceinline__ void lds_putf(T* a, int pitch, int i, const Cons<T>& v) {
  a[i] = v.d;
  a[pitch + i] = v.mx;
  a[2 * pitch + i] = v.my;
  a[3 * pitch + i] = v.mz;
  a[4 * pitch + i] = v.e;
}
// Exchange the x and y components (a y face evaluated in the frame of an x face).
template <typename T>
__device__ __forceinline__ Prim<T> swap_xy(const Prim<T>& w, bool on) {
  Prim<T> o = w;
  o.vx = on ? w.vy : w.vx;
  o.vy = on ? w.vx : w.vy;
  return o;
}
template <typename T>
__device__ __forceinline__ Cons<T> swap_xy(const Cons<T>& f, bool on) {
  Cons<T> o = f;
  o.mx = on ? f.my : f.mx;
  o.my = on ? f.mx : f.my;
  return o;
}

#define VH_PUT5(arr, idx, v) \
  do {                       \
    arr[0] idx = (v).r;      \
    arr[1] idx = (v).vx;     \
    arr[2] idx = (v).vy;     \
    arr[3] idx = (v).vz;     \
    arr[4] idx = (v).p;      \
  } while (0)

// ParentGhost state (src/driver.jl:127-134) of the face on `side` of `axis`
// (0-based) of the boundary cell with face-plane index idx: 5 primitives per
// boundary-face slot in the boundary-flux register's slot order.
template <typename T>
__device__ __forceinline__ Prim<T> parent_ghost(const GridDev& G, int axis, int side, long long idx) {
  const T* q = (const T*)G.pg + 5 * (G.bbase[2 * axis + side] + idx);
  Prim<T> o;
  o.r = q[0];
  o.vx = q[1];
  o.vy = q[2];
  o.vz = q[3];
  o.p = q[4];
  return o;
}

// _boundary_ghost (src/driver.jl:127-134) at a domain-boundary side: the
// ParentGhost state (PG: the kernel form that reads them), or ghost_state of w
// (the cell's primitive for a slope, its face value for a face).
template <bool PG, typename T>
__device__ __forceinline__ Prim<T> bghost(const Prim<T>& w, const GridDev& G, int axis, int side, long long idx) {
  const int bc = G.bc[2 * axis + side];
  if constexpr (PG) {
    if (bc == BC_PARENT) return parent_ghost<T>(G, axis, side, idx);
  }
  return ghost(w, bc, axis + 1);
}

// State of a column cell at plane index kk (any integer) with the z boundary
// rule applied: periodic wrap, or the ghost of the nearest boundary cell (zidx:
// the column's index in the z boundary planes).
template <typename T, bool FLOOR64, bool PER, bool PG>
__device__ __forceinline__ Prim<T> column_w(const Cons<T>& u, const GridDev& G, int kk, T g, long long zidx) {
  Prim<T> w = cons2prim<T, FLOOR64>(u, g, G.gamma);
  if constexpr (!PER) {
    int gh;
    (void)resolve(kk, G.n[2], G.bc[4], gh);
    if (gh) w = bghost<PG>(w, G, 2, gh - 1, zidx);
  }
  return w;
}
template <bool PER>
__device__ __forceinline__ long long column_src(const GridDev& G, long long col, int kk) {
  int gh;
  const int src = resolve_axis<PER>(kk, G.n[2], G.bc[4], gh);
  return col + (long long)src * ((long long)G.n[0] * G.n[1]);
}

// Flux of the face between cells gpos-1 and gpos along an axis with n cells,
// given the hi state of gpos-1 (L) and the lo state of gpos (R); applies the
// domain-boundary rule of _flux_face! (src/driver.jl:255-281).
template <bool PER, typename T>
__device__ __forceinline__ Cons<T> axis_face(int gpos, int n, int bclo, int bchi, Prim<T> L, Prim<T> R, T g,
                                             T igm1, int dim) {
  if (!PER && bclo != BC_PERIODIC) {
    if (gpos == 0) L = ghost(R, bclo, dim);
    else if (gpos == n) R = ghost(L, bchi, dim);
  }
  return hllc(L, R, g, igm1, dim);
}

// The same on the uniform grid, where a side may be ParentGhost: the outer
// state is then the caller's state of the boundary cell (plane index idx along
// grid axis `axis`), in the frame of the face (x and y exchanged with swap).
template <bool PER, bool PG, typename T>
__device__ __forceinline__ Cons<T> axis_face_g(int gpos, int n, int bclo, int bchi, Prim<T> L, Prim<T> R, T g,
                                               T igm1, int dim, const GridDev& G, int axis, long long idx,
                                               bool swap) {
  if constexpr (!PG) return axis_face<PER>(gpos, n, bclo, bchi, L, R, g, igm1, dim);
  if (!PER && bclo != BC_PERIODIC) {
    if (gpos == 0) L = bclo == BC_PARENT ? swap_xy(parent_ghost<T>(G, axis, 0, idx), swap) : ghost(R, bclo, dim);
    else if (gpos == n) R = bchi == BC_PARENT ? swap_xy(parent_ghost<T>(G, axis, 1, idx), swap) : ghost(L, bchi, dim);
  }
  return hllc(L, R, g, igm1, dim);
}

// Boundary-flux register (src/reflux.jl:105-132): b[slot] += (s·F)·area for a
// domain-boundary face, s = ½dt per RK stage. Slots: planes ordered (axis 0 lo,
// axis 0 hi, axis 1 lo, ...), each column-major over the other axes; idx is the
// face's index in its plane; G.bbase holds each plane's first slot. Each
// boundary face is evaluated by one lane per stage, so the read-modify-write
// needs no atomics.
template <typename T>
__device__ __forceinline__ void bflux_add(const GridDev& G, int axis, int side, long long idx, const Cons<T>& F,
                                          T s, T a) {
  T* b = (T*)G.bflux + 5 * (G.bbase[2 * axis + side] + idx);
  b[0] = b[0] + (s * F.d) * a;
  b[1] = b[1] + (s * F.mx) * a;
  b[2] = b[2] + (s * F.my) * a;
  b[3] = b[3] + (s * F.mz) * a;
  b[4] = b[4] + (s * F.e) * a;
}

// One RK stage. STAGE2 = false: Uout = Us − (dt/V)·acc(Us).
// STAGE2 = true: Uout = ½·U0 + ½·(Us − (dt/V)·acc(Us)) written over U0 (point
// reads only), plus the per-workgroup max of the CFL term of the new state.
// PER: every axis of the grid is periodic (no ghosts, no boundary faces, no
// boundary-flux register), resolved at compile time. GRAV: self-gravity on
// (G.phi set): the gravity source joins the flux sum and the free-fall limiter
// the CFL maximum.
template <typename T, int RANK, int TX, int TY, bool STAGE2, bool FLOOR64, bool PER, bool GRAV, bool PG = false>
__global__ void __launch_bounds__(TX* TY, sizeof(T) == 4 ? VH_MIN_WAVES_F32 : VH_MIN_WAVES) rk_stage_kernel(const T* __restrict__ Us, const T* U0,
                                                          T* Uout, const double* __restrict__ dtp,
                                                          double* __restrict__ partials, GridDev G) {
  constexpr int NT = TX * TY;
  __shared__ PlaneLDS<T, TX, TY, RANK> S;
  using L_ = PlaneLDS<T, TX, TY, RANK>;
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
  const int sxc = resolve_axis<PER>(gx, nx, G.bc[0], ghx);
  const int syc = resolve_axis<PER>(gy, ny, G.bc[2], ghy);
  const long long col = (long long)syc * nx + sxc;

  // ---- z pipeline prologue ----
  Prim<T> wz0, wz1, wz2, sz0;
  Cons<T> fzlo;
  if constexpr (RANK == 3) {
    const Cons<T> um2 = load_cons(Us, n, column_src<PER>(G, col, kz0 - 2));
    const Cons<T> um1 = load_cons(Us, n, column_src<PER>(G, col, kz0 - 1));
    const Cons<T> u0 = load_cons(Us, n, column_src<PER>(G, col, kz0));
    const Cons<T> u1 = load_cons(Us, n, column_src<PER>(G, col, kz0 + 1));
    const Cons<T> u2 = load_cons(Us, n, column_src<PER>(G, col, kz0 + 2));
    const Prim<T> wm2 = column_w<T, FLOOR64, PER, PG>(um2, G, kz0 - 2, g, col);
    const Prim<T> wm1 = column_w<T, FLOOR64, PER, PG>(um1, G, kz0 - 1, g, col);
    wz0 = column_w<T, FLOOR64, PER, PG>(u0, G, kz0, g, col);
    wz1 = column_w<T, FLOOR64, PER, PG>(u1, G, kz0 + 1, g, col);
    wz2 = column_w<T, FLOOR64, PER, PG>(u2, G, kz0 + 2, g, col);
    // The slope of the ghost plane kz0-1 is never used when kz0-1 is outside a
    // non-periodic domain: that face is a boundary face (lo state = ghost).
    const Prim<T> sm1 = slope(wm2, wm1, wz0);
    sz0 = slope(wm1, wz0, wz1);
    const int bz = G.bc[4];
    if (!PER && bz != BC_PERIODIC && kz0 == 0) {
      const Prim<T> R = face_value(wz0, sz0, T(-0.5));
      fzlo = hllc(bghost<PG>(R, G, 2, 0, col), R, g, igm1, 3);
      if (G.bflux && active) bflux_add(G, 2, 0, (long long)gy * nx + gx, fzlo, T(0.5 * dt), a2);
    } else {
      fzlo = hllc(face_value(wm1, sm1, T(0.5)), face_value(wz0, sz0, T(-0.5)), g, igm1, 3);
    }
  }

  CflMax<T, RANK> cfl;

  // Halo cells (cross-shaped, two deep in x and y) of a plane: item h of
  // NHX + NHY, staged by lane h counted from the plane's rotating first wave.
  constexpr int NHX = 4 * TY;
  constexpr int NHY = (RANK > 1) ? 4 * TX : 0;
  constexpr bool PREFETCH = VH_HALO_PREFETCH && (NHX + NHY <= NT);
  auto lane_item = [&](int kk) { return (tid + 64 * ((kk + (int)blockIdx.x) % (NT / 64))) % NT; };
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
  // Stage a halo cell of plane kk given its conserved state (ghost rule applied here).
  auto stage_halo = [&](int lx, int ly, int kk, const Cons<T>& u) {
    int hgx, hgy;
    const int sx = resolve_axis<PER>(x0 + lx, nx, G.bc[0], hgx);
    const int sy = resolve_axis<PER>(y0 + ly, ny, G.bc[2], hgy);
    Prim<T> w = cons2prim<T, FLOOR64>(u, g, G.gamma);
    if (hgx) w = bghost<PG>(w, G, 0, hgx - 1, (long long)kk * ny + sy);
    if (hgy) w = bghost<PG>(w, G, 1, hgy - 1, (long long)kk * nx + sx);
    VH_PUT5(S.w, [ly + L_::HY][lx + 2], w);
  };
  auto halo_src = [&](int lx, int ly, int kk) -> long long {
    int hgx, hgy;
    const int sx = resolve_axis<PER>(x0 + lx, nx, G.bc[0], hgx);
    const int sy = resolve_axis<PER>(y0 + ly, ny, G.bc[2], hgy);
    return (long long)kk * plane + (long long)sy * nx + sx;
  };
  Cons<T> uhalo;
  if constexpr (PREFETCH) {
    const int h = lane_item(kz0);
    if (h < NHX + NHY) {
      int lx, ly;
      halo_pos(h, lx, ly);
      uhalo = load_cons(Us, n, halo_src(lx, ly, kz0));
    }
  }

  // Periodic form: the z face of the plane is computed in P0 behind the halo
  // loads (stage 1 -3 to -8%, stage 2 -1.5% at 256³ FP64); the boundary forms
  // keep it in PB, where they have the registers for it.
  constexpr bool ZEARLY = PER && !PREFETCH && NHX + NHY <= NT;
  for (int k = kz0; k < kz1; ++k) {
    const long long c = (long long)k * plane + (long long)gy * nx + gx;
    Cons<T> fzhi;   // z face at k+½ of the lane's column
    Prim<T> sz1;    // its z slope at k+1
    // the z face at k+½ of the lane's column and the z slope of plane k+1
    auto z_face = [&]() {
      if constexpr (RANK == 3) {
        sz1 = slope(wz0, wz1, wz2);
        const int bz = G.bc[4];
        const Prim<T> L = face_value(wz0, sz0, T(0.5));
        if (!PER && bz != BC_PERIODIC && k == nz - 1) {
          fzhi = hllc(L, bghost<PG>(L, G, 2, 1, col), g, igm1, 3);
          if (G.bflux && active) bflux_add(G, 2, 1, (long long)gy * nx + gx, fzhi, T(0.5 * dt), a2);
        } else
          fzhi = hllc(L, face_value(wz1, sz1, T(-0.5)), g, igm1, 3);
      }
    };
    // Early loads: the next z plane of the column and this cell's own state.
    Cons<T> unext, uown, u0own;
    if constexpr (RANK == 3) {
      if (k + 1 < kz1) unext = load_cons(Us, n, column_src<PER>(G, col, k + 3));
    }
    constexpr bool EARLY = VH_EARLY_OWN && PER && !STAGE2 && sizeof(T) == 8;
    if constexpr (EARLY) {
      if (active) {
        uown = load_cons(Us, n, c);
        if constexpr (STAGE2) u0own = load_cons_stream(U0, n, c);
      }
    }

    // ---- P0: stage the plane's primitive state in LDS ----
    Prim<T> wc;
    if constexpr (RANK == 3) {
      wc = wz0;
    } else {
      wc = cons2prim<T, FLOOR64>(load_cons(Us, n, col), g, G.gamma);
    }
    {
      Prim<T> w = wc;
      if (ghx) w = bghost<PG>(w, G, 0, ghx - 1, (long long)k * ny + syc);
      if (ghy) w = bghost<PG>(w, G, 1, ghy - 1, (long long)k * nx + sxc);
      VH_PUT5(S.w, [ty + L_::HY][tx + 2], w);
      wc = w;
    }
    // The halo, rim and edge work below goes to lanes counted from wave `rot`,
    // which moves by one wave per plane so every SIMD gets its share.
    const int lt = lane_item(k);
    if constexpr (PREFETCH) {
      if (lt < NHX + NHY) {
        int lx, ly;
        halo_pos(lt, lx, ly);
        stage_halo(lx, ly, k, uhalo);
      }
      // the next plane's halo cell of this lane, loaded while this plane runs
      const int hn = lane_item(k + 1);
      if (k + 1 < kz1 && hn < NHX + NHY) {
        int lx, ly;
        halo_pos(hn, lx, ly);
        uhalo = load_cons(Us, n, halo_src(lx, ly, k + 1));
      }
    } else {
      if constexpr (ZEARLY) {
        // At most one halo item per lane: its loads are issued first and the
        // z face of the lane's column (registers only) is computed while they
        // are in flight.
        int lx = 0, ly = 0;
        Cons<T> uh;
        const bool hv = lt < NHX + NHY;
        if (hv) {
          halo_pos(lt, lx, ly);
          uh = load_cons(Us, n, halo_src(lx, ly, k));
        }
        z_face();
        if (hv) stage_halo(lx, ly, k, uh);
      } else {
        for (int h = lt; h < NHX + NHY; h += NT) {
          int lx, ly;
          halo_pos(h, lx, ly);
          stage_halo(lx, ly, k, load_cons(Us, n, halo_src(lx, ly, k)));
        }
      }
    }
    __syncthreads();

    // ---- PA: slopes once per cell; publish hi face states ----
    const T* wf = &S.w[0][0][0];
    T* ff = &S.f[0][0];
    Prim<T> xlo, ylo;
    {
      const int ic = (ty + L_::HY) * L_::WX + tx + 2;
      const Prim<T> sx = slope(lds_get5(wf, L_::WS, ic - 1), wc, lds_get5(wf, L_::WS, ic + 1));
      xlo = face_value(wc, sx, T(-0.5));
      lds_put5(ff, L_::FS, L_::xs(ty, tx + 1), face_value(wc, sx, T(0.5)));
      if constexpr (RANK > 1) {
        const Prim<T> sy = slope(lds_get5(wf, L_::WS, ic - L_::WX), wc, lds_get5(wf, L_::WS, ic + L_::WX));
        ylo = face_value(wc, sy, T(-0.5));
        lds_put5(ff, L_::FS, L_::ys(ty + 1, tx), face_value(wc, sy, T(0.5)));
      }
    }
    // Halo rim, one code path for the four kinds: x hi state of cell -1 and lo
    // state of cell TX per row, y hi state of cell -1 and lo state of cell TY
    // per column.
    constexpr int NRX = 2 * TY;
    constexpr int NRY = (RANK > 1) ? 2 * TX : 0;
    for (int h = lt; h < NRX + NRY; h += NT) {
      const bool isy = h >= NRX;
      const int q = isy ? h - NRX : h;
      const int line = q >> 1;
      const bool hi = q & 1;   // the lo state of the cell past the tile's hi edge
      int ic, dst;
      if (!isy) {
        ic = (line + L_::HY) * L_::WX + (hi ? TX + 2 : 1);
        dst = L_::xs(line, hi ? TX + 1 : 0);
      } else {
        ic = (hi ? TY + 2 : 1) * L_::WX + line + 2;
        dst = L_::ys(hi ? TY + 1 : 0, line);
      }
      const int d = isy ? L_::WX : 1;
      const Prim<T> wm = lds_get5(wf, L_::WS, ic);
      const Prim<T> sl = slope(lds_get5(wf, L_::WS, ic - d), wm, lds_get5(wf, L_::WS, ic + d));
      lds_put5(ff, L_::FS, dst, face_value(wm, sl, hi ? T(-0.5) : T(0.5)));
    }
    __syncthreads();

    // ---- PB: face fluxes (each face once), written over the consumed state ----
    const int bx = G.bc[0], by = G.bc[2];
    if (gy < ny && gx <= nx) {
      const int sl = L_::xs(ty, tx);
      const Cons<T> F = axis_face_g<PER, PG>(gx, nx, bx, G.bc[1], lds_get5(ff, L_::FS, sl), xlo, g, igm1, 1, G, 0,
                                         (long long)k * ny + gy, false);
      lds_putf(ff, L_::FS, sl, F);
      if (!PER && G.bflux && bx != BC_PERIODIC && (gx == 0 || gx == nx))
        bflux_add(G, 0, gx == nx, (long long)k * ny + gy, F, T(0.5 * dt), a0);
    }
    if constexpr (RANK > 1) {
      if (gx < nx && gy <= ny) {
        const int sl = L_::ys(ty, tx);
        const Cons<T> F = axis_face_g<PER, PG>(gy, ny, by, G.bc[3], lds_get5(ff, L_::FS, sl), ylo, g, igm1, 2, G, 1,
                                           (long long)k * nx + gx, false);
        lds_putf(ff, L_::FS, sl, F);
        if (!PER && G.bflux && by != BC_PERIODIC && (gy == 0 || gy == ny))
          bflux_add(G, 1, gy == ny, (long long)k * nx + gx, F, T(0.5 * dt), a1);
      }
    }
    // Tile hi-edge faces, x face TX of each row and y face TY of each column,
    // through one code path: a y face is evaluated in the frame of an x face
    // (x and y components exchanged on the way in and out).
    constexpr int NEX = TY;
    constexpr int NEY = (RANK > 1) ? TX : 0;
    for (int h = lt; h < NEX + NEY; h += NT) {
      const bool isy = h >= NEX;
      const int line = isy ? h - NEX : h;
      const int gp = isy ? y0 + TY : x0 + TX;
      const int na = isy ? ny : nx;
      const bool inside = isy ? (x0 + line < nx) : (y0 + line < ny);
      if (inside && gp <= na) {
        const int sl = isy ? L_::ys(TY, line) : L_::xs(line, TX);
        const int sr = isy ? L_::ys(TY + 1, line) : L_::xs(line, TX + 1);
        const Prim<T> L = swap_xy(lds_get5(ff, L_::FS, sl), isy);
        const Prim<T> R = swap_xy(lds_get5(ff, L_::FS, sr), isy);
        const Cons<T> F = swap_xy(axis_face_g<PER, PG>(gp, na, isy ? by : bx, isy ? G.bc[3] : G.bc[1], L, R, g, igm1, 1,
                                                   G, isy ? 1 : 0,
                                                   isy ? (long long)k * nx + x0 + line : (long long)k * ny + y0 + line,
                                                   isy),
                                  isy);
        lds_putf(ff, L_::FS, sl, F);
        if (!PER && G.bflux && (isy ? by : bx) != BC_PERIODIC && gp == na)
          bflux_add(G, isy ? 1 : 0, 1, isy ? (long long)k * nx + x0 + line : (long long)k * ny + y0 + line, F,
                    T(0.5 * dt), isy ? a1 : a0);
      }
    }

    // z face at k+½ for the lane's column (here unless computed in P0).
    if constexpr (!ZEARLY) {
      z_face();
    }
    __syncthreads();

    // ---- PC: gather the six face fluxes in the reference order and update ----
    if (active) {
      if constexpr (!EARLY) {
        uown = load_cons(Us, n, c);
        if constexpr (STAGE2) u0own = load_cons_stream(U0, n, c);
      }
      Cons<T> fxlo, fxhi, fylo, fyhi;
      fxlo = lds_getf(ff, L_::FS, L_::xs(ty, tx));
      fxhi = lds_getf(ff, L_::FS, L_::xs(ty, tx + 1));
      if constexpr (RANK > 1) {
        fylo = lds_getf(ff, L_::FS, L_::ys(ty, tx));
        fyhi = lds_getf(ff, L_::FS, L_::ys(ty + 1, tx));
      }
      // Lo-face kinds: 0 interior (partner earlier), 1 domain boundary, 2 periodic wrap.
      const int kx = gx > 0 ? 0 : (PER || bx == BC_PERIODIC ? 2 : 1);
      const int ky = gy > 0 ? 0 : (PER || by == BC_PERIODIC ? 2 : 1);
      const int kz = k > 0 ? 0 : (PER || G.bc[4] == BC_PERIODIC ? 2 : 1);
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

      if constexpr (GRAV) grav_source_add<T, RANK>(acc, uown, (const T*)G.phi, G, c, gx, gy, k);

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
        cfl.add(cons2prim<T, FLOOR64>(o, g, G.gamma), g);
      }
      store_cons_stream(Uout, n, c, o);
    }

    if constexpr (RANK == 3) {
      fzlo = fzhi;
      sz0 = sz1;
      wz0 = wz1;
      wz1 = wz2;
      if (k + 1 < kz1) wz2 = column_w<T, FLOOR64, PER, PG>(unext, G, k + 3, g, col);
    }
  }

  if constexpr (STAGE2) {
    static_assert(sizeof(S.w) >= (NT / 64) * sizeof(double), "reduction scratch");
    const double m = block_max<NT>(cfl.value(G.w, GRAV, G.gravG), reinterpret_cast<double*>(&S.w[0][0][0]));
    if (tid == 0) partials[blockIdx.x] = m;
  }
}

// Per-workgroup max of the CFL term over a grid-stride range of cells.
template <typename T, int RANK, bool FLOOR64>
__global__ void __launch_bounds__(256) cfl_kernel(const T* __restrict__ U, GridDev G,
                                                  double* __restrict__ partials) {
  __shared__ double red[16];
  const T g = T(G.gamma);
  CflMax<T, RANK> cfl;
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < G.ncell; c += (long long)gridDim.x * 256)
    cfl.add(cons2prim<T, FLOOR64>(load_cons(U, G.ncell, c), g, G.gamma), g);
  const double m = block_max<256>(cfl.value(G.w, G.phi != nullptr, G.gravG), red);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
