// Reverse mode of one SSP-RK2 step on a uniform grid (Float64, dt fixed).
//
// Forward (invV = dt/V): U1 = U − invV·A(U),  U2 = ½U + ½(U1 − invV·A(U1)).
// Reverse with w = ∂L/∂U2:
//   λ1 = ½w − ½·invV·A'(U1)ᵀw,   ∂L/∂U = ½w + λ1 − invV·A'(U)ᵀλ1.
// A'(U)ᵀλ is evaluated in three passes:
//   adj_prim_kernel   W = cons2prim(U) per cell;
//   adj_face_kernel   one lane per face: the face states (PLM-minmod), the
//                     flux VJP applied to area·(λ_left − λ_right) (one side at a
//                     domain boundary, with the ghost folded back), giving the
//                     adjoints of the two face states;
//   adj_cell_kernel   one lane per cell: gathers the face-state adjoints of its
//                     own faces and the slope adjoints of itself and its two
//                     neighbours per axis (minmod active branch), maps the
//                     primitive adjoint to conserved variables and writes
//                     a1·base1 + a2·base2 + beta·invV·A'ᵀλ.
// Non-smooth points take the active branch: minmod's selected argument (zero
// when the signs differ), the Davis min/max active argument, the upwind HLLC
// branch, the pressure floor when it is the larger argument.
#pragma once

#include "hydro.h"

namespace vh {
namespace VH_PROFILE_NS {

using D5 = Prim<double>;

// Division and the sound speed: correctly rounded in the exact profile, refined
// hardware reciprocal / rsqrt in the fast profile (the reverse mode needs no bit
// parity; it is tested to 1e-12 against a scatter-form CPU restatement).
__device__ __forceinline__ double adiv(double a, double b) {
#if VH_FAST_PROFILE
  return fast_div(a, b);
#else
  return a / b;
#endif
}
__device__ __forceinline__ double asound(double g, double p, double r) { return sound(g, p, r); }

template <int A>
__device__ __forceinline__ double& ncomp(D5& w) {
  if constexpr (A == 0) return w.vx;
  else if constexpr (A == 1) return w.vy;
  else return w.vz;
}
template <int A>
__device__ __forceinline__ double ncomp(const D5& w) {
  if constexpr (A == 0) return w.vx;
  else if constexpr (A == 1) return w.vy;
  else return w.vz;
}

__device__ __forceinline__ void d5_add(D5& a, const D5& b) {
  a.r += b.r; a.vx += b.vx; a.vy += b.vy; a.vz += b.vz; a.p += b.p;
}
__device__ __forceinline__ D5 d5_zero() {
  D5 z;
  z.r = z.vx = z.vy = z.vz = z.p = 0.0;
  return z;
}

// Transpose of the ghost map (outflow: identity; reflecting: flip the normal velocity).
template <int A>
__device__ __forceinline__ D5 ghostT(const D5& b, int bc) {
  D5 o = b;
  if (bc == BC_REFLECTING) ncomp<A>(o) = -ncomp<A>(b);
  return o;
}

// cons2prim adjoint at U: Wb → Ub (src/eos.jl:8-25).
__device__ __forceinline__ Cons<double> c2p_vjp(const Cons<double>& U, double g, const D5& Wb) {
  const double rho = U.d;
  const double inv = recip(rho);
  const double vx = U.mx * inv, vy = U.my * inv, vz = U.mz * inv;
  const double v2 = (vx * vx + vy * vy) + vz * vz;
  const double kin = (0.5 * rho) * v2;
  const double pc = (g - 1.0) * (U.e - kin);
  const double pf = 1e-12 * kin;
  double kinb, Eb;
  if (pf > pc) {
    kinb = Wb.p * 1e-12;
    Eb = 0.0;
  } else {
    kinb = -Wb.p * (g - 1.0);
    Eb = Wb.p * (g - 1.0);
  }
  const double vbx = Wb.vx + kinb * rho * vx, vby = Wb.vy + kinb * rho * vy, vbz = Wb.vz + kinb * rho * vz;
  const double invb = vbx * U.mx + vby * U.my + vbz * U.mz;
  Cons<double> o;
  o.d = Wb.r + kinb * 0.5 * v2 - invb * inv * inv;
  o.mx = vbx * inv;
  o.my = vby * inv;
  o.mz = vbz * inv;
  o.e = Eb;
  return o;
}

__device__ __forceinline__ double energy64(const D5& W, double g) {
  return adiv(W.p, g - 1.0) + (0.5 * W.r) * ((W.vx * W.vx + W.vy * W.vy) + W.vz * W.vz);
}

__device__ __forceinline__ void energy_vjp(const D5& W, double g, double Eb, D5& Wb) {
  Wb.r += Eb * 0.5 * ((W.vx * W.vx + W.vy * W.vy) + W.vz * W.vz);
  Wb.vx += Eb * W.r * W.vx;
  Wb.vy += Eb * W.r * W.vy;
  Wb.vz += Eb * W.r * W.vz;
  Wb.p += adiv(Eb, g - 1.0);
}

// euler_flux adjoint (src/riemann.jl:9-20) along axis A (0-based).
template <int A>
__device__ __forceinline__ void eflux_vjp(const D5& W, double g, const Cons<double>& Fb, D5& Wb) {
  const double r = W.r, p = W.p, un = ncomp<A>(W);
  const double E = energy64(W, g);
  double rb = Fb.d * un, unb = Fb.d * r, pb = 0.0;
  const double mbx = Fb.mx * un, mby = Fb.my * un, mbz = Fb.mz * un;
  unb += Fb.mx * (r * W.vx) + Fb.my * (r * W.vy) + Fb.mz * (r * W.vz);
  pb += A == 0 ? Fb.mx : (A == 1 ? Fb.my : Fb.mz);
  rb += mbx * W.vx + mby * W.vy + mbz * W.vz;
  Wb.vx += mbx * r;
  Wb.vy += mby * r;
  Wb.vz += mbz * r;
  const double Eb = Fb.e * un;
  pb += Fb.e * un;
  unb += Fb.e * (E + p);
  energy_vjp(W, g, Eb, Wb);
  Wb.r += rb;
  Wb.p += pb;
  ncomp<A>(Wb) += unb;
}

// hllc_flux adjoint (src/riemann.jl:33-88) along axis A: accumulates WLb, WRb.
template <int A>
__device__ void hllc_vjp(const D5& WL, const D5& WR, double g, const Cons<double>& ob, D5& WLb, D5& WRb) {
  const double rL = WL.r, rR = WR.r, pL = WL.p, pR = WR.p;
  const double unL = ncomp<A>(WL), unR = ncomp<A>(WR);
  const double cL = asound(g, pL, rL), cR = asound(g, pR, rR);
  const double xL = cL * cL, xR = cR * cR;
  const double aL = unL - cL, aR = unR - cR, bL = unL + cL, bR = unR + cR;
  const double SL = vmin(aL, aR), SR = vmax(bL, bR);
  if (SL >= 0.0) {
    eflux_vjp<A>(WL, g, ob, WLb);
    return;
  }
  if (SR <= 0.0) {
    eflux_vjp<A>(WR, g, ob, WRb);
    return;
  }
  const double tL = rL * unL, tR = rR * unR;
  const double num = pR - pL + tL * (SL - unL) - tR * (SR - unR);
  const double den = rL * (SL - unL) - rR * (SR - unR);
#if VH_FAST_PROFILE
  // one refined reciprocal per denominator, shared by the value and its adjoint
  const double deninv = recip(den);
  const double Ss = num * deninv;
#else
  const double Ss = adiv(num, den);
#endif
  const bool useL = Ss >= 0.0;
  const D5 W = useL ? WL : WR;
  D5 Wb = d5_zero();
  const double SK = useL ? SL : SR;
  const double r = W.r, un = ncomp<A>(W), p = W.p;
  const double E = energy64(W, g);
  const double q = r * (SK - un);
  const double d3 = SK - Ss;
  const double rinv = recip(r);
  const double qinv = recip(q);
  const double d3inv = recip(d3);
#if VH_FAST_PROFILE
  const double factor = q * d3inv;
  const double vs0 = A == 0 ? Ss : W.vx, vs1 = A == 1 ? Ss : W.vy, vs2 = A == 2 ? Ss : W.vz;
  const double Aa = Ss - un, Bb = Ss + p * qinv;
  const double Es = E * rinv + Aa * Bb;
#else
  const double factor = adiv(q, d3);
  const double vs0 = A == 0 ? Ss : W.vx, vs1 = A == 1 ? Ss : W.vy, vs2 = A == 2 ? Ss : W.vz;
  const double Aa = Ss - un, Bb = Ss + adiv(p, q);
  const double Es = adiv(E, r) + Aa * Bb;
#endif
  // out = F(W) + SK·(Us − Uk), Us = factor·(1, vs, Es), Uk = (r, r·v, E)
  const double SKb0 = ob.d * (factor - r) + ob.mx * (factor * vs0 - r * W.vx) + ob.my * (factor * vs1 - r * W.vy) +
                      ob.mz * (factor * vs2 - r * W.vz) + ob.e * (factor * Es - E);
  double SKb = SKb0;
  eflux_vjp<A>(W, g, ob, Wb);
  const double u0 = SK * ob.d, u1 = SK * ob.mx, u2 = SK * ob.my, u3 = SK * ob.mz, u4 = SK * ob.e;
  const double factorb = u0 + u1 * vs0 + u2 * vs1 + u3 * vs2 + u4 * Es;
  const double vsb0 = u1 * factor, vsb1 = u2 * factor, vsb2 = u3 * factor;
  const double Esb = u4 * factor;
  double rb = 0.0, unb = 0.0, pb = 0.0, Eb = 0.0, Ssb = 0.0, qb = 0.0;
  // Uk = prim2cons(W) with adjoint −SK·ob
  rb -= u0;
  rb -= u1 * W.vx + u2 * W.vy + u3 * W.vz;
  Wb.vx -= u1 * r;
  Wb.vy -= u2 * r;
  Wb.vz -= u3 * r;
  Eb -= u4;
  // Es = E/r + Aa·Bb
  Eb += Esb * rinv;
  rb -= Esb * E * rinv * rinv;
  const double Ab = Esb * Bb, Bbb = Esb * Aa;
  Ssb += Ab;
  unb -= Ab;
  Ssb += Bbb;
  pb += Bbb * qinv;
  qb -= Bbb * p * qinv * qinv;
  if (A == 0) { Ssb += vsb0; } else { Wb.vx += vsb0; }
  if (A == 1) { Ssb += vsb1; } else { Wb.vy += vsb1; }
  if (A == 2) { Ssb += vsb2; } else { Wb.vz += vsb2; }
  // factor = q/(SK − Ss)
  qb += factorb * d3inv;
  const double d3b = -factorb * factor * d3inv;
  SKb += d3b;
  Ssb -= d3b;
  // q = r·(SK − un)
  rb += qb * (SK - un);
  SKb += qb * r;
  unb -= qb * r;
  energy_vjp(W, g, Eb, Wb);
  Wb.r += rb;
  Wb.p += pb;
  ncomp<A>(Wb) += unb;
  if (useL) d5_add(WLb, Wb);
  else d5_add(WRb, Wb);
  double SLb = useL ? SKb : 0.0, SRb = useL ? 0.0 : SKb;
  // Ss = num/den
#if !VH_FAST_PROFILE
  const double deninv = recip(den);
#endif
  const double numb = Ssb * deninv, denb = -Ssb * Ss * deninv;
  double rLb = 0, rRb = 0, pLb = 0, pRb = 0, unLb = 0, unRb = 0, cLb = 0, cRb = 0;
  pRb += numb;
  pLb -= numb;
  const double tLb = numb * (SL - unL);
  rLb += tLb * unL;
  unLb += tLb * rL;
  SLb += numb * tL;
  unLb -= numb * tL;
  const double tRb = -numb * (SR - unR);
  rRb += tRb * unR;
  unRb += tRb * rR;
  SRb -= numb * tR;
  unRb += numb * tR;
  rLb += denb * (SL - unL);
  SLb += denb * rL;
  unLb -= denb * rL;
  rRb -= denb * (SR - unR);
  SRb -= denb * rR;
  unRb += denb * rR;
  if (aL <= aR) { unLb += SLb; cLb -= SLb; } else { unRb += SLb; cRb -= SLb; }
  if (bL >= bR) { unLb += SRb; cLb += SRb; } else { unRb += SRb; cRb += SRb; }
  const double xLb = adiv(cLb, 2.0 * cL), xRb = adiv(cRb, 2.0 * cR);
  const double rLi = recip(rL), rRi = recip(rR);
  pLb += xLb * g * rLi;
  rLb -= xLb * xL * rLi;
  pRb += xRb * g * rRi;
  rRb -= xRb * xR * rRi;
  WLb.r += rLb;
  WLb.p += pLb;
  ncomp<A>(WLb) += unLb;
  WRb.r += rRb;
  WRb.p += pRb;
  ncomp<A>(WRb) += unRb;
}

// minmod(c − l, r − c) adjoint per component: sb → (lb, cb, rb) contributions.
__device__ __forceinline__ void minmod_vjp1(double l, double c, double r, double sb, double& lb, double& cb,
                                            double& rb) {
  const double a = c - l, b = r - c;
  if (a * b <= 0.0) return;
  if (vabs(a) < vabs(b)) {
    cb += sb;
    lb -= sb;
  } else {
    rb += sb;
    cb -= sb;
  }
}
__device__ __forceinline__ void slope_vjp(const D5& l, const D5& c, const D5& r, const D5& sb, D5& lb, D5& cb,
                                          D5& rb) {
  minmod_vjp1(l.r, c.r, r.r, sb.r, lb.r, cb.r, rb.r);
  minmod_vjp1(l.vx, c.vx, r.vx, sb.vx, lb.vx, cb.vx, rb.vx);
  minmod_vjp1(l.vy, c.vy, r.vy, sb.vy, lb.vy, cb.vy, rb.vy);
  minmod_vjp1(l.vz, c.vz, r.vz, sb.vz, lb.vz, cb.vz, rb.vz);
  minmod_vjp1(l.p, c.p, r.p, sb.p, lb.p, cb.p, rb.p);
}

__device__ __forceinline__ D5 load_w(const double* __restrict__ W, long long n, long long c) {
  D5 w;
  w.r = W[c]; w.vx = W[n + c]; w.vy = W[2 * n + c]; w.vz = W[3 * n + c]; w.p = W[4 * n + c];
  return w;
}
__device__ __forceinline__ void store_w(double* __restrict__ W, long long n, long long c, const D5& w) {
  W[c] = w.r; W[n + c] = w.vx; W[2 * n + c] = w.vy; W[3 * n + c] = w.vz; W[4 * n + c] = w.p;
}

__global__ void __launch_bounds__(256) adj_prim_kernel(const double* __restrict__ U, double* __restrict__ W,
                                                       GridDev G) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.ncell) return;
  store_w(W, G.ncell, c, cons2prim<double, false>(load_cons(U, G.ncell, c), G.gamma, G.gamma));
}

// Strides and face-slot geometry along axis A.
struct AxisGeo {
  long long stride;   // cell index stride along A
  int n;              // cells along A
};
template <int A>
__device__ __forceinline__ AxisGeo axis_geo(const GridDev& G) {
  AxisGeo a;
  a.n = G.n[A];
  a.stride = A == 0 ? 1 : (A == 1 ? (long long)G.n[0] : (long long)G.n[0] * G.n[1]);
  return a;
}

// Face value of cell c (coordinate p along A) toward side h (±½) with the
// neighbours resolved: periodic wrap, else the ghost of the cell itself.
template <int A>
__device__ __forceinline__ D5 face_state(const double* __restrict__ W, const GridDev& G, long long c, int p, double h) {
  const AxisGeo ax = axis_geo<A>(G);
  const long long n = G.ncell;
  const D5 wc = load_w(W, n, c);
  const bool per = G.bc[2 * A] == BC_PERIODIC;
  D5 wl, wr;
  if (p > 0) wl = load_w(W, n, c - ax.stride);
  else if (per) wl = load_w(W, n, c + (long long)(ax.n - 1) * ax.stride);
  else wl = ghost(wc, G.bc[2 * A], A + 1);
  if (p < ax.n - 1) wr = load_w(W, n, c + ax.stride);
  else if (per) wr = load_w(W, n, c - (long long)(ax.n - 1) * ax.stride);
  else wr = ghost(wc, G.bc[2 * A + 1], A + 1);
  return face_value(wc, slope(wl, wc, wr), h);
}

// Face slots along A: (n_A + 1) per line, slot coordinate q ∈ [0, n_A] is the
// face between cells q−1 and q (periodic axes use q ∈ [0, n_A) with q = 0 the
// wrapped face). Output gL/gR: 5 SoA fields over the slots.
template <int A>
__global__ void __launch_bounds__(256) adj_face_kernel(const double* __restrict__ W, const double* __restrict__ lam,
                                                       GridDev G, double* __restrict__ gL, double* __restrict__ gR) {
  const int m0 = A == 0 ? G.n[0] + 1 : G.n[0];
  const int m1 = A == 1 ? G.n[1] + 1 : G.n[1];
  const int m2 = A == 2 ? G.n[2] + 1 : G.n[2];
  const long long ns = (long long)m0 * m1 * m2;
  const long long s = (long long)blockIdx.x * 256 + threadIdx.x;
  if (s >= ns) return;
  const int i = (int)(s % m0), j = (int)((s / m0) % m1), k = (int)(s / ((long long)m0 * m1));
  const int q = A == 0 ? i : (A == 1 ? j : k);
  const AxisGeo ax = axis_geo<A>(G);
  const bool per = G.bc[2 * A] == BC_PERIODIC;
  const long long n = G.ncell;
  D5 bL = d5_zero(), bR = d5_zero();
  if (per && q == ax.n) {   // unused slot of a periodic axis
    store_w(gL, ns, s, bL);
    store_w(gR, ns, s, bR);
    return;
  }
  const int ci = A == 0 ? (q < ax.n ? q : ax.n - 1) : i;
  const int cj = A == 1 ? (q < ax.n ? q : ax.n - 1) : j;
  const int ck = A == 2 ? (q < ax.n ? q : ax.n - 1) : k;
  const long long cell = ((long long)ck * G.n[1] + cj) * G.n[0] + ci;   // cell q (or the last cell at q = n)
  const double a = G.area[A];
  const double g = G.gamma;
  if (q == 0 && !per) {
    // lo boundary face: F = hllc(ghost(WR), WR), acc[0] −= a·F
    const D5 WR = face_state<A>(W, G, cell, 0, -0.5);
    const D5 Wg = ghost(WR, G.bc[2 * A], A + 1);
    const Cons<double> lr = load_cons(lam, n, cell);
    Cons<double> mu;
    mu.d = -a * lr.d; mu.mx = -a * lr.mx; mu.my = -a * lr.my; mu.mz = -a * lr.mz; mu.e = -a * lr.e;
    D5 gb = d5_zero();
    hllc_vjp<A>(Wg, WR, g, mu, gb, bR);
    d5_add(bR, ghostT<A>(gb, G.bc[2 * A]));
  } else if (q == ax.n) {
    // hi boundary face: F = hllc(WL, ghost(WL)), acc[n−1] += a·F
    const D5 WL = face_state<A>(W, G, cell, ax.n - 1, 0.5);
    const D5 Wg = ghost(WL, G.bc[2 * A + 1], A + 1);
    const Cons<double> ll = load_cons(lam, n, cell);
    Cons<double> mu;
    mu.d = a * ll.d; mu.mx = a * ll.mx; mu.my = a * ll.my; mu.mz = a * ll.mz; mu.e = a * ll.e;
    D5 gb = d5_zero();
    hllc_vjp<A>(WL, Wg, g, mu, bL, gb);
    d5_add(bL, ghostT<A>(gb, G.bc[2 * A + 1]));
  } else {
    // interior (or wrapped) face between cells q−1 and q
    const long long cl = q > 0 ? cell - ax.stride : cell + (long long)(ax.n - 1) * ax.stride;
    const int ql = q > 0 ? q - 1 : ax.n - 1;
    const D5 WL = face_state<A>(W, G, cl, ql, 0.5);
    const D5 WR = face_state<A>(W, G, cell, q, -0.5);
    const Cons<double> ll = load_cons(lam, n, cl), lr = load_cons(lam, n, cell);
    Cons<double> mu;
    mu.d = a * ll.d - a * lr.d;
    mu.mx = a * ll.mx - a * lr.mx;
    mu.my = a * ll.my - a * lr.my;
    mu.mz = a * ll.mz - a * lr.mz;
    mu.e = a * ll.e - a * lr.e;
    hllc_vjp<A>(WL, WR, g, mu, bL, bR);
  }
  store_w(gL, ns, s, bL);
  store_w(gR, ns, s, bR);
}

// Primitive adjoint gathered by cell c from axis A.
template <int A>
__device__ __forceinline__ void gather_axis(const double* __restrict__ W, const double* __restrict__ gL,
                                            const double* __restrict__ gR, const GridDev& G, long long c, int i, int j,
                                            int k, D5& Wb) {
  const AxisGeo ax = axis_geo<A>(G);
  const int m0 = A == 0 ? G.n[0] + 1 : G.n[0];
  const int m1 = A == 1 ? G.n[1] + 1 : G.n[1];
  const int m2 = A == 2 ? G.n[2] + 1 : G.n[2];
  const long long ns = (long long)m0 * m1 * m2;
  const int p = A == 0 ? i : (A == 1 ? j : k);
  const int nA = ax.n;
  const bool per = G.bc[2 * A] == BC_PERIODIC;
  const long long n = G.ncell;
  // slot of face coordinate qf along A (wrapped on periodic axes)
  auto slot = [&](int qf) -> long long {
    if (per) qf = qf < 0 ? qf + nA : (qf >= nA ? qf - nA : qf);
    const int si = A == 0 ? qf : i, sj = A == 1 ? qf : j, sk = A == 2 ? qf : k;
    return ((long long)sk * m1 + sj) * m0 + si;
  };
  auto cell_at = [&](int pp) -> long long {   // cell at coordinate pp (wrapped when periodic)
    if (per) pp = pp < 0 ? pp + nA : (pp >= nA ? pp - nA : pp);
    return c + (long long)(pp - p) * ax.stride;
  };
  const D5 wc = load_w(W, n, c);
  // own faces
  const D5 ghi = load_w(gL, ns, slot(p + 1));
  const D5 glo = load_w(gR, ns, slot(p));
  d5_add(Wb, ghi);
  d5_add(Wb, glo);
  // own slope
  {
    const bool hl = p > 0 || per, hr = p < nA - 1 || per;
    const D5 wl = hl ? load_w(W, n, cell_at(p - 1)) : ghost(wc, G.bc[2 * A], A + 1);
    const D5 wr = hr ? load_w(W, n, cell_at(p + 1)) : ghost(wc, G.bc[2 * A + 1], A + 1);
    D5 sb;
    sb.r = 0.5 * ghi.r - 0.5 * glo.r;
    sb.vx = 0.5 * ghi.vx - 0.5 * glo.vx;
    sb.vy = 0.5 * ghi.vy - 0.5 * glo.vy;
    sb.vz = 0.5 * ghi.vz - 0.5 * glo.vz;
    sb.p = 0.5 * ghi.p - 0.5 * glo.p;
    D5 lb = d5_zero(), cb = d5_zero(), rb = d5_zero();
    slope_vjp(wl, wc, wr, sb, lb, cb, rb);
    d5_add(Wb, cb);
    if (!hl) d5_add(Wb, ghostT<A>(lb, G.bc[2 * A]));
    if (!hr) d5_add(Wb, ghostT<A>(rb, G.bc[2 * A + 1]));
  }
  // slope of the hi neighbour h (c is its lo neighbour)
  if (p < nA - 1 || per) {
    const long long h = cell_at(p + 1);
    const D5 wh = load_w(W, n, h);
    const bool hr = p + 1 < nA - 1 || per;
    const D5 whr = hr ? load_w(W, n, cell_at(p + 2)) : ghost(wh, G.bc[2 * A + 1], A + 1);
    const D5 a = load_w(gL, ns, slot(p + 2)), b = load_w(gR, ns, slot(p + 1));
    D5 sb;
    sb.r = 0.5 * a.r - 0.5 * b.r;
    sb.vx = 0.5 * a.vx - 0.5 * b.vx;
    sb.vy = 0.5 * a.vy - 0.5 * b.vy;
    sb.vz = 0.5 * a.vz - 0.5 * b.vz;
    sb.p = 0.5 * a.p - 0.5 * b.p;
    D5 lb = d5_zero(), cb = d5_zero(), rb = d5_zero();
    slope_vjp(wc, wh, whr, sb, lb, cb, rb);
    d5_add(Wb, lb);
  }
  // slope of the lo neighbour l (c is its hi neighbour)
  if (p > 0 || per) {
    const long long l = cell_at(p - 1);
    const D5 wl = load_w(W, n, l);
    const bool hl = p - 1 > 0 || per;
    const D5 wll = hl ? load_w(W, n, cell_at(p - 2)) : ghost(wl, G.bc[2 * A], A + 1);
    const D5 a = load_w(gL, ns, slot(p)), b = load_w(gR, ns, slot(p - 1));
    D5 sb;
    sb.r = 0.5 * a.r - 0.5 * b.r;
    sb.vx = 0.5 * a.vx - 0.5 * b.vx;
    sb.vy = 0.5 * a.vy - 0.5 * b.vy;
    sb.vz = 0.5 * a.vz - 0.5 * b.vz;
    sb.p = 0.5 * a.p - 0.5 * b.p;
    D5 lb = d5_zero(), cb = d5_zero(), rb = d5_zero();
    slope_vjp(wll, wl, wc, sb, lb, cb, rb);
    d5_add(Wb, rb);
  }
}

struct AdjFaces {
  const double* gL[3];
  const double* gR[3];
};

// out = a1·base1 + a2·base2 + (beta·dt/V)·A'(U)ᵀλ (base2 may be null). out may
// alias base1 or base2 (each lane reads and writes its own cell only).
template <int RANK>
__global__ void __launch_bounds__(256) adj_cell_kernel(const double* __restrict__ U, const double* __restrict__ W,
                                                       AdjFaces F, const double* base1, double a1,
                                                       const double* base2, double a2, double beta,
                                                       const double* __restrict__ dtp, GridDev G, double* out) {
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= G.ncell) return;
  const int i = (int)(c % G.n[0]), j = (int)((c / G.n[0]) % G.n[1]), k = (int)(c / ((long long)G.n[0] * G.n[1]));
  D5 Wb = d5_zero();
  gather_axis<0>(W, F.gL[0], F.gR[0], G, c, i, j, k, Wb);
  if constexpr (RANK > 1) gather_axis<1>(W, F.gL[1], F.gR[1], G, c, i, j, k, Wb);
  if constexpr (RANK > 2) gather_axis<2>(W, F.gL[2], F.gR[2], G, c, i, j, k, Wb);
  const Cons<double> Ub = c2p_vjp(load_cons(U, G.ncell, c), G.gamma, Wb);
  const double s = beta * (*dtp / (1.0 * G.V));
  const long long n = G.ncell;
  const double ub[5] = {Ub.d, Ub.mx, Ub.my, Ub.mz, Ub.e};
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    double v = a1 * base1[f * n + c];
    if (base2) v += a2 * base2[f * n + c];
    out[f * n + c] = v + s * ub[f];
  }
}

// ---------------------------------------------------------------------------
// Fused A'(U)ᵀλ for 2D grids periodic on both axes: one workgroup per 32×8 tile
// computes W for the tile and a cross-shaped halo of three cells, every face
// adjoint its cells need (the tile's faces plus one more on each side), and the
// gather, in LDS; one launch instead of four, no face buffers in HBM.
constexpr int JTX = 32, JTY = 16, JNT = 512;
constexpr int JNXF = JTY * (JTX + 3);   // x faces of a tile (one extra on the lo side, two on the hi)
constexpr int JNYF = (JTY + 3) * JTX;   // y faces of a tile

// One 32×16 tile per workgroup of 512 lanes (one workgroup per compute unit:
// 256 VGPRs per lane, 150 KiB of LDS).
struct JatLDS {
  double Wx[5][JTY][JTX + 6];   // W, rows y0..y0+15, cols x0-3..x0+34
  double Wy[5][6][JTX];         // W, rows y0-3..y0-1 and y0+16..y0+18, cols x0..x0+31
  double Lx[5][JTY][JTX + 4];   // λ, rows y0..y0+15, cols x0-2..x0+33
  double Ly[5][4][JTX];         // λ, rows y0-2, y0-1, y0+16, y0+17
  double Gx[10][JNXF];          // x-face state adjoints (left 0..4, right 5..9)
  double Gy[10][JNYF];          // y-face state adjoints
};

// Flat LDS offsets of the cross-shaped W and λ halos: (row, col) in tile
// coordinates, row ∈ [−3, JTY+2] with col ∈ [0, JTX) or row ∈ [0, JTY) with
// col ∈ [−3, JTX+2]; returns the address of field 0 and the field pitch.
__device__ __forceinline__ const double* jat_w_at(const JatLDS& S, int row, int col, int& pitch) {
  if (row >= 0 && row < JTY) {
    pitch = JTY * (JTX + 6);
    return &S.Wx[0][row][col + 3];
  }
  pitch = 6 * JTX;
  return &S.Wy[0][row < 0 ? row + 3 : 3 + row - JTY][col];
}
__device__ __forceinline__ const double* jat_l_at(const JatLDS& S, int row, int col, int& pitch) {
  if (row >= 0 && row < JTY) {
    pitch = JTY * (JTX + 4);
    return &S.Lx[0][row][col + 2];
  }
  pitch = 4 * JTX;
  return &S.Ly[0][row < 0 ? row + 2 : 2 + row - JTY][col];
}
__device__ __forceinline__ D5 jat_w(const JatLDS& S, int row, int col) {
  int p;
  const double* a = jat_w_at(S, row, col, p);
  D5 w;
  w.r = a[0]; w.vx = a[p]; w.vy = a[2 * p]; w.vz = a[3 * p]; w.p = a[4 * p];
  return w;
}
__device__ __forceinline__ Cons<double> jat_l(const JatLDS& S, int row, int col) {
  int p;
  const double* a = jat_l_at(S, row, col, p);
  Cons<double> l;
  l.d = a[0]; l.mx = a[p]; l.my = a[2 * p]; l.mz = a[3 * p]; l.e = a[4 * p];
  return l;
}
__device__ __forceinline__ D5 jat_gx(const JatLDS& S, int side, int i) {
  const int o = 5 * side;
  D5 w;
  w.r = S.Gx[o][i]; w.vx = S.Gx[o + 1][i]; w.vy = S.Gx[o + 2][i]; w.vz = S.Gx[o + 3][i]; w.p = S.Gx[o + 4][i];
  return w;
}
__device__ __forceinline__ D5 jat_gy(const JatLDS& S, int side, int i) {
  const int o = 5 * side;
  D5 w;
  w.r = S.Gy[o][i]; w.vx = S.Gy[o + 1][i]; w.vy = S.Gy[o + 2][i]; w.vz = S.Gy[o + 3][i]; w.p = S.Gy[o + 4][i];
  return w;
}
__device__ __forceinline__ D5 half_diff(const D5& a, const D5& b) {
  D5 s;
  s.r = 0.5 * a.r - 0.5 * b.r;
  s.vx = 0.5 * a.vx - 0.5 * b.vx;
  s.vy = 0.5 * a.vy - 0.5 * b.vy;
  s.vz = 0.5 * a.vz - 0.5 * b.vz;
  s.p = 0.5 * a.p - 0.5 * b.p;
  return s;
}
__device__ __forceinline__ Cons<double> mu_of(double a, const Cons<double>& l, const Cons<double>& r) {
  Cons<double> m;
  m.d = a * l.d - a * r.d;
  m.mx = a * l.mx - a * r.mx;
  m.my = a * l.my - a * r.my;
  m.mz = a * l.mz - a * r.mz;
  m.e = a * l.e - a * r.e;
  return m;
}
__device__ __forceinline__ D5 d5_swap_xy(const D5& w, bool on) {
  D5 o = w;
  o.vx = on ? w.vy : w.vx;
  o.vy = on ? w.vx : w.vy;
  return o;
}
__device__ __forceinline__ Cons<double> c_swap_xy(const Cons<double>& f, bool on) {
  Cons<double> o = f;
  o.mx = on ? f.my : f.mx;
  o.my = on ? f.mx : f.my;
  return o;
}

// Slope adjoint terms gathered by a cell from its own slope and its two
// neighbours' slopes along one axis. w[0..4] = W at offsets −2..+2; gl[k], gr[k]
// = left/right face-state adjoints of the faces at offsets −3/2.. +3/2 (k = 0..3).
__device__ __forceinline__ void gather_line(const D5 w[5], const D5 gl[4], const D5 gr[4], D5& Wb) {
  // own faces: hi face (k = 2) left state, lo face (k = 1) right state
  d5_add(Wb, gl[2]);
  d5_add(Wb, gr[1]);
  D5 lb = d5_zero(), cb = d5_zero(), rb = d5_zero();
  slope_vjp(w[1], w[2], w[3], half_diff(gl[2], gr[1]), lb, cb, rb);
  d5_add(Wb, cb);
  D5 lb2 = d5_zero(), cb2 = d5_zero(), rb2 = d5_zero();
  slope_vjp(w[2], w[3], w[4], half_diff(gl[3], gr[2]), lb2, cb2, rb2);   // hi neighbour: this cell is its lo
  d5_add(Wb, lb2);
  D5 lb3 = d5_zero(), cb3 = d5_zero(), rb3 = d5_zero();
  slope_vjp(w[0], w[1], w[2], half_diff(gl[1], gr[0]), lb3, cb3, rb3);   // lo neighbour: this cell is its hi
  d5_add(Wb, rb3);
}

// out = a1·base1 + a2·base2 + beta·(dt/V)·A'(U)ᵀλ for a 2D grid periodic on both
// axes, one launch. Phases (two barriers): load the tile's W (halo 3) and λ (halo
// 2) in a cross shape; evaluate every x and y face VJP of the tile plus the faces
// its cells' slope adjoints reach (560 + 608 faces, three passes over the 512
// lanes; a y face runs in the frame of an x face, x and y exchanged); gather
// both axes per cell and map the primitive adjoint to conserved variables.
__global__ void __launch_bounds__(JNT, 1) adj_jat2d_periodic_kernel(const double* __restrict__ U,
                                                                 const double* __restrict__ lam,
                                                                 const double* base1, double a1,
                                                                 const double* base2, double a2, double beta,
                                                                 const double* __restrict__ dtp, GridDev G,
                                                                 double* out) {
  __shared__ JatLDS S;
  const int nx = G.n[0], ny = G.n[1];
  const long long n = G.ncell;
  const int tid = threadIdx.x;
  const int tx = tid % JTX, ty = tid / JTX;
  const int x0 = blockIdx.x * JTX, y0 = blockIdx.y * JTY;
  const double g = G.gamma;
  auto wrapx = [&](int x) { return x < 0 ? x + nx : (x >= nx ? x - nx : x); };
  auto wrapy = [&](int y) { return y < 0 ? y + ny : (y >= ny ? y - ny : y); };
  auto cidx = [&](int x, int y) { return (long long)wrapy(y) * nx + wrapx(x); };
  // ---- load: own cell, then the cross-shaped halos of W (3) and λ (2) ----
  const long long c = (long long)(y0 + ty) * nx + (x0 + tx);
  const Cons<double> Uc = load_cons(U, n, c);
  {
    const D5 w = cons2prim<double, false>(Uc, g, g);
    S.Wx[0][ty][tx + 3] = w.r; S.Wx[1][ty][tx + 3] = w.vx; S.Wx[2][ty][tx + 3] = w.vy;
    S.Wx[3][ty][tx + 3] = w.vz; S.Wx[4][ty][tx + 3] = w.p;
    const Cons<double> l = load_cons(lam, n, c);
    S.Lx[0][ty][tx + 2] = l.d; S.Lx[1][ty][tx + 2] = l.mx; S.Lx[2][ty][tx + 2] = l.my;
    S.Lx[3][ty][tx + 2] = l.mz; S.Lx[4][ty][tx + 2] = l.e;
  }
  constexpr int NWX = 6 * JTY, NLX = 4 * JTY, NWY = 6 * JTX, NLY = 4 * JTX;
  static_assert(NWX + NLX + NWY + NLY <= JNT, "one halo item per lane");
  {
    const int h = tid;
    if (h < NWX) {
      const int r = h / 6, q = h % 6;
      const int lc = q < 3 ? q : JTX + q;   // 0,1,2 | JTX+3..JTX+5
      const D5 w = cons2prim<double, false>(load_cons(U, n, cidx(x0 - 3 + lc, y0 + r)), g, g);
      S.Wx[0][r][lc] = w.r; S.Wx[1][r][lc] = w.vx; S.Wx[2][r][lc] = w.vy; S.Wx[3][r][lc] = w.vz; S.Wx[4][r][lc] = w.p;
    } else if (h < NWX + NLX) {
      const int hh = h - NWX;
      const int r = hh / 4, q = hh % 4;
      const int lc = q < 2 ? q : JTX + q;   // 0,1 | JTX+2, JTX+3
      const Cons<double> l = load_cons(lam, n, cidx(x0 - 2 + lc, y0 + r));
      S.Lx[0][r][lc] = l.d; S.Lx[1][r][lc] = l.mx; S.Lx[2][r][lc] = l.my; S.Lx[3][r][lc] = l.mz; S.Lx[4][r][lc] = l.e;
    } else if (h < NWX + NLX + NWY) {
      const int hh = h - NWX - NLX;
      const int i = hh / JTX, cc = hh % JTX;
      const int y = i < 3 ? y0 - 3 + i : y0 + JTY + (i - 3);
      const D5 w = cons2prim<double, false>(load_cons(U, n, cidx(x0 + cc, y)), g, g);
      S.Wy[0][i][cc] = w.r; S.Wy[1][i][cc] = w.vx; S.Wy[2][i][cc] = w.vy; S.Wy[3][i][cc] = w.vz; S.Wy[4][i][cc] = w.p;
    } else if (h < NWX + NLX + NWY + NLY) {
      const int hh = h - NWX - NLX - NWY;
      const int i = hh / JTX, cc = hh % JTX;
      const int y = i < 2 ? y0 - 2 + i : y0 + JTY + (i - 2);
      const Cons<double> l = load_cons(lam, n, cidx(x0 + cc, y));
      S.Ly[0][i][cc] = l.d; S.Ly[1][i][cc] = l.mx; S.Ly[2][i][cc] = l.my; S.Ly[3][i][cc] = l.mz; S.Ly[4][i][cc] = l.e;
    }
  }
  __syncthreads();

  // ---- face VJPs, x faces then y faces in one item list ----
  //   x face (r, ql), ql = 0..JTX+2: between columns ql−2 and ql−1 of row r;
  //   y face (ql, cc), ql = 0..JTY+2: between rows ql−2 and ql−1 of column cc.
  for (int h = tid; h < JNXF + JNYF; h += JNT) {
    const bool isy = h >= JNXF;
    const int q = isy ? h - JNXF : h;
    int row, col, dr, dc;
    if (!isy) {
      row = q / (JTX + 3);
      col = q % (JTX + 3) - 3;   // cell ql−3 (the face's w0)
      dr = 0;
      dc = 1;
    } else {
      row = q / JTX - 3;
      col = q % JTX;
      dr = 1;
      dc = 0;
    }
    const D5 w0 = jat_w(S, row, col), w1 = jat_w(S, row + dr, col + dc), w2 = jat_w(S, row + 2 * dr, col + 2 * dc),
             w3 = jat_w(S, row + 3 * dr, col + 3 * dc);
    const D5 WL = d5_swap_xy(face_value(w1, slope(w0, w1, w2), 0.5), isy);
    const D5 WR = d5_swap_xy(face_value(w2, slope(w1, w2, w3), -0.5), isy);
    const Cons<double> mu =
        c_swap_xy(mu_of(isy ? G.area[1] : G.area[0], jat_l(S, row + dr, col + dc), jat_l(S, row + 2 * dr, col + 2 * dc)),
                  isy);
    D5 bL = d5_zero(), bR = d5_zero();
    hllc_vjp<0>(WL, WR, g, mu, bL, bR);
    bL = d5_swap_xy(bL, isy);
    bR = d5_swap_xy(bR, isy);
    double* gb = isy ? &S.Gy[0][q] : &S.Gx[0][q];
    const int gp = isy ? JNYF : JNXF;
    gb[0] = bL.r; gb[gp] = bL.vx; gb[2 * gp] = bL.vy; gb[3 * gp] = bL.vz; gb[4 * gp] = bL.p;
    gb[5 * gp] = bR.r; gb[6 * gp] = bR.vx; gb[7 * gp] = bR.vy; gb[8 * gp] = bR.vz; gb[9 * gp] = bR.p;
  }
  __syncthreads();

  // ---- gather both axes per cell ----
  D5 Wb = d5_zero();
  {
    D5 w[5], gl[4], gr[4];
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = jat_w(S, ty, tx - 2 + k);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gl[k] = jat_gx(S, 0, ty * (JTX + 3) + tx + k);
      gr[k] = jat_gx(S, 1, ty * (JTX + 3) + tx + k);
    }
    gather_line(w, gl, gr, Wb);
  }
  {
    D5 w[5], gl[4], gr[4];
#pragma unroll
    for (int k = 0; k < 5; ++k) w[k] = jat_w(S, ty - 2 + k, tx);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gl[k] = jat_gy(S, 0, (ty + k) * JTX + tx);
      gr[k] = jat_gy(S, 1, (ty + k) * JTX + tx);
    }
    gather_line(w, gl, gr, Wb);
  }
  const Cons<double> Ub = c2p_vjp(Uc, g, Wb);
  const double sc = beta * (*dtp / (1.0 * G.V));
  const double ub[5] = {Ub.d, Ub.mx, Ub.my, Ub.mz, Ub.e};
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    double v = a1 * base1[f * n + c];
    if (base2) v += a2 * base2[f * n + c];
    out[f * n + c] = v + sc * ub[f];
  }
}

}  // namespace VH_PROFILE_NS
}  // namespace vh
