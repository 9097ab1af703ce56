/*
 * jvp_oracle.cc — forward-mode (Jacobian-vector product) instantiation of the
 * oracle's SSP-RK2 step (TEST INFRASTRUCTURE ONLY; linked into liboracle.so).
 *
 * oracle_impl.h is the operation-for-operation restatement of the RefMesh step
 * (src/eos.jl:8-32, src/reconstruct.jl:6-16, src/riemann.jl:9-88,
 * src/driver.jl:145-359). Here it is compiled once more with its field type OT
 * set to a dual number that carries up to VO_NDIR tangents beside the value:
 * every +, −, ×, ÷, √ and |·| propagates the tangents by the chain rule, every
 * comparison looks at the value only. The value part therefore follows exactly
 * the double-precision oracle (same operations, same order, no contraction:
 * bit-identical states), and the tangent part is J·v of that same evaluation.
 *
 * The step is not differentiable everywhere. Where it is not, the tangent takes
 * the branch the primal takes, with the same rules as the reverse mode
 * (adjoint_oracle.c and the GPU adjoint): minmod gives 0 when a·b ≤ 0, else the
 * tangent of the argument of smaller magnitude (b on ties); the Davis speeds
 * min(a, b) / max(a, b) and the pressure floor max(p, 1e-12·K) take the first
 * argument's tangent on ties; the HLLC branch follows the signs of SL, SR, S*.
 * With both modes using one linear map per evaluation point,
 * ⟨VJP(w), v⟩ = ⟨w, JVP(v)⟩ holds to round-off even on minmod and HLLC kinks,
 * which is what pins the reverse mode where finite differences cannot.
 * dt is held fixed per step, as in the reverse mode.
 */
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "grid_impl.h"

namespace {

constexpr int VO_NDIR = 16;

struct Dual {
    double v;
    double d[VO_NDIR];
    Dual() = default;
    Dual(double x) : v(x) {  // NOLINT: constants and geometry enter as values
        for (int k = 0; k < VO_NDIR; ++k) d[k] = 0.0;
    }
    explicit operator double() const { return v; }
};

inline Dual operator+(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v + b.v;
    for (int k = 0; k < VO_NDIR; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v - b.v;
    for (int k = 0; k < VO_NDIR; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline Dual operator-(const Dual& a) {
    Dual r;
    r.v = -a.v;
    for (int k = 0; k < VO_NDIR; ++k) r.d[k] = -a.d[k];
    return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
    for (int k = 0; k < VO_NDIR; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
    const double q = r.v;
    for (int k = 0; k < VO_NDIR; ++k) r.d[k] = (a.d[k] - q * b.d[k]) / b.v;
    return r;
}
inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }

inline Dual dsqrt(const Dual& a) {
    Dual r;
    r.v = std::sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int k = 0; k < VO_NDIR; ++k) r.d[k] = a.d[k] * h;
    return r;
}
inline Dual dfabs(const Dual& a) { return std::signbit(a.v) ? -a : a; }

/* Julia's min/max for the value (NaN-propagating, −0 < +0); on a tie of values
 * the first argument's tangent, the reverse mode's rule (Davis speeds:
 * `aL <= aR` / `bL >= bR` pick the left state; the floor is active only when
 * strictly larger). */
inline Dual jmin_dual(const Dual& x, const Dual& y) {
    Dual r = y.v < x.v ? y : x;
    r.v = vo_jmin_f64_(x.v, y.v);
    return r;
}
inline Dual jmax_dual(const Dual& x, const Dual& y) {
    Dual r = y.v > x.v ? y : x;
    r.v = vo_jmax_f64_(x.v, y.v);
    return r;
}

#define OIMPL_OWN_MINMAX
#define OT Dual
#define OPG double
#define OSUF dual
#define OSQRT dsqrt
#define OFABS dfabs
#include "oracle_impl.h"
#undef OT
#undef OSUF
#undef OSQRT
#undef OFABS
#undef OIMPL_OWN_MINMAX

/* Dual state from U (5n values) and ndir tangent arrays V[k*5n + i]. */
void load_dual(int64_t nv, const double* U, int ndir, const double* V, Dual* X) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nv; ++i) {
        X[i] = Dual(U[i]);
        for (int k = 0; k < ndir; ++k) X[i].d[k] = V[(int64_t)k * nv + i];
    }
}

void store_dual(int64_t nv, const Dual* X, int ndir, double* U, double* JV) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nv; ++i) {
        if (U) U[i] = X[i].v;
        for (int k = 0; k < ndir; ++k) JV[(int64_t)k * nv + i] = X[i].d[k];
    }
}

}  // namespace

extern "C" {

int vo_jvp_max_directions(void) { return VO_NDIR; }

/* nsteps SSP-RK2 steps with the given dts (held fixed) from U, carrying ndir
 * tangent directions V (ndir × 5n, direction-major): Unext = the state after the
 * steps (bit-identical to the oracle's steps), JV = J·V of the whole rollout.
 * Gather form over cells, OpenMP (threads <= 0: runtime default). */
int vo_rollout_jvp_f64(const vo_grid* m, const double* U, int64_t nsteps, const double* dts, int ndir,
                       const double* V, int threads, double* Unext, double* JV) {
    if (vo_check(m) || ndir < 1 || ndir > VO_NDIR || nsteps < 0) return -1;
    if (threads > 0) omp_set_num_threads(threads);
    const int64_t n = vo_ncell(m);
    Dual* X = (Dual*)std::malloc(sizeof(Dual) * 5 * n);
    Dual* work = (Dual*)std::malloc(sizeof(Dual) * 10 * n);
    if (!X || !work) {
        std::free(X);
        std::free(work);
        return -1;
    }
    load_dual(5 * n, U, ndir, V, X);
    for (int64_t s = 0; s < nsteps; ++s) vo_step_omp_dual(m, X, dts[s], work, 0, threads);
    store_dual(5 * n, X, ndir, Unext, JV);
    std::free(X);
    std::free(work);
    return 0;
}

/* One step (dt fixed): the rollout above with nsteps = 1. */
int vo_step_jvp_f64(const vo_grid* m, const double* U, double dt, int ndir, const double* V, int threads,
                    double* Unext, double* JV) {
    return vo_rollout_jvp_f64(m, U, 1, &dt, ndir, V, threads, Unext, JV);
}

/* J·v of cons2prim and of the HLLC flux for single states (the scalar pieces,
 * one direction): the tangents of the outputs for input tangents dU / (dWL, dWR). */
void vo_c2p_jvp_f64(const double* U, double gamma, const double* dU, double* W, double* dW) {
    Dual Ud[5], Wd[5];
    for (int f = 0; f < 5; ++f) {
        Ud[f] = Dual(U[f]);
        Ud[f].d[0] = dU[f];
    }
    c2p_dual(Ud, Dual(gamma), gamma, 0, Wd);
    for (int f = 0; f < 5; ++f) {
        W[f] = Wd[f].v;
        dW[f] = Wd[f].d[0];
    }
}

void vo_hllc_jvp_f64(const double* WL, const double* WR, double gamma, int dim, const double* dWL,
                     const double* dWR, double* F, double* dF) {
    Dual L[5], R[5], o[5];
    for (int f = 0; f < 5; ++f) {
        L[f] = Dual(WL[f]);
        L[f].d[0] = dWL[f];
        R[f] = Dual(WR[f]);
        R[f].d[0] = dWR[f];
    }
    hllc_dual(L, R, Dual(gamma), dim, o);
    for (int f = 0; f < 5; ++f) {
        F[f] = o[f].v;
        dF[f] = o[f].d[0];
    }
}

}  // extern "C"
