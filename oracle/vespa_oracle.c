/*
 * vespa_oracle.c — CPU oracle for the ghost-free flux-divergence hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: tests/, the smoke
 * entry point and the benchmark's cpu_baseline leg load it to compare the HIP
 * library against. Nothing in the product (vespa.jl_amd/) links, loads or calls it.
 *
 * It restates, operation for operation, Vespa.jl's RefMesh path:
 *   src/eos.jl (cons2prim/prim2cons/sound_speed), src/reconstruct.jl (minmod),
 *   src/riemann.jl (euler_flux, hllc_flux, _hllc_star), src/driver.jl
 *   (_plm_slope, _face_value, _flux_face!, accumulate_flux!, _euler_apply!,
 *   _rk2_combine!, step!, compute_dt, evolve!), src/diagnostics.jl
 *   (conserved_totals), src/exact_riemann.jl (Toro exact solver), and
 *   lib/RefMesh/src/RefMesh.jl (neighbor, face_area, cell_volume, for_each_face),
 *   and src/gravity.jl (Poisson CG, g = −∇φ, gravity source, free-fall limiter).
 *
 * The reference itself is Julia and no Julia toolchain exists here, so the
 * restatement is pinned by closed forms and invariants in tests/ (exact Riemann
 * star state, HLLC special cases, minmod table, round trips, conservation,
 * symmetry) and by an independent numpy restatement on small grids.
 *
 * Build: see oracle/Makefile (gcc, -ffp-contract=off: Julia never contracts).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { VO_BC_PERIODIC = 0, VO_BC_OUTFLOW = 1, VO_BC_REFLECTING = 2 };

/* Uniform mesh description (lib/RefMesh/src/RefMesh.jl:89-100). Axes beyond
 * rank have dims = 1 and take no faces. bc[2*axis + side], side 0 = lo, 1 = hi. */
typedef struct {
    int32_t rank;
    int32_t bc[6];
    int64_t dims[3];
    double lo[3];
    double hi[3];
    double gamma;
} vo_grid;

static int vo_check(const vo_grid* m) {
    if (m->rank < 1 || m->rank > 3) return 1;
    for (int d = 0; d < 3; ++d) {
        if (m->dims[d] < 1) return 1;
        if (d >= m->rank && m->dims[d] != 1) return 1;
    }
    for (int d = 0; d < m->rank; ++d) {
        int lo = m->bc[2 * d], hi = m->bc[2 * d + 1];
        if (lo < 0 || lo > 2 || hi < 0 || hi > 2) return 1;
        if ((lo == VO_BC_PERIODIC) != (hi == VO_BC_PERIODIC)) return 1;
    }
    return 0;
}

static inline int64_t vo_ncell(const vo_grid* m) { return m->dims[0] * m->dims[1] * m->dims[2]; }

/* lib/RefMesh/src/RefMesh.jl:122-123 _width = (hi - lo)/dims in Tg = Float64. */
static inline double vo_cell_width(const vo_grid* m, int d) {
    return (m->hi[d] - m->lo[d]) / (double)m->dims[d];
}
/* lib/RefMesh/src/RefMesh.jl:127 cell_volume = prod(widths), left fold. */
static inline double vo_cell_volume(const vo_grid* m) {
    double v = vo_cell_width(m, 0);
    for (int d = 1; d < m->rank; ++d) v = v * vo_cell_width(m, d);
    return v;
}
/* lib/RefMesh/src/RefMesh.jl:135-142 face_area: a = one; a *= w[d] for d != axis. */
static inline double vo_face_area(const vo_grid* m, int axis) {
    double a = 1.0;
    for (int d = 0; d < m->rank; ++d)
        if (d != axis) a *= vo_cell_width(m, d);
    return a;
}

static inline int64_t vo_stride(const vo_grid* m, int d) {
    return d == 0 ? 1 : (d == 1 ? m->dims[0] : m->dims[0] * m->dims[1]);
}
static inline int64_t vo_coord(const vo_grid* m, int64_t c, int d) {
    return (c / vo_stride(m, d)) % m->dims[d];
}

/* lib/RefMesh/src/RefMesh.jl:147-161 neighbor: returns 1 and the neighbor index for
 * an interior (or periodic-wrapped) neighbor, 0 for a domain boundary. */
static inline int vo_neighbor(const vo_grid* m, int64_t c, int d, int side, int64_t* nb) {
    int64_t n = m->dims[d], s = vo_stride(m, d);
    int64_t i = vo_coord(m, c, d);
    int64_t j = i + (side ? 1 : -1);
    if (j >= 0 && j < n) { *nb = c + (j - i) * s; return 1; }
    if (m->bc[2 * d + side] == VO_BC_PERIODIC) {
        int64_t jw = j < 0 ? n - 1 : 0;
        *nb = c + (jw - i) * s;
        return 1;
    }
    return 0;
}

/* Slot of a boundary face in the flat boundary-flux array: planes ordered
 * (axis 0 lo, axis 0 hi, axis 1 lo, ...), each column-major over the other axes. */
static int64_t vo_bface_slot(const vo_grid* m, int64_t c, int d, int side) {
    int64_t n = vo_ncell(m);
    int64_t base = 0;
    for (int e = 0; e < d; ++e) base += 2 * (n / m->dims[e]);
    if (side) base += n / m->dims[d];
    int64_t idx = 0, mul = 1;
    for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        idx += vo_coord(m, c, e) * mul;
        mul *= m->dims[e];
    }
    return base + idx;
}

int64_t vo_bface_count(const vo_grid* m) {
    int64_t n = vo_ncell(m), t = 0;
    for (int d = 0; d < m->rank; ++d) t += 2 * (n / m->dims[d]);
    return t;
}

static inline double vo_jmax_f64_(double x, double y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y > x) return y;
    if (x > y) return x;
    return signbit(y) ? x : y;
}
static inline double vo_jmin_f64_(double x, double y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y < x) return y;
    if (x < y) return x;
    return signbit(y) ? y : x;
}

#define OT double
#define OSUF f64
#define OSQRT sqrt
#define OFABS fabs
#include "oracle_impl.h"
#include "gravity_impl.h"
#undef OT
#undef OSUF
#undef OSQRT
#undef OFABS

#define OT float
#define OSUF f32
#define OSQRT sqrtf
#define OFABS fabsf
#include "oracle_impl.h"
#include "gravity_impl.h"
#undef OT
#undef OSUF
#undef OSQRT
#undef OFABS

/* ---------------- exact Riemann solver (src/exact_riemann.jl) ---------------- */

static inline double er_sound(double rho, double p, double g) { return sqrt(g * p / rho); }

/* src/exact_riemann.jl:11-22 pressure function f_K and its derivative. */
static void er_pressure_fn(double p, double rK, double pK, double cK, double g, double* f,
                           double* df) {
    if (p > pK) {
        double AK = 2 / ((g + 1) * rK);
        double BK = (g - 1) / (g + 1) * pK;
        *f = (p - pK) * sqrt(AK / (p + BK));
        *df = sqrt(AK / (BK + p)) * (1 - (p - pK) / (2 * (BK + p)));
    } else {
        *f = 2 * cK / (g - 1) * (pow(p / pK, (g - 1) / (2 * g)) - 1);
        *df = 1 / (rK * cK) * pow(p / pK, -(g + 1) / (2 * g));
    }
}

/* src/exact_riemann.jl:25-48 _star_state: Newton on p* (tol 1e-12, ≤100 it). */
void vo_star_state(const double* WL, const double* WR, double g, double* pstar, double* ustar) {
    const double tol = 1e-12;
    double rL = WL[0], uL = WL[1], pL = WL[2];
    double rR = WR[0], uR = WR[1], pR = WR[2];
    double cL = er_sound(rL, pL, g), cR = er_sound(rR, pR, g);
    double p = fmax(tol, 0.5 * (pL + pR));
    for (int it = 0; it < 100; ++it) {
        double fL, dfL, fR, dfR;
        er_pressure_fn(p, rL, pL, cL, g, &fL, &dfL);
        er_pressure_fn(p, rR, pR, cR, g, &fR, &dfR);
        double f = fL + fR + (uR - uL);
        double pn = p - f / (dfL + dfR);
        pn = fmax(tol, pn);
        if (fabs(pn - p) / (0.5 * (pn + p)) < tol) { p = pn; break; }
        p = pn;
    }
    double fL, fR, dd;
    er_pressure_fn(p, rL, pL, cL, g, &fL, &dd);
    er_pressure_fn(p, rR, pR, cR, g, &fR, &dd);
    *pstar = p;
    *ustar = 0.5 * (uL + uR) + 0.5 * (fR - fL);
}

/* src/exact_riemann.jl:56-120 exact_riemann_sample at s = x/t; out = (ρ, u, p). */
void vo_exact_riemann_sample(const double* WL, const double* WR, double g, double s, double* out) {
    double rL = WL[0], uL = WL[1], pL = WL[2];
    double rR = WR[0], uR = WR[1], pR = WR[2];
    double cL = er_sound(rL, pL, g), cR = er_sound(rR, pR, g);
    double ps, us;
    vo_star_state(WL, WR, g, &ps, &us);
    double r, u, p;
    if (s <= us) {
        if (ps > pL) {
            double SL = uL - cL * sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g));
            if (s <= SL) { r = rL; u = uL; p = pL; }
            else {
                r = rL * (ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1);
                u = us; p = ps;
            }
        } else {
            double SHL = uL - cL;
            double cs = cL * pow(ps / pL, (g - 1) / (2 * g));
            double STL = us - cs;
            if (s <= SHL) { r = rL; u = uL; p = pL; }
            else if (s >= STL) { r = rL * pow(ps / pL, 1 / g); u = us; p = ps; }
            else {
                u = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + s);
                double c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - s));
                r = rL * pow(c / cL, 2 / (g - 1));
                p = pL * pow(c / cL, 2 * g / (g - 1));
            }
        }
    } else {
        if (ps > pR) {
            double SR = uR + cR * sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g));
            if (s >= SR) { r = rR; u = uR; p = pR; }
            else {
                r = rR * (ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1);
                u = us; p = ps;
            }
        } else {
            double SHR = uR + cR;
            double cs = cR * pow(ps / pR, (g - 1) / (2 * g));
            double STR = us + cs;
            if (s >= SHR) { r = rR; u = uR; p = pR; }
            else if (s <= STR) { r = rR * pow(ps / pR, 1 / g); u = us; p = ps; }
            else {
                u = 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + s);
                double c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - s));
                r = rR * pow(c / cR, 2 / (g - 1));
                p = pR * pow(c / cR, 2 * g / (g - 1));
            }
        }
    }
    out[0] = r; out[1] = u; out[2] = p;
}

int vo_abi_version(void) { return 1; }

/* Block-structured AMR restatement (shares the scalar kernels above). */
#include "blocks_oracle.c"
#include "adjoint_oracle.c"
#include "muscl_oracle.c"
