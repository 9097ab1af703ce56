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

#include "grid_impl.h"

int64_t vo_bface_count(const vo_grid* m) {
    int64_t n = vo_ncell(m), t = 0;
    for (int d = 0; d < m->rank; ++d) t += 2 * (n / m->dims[d]);
    return t;
}

#define OT double
#define OPG double
#define OSUF f64
#define OSQRT sqrt
#define OFABS fabs
#include "oracle_impl.h"
#include "gravity_impl.h"
#undef OT
#undef OPG
#undef OSUF
#undef OSQRT
#undef OFABS

#define OT float
#define OPG float
#define OSUF f32
#define OSQRT sqrtf
#define OFABS fabsf
#include "oracle_impl.h"
#include "gravity_impl.h"
#undef OT
#undef OPG
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
