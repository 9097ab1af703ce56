/* Precision-templated self-gravity restatement (TEST INFRASTRUCTURE ONLY).
 *
 * Included by vespa_oracle.c after oracle_impl.h, once per field precision OT.
 * It restates, for the uniform mesh, the reference's gravity path:
 *   src/gravity.jl:99-128   apply_laplacian!  (A = −∇²·V, homogeneous Neumann at
 *                           non-periodic φ boundaries) in the CSR gather form of
 *                           _lap_csr_cpu! (src/gravity.jl:326-336);
 *   src/gravity.jl:212-246  Poisson right-hand side b = −4πG(ρ − ρ̄)·V;
 *   src/gravity.jl:248-322  solve_poisson! → _cg_csr_run! (plain CG, f64 dots,
 *                           warm start, volume-weighted mean projection);
 *   src/gravity.jl:353-410  _grav_accel / _gax (g = −∇φ, central differences);
 *   src/gravity.jl:388-447  _grav_src_tuple / apply_gravity_source! (−ρg·V on
 *                           momentum, −(ρv·g)·V on energy, added to the flux sum);
 *   src/gravity.jl:449-452  free-fall limiter √(4πGρ)/η folded into compute_dt;
 *   src/driver.jl:333-359, 483-508  step! / evolve! with gravity on.
 *
 * gper is a bit mask of the axes along which φ is periodic (its own boundary
 * conditions, independent of the gas). */

#define GCAT2(a, b) a##_##b
#define GCAT(a, b) GCAT2(a, b)
#define ON(name) GCAT(name, OSUF)

/* CSR row of a uniform cell in the order the face walk pushes its edges
 * (src/gravity.jl:540-556 counting sort over the for_each_face edge list): lo
 * neighbours that precede the cell (z, y, x), the cell's own hi neighbours
 * (x, y, z), then lo neighbours wrapped by a periodic φ boundary (x, y, z).
 * Returns the number of entries; nb[] the neighbours, ax[] their axes. */
static int ON(grav_row)(const vo_grid* m, int gper, int64_t c, int64_t nb[6], int ax[6]) {
    int k = 0;
    for (int d = m->rank - 1; d >= 0; --d) {
        int64_t i = vo_coord(m, c, d);
        if (i > 0) { nb[k] = c - vo_stride(m, d); ax[k++] = d; }
    }
    for (int d = 0; d < m->rank; ++d) {
        int64_t i = vo_coord(m, c, d), n = m->dims[d];
        if (i + 1 < n) { nb[k] = c + vo_stride(m, d); ax[k++] = d; }
        else if (gper >> d & 1) { nb[k] = c - (n - 1) * vo_stride(m, d); ax[k++] = d; }
    }
    for (int d = 0; d < m->rank; ++d) {
        int64_t i = vo_coord(m, c, d), n = m->dims[d];
        if (i == 0 && (gper >> d & 1)) { nb[k] = c + (n - 1) * vo_stride(m, d); ax[k++] = d; }
    }
    return k;
}

/* out = A·x, A = −∇²·V: Σ coef·(x_c − x_nb) in Float64, coef = area/d with the
 * face distance d = ½(w + w) (src/gravity.jl:89-93). */
void ON(vo_poisson_apply)(const vo_grid* m, int gper, const OT* x, OT* out) {
    int64_t n = vo_ncell(m);
    double coef[3];
    for (int d = 0; d < m->rank; ++d) {
        double w = vo_cell_width(m, d);
        coef[d] = vo_face_area(m, d) / (0.5 * (w + w));
    }
    for (int64_t c = 0; c < n; ++c) {
        int64_t nb[6];
        int ax[6];
        int k = ON(grav_row)(m, gper, c, nb, ax);
        double xc = (double)x[c], s = 0.0;
        for (int q = 0; q < k; ++q) s += coef[ax[q]] * (xc - (double)x[nb[q]]);
        out[c] = (OT)s;
    }
}

static double ON(dot64)(const OT* x, const OT* y, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (double)x[i] * (double)y[i];
    return s;
}

/* b[i] = Tf(−4πG(ρ_i − ρ̄)·V_i), ρ̄ = Σρ·V/ΣV when the mean is projected
 * (src/gravity.jl:230-246, inv_a = 1). */
void ON(vo_poisson_rhs)(const vo_grid* m, const OT* U, double G, int project_mean, OT* b) {
    int64_t n = vo_ncell(m);
    double vol = vo_cell_volume(m);
    double rbar = 0.0;
    if (project_mean) {
        double s = 0.0, vs = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            s += (double)U[i] * vol;
            vs += vol;
        }
        rbar = s / vs;
    }
    const double fourpi = 4.0 * 3.141592653589793;
    for (int64_t i = 0; i < n; ++i) b[i] = (OT)(-fourpi * G * ((double)U[i] - rbar) * vol * 1.0);
}

/* solve_poisson! (src/gravity.jl:248-322): CG from the warm start phi; work holds
 * 4·ncell values (r, p, Ap, b). Returns iterations; *relres the final residual. */
int64_t ON(vo_poisson_solve)(const vo_grid* m, int gper, const OT* U, double G, int project_mean,
                             int64_t maxiter, double tol, OT* phi, OT* work, double* relres_out) {
    int64_t n = vo_ncell(m);
    OT *r = work, *p = work + n, *Ap = work + 2 * n, *b = work + 3 * n;
    ON(vo_poisson_rhs)(m, U, G, project_mean, b);
    ON(vo_poisson_apply)(m, gper, phi, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    memcpy(p, r, sizeof(OT) * n);
    double rsold = ON(dot64)(r, r, n);
    double bnorm = sqrt(ON(dot64)(b, b, n));
    if (bnorm == 0.0) bnorm = 1.0;
    int64_t iters = 0;
    double relres = sqrt(rsold) / bnorm;
    for (int64_t k = 1; k <= maxiter; ++k) {
        iters = k;
        ON(vo_poisson_apply)(m, gper, p, Ap);
        double pAp = ON(dot64)(p, Ap, n);
        if (pAp <= 0.0) break;
        double alpha = rsold / pAp;
        for (int64_t i = 0; i < n; ++i) {
            phi[i] += (OT)alpha * p[i];
            r[i] -= (OT)alpha * Ap[i];
        }
        double rsnew = ON(dot64)(r, r, n);
        relres = sqrt(rsnew) / bnorm;
        if (relres <= tol) break;
        double beta = rsnew / rsold;
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + (OT)beta * p[i];
        rsold = rsnew;
    }
    if (project_mean) {
        double vol = vo_cell_volume(m), s = 0.0, vs = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            s += vol * (double)phi[i];
            vs += vol;
        }
        OT mu = (OT)(s / vs);
        for (int64_t i = 0; i < n; ++i) phi[i] -= mu;
    }
    *relres_out = relres;
    return iters;
}

/* _grav_accel / _gax (src/gravity.jl:372-410): g_d = −(φ_hi − φ_lo)/Tf(d_lo + d_hi);
 * a side on a non-periodic φ boundary takes φ_c and the half width. */
static void ON(grav_accel)(const vo_grid* m, int gper, const OT* phi, int64_t c, OT g[3]) {
    OT pc = phi[c];
    for (int d = 0; d < 3; ++d) {
        if (d >= m->rank) { g[d] = (OT)0; continue; }
        int64_t i = vo_coord(m, c, d), n = m->dims[d], s = vo_stride(m, d);
        double w = vo_cell_width(m, d);
        OT wlo, whi;
        double dlo, dhi;
        if (i > 0) { wlo = phi[c - s]; dlo = 0.5 * (w + w); }
        else if (gper >> d & 1) { wlo = phi[c + (n - 1) * s]; dlo = 0.5 * (w + w); }
        else { wlo = pc; dlo = 0.5 * w; }
        if (i + 1 < n) { whi = phi[c + s]; dhi = 0.5 * (w + w); }
        else if (gper >> d & 1) { whi = phi[c - (n - 1) * s]; dhi = 0.5 * (w + w); }
        else { whi = pc; dhi = 0.5 * w; }
        g[d] = -(whi - wlo) / (OT)(dlo + dhi);
    }
}

void ON(vo_grav_accel)(const vo_grid* m, int gper, const OT* phi, OT* g) {
    int64_t n = vo_ncell(m);
    for (int64_t c = 0; c < n; ++c) {
        OT gg[3];
        ON(grav_accel)(m, gper, phi, c, gg);
        for (int d = 0; d < 3; ++d) g[d * n + c] = gg[d];
    }
}

/* apply_gravity_source! (src/gravity.jl:388-447): acc += (0, −ρg_x·V, −ρg_y·V,
 * −ρg_z·V, −(ρv·g)·V) for the state U the stage's fluxes were computed from. */
static void ON(grav_source)(const vo_grid* m, int gper, const OT* phi, const OT* U, OT* acc) {
    int64_t n = vo_ncell(m);
    OT V = (OT)vo_cell_volume(m);
    for (int64_t c = 0; c < n; ++c) {
        OT g[3], Uc[5];
        ON(grav_accel)(m, gper, phi, c, g);
        ON(getU)(U, n, c, Uc);
        OT rho = Uc[0];
        OT es = (OT)0;
        for (int d = 0; d < 3; ++d) es += Uc[1 + d] * g[d];
        OT esV = es * V;
        acc[c] = acc[c] + (OT)0;
        for (int d = 0; d < 3; ++d) acc[(1 + d) * n + c] = acc[(1 + d) * n + c] + (-(rho * g[d] * V));
        acc[4 * n + c] = acc[4 * n + c] + (-esV);
    }
}

/* step! with gravity (src/driver.jl:333-359): the source after each flux sum, φ
 * (hence g) held over the step. work holds 10·ncell values. */
int ON(vo_step_grav)(const vo_grid* m, OT* U, double dt, OT* work, int floor64, int gper, const OT* phi) {
    if (vo_check(m)) return -1;
    int64_t n = vo_ncell(m);
    OT* u0 = work;
    OT* acc = work + 5 * n;
    OT invV = (OT)dt / ((OT)1.0 * (OT)vo_cell_volume(m));
    OT half = (OT)0.5;
    memcpy(u0, U, sizeof(OT) * 5 * n);
    ON(accumulate)(m, U, acc, floor64, NULL, (OT)0);
    ON(grav_source)(m, gper, phi, U, acc);
    for (int64_t i = 0; i < 5 * n; ++i) U[i] = u0[i] - invV * acc[i];
    ON(accumulate)(m, U, acc, floor64, NULL, (OT)0);
    ON(grav_source)(m, gper, phi, U, acc);
    for (int64_t i = 0; i < 5 * n; ++i) U[i] = half * u0[i] + half * (U[i] - invV * acc[i]);
    return 0;
}

/* compute_dt with the free-fall limiter (src/driver.jl:169-199, src/gravity.jl:449-452):
 * invdt = max(CFL term, max over cells of √(4πGρ)/0.5). */
double ON(vo_compute_dt_grav)(const vo_grid* m, const OT* U, double cfl, int floor64, double G) {
    if (vo_check(m)) return NAN;
    int64_t n = vo_ncell(m);
    OT g = (OT)m->gamma;
    double invdt = 0.0, ginv = 0.0;
    const double fourpi = 4.0 * 3.141592653589793;
    for (int64_t c = 0; c < n; ++c) {
        OT Uc[5], W[5];
        ON(getU)(U, n, c, Uc);
        ON(c2p)(Uc, g, m->gamma, floor64, W);
        OT cs = OSQRT(g * W[4] / W[0]);
        for (int d = 0; d < m->rank; ++d) {
            double w = vo_cell_width(m, d);
            invdt = vo_jmax_f64_(invdt, (double)(OFABS(W[d + 1]) + cs) / w);
        }
        double rho = (double)W[0];
        ginv = vo_jmax_f64_(ginv, rho > 0.0 ? sqrt(fourpi * G * rho) / 0.5 : 0.0);
    }
    invdt = vo_jmax_f64_(invdt, ginv);
    if (invdt == 0.0) return INFINITY;
    return cfl / invdt;
}

#undef ON
#undef GCAT
#undef GCAT2
