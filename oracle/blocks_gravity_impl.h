/* Precision-templated self-gravity on block meshes (TEST INFRASTRUCTURE ONLY).
 * Included by blocks_gravity.c with BT / BSUF / BSQRT / BFABS; follows
 * gravity_impl.h on the composite operator of bo_poisson. */

#define QCAT2(a, b) a##_##b
#define QCAT(a, b) QCAT2(a, b)
#define QN(name) QCAT(name, BSUF)

/* out = A·x (src/gravity.jl:326-336 _lap_csr_cpu!): per row, Σ coef·(x_c − x_nb)
 * in Float64 in CSR order, stored at BT. */
void QN(bo_poisson_apply)(const bo_poisson* P, const BT* x, BT* out) {
    for (int64_t c = 0; c < P->n; ++c) {
        double xc = (double)x[c], s = 0.0;
        for (int64_t k = P->rowptr[c]; k < P->rowptr[c + 1]; ++k) s += P->coef[k] * (xc - (double)x[P->colidx[k]]);
        out[c] = (BT)s;
    }
}

static double QN(bo_dot64)(const BT* x, const BT* y, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (double)x[i] * (double)y[i];
    return s;
}

/* b[i] = BT(−4πG(ρ_i − ρ̄)·V_i), ρ̄ = Σρ·V/ΣV in cell order (src/gravity.jl:230-246). */
void QN(bo_poisson_rhs)(const bo_poisson* P, const BT* U, double G, int project_mean, BT* b) {
    double rbar = 0.0;
    if (project_mean) {
        double s = 0.0, vs = 0.0;
        for (int64_t i = 0; i < P->n; ++i) {
            s += (double)U[i] * P->vol[i];
            vs += P->vol[i];
        }
        rbar = s / vs;
    }
    const double fourpi = 4.0 * 3.141592653589793;
    for (int64_t i = 0; i < P->n; ++i) b[i] = (BT)(-fourpi * G * ((double)U[i] - rbar) * P->vol[i] * 1.0);
}

/* _cg_csr_run! (src/gravity.jl:273-322) from the warm start phi; work holds 4n. */
int64_t QN(bo_poisson_solve)(const bo_poisson* P, const BT* U, double G, int project_mean, int64_t maxiter,
                             double tol, BT* phi, BT* work, double* relres_out) {
    int64_t n = P->n;
    BT *r = work, *p = work + n, *Ap = work + 2 * n, *b = work + 3 * n;
    QN(bo_poisson_rhs)(P, U, G, project_mean, b);
    QN(bo_poisson_apply)(P, phi, Ap);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
    memcpy(p, r, sizeof(BT) * n);
    double rsold = QN(bo_dot64)(r, r, n);
    double bnorm = sqrt(QN(bo_dot64)(b, b, n));
    if (bnorm == 0.0) bnorm = 1.0;
    int64_t iters = 0;
    double relres = sqrt(rsold) / bnorm;
    for (int64_t k = 1; k <= maxiter; ++k) {
        iters = k;
        QN(bo_poisson_apply)(P, p, Ap);
        double pAp = QN(bo_dot64)(p, Ap, n);
        if (pAp <= 0.0) break;
        double alpha = rsold / pAp;
        for (int64_t i = 0; i < n; ++i) {
            phi[i] += (BT)alpha * p[i];
            r[i] -= (BT)alpha * Ap[i];
        }
        double rsnew = QN(bo_dot64)(r, r, n);
        relres = sqrt(rsnew) / bnorm;
        if (relres <= tol) break;
        double beta = rsnew / rsold;
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + (BT)beta * p[i];
        rsold = rsnew;
    }
    if (project_mean) {
        double s = 0.0, vs = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            s += P->vol[i] * (double)phi[i];
            vs += P->vol[i];
        }
        BT mu = (BT)(s / vs);
        for (int64_t i = 0; i < n; ++i) phi[i] -= mu;
    }
    *relres_out = relres;
    return iters;
}

/* _gax (src/gravity.jl:403-410): g_d = −(φ_hi − φ_lo)/BT(d_lo + d_hi). */
static void QN(bo_gcell)(const bo_mesh* m, const bo_poisson* P, const BT* phi, int64_t c, BT g[3]) {
    for (int d = 0; d < 3; ++d) {
        if (d >= m->rank) { g[d] = (BT)0; continue; }
        BT wlo = phi[P->glo[3 * c + d]], whi = phi[P->ghi[3 * c + d]];
        g[d] = -(whi - wlo) / (BT)P->gsum[3 * c + d];
    }
}

void QN(bo_grav_accel)(const bo_mesh* m, const bo_poisson* P, const BT* phi, BT* g) {
    for (int64_t c = 0; c < P->n; ++c) {
        BT gg[3];
        QN(bo_gcell)(m, P, phi, c, gg);
        for (int d = 0; d < 3; ++d) g[d * P->n + c] = gg[d];
    }
}

/* apply_gravity_source! (src/gravity.jl:412-447) on every leaf. */
static void QN(bo_grav_source)(const bo_mesh* m, const bo_poisson* P, const BT* phi, const BT* U, BT* acc) {
    int64_t n = P->n;
    for (int b = 0; b < m->nblocks; ++b) {
        BT V = (BT)bo_volume(m, b);
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c) {
            BT g[3];
            QN(bo_gcell)(m, P, phi, c, g);
            BT rho = U[c];
            BT es = (BT)0;
            for (int d = 0; d < 3; ++d) es += U[(1 + d) * n + c] * g[d];
            BT esV = es * V;
            acc[c] = acc[c] + (BT)0;
            for (int d = 0; d < 3; ++d) acc[(1 + d) * n + c] = acc[(1 + d) * n + c] + (-(rho * g[d] * V));
            acc[4 * n + c] = acc[4 * n + c] + (-esV);
        }
    }
}

/* step! with gravity on a block mesh, every leaf (src/driver.jl:333-359): the
 * source after each flux sum, φ held over the step. work holds 10·ncell values. */
int QN(bo_step_grav)(const bo_mesh* m, const bo_poisson* P, BT* U, double dt, BT* work, int floor64, const BT* phi) {
    int64_t n = m->ncell;
    BT* u0 = work;
    BT* acc = work + 5 * n;
    memcpy(u0, U, sizeof(BT) * 5 * n);
    BT half = (BT)0.5;
    for (int stage = 0; stage < 2; ++stage) {
        QCAT(bo_accumulate, BSUF)(m, U, acc, floor64, NULL, 0);
        QN(bo_grav_source)(m, P, phi, U, acc);
        for (int b = 0; b < m->nblocks; ++b) {
            BT invV = (BT)dt / ((BT)1.0 * (BT)bo_volume(m, b));
            for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c)
                for (int f = 0; f < 5; ++f) {
                    int64_t i = f * n + c;
                    if (stage == 0) U[i] = u0[i] - invV * acc[i];
                    else U[i] = half * u0[i] + half * (U[i] - invV * acc[i]);
                }
        }
    }
    return 0;
}

/* compute_dt over every leaf with the free-fall limiter √(4πGρ)/0.5. */
double QN(bo_compute_dt_grav)(const bo_mesh* m, const BT* U, double cfl, int floor64, double G) {
    int64_t n = m->ncell;
    BT gT = (BT)m->gamma;
    double invdt = 0.0, ginv = 0.0;
    const double fourpi = 4.0 * 3.141592653589793;
    for (int b = 0; b < m->nblocks; ++b)
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c) {
            BT Uc[5], W[5];
            for (int f = 0; f < 5; ++f) Uc[f] = U[f * n + c];
            QCAT(c2p, BSUF)(Uc, gT, m->gamma, floor64, W);
            BT cs = BSQRT(gT * W[4] / W[0]);
            for (int d = 0; d < m->rank; ++d)
                invdt = vo_jmax_f64_(invdt, (double)(BFABS(W[d + 1]) + cs) / bo_width(m, b, d));
            double rho = (double)W[0];
            ginv = vo_jmax_f64_(ginv, rho > 0.0 ? sqrt(fourpi * G * rho) / 0.5 : 0.0);
        }
    invdt = vo_jmax_f64_(invdt, ginv);
    if (invdt == 0.0) return INFINITY;
    return cfl / invdt;
}

#undef QN
#undef QCAT
#undef QCAT2
