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

/* apply_gravity_source! (src/gravity.jl:412-447) on the leaves of a level
 * (level < 0: every leaf). */
static void QN(bo_grav_source)(const bo_mesh* m, const bo_poisson* P, const BT* phi, const BT* U, BT* acc, int level) {
    int64_t n = P->n;
    for (int b = 0; b < m->nblocks; ++b) {
        if (level >= 0 && m->level[b] != level) continue;
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

/* The free-fall limiter √(4πGρ)/0.5 over the leaves of a level (_gravity_invdt,
 * src/gravity.jl:452; compute_dt folds it in, src/driver.jl:186-188). */
static double QN(bo_grav_invdt)(const bo_mesh* m, const BT* U, int level, int floor64, double G) {
    int64_t n = m->ncell;
    BT gT = (BT)m->gamma;
    double ginv = 0.0;
    const double fourpi = 4.0 * 3.141592653589793;
    for (int b = 0; b < m->nblocks; ++b) {
        if (level >= 0 && m->level[b] != level) continue;
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c) {
            BT Uc[5], W[5];
            for (int f = 0; f < 5; ++f) Uc[f] = U[f * n + c];
            QCAT(c2p, BSUF)(Uc, gT, m->gamma, floor64, W);
            double rho = (double)W[0];
            ginv = vo_jmax_f64_(ginv, rho > 0.0 ? sqrt(fourpi * G * rho) / 0.5 : 0.0);
        }
    }
    return ginv;
}

/* The gravity hook of the block stepping: φ held over a (root) step; root()
 * re-solves the Poisson equation warm-started from φ. */
typedef struct {
    const bo_mesh* m;
    const bo_poisson* P;
    BT* phi;
    double G, tol;
    int64_t maxiter;
    int project_mean;
    BT* work;          /* 4n: the CG vectors */
    int64_t solves, iters;
    double worst_relres;
} QN(bo_gctx);

static void QN(bo_ghook_add)(const void* ctx, const BT* U, BT* acc, int level) {
    const QN(bo_gctx)* g = (const QN(bo_gctx)*)ctx;
    QN(bo_grav_source)(g->m, g->P, g->phi, U, acc, level);
}

static double QN(bo_ghook_invdt)(const void* ctx, const BT* U, int level, int fl64) {
    const QN(bo_gctx)* g = (const QN(bo_gctx)*)ctx;
    return QN(bo_grav_invdt)(g->m, U, level, fl64, g->G);
}

static void QN(bo_ghook_root)(const void* ctx, const BT* U) {
    QN(bo_gctx)* g = (QN(bo_gctx)*)ctx;
    double rr = 0.0;
    g->iters += QN(bo_poisson_solve)(g->P, U, g->G, g->project_mean, g->maxiter, g->tol, g->phi, g->work, &rr);
    g->solves += 1;
    if (rr > g->worst_relres) g->worst_relres = rr;
}

/* step! with gravity on a block mesh, every leaf (src/driver.jl:333-359): the
 * source after each flux sum, φ held over the step. work holds 10·ncell values. */
int QN(bo_step_grav)(const bo_mesh* m, const bo_poisson* P, BT* U, double dt, BT* work, int floor64, const BT* phi) {
    QN(bo_gctx) g = {m, P, (BT*)phi, 0.0, 0.0, 0, 0, NULL, 0, 0, 0.0};
    QCAT(bo_hook, BSUF) hk = {QN(bo_ghook_add), NULL, NULL, &g};
    QCAT(bo_step_regs, BSUF)(m, U, dt, -1, work, floor64, NULL, NULL, 0, &hk);
    return 0;
}

/* compute_dt over every leaf with the free-fall limiter. */
double QN(bo_compute_dt_grav)(const bo_mesh* m, const BT* U, double cfl, int floor64, double G) {
    QN(bo_gctx) g = {m, NULL, NULL, G, 0.0, 0, 0, NULL, 0, 0, 0.0};
    QCAT(bo_hook, BSUF) hk = {NULL, QN(bo_ghook_invdt), NULL, &g};
    return QCAT(bo_compute_dt_h, BSUF)(m, U, cfl, -1, floor64, &hk);
}

/* evolve! with self-gravity (src/driver.jl:483-508, 581-605): a Poisson solve
 * (warm-started from phi) before each root step, g held over it and, when
 * subcycling, over every fine substep; the source in each stage of the stepped
 * level; the free-fall limiter in compute_dt of each level. Returns the root
 * steps taken; *iters (may be NULL) the CG iterations summed over the solves,
 * *relres (may be NULL) the largest final relative residual. */
int64_t QN(bo_evolve_grav)(const bo_mesh* m, const bo_poisson* P, BT* U, double tfinal, double cfl, int64_t max_steps,
                           int subcycle, int floor64, double G, int project_mean, int64_t maxiter, double tol, BT* phi,
                           double* t_io, int64_t* iters, double* relres) {
    BT* work = (BT*)malloc(sizeof(BT) * 4 * P->n);
    QN(bo_gctx) g = {m, P, phi, G, tol, maxiter, project_mean, work, 0, 0, 0.0};
    QCAT(bo_hook, BSUF) hk = {QN(bo_ghook_add), QN(bo_ghook_invdt), QN(bo_ghook_root), &g};
    int64_t k = QCAT(bo_evolve_h, BSUF)(m, U, tfinal, cfl, max_steps, subcycle, floor64, t_io, &hk);
    free(work);
    if (iters) *iters = g.iters;
    if (relres) *relres = g.worst_relres;
    return k;
}

#undef QN
#undef QCAT
#undef QCAT2
