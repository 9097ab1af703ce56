/* Precision-templated body of the block-mesh oracle (TEST INFRASTRUCTURE ONLY).
 * Included by blocks_oracle.c with BT / BSUF / BSQRT / BFABS; uses the scalar
 * kernels of oracle_impl.h (c2p, hllc, ghost, minmod) of the same precision. */

#define BCAT2(a, b) a##_##b
#define BCAT(a, b) BCAT2(a, b)
#define BN(name) BCAT(name, BSUF)

/* Flux register (src/reflux.jl:51-103): dense per-cell accumulator with a
 * presence flag (the reference keys a Dict by coarse leaf). */
typedef struct {
    BT* delta;       /* 5 * ncell */
    uint8_t* has;    /* ncell */
    int32_t coarse_level;
    BT scale;
} BN(bo_reg);

/* A source term hooked into the stages (self-gravity, src/driver.jl:347-375):
 * add() after each flux sum over the leaves of the stepped level (level < 0:
 * every leaf); invdt() is the limiter folded into compute_dt of that level;
 * root() runs before each root step of evolve! (the Poisson solve,
 * src/driver.jl:498,592). Any member may be NULL. */
typedef struct {
    void (*add)(const void* ctx, const BT* U, BT* acc, int level);
    double (*invdt)(const void* ctx, const BT* U, int level, int fl64);
    void (*root)(const void* ctx, const BT* U);
    const void* ctx;
} BN(bo_hook);

static void BN(bo_getU)(const BT* U, int64_t n, int64_t c, BT out[5]) {
    for (int f = 0; f < 5; ++f) out[f] = U[f * n + c];
}

/* _face_value (src/driver.jl:227-233) with the representative neighbours. */
static void BN(bo_face_value)(const bo_mesh* m, const BT* U, int64_t g, int d, int side, int fl64, BT out[5]) {
    int64_t n = m->ncell;
    BT gT = (BT)m->gamma;
    BT Uc[5], Wc[5], Wl[5], Wr[5], Un[5];
    BN(bo_getU)(U, n, g, Uc);
    BCAT(c2p, BSUF)(Uc, gT, m->gamma, fl64, Wc);
    int64_t ijk[3];
    int b = bo_locate(m, g, ijk);
    for (int s = 0; s < 2; ++s) {
        int64_t rep;
        BT* W = s ? Wr : Wl;
        int k = bo_neighbor(m, g, d, s, &rep, NULL, NULL);
        if (k == NB_BOUNDARY) {
            BCAT(ghost, BSUF)(Wc, m->bc[6 * b + 2 * d + s], d, W);
        } else {
            BN(bo_getU)(U, n, rep, Un);
            BCAT(c2p, BSUF)(Un, gT, m->gamma, fl64, W);
        }
    }
    BT h = side ? (BT)0.5 : (BT)-0.5;
    for (int f = 0; f < 5; ++f) {
        BT sl = BCAT(minmod, BSUF)(Wc[f] - Wl[f], Wr[f] - Wc[f]);
        out[f] = Wc[f] + h * sl;
    }
}

static int BN(bo_level_of)(const bo_mesh* m, int64_t g) {
    int64_t ijk[3];
    return m->level[bo_locate(m, g, ijk)];
}

/* _reflux_capture! (src/reflux.jl:70-89) for one emitted interior face (i left, j right). */
static void BN(bo_capture)(const bo_mesh* m, BN(bo_reg) * *regs, int nregs, int64_t i, int64_t j, const BT F[5],
                           BT aT) {
    int64_t n = m->ncell;
    int li = BN(bo_level_of)(m, i), lj = BN(bo_level_of)(m, j);
    if (li == lj) return;
    for (int r = 0; r < nregs; ++r) {
        BN(bo_reg)* R = regs[r];
        int cl = R->coarse_level;
        BT s = R->scale;
        if (li == cl && lj == cl + 1) {
            for (int f = 0; f < 5; ++f) R->delta[f * n + i] = R->delta[f * n + i] + (s * F[f]) * aT;
            R->has[i] = 1;
        } else if (lj == cl && li == cl + 1) {
            for (int f = 0; f < 5; ++f) R->delta[f * n + j] = R->delta[f * n + j] + ((-s) * F[f]) * aT;
            R->has[j] = 1;
        }
    }
}

/* accumulate_flux! over HGBackend's face enumeration (see the file header). */
static void BN(bo_accumulate)(const bo_mesh* m, const BT* U, BT* acc, int fl64, BN(bo_reg) * *regs, int nregs) {
    int64_t n = m->ncell;
    BT gT = (BT)m->gamma;
    for (int64_t i = 0; i < 5 * n; ++i) acc[i] = (BT)0;
    for (int64_t g = 0; g < n; ++g) {
        int64_t ijk[3];
        int b = bo_locate(m, g, ijk);
        for (int d = 0; d < m->rank; ++d) {
            BT aT = (BT)bo_area(m, b, d);
            BT WL[5], WR[5], Wg[5], F[5];
            int64_t rep;
            int k = bo_neighbor(m, g, d, 1, &rep, NULL, NULL);
            if (k == NB_BOUNDARY) {
                BN(bo_face_value)(m, U, g, d, 1, fl64, WL);
                BCAT(ghost, BSUF)(WL, m->bc[6 * b + 2 * d + 1], d, Wg);
                BCAT(hllc, BSUF)(WL, Wg, gT, d + 1, F);
                for (int f = 0; f < 5; ++f) acc[f * n + g] += aT * F[f];
            } else if (k == NB_SAME || k == NB_COARSER) {
                BN(bo_face_value)(m, U, g, d, 1, fl64, WL);
                BN(bo_face_value)(m, U, rep, d, 0, fl64, WR);
                BCAT(hllc, BSUF)(WL, WR, gT, d + 1, F);
                for (int f = 0; f < 5; ++f) acc[f * n + g] += aT * F[f];
                for (int f = 0; f < 5; ++f) acc[f * n + rep] += (-aT) * F[f];
                if (nregs) BN(bo_capture)(m, regs, nregs, g, rep, F, aT);
            }
            k = bo_neighbor(m, g, d, 0, &rep, NULL, NULL);
            if (k == NB_BOUNDARY) {
                BN(bo_face_value)(m, U, g, d, 0, fl64, WR);
                BCAT(ghost, BSUF)(WR, m->bc[6 * b + 2 * d], d, Wg);
                BCAT(hllc, BSUF)(Wg, WR, gT, d + 1, F);
                for (int f = 0; f < 5; ++f) acc[f * n + g] += (-aT) * F[f];
            } else if (k == NB_COARSER) {
                BN(bo_face_value)(m, U, rep, d, 1, fl64, WL);
                BN(bo_face_value)(m, U, g, d, 0, fl64, WR);
                BCAT(hllc, BSUF)(WL, WR, gT, d + 1, F);
                for (int f = 0; f < 5; ++f) acc[f * n + rep] += aT * F[f];
                for (int f = 0; f < 5; ++f) acc[f * n + g] += (-aT) * F[f];
                if (nregs) BN(bo_capture)(m, regs, nregs, rep, g, F, aT);
            }
        }
    }
}

int BN(bo_accumulate_flux)(const bo_mesh* m, const BT* U, BT* acc, int floor64) {
    BN(bo_accumulate)(m, U, acc, floor64, NULL, 0);
    return 0;
}

static int BN(bo_in_level)(const bo_mesh* m, int b, int level) { return level < 0 || m->level[b] == level; }

/* step_level! / step!(sim, dt; level) (src/driver.jl:333-378): level < 0 updates
 * every leaf. work: 10*ncell values. */
static void BN(bo_step_regs)(const bo_mesh* m, BT* U, double dt, int level, BT* work, int fl64, BN(bo_reg) * *regs,
                             const double* signs, int nregs, const BN(bo_hook) * hk) {
    int64_t n = m->ncell;
    BT* u0 = work;
    BT* acc = work + 5 * n;
    for (int b = 0; b < m->nblocks; ++b) {
        if (!BN(bo_in_level)(m, b, level)) continue;
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c)
            for (int f = 0; f < 5; ++f) u0[f * n + c] = U[f * n + c];
    }
    BT half = (BT)0.5;
    for (int stage = 0; stage < 2; ++stage) {
        for (int r = 0; r < nregs; ++r) regs[r]->scale = (BT)(signs[r] * (0.5 * dt));
        BN(bo_accumulate)(m, U, acc, fl64, regs, nregs);
        if (hk && hk->add) hk->add(hk->ctx, U, acc, level);
        for (int b = 0; b < m->nblocks; ++b) {
            if (!BN(bo_in_level)(m, b, level)) continue;
            BT invV = (BT)dt / ((BT)1.0 * (BT)bo_volume(m, b));
            for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c)
                for (int f = 0; f < 5; ++f) {
                    int64_t i = f * n + c;
                    if (stage == 0) U[i] = u0[i] - invV * acc[i];
                    else U[i] = half * u0[i] + half * (U[i] - invV * acc[i]);
                }
        }
    }
}

int BN(bo_step)(const bo_mesh* m, BT* U, double dt, int level, BT* work, int floor64) {
    BN(bo_step_regs)(m, U, dt, level, work, floor64, NULL, NULL, 0, NULL);
    return 0;
}

/* compute_dt(sim; level) (src/driver.jl:169-199), the hook's limiter folded in
 * after the wave-speed maximum. */
static double BN(bo_compute_dt_h)(const bo_mesh* m, const BT* U, double cfl, int level, int floor64,
                                  const BN(bo_hook) * hk) {
    int64_t n = m->ncell;
    BT gT = (BT)m->gamma;
    double invdt = 0.0;
    for (int b = 0; b < m->nblocks; ++b) {
        if (!BN(bo_in_level)(m, b, level)) continue;
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c) {
            BT Uc[5], W[5];
            BN(bo_getU)(U, n, c, Uc);
            BCAT(c2p, BSUF)(Uc, gT, m->gamma, floor64, W);
            BT cs = BSQRT(gT * W[4] / W[0]);
            for (int d = 0; d < m->rank; ++d)
                invdt = vo_jmax_f64_(invdt, (double)(BFABS(W[d + 1]) + cs) / bo_width(m, b, d));
        }
    }
    if (hk && hk->invdt) invdt = vo_jmax_f64_(invdt, hk->invdt(hk->ctx, U, level, floor64));
    if (invdt == 0.0) return INFINITY;
    return cfl / invdt;
}

double BN(bo_compute_dt)(const bo_mesh* m, const BT* U, double cfl, int level, int floor64) {
    return BN(bo_compute_dt_h)(m, U, cfl, level, floor64, NULL);
}

void BN(bo_totals)(const bo_mesh* m, const BT* U, double out[5]) {
    int64_t n = m->ncell;
    for (int f = 0; f < 5; ++f) out[f] = 0.0;
    for (int b = 0; b < m->nblocks; ++b) {
        double v = bo_volume(m, b);
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c)
            for (int f = 0; f < 5; ++f) out[f] += (double)U[f * n + c] * v;
    }
}

/* density_gradient_indicator (src/driver.jl:400-414): per leaf, the largest
 * |ρn − ρc| / max(ρc, ρn, eps()) over the representative neighbours across the
 * 2·rank faces that are Interior (a domain boundary is skipped; a periodic side
 * wraps inside the block). The difference is taken in BT and the ratio in double,
 * as Julia promotes a Float32 difference against eps()::Float64. block_max[b] is
 * the largest value over the leaves of block b; cell_g may be NULL. */
void BN(bo_indicator)(const bo_mesh* m, const BT* U, double* cell_g, double* block_max) {
    for (int b = 0; b < m->nblocks; ++b) {
        double gb = 0.0;
        for (int64_t c = m->offset[b]; c < m->offset[b + 1]; ++c) {
            BT rc = U[c];
            double g = 0.0;
            for (int d = 0; d < m->rank; ++d)
                for (int side = 0; side < 2; ++side) {
                    int64_t rep;
                    if (bo_neighbor(m, c, d, side, &rep, NULL, NULL) == NB_BOUNDARY) continue;
                    BT rn = U[rep];
                    double num = (double)BFABS(rn - rc);
                    double den = vo_jmax_f64_(vo_jmax_f64_((double)rc, (double)rn), DBL_EPSILON);
                    g = vo_jmax_f64_(g, num / den);
                }
            if (cell_g) cell_g[c] = g;
            gb = vo_jmax_f64_(gb, g);
        }
        block_max[b] = gb;
    }
}

static int BN(bo_max_level)(const bo_mesh* m) {
    int l = 0;
    for (int b = 0; b < m->nblocks; ++b)
        if (m->level[b] > l) l = m->level[b];
    return l;
}

/* _reflux_apply! (src/reflux.jl:94-103). */
static void BN(bo_reflux_apply)(const bo_mesh* m, BT* U, BN(bo_reg) * R) {
    int64_t n = m->ncell;
    for (int64_t c = 0; c < n; ++c) {
        if (!R->has[c]) continue;
        int64_t ijk[3];
        int b = bo_locate(m, c, ijk);
        BT invV = (BT)1 / (BT)bo_volume(m, b);
        for (int f = 0; f < 5; ++f) U[f * n + c] = U[f * n + c] - invV * R->delta[f * n + c];
        for (int f = 0; f < 5; ++f) R->delta[f * n + c] = (BT)0;
        R->has[c] = 0;
    }
}

/* evolve_level! (src/driver.jl:553-578). */
static void BN(bo_evolve_level)(const bo_mesh* m, BT* U, BT* work, int level, double dt_target, double cfl, int fl64,
                                BN(bo_reg) * parent, const BN(bo_hook) * hk) {
    double dt_stable = BN(bo_compute_dt_h)(m, U, cfl, level, fl64, hk);
    if (!isfinite(dt_stable)) return;
    double q = dt_target / dt_stable * (1 - 1e-12);
    int64_t ns = (int64_t)ceil(q);
    if (ns < 1) ns = 1;
    double dt_sub = dt_target / (double)ns;
    int has_finer = level < BN(bo_max_level)(m);
    BN(bo_reg) own = {0};
    if (has_finer) {
        own.delta = (BT*)calloc(5 * m->ncell, sizeof(BT));
        own.has = (uint8_t*)calloc(m->ncell, 1);
        own.coarse_level = level;
    }
    BN(bo_reg)* regs[2];
    double signs[2];
    int nregs = 0;
    if (parent) { regs[nregs] = parent; signs[nregs++] = 1.0; }
    if (has_finer) { regs[nregs] = &own; signs[nregs++] = -1.0; }
    for (int64_t s = 0; s < ns; ++s) {
        if (has_finer) {
            memset(own.delta, 0, sizeof(BT) * 5 * m->ncell);
            memset(own.has, 0, m->ncell);
        }
        BN(bo_step_regs)(m, U, dt_sub, level, work, fl64, regs, signs, nregs, hk);
        if (has_finer) {
            BN(bo_evolve_level)(m, U, work, level + 1, dt_sub, cfl, fl64, &own, hk);
            BN(bo_reflux_apply)(m, U, &own);
        }
    }
    free(own.delta);
    free(own.has);
}

/* evolve! (src/driver.jl:483-508, 581-605): single rate (global dt over every
 * leaf) or subcycled. Returns the number of root steps taken. */
static int64_t BN(bo_evolve_h)(const bo_mesh* m, BT* U, double tfinal, double cfl, int64_t max_steps, int subcycle,
                               int fl64, double* t_io, const BN(bo_hook) * hk) {
    BT* work = (BT*)malloc(sizeof(BT) * 10 * m->ncell);
    double t = *t_io;
    int64_t k = 0;
    while (t < tfinal * (1 - 1e-12) && (max_steps <= 0 || k < max_steps)) {
        if (hk && hk->root) hk->root(hk->ctx, U);
        if (subcycle) {
            double dt = vo_jmin_f64_(BN(bo_compute_dt_h)(m, U, cfl, 0, fl64, hk), tfinal - t);
            BN(bo_evolve_level)(m, U, work, 0, dt, cfl, fl64, NULL, hk);
            t += dt;
        } else {
            double dt = vo_jmin_f64_(BN(bo_compute_dt_h)(m, U, cfl, -1, fl64, hk), tfinal - t);
            BN(bo_step_regs)(m, U, dt, -1, work, fl64, NULL, NULL, 0, hk);
            t += dt;
        }
        ++k;
    }
    free(work);
    *t_io = t;
    return k;
}

int64_t BN(bo_evolve)(const bo_mesh* m, BT* U, double tfinal, double cfl, int64_t max_steps, int subcycle, int fl64,
                      double* t_io) {
    return BN(bo_evolve_h)(m, U, tfinal, cfl, max_steps, subcycle, fl64, t_io, NULL);
}

#undef BN
#undef BCAT
#undef BCAT2
