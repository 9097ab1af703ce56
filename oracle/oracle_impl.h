/* Precision-templated body of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Included twice by vespa_oracle.c with
 *   OT  = field precision Tf (double or float),
 *   OSUF = symbol suffix (f64 / f32),
 *   OSQRT / OFABS = the Tf sqrt and fabs.
 * Geometry/clock (Tg) and reductions (Tr) are always double, which is the
 * reference's decoupled precision split (src/driver.jl:48-52, test/test_sod.jl:64-70).
 *
 * Every arithmetic expression below keeps the reference's evaluation order
 * (Julia folds `a*b*c` and `a+b+c` left to right and never contracts to FMA);
 * the file is compiled with -ffp-contract=off.
 */

#define OCAT2(a, b) a##_##b
#define OCAT(a, b) OCAT2(a, b)
#define ON(name) OCAT(name, OSUF)

/* Julia min/max on IEEE floats: NaN-propagating, -0.0 < +0.0. An includer
 * that defines OIMPL_OWN_MINMAX supplies ON(jmin)/ON(jmax) itself (the
 * dual-number instantiation, which also fixes which argument's tangent a tie
 * takes). */
#ifndef OIMPL_OWN_MINMAX
static inline OT ON(jmin)(OT x, OT y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y < x) return y;
    if (x < y) return x;
    return signbit(y) ? y : x;
}
static inline OT ON(jmax)(OT x, OT y) {
    if (x != x) return x;
    if (y != y) return y;
    if (y > x) return y;
    if (x > y) return x;
    return signbit(y) ? x : y;
}
#endif

/* src/eos.jl:8-25 cons2prim. floor64 != 0 selects the FP32 form's deviation:
 * the pressure and its positivity floor are evaluated in double from the Tf
 * energy and kinetic terms (the reference floors in Tf, src/eos.jl:24). */
static inline void ON(c2p)(const OT U[5], OT g, double gd, int floor64, OT W[5]) {
    OT rho = U[0];
    OT inv = (OT)1 / rho;
    OT vx = U[1] * inv;
    OT vy = U[2] * inv;
    OT vz = U[3] * inv;
    OT kinetic = ((OT)0.5 * rho) * ((vx * vx + vy * vy) + vz * vz);
    OT p;
    if (floor64) {
        double pd = (gd - 1.0) * ((double)U[4] - (double)kinetic);
        double fl = 1e-12 * (double)kinetic;
        p = (OT)vo_jmax_f64_(pd, fl);
    } else {
        p = (g - (OT)1) * (U[4] - kinetic);
        p = ON(jmax)(p, (OT)1e-12 * kinetic);
    }
    W[0] = rho; W[1] = vx; W[2] = vy; W[3] = vz; W[4] = p;
}

/* src/eos.jl:28-32 prim2cons: E = p/(γ-1) + (½ρ)·((vx²+vy²)+vz²) (a division). */
static inline OT ON(energy)(const OT W[5], OT g) {
    return W[4] / (g - (OT)1) +
           ((OT)0.5 * W[0]) * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]);
}
static inline void ON(p2c)(const OT W[5], OT g, OT U[5]) {
    OT E = ON(energy)(W, g);
    U[0] = W[0]; U[1] = W[0] * W[1]; U[2] = W[0] * W[2]; U[3] = W[0] * W[3]; U[4] = E;
}

/* src/reconstruct.jl:6-12 minmod (ties → b). */
static inline OT ON(minmod)(OT a, OT b) {
    if (a * b <= 0) return (OT)0;
    return OFABS(a) < OFABS(b) ? a : b;
}

/* src/riemann.jl:9-20 euler_flux; dim is 1-based (1..3). */
static inline void ON(eflux)(const OT W[5], OT g, int dim, OT F[5]) {
    OT rho = W[0], p = W[4];
    OT un = W[dim];
    OT E = ON(energy)(W, g);
    OT m1 = rho * W[1], m2 = rho * W[2], m3 = rho * W[3];
    OT z = (OT)0;
    F[0] = rho * un;
    F[1] = m1 * un + (dim == 1 ? p : z);
    F[2] = m2 * un + (dim == 2 ? p : z);
    F[3] = m3 * un + (dim == 3 ? p : z);
    F[4] = (E + p) * un;
}

/* src/riemann.jl:69-88 _hllc_star. */
static inline void ON(hstar)(const OT W[5], const OT U[5], const OT F[5], OT SK, OT Ss,
                             int dim, OT out[5]) {
    OT rho = W[0];
    OT un = W[dim];
    OT p = W[4];
    OT E = U[4];
    OT factor = rho * (SK - un) / (SK - Ss);
    OT vs[3] = {W[1], W[2], W[3]};
    vs[dim - 1] = Ss;
    OT Es = E / rho + (Ss - un) * (Ss + p / (rho * (SK - un)));
    OT Us[5] = {factor, factor * vs[0], factor * vs[1], factor * vs[2], factor * Es};
    for (int i = 0; i < 5; ++i) out[i] = F[i] + SK * (Us[i] - U[i]);
}

/* src/riemann.jl:33-66 hllc_flux. */
static void ON(hllc)(const OT WL[5], const OT WR[5], OT g, int dim, OT out[5]) {
    OT rL = WL[0], rR = WR[0], pL = WL[4], pR = WR[4];
    OT unL = WL[dim], unR = WR[dim];
    OT cL = OSQRT(g * pL / rL);
    OT cR = OSQRT(g * pR / rR);
    OT SL = ON(jmin)(unL - cL, unR - cR);
    OT SR = ON(jmax)(unL + cL, unR + cR);
    OT UL[5], UR[5], FL[5], FR[5];
    ON(p2c)(WL, g, UL);
    ON(p2c)(WR, g, UR);
    ON(eflux)(WL, g, dim, FL);
    ON(eflux)(WR, g, dim, FR);
    if (SL >= 0) { memcpy(out, FL, sizeof(OT) * 5); return; }
    if (SR <= 0) { memcpy(out, FR, sizeof(OT) * 5); return; }
    OT Ss = (pR - pL + rL * unL * (SL - unL) - rR * unR * (SR - unR)) /
            (rL * (SL - unL) - rR * (SR - unR));
    if (Ss >= 0) ON(hstar)(WL, UL, FL, SL, Ss, dim, out);
    else ON(hstar)(WR, UR, FR, SR, Ss, dim, out);
}

/* ---- exported scalar kernels ---- */
void ON(vo_cons2prim)(const OT* U, double gamma, int floor64, OT* W) {
    ON(c2p)(U, (OT)gamma, gamma, floor64, W);
}
void ON(vo_prim2cons)(const OT* W, double gamma, OT* U) { ON(p2c)(W, (OT)gamma, U); }
OT ON(vo_minmod)(OT a, OT b) { return ON(minmod)(a, b); }
void ON(vo_euler_flux)(const OT* W, double gamma, int dim, OT* F) { ON(eflux)(W, (OT)gamma, dim, F); }
void ON(vo_hllc_flux)(const OT* WL, const OT* WR, double gamma, int dim, OT* F) {
    ON(hllc)(WL, WR, (OT)gamma, dim, F);
}

/* src/ka_flux.jl:48-53 compute_face_fluxes!(::HostBatchedFlux, …): AoS [nf][5]. */
void ON(vo_face_flux)(const OT* WL, const OT* WR, const int32_t* axis, int64_t nf, double gamma,
                      OT* F) {
    OT g = (OT)gamma;
    for (int64_t f = 0; f < nf; ++f) ON(hllc)(WL + 5 * f, WR + 5 * f, g, axis[f], F + 5 * f);
}

/* ---- uniform mesh (lib/RefMesh/src/RefMesh.jl) ---- */

/* One cell's conserved tuple from the SoA store (field-major, x fastest). */
static inline void ON(getU)(const OT* U, int64_t n, int64_t c, OT out[5]) {
    for (int f = 0; f < 5; ++f) out[f] = U[f * n + c];
}

/* src/driver.jl:117-122 ghost_state: Outflow copies, Reflecting flips the primitive
 * normal velocity (momentum_indices(model)[axis]). axis is 0-based here. */
static inline void ON(ghost)(const OT W[5], int bc, int axis, OT out[5]) {
    for (int f = 0; f < 5; ++f) out[f] = W[f];
    if (bc == VO_BC_REFLECTING) out[axis + 1] = -W[axis + 1];
}

/* _boundary_ghost (src/driver.jl:127-134) at the face on `side` of `axis` of
 * boundary cell c, given the state W the synthesizing rules start from (the
 * cell's primitive for a slope, its face value for a face): ParentGhost ignores
 * W and returns the closure's state for (axis, side, c) (OPG: its element
 * type); the other rules are ghost_state. */
static inline void ON(bghost)(const vo_grid* m, const OT W[5], int axis, int side, int64_t c, OT out[5]) {
    const int bc = m->bc[2 * axis + side];
    if (bc == VO_BC_PARENT) {
        const OPG* p = (const OPG*)m->pg + 5 * vo_bface_slot(m, c, axis, side);
        for (int f = 0; f < 5; ++f) out[f] = (OT)p[f];
        return;
    }
    ON(ghost)(W, bc, axis, out);
}

/* src/driver.jl:145-159 _plm_slope + src/driver.jl:227-233 _face_value.
 * side: 0 = :lo (h = -½), 1 = :hi (h = +½). */
static void ON(face_value)(const vo_grid* m, const OT* U, int64_t c, int axis, int side, int fl64,
                           OT out[5]) {
    int64_t n = vo_ncell(m);
    OT g = (OT)m->gamma;
    OT Uc[5], Wc[5], Wl[5], Wr[5], Un[5];
    ON(getU)(U, n, c, Uc);
    ON(c2p)(Uc, g, m->gamma, fl64, Wc);
    int64_t nb;
    if (vo_neighbor(m, c, axis, 0, &nb)) {
        ON(getU)(U, n, nb, Un);
        ON(c2p)(Un, g, m->gamma, fl64, Wl);
    } else {
        ON(bghost)(m, Wc, axis, 0, c, Wl);
    }
    if (vo_neighbor(m, c, axis, 1, &nb)) {
        ON(getU)(U, n, nb, Un);
        ON(c2p)(Un, g, m->gamma, fl64, Wr);
    } else {
        ON(bghost)(m, Wc, axis, 1, c, Wr);
    }
    OT h = side ? (OT)0.5 : (OT)-0.5;
    for (int f = 0; f < 5; ++f) {
        OT s = ON(minmod)(Wc[f] - Wl[f], Wr[f] - Wc[f]);
        out[f] = Wc[f] + h * s;
    }
}

static inline void ON(acc_add)(OT* acc, int64_t n, int64_t c, const OT F[5], OT s) {
    for (int f = 0; f < 5; ++f) acc[f * n + c] += s * F[f];
}

/* src/driver.jl:209-281 accumulate_flux! on RefMesh's for_each_face
 * (lib/RefMesh/src/RefMesh.jl:187-197): cells column-major, axes in order, the
 * hi face first, then the lo face iff it is a domain boundary. Optional capture of
 * the boundary-face fluxes (src/reflux.jl:125-132, scale s): bflux[(2*axis+side)*plane
 * + boundary cell index within the face plane][5]. */
static void ON(accumulate)(const vo_grid* m, const OT* U, OT* acc, int fl64, OT* bflux, OT bscale) {
    int64_t n = vo_ncell(m);
    OT g = (OT)m->gamma;
    for (int64_t i = 0; i < 5 * n; ++i) acc[i] = (OT)0;
    for (int64_t c = 0; c < n; ++c) {
        for (int d = 0; d < m->rank; ++d) {
            OT aT = (OT)vo_face_area(m, d);
            OT WL[5], WR[5], Wg[5], F[5];
            int64_t j;
            ON(face_value)(m, U, c, d, 1, fl64, WL);
            if (vo_neighbor(m, c, d, 1, &j)) {
                ON(face_value)(m, U, j, d, 0, fl64, WR);
                ON(hllc)(WL, WR, g, d + 1, F);
                ON(acc_add)(acc, n, c, F, aT);
                ON(acc_add)(acc, n, j, F, -aT);
            } else {
                ON(bghost)(m, WL, d, 1, c, Wg);
                ON(hllc)(WL, Wg, g, d + 1, F);
                ON(acc_add)(acc, n, c, F, aT);
                if (bflux) {
                    OT* b = bflux + 5 * vo_bface_slot(m, c, d, 1);
                    for (int f = 0; f < 5; ++f) b[f] += (bscale * F[f]) * aT;
                }
            }
            if (!vo_neighbor(m, c, d, 0, &j)) {
                ON(face_value)(m, U, c, d, 0, fl64, WR);
                ON(bghost)(m, WR, d, 0, c, Wg);
                ON(hllc)(Wg, WR, g, d + 1, F);
                if (bflux) {
                    OT* b = bflux + 5 * vo_bface_slot(m, c, d, 0);
                    for (int f = 0; f < 5; ++f) b[f] += (bscale * F[f]) * aT;
                }
                ON(acc_add)(acc, n, c, F, -aT);
            }
        }
    }
}

int ON(vo_accumulate_flux)(const vo_grid* m, const OT* U, OT* acc, int floor64) {
    if (vo_check(m)) return -1;
    ON(accumulate)(m, U, acc, floor64, NULL, (OT)0);
    return 0;
}

/* src/driver.jl:333-359 step! (SSP-RK2, aexp = 1): u0 ← U; acc ← L(U);
 * U ← u0 − invV·acc (src/driver.jl:297-307); acc ← L(U);
 * U ← ½u0 + ½(U − invV·acc) (src/driver.jl:310-321), invV = Tf(dt)/(Tf(1)·Tf(V)).
 * work holds 10·ncell values (u0 and acc). bflux (optional) records
 * Σ_stages ½dt·F·area per boundary face (src/driver.jl:344). */
int ON(vo_step)(const vo_grid* m, OT* U, double dt, OT* work, int floor64, OT* bflux) {
    if (vo_check(m)) return -1;
    int64_t n = vo_ncell(m);
    OT* u0 = work;
    OT* acc = work + 5 * n;
    OT invV = (OT)dt / ((OT)1.0 * (OT)vo_cell_volume(m));
    OT half = (OT)0.5;
    OT bs = (OT)(0.5 * dt);
    memcpy(u0, U, sizeof(OT) * 5 * n);
    ON(accumulate)(m, U, acc, floor64, bflux, bs);
    for (int64_t i = 0; i < 5 * n; ++i) U[i] = u0[i] - invV * acc[i];
    ON(accumulate)(m, U, acc, floor64, bflux, bs);
    for (int64_t i = 0; i < 5 * n; ++i) U[i] = half * u0[i] + half * (U[i] - invV * acc[i]);
    return 0;
}

/* src/driver.jl:169-199 compute_dt: invdt = max over cells and axes of
 * Tr(|v_d| + c)/Tr(w_d), Tr = Float64; dt = cfl/invdt; +Inf when invdt == 0. */
double ON(vo_compute_dt)(const vo_grid* m, const OT* U, double cfl, int floor64) {
    if (vo_check(m)) return NAN;
    int64_t n = vo_ncell(m);
    OT g = (OT)m->gamma;
    double invdt = 0.0;
    for (int64_t c = 0; c < n; ++c) {
        OT Uc[5], W[5];
        ON(getU)(U, n, c, Uc);
        ON(c2p)(Uc, g, m->gamma, floor64, W);
        OT cs = OSQRT(g * W[4] / W[0]);
        for (int d = 0; d < m->rank; ++d) {
            double w = vo_cell_width(m, d);
            invdt = vo_jmax_f64_(invdt, (double)(OFABS(W[d + 1]) + cs) / w);
        }
    }
    if (invdt == 0.0) return INFINITY;
    return cfl / invdt;
}

/* src/diagnostics.jl:16-31 conserved_totals: Σ U·V in Float64, cell order. */
void ON(vo_totals)(const vo_grid* m, const OT* U, double out[5]) {
    int64_t n = vo_ncell(m);
    double v = vo_cell_volume(m);
    for (int f = 0; f < 5; ++f) out[f] = 0.0;
    for (int64_t c = 0; c < n; ++c)
        for (int f = 0; f < 5; ++f) out[f] += (double)U[f * n + c] * v;
}

/* src/driver.jl:483-508 evolve! (no policy/gravity/cosmology): while
 * t < tfinal·(1−1e-12): dt = min(compute_dt, tfinal − t); step!; t += dt.
 * max_steps <= 0 means unbounded. Returns the number of steps taken. */
int64_t ON(vo_evolve)(const vo_grid* m, OT* U, double tfinal, double cfl, int64_t max_steps,
                      int floor64, double* t_io) {
    if (vo_check(m)) return -1;
    int64_t n = vo_ncell(m);
    OT* work = (OT*)malloc(sizeof(OT) * 10 * n);
    if (!work) return -1;
    double t = *t_io;
    int64_t k = 0;
    while (t < tfinal * (1 - 1e-12) && (max_steps <= 0 || k < max_steps)) {
        double dt = vo_jmin_f64_(ON(vo_compute_dt)(m, U, cfl, floor64), tfinal - t);
        ON(vo_step)(m, U, dt, work, floor64, NULL);
        t += dt;
        ++k;
    }
    free(work);
    *t_io = t;
    return k;
}

/* Fixed step count with the CFL dt recomputed before each step (the benchmark
 * loop: no tfinal clip). dts (optional) receives each step's dt. */
int ON(vo_run_steps)(const vo_grid* m, OT* U, int64_t nsteps, double cfl, int floor64,
                     double* dts) {
    if (vo_check(m)) return -1;
    int64_t n = vo_ncell(m);
    OT* work = (OT*)malloc(sizeof(OT) * 10 * n);
    if (!work) return -1;
    for (int64_t k = 0; k < nsteps; ++k) {
        double dt = ON(vo_compute_dt)(m, U, cfl, floor64);
        if (dts) dts[k] = dt;
        ON(vo_step)(m, U, dt, work, floor64, NULL);
    }
    free(work);
    return 0;
}

/* ---- gather form (parallel over cells) ----
 * Each cell computes its own 2·rank face fluxes and adds them in the order the
 * literal face walk above produces for that cell:
 *   1. lo faces whose partner precedes the cell (non-wrapped interior), z, y, x;
 *   2. per axis: the hi face, then the lo face if it is a domain boundary;
 *   3. lo faces wrapped by a periodic boundary, x, y, z.
 * Equality with the literal walk is checked bit for bit by the test suite; this
 * form exists to time the restatement on all host cores and to check large grids. */
static void ON(face_flux_of)(const vo_grid* m, const OT* U, int64_t c, int d, int side, int fl64,
                             OT F[5], int* kind) {
    OT g = (OT)m->gamma;
    OT WL[5], WR[5], Wg[5];
    int64_t j;
    int inner = vo_neighbor(m, c, d, side, &j);
    if (side) {
        *kind = 0;
        ON(face_value)(m, U, c, d, 1, fl64, WL);
        if (inner) {
            ON(face_value)(m, U, j, d, 0, fl64, WR);
            ON(hllc)(WL, WR, g, d + 1, F);
        } else {
            ON(bghost)(m, WL, d, 1, c, Wg);
            ON(hllc)(WL, Wg, g, d + 1, F);
        }
    } else {
        ON(face_value)(m, U, c, d, 0, fl64, WR);
        if (inner) {
            *kind = vo_coord(m, c, d) == 0 ? 2 : 0;
            ON(face_value)(m, U, j, d, 1, fl64, WL);
            ON(hllc)(WL, WR, g, d + 1, F);
        } else {
            *kind = 1;
            ON(bghost)(m, WR, d, 0, c, Wg);
            ON(hllc)(Wg, WR, g, d + 1, F);
        }
    }
}

static void ON(accumulate_gather)(const vo_grid* m, const OT* U, OT* acc, int fl64) {
    int64_t n = vo_ncell(m);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; ++c) {
        OT Flo[3][5], Fhi[3][5];
        int klo[3], khi;
        for (int d = 0; d < m->rank; ++d) {
            ON(face_flux_of)(m, U, c, d, 0, fl64, Flo[d], &klo[d]);
            ON(face_flux_of)(m, U, c, d, 1, fl64, Fhi[d], &khi);
        }
        OT a[5] = {0, 0, 0, 0, 0};
        for (int d = m->rank - 1; d >= 0; --d) {
            if (klo[d] != 0) continue;
            OT s = -(OT)vo_face_area(m, d);
            for (int f = 0; f < 5; ++f) a[f] += s * Flo[d][f];
        }
        for (int d = 0; d < m->rank; ++d) {
            OT s = (OT)vo_face_area(m, d);
            for (int f = 0; f < 5; ++f) a[f] += s * Fhi[d][f];
            if (klo[d] == 1)
                for (int f = 0; f < 5; ++f) a[f] += (-s) * Flo[d][f];
        }
        for (int d = 0; d < m->rank; ++d) {
            if (klo[d] != 2) continue;
            OT s = -(OT)vo_face_area(m, d);
            for (int f = 0; f < 5; ++f) a[f] += s * Flo[d][f];
        }
        for (int f = 0; f < 5; ++f) acc[f * n + c] = a[f];
    }
}

int ON(vo_accumulate_flux_gather)(const vo_grid* m, const OT* U, OT* acc, int floor64) {
    if (vo_check(m)) return -1;
    ON(accumulate_gather)(m, U, acc, floor64);
    return 0;
}

/* step! in gather form, parallel over cells with OpenMP (threads <= 0: runtime default). */
int ON(vo_step_omp)(const vo_grid* m, OT* U, double dt, OT* work, int floor64, int threads) {
    if (vo_check(m)) return -1;
    if (threads > 0) omp_set_num_threads(threads);
    int64_t n = vo_ncell(m);
    OT* u0 = work;
    OT* acc = work + 5 * n;
    OT invV = (OT)dt / ((OT)1.0 * (OT)vo_cell_volume(m));
    OT half = (OT)0.5;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 5 * n; ++i) u0[i] = U[i];
    ON(accumulate_gather)(m, U, acc, floor64);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 5 * n; ++i) U[i] = u0[i] - invV * acc[i];
    ON(accumulate_gather)(m, U, acc, floor64);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 5 * n; ++i) U[i] = half * u0[i] + half * (U[i] - invV * acc[i]);
    return 0;
}

double ON(vo_compute_dt_omp)(const vo_grid* m, const OT* U, double cfl, int floor64, int threads) {
    if (vo_check(m)) return NAN;
    if (threads > 0) omp_set_num_threads(threads);
    int64_t n = vo_ncell(m);
    OT g = (OT)m->gamma;
    double invdt = 0.0;
#pragma omp parallel for schedule(static) reduction(max : invdt)
    for (int64_t c = 0; c < n; ++c) {
        OT Uc[5], W[5];
        ON(getU)(U, n, c, Uc);
        ON(c2p)(Uc, g, m->gamma, floor64, W);
        OT cs = OSQRT(g * W[4] / W[0]);
        for (int d = 0; d < m->rank; ++d) {
            double v = (double)(OFABS(W[d + 1]) + cs) / vo_cell_width(m, d);
            if (v > invdt) invdt = v;
        }
    }
    if (invdt == 0.0) return INFINITY;
    return cfl / invdt;
}

#undef ON
#undef OCAT
#undef OCAT2
