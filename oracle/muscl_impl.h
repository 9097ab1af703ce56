/* Precision-templated body of the MUSCL oracle (TEST INFRASTRUCTURE ONLY).
 * Included by muscl_oracle.c with MT / MSUF / MSQRT / MFABS. Every operation is
 * taken at the state's precision MT, in the reference's evaluation order
 * (Julia never contracts a*b+c); max/min follow Julia (NaN wins). */

#define MCAT2(a, b) a##_##b
#define MCAT(a, b) MCAT2(a, b)
#define MN(name) MCAT(name, MSUF)

static inline MT MN(mu_max)(MT a, MT b) { return (MT)vo_jmax_f64_((double)a, (double)b); }
static inline MT MN(mu_min)(MT a, MT b) { return (MT)vo_jmin_f64_((double)a, (double)b); }

/* Enzo sign(): sign(0) = -1 (lib/PPMKernels/src/muscl.jl:13). */
static inline MT MN(mu_sign)(MT x) { return x > (MT)0 ? (MT)1 : (MT)-1; }

/* 3-argument minmod (lib/PPMKernels/src/muscl.jl:16-19). */
static inline MT MN(mu_minmod3)(MT a, MT b, MT c) {
    MT sa = MN(mu_sign)(a), sb = MN(mu_sign)(b), sc = MN(mu_sign)(c);
    return (MT)0.25 * (sa + sb) * MFABS(sa + sc) * MN(mu_min)(MFABS(a), MN(mu_min)(MFABS(b), MFABS(c)));
}

/* Right-face value of the middle cell, v + ½·minmod((v−vm1)θ, (vp1−v)θ, ½(vp1−vm1))
 * (lib/PPMKernels/src/muscl.jl:24-27). Called with the arguments reversed it gives
 * the left-face value, as the reference does for the R state. */
static inline MT MN(mu_plm_pt)(MT vm1, MT v, MT vp1, MT th) {
    MT s = MN(mu_minmod3)((v - vm1) * th, (vp1 - v) * th, (MT)0.5 * (vp1 - vm1));
    return v + (MT)0.5 * s;
}

static inline MT MN(mu_hll)(MT ap, MT am, MT fl, MT fr, MT ul, MT ur) {
    return (ap * fl + am * fr - ap * am * (ur - ul)) / (ap + am);
}

/* HLL flux of the hydro fields from L/R primitive states (ρ, eint, u, v, w), u the
 * normal velocity (lib/PPMKernels/src/muscl.jl:37-58, the first five of _hll6). */
static void MN(mu_hll5)(MT rl, MT el, MT ul, MT vl, MT wl, MT rr, MT er, MT ur, MT vr, MT wr, MT g, MT gm1,
                        MT F[6]) {
    MT h = (MT)0.5;
    MT v2l = ul * ul + vl * vl + wl * wl, pl = gm1 * rl * el, csl = MSQRT(g * pl / rl), etl = el + h * v2l;
    MT UlD = rl, UlS1 = rl * ul, UlS2 = rl * vl, UlS3 = rl * wl, UlE = rl * etl;
    MT FlD = rl * ul, FlS1 = UlS1 * ul + pl, FlS2 = UlS2 * ul, FlS3 = UlS3 * ul, FlE = (UlE + pl) * ul;
    MT lpl = ul + csl, lml = ul - csl;
    MT v2r = ur * ur + vr * vr + wr * wr, pr = gm1 * rr * er, csr = MSQRT(g * pr / rr), etr = er + h * v2r;
    MT UrD = rr, UrS1 = rr * ur, UrS2 = rr * vr, UrS3 = rr * wr, UrE = rr * etr;
    MT FrD = rr * ur, FrS1 = UrS1 * ur + pr, FrS2 = UrS2 * ur, FrS3 = UrS3 * ur, FrE = (UrE + pr) * ur;
    MT lpr = ur + csr, lmr = ur - csr;
    MT ap = MN(mu_max)((MT)0, MN(mu_max)(lpl, lpr));
    MT am = MN(mu_max)((MT)0, MN(mu_max)(-lml, -lmr));
    F[0] = MN(mu_hll)(ap, am, FlD, FrD, UlD, UrD);
    F[1] = MN(mu_hll)(ap, am, FlS1, FrS1, UlS1, UrS1);
    F[2] = MN(mu_hll)(ap, am, FlS2, FrS2, UlS2, UrS2);
    F[3] = MN(mu_hll)(ap, am, FlS3, FrS3, UlS3, UrS3);
    F[4] = MN(mu_hll)(ap, am, FlE, FrE, UlE, UrE);
    /* the gas-energy flux of _hll6 (dual-energy formalism): U = ρe, F = ρe·u */
    MT UlG = rl * el, UrG = rr * er;
    F[5] = MN(mu_hll)(ap, am, UlG * ul, UrG * ur, UlG, UrG);
}

/* One interface of a pencil with stride st: left cell cl (lib/PPMKernels/src/muscl.jl:181-206). */
static void MN(mu_face)(const MT* rho, const MT* eint, const MT* vu, const MT* vv, const MT* vw, int64_t cl,
                        int64_t st, MT g, MT gm1, MT th, MT sr, MT F[6]) {
    int64_t a = cl - st, b = cl, c = cl + st, d = cl + 2 * st;
    MT rl = MN(mu_max)(MN(mu_plm_pt)(rho[a], rho[b], rho[c], th), sr);
    MT el = MN(mu_plm_pt)(eint[a], eint[b], eint[c], th);
    MT ul = MN(mu_plm_pt)(vu[a], vu[b], vu[c], th);
    MT vl = MN(mu_plm_pt)(vv[a], vv[b], vv[c], th);
    MT wl = MN(mu_plm_pt)(vw[a], vw[b], vw[c], th);
    MT rr = MN(mu_max)(MN(mu_plm_pt)(rho[d], rho[c], rho[b], th), sr);
    MT er = MN(mu_plm_pt)(eint[d], eint[c], eint[b], th);
    MT ur = MN(mu_plm_pt)(vu[d], vu[c], vu[b], th);
    MT vr = MN(mu_plm_pt)(vv[d], vv[c], vv[b], th);
    MT wr = MN(mu_plm_pt)(vw[d], vw[c], vw[b], th);
    MN(mu_hll5)(rl, el, ul, vl, wl, rr, er, ur, vr, wr, g, gm1, F);
}

/* muscl_flux_line! (lib/PPMKernels/src/muscl.jl:223-239): fluxes at the active+1
 * interfaces of jdim contiguous pencils of ncells cells; F arrays hold (active+1)·jdim;
 * fge (may be NULL) receives the gas-energy flux of the dual-energy formalism. */
void MN(vo_muscl_flux_line)(const MT* rho, const MT* eint, const MT* vx, const MT* vy, const MT* vz,
                            int64_t ncells, int64_t nghost, int64_t jdim, double gamma, double theta,
                            double small_rho, MT* fd, MT* fs1, MT* fs2, MT* fs3, MT* fe, MT* fge) {
    MT g = (MT)gamma, gm1 = g - (MT)1, th = (MT)theta, sr = (MT)small_rho;
    int64_t nfi = ncells - 2 * nghost + 1;
    for (int64_t j = 0; j < jdim; ++j)
        for (int64_t gi = 0; gi < nfi; ++gi) {
            MT F[6];
            MN(mu_face)(rho, eint, vx, vy, vz, j * ncells + nghost + gi - 1, 1, g, gm1, th, sr, F);
            int64_t o = j * nfi + gi;
            fd[o] = F[0]; fs1[o] = F[1]; fs2[o] = F[2]; fs3[o] = F[3]; fe[o] = F[4];
            if (fge) fge[o] = F[5];
        }
}

/* _dual_eint (lib/PPMKernels/src/muscl_grid.jl:41-47): the total-energy eint where
 * pressure is dynamically significant, else the advected one; floored. */
static inline MT MN(mu_dual_eint)(MT tau, MT d, MT ge, MT v2, MT g, MT gm1, MT eta1, MT small) {
    MT eint_tot = tau / d - (MT)0.5 * v2;
    MT eint_dual = ge / d;
    MT cs2 = eint_tot > (MT)0 ? g * gm1 * eint_tot : (MT)0;
    int use_tot = (cs2 > eta1 * v2) && (eint_tot > (MT)0.5 * eint_dual);
    return MN(mu_max)(use_tot ? eint_tot : eint_dual, small);
}

/* dU(U) of the unsplit driver (_muscl_L!, lib/PPMKernels/src/muscl_grid.jl:212-295):
 * primitives over the whole grid (eint floored at small_rho), then per axis the HLL
 * fluxes of every pencil (transverse ghosts included) and their divergence on the
 * cells active along that axis, summed into dU as ((0 + x) + y) + z where the y and
 * z terms of a cell not active along that axis are +0. frec (15·N or NULL) receives
 * w·F at the cell whose lo face the flux crosses (grid-order momenta). */
static void MN(mu_L)(const MT* U, int64_t N, const int64_t dims[3], int64_t ng, MT dtdx, MT g, MT th, MT sr,
                     int nf, MT eta1, MT* prim, MT* dU, MT* frec, MT w) {
    MT* rho = prim;
    MT* eint = prim + N;
    MT* vel[3] = {prim + 2 * N, prim + 3 * N, prim + 4 * N};
    MT gm1 = g - (MT)1;
    for (int64_t i = 0; i < N; ++i) {
        MT d = U[i], u = U[N + i] / d, v = U[2 * N + i] / d, ww = U[3 * N + i] / d;
        rho[i] = d;
        vel[0][i] = u;
        vel[1][i] = v;
        vel[2][i] = ww;
        if (nf == 6)
            eint[i] = MN(mu_dual_eint)(U[4 * N + i], d, U[5 * N + i], u * u + v * v + ww * ww, g, gm1, eta1, sr);
        else
            eint[i] = MN(mu_max)(U[4 * N + i] / d - (MT)0.5 * (u * u + v * v + ww * ww), sr);
    }
    for (int64_t i = 0; i < nf * N; ++i) dU[i] = (MT)0;
    const int64_t str[3] = {1, dims[0], dims[0] * dims[1]};
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;   /* cyclic roles: normal a, t1 b, t2 c */
        const int64_t na = dims[a], active = na - 2 * ng, nfi = active + 1;
        const int64_t sa = str[a];
        const int tgt[6] = {0, 1 + a, 1 + b, 1 + c, 4, 5};
        MT* Fl = (MT*)malloc(sizeof(MT) * 6 * (size_t)nfi);
        for (int64_t q = 0; q < dims[b] * dims[c]; ++q) {
            const int64_t ib = q % dims[b], ic = q / dims[b];
            const int64_t base = ib * str[b] + ic * str[c];
            for (int64_t gi = 0; gi < nfi; ++gi)
                MN(mu_face)(rho, eint, vel[a], vel[b], vel[c], base + (ng + gi - 1) * sa, sa, g, gm1, th, sr,
                            Fl + 6 * gi);
            for (int64_t p = 0; p < na; ++p) {
                const int64_t cell = base + p * sa;
                const int act = p >= ng && p < na - ng;
                for (int f = 0; f < nf; ++f) {
                    MT t = (MT)0;
                    if (act) t = -(Fl[6 * (p - ng + 1) + f] - Fl[6 * (p - ng) + f]) * dtdx;
                    if (a == 0) {
                        if (act) dU[tgt[f] * N + cell] += t;
                    } else {
                        dU[tgt[f] * N + cell] += t;
                    }
                    if (frec) {
                        MT s = (MT)0;
                        if (p >= ng && p <= na - ng) s = w * Fl[6 * (p - ng) + f];
                        frec[(int64_t)a * nf * N + tgt[f] * N + cell] += s;
                    }
                }
            }
        }
        free(Fl);
    }
}

/* muscl_step_3d! (lib/PPMKernels/src/muscl_grid.jl:314-370): SSP-RK2, U1 = U0 + dU(U0),
 * Unew = ½((U0 + U1) + dU(U1)), over every cell (ghosts get dU = 0). nf = 5, or 6 with
 * the gas energy Ge = ρe as field 5 (the dual-energy formalism, eta1): eint selected by
 * _dual_eint, Ge advected with the HLL flux ρe·u, and the end-of-step sync
 * (dual_energy_sync!, lib/PPMKernels/src/muscl_grid.jl:97-115). frec (3·nf·N:
 * [axis][field][cell], or NULL) is zeroed and receives ½F per stage. */
/* muscl_step_3d! with a bc! hook: bc(user, U, stage) refills the ghost zones of
 * the nf fields before each stage (stage 0: U holds the step's input; stage 1:
 * the stage-1 state), after the copy U0 of the unfilled input is taken
 * (lib/PPMKernels/src/muscl_grid.jl:316-322, 329, 339, 350). bc == NULL is
 * bc! = nothing. */
void MN(vo_muscl_step_3d_bc)(MT* U, const int64_t dims[3], int64_t ng, double dt, double gamma, double theta,
                             double dx, double small_rho, int nf, double eta1, MT* frec,
                             void (*bc)(void* user, MT* U, int stage), void* user) {
    const int64_t N = dims[0] * dims[1] * dims[2];
    MT* U0 = (MT*)malloc(sizeof(MT) * nf * (size_t)N);
    MT* dU = (MT*)malloc(sizeof(MT) * nf * (size_t)N);
    MT* prim = (MT*)malloc(sizeof(MT) * 5 * (size_t)N);
    memcpy(U0, U, sizeof(MT) * nf * (size_t)N);
    const MT half = (MT)0.5, dtdx = (MT)dt / (MT)dx;
    const MT g = (MT)gamma, e1 = (MT)eta1, sr = (MT)small_rho;
    if (frec)
        for (int64_t i = 0; i < 3 * nf * N; ++i) frec[i] = (MT)0;
    if (bc) bc(user, U, 0);
    MN(mu_L)(U, N, dims, ng, dtdx, g, (MT)theta, sr, nf, e1, prim, dU, frec, half);
    for (int64_t i = 0; i < nf * N; ++i) U[i] = U0[i] + dU[i];
    if (bc) bc(user, U, 1);
    MN(mu_L)(U, N, dims, ng, dtdx, g, (MT)theta, sr, nf, e1, prim, dU, frec, half);
    for (int64_t i = 0; i < nf * N; ++i) U[i] = half * (U0[i] + U[i] + dU[i]);
    if (nf == 6)
        for (int64_t i = 0; i < N; ++i) {
            MT d = U[i], u = U[N + i] / d, v = U[2 * N + i] / d, w = U[3 * N + i] / d;
            MT v2 = u * u + v * v + w * w;
            MT e = MN(mu_dual_eint)(U[4 * N + i], d, U[5 * N + i], v2, g, g - (MT)1, e1, sr);
            U[5 * N + i] = d * e;
            U[4 * N + i] = d * (e + (MT)0.5 * v2);
        }
    free(U0);
    free(dU);
    free(prim);
}

void MN(vo_muscl_step_3d)(MT* U, const int64_t dims[3], int64_t ng, double dt, double gamma, double theta,
                          double dx, double small_rho, int nf, double eta1, MT* frec) {
    MN(vo_muscl_step_3d_bc)(U, dims, ng, dt, gamma, theta, dx, small_rho, nf, eta1, frec, NULL, NULL);
}

/* fill_periodic! for one field (lib/PPMKernels/src/muscl_grid.jl:644-670). */
void MN(vo_muscl_fill_periodic)(MT* f, const int64_t dims[3], int64_t ng) {
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t nix = nx - 2 * ng, niy = ny - 2 * ng, niz = nz - 2 * ng;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int gh = i < ng || i >= nx - ng || j < ng || j >= ny - ng || k < ng || k >= nz - ng;
                if (!gh) continue;
                int64_t si = i < ng ? i + nix : (i >= nx - ng ? i - nix : i);
                int64_t sj = j < ng ? j + niy : (j >= ny - ng ? j - niy : j);
                int64_t sk = k < ng ? k + niz : (k >= nz - ng ? k - niz : k);
                f[(k * ny + j) * nx + i] = f[(sk * ny + sj) * nx + si];
            }
}

/* max_wavespeed (lib/PPMKernels/src/muscl_grid.jl:672-695): max over every cell of
 * sqrt(max(γ(γ−1)·eint, 0)) + max(|u|, |v|, |w|), eint unfloored. */
double MN(vo_muscl_max_wavespeed)(const MT* U, int64_t N, double gamma) {
    MT g = (MT)gamma, m = (MT)0;
    int first = 1;
    for (int64_t i = 0; i < N; ++i) {
        MT d = U[i], u = U[N + i] / d, v = U[2 * N + i] / d, w = U[3 * N + i] / d;
        MT eint = U[4 * N + i] / d - (MT)0.5 * (u * u + v * v + w * w);
        MT cs = MSQRT(MN(mu_max)(g * (g - (MT)1) * eint, (MT)0));
        MT s = cs + MN(mu_max)(MFABS(u), MN(mu_max)(MFABS(v), MFABS(w)));
        m = first ? s : MN(mu_max)(m, s);
        first = 0;
    }
    return (double)m;
}

#undef MN
#undef MCAT
#undef MCAT2
