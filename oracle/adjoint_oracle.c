/*
 * adjoint_oracle.c — reverse-mode (vector-Jacobian) restatement of one SSP-RK2
 * step on a uniform grid, Float64 (TEST INFRASTRUCTURE ONLY; included by
 * vespa_oracle.c).
 *
 * The reference has no derivative of its step; this is the derivative of the
 * oracle's own step (oracle_impl.h, which follows src/driver.jl:209-359,
 * src/riemann.jl:33-88, src/reconstruct.jl:6-16, src/eos.jl:8-32) with dt held
 * fixed. Non-smooth points take the active branch: minmod's selected argument
 * (zero when the signs differ), the active argument of the Davis min/max wave
 * speeds, the upwind HLLC branch, and the pressure floor when it is the larger
 * argument. Written in scatter form over the same face walk as the forward
 * accumulate; pinned by central finite differences (tests/test_adjoint_oracle.py).
 *
 * Forward step (invV = dt/V):  U1 = U − invV·A(U),  U2 = ½U + ½(U1 − invV·A(U1)).
 * Reverse with w = ∂L/∂U2:
 *   λ1 = ½w − ½·invV·A'(U1)ᵀw,   ∂L/∂U = ½w + λ1 − invV·A'(U)ᵀλ1,
 * where A'(U)ᵀλ sums, over every emitted face, the face flux's VJP applied to
 * area·(λ_left − λ_right) (one side only at a domain boundary).
 */

/* ∂(cons2prim)ᵀ: Wb (adjoint of W) → Ub (adjoint of U) at state U. */
static void ad_c2p_vjp(const double U[5], double g, const double Wb[5], double Ub[5]) {
    double rho = U[0];
    double inv = 1.0 / rho;
    double v[3] = {U[1] * inv, U[2] * inv, U[3] * inv};
    double v2 = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    double kin = (0.5 * rho) * v2;
    double pc = (g - 1.0) * (U[4] - kin);
    double pf = 1e-12 * kin;
    double pb = Wb[4];
    double kinb, Eb;
    if (pf > pc) { kinb = pb * 1e-12; Eb = 0.0; }
    else { kinb = -pb * (g - 1.0); Eb = pb * (g - 1.0); }
    double rhob = Wb[0] + kinb * 0.5 * v2;
    double invb = 0.0;
    for (int k = 0; k < 3; ++k) {
        double vb = Wb[k + 1] + kinb * rho * v[k];
        Ub[k + 1] = vb * inv;
        invb += vb * U[k + 1];
    }
    rhob += invb * (-inv * inv);
    Ub[0] = rhob;
    Ub[4] = Eb;
}

/* E(W) adjoint: Eb → Wb (E = p/(γ−1) + (½ρ)|v|²). */
static void ad_energy_vjp(const double W[5], double g, double Eb, double Wb[5]) {
    double v2 = (W[1] * W[1] + W[2] * W[2]) + W[3] * W[3];
    Wb[0] += Eb * 0.5 * v2;
    for (int k = 0; k < 3; ++k) Wb[k + 1] += Eb * W[0] * W[k + 1];
    Wb[4] += Eb / (g - 1.0);
}

/* euler_flux(W) adjoint (src/riemann.jl:9-20), dim 1-based; accumulates into Wb. */
static void ad_eflux_vjp(const double W[5], double g, int dim, const double Fb[5], double Wb[5]) {
    double r = W[0], p = W[4], un = W[dim];
    double E = energy_f64(W, g);
    double rb = Fb[0] * un, unb = Fb[0] * r, pb = 0.0;
    for (int k = 0; k < 3; ++k) {
        double m = r * W[k + 1];
        double mb = Fb[k + 1] * un;
        unb += Fb[k + 1] * m;
        if (dim == k + 1) pb += Fb[k + 1];
        rb += mb * W[k + 1];
        Wb[k + 1] += mb * r;
    }
    double Eb = Fb[4] * un;
    pb += Fb[4] * un;
    unb += Fb[4] * (E + p);
    ad_energy_vjp(W, g, Eb, Wb);
    Wb[0] += rb;
    Wb[4] += pb;
    Wb[dim] += unb;
}

/* hllc_flux adjoint (src/riemann.jl:33-88): ob = ∂L/∂F → accumulates WLb, WRb. */
static void ad_hllc_vjp(const double WL[5], const double WR[5], double g, int dim, const double ob[5],
                        double WLb[5], double WRb[5]) {
    double rL = WL[0], rR = WR[0], pL = WL[4], pR = WR[4];
    double unL = WL[dim], unR = WR[dim];
    double xL = g * pL / rL, xR = g * pR / rR;
    double cL = sqrt(xL), cR = sqrt(xR);
    double aL = unL - cL, aR = unR - cR, bL = unL + cL, bR = unR + cR;
    double SL = jmin_f64(aL, aR), SR = jmax_f64(bL, bR);
    if (SL >= 0) { ad_eflux_vjp(WL, g, dim, ob, WLb); return; }
    if (SR <= 0) { ad_eflux_vjp(WR, g, dim, ob, WRb); return; }
    double tL = rL * unL, tR = rR * unR;
    double num = pR - pL + tL * (SL - unL) - tR * (SR - unR);
    double den = rL * (SL - unL) - rR * (SR - unR);
    double Ss = num / den;
    int useL = Ss >= 0;
    const double* W = useL ? WL : WR;
    double* Wb = useL ? WLb : WRb;
    double SK = useL ? SL : SR;
    double r = W[0], un = W[dim], p = W[4];
    double E = energy_f64(W, g);
    double q = r * (SK - un);
    double d3 = SK - Ss;
    double factor = q / d3;
    double vs[3] = {W[1], W[2], W[3]};
    vs[dim - 1] = Ss;
    double A = Ss - un, B = Ss + p / q;
    double Es = E / r + A * B;
    double Us[5] = {factor, factor * vs[0], factor * vs[1], factor * vs[2], factor * Es};
    double Uk[5] = {r, r * W[1], r * W[2], r * W[3], E};
    /* out = F(W) + SK·(Us − Uk) */
    double SKb = 0.0, Usb[5], Ukb[5];
    for (int i = 0; i < 5; ++i) {
        SKb += ob[i] * (Us[i] - Uk[i]);
        Usb[i] = SK * ob[i];
        Ukb[i] = -SK * ob[i];
    }
    ad_eflux_vjp(W, g, dim, ob, Wb);
    double factorb = Usb[0] + Usb[1] * vs[0] + Usb[2] * vs[1] + Usb[3] * vs[2] + Usb[4] * Es;
    double vsb[3] = {Usb[1] * factor, Usb[2] * factor, Usb[3] * factor};
    double Esb = Usb[4] * factor;
    double rb = 0.0, unb = 0.0, pb = 0.0, Eb = 0.0, Ssb = 0.0, qb = 0.0;
    /* Uk = prim2cons(W) */
    rb += Ukb[0];
    for (int k = 0; k < 3; ++k) {
        rb += Ukb[k + 1] * W[k + 1];
        Wb[k + 1] += Ukb[k + 1] * r;
    }
    Eb += Ukb[4];
    /* Es = E/r + A·B */
    Eb += Esb / r;
    rb -= Esb * E / (r * r);
    double Ab = Esb * B, Bb = Esb * A;
    Ssb += Ab;
    unb -= Ab;
    Ssb += Bb;
    pb += Bb / q;
    qb -= Bb * p / (q * q);
    for (int k = 0; k < 3; ++k) {
        if (k == dim - 1) Ssb += vsb[k];
        else Wb[k + 1] += vsb[k];
    }
    /* factor = q / (SK − Ss) */
    qb += factorb / d3;
    double d3b = -factorb * factor / d3;
    SKb += d3b;
    Ssb -= d3b;
    /* q = r·(SK − un) */
    rb += qb * (SK - un);
    SKb += qb * r;
    unb -= qb * r;
    ad_energy_vjp(W, g, Eb, Wb);
    Wb[0] += rb;
    Wb[4] += pb;
    Wb[dim] += unb;
    double SLb = useL ? SKb : 0.0, SRb = useL ? 0.0 : SKb;
    /* Ss = num/den */
    double numb = Ssb / den, denb = -Ssb * Ss / den;
    double rLb = 0, rRb = 0, pLb = 0, pRb = 0, unLb = 0, unRb = 0, cLb = 0, cRb = 0;
    pRb += numb;
    pLb -= numb;
    double tLb = numb * (SL - unL);
    rLb += tLb * unL;
    unLb += tLb * rL;
    SLb += numb * tL;
    unLb -= numb * tL;
    double tRb = -numb * (SR - unR);
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
    /* Davis speeds: active argument of min / max */
    if (aL <= aR) { unLb += SLb; cLb -= SLb; } else { unRb += SLb; cRb -= SLb; }
    if (bL >= bR) { unLb += SRb; cLb += SRb; } else { unRb += SRb; cRb += SRb; }
    /* c = sqrt(γp/ρ) */
    double xLb = cLb / (2.0 * cL), xRb = cRb / (2.0 * cR);
    pLb += xLb * g / rL;
    rLb -= xLb * xL / rL;
    pRb += xRb * g / rR;
    rRb -= xRb * xR / rR;
    WLb[0] += rLb;
    WLb[4] += pLb;
    WLb[dim] += unLb;
    WRb[0] += rRb;
    WRb[4] += pRb;
    WRb[dim] += unRb;
}

/* Reflecting/outflow ghost is linear: its transpose is itself. */
static void ad_ghost_add(const double Wgb[5], int bc, int axis, double Wb[5]) {
    for (int f = 0; f < 5; ++f) Wb[f] += Wgb[f];
    if (bc == VO_BC_REFLECTING) Wb[axis + 1] -= 2.0 * Wgb[axis + 1];
}

/* Adjoint of _face_value(c, axis, side) = Wc + h·minmod(Wc − Wl, Wr − Wc):
 * adds to Wbar (5·n, SoA) for c and its neighbours (ghosts map back to c). */
static void ad_face_value_vjp(const vo_grid* m, const double* W, int64_t c, int axis, int side,
                              const double fb[5], double* Wbar) {
    int64_t n = vo_ncell(m);
    double Wc[5], Wl[5], Wr[5];
    for (int f = 0; f < 5; ++f) Wc[f] = W[f * n + c];
    int64_t jl = -1, jr = -1;
    int hl = vo_neighbor(m, c, axis, 0, &jl), hr = vo_neighbor(m, c, axis, 1, &jr);
    if (hl) for (int f = 0; f < 5; ++f) Wl[f] = W[f * n + jl];
    else ghost_f64(Wc, m->bc[2 * axis], axis, Wl);
    if (hr) for (int f = 0; f < 5; ++f) Wr[f] = W[f * n + jr];
    else ghost_f64(Wc, m->bc[2 * axis + 1], axis, Wr);
    double h = side ? 0.5 : -0.5;
    double cb[5], lb[5] = {0, 0, 0, 0, 0}, rb[5] = {0, 0, 0, 0, 0};
    for (int f = 0; f < 5; ++f) {
        cb[f] = fb[f];
        double a = Wc[f] - Wl[f], b = Wr[f] - Wc[f];
        double sb = h * fb[f];
        if (a * b <= 0) continue;
        if (fabs(a) < fabs(b)) { cb[f] += sb; lb[f] -= sb; }
        else { rb[f] += sb; cb[f] -= sb; }
    }
    if (hl) for (int f = 0; f < 5; ++f) Wbar[f * n + jl] += lb[f];
    else ad_ghost_add(lb, m->bc[2 * axis], axis, cb);
    if (hr) for (int f = 0; f < 5; ++f) Wbar[f * n + jr] += rb[f];
    else ad_ghost_add(rb, m->bc[2 * axis + 1], axis, cb);
    for (int f = 0; f < 5; ++f) Wbar[f * n + c] += cb[f];
}

/* out = A'(U)ᵀ lam (A = accumulate_flux, src/driver.jl:209-281). */
int vo_accumulate_vjp_f64(const vo_grid* m, const double* U, const double* lam, double* out) {
    if (vo_check(m) || vo_has_parent(m)) return -1;
    int64_t n = vo_ncell(m);
    double g = m->gamma;
    double* W = (double*)malloc(sizeof(double) * 5 * n);
    double* Wbar = (double*)calloc(5 * n, sizeof(double));
    if (!W || !Wbar) { free(W); free(Wbar); return -1; }
    for (int64_t c = 0; c < n; ++c) {
        double Uc[5], Wc[5];
        getU_f64(U, n, c, Uc);
        c2p_f64(Uc, g, g, 0, Wc);
        for (int f = 0; f < 5; ++f) W[f * n + c] = Wc[f];
    }
    for (int64_t c = 0; c < n; ++c) {
        for (int d = 0; d < m->rank; ++d) {
            double aT = vo_face_area(m, d);
            double WL[5], WR[5], Wg[5], mu[5];
            int64_t j;
            face_value_f64(m, U, c, d, 1, 0, WL);
            if (vo_neighbor(m, c, d, 1, &j)) {
                double WLb[5] = {0, 0, 0, 0, 0}, WRb[5] = {0, 0, 0, 0, 0};
                face_value_f64(m, U, j, d, 0, 0, WR);
                for (int f = 0; f < 5; ++f) mu[f] = aT * lam[f * n + c] - aT * lam[f * n + j];
                ad_hllc_vjp(WL, WR, g, d + 1, mu, WLb, WRb);
                ad_face_value_vjp(m, W, c, d, 1, WLb, Wbar);
                ad_face_value_vjp(m, W, j, d, 0, WRb, Wbar);
            } else {
                double WLb[5] = {0, 0, 0, 0, 0}, Wgb[5] = {0, 0, 0, 0, 0};
                ghost_f64(WL, m->bc[2 * d + 1], d, Wg);
                for (int f = 0; f < 5; ++f) mu[f] = aT * lam[f * n + c];
                ad_hllc_vjp(WL, Wg, g, d + 1, mu, WLb, Wgb);
                ad_ghost_add(Wgb, m->bc[2 * d + 1], d, WLb);
                ad_face_value_vjp(m, W, c, d, 1, WLb, Wbar);
            }
            if (!vo_neighbor(m, c, d, 0, &j)) {
                double WRb[5] = {0, 0, 0, 0, 0}, Wgb[5] = {0, 0, 0, 0, 0};
                face_value_f64(m, U, c, d, 0, 0, WR);
                ghost_f64(WR, m->bc[2 * d], d, Wg);
                for (int f = 0; f < 5; ++f) mu[f] = -aT * lam[f * n + c];
                ad_hllc_vjp(Wg, WR, g, d + 1, mu, Wgb, WRb);
                ad_ghost_add(Wgb, m->bc[2 * d], d, WRb);
                ad_face_value_vjp(m, W, c, d, 0, WRb, Wbar);
            }
        }
    }
    for (int64_t c = 0; c < n; ++c) {
        double Uc[5], Wb[5], Ub[5];
        getU_f64(U, n, c, Uc);
        for (int f = 0; f < 5; ++f) Wb[f] = Wbar[f * n + c];
        ad_c2p_vjp(Uc, g, Wb, Ub);
        for (int f = 0; f < 5; ++f) out[f * n + c] = Ub[f];
    }
    free(W);
    free(Wbar);
    return 0;
}

/* ∂L/∂U for L depending on U2 = step(U, dt) through w = ∂L/∂U2 (dt fixed). */
int vo_step_vjp_f64(const vo_grid* m, const double* U, double dt, const double* wbar, double* Ubar) {
    if (vo_check(m) || vo_has_parent(m)) return -1;
    int64_t n = vo_ncell(m);
    double invV = dt / (1.0 * vo_cell_volume(m));
    double* U1 = (double*)malloc(sizeof(double) * 5 * n);
    double* J = (double*)malloc(sizeof(double) * 5 * n);
    double* lam1 = (double*)malloc(sizeof(double) * 5 * n);
    if (!U1 || !J || !lam1) { free(U1); free(J); free(lam1); return -1; }
    vo_accumulate_flux_f64(m, U, J, 0);
    for (int64_t i = 0; i < 5 * n; ++i) U1[i] = U[i] - invV * J[i];
    vo_accumulate_vjp_f64(m, U1, wbar, J);
    for (int64_t i = 0; i < 5 * n; ++i) lam1[i] = 0.5 * wbar[i] + (-0.5 * invV) * J[i];
    vo_accumulate_vjp_f64(m, U, lam1, J);
    for (int64_t i = 0; i < 5 * n; ++i) Ubar[i] = (0.5 * wbar[i] + lam1[i]) + (-invV) * J[i];
    free(U1);
    free(J);
    free(lam1);
    return 0;
}

/* Scalar pieces for the tests: VJPs of cons2prim and of the HLLC flux. */
void vo_c2p_vjp_f64(const double* U, double gamma, const double* Wb, double* Ub) { ad_c2p_vjp(U, gamma, Wb, Ub); }
void vo_hllc_vjp_f64(const double* WL, const double* WR, double gamma, int dim, const double* Fb, double* WLb,
                     double* WRb) {
    for (int f = 0; f < 5; ++f) WLb[f] = WRb[f] = 0.0;
    ad_hllc_vjp(WL, WR, gamma, dim, Fb, WLb, WRb);
}
