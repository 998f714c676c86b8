/*
 * orc_response.c -- oracle: the surface-load response operator (null,
 * elastic, viscoelastic field driver with 1-D and pseudo-spectral 3-D
 * Maxwell memory).
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Restates src/vilma_response.f90:
 *   :386-420 (deg1_to_cm), :422-489 (elastic), :542-677 (init_ve),
 *   :694-826 (solve_drift + skip gate), :842-864 (apply), :944-992 (FE
 *   advance), :994-1053 (enable lateral visc), :1171-1319 (3-D advance),
 *   :1665-1811 (set_dt, save/restore, max_rate, memory_norm), and the
 *   toroidal degree of freedom (:694-809 tor branch, :966-991, :1055-1109,
 *   :1227-1319, :1371-1445, :1524-1596).
 * Memory arrays keep the reference layout (nlam, ne, nk) column-major, nlam =
 * 4, or 6 once the toroidal channels are carried.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

static void deg1_to_cm(const orc_radop* op, double* x) {
    if (!op->nullmode) return;
    double fa_n = op->nullmode[idx_f(op->nr - 1)];
    if (fa_n == 0.0) return;
    double c = -x[idx_f(op->nr - 1)] / fa_n;
    for (int i = 0; i < op->ndof; i++) x[i] = x[i] + c * op->nullmode[i];
}

int orc_resp_init_elastic(orc_resp* r, const orc_earth* e, const orc_mesh* m, int lmax, char* err) {
    int deg1_cm = r->deg1_cm;
    memset(r, 0, sizeof(*r));
    r->deg1_cm = deg1_cm;
    r->kind = RESP_ELASTIC; r->lmax = lmax;
    r->a = e->r_earth; r->g = orc_gravity_at(e, e->r_earth);
    r->ugain = calloc(lmax + 1, 8); r->ngain = calloc(lmax + 1, 8); r->vgain = calloc(lmax + 1, 8);
    r->ngain[0] = 4.0 * ORC_PI * ORC_G * r->a / r->g;
    int nd = 4 * m->nr - 1, fail = 0;
#pragma omp parallel for schedule(dynamic)
    for (int l = 1; l <= lmax; l++) {
        orc_radop op; char e2[256];
        if (orc_radop_assemble(&op, e, m, l, e2)) { fail = 1; continue; }
        double* b = malloc(nd * 8); double* x = malloc(nd * 8);
        orc_radop_load_rhs(&op, 1.0, b);
        orc_radop_solve(&op, b, x, NULL);
        if (l == 1 && deg1_cm) deg1_to_cm(&op, x);
        r->ugain[l] = x[idx_u(op.nr - 1)];
        r->ngain[l] = -x[idx_f(op.nr - 1)] / r->g;
        r->vgain[l] = x[idx_v(op.nr - 1)];
        free(b); free(x); orc_radop_free(&op);
    }
    if (fail) { snprintf(err, 256, "bordered_band_factor: band LU factorization failed"); return 1; }
    if (lmax >= 1 && !deg1_cm) r->ngain[1] = 0.0;
    return 0;
}

int orc_resp_init_ve(orc_resp* r, const orc_earth* e, const orc_mesh* m, const orc_sht* s, double dt, char* err) {
    int deg1_cm = r->deg1_cm, toroidal = r->toroidal;
    double skip_tol = r->skip_tol, visc3d_tol = r->visc3d_tol;
    memset(r, 0, sizeof(*r));
    r->deg1_cm = deg1_cm; r->skip_tol = skip_tol; r->visc3d_tol = visc3d_tol; r->toroidal = toroidal;
    r->kind = RESP_VE;
    r->nlam = 4;
    int L = s->lmax;
    r->lmax = L; r->nlm = s->nlm; r->nr = m->nr; r->ne = m->ne; r->ndof = 4 * m->nr - 1;
    r->dt = dt; r->time = 0.0;
    r->a = e->r_earth; r->g = orc_gravity_at(e, e->r_earth);
    int nr = r->nr, ne = r->ne, nd = r->ndof;
    r->r = malloc(nr * 8); memcpy(r->r, m->r, nr * 8);
    r->mu = malloc(ne * 8); r->Mk = malloc(ne * 8); r->MkPerDt = malloc(ne * 8);
    r->active1d = malloc(ne);
    for (int el = 0; el < ne; el++) {
        const orc_layer* ly = &e->L[m->lay[el]];
        r->mu[el] = ly->mu; r->active1d[el] = 1;
        if (ly->rheology == RHEOL_MAXWELL) {
            r->Mk[el] = r->mu[el] * dt / ly->eta;
            r->MkPerDt[el] = r->mu[el] / ly->eta;
        } else {
            r->Mk[el] = 0.0; r->MkPerDt[el] = 0.0;
        }
    }
    r->Jr = calloc(L + 1, 8);
    r->nrmc = calloc(L + 1, sizeof(*r->nrmc));
    r->sa = calloc(L + 1, sizeof(*r->sa)); r->sb = calloc(L + 1, sizeof(*r->sb)); r->sc = calloc(L + 1, sizeof(*r->sc));
    r->gu = calloc(L + 1, 8); r->gn = calloc(L + 1, 8); r->gv = calloc(L + 1, 8);
    r->xUn = calloc((size_t)L * nr, 8); r->xVn = calloc((size_t)L * nr, 8);
    r->ops = calloc(L, sizeof(orc_radop));
    r->gn[0] = 4.0 * ORC_PI * ORC_G * r->a / r->g;
    int fail = 0;
#pragma omp parallel for schedule(dynamic)
    for (int l = 1; l <= L; l++) {
        char e2[256];
        r->Jr[l] = (double)l * (double)(l + 1);
        orc_strain_constants(r->Jr[l], r->nrmc[l], r->sa[l], r->sb[l], r->sc[l]);
        if (orc_radop_assemble(&r->ops[l - 1], e, m, l, e2)) { fail = 1; continue; }
        double* b = malloc(nd * 8); double* x = malloc(nd * 8);
        orc_radop_load_rhs(&r->ops[l - 1], 1.0, b);
        orc_radop_solve(&r->ops[l - 1], b, x, NULL);
        if (l == 1 && deg1_cm) deg1_to_cm(&r->ops[0], x);
        r->gu[l] = x[idx_u(nr - 1)];
        r->gn[l] = -x[idx_f(nr - 1)] / r->g;
        r->gv[l] = x[idx_v(nr - 1)];
        for (int k = 0; k < nr; k++) {
            r->xUn[(size_t)(l - 1) * nr + k] = x[idx_u(k)];
            r->xVn[(size_t)(l - 1) * nr + k] = x[idx_v(k)];
        }
        free(b); free(x);
    }
    if (fail) { snprintf(err, 256, "bordered_band_factor: band LU factorization failed"); return 1; }
    if (L >= 1 && !deg1_cm) r->gn[1] = 0.0;
    /* degree-grouped slot map (src/vilma_response.f90:629-648) */
    r->nk = L * (L + 3) / 2;
    r->k2lm = malloc(r->nk * sizeof(int)); r->kdeg = malloc(r->nk * sizeof(int));
    r->kbeg = malloc((L + 2) * sizeof(int));
    int k = 0;
    for (int l = 1; l <= L; l++) {
        r->kbeg[l] = k;
        for (int mm = 0; mm <= l; mm++) { r->k2lm[k] = orc_lmidx(s, l, mm); r->kdeg[k] = l; k++; }
    }
    r->kbeg[L + 1] = k;
    size_t nmem = (size_t)4 * ne * r->nk;
    r->Are = calloc(nmem, 8); r->Aim = calloc(nmem, 8); r->Bre = calloc(nmem, 8);
    r->Bim = calloc(nmem, 8); r->Cre = calloc(nmem, 8); r->Cim = calloc(nmem, 8);
    r->dUa = calloc(r->nk, sizeof(cplx)); r->dFa = calloc(r->nk, sizeof(cplx)); r->dVa = calloc(r->nk, sizeof(cplx));
    size_t nn = (size_t)nr * r->nk;
    r->dUn_re = calloc(nn, 8); r->dUn_im = calloc(nn, 8); r->dVn_re = calloc(nn, 8); r->dVn_im = calloc(nn, 8);
    r->mnorm = calloc(r->nk, 8);
    return 0;
}

void orc_resp_free(orc_resp* r) {
    free(r->ugain); free(r->ngain); free(r->vgain);
    if (r->ops) for (int l = 1; l <= r->lmax; l++) orc_radop_free(&r->ops[l - 1]);
    free(r->ops); free(r->r); free(r->mu); free(r->Mk); free(r->MkPerDt); free(r->active1d);
    free(r->Jr); free(r->nrmc); free(r->sa); free(r->sb); free(r->sc);
    free(r->gu); free(r->gn); free(r->gv); free(r->xUn); free(r->xVn);
    free(r->k2lm); free(r->kdeg); free(r->kbeg);
    free(r->Are); free(r->Aim); free(r->Bre); free(r->Bim); free(r->Cre); free(r->Cim);
    free(r->dUa); free(r->dFa); free(r->dVa);
    free(r->dUn_re); free(r->dUn_im); free(r->dVn_re); free(r->dVn_im); free(r->mnorm);
    free(r->Mk3); free(r->MkPerDt3); free(r->e3d);
    if (r->tsh_ready) orc_tsh_free(&r->tsh);
    free(r->Are_s); free(r->Aim_s); free(r->Bre_s); free(r->Bim_s); free(r->Cre_s); free(r->Cim_s);
    free(r->sigma_n); free(r->sigma_next); free(r->sigma_n_s); free(r->dUa_prev);
    free(r->Are0); free(r->Aim0); free(r->Bre0); free(r->Bim0); free(r->Cre0); free(r->Cim0);
    free(r->edUn_re); free(r->edUn_im); free(r->edVn_re); free(r->edVn_im);
    free(r->Are_c); free(r->Aim_c); free(r->Bre_c); free(r->Bim_c); free(r->Cre_c); free(r->Cim_c);
    if (r->tops) for (int l = 1; l <= r->lmax; l++) orc_torop_free(&r->tops[l - 1]);
    free(r->tops); free(r->nrmt); free(r->sat); free(r->sbt); free(r->sct);
    free(r->dWa); free(r->dWa_prev); free(r->dWn_re); free(r->dWn_im); free(r->edWn_re); free(r->edWn_im);
    memset(r, 0, sizeof(*r));
}

/* max that keeps a NaN (fmax drops it): the memory norm feeds the stability
 * guard, whose finite test (src/vilma_timestep.f90:195, "false for NaN /
 * +Inf") must see a NaN memory value */
static inline double nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

/* absmax6 (src/vilma_response.f90:811-826) */
static double absmax6(size_t n, const double* x1, const double* x2, const double* x3,
                      const double* x4, const double* x5, const double* x6) {
    double a = 0.0;
    for (size_t i = 0; i < n; i++) {
        a = nanmax(a, fabs(x1[i])); a = nanmax(a, fabs(x2[i])); a = nanmax(a, fabs(x3[i]));
        a = nanmax(a, fabs(x4[i])); a = nanmax(a, fabs(x5[i])); a = nanmax(a, fabs(x6[i]));
    }
    return a;
}

/* solve_drift (src/vilma_response.f90:694-809) into the nodal targets
 * (U, V re/im per slot and node; W too once the toroidal channels are
 * carried) and the report drift dUa / dFa / dVa (and dWa) */
static void resp_drift(orc_resp* r, double* tUr, double* tUi, double* tVr, double* tVi, double* tWr, double* tWi) {
    int ne = r->ne, nr = r->nr, nd = r->ndof, L = r->lmax;
    const int ld = r->nlam, tor = r->nlam > 4;
    size_t nb = (size_t)ld * ne;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < r->nk; k++)
        r->mnorm[k] = absmax6(nb, r->Are + k * nb, r->Bre + k * nb, r->Cre + k * nb,
                              r->Aim + k * nb, r->Bim + k * nb, r->Cim + k * nb);
    double mx = 0.0;
    for (int k = 0; k < r->nk; k++) mx = nanmax(mx, r->mnorm[k]);
    double thr = r->skip_tol * mx;
#pragma omp parallel
    {
        double *fre = malloc(nd * 8), *fim = malloc(nd * 8), *xre = malloc(nd * 8), *xim = malloc(nd * 8);
        double *gre = malloc(nr * 8), *gim = malloc(nr * 8);
#pragma omp for schedule(dynamic)
        for (int l = 1; l <= L; l++) {
            for (int k = r->kbeg[l]; k < r->kbeg[l + 1]; k++) {
                double* Ur = tUr + (size_t)k * nr; double* Ui = tUi + (size_t)k * nr;
                double* Vr = tVr + (size_t)k * nr; double* Vi = tVi + (size_t)k * nr;
                if (r->mnorm[k] <= thr) {
                    r->dUa[k] = 0.0; r->dFa[k] = 0.0; r->dVa[k] = 0.0;
                    memset(Ur, 0, nr * 8); memset(Ui, 0, nr * 8); memset(Vr, 0, nr * 8); memset(Vi, 0, nr * 8);
                    if (tor) {
                        r->dWa[k] = 0.0;
                        memset(tWr + (size_t)k * nr, 0, nr * 8); memset(tWi + (size_t)k * nr, 0, nr * 8);
                    }
                    continue;
                }
                memset(fre, 0, nd * 8); memset(fim, 0, nd * 8);
                orc_dissipative_rhs(ne, r->r, r->sa[l], r->sb[l], r->sc[l], r->nrmc[l],
                                    r->Are + k * nb, r->Bre + k * nb, r->Cre + k * nb, ld, fre);
                orc_dissipative_rhs(ne, r->r, r->sa[l], r->sb[l], r->sc[l], r->nrmc[l],
                                    r->Aim + k * nb, r->Bim + k * nb, r->Cim + k * nb, ld, fim);
                orc_radop_solve(&r->ops[l - 1], fre, xre, NULL);
                orc_radop_solve(&r->ops[l - 1], fim, xim, NULL);
                if (l == 1 && r->deg1_cm) { deg1_to_cm(&r->ops[0], xre); deg1_to_cm(&r->ops[0], xim); }
                r->dUa[k] = xre[idx_u(nr - 1)] + I * xim[idx_u(nr - 1)];
                r->dFa[k] = xre[idx_f(nr - 1)] + I * xim[idx_f(nr - 1)];
                r->dVa[k] = xre[idx_v(nr - 1)] + I * xim[idx_v(nr - 1)];
                if (l == 1 && !r->deg1_cm) r->dFa[k] = 0.0;
                for (int n = 0; n < nr; n++) {
                    Ur[n] = xre[idx_u(n)]; Ui[n] = xim[idx_u(n)];
                    Vr[n] = xre[idx_v(n)]; Vi[n] = xim[idx_v(n)];
                }
                if (tor) {
                    /* the toroidal drift: its own forcing and operator, no
                     * frame projection (the l = 1 border fixes it) */
                    double* Wr = tWr + (size_t)k * nr; double* Wi = tWi + (size_t)k * nr;
                    memset(gre, 0, nr * 8); memset(gim, 0, nr * 8);
                    orc_dissipative_rhs_tor(ne, r->r, r->sat[l], r->sbt[l], r->sct[l], r->nrmt[l],
                                            r->Are + k * nb, r->Bre + k * nb, r->Cre + k * nb, gre);
                    orc_dissipative_rhs_tor(ne, r->r, r->sat[l], r->sbt[l], r->sct[l], r->nrmt[l],
                                            r->Aim + k * nb, r->Bim + k * nb, r->Cim + k * nb, gim);
                    orc_torop_solve(&r->tops[l - 1], gre, Wr);
                    orc_torop_solve(&r->tops[l - 1], gim, Wi);
                    r->dWa[k] = Wr[nr - 1] + I * Wi[nr - 1];
                }
            }
        }
        free(fre); free(fim); free(xre); free(xim); free(gre); free(gim);
    }
}

/* response_begin_step: the drift of the entering memory into dUn / dVn */
void orc_resp_begin(orc_resp* r) {
    if (r->kind != RESP_VE) return;
    resp_drift(r, r->dUn_re, r->dUn_im, r->dVn_re, r->dVn_im, r->dWn_re, r->dWn_im);
}

/* response_apply for every kind (src/vilma_response.f90:241-252, :376-384, :491-511, :842-864) */
void orc_resp_apply(const orc_resp* r, const orc_sht* s, const cplx* sigma, cplx* u, cplx* n) {
    memset(u, 0, s->nlm * sizeof(cplx)); memset(n, 0, s->nlm * sizeof(cplx));
    if (r->kind == RESP_NULL) return;
    if (r->kind == RESP_ELASTIC) {
        int lcap = r->lmax < s->lmax ? r->lmax : s->lmax;
        for (int m = 0; m <= s->lmax; m++)
            for (int l = m; l <= lcap; l++) {
                int lm = orc_lmidx(s, l, m);
                u[lm] = r->ugain[l] * sigma[lm]; n[lm] = r->ngain[l] * sigma[lm];
            }
        return;
    }
    int lm0 = orc_lmidx(s, 0, 0);
    u[lm0] = r->gu[0] * sigma[lm0];
    n[lm0] = r->gn[0] * sigma[lm0];
    for (int k = 0; k < r->nk; k++) {
        int l = r->kdeg[k], lm = r->k2lm[k];
        u[lm] = r->gu[l] * sigma[lm] + r->dUa[k];
        n[lm] = r->gn[l] * sigma[lm] - r->dFa[k] / r->g;
    }
}

/* gather + advance_shape_tensor for one element (src/vilma_response.f90:1227-1319) */
static void advance_element_3d(orc_resp* r, const orc_sht* s, const cplx* sigma, int e,
                               cplx* cm, cplx* ce, cplx* cdel, double* dtau, double* deps) {
    int nlm = s->nlm, nr = r->nr, ne = r->ne;
    const int nc = r->nlam;
    size_t nb = (size_t)nc * ne;
    size_t ng = (size_t)s->nphi * s->nlat;
    /* cm/ce: 3 shapes x [lam][lm], nc channels */
    memset(cm, 0, 3 * (size_t)nc * nlm * sizeof(cplx)); memset(ce, 0, 3 * (size_t)nc * nlm * sizeof(cplx));
    for (int k = 0; k < r->nk; k++) {
        int l = r->kdeg[k], lm = r->k2lm[k];
        double sre = creal(sigma[lm]), sim = cimag(sigma[lm]);
        const double* xU = r->xUn + (size_t)(l - 1) * nr; const double* xV = r->xVn + (size_t)(l - 1) * nr;
        size_t o = (size_t)k * nr;
        double Ur = sre * xU[e] + r->dUn_re[o + e], Ui = sim * xU[e] + r->dUn_im[o + e];
        double Ur1 = sre * xU[e + 1] + r->dUn_re[o + e + 1], Ui1 = sim * xU[e + 1] + r->dUn_im[o + e + 1];
        double Vr = sre * xV[e] + r->dVn_re[o + e], Vi = sim * xV[e] + r->dVn_im[o + e];
        double Vr1 = sre * xV[e + 1] + r->dVn_re[o + e + 1], Vi1 = sim * xV[e + 1] + r->dVn_im[o + e + 1];
        double ar[4], br[4], cr[4], ai[4], bi[4], ci[4];
        orc_strain_coeffs(Ur, Ur1, Vr, Vr1, r->Jr[l], ar, br, cr);
        orc_strain_coeffs(Ui, Ui1, Vi, Vi1, r->Jr[l], ai, bi, ci);
        size_t mo = k * nb + (size_t)e * nc;
        for (int lam = 0; lam < 4; lam++) {
            ce[(0 * nc + lam) * nlm + lm] = ar[lam] + I * ai[lam];
            ce[(1 * nc + lam) * nlm + lm] = br[lam] + I * bi[lam];
            ce[(2 * nc + lam) * nlm + lm] = cr[lam] + I * ci[lam];
        }
        if (nc == 6) {
            /* channels 4, 5: the strain of the nodal W drift (no load term) */
            double at[2], bt[2], ct[2], ati[2], bti[2], cti[2];
            orc_strain_coeffs_tor(r->dWn_re[o + e], r->dWn_re[o + e + 1], r->Jr[l], at, bt, ct);
            orc_strain_coeffs_tor(r->dWn_im[o + e], r->dWn_im[o + e + 1], r->Jr[l], ati, bti, cti);
            for (int q = 0; q < 2; q++) {
                ce[(0 * nc + 4 + q) * nlm + lm] = at[q] + I * ati[q];
                ce[(1 * nc + 4 + q) * nlm + lm] = bt[q] + I * bti[q];
                ce[(2 * nc + 4 + q) * nlm + lm] = ct[q] + I * cti[q];
            }
        }
        for (int lam = 0; lam < nc; lam++) {
            cm[(0 * nc + lam) * nlm + lm] = r->Are[mo + lam] + I * r->Aim[mo + lam];
            cm[(1 * nc + lam) * nlm + lm] = r->Bre[mo + lam] + I * r->Bim[mo + lam];
            cm[(2 * nc + lam) * nlm + lm] = r->Cre[mo + lam] + I * r->Cim[mo + lam];
        }
    }
    double twoMu = 2.0 * r->mu[e];
    const double* M3 = r->Mk3 + (size_t)e * ng;
    for (int sh = 0; sh < 3; sh++) {
        cplx* c = cm + (size_t)sh * nc * nlm;
        orc_tsh_synth(&r->tsh, s, c, nc, dtau);
        orc_tsh_synth(&r->tsh, s, ce + (size_t)sh * nc * nlm, nc, deps);
        for (int p = 0; p < 6; p++)
            for (size_t id = 0; id < ng; id++)
                dtau[p * ng + id] = -M3[id] * (dtau[p * ng + id] + twoMu * deps[p * ng + id]);
        orc_tsh_analysis(&r->tsh, s, dtau, nc, cdel);
        for (int q = 0; q < nc * nlm; q++) c[q] = c[q] + cdel[q];
    }
    for (int k = 0; k < r->nk; k++) {
        int lm = r->k2lm[k];
        size_t mo = k * nb + (size_t)e * nc;
        for (int lam = 0; lam < nc; lam++) {
            r->Are[mo + lam] = creal(cm[(0 * nc + lam) * nlm + lm]); r->Aim[mo + lam] = cimag(cm[(0 * nc + lam) * nlm + lm]);
            r->Bre[mo + lam] = creal(cm[(1 * nc + lam) * nlm + lm]); r->Bim[mo + lam] = cimag(cm[(1 * nc + lam) * nlm + lm]);
            r->Cre[mo + lam] = creal(cm[(2 * nc + lam) * nlm + lm]); r->Cim[mo + lam] = cimag(cm[(2 * nc + lam) * nlm + lm]);
        }
    }
}

/* vilma_advance (src/vilma_response.f90:944-992) with advance_memory_3d (:1171-1225) */
void orc_resp_advance(orc_resp* r, const orc_sht* s, const cplx* sigma) {
    if (r->kind != RESP_VE) return;
    int nr = r->nr, ne = r->ne;
    const int ld = r->nlam;
    size_t nb = (size_t)ld * ne;
    if (r->lat_visc && r->ne3d > 0) {
        size_t ng = (size_t)s->nphi * s->nlat;
        int nlm = s->nlm;
#pragma omp parallel
        {
            cplx* cm = malloc((size_t)18 * nlm * sizeof(cplx));
            cplx* ce = malloc((size_t)18 * nlm * sizeof(cplx));
            cplx* cdel = malloc((size_t)6 * nlm * sizeof(cplx));
            double* dtau = malloc(6 * ng * 8); double* deps = malloc(6 * ng * 8);
#pragma omp for schedule(dynamic)
            for (int ei = 0; ei < r->ne3d; ei++) advance_element_3d(r, s, sigma, r->e3d[ei], cm, ce, cdel, dtau, deps);
            free(cm); free(ce); free(cdel); free(dtau); free(deps);
        }
    }
#pragma omp parallel
    {
        double *Ure = malloc(nr * 8), *Uim = malloc(nr * 8), *Vre = malloc(nr * 8), *Vim = malloc(nr * 8);
#pragma omp for schedule(static)
        for (int k = 0; k < r->nk; k++) {
            int l = r->kdeg[k], lm = r->k2lm[k];
            double sre = creal(sigma[lm]), sim = cimag(sigma[lm]);
            const double* xU = r->xUn + (size_t)(l - 1) * nr; const double* xV = r->xVn + (size_t)(l - 1) * nr;
            size_t o = (size_t)k * nr;
            for (int n = 0; n < nr; n++) {
                Ure[n] = sre * xU[n] + r->dUn_re[o + n];
                Uim[n] = sim * xU[n] + r->dUn_im[o + n];
                Vre[n] = sre * xV[n] + r->dVn_re[o + n];
                Vim[n] = sim * xV[n] + r->dVn_im[o + n];
            }
            orc_advance_memory_fe(ne, r->mu, r->Mk, Ure, Vre, r->Jr[l], r->Are + k * nb, r->Bre + k * nb, r->Cre + k * nb, ld, r->active1d);
            orc_advance_memory_fe(ne, r->mu, r->Mk, Uim, Vim, r->Jr[l], r->Aim + k * nb, r->Bim + k * nb, r->Cim + k * nb, ld, r->active1d);
            if (ld > 4) {   /* W is pure drift: no sigma term */
                orc_advance_memory_tor(ne, r->mu, r->Mk, r->dWn_re + o, NULL, r->Jr[l], r->Are + k * nb, r->Bre + k * nb,
                                       r->Cre + k * nb, r->active1d);
                orc_advance_memory_tor(ne, r->mu, r->Mk, r->dWn_im + o, NULL, r->Jr[l], r->Aim + k * nb, r->Bim + k * nb,
                                       r->Cim + k * nb, r->active1d);
            }
        }
        free(Ure); free(Uim); free(Vre); free(Vim);
    }
}

void orc_resp_set_dt(orc_resp* r, double dt) {
    r->dt = dt;
    if (r->kind != RESP_VE) return;
    for (int e = 0; e < r->ne; e++) r->Mk[e] = r->MkPerDt[e] * dt;
    if (r->lat_visc) {
        size_t n = (size_t)r->ne * r->tsh.nphi * r->tsh.nlat;
        for (size_t i = 0; i < n; i++) r->Mk3[i] = r->MkPerDt3[i] * dt;
    }
}

double orc_resp_max_rate(const orc_resp* r) {
    double rate = 0.0;
    if (r->kind != RESP_VE) return 0.0;
    for (int e = 0; e < r->ne; e++) rate = fmax(rate, r->MkPerDt[e]);
    if (r->lat_visc) {
        size_t n = (size_t)r->ne * r->tsh.nphi * r->tsh.nlat;
        for (size_t i = 0; i < n; i++) rate = fmax(rate, r->MkPerDt3[i]);
    }
    return rate;
}

double orc_resp_memory_norm(const orc_resp* r) {
    if (r->kind != RESP_VE || !r->Are) return 0.0;
    size_t nb = (size_t)r->nlam * r->ne;
    /* response_memory_norm (src/vilma_response.f90:1795-1811): per-slot
     * maxima in parallel, then one ordered NaN-keeping max */
    double* sk = malloc((size_t)r->nk * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < r->nk; k++)
        sk[k] = absmax6(nb, r->Are + k * nb, r->Aim + k * nb, r->Bre + k * nb,
                        r->Bim + k * nb, r->Cre + k * nb, r->Cim + k * nb);
    double nrm = 0.0;
    for (int k = 0; k < r->nk; k++) nrm = nanmax(nrm, sk[k]);
    free(sk);
    return nrm;
}

void orc_resp_save(orc_resp* r) {
    if (r->kind != RESP_VE) { r->time_s = r->time; return; }
    size_t n = (size_t)r->nlam * r->ne * r->nk;
    if (!r->Are_s) {
        r->Are_s = malloc(n * 8); r->Aim_s = malloc(n * 8); r->Bre_s = malloc(n * 8);
        r->Bim_s = malloc(n * 8); r->Cre_s = malloc(n * 8); r->Cim_s = malloc(n * 8);
    }
    memcpy(r->Are_s, r->Are, n * 8); memcpy(r->Aim_s, r->Aim, n * 8); memcpy(r->Bre_s, r->Bre, n * 8);
    memcpy(r->Bim_s, r->Bim, n * 8); memcpy(r->Cre_s, r->Cre, n * 8); memcpy(r->Cim_s, r->Cim, n * 8);
    r->time_s = r->time;
    /* response_save_state also keeps sigma_n (src/vilma_response.f90:1695-1723) */
    if (r->sigma_n) memcpy(r->sigma_n_s, r->sigma_n, (size_t)r->nlm * sizeof(cplx));
    r->sigma_primed_s = r->sigma_primed;
}

void orc_resp_restore(orc_resp* r) {
    r->time = r->time_s;
    if (r->kind != RESP_VE) return;
    size_t n = (size_t)r->nlam * r->ne * r->nk;
    memcpy(r->Are, r->Are_s, n * 8); memcpy(r->Aim, r->Aim_s, n * 8); memcpy(r->Bre, r->Bre_s, n * 8);
    memcpy(r->Bim, r->Bim_s, n * 8); memcpy(r->Cre, r->Cre_s, n * 8); memcpy(r->Cim, r->Cim_s, n * 8);
    if (r->sigma_n) memcpy(r->sigma_n, r->sigma_n_s, (size_t)r->nlm * sizeof(cplx));
    r->sigma_primed = r->sigma_primed_s;
}

/* (nlam, ne, nk) -> (6, ne, nk), spheroidal channels kept, the new ones zero
 * (enable_toroidal's widen, src/vilma_response.f90:1092-1108) */
static void widen(double** x, int ne, int nk) {
    if (!*x) return;
    double* y = calloc((size_t)6 * ne * nk, sizeof(double));
    for (size_t ke = 0; ke < (size_t)ne * nk; ke++) memcpy(y + ke * 6, *x + ke * 4, 4 * sizeof(double));
    free(*x);
    *x = y;
}

/* enable_toroidal (src/vilma_response.f90:1055-1109): per-degree W operators
 * and constants, the W drift arrays, and every memory array and state buffer
 * widened to six channels */
static int enable_toroidal(orc_resp* r, char* err) {
    int L = r->lmax, fail = 0;
    r->tops = calloc(L, sizeof(orc_torop));
    r->nrmt = calloc(L + 1, sizeof(*r->nrmt));
    r->sat = calloc(L + 1, sizeof(*r->sat)); r->sbt = calloc(L + 1, sizeof(*r->sbt)); r->sct = calloc(L + 1, sizeof(*r->sct));
#pragma omp parallel for schedule(dynamic)
    for (int l = 1; l <= L; l++) {
        char e2[256];
        if (orc_torop_assemble(&r->tops[l - 1], r->r, r->mu, r->ne, l, e2)) fail = 1;
        orc_strain_constants_tor(r->Jr[l], r->nrmt[l], r->sat[l], r->sbt[l], r->sct[l]);
    }
    if (fail) { snprintf(err, 256, "toroidal_operator_assemble: band LU factorization failed"); return 1; }
    size_t nn = (size_t)r->nr * r->nk;
    r->dWa = calloc(r->nk ? r->nk : 1, sizeof(cplx));
    r->dWn_re = calloc(nn, 8); r->dWn_im = calloc(nn, 8);
    r->nlam = 6;
    double** arrs[18] = {&r->Are, &r->Aim, &r->Bre, &r->Bim, &r->Cre, &r->Cim,
                         &r->Are0, &r->Aim0, &r->Bre0, &r->Bim0, &r->Cre0, &r->Cim0,
                         &r->Are_s, &r->Aim_s, &r->Bre_s, &r->Bim_s, &r->Cre_s, &r->Cim_s};
    for (int q = 0; q < 18; q++) widen(arrs[q], r->ne, r->nk);
    double** arrc[6] = {&r->Are_c, &r->Aim_c, &r->Bre_c, &r->Bim_c, &r->Cre_c, &r->Cim_c};
    for (int q = 0; q < 6; q++) widen(arrc[q], r->ne, r->nk);
    if (r->Are0) {   /* ensure_commit_scratch_tor (:1111-1118) */
        r->edWn_re = calloc(nn, 8); r->edWn_im = calloc(nn, 8);
        r->dWa_prev = calloc(r->nk ? r->nk : 1, sizeof(cplx));
    }
    return 0;
}

/* response_enable_lateral_visc (src/vilma_response.f90:994-1053); with the
 * toroidal flag, the first genuinely 3-D element turns the toroidal channels
 * on for good. pert is [e][nlat][nphi]. */
int orc_resp_enable_lateral(orc_resp* r, const orc_sht* s, const double* pert, char* err) {
    if (r->kind != RESP_VE) { snprintf(err, 256, "solid_earth_enable_visc_3d: lateral viscosity needs earth_response=ve"); return 1; }
    size_t ng = (size_t)s->nphi * s->nlat;
    if (!r->tsh_ready) { orc_tsh_init(&r->tsh, s); r->tsh_ready = 1; }
    free(r->Mk3); free(r->MkPerDt3); free(r->e3d);
    r->MkPerDt3 = malloc(ng * r->ne * 8); r->Mk3 = malloc(ng * r->ne * 8);
    for (int e = 0; e < r->ne; e++)
        for (size_t i = 0; i < ng; i++) r->MkPerDt3[e * ng + i] = r->MkPerDt[e] * pow(10.0, -pert[e * ng + i]);
    r->ne3d = 0;
    for (int e = 0; e < r->ne; e++) {
        r->active1d[e] = 1;
        if (r->MkPerDt[e] == 0.0) continue;
        double mx = -INFINITY, mn = INFINITY, sum = 0.0;
        for (size_t i = 0; i < ng; i++) { double v = pert[e * ng + i]; mx = fmax(mx, v); mn = fmin(mn, v); }
        if (mx - mn > r->visc3d_tol) { r->active1d[e] = 0; r->ne3d++; }
        else {
            /* Fortran sum(pert(:,:,e)) order: phi fastest, then latitude */
            for (size_t i = 0; i < ng; i++) sum += pert[e * ng + i];
            r->MkPerDt[e] = r->MkPerDt[e] * pow(10.0, -(sum / (double)ng));
        }
    }
    for (int e = 0; e < r->ne; e++) r->Mk[e] = r->MkPerDt[e] * r->dt;
    for (size_t i = 0; i < ng * r->ne; i++) r->Mk3[i] = r->MkPerDt3[i] * r->dt;
    r->e3d = malloc((r->ne3d > 0 ? r->ne3d : 1) * sizeof(int));
    int q = 0;
    for (int e = 0; e < r->ne; e++) if (!r->active1d[e]) r->e3d[q++] = e;
    r->lat_visc = 1;
    if (r->toroidal && r->ne3d > 0 && r->nlam == 4) return enable_toroidal(r, err);
    return 0;
}

/* ------------------------------------------------------------------ TRAP
 * The implicit (trapezoidal, Crank-Nicolson) memory rule with the
 * sea-level <-> memory co-convergence (src/vilma_response.f90:1514-1663) and
 * the step-doubling controller's buffers (:1682-1788). */

static double* dupz(size_t n) { return calloc(n ? n : 1, sizeof(double)); }

void orc_resp_set_implicit(orc_resp* r, int nlm) {
    r->implicit = 1;
    r->couple_tol = 1.0e-6;   /* src/vilma_response.f90:177 */
    r->nlm = nlm;
    if (r->kind != RESP_VE || r->Are0) return;
    size_t n = (size_t)r->nlam * r->ne * r->nk, nn = (size_t)r->nk * r->nr;
    r->Are0 = dupz(n); r->Aim0 = dupz(n); r->Bre0 = dupz(n); r->Bim0 = dupz(n); r->Cre0 = dupz(n); r->Cim0 = dupz(n);
    r->Are_c = dupz(n); r->Aim_c = dupz(n); r->Bre_c = dupz(n); r->Bim_c = dupz(n); r->Cre_c = dupz(n); r->Cim_c = dupz(n);
    if (!r->Are_s) {
        r->Are_s = dupz(n); r->Aim_s = dupz(n); r->Bre_s = dupz(n); r->Bim_s = dupz(n); r->Cre_s = dupz(n); r->Cim_s = dupz(n);
    }
    r->edUn_re = dupz(nn); r->edUn_im = dupz(nn); r->edVn_re = dupz(nn); r->edVn_im = dupz(nn);
    r->sigma_n = calloc(nlm, sizeof(cplx)); r->sigma_next = calloc(nlm, sizeof(cplx));
    r->sigma_n_s = calloc(nlm, sizeof(cplx)); r->dUa_prev = calloc(r->nk ? r->nk : 1, sizeof(cplx));
    if (r->nlam > 4) {
        r->edWn_re = dupz(nn); r->edWn_im = dupz(nn);
        r->dWa_prev = calloc(r->nk ? r->nk : 1, sizeof(cplx));
    }
}

/* ve_response_prepare_endpoint (:1598-1616): snapshot tau_n, seed the
 * endpoint drift with the entering drift, keep the report drift */
void orc_resp_prepare_endpoint(orc_resp* r) {
    r->couple_pass = 0;
    r->couple_done = 0;
    if (!r->implicit || r->kind != RESP_VE) return;
    size_t n = (size_t)r->nlam * r->ne * r->nk, nn = (size_t)r->nk * r->nr;
    memcpy(r->Are0, r->Are, n * 8); memcpy(r->Aim0, r->Aim, n * 8); memcpy(r->Bre0, r->Bre, n * 8);
    memcpy(r->Bim0, r->Bim, n * 8); memcpy(r->Cre0, r->Cre, n * 8); memcpy(r->Cim0, r->Cim, n * 8);
    memcpy(r->edUn_re, r->dUn_re, nn * 8); memcpy(r->edUn_im, r->dUn_im, nn * 8);
    memcpy(r->edVn_re, r->dVn_re, nn * 8); memcpy(r->edVn_im, r->dVn_im, nn * 8);
    memcpy(r->dUa_prev, r->dUa, (size_t)r->nk * sizeof(cplx));
    if (r->nlam > 4) {
        memcpy(r->edWn_re, r->dWn_re, nn * 8); memcpy(r->edWn_im, r->dWn_im, nn * 8);
        memcpy(r->dWa_prev, r->dWa, (size_t)r->nk * sizeof(cplx));
    }
}

/* advance_shape_tensor_trap for the three shapes of element e
 * (src/vilma_response.f90:1371-1478): tau_n from the snapshot, eps_n from
 * (sigma_n, dUn), eps_{n+1} from (sigma_{n+1}, edUn); pointwise
 * tau+ - tau_n = -M/(1 + M/2) tau_n - mu M/(1 + M/2) (eps_n + eps_{n+1}) */
static void advance_element_3d_trap(orc_resp* r, const orc_sht* s, const cplx* sigma, const cplx* sig_n, int e,
                                    cplx* cm, cplx* cn, cplx* c1, cplx* cdel, double* dt0, double* den, double* de1) {
    int nlm = s->nlm, nr = r->nr, ne = r->ne;
    const int nc = r->nlam;
    size_t nb = (size_t)nc * ne, ng = (size_t)s->nphi * s->nlat;
    memset(cm, 0, 3 * (size_t)nc * nlm * sizeof(cplx)); memset(cn, 0, 3 * (size_t)nc * nlm * sizeof(cplx));
    memset(c1, 0, 3 * (size_t)nc * nlm * sizeof(cplx));
    for (int k = 0; k < r->nk; k++) {
        int l = r->kdeg[k], lm = r->k2lm[k];
        double sre = creal(sigma[lm]), sim = cimag(sigma[lm]);
        double srn = creal(sig_n[lm]), sin_ = cimag(sig_n[lm]);
        const double* xU = r->xUn + (size_t)(l - 1) * nr; const double* xV = r->xVn + (size_t)(l - 1) * nr;
        size_t o = (size_t)k * nr;
        double a[4][4], b[4][4], c[4][4];   /* [eps_n re, eps_n im, eps_1 re, eps_1 im][lam] */
        orc_strain_coeffs(srn * xU[e] + r->dUn_re[o + e], srn * xU[e + 1] + r->dUn_re[o + e + 1],
                          srn * xV[e] + r->dVn_re[o + e], srn * xV[e + 1] + r->dVn_re[o + e + 1], r->Jr[l], a[0], b[0], c[0]);
        orc_strain_coeffs(sin_ * xU[e] + r->dUn_im[o + e], sin_ * xU[e + 1] + r->dUn_im[o + e + 1],
                          sin_ * xV[e] + r->dVn_im[o + e], sin_ * xV[e + 1] + r->dVn_im[o + e + 1], r->Jr[l], a[1], b[1], c[1]);
        orc_strain_coeffs(sre * xU[e] + r->edUn_re[o + e], sre * xU[e + 1] + r->edUn_re[o + e + 1],
                          sre * xV[e] + r->edVn_re[o + e], sre * xV[e + 1] + r->edVn_re[o + e + 1], r->Jr[l], a[2], b[2], c[2]);
        orc_strain_coeffs(sim * xU[e] + r->edUn_im[o + e], sim * xU[e + 1] + r->edUn_im[o + e + 1],
                          sim * xV[e] + r->edVn_im[o + e], sim * xV[e + 1] + r->edVn_im[o + e + 1], r->Jr[l], a[3], b[3], c[3]);
        size_t mo = k * nb + (size_t)e * nc;
        for (int lam = 0; lam < 4; lam++) {
            cn[(0 * nc + lam) * nlm + lm] = a[0][lam] + I * a[1][lam];
            cn[(1 * nc + lam) * nlm + lm] = b[0][lam] + I * b[1][lam];
            cn[(2 * nc + lam) * nlm + lm] = c[0][lam] + I * c[1][lam];
            c1[(0 * nc + lam) * nlm + lm] = a[2][lam] + I * a[3][lam];
            c1[(1 * nc + lam) * nlm + lm] = b[2][lam] + I * b[3][lam];
            c1[(2 * nc + lam) * nlm + lm] = c[2][lam] + I * c[3][lam];
        }
        if (nc == 6) {
            /* W: eps_n from the tau_n drift dWn, eps_{n+1} from the endpoint's edWn */
            double ta[4][2], tb[4][2], tc[4][2];
            orc_strain_coeffs_tor(r->dWn_re[o + e], r->dWn_re[o + e + 1], r->Jr[l], ta[0], tb[0], tc[0]);
            orc_strain_coeffs_tor(r->dWn_im[o + e], r->dWn_im[o + e + 1], r->Jr[l], ta[1], tb[1], tc[1]);
            orc_strain_coeffs_tor(r->edWn_re[o + e], r->edWn_re[o + e + 1], r->Jr[l], ta[2], tb[2], tc[2]);
            orc_strain_coeffs_tor(r->edWn_im[o + e], r->edWn_im[o + e + 1], r->Jr[l], ta[3], tb[3], tc[3]);
            for (int q = 0; q < 2; q++) {
                cn[(0 * nc + 4 + q) * nlm + lm] = ta[0][q] + I * ta[1][q];
                cn[(1 * nc + 4 + q) * nlm + lm] = tb[0][q] + I * tb[1][q];
                cn[(2 * nc + 4 + q) * nlm + lm] = tc[0][q] + I * tc[1][q];
                c1[(0 * nc + 4 + q) * nlm + lm] = ta[2][q] + I * ta[3][q];
                c1[(1 * nc + 4 + q) * nlm + lm] = tb[2][q] + I * tb[3][q];
                c1[(2 * nc + 4 + q) * nlm + lm] = tc[2][q] + I * tc[3][q];
            }
        }
        for (int lam = 0; lam < nc; lam++) {
            cm[(0 * nc + lam) * nlm + lm] = r->Are0[mo + lam] + I * r->Aim0[mo + lam];
            cm[(1 * nc + lam) * nlm + lm] = r->Bre0[mo + lam] + I * r->Bim0[mo + lam];
            cm[(2 * nc + lam) * nlm + lm] = r->Cre0[mo + lam] + I * r->Cim0[mo + lam];
        }
    }
    const double* M3 = r->Mk3 + (size_t)e * ng;
    for (int sh = 0; sh < 3; sh++) {
        cplx* cc = cm + (size_t)sh * nc * nlm;
        orc_tsh_synth(&r->tsh, s, cc, nc, dt0);
        orc_tsh_synth(&r->tsh, s, cn + (size_t)sh * nc * nlm, nc, den);
        orc_tsh_synth(&r->tsh, s, c1 + (size_t)sh * nc * nlm, nc, de1);
        for (size_t id = 0; id < ng; id++) {
            const double wt = M3[id] / (1.0 + 0.5 * M3[id]), we = r->mu[e] * M3[id] / (1.0 + 0.5 * M3[id]);
            for (int p = 0; p < 6; p++)
                dt0[p * ng + id] = -wt * dt0[p * ng + id] - we * (den[p * ng + id] + de1[p * ng + id]);
        }
        orc_tsh_analysis(&r->tsh, s, dt0, nc, cdel);
        for (int q = 0; q < nc * nlm; q++) cc[q] = cc[q] + cdel[q];
    }
    for (int k = 0; k < r->nk; k++) {
        int lm = r->k2lm[k];
        size_t mo = k * nb + (size_t)e * nc;
        for (int lam = 0; lam < nc; lam++) {
            r->Are[mo + lam] = creal(cm[(0 * nc + lam) * nlm + lm]); r->Aim[mo + lam] = cimag(cm[(0 * nc + lam) * nlm + lm]);
            r->Bre[mo + lam] = creal(cm[(1 * nc + lam) * nlm + lm]); r->Bim[mo + lam] = cimag(cm[(1 * nc + lam) * nlm + lm]);
            r->Cre[mo + lam] = creal(cm[(2 * nc + lam) * nlm + lm]); r->Cim[mo + lam] = cimag(cm[(2 * nc + lam) * nlm + lm]);
        }
    }
}

/* trapezoid_advance_all (:1524-1596): every slot reset to tau_n and the
 * 1-D-effective elements advanced spectrally, then the 3-D elements from
 * the snapshot */
static void trapezoid_advance_all(orc_resp* r, const orc_sht* s, const cplx* sigma) {
    int nr = r->nr, ne = r->ne;
    const int ld = r->nlam;
    size_t nb = (size_t)ld * ne;
    const cplx* sig_n = r->sigma_primed ? r->sigma_n : sigma;   /* first step: eps_n from sigma_{n+1} */
#pragma omp parallel
    {
        double *Ur = malloc(nr * 8), *Ui = malloc(nr * 8), *Vr = malloc(nr * 8), *Vi = malloc(nr * 8);
        double *Urn = malloc(nr * 8), *Uin = malloc(nr * 8), *Vrn = malloc(nr * 8), *Vin = malloc(nr * 8);
#pragma omp for schedule(static)
        for (int k = 0; k < r->nk; k++) {
            int l = r->kdeg[k], lm = r->k2lm[k];
            double sre = creal(sigma[lm]), sim = cimag(sigma[lm]);
            double srn = creal(sig_n[lm]), sin_ = cimag(sig_n[lm]);
            const double* xU = r->xUn + (size_t)(l - 1) * nr; const double* xV = r->xVn + (size_t)(l - 1) * nr;
            size_t o = (size_t)k * nr;
            for (int n = 0; n < nr; n++) {
                Urn[n] = srn * xU[n] + r->dUn_re[o + n]; Vrn[n] = srn * xV[n] + r->dVn_re[o + n];
                Uin[n] = sin_ * xU[n] + r->dUn_im[o + n]; Vin[n] = sin_ * xV[n] + r->dVn_im[o + n];
                Ur[n] = sre * xU[n] + r->edUn_re[o + n]; Vr[n] = sre * xV[n] + r->edVn_re[o + n];
                Ui[n] = sim * xU[n] + r->edUn_im[o + n]; Vi[n] = sim * xV[n] + r->edVn_im[o + n];
            }
            memcpy(r->Are + k * nb, r->Are0 + k * nb, nb * 8); memcpy(r->Bre + k * nb, r->Bre0 + k * nb, nb * 8);
            memcpy(r->Cre + k * nb, r->Cre0 + k * nb, nb * 8);
            orc_advance_memory_trap(ne, r->mu, r->Mk, Ur, Vr, Urn, Vrn, r->Jr[l], r->Are + k * nb, r->Bre + k * nb,
                                    r->Cre + k * nb, ld, r->active1d);
            memcpy(r->Aim + k * nb, r->Aim0 + k * nb, nb * 8); memcpy(r->Bim + k * nb, r->Bim0 + k * nb, nb * 8);
            memcpy(r->Cim + k * nb, r->Cim0 + k * nb, nb * 8);
            orc_advance_memory_trap(ne, r->mu, r->Mk, Ui, Vi, Uin, Vin, r->Jr[l], r->Aim + k * nb, r->Bim + k * nb,
                                    r->Cim + k * nb, ld, r->active1d);
            if (ld > 4) {   /* W: eps_n from tau_n's drift, eps_{n+1} from the endpoint's */
                orc_advance_memory_tor(ne, r->mu, r->Mk, r->edWn_re + o, r->dWn_re + o, r->Jr[l], r->Are + k * nb,
                                       r->Bre + k * nb, r->Cre + k * nb, r->active1d);
                orc_advance_memory_tor(ne, r->mu, r->Mk, r->edWn_im + o, r->dWn_im + o, r->Jr[l], r->Aim + k * nb,
                                       r->Bim + k * nb, r->Cim + k * nb, r->active1d);
            }
        }
        free(Ur); free(Ui); free(Vr); free(Vi); free(Urn); free(Uin); free(Vrn); free(Vin);
    }
    if (r->lat_visc && r->ne3d > 0) {
        size_t ng = (size_t)s->nphi * s->nlat;
        int nlm = s->nlm;
#pragma omp parallel
        {
            cplx* cm = malloc((size_t)18 * nlm * sizeof(cplx));
            cplx* cn = malloc((size_t)18 * nlm * sizeof(cplx));
            cplx* c1 = malloc((size_t)18 * nlm * sizeof(cplx));
            cplx* cdel = malloc((size_t)6 * nlm * sizeof(cplx));
            double *dt0 = malloc(6 * ng * 8), *den = malloc(6 * ng * 8), *de1 = malloc(6 * ng * 8);
#pragma omp for schedule(dynamic)
            for (int ei = 0; ei < r->ne3d; ei++)
                advance_element_3d_trap(r, s, sigma, sig_n, r->e3d[ei], cm, cn, c1, cdel, dt0, den, de1);
            free(cm); free(cn); free(c1); free(cdel); free(dt0); free(den); free(de1);
        }
    }
}

/* ve_response_advance_endpoint (:1618-1646): FE one Maxwell update; TRAP one
 * trapezoid pass from tau_n, the endpoint and report drift re-solved from the
 * new tau_{n+1}, converged once the surface drift settles (pass >= 2) */
void orc_resp_advance_endpoint(orc_resp* r, const orc_sht* s, const cplx* sigma) {
    if (!r->implicit || r->kind != RESP_VE) {
        orc_resp_advance(r, s, sigma);
        r->couple_pass = 1;
        r->couple_done = 1;
        return;
    }
    trapezoid_advance_all(r, s, sigma);
    memcpy(r->sigma_next, sigma, (size_t)r->nlm * sizeof(cplx));
    r->have_next = 1;
    resp_drift(r, r->edUn_re, r->edUn_im, r->edVn_re, r->edVn_im, r->edWn_re, r->edWn_im);
    r->couple_pass++;
    /* drift_change (:1499-1512), complex moduli; W(a) joins once carried */
    double cnorm = 0.0, snorm = 0.0;
    for (int k = 0; k < r->nk; k++) {
        cnorm = fmax(cnorm, cabs(r->dUa[k] - r->dUa_prev[k]));
        snorm = fmax(snorm, cabs(r->dUa[k]));
    }
    if (r->nlam > 4)
        for (int k = 0; k < r->nk; k++) {
            cnorm = fmax(cnorm, cabs(r->dWa[k] - r->dWa_prev[k]));
            snorm = fmax(snorm, cabs(r->dWa[k]));
        }
    memcpy(r->dUa_prev, r->dUa, (size_t)r->nk * sizeof(cplx));
    if (r->nlam > 4) memcpy(r->dWa_prev, r->dWa, (size_t)r->nk * sizeof(cplx));
    r->couple_done = r->couple_pass >= 2 && cnorm <= r->couple_tol * fmax(snorm, DBL_MIN);
}

/* ve_response_finalize_step (:1653-1663) */
void orc_resp_finalize_step(orc_resp* r) {
    if (r->implicit && r->have_next) {
        memcpy(r->sigma_n, r->sigma_next, (size_t)r->nlm * sizeof(cplx));
        r->sigma_primed = 1;
    }
    r->time += r->dt;
}

/* response_prime_sigma (:1838-1847) */
void orc_resp_prime_sigma(orc_resp* r, const cplx* sigma) {
    if (!r->sigma_n) return;
    memcpy(r->sigma_n, sigma, (size_t)r->nlm * sizeof(cplx));
    r->sigma_primed = 1;
}

/* response_stash_coarse (:1742-1757) */
void orc_resp_stash_coarse(orc_resp* r) {
    if (r->kind != RESP_VE) return;
    size_t n = (size_t)r->nlam * r->ne * r->nk;
    memcpy(r->Are_c, r->Are, n * 8); memcpy(r->Aim_c, r->Aim, n * 8); memcpy(r->Bre_c, r->Bre, n * 8);
    memcpy(r->Bim_c, r->Bim, n * 8); memcpy(r->Cre_c, r->Cre, n * 8); memcpy(r->Cim_c, r->Cim, n * 8);
}

/* response_coarse_fine_error (:1759-1788): max |tau_fine - tau_coarse| and
 * max |tau_fine| over every slot, element and coefficient (NaN kept) */
void orc_resp_coarse_fine_error(const orc_resp* r, double* err_inf, double* tau_inf) {
    *err_inf = 0.0; *tau_inf = 0.0;
    if (r->kind != RESP_VE) return;
    size_t n = (size_t)r->nlam * r->ne * r->nk;
    const double* f[6] = {r->Are, r->Aim, r->Bre, r->Bim, r->Cre, r->Cim};
    const double* c[6] = {r->Are_c, r->Aim_c, r->Bre_c, r->Bim_c, r->Cre_c, r->Cim_c};
    double e = 0.0, t = 0.0;
    for (int q = 0; q < 6; q++)
        for (size_t i = 0; i < n; i++) {
            e = nanmax(e, fabs(f[q][i] - c[q][i]));
            t = nanmax(t, fabs(f[q][i]));
        }
    *err_inf = e; *tau_inf = t;
}
