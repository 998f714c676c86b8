/*
 * orc_radial.c -- oracle: earth structure, radial mesh, P1 element integrals,
 * per-degree saddle-point operator, equilibrated (bordered) band LU.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Restates:
 *   src/vilma_earth_structure.f90:62-116, :168-238, :284-315
 *   src/vilma_radial_fe.f90:140-224 (mesh), :252-270 (R_k), :274-443 (operator),
 *     :447-605 (assemble/equilibrate/factor/solve), :616-645 (RHS)
 *   src/vilma_radial_integrals.f90:28-126 (element integrals)
 *   src/vilma_band.f90:28-142 (dgbtf2/dgbtrs-style band LU)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* ------------------------------------------------------------------ earth */

/* Incompressible PREM density / rigidity at radius r
 * (src/vilma_earth_structure.f90:118-183; Dziewonski & Anderson 1981, Table I). */
static void prem_rho_mu(double r, double* rho, double* mu) {
    double x = r / 6371.0e3, vs;
    if (r <= 1221.5e3) { *rho = 13088.5 - 8838.1 * x * x; *mu = 7.0363e10; return; }
    if (r <= 3480.0e3) { *rho = 12581.5 - 1263.8 * x - 3642.6 * x * x - 5528.1 * x * x * x; *mu = 7.0363e10; return; }
    if (r <= 3630.0e3) {
        vs = 6925.4 + 1467.2 * x - 2083.4 * x * x + 978.3 * x * x * x;
        *rho = 7956.5 - 6476.1 * x + 5528.3 * x * x - 3080.7 * x * x * x;
    } else if (r <= 5600.0e3) {
        vs = 11167.1 - 13781.8 * x + 17457.5 * x * x - 9277.7 * x * x * x;
        *rho = 7956.5 - 6476.1 * x + 5528.3 * x * x - 3080.7 * x * x * x;
    } else if (r <= 5701.0e3) {
        vs = 22345.9 - 17247.3 * x - 2083.4 * x * x + 978.3 * x * x * x;
        *rho = 7956.5 - 6476.1 * x + 5528.3 * x * x - 3080.7 * x * x * x;
    } else if (r <= 5771.0e3) {
        vs = 9983.9 - 4932.4 * x; *rho = 5319.7 - 1483.6 * x;
    } else if (r <= 5971.0e3) {
        vs = 22351.2 - 18585.6 * x; *rho = 11249.4 - 8029.8 * x;
    } else if (r <= 6151.0e3) {
        vs = 8949.6 - 4459.7 * x; *rho = 7108.9 - 3804.5 * x;
    } else if (r <= 6346.6e3) {
        vs = 2151.9 + 2348.1 * x; *rho = 2691.0 + 692.4 * x;
    } else if (r <= 6356.0e3) {
        vs = 3900.0; *rho = 2900.0;
    } else {
        vs = 3200.0; *rho = 2600.0;
    }
    *mu = vs * vs * (*rho);
}

/* r^2-weighted Simpson shell average (src/vilma_earth_structure.f90:185-218). */
static void prem_shell_avg(double rb, double rt, double* rho, double* mu) {
    const int n = 200;
    double h = (rt - rb) / n, nr = 0, nm = 0, den = 0;
    for (int i = 0; i <= n; i++) {
        double w = (i == 0 || i == n) ? 1.0 : ((i % 2) ? 4.0 : 2.0);
        double r = rb + i * h, ww = w * r * r, ri, mi;
        prem_rho_mu(r, &ri, &mi);
        nr += ww * ri; nm += ww * mi; den += ww;
    }
    *rho = nr / den; *mu = nm / den;
}

int orc_build_earth(const vh_params* p, orc_earth* e, char* err) {
    memset(e, 0, sizeof(*e));
    if (strcmp(p->earth, "M3-L70-V01") == 0) {
        /* src/vilma_earth_structure.f90:220-238 */
        const double km = 1.0e3;
        e->n = 5; e->r_earth = 6371.0 * km; e->r_core = 3480.0 * km;
        orc_layer L[5] = {
            {6301.0 * km, 6371.0 * km, 3037.0, 5.0605e10, 1.79769313486231571e308, RHEOL_ELASTIC},
            {5951.0 * km, 6301.0 * km, 3438.0, 7.0363e10, 1.0e21, RHEOL_MAXWELL},
            {5701.0 * km, 5951.0 * km, 3871.0, 1.0549e11, 1.0e21, RHEOL_MAXWELL},
            {3480.0 * km, 5701.0 * km, 4978.0, 2.2834e11, 2.0e21, RHEOL_MAXWELL},
            {0.0, 3480.0 * km, 10750.0, 0.0, 0.0, RHEOL_FLUID}};
        memcpy(e->L, L, sizeof(L));
        return 0;
    }
    int prem = strcmp(p->earth, "PREM") == 0;
    if (!prem && strcmp(p->earth, "custom") != 0) {
        snprintf(err, 256, "build_earth: unknown earth model \"%s\"", p->earth);
        return 1;
    }
    if (p->n_layer < 1 || p->n_layer > VH_MAX_LAYER) {
        snprintf(err, 256, "build_layered: n_layer out of range for a layered earth model");
        return 1;
    }
    e->n = p->n_layer; e->r_earth = p->r_earth; e->r_core = p->r_core;
    for (int k = 0; k < e->n; k++) {
        double rho = p->rho[k], mu = p->mu[k];
        if (prem) {
            prem_shell_avg(p->r_bot[k], p->r_top[k], &rho, &mu);
            if (p->rheology[k] == RHEOL_FLUID) mu = 0.0;
        }
        e->L[k] = (orc_layer){p->r_bot[k], p->r_top[k], rho, mu, p->eta[k], p->rheology[k]};
    }
    return 0;
}

static double mass_below(const orc_earth* e, double r) {
    double m = 0.0;
    for (int i = 0; i < e->n; i++) {
        double lo = e->L[i].r_bot, hi = fmin(e->L[i].r_top, r);
        if (hi > lo) m += (4.0 / 3.0) * ORC_PI * e->L[i].rho * (hi * hi * hi - lo * lo * lo);
    }
    return m;
}

/* g(r) = G M(<r)/r^2 (src/vilma_earth_structure.f90:309-315). */
double orc_gravity_at(const orc_earth* e, double r) {
    return (r > 0.0) ? ORC_G * mass_below(e, r) / (r * r) : 0.0;
}

/* ------------------------------------------------------------------- mesh */

static double dr_target(double depth) {
    if (depth <= 70.0e3) return 5.0e3;      /* src/vilma_radial_fe.f90:46-50 */
    if (depth <= 670.0e3) return 10.0e3;
    return 40.0e3;
}

static int layer_nelem(const orc_earth* e, int i, const int* n_elem_layer) {
    if (n_elem_layer && n_elem_layer[i] > 0) return n_elem_layer[i];
    double r0 = e->L[i].r_bot, r1 = e->L[i].r_top;
    double dr = dr_target(e->r_earth - 0.5 * (r0 + r1));
    int n = (int)ceil((r1 - r0) / dr);
    return n < 1 ? 1 : n;
}

/* Whole-sphere P1 mesh, nodes pinned to interfaces
 * (src/vilma_radial_fe.f90:153-224). n_elem_layer (optional) replaces the
 * depth-target count per layer by an explicit one. */
int orc_mesh_build(const orc_earth* e, const int* n_elem_layer, orc_mesh* m, char* err) {
    memset(m, 0, sizeof(*m));
    int n = e->n;
    if (fabs(e->L[0].r_top - e->r_earth) > 1.0) {
        snprintf(err, 256, "radial_mesh_build: r_top(1) must equal r_earth (layers are surface-first)"); return 1;
    }
    if (e->L[n - 1].r_bot != 0.0) {
        snprintf(err, 256, "radial_mesh_build: the innermost layer must reach r = 0"); return 1;
    }
    for (int i = 0; i < n; i++) {
        if (e->L[i].r_top <= e->L[i].r_bot) {
            snprintf(err, 256, "radial_mesh_build: every layer needs r_top > r_bot (surface-first ordering)"); return 1;
        }
        if (i < n - 1 && fabs(e->L[i].r_bot - e->L[i + 1].r_top) > 1.0) {
            snprintf(err, 256, "radial_mesh_build: layers must be contiguous, r_bot(k) = r_top(k+1)"); return 1;
        }
    }
    int ntot = 0;
    for (int i = 0; i < n; i++) ntot += layer_nelem(e, i, n_elem_layer);
    m->ne = ntot; m->nr = ntot + 1;
    m->r = (double*)calloc(m->nr, sizeof(double));
    m->lay = (int*)calloc(m->ne, sizeof(int));
    m->r[0] = 0.0;
    int off = 0;   /* index of the last node placed */
    for (int i = n - 1; i >= 0; i--) {
        double r0 = e->L[i].r_bot, r1 = e->L[i].r_top;
        int nl = layer_nelem(e, i, n_elem_layer);
        double h = (r1 - r0) / (double)nl;
        for (int k = 1; k <= nl; k++) {
            m->r[off + k] = r0 + (double)k * h;
            m->lay[off + k - 1] = i;
        }
        off += nl;
    }
    m->r[m->nr - 1] = e->r_earth;
    return 0;
}

void orc_mesh_free(orc_mesh* m) { free(m->r); free(m->lay); memset(m, 0, sizeof(*m)); }

/* ------------------------------------------------------ element integrals */
/* Martinec (2000) App. C closed forms (src/vilma_radial_integrals.f90:28-126).
 * M[a][b], a,b in {0,1} the local nodes (k, k+1). */

static void el_i1(double rk, double rk1, double m[2][2]) {
    double h = rk1 - rk, d = (rk1 * rk1 + rk1 * rk + rk * rk) / (3.0 * h);
    m[0][0] = d; m[0][1] = -d; m[1][0] = -d; m[1][1] = d;
}
static void el_i2(double rk, double rk1, double m[2][2]) {
    m[0][0] = -(rk1 * rk1 + 2.0 * rk1 * rk + 3.0 * rk * rk) / 12.0;
    m[0][1] = -m[0][0];
    m[1][1] = (3.0 * rk1 * rk1 + 2.0 * rk1 * rk + rk * rk) / 12.0;
    m[1][0] = -m[1][1];
}
static void el_i3(double rk, double rk1, double m[2][2]) {
    m[0][0] = -(rk1 + 2.0 * rk) / 6.0;
    m[0][1] = -m[0][0];
    m[1][1] = (2.0 * rk1 + rk) / 6.0;
    m[1][0] = -m[1][1];
}
static void el_i4(double rk, double rk1, double m[2][2]) {
    double h = rk1 - rk;
    m[0][0] = h / 30.0 * (rk1 * rk1 + 3.0 * rk1 * rk + 6.0 * rk * rk);
    m[0][1] = h / 60.0 * (3.0 * rk1 * rk1 + 4.0 * rk1 * rk + 3.0 * rk * rk);
    m[1][0] = m[0][1];
    m[1][1] = h / 30.0 * (6.0 * rk1 * rk1 + 3.0 * rk1 * rk + rk * rk);
}
static void el_i5(double rk, double rk1, double m[2][2]) {
    double h = rk1 - rk;
    m[0][0] = h / 12.0 * (rk1 + 3.0 * rk);
    m[0][1] = h / 12.0 * (rk1 + rk);
    m[1][0] = m[0][1];
    m[1][1] = h / 12.0 * (3.0 * rk1 + rk);
}
static void el_i6(double rk, double rk1, double m[2][2]) {
    double h = rk1 - rk;
    m[0][0] = h / 3.0; m[0][1] = h / 6.0; m[1][0] = h / 6.0; m[1][1] = h / 3.0;
}
static void el_i7(double rk, double rk1, double m[2][2]) {
    double h = rk1 - rk, lg = log(rk1 / rk);
    m[0][0] = (rk1 * rk1 / h * lg - 2.0 * rk1 + 0.5 * (rk1 + rk)) / h;
    m[0][1] = (-rk1 * rk / h * lg + 0.5 * (rk1 + rk)) / h;
    m[1][0] = m[0][1];
    m[1][1] = (rk * rk / h * lg - 2.0 * rk + 0.5 * (rk1 + rk)) / h;
}
static void el_k1(double rk, double rk1, double v[2]) {
    v[0] = -(rk1 * rk1 + rk1 * rk + rk * rk) / 3.0; v[1] = -v[0];
}
static void el_k2(double rk, double rk1, double v[2]) {
    double h = rk1 - rk;
    v[0] = h / 6.0 * (rk1 + 2.0 * rk); v[1] = h / 6.0 * (2.0 * rk1 + rk);
}
static void el_k3(double rk, double rk1, double v[2]) {
    double h = rk1 - rk;
    v[0] = h / 12.0 * (rk1 * rk1 + 2.0 * rk1 * rk + 3.0 * rk * rk);
    v[1] = h / 12.0 * (3.0 * rk1 * rk1 + 2.0 * rk1 * rk + rk * rk);
}

/* exported for the integral unit tests */
void orc_elem_integrals(double rk, double rk1, double* out /* 7*4 + 3*2 */) {
    double m[2][2], v[2];
    void (*f[7])(double, double, double[2][2]) = {el_i1, el_i2, el_i3, el_i4, el_i5, el_i6, el_i7};
    for (int q = 0; q < 7; q++) {
        if (q == 6 && rk == 0.0) { memset(out + 24, 0, 4 * sizeof(double)); continue; }
        f[q](rk, rk1, m);
        out[4 * q + 0] = m[0][0]; out[4 * q + 1] = m[0][1];
        out[4 * q + 2] = m[1][0]; out[4 * q + 3] = m[1][1];
    }
    el_k1(rk, rk1, v); out[28] = v[0]; out[29] = v[1];
    el_k2(rk, rk1, v); out[30] = v[0]; out[31] = v[1];
    el_k3(rk, rk1, v); out[32] = v[0]; out[33] = v[1];
}

/* ----------------------------------------------------------- operator */

/* R_k per element (src/vilma_radial_fe.f90:252-270). */
static void shell_Rk(const orc_earth* e, const orc_mesh* m, double* Rk) {
    Rk[0] = 0.0;
    for (int k = 1; k < m->ne; k++) {
        double rho_k = e->L[m->lay[k]].rho, rho_km1 = e->L[m->lay[k - 1]].rho;
        Rk[k] = Rk[k - 1] + (rho_km1 - rho_k) * m->r[k] * m->r[k] * m->r[k];
    }
}

/* Degree-1 rigid-translation weight (src/vilma_radial_fe.f90:425-443). */
static void uniq_weight(const orc_mesh* m, double* w) {
    int nd = 4 * m->nr - 1;
    memset(w, 0, nd * sizeof(double));
    double k3[2];
    for (int e = 0; e < m->ne; e++) {
        el_k3(m->r[e], m->r[e + 1], k3);
        w[idx_u(e)] += k3[0];
        w[idx_u(e + 1)] += k3[1];
        w[idx_v(e)] += 2.0 * k3[0];
        w[idx_v(e + 1)] += 2.0 * k3[1];
    }
}

/* Dense per-degree operator, column-major nd x nd
 * (src/vilma_radial_fe.f90:274-423). Local dof order [U1 V1 F1 U2 V2 F2 P]. */
double* orc_build_dense_operator(const orc_earth* e, const orc_mesh* m, int j, int with_uniq) {
    int nr = m->nr, ne = m->ne, nd = 4 * nr - 1;
    double Jr = (double)j * (double)(j + 1);
    double fourpiG = 4.0 * ORC_PI * ORC_G;
    double* A = (double*)calloc((size_t)nd * nd, sizeof(double));
    double* Rk = (double*)calloc(ne, sizeof(double));
    shell_Rk(e, m, Rk);
    const int lU[2] = {0, 3}, lV[2] = {1, 4}, lF[2] = {2, 5}, lP = 6;
    for (int el = 0; el < ne; el++) {
        double rlo = m->r[el], rhi = m->r[el + 1];
        double mu_k = e->L[m->lay[el]].mu, rho_k = e->L[m->lay[el]].rho;
        double i1[2][2], i2[2][2], i3[2][2], i4[2][2], i5[2][2], i6[2][2], i7[2][2], k1[2], k2[2];
        el_i1(rlo, rhi, i1); el_i2(rlo, rhi, i2); el_i3(rlo, rhi, i3);
        el_i4(rlo, rhi, i4); el_i5(rlo, rhi, i5); el_i6(rlo, rhi, i6);
        if (Rk[el] != 0.0) el_i7(rlo, rhi, i7); else memset(i7, 0, sizeof(i7));
        el_k1(rlo, rhi, k1); el_k2(rlo, rhi, k2);
        double Al[7][7];
        memset(Al, 0, sizeof(Al));
#define AL(r, c) Al[r][c]
        for (int ia = 0; ia < 2; ia++) {
            for (int ib = 0; ib < 2; ib++) {
                /* shear (eq 80), W dropped */
                AL(lU[ib], lU[ia]) += mu_k * (2.0 * i1[ia][ib]);
                AL(lV[ib], lV[ia]) += mu_k * (Jr * i1[ia][ib]);
                AL(lV[ib], lU[ia]) += mu_k * (Jr * i3[ia][ib]);
                AL(lV[ib], lV[ia]) -= mu_k * (Jr * i3[ia][ib]);
                AL(lV[ib], lV[ia]) -= mu_k * (Jr * i3[ib][ia]);
                AL(lU[ib], lV[ia]) += mu_k * (Jr * i3[ib][ia]);
                AL(lU[ib], lU[ia]) += mu_k * (Jr * i6[ia][ib]);
                AL(lU[ib], lV[ia]) -= mu_k * (Jr * i6[ia][ib]);
                AL(lV[ib], lU[ia]) -= mu_k * (Jr * i6[ia][ib]);
                AL(lV[ib], lV[ia]) += mu_k * (Jr * i6[ia][ib]);
                AL(lU[ib], lU[ia]) += mu_k * (4.0 * i6[ia][ib]);
                AL(lV[ib], lU[ia]) -= mu_k * (2.0 * Jr * i6[ia][ib]);
                AL(lU[ib], lV[ia]) -= mu_k * (2.0 * Jr * i6[ia][ib]);
                AL(lV[ib], lV[ia]) += mu_k * (Jr * Jr * i6[ia][ib]);
                AL(lV[ib], lV[ia]) += mu_k * (Jr * (Jr - 2.0) * i6[ia][ib]);
                /* self-gravity (eq 81) */
                double gg = (fourpiG / 3.0) * (rho_k * i4[ia][ib] + Rk[el] * i7[ia][ib]);
                AL(lU[ib], lU[ia]) += rho_k * (-4.0 * gg + fourpiG * rho_k * i4[ia][ib]);
                AL(lU[ib], lV[ia]) += rho_k * (Jr * gg);
                AL(lU[ib], lF[ia]) += rho_k * (i2[ib][ia]);
                AL(lV[ib], lU[ia]) += rho_k * (Jr * gg);
                AL(lV[ib], lF[ia]) += rho_k * (Jr * i5[ia][ib]);
                AL(lF[ib], lF[ia]) += (i1[ia][ib] + Jr * i6[ia][ib]) / fourpiG;
                AL(lF[ib], lU[ia]) += rho_k * (i2[ia][ib]);
                AL(lF[ib], lV[ia]) += rho_k * (Jr * i5[ia][ib]);
            }
        }
        /* pressure coupling (eq 82) */
        for (int ia = 0; ia < 2; ia++) {
            AL(lP, lU[ia]) += (k1[ia] + 2.0 * k2[ia]);
            AL(lU[ia], lP) += (k1[ia] + 2.0 * k2[ia]);
            AL(lP, lV[ia]) -= (Jr * k2[ia]);
            AL(lV[ia], lP) -= (Jr * k2[ia]);
        }
#undef AL
        int g[7] = {idx_u(el), idx_v(el), idx_f(el), idx_u(el + 1), idx_v(el + 1), idx_f(el + 1), idx_p(el)};
        for (int c = 0; c < 7; c++)
            for (int r = 0; r < 7; r++) A[g[r] + (size_t)g[c] * nd] += Al[r][c];
    }
    /* surface exterior-potential match (eq 84) */
    A[idx_f(nr - 1) + (size_t)idx_f(nr - 1) * nd] += e->r_earth / fourpiG * (double)(j + 1);
    if (j == 1 && with_uniq) {
        double* w = (double*)malloc(nd * sizeof(double));
        uniq_weight(m, w);
        for (int ib = 0; ib < nd; ib++) {
            if (w[ib] == 0.0) continue;
            for (int ia = 0; ia < nd; ia++) {
                if (w[ia] == 0.0) continue;
                A[ib + (size_t)ia * nd] += (4.0 * ORC_PI / 3.0) * w[ib] * w[ia];
            }
        }
        free(w);
    }
    free(Rk);
    return A;
}

/* ------------------------------------------------------------- band LU */

int orc_band_build(orc_band* b, int n, int nz, const int* rows, const int* cols, const double* vals) {
    int kl = 0, ku = 0;
    for (int t = 0; t < nz; t++) {
        if (vals[t] == 0.0) continue;
        int d = rows[t] - cols[t];
        if (d > kl) kl = d;
        if (-d > ku) ku = -d;
    }
    b->n = n; b->kl = kl; b->ku = ku; b->kv = kl + ku; b->ldab = 2 * kl + ku + 1;
    b->ab = (double*)calloc((size_t)b->ldab * n, sizeof(double));
    b->ipiv = (int*)calloc(n, sizeof(int));
    int kv = b->kv, ldab = b->ldab;
    double* ab = b->ab;
    for (int t = 0; t < nz; t++) {
        if (vals[t] == 0.0) continue;
        ab[(kv + rows[t] - cols[t]) + (size_t)cols[t] * ldab] += vals[t];
    }
    /* unblocked pivoted band LU (src/vilma_band.f90:61-109) */
    int ju = 0;
#define AB(r, c) ab[(r) + (size_t)(c) * ldab]
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = 0; i < kl; i++) AB(i, j + kv) = 0.0;
        int km = kl < (n - 1 - j) ? kl : (n - 1 - j);
        int jp = 0;
        double piv = fabs(AB(kv, j));
        for (int p = 1; p <= km; p++)
            if (fabs(AB(kv + p, j)) > piv) { piv = fabs(AB(kv + p, j)); jp = p; }
        b->ipiv[j] = j + jp;
        if (AB(kv + jp, j) == 0.0) return 1;
        int jmax = j + ku + jp < n - 1 ? j + ku + jp : n - 1;
        if (jmax > ju) ju = jmax;
        if (jp != 0) {
            for (int jj = j; jj <= ju; jj++) {
                double t = AB(kv + j - jj, jj);
                AB(kv + j - jj, jj) = AB(kv + j + jp - jj, jj);
                AB(kv + j + jp - jj, jj) = t;
            }
        }
        if (km > 0) {
            double rp = 1.0 / AB(kv, j);
            for (int p = 1; p <= km; p++) AB(kv + p, j) *= rp;
            for (int jj = j + 1; jj <= ju; jj++) {
                double t = AB(kv + j - jj, jj);
                if (t != 0.0)
                    for (int p = 1; p <= km; p++) AB(kv + j + p - jj, jj) -= AB(kv + p, j) * t;
            }
        }
    }
#undef AB
    return 0;
}

/* dgbtrs, no transpose (src/vilma_band.f90:111-142). */
void orc_band_solve(const orc_band* b, const double* rhs, double* x) {
    int n = b->n, kl = b->kl, kv = b->kv, ldab = b->ldab;
    const double* ab = b->ab;
    if (x != rhs) memcpy(x, rhs, n * sizeof(double));
    for (int j = 0; j < n - 1; j++) {
        int lm = kl < (n - 1 - j) ? kl : (n - 1 - j);
        int l = b->ipiv[j];
        if (l != j) { double t = x[l]; x[l] = x[j]; x[j] = t; }
        double t = x[j];
        for (int i = 1; i <= lm; i++) x[j + i] -= ab[(kv + i) + (size_t)j * ldab] * t;
    }
    for (int j = n - 1; j >= 0; j--) {
        x[j] = x[j] / ab[kv + (size_t)j * ldab];
        double t = x[j];
        int lm = kv < j ? kv : j;
        for (int i = 1; i <= lm; i++) x[j - i] -= ab[(kv - i) + (size_t)j * ldab] * t;
    }
}

void orc_band_free(orc_band* b) { free(b->ab); free(b->ipiv); memset(b, 0, sizeof(*b)); }

/* ------------------------------------------------- bordered band system */

static double colnorm(const double* col, int n, int stride) {
    double m = 0.0;
    for (int i = 0; i < n; i++) { double a = fabs(col[(size_t)i * stride]); if (a > m) m = a; }
    return m > 0.0 ? 1.0 / sqrt(m) : 1.0;
}

/* Geometric-mean equilibration + band factor, optional KKT border of nb
 * columns w (src/vilma_radial_fe.f90:497-576). A is dense column-major. */
static int bband_factor(orc_bband* s, const double* A, int nd, const double* w, int nb) {
    s->nd = nd; s->nb = nb; s->ns = nd + nb;
    s->w = (double*)calloc((size_t)nd * (nb > 0 ? nb : 1), sizeof(double));
    if (nb > 0) memcpy(s->w, w, (size_t)nd * nb * sizeof(double));
    s->dc = (double*)calloc(nd, sizeof(double));
    s->dr = (double*)calloc(nd, sizeof(double));
    double dc_b[4], dr_b[4];
    for (int k = 0; k < nd; k++) {
        s->dc[k] = colnorm(A + (size_t)k * nd, nd, 1);
        for (int b = 0; b < nb; b++)
            if (fabs(s->w[k + b * nd]) > 1.0 / (s->dc[k] * s->dc[k])) s->dc[k] = 1.0 / sqrt(fabs(s->w[k + b * nd]));
    }
    for (int b = 0; b < nb; b++) dc_b[b] = colnorm(s->w + (size_t)b * nd, nd, 1);
    for (int i = 0; i < nd; i++) {
        double m = 0.0;
        for (int k = 0; k < nd; k++) { double a = fabs(A[i + (size_t)k * nd] * s->dc[k]); if (a > m) m = a; }
        s->dr[i] = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
        for (int b = 0; b < nb; b++)
            if (fabs(s->w[i + b * nd]) * dc_b[b] > 1.0 / (s->dr[i] * s->dr[i]))
                s->dr[i] = 1.0 / sqrt(fabs(s->w[i + b * nd]) * dc_b[b]);
    }
    for (int b = 0; b < nb; b++) {
        double m = 0.0;
        for (int k = 0; k < nd; k++) { double a = fabs(s->w[k + b * nd] * s->dc[k]); if (a > m) m = a; }
        dr_b[b] = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
    }
    int nnz = 0;
    for (size_t t = 0; t < (size_t)nd * nd; t++) nnz += (A[t] != 0.0);
    for (int t = 0; t < nd * nb; t++) nnz += 2 * (s->w[t] != 0.0);
    int* rows = (int*)malloc(nnz * sizeof(int));
    int* cols = (int*)malloc(nnz * sizeof(int));
    double* vals = (double*)malloc(nnz * sizeof(double));
    int p = 0;
    for (int k = 0; k < nd; k++)
        for (int i = 0; i < nd; i++) {
            double a = A[i + (size_t)k * nd];
            if (a == 0.0) continue;
            rows[p] = i; cols[p] = k; vals[p] = s->dr[i] * a * s->dc[k]; p++;
        }
    for (int b = 0; b < nb; b++) {
        const double* wb = s->w + (size_t)b * nd;
        for (int i = 0; i < nd; i++) {
            if (wb[i] == 0.0) continue;
            rows[p] = i; cols[p] = nd + b; vals[p] = s->dr[i] * wb[i] * dc_b[b]; p++;
        }
        for (int k = 0; k < nd; k++) {
            if (wb[k] == 0.0) continue;
            rows[p] = nd + b; cols[p] = k; vals[p] = dr_b[b] * wb[k] * s->dc[k]; p++;
        }
    }
    int rc = orc_band_build(&s->band, s->ns, p, rows, cols, vals);
    free(rows); free(cols); free(vals);
    return rc;
}

/* src/vilma_radial_fe.f90:578-605 */
static void bband_solve(const orc_bband* s, const double* b, double* x, const double* border) {
    double* bs = (double*)calloc(s->ns, sizeof(double));
    double* y = (double*)malloc(s->ns * sizeof(double));
    for (int i = 0; i < s->nd; i++) bs[i] = s->dr[i] * b[i];
    if (border) for (int i = 0; i < s->nb; i++) bs[s->nd + i] = border[i];
    orc_band_solve(&s->band, bs, y);
    for (int i = 0; i < s->nd; i++) x[i] = s->dc[i] * y[i];
    free(bs); free(y);
}

static void bband_free(orc_bband* s) {
    orc_band_free(&s->band); free(s->dr); free(s->dc); free(s->w);
    memset(s, 0, sizeof(*s));
}

/* src/vilma_radial_fe.f90:447-495 */
int orc_radop_assemble(orc_radop* op, const orc_earth* e, const orc_mesh* m, int j, char* err) {
    memset(op, 0, sizeof(*op));
    int nd = 4 * m->nr - 1;
    double* A = orc_build_dense_operator(e, m, j, 0);
    op->j = j; op->nr = m->nr; op->ne = m->ne; op->ndof = nd;
    op->r_earth = e->r_earth; op->g_surf = orc_gravity_at(e, e->r_earth);
    int rc;
    if (j == 1) {
        double* w = (double*)malloc(nd * sizeof(double));
        uniq_weight(m, w);
        rc = bband_factor(&op->sys, A, nd, w, 1);
        free(w);
    } else {
        rc = bband_factor(&op->sys, A, nd, NULL, 0);
    }
    free(A);
    if (rc) { snprintf(err, 256, "bordered_band_factor: band LU factorization failed"); return 1; }
    if (op->sys.nb > 0) {
        double* zero = (double*)calloc(nd, sizeof(double));
        op->nullmode = (double*)malloc(nd * sizeof(double));
        double one = 1.0;
        bband_solve(&op->sys, zero, op->nullmode, &one);
        free(zero);
    }
    return 0;
}

void orc_radop_solve(const orc_radop* op, const double* b, double* x, const double* border) {
    bband_solve(&op->sys, b, x, (border && op->sys.nb > 0) ? border : NULL);
}

/* src/vilma_radial_fe.f90:616-625 */
void orc_radop_load_rhs(const orc_radop* op, double sigma, double* b) {
    memset(b, 0, op->ndof * sizeof(double));
    b[idx_u(op->nr - 1)] = -op->r_earth * op->r_earth * sigma * op->g_surf;
    b[idx_f(op->nr - 1)] = -op->r_earth * op->r_earth * sigma;
}

/* src/vilma_radial_fe.f90:627-645 */
void orc_radop_tidal_rhs(const orc_radop* op, double phi_t, double* b) {
    memset(b, 0, op->ndof * sizeof(double));
    b[idx_f(op->nr - 1)] = -op->r_earth / (4.0 * ORC_PI * ORC_G) * (double)(2 * op->j + 1) * phi_t;
}

void orc_radop_free(orc_radop* op) { bband_free(&op->sys); free(op->nullmode); memset(op, 0, sizeof(*op)); }

/* ---------------------------------------------------- Maxwell kernels */

/* src/vilma_viscoelastic.f90:329-359 */
void orc_strain_coeffs(double u1, double u2, double v1, double v2, double Jr,
                       double a[4], double b[4], double c[4]) {
    a[0] = -u1 + u2; b[0] = 0.0; c[0] = 0.0;
    a[1] = -v1 + v2; b[1] = u1 - v1; c[1] = u2 - v2;
    a[2] = 0.0; b[2] = -u1 / Jr + 0.5 * v1; c[2] = -u2 / Jr + 0.5 * v2;
    a[3] = 0.0; b[3] = 0.5 * v1; c[3] = 0.5 * v2;
    if (Jr <= 2.0) { b[3] = 0.0; c[3] = 0.0; }
}

/* src/vilma_viscoelastic.f90:361-371; sa[t][lam] is the strain of unit test dof t */
void orc_strain_constants(double Jr, double norm[4], double sa[4][4], double sb[4][4], double sc[4][4]) {
    norm[0] = 1.0; norm[1] = 0.5 * Jr; norm[2] = 2.0 * Jr * Jr; norm[3] = 2.0 * Jr * (Jr - 2.0);
    orc_strain_coeffs(1, 0, 0, 0, Jr, sa[0], sb[0], sc[0]);
    orc_strain_coeffs(0, 1, 0, 0, Jr, sa[1], sb[1], sc[1]);
    orc_strain_coeffs(0, 0, 1, 0, Jr, sa[2], sb[2], sc[2]);
    orc_strain_coeffs(0, 0, 0, 1, Jr, sa[3], sb[3], sc[3]);
}

/* -int tau:delta eps, 2-point radial Gauss (src/vilma_viscoelastic.f90:441-480).
 * A/B/C are [e][ld] (ld = 4, or 6 with the toroidal channels appended). */
void orc_dissipative_rhs(int ne, const double* r, double sa[4][4], double sb[4][4], double sc[4][4],
                         const double norm[4], const double* A, const double* B, const double* C,
                         int ld, double* f) {
    const double xg = 0.5773502691896257;
    const double gp[2] = {-xg, xg};
    for (int e = 0; e < ne; e++) {
        double rk = r[e], rk1 = r[e + 1], h = rk1 - rk;
        double floc[4] = {0, 0, 0, 0};
        for (int ig = 0; ig < 2; ig++) {
            double ra = 0.5 * (h * gp[ig] + rk + rk1);
            double psik = (rk1 - ra) / h, psik1 = (ra - rk) / h;
            double tauV[4];
            for (int m = 0; m < 4; m++)
                tauV[m] = A[e * ld + m] / h + B[e * ld + m] * psik / ra + C[e * ld + m] * psik1 / ra;
            for (int t = 0; t < 4; t++) {
                double D = 0.0;
                for (int m = 0; m < 4; m++) {
                    double deps = sa[t][m] / h + sb[t][m] * psik / ra + sc[t][m] * psik1 / ra;
                    D = D + norm[m] * tauV[m] * deps;
                }
                floc[t] = floc[t] - D * ra * ra * h * 0.5;
            }
        }
        f[idx_u(e)] += floc[0];
        f[idx_u(e + 1)] += floc[1];
        f[idx_v(e)] += floc[2];
        f[idx_v(e + 1)] += floc[3];
    }
}

/* forward-Euler memory update (src/vilma_viscoelastic.f90:482-589, FE at :577-584) */
void orc_advance_memory_fe(int ne, const double* mu, const double* Mk, const double* Un,
                           const double* Vn, double Jr, double* A, double* B, double* C, int ld,
                           const unsigned char* active) {
    for (int e = 0; e < ne; e++) {
        if (active && !active[e]) continue;
        double a[4], b[4], c[4];
        orc_strain_coeffs(Un[e], Un[e + 1], Vn[e], Vn[e + 1], Jr, a, b, c);
        double om = 1.0 - Mk[e], two_muM = 2.0 * mu[e] * Mk[e];
        for (int m = 0; m < 4; m++) {
            A[e * ld + m] = om * A[e * ld + m] - two_muM * a[m];
            B[e * ld + m] = om * B[e * ld + m] - two_muM * b[m];
            C[e * ld + m] = om * C[e * ld + m] - two_muM * c[m];
        }
    }
}

/* advance_memory, trapezoidal rule (src/vilma_viscoelastic.f90:551-565):
 * tau+ = [(1 - M/2) tau_n - mu M (eps_n + eps_{n+1})] / (1 + M/2); Un/Vn the
 * endpoint nodal fields, Up/Vp the start-of-step ones. A/B/C hold tau_n on
 * entry. */
void orc_advance_memory_trap(int ne, const double* mu, const double* Mk, const double* Un, const double* Vn,
                             const double* Up, const double* Vp, double Jr, double* A, double* B, double* C,
                             int ld, const unsigned char* active) {
    for (int e = 0; e < ne; e++) {
        if (active && !active[e]) continue;
        double a[4], b[4], c[4], ap[4], bp[4], cp[4];
        orc_strain_coeffs(Un[e], Un[e + 1], Vn[e], Vn[e + 1], Jr, a, b, c);
        orc_strain_coeffs(Up[e], Up[e + 1], Vp[e], Vp[e + 1], Jr, ap, bp, cp);
        double den = 1.0 + 0.5 * Mk[e];
        double c_old = (1.0 - 0.5 * Mk[e]) / den, w_eps = mu[e] * Mk[e] / den;
        for (int m = 0; m < 4; m++) {
            A[e * ld + m] = c_old * A[e * ld + m] - w_eps * (a[m] + ap[m]);
            B[e * ld + m] = c_old * B[e * ld + m] - w_eps * (b[m] + bp[m]);
            C[e * ld + m] = c_old * C[e * ld + m] - w_eps * (c[m] + cp[m]);
        }
    }
}

/* ------------------------------------------------ toroidal channels */

/* int psi_a r^3 dr (src/vilma_radial_integrals.f90:117-126) */
static void el_k4(double rk, double rk1, double v[2]) {
    double h = rk1 - rk;
    v[0] = h / 20.0 * (rk1 * rk1 * rk1 + 2.0 * rk1 * rk1 * rk + 3.0 * rk1 * rk * rk + 4.0 * rk * rk * rk);
    v[1] = h / 20.0 * (4.0 * rk1 * rk1 * rk1 + 3.0 * rk1 * rk1 * rk + 2.0 * rk1 * rk * rk + rk * rk * rk);
}

/* build_toroidal_operator (src/vilma_radial_fe.f90:699-726): the eq-80 W
 * block of degree j, dense column-major nr x nr,
 *   mu_k { J [I1 - I3(b,a) - I3(a,b) + I6] + J (J - 2) I6 } W^a dW^b */
double* orc_build_tor_operator(const double* r, const double* mu, int ne, int j) {
    int nr = ne + 1;
    double Jr = (double)j * (double)(j + 1);
    double* A = calloc((size_t)nr * nr, sizeof(double));
    for (int e = 0; e < ne; e++) {
        if (mu[e] == 0.0) continue;
        double i1[2][2], i3[2][2], i6[2][2];
        el_i1(r[e], r[e + 1], i1); el_i3(r[e], r[e + 1], i3); el_i6(r[e], r[e + 1], i6);
        int gmap[2] = {e, e + 1};
        for (int ia = 0; ia < 2; ia++)
            for (int ib = 0; ib < 2; ib++)
                A[gmap[ib] + (size_t)gmap[ia] * nr] += mu[e] * (Jr * (i1[ia][ib] - i3[ib][ia] - i3[ia][ib] + i6[ia][ib])
                                                                + Jr * (Jr - 2.0) * i6[ia][ib]);
    }
    return A;
}

/* toroidal_operator_assemble (src/vilma_radial_fe.f90:773-798): pin the nodes
 * all of whose elements are fluid (:728-740), border degree 1 by one
 * net-rotation constraint per solid shell (:742-771), factor once. */
int orc_torop_assemble(orc_torop* op, const double* r, const double* mu, int ne, int j, char* err) {
    memset(op, 0, sizeof(*op));
    int nr = ne + 1;
    op->j = j; op->nr = nr;
    double* A = orc_build_tor_operator(r, mu, ne, j);
    op->pinned = malloc(nr);
    memset(op->pinned, 1, nr);
    for (int e = 0; e < ne; e++)
        if (mu[e] > 0.0) { op->pinned[e] = 0; op->pinned[e + 1] = 0; }
    for (int k = 0; k < nr; k++)
        if (op->pinned[k]) A[k + (size_t)k * nr] = 1.0;
    int rc;
    if (j == 1) {
        int* shell = calloc(ne, sizeof(int));
        int ns = 0, prev = 0;
        for (int e = 0; e < ne; e++) {
            int solid = mu[e] > 0.0;
            if (solid && !prev) ns++;
            if (solid) shell[e] = ns;
            prev = solid;
        }
        if (ns > 4) { free(shell); free(A); snprintf(err, 256, "toroidal_operator_assemble: more than 4 solid shells"); return 1; }
        double* w = calloc((size_t)nr * (ns > 0 ? ns : 1), sizeof(double));
        for (int e = 0; e < ne; e++) {
            if (!shell[e]) continue;
            double k4[2];
            el_k4(r[e], r[e + 1], k4);
            w[e + (size_t)(shell[e] - 1) * nr] += k4[0];
            w[e + 1 + (size_t)(shell[e] - 1) * nr] += k4[1];
        }
        rc = bband_factor(&op->sys, A, nr, w, ns);
        free(w); free(shell);
    } else {
        rc = bband_factor(&op->sys, A, nr, NULL, 0);
    }
    free(A);
    if (rc) { snprintf(err, 256, "bordered_band_factor: band LU factorization failed"); return 1; }
    return 0;
}

/* toroidal_operator_solve_vec (src/vilma_radial_fe.f90:800-810): pinned dofs
 * take zero whatever b holds there */
void orc_torop_solve(const orc_torop* op, const double* b, double* x) {
    double* bd = malloc(op->nr * sizeof(double));
    for (int k = 0; k < op->nr; k++) bd[k] = op->pinned[k] ? 0.0 : b[k];
    bband_solve(&op->sys, bd, x, NULL);
    free(bd);
}

void orc_torop_free(orc_torop* op) { bband_free(&op->sys); free(op->pinned); memset(op, 0, sizeof(*op)); }

/* strain_coeffs_tor (src/vilma_viscoelastic.f90:373-393): lambda = 3, 4 */
void orc_strain_coeffs_tor(double w1, double w2, double Jr, double a[2], double b[2], double c[2]) {
    a[0] = -w1 + w2; b[0] = -w1; c[0] = -w2;
    a[1] = 0.0; b[1] = w1; c[1] = w2;
    if (Jr <= 2.0) { b[1] = 0.0; c[1] = 0.0; }
}

/* ve_strain_constants_tor (src/vilma_viscoelastic.f90:395-405) */
void orc_strain_constants_tor(double Jr, double norm[2], double sa[2][2], double sb[2][2], double sc[2][2]) {
    norm[0] = 0.5 * Jr; norm[1] = 0.5 * Jr * (Jr - 2.0);
    orc_strain_coeffs_tor(1.0, 0.0, Jr, sa[0], sb[0], sc[0]);
    orc_strain_coeffs_tor(0.0, 1.0, Jr, sa[1], sb[1], sc[1]);
}

/* dissipative_rhs_tor (src/vilma_viscoelastic.f90:407-439): the forcing on
 * the nodal W test dofs from channels 4, 5 of the 6-channel memory
 * A/B/C [e][6]; f has nr entries */
void orc_dissipative_rhs_tor(int ne, const double* r, double sa[2][2], double sb[2][2], double sc[2][2],
                             const double norm[2], const double* A, const double* B, const double* C, double* f) {
    const double xg = 0.5773502691896257;
    const double gp[2] = {-xg, xg};
    for (int e = 0; e < ne; e++) {
        double rk = r[e], rk1 = r[e + 1], h = rk1 - rk;
        double floc[2] = {0, 0};
        for (int ig = 0; ig < 2; ig++) {
            double ra = 0.5 * (h * gp[ig] + rk + rk1);
            double psik = (rk1 - ra) / h, psik1 = (ra - rk) / h;
            double tauV[2];
            for (int m = 0; m < 2; m++)
                tauV[m] = A[e * 6 + 4 + m] / h + B[e * 6 + 4 + m] * psik / ra + C[e * 6 + 4 + m] * psik1 / ra;
            for (int t = 0; t < 2; t++) {
                double D = 0.0;
                for (int m = 0; m < 2; m++) {
                    double deps = sa[t][m] / h + sb[t][m] * psik / ra + sc[t][m] * psik1 / ra;
                    D = D + norm[m] * tauV[m] * deps;
                }
                floc[t] = floc[t] - D * ra * ra * h * 0.5;
            }
        }
        f[e] += floc[0];
        f[e + 1] += floc[1];
    }
}

/* advance_memory_tor (src/vilma_viscoelastic.f90:591-629) for the forward-
 * Euler (Wp == NULL) and trapezoidal (Wp = start-of-step W) rules, channels
 * 4, 5 of A/B/C [e][6] */
void orc_advance_memory_tor(int ne, const double* mu, const double* Mk, const double* Wn, const double* Wp,
                            double Jr, double* A, double* B, double* C, const unsigned char* active) {
    for (int e = 0; e < ne; e++) {
        if (active && !active[e]) continue;
        double a[2], b[2], c[2], ap[2], bp[2], cp[2];
        orc_strain_coeffs_tor(Wn[e], Wn[e + 1], Jr, a, b, c);
        double* At = A + e * 6 + 4; double* Bt = B + e * 6 + 4; double* Ct = C + e * 6 + 4;
        if (Wp) {
            orc_strain_coeffs_tor(Wp[e], Wp[e + 1], Jr, ap, bp, cp);
            double den = 1.0 + 0.5 * Mk[e];
            double c_old = (1.0 - 0.5 * Mk[e]) / den, w_eps = mu[e] * Mk[e] / den;
            for (int m = 0; m < 2; m++) {
                At[m] = c_old * At[m] - w_eps * (a[m] + ap[m]);
                Bt[m] = c_old * Bt[m] - w_eps * (b[m] + bp[m]);
                Ct[m] = c_old * Ct[m] - w_eps * (c[m] + cp[m]);
            }
        } else {
            double om = 1.0 - Mk[e], two_muM = 2.0 * mu[e] * Mk[e];
            for (int m = 0; m < 2; m++) {
                At[m] = om * At[m] - two_muM * a[m];
                Bt[m] = om * Bt[m] - two_muM * b[m];
                Ct[m] = om * Ct[m] - two_muM * c[m];
            }
        }
    }
}
