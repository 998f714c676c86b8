/*
 * orc_rotation.c -- oracle: degree-2 Liouville polar motion with two compact
 * Maxwell channels (loading and tidal).
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Restates src/vilma_rotation.f90:
 *   :114-152 (init), :154-168 (begin_step), :170-190 (solve_m),
 *   :192-222 (s_rot), :224-249 (commit/update), :335-363 (inertia21),
 *   :375-404 (fluid_tidal_k), :408-487 (channels).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

static int ch_init(orc_ch2* c, const orc_earth* e, const orc_mesh* m, double dt, int tidal, char* err) {
    memset(c, 0, sizeof(*c));
    c->tidal = tidal; c->nr = m->nr; c->ne = m->ne; c->ndof = 4 * m->nr - 1;
    c->Jr = 6.0; c->dt = dt;
    int nr = c->nr, ne = c->ne, nd = c->ndof;
    c->r = malloc(nr * 8); memcpy(c->r, m->r, nr * 8);
    c->mu = malloc(ne * 8); c->Mk = malloc(ne * 8); c->MkPerDt = malloc(ne * 8);
    for (int el = 0; el < ne; el++) {
        const orc_layer* ly = &e->L[m->lay[el]];
        c->mu[el] = ly->mu;
        c->MkPerDt[el] = ly->eta > 0.0 ? c->mu[el] / ly->eta : 0.0;
        c->Mk[el] = c->MkPerDt[el] * dt;
    }
    orc_strain_constants(c->Jr, c->norm, c->sa, c->sb, c->sc);
    if (orc_radop_assemble(&c->op, e, m, 2, err)) return 1;
    double *b = malloc(nd * 8), *x = malloc(nd * 8);
    if (tidal) orc_radop_tidal_rhs(&c->op, 1.0, b); else orc_radop_load_rhs(&c->op, 1.0, b);
    orc_radop_solve(&c->op, b, x, NULL);
    c->Fe = x[idx_f(nr - 1)]; c->Ue = x[idx_u(nr - 1)];
    c->xUn = malloc(nr * 8); c->xVn = malloc(nr * 8);
    for (int n = 0; n < nr; n++) { c->xUn[n] = x[idx_u(n)]; c->xVn[n] = x[idx_v(n)]; }
    free(b); free(x);
    c->Are = calloc(4 * ne, 8); c->Aim = calloc(4 * ne, 8); c->Bre = calloc(4 * ne, 8);
    c->Bim = calloc(4 * ne, 8); c->Cre = calloc(4 * ne, 8); c->Cim = calloc(4 * ne, 8);
    c->dUn_re = calloc(nr, 8); c->dUn_im = calloc(nr, 8); c->dVn_re = calloc(nr, 8); c->dVn_im = calloc(nr, 8);
    return 0;
}

static void ch_free(orc_ch2* c) {
    orc_radop_free(&c->op);
    free(c->r); free(c->mu); free(c->Mk); free(c->MkPerDt); free(c->xUn); free(c->xVn);
    free(c->Are); free(c->Aim); free(c->Bre); free(c->Bim); free(c->Cre); free(c->Cim);
    free(c->dUn_re); free(c->dUn_im); free(c->dVn_re); free(c->dVn_im);
    memset(c, 0, sizeof(*c));
}

static void ch_begin(orc_ch2* c) {
    int nd = c->ndof, nr = c->nr;
    double *fre = calloc(nd, 8), *fim = calloc(nd, 8), *xre = malloc(nd * 8), *xim = malloc(nd * 8);
    orc_dissipative_rhs(c->ne, c->r, c->sa, c->sb, c->sc, c->norm, c->Are, c->Bre, c->Cre, 4, fre);
    orc_dissipative_rhs(c->ne, c->r, c->sa, c->sb, c->sc, c->norm, c->Aim, c->Bim, c->Cim, 4, fim);
    orc_radop_solve(&c->op, fre, xre, NULL);
    orc_radop_solve(&c->op, fim, xim, NULL);
    c->dF = xre[idx_f(nr - 1)] + I * xim[idx_f(nr - 1)];
    c->dU = xre[idx_u(nr - 1)] + I * xim[idx_u(nr - 1)];
    for (int n = 0; n < nr; n++) {
        c->dUn_re[n] = xre[idx_u(n)]; c->dUn_im[n] = xim[idx_u(n)];
        c->dVn_re[n] = xre[idx_v(n)]; c->dVn_im[n] = xim[idx_v(n)];
    }
    free(fre); free(fim); free(xre); free(xim);
}

static void ch_commit(orc_ch2* c, cplx coeff) {
    int nr = c->nr;
    double *Ure = malloc(nr * 8), *Uim = malloc(nr * 8), *Vre = malloc(nr * 8), *Vim = malloc(nr * 8);
    double cr = creal(coeff), ci = cimag(coeff);
    for (int n = 0; n < nr; n++) {
        Ure[n] = cr * c->xUn[n] + c->dUn_re[n]; Uim[n] = ci * c->xUn[n] + c->dUn_im[n];
        Vre[n] = cr * c->xVn[n] + c->dVn_re[n]; Vim[n] = ci * c->xVn[n] + c->dVn_im[n];
    }
    orc_advance_memory_fe(c->ne, c->mu, c->Mk, Ure, Vre, c->Jr, c->Are, c->Bre, c->Cre, 4, NULL);
    orc_advance_memory_fe(c->ne, c->mu, c->Mk, Uim, Vim, c->Jr, c->Aim, c->Bim, c->Cim, 4, NULL);
    free(Ure); free(Uim); free(Vre); free(Vim);
}

/* relaxed degree-2 tidal Love number: Maxwell layers fluidised (:375-404) */
static int fluid_tidal_k(const orc_earth* e, const orc_mesh* m, double* ks, char* err) {
    orc_earth ef = *e;
    for (int i = 0; i < ef.n; i++)
        if (ef.L[i].rheology == RHEOL_MAXWELL) { ef.L[i].mu = 0.0; ef.L[i].rheology = RHEOL_FLUID; }
    orc_radop op;
    if (orc_radop_assemble(&op, &ef, m, 2, err)) return 1;
    double *b = malloc(op.ndof * 8), *x = malloc(op.ndof * 8);
    orc_radop_tidal_rhs(&op, 1.0, b);
    orc_radop_solve(&op, b, x, NULL);
    *ks = -x[idx_f(m->nr - 1)] / 1.0 - 1.0;
    free(b); free(x); orc_radop_free(&op);
    return 0;
}

int orc_rot_init(orc_rot* rt, const orc_earth* e, const orc_mesh* m, double dt, char* err) {
    memset(rt, 0, sizeof(*rt));
    rt->a = e->r_earth; rt->g = orc_gravity_at(e, e->r_earth);
    rt->CminusA = 2.63e35; rt->Omega = 7.292115e-5;
    rt->dt_fe_max = 1.79769313486231571e308;
    if (ch_init(&rt->load_ch, e, m, dt, 0, err)) return 1;
    if (ch_init(&rt->tidal_ch, e, m, dt, 1, err)) return 1;
    rt->kTe = -rt->tidal_ch.Fe - 1.0;
    rt->hTe = orc_gravity_at(e, e->r_earth) * rt->tidal_ch.Ue;
    if (fluid_tidal_k(e, m, &rt->k_s_fluid, err)) return 1;
    rt->k_s_flat = 3.0 * ORC_G * rt->CminusA / (pow(rt->a, 5) * rt->Omega * rt->Omega);
    double mx = 0.0;
    for (int el = 0; el < rt->load_ch.ne; el++) mx = fmax(mx, rt->load_ch.MkPerDt[el]);
    if (mx > 0.0) rt->dt_fe_max = 1.0 / mx;
    rt->k_s = rt->k_s_fluid;
    rt->m = 0.0; rt->time = 0.0;
    return 0;
}

void orc_rot_free(orc_rot* rt) { ch_free(&rt->load_ch); ch_free(&rt->tidal_ch); memset(rt, 0, sizeof(*rt)); }

static void ch_set_dt(orc_ch2* c, double dt) {
    c->dt = dt;
    for (int e = 0; e < c->ne; e++) c->Mk[e] = c->MkPerDt[e] * dt;
}

void orc_rot_begin(orc_rot* rt, double dt) {
    if (!rt->enabled) return;
    if (dt != rt->load_ch.dt) { ch_set_dt(&rt->load_ch, dt); ch_set_dt(&rt->tidal_ch, dt); }
    ch_begin(&rt->load_ch);
    ch_begin(&rt->tidal_ch);
}

/* I13 + i I23 by direct quadrature (:335-363) */
static cplx inertia21(const orc_sht* s, const double* load, double a) {
    int np = s->nphi, nl = s->nlat;
    double* w13 = malloc((size_t)np * nl * 8); double* w23 = malloc((size_t)np * nl * 8);
    for (int j = 0; j < nl; j++) {
        double st = sin(s->colat[j]), ct = cos(s->colat[j]), sc2 = st * ct;
        for (int i = 0; i < np; i++) {
            w13[(size_t)j * np + i] = load[(size_t)j * np + i] * sc2 * cos(s->lon[i]);
            w23[(size_t)j * np + i] = load[(size_t)j * np + i] * sc2 * sin(s->lon[i]);
        }
    }
    double a4 = pow(a, 4);
    cplx r = -a4 * orc_surface_integral(s, w13) + I * (-a4 * orc_surface_integral(s, w23));
    free(w13); free(w23);
    return r;
}

void orc_rot_solve_m(orc_rot* rt, const orc_sht* s, const double* load) {
    if (!rt->enabled) return;
    cplx Irig = inertia21(s, load, rt->a);
    double scl = 5.0 / (4.0 * ORC_PI * ORC_G * rt->a);
    rt->cload = Irig * scl;
    cplx Itot = -(rt->load_ch.Fe * rt->cload + rt->load_ch.dF);
    cplx psiL = Itot / rt->CminusA;
    rt->m = (psiL - rt->tidal_ch.dF / rt->k_s) / (1.0 - rt->kTe / rt->k_s);
}

void orc_rot_s_rot(const orc_rot* rt, const orc_sht* s, double* srot) {
    int np = s->nphi, nl = s->nlat;
    if (!rt->enabled) { memset(srot, 0, (size_t)np * nl * 8); return; }
    cplx cN = rt->m + (rt->kTe * rt->m - rt->tidal_ch.dF);
    cplx cU = rt->tidal_ch.Ue * rt->m + rt->tidal_ch.dU;
    double kN = rt->Omega * rt->Omega * rt->a * rt->a / rt->g;
    double ku = rt->Omega * rt->Omega * rt->a * rt->a;
    for (int j = 0; j < nl; j++) {
        double gam = sin(s->colat[j]) * cos(s->colat[j]);
        for (int i = 0; i < np; i++) {
            double cphi = cos(s->lon[i]), sphi = sin(s->lon[i]);
            srot[(size_t)j * np + i] = gam * (kN * (creal(cN) * cphi + cimag(cN) * sphi)
                                              - ku * (creal(cU) * cphi + cimag(cU) * sphi));
        }
    }
}

void orc_rot_update(orc_rot* rt, const orc_sht* s, const double* load, double dt) {
    if (!rt->enabled) return;
    orc_rot_begin(rt, dt);
    orc_rot_solve_m(rt, s, load);
    ch_commit(&rt->load_ch, rt->cload);
    ch_commit(&rt->tidal_ch, rt->m);
    rt->time += rt->load_ch.dt;
}

/* The rotation module on its own, as the reference's rotation test drives it
 * (tests/test_rotation.f90:73-117): a fixed surface load on the grid, the
 * channels of the earth in p on its own mesh, nsteps + 1 updates of dt
 * seconds. m_out[2i], m_out[2i+1] = m after update i (the polar motion at
 * t = i dt); info = kTe, k_s, and the rigid (2,1) inertia of the load as
 * (re, im). Test helper: the oracle's rotation checked against the Spada et
 * al. (2011) benchmark. */
int orc_rotation_run(const vh_params* p, int lmax, int nlat, int nphi, const double* load, double dt,
                     int nsteps, double* m_out, double* info, char* err) {
    orc_earth e; orc_mesh m; orc_sht s; orc_rot rt;
    if (orc_build_earth(p, &e, err)) return 1;
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return 1;
    if (orc_sht_init(&s, lmax, nlat, nphi, err)) { orc_mesh_free(&m); return 1; }
    if (orc_rot_init(&rt, &e, &m, dt, err)) { orc_sht_free(&s); orc_mesh_free(&m); return 1; }
    rt.enabled = 1;
    cplx I21 = inertia21(&s, load, rt.a);
    info[0] = rt.kTe; info[1] = rt.k_s; info[2] = creal(I21); info[3] = cimag(I21);
    for (int i = 0; i <= nsteps; i++) {
        orc_rot_update(&rt, &s, load, dt);
        m_out[2 * i] = creal(rt.m); m_out[2 * i + 1] = cimag(rt.m);
    }
    orc_rot_free(&rt); orc_sht_free(&s); orc_mesh_free(&m);
    return 0;
}
