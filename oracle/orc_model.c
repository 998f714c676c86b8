/*
 * orc_model.c -- oracle: parameters, the forward-Euler sub-stepping
 * controller, the coupling step (solid_earth_init/update) and the C entry
 * points the tests and the CPU baseline load through ctypes.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Restates:
 *   input/vilma_defaults.nml + src/vilma_params.f90:26-205 (defaults)
 *   src/vilma_coupling.f90:144-232 (init), :252-318 (build_solver),
 *     :336-437 (update), :572-597 (update_bsl), :644-662 (build_sht)
 *   src/vilma_timestep.f90:79-227, :314-326 (FE stepper, solve_at)
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

int orc_params_defaults(vh_params* p) {
    memset(p, 0, sizeof(*p));
    p->lmax = 64; p->nlat = 0; p->nphi = 0; p->mmax = -1; p->mres = 1; p->eps_polar = -1.0;
    strcpy(p->earth, "M3-L70-V01");
    p->n_layer = 5; p->r_earth = 6371.0e3; p->r_core = 3480.0e3;
    const double rb[5] = {6301.0e3, 5951.0e3, 5701.0e3, 3480.0e3, 0.0};
    const double rt[5] = {6371.0e3, 6301.0e3, 5951.0e3, 5701.0e3, 3480.0e3};
    const double rho[5] = {3037.0, 3438.0, 3871.0, 4978.0, 10750.0};
    const double mu[5] = {5.0605e10, 7.0363e10, 1.0549e11, 2.2834e11, 0.0};
    const double eta[5] = {1.0e40, 1.0e21, 1.0e21, 2.0e21, 0.0};
    const int rh[5] = {0, 1, 1, 1, 2};
    for (int i = 0; i < VH_MAX_LAYER; i++) p->rheology[i] = 1;
    for (int i = 0; i < 5; i++) {
        p->r_bot[i] = rb[i]; p->r_top[i] = rt[i]; p->rho[i] = rho[i];
        p->mu[i] = mu[i]; p->eta[i] = eta[i]; p->rheology[i] = rh[i];
    }
    strcpy(p->scheme, "fe"); p->max_couple_iter = 20;
    strcpy(p->earth_response, "ve");
    p->sle_n_outer = 3; p->sle_n_inner = 20; p->sle_tol = 1.0e-7; p->sle_max_mem_iter = 20;
    p->sle_fixed_ocean = 0; p->sle_subgrid = 1;
    p->dt_init = 0.0; p->dt_min = 0.0; p->dt_max = 1.0e30 * ORC_SEC_PER_YEAR;
    p->rtol = 1.0e-4; p->atol = 1.0e-3; p->safety = 0.9; p->grow_max = 5.0; p->shrink_min = 0.2; p->cfl = 1.0;
    p->rotation = 1;
    strcpy(p->deg1_frame, "cm");
    p->l_visc_3d = 0; p->visc3d_tol = 1.0e-3; p->l_toroidal = 0;
    p->visc_log10_min = 19.5; p->visc_log10_max = 30.0; p->visc3d_lid_depth = 0.0; p->visc3d_lid_log10max = 22.0;
    p->skip_tol = 1.0e-4;
    return 0;
}

void orc_destroy(orc_model* m);

int orc_create(const vh_params* p, int nphi, int nlat, const double* z_bed_eq, const double* h_ice_eq,
               const double* h_ice_init, orc_model** out, char* err) {
    orc_model* m = calloc(1, sizeof(orc_model));
    *out = m;
    m->par = *p;
    if (strcmp(p->scheme, "fe") != 0 && strcmp(p->scheme, "trap") != 0) {
        snprintf(err, 256, "scheme \"%s\" not supported (fe|trap)", p->scheme); return 1;
    }
    if ((p->mmax >= 0 && p->mmax != p->lmax) || p->mres != 1) { snprintf(err, 256, "mmax/mres other than lmax/1 not supported"); return 1; }
    int nl = p->nlat > 0 ? p->nlat : 2 * p->lmax + 2;
    int np = p->nphi > 0 ? p->nphi : 4 * p->lmax;
    if (nl < 32) nl = 32;                               /* src/vilma_sht.f90:72 NLAT_MIN */
    if (nl != nlat || np != nphi) { snprintf(err, 256, "grid (%d,%d) does not match the model Gauss grid (%d,%d)", nphi, nlat, np, nl); return 1; }
    if (orc_sht_init(&m->sht, p->lmax, nl, np, err)) return 1;
    m->np = np; m->nl = nl;
    if (orc_build_earth(p, &m->earth, err)) return 1;
    if (orc_mesh_build(&m->earth, p->n_elem_layer, &m->mesh, err)) return 1;
    double dt0 = p->dt_init > 0.0 ? p->dt_init : 1000.0 * ORC_SEC_PER_YEAR;
    m->resp.deg1_cm = strcmp(p->deg1_frame, "cm") == 0;
    m->resp.skip_tol = p->skip_tol;
    m->resp.visc3d_tol = p->visc3d_tol;
    m->resp.toroidal = p->l_toroidal != 0;   /* src/vilma_coupling.f90:282 */
    if (strcmp(p->earth_response, "ve") == 0) {
        if (orc_resp_init_ve(&m->resp, &m->earth, &m->mesh, &m->sht, dt0, err)) return 1;
    } else if (strcmp(p->earth_response, "elastic") == 0) {
        if (orc_resp_init_elastic(&m->resp, &m->earth, &m->mesh, p->lmax, err)) return 1;
    } else if (strcmp(p->earth_response, "null") == 0) {
        m->resp.kind = RESP_NULL;
    } else {
        snprintf(err, 256, "solid_earth_init: unknown earth_response (use ve|elastic|null)"); return 1;
    }
    m->resp.visc3d_tol = p->visc3d_tol;
    /* the scheme reaches the response only for ve (src/vilma_coupling.f90:269-281) */
    if (m->resp.kind == RESP_VE && strcmp(p->scheme, "trap") == 0) orc_resp_set_implicit(&m->resp, m->sht.nlm);
    m->sle.n_outer = p->sle_n_outer; m->sle.n_inner = p->sle_n_inner; m->sle.tol = p->sle_tol;
    m->sle.max_mem_iter = p->sle_max_mem_iter; m->sle.fixed_ocean = p->sle_fixed_ocean;
    m->sle.subgrid = p->sle_subgrid; m->sle.warm_start = 1;
    m->cfl = p->cfl; m->dt_min = p->dt_min; m->dt_max = p->dt_max; m->guard_growth = 1.0e3;
    m->rtol = p->rtol; m->atol = p->atol; m->safety = p->safety; m->grow_max = p->grow_max;
    m->shrink_min = p->shrink_min; m->dt_try = p->dt_init;   /* src/vilma_coupling.f90:300-308 */
    size_t ng = (size_t)np * nl;
    m->s_rot = calloc(ng, 8);
    if (p->rotation) {
        if (orc_rot_init(&m->rot, &m->earth, &m->mesh, dt0, err)) return 1;
        m->rot.enabled = 1;
    }
    m->z_bed_eq = malloc(ng * 8); memcpy(m->z_bed_eq, z_bed_eq, ng * 8);
    m->h_ice_eq = malloc(ng * 8); memcpy(m->h_ice_eq, h_ice_eq, ng * 8);
    m->h_ice = malloc(ng * 8); memcpy(m->h_ice, h_ice_init ? h_ice_init : h_ice_eq, ng * 8);
    m->rsl = calloc(ng, 8);
    m->z_bed = malloc(ng * 8); memcpy(m->z_bed, z_bed_eq, ng * 8);
    m->C0 = malloc(ng * 8); m->C = malloc(ng * 8);
    orc_ocean_function((int)ng, m->z_bed_eq, m->h_ice_eq, m->C0);
    memcpy(m->C, m->C0, ng * 8);
    m->sigma_lm = calloc(m->sht.nlm, sizeof(cplx));
    m->u_lm = calloc(m->sht.nlm, sizeof(cplx)); m->n_lm = calloc(m->sht.nlm, sizeof(cplx));
    return 0;
}

void orc_destroy(orc_model* m) {
    if (!m) return;
    orc_resp_free(&m->resp);
    if (m->rot.load_ch.r) orc_rot_free(&m->rot);
    orc_mesh_free(&m->mesh);
    orc_sht_free(&m->sht);
    free(m->z_bed_eq); free(m->h_ice_eq); free(m->h_ice); free(m->rsl); free(m->z_bed);
    free(m->C); free(m->C0); free(m->s_rot); free(m->sigma_lm); free(m->u_lm); free(m->n_lm);
    free(m);
}

int orc_enable_visc_3d(orc_model* m, const double* pert) {
    return orc_resp_enable_lateral(&m->resp, &m->sht, pert, m->err);
}

/* response_enable_lateral_visc_from_nodes (src/vilma_response.f90:1120-1169) */
int orc_enable_visc_3d_nodes(orc_model* m, const double* visc_node, double lid_depth,
                             double lid_log10max, double log10_cap) {
    orc_resp* r = &m->resp;
    if (r->kind != RESP_VE) { snprintf(m->err, 256, "solid_earth_enable_visc_3d: lateral viscosity needs earth_response=ve"); return 1; }
    size_t ng = (size_t)m->np * m->nl;
    double r_lid = DBL_MAX;
    if (lid_depth > 0.0) r_lid = r->r[r->nr - 1] - lid_depth;
    double* pert = calloc(ng * r->ne, 8);
    for (int e = 0; e < r->ne; e++) {
        if (r->MkPerDt[e] == 0.0) continue;
        double logeta_ref = log10(r->mu[e] / r->MkPerDt[e]);
        int in_lid = r->r[e] >= r_lid - 1.0e-6 * r->r[r->nr - 1];
        for (size_t sp = 0; sp < ng; sp++) {
            double ea = 0.5 * (visc_node[sp + (size_t)e * ng] + visc_node[sp + (size_t)(e + 1) * ng]);
            if (in_lid && ea > lid_log10max) ea = log10_cap;
            pert[(size_t)e * ng + sp] = ea - logeta_ref;
        }
    }
    int rc = orc_resp_enable_lateral(r, &m->sht, pert, m->err);
    free(pert);
    return rc;
}

/* one SLE solve with the ice interpolated at t_eval (solve_at,
 * src/vilma_timestep.f90:312-326) */
static void solve_at(orc_model* m, const double* ice0, const double* ice1, double t0, double span, double t_eval,
                     const double* s_rot, double* ice_now, double* dice_now) {
    size_t ng = (size_t)m->np * m->nl;
    orc_sle_result res;
    double frac = (t_eval - t0) / span;
    for (size_t i = 0; i < ng; i++) {
        ice_now[i] = ice0[i] + frac * (ice1[i] - ice0[i]);
        dice_now[i] = ice_now[i] - m->h_ice_eq[i];
    }
    m->n_solve++;
    orc_sle_solve(&m->sle, &m->sht, &m->resp, dice_now, ice_now, m->z_bed_eq, m->rsl, m->C, &res, 0, m->sigma_lm, s_rot);
    m->worst_mass_resid = fmax(m->worst_mass_resid, res.mass_resid);
}

/* stepper_advance, implicit (trapezoidal) branch with the step-doubling
 * controller (src/vilma_timestep.f90:229-307) */
static int stepper_trap(orc_model* m, const double* ice0, const double* ice1, double t0, double t1,
                        const double* s_rot, double* ice_now, double* dice_now, double* rsl_n) {
    size_t ng = (size_t)m->np * m->nl;
    double span = t1 - t0;
    orc_resp* r = &m->resp;
    if (!r->sigma_primed) {   /* the start-of-step load sigma_0 from a report-only solve */
        orc_sle_result res;
        cplx* sig0 = calloc(m->sht.nlm, sizeof(cplx));
        memcpy(ice_now, ice0, ng * 8);
        for (size_t i = 0; i < ng; i++) dice_now[i] = ice0[i] - m->h_ice_eq[i];
        orc_sle_solve(&m->sle, &m->sht, r, dice_now, ice_now, m->z_bed_eq, m->rsl, m->C, &res, 1, sig0, s_rot);
        orc_resp_prime_sigma(r, sig0);
        free(sig0);
    }
    const int p = 2;                                   /* scheme_order(trap) */
    const double ricfac = (double)((1 << p) - 1), expo = 1.0 / (double)(p + 1);
    double t = t0;
    if (m->dt_try <= 0.0) m->dt_try = span;
    for (;;) {
        if (t >= t1 - 1.0e-9 * span) break;
        double dt = fmin(m->dt_try, t1 - t);
        int at_floor = m->dt_min > 0.0 && dt <= m->dt_min * (1.0 + 1.0e-9);
        memcpy(rsl_n, m->rsl, ng * 8);
        orc_resp_save(r);
        orc_resp_set_dt(r, dt);
        solve_at(m, ice0, ice1, t0, span, t + dt, s_rot, ice_now, dice_now);            /* coarse */
        orc_resp_stash_coarse(r);
        orc_resp_restore(r);
        memcpy(m->rsl, rsl_n, ng * 8);
        orc_resp_set_dt(r, 0.5 * dt);
        solve_at(m, ice0, ice1, t0, span, t + 0.5 * dt, s_rot, ice_now, dice_now);      /* fine 1 */
        solve_at(m, ice0, ice1, t0, span, t + dt, s_rot, ice_now, dice_now);            /* fine 2 */
        double err_inf, tau_inf;
        orc_resp_coarse_fine_error(r, &err_inf, &tau_inf);
        orc_resp_set_dt(r, dt);
        double errsc = (err_inf / ricfac) / (m->atol + m->rtol * tau_inf);
        if (!(errsc <= DBL_MAX)) {
            snprintf(m->err, 256, "vilma_timestep: implicit local-error estimate went non-finite "
                                  "(viscosity too stiff, or the state has already diverged)");
            return 1;
        }
        int accept = errsc <= 1.0 || at_floor;
        if (accept) {
            t = t + dt; m->n_accept++;
            if (at_floor && errsc > 1.0) m->n_floor++;
        } else {
            orc_resp_restore(r);
            memcpy(m->rsl, rsl_n, ng * 8);
            m->n_reject++;
        }
        double fac = errsc <= 0.0 ? m->grow_max : m->safety * pow(errsc, -expo);
        fac = fmin(m->grow_max, fmax(m->shrink_min, fac));
        m->dt_try = fmin(m->dt_max, fmax(m->dt_min, dt * fac));
        if (m->dt_try <= 1.0e-12 * span) {
            snprintf(m->err, 256, "vilma_timestep: implicit adaptive step collapsed (tolerance unreachable)");
            return 1;
        }
    }
    return 0;
}

/* stepper_advance, forward-Euler branch (src/vilma_timestep.f90:79-227) */
static int stepper_advance(orc_model* m, const double* ice0, const double* ice1, double t0, double t1,
                           const double* s_rot) {
    size_t ng = (size_t)m->np * m->nl;
    double span = t1 - t0;
    orc_sle_result res;
    m->worst_mass_resid = 0.0;
    double *ice_now = malloc(ng * 8), *dice_now = malloc(ng * 8), *rsl_n = malloc(ng * 8);
    if (span <= 0.0) {
        memcpy(ice_now, ice0, ng * 8);
        for (size_t i = 0; i < ng; i++) dice_now[i] = ice0[i] - m->h_ice_eq[i];
        orc_sle_solve(&m->sle, &m->sht, &m->resp, dice_now, ice_now, m->z_bed_eq, m->rsl, m->C, &res, 1, m->sigma_lm, s_rot);
        m->worst_mass_resid = res.mass_resid;
        free(ice_now); free(dice_now); free(rsl_n);
        return 0;
    }
    if (m->resp.implicit) {
        int rc = stepper_trap(m, ice0, ice1, t0, t1, s_rot, ice_now, dice_now, rsl_n);
        free(ice_now); free(dice_now); free(rsl_n);
        return rc;
    }
    double rate = orc_resp_max_rate(&m->resp);
    double dt_stab = rate > 0.0 ? m->cfl / rate : span;
    dt_stab = fmin(dt_stab, m->dt_max);
    if (m->dt_min > 0.0) dt_stab = fmax(dt_stab, m->dt_min);
    int n_sub = (int)ceil(span / dt_stab - 1.0e-9);
    if (n_sub < 1) n_sub = 1;
    double dt = span / (double)n_sub, t = t0;
    double tau_run = orc_resp_memory_norm(&m->resp);
    for (;;) {
        if (t >= t1 - 1.0e-9 * span) break;
        dt = fmin(dt, t1 - t);
        memcpy(rsl_n, m->rsl, ng * 8);
        double tau0 = orc_resp_memory_norm(&m->resp);
        orc_resp_save(&m->resp);
        orc_resp_set_dt(&m->resp, dt);
        double frac = ((t + dt) - t0) / span;
        for (size_t i = 0; i < ng; i++) {
            ice_now[i] = ice0[i] + frac * (ice1[i] - ice0[i]);
            dice_now[i] = ice_now[i] - m->h_ice_eq[i];
        }
        m->n_solve++;
        orc_sle_solve(&m->sle, &m->sht, &m->resp, dice_now, ice_now, m->z_bed_eq, m->rsl, m->C, &res, 0, m->sigma_lm, s_rot);
        m->worst_mass_resid = fmax(m->worst_mass_resid, res.mass_resid);
        double err_inf = orc_resp_memory_norm(&m->resp);
        int finite = err_inf <= DBL_MAX;
        double scale = fmax(tau0, tau_run);
        int accept = scale <= 0.0 ? finite : (finite && err_inf <= m->guard_growth * scale);
        if (accept) {
            tau_run = fmax(tau_run, err_inf);
            t = t + dt; m->n_accept++;
            dt = fmin(span / (double)n_sub, 2.0 * dt);
        } else {
            orc_resp_restore(&m->resp);
            memcpy(m->rsl, rsl_n, ng * 8);
            m->n_reject++;
            dt = 0.5 * dt;
            if (dt <= 1.0e-12 * span) {
                snprintf(m->err, 256, "vilma_timestep: explicit guard sub-step collapsed (viscosity too stiff for FE)");
                free(ice_now); free(dice_now); free(rsl_n);
                return 1;
            }
        }
    }
    free(ice_now); free(dice_now); free(rsl_n);
    return 0;
}

/* update_bsl (src/vilma_coupling.f90:572-597) */
static void update_bsl(orc_model* m) {
    size_t ng = (size_t)m->np * m->nl;
    double c_int = orc_surface_integral(&m->sht, m->C);
    if (c_int > 0.0) {
        double* t = malloc(ng * 8);
        for (size_t i = 0; i < ng; i++) t[i] = m->h_ice[i] * (1.0 - m->C[i]) - m->h_ice_eq[i] * (1.0 - m->C0[i]);
        m->bsl = -(ORC_RHO_ICE / ORC_RHO_WATER) * orc_surface_integral(&m->sht, t) / c_int;
        free(t);
    } else {
        m->bsl = 0.0;
    }
}

/* solid_earth_update (src/vilma_coupling.f90:336-437), passthrough grid */
int orc_update(orc_model* m, const double* h_ice, double dt_yr) {
    size_t ng = (size_t)m->np * m->nl;
    double dt = dt_yr * ORC_SEC_PER_YEAR;
    double t0 = m->time * ORC_SEC_PER_YEAR, t1 = t0 + dt;
    double* ice_new = malloc(ng * 8);
    memcpy(ice_new, h_ice, ng * 8);
    int rc;
    if (m->rot.enabled) {
        orc_rot_begin(&m->rot, dt);
        orc_rot_s_rot(&m->rot, &m->sht, m->s_rot);
        rc = stepper_advance(m, m->h_ice, ice_new, t0, t1, m->s_rot);
        if (rc) { free(ice_new); return rc; }
        double* load = malloc(ng * 8);
        orc_sht_synthesis(&m->sht, m->sigma_lm, load);
        int n_sub = (int)ceil(dt / m->rot.dt_fe_max);
        if (n_sub < 1) n_sub = 1;
        double dt_sub = dt / (double)n_sub;
        for (int k = 0; k < n_sub; k++) orc_rot_update(&m->rot, &m->sht, load, dt_sub);
        free(load);
    } else {
        rc = stepper_advance(m, m->h_ice, ice_new, t0, t1, NULL);
        if (rc) { free(ice_new); return rc; }
    }
    memcpy(m->h_ice, ice_new, ng * 8);
    for (size_t i = 0; i < ng; i++) m->z_bed[i] = m->z_bed_eq[i] - m->rsl[i];
    m->time += dt_yr;
    update_bsl(m);
    free(ice_new);
    return 0;
}

/* ---------------------------------------------------------------- getters */

int orc_get_field(orc_model* m, int field, double* out) {
    size_t ng = (size_t)m->np * m->nl;
    const double* src = NULL;
    switch (field) {
        case VH_F_RSL: src = m->rsl; break;
        case VH_F_ZBED: src = m->z_bed; break;
        case VH_F_C: src = m->C; break;
        case VH_F_C0: src = m->C0; break;
        case VH_F_SROT: src = m->s_rot; break;
        case VH_F_HICE: src = m->h_ice; break;
        default: return 1;
    }
    memcpy(out, src, ng * 8);
    return 0;
}

int orc_get_scalar(orc_model* m, int which, double* out) {
    switch (which) {
        case VH_S_TIME: *out = m->time; break;
        case VH_S_BSL: *out = m->bsl; break;
        case VH_S_WORST_MASS_RESID: *out = m->worst_mass_resid; break;
        case VH_S_N_ACCEPT: *out = (double)m->n_accept; break;
        case VH_S_N_REJECT: *out = (double)m->n_reject; break;
        case VH_S_N_SOLVE: *out = (double)m->n_solve; break;
        case VH_S_N_INNER_TOT: *out = (double)m->sle.n_inner_tot; break;
        case VH_S_N_OUTER_TOT: *out = (double)m->sle.n_outer_tot; break;
        case VH_S_POLAR_M_RE: *out = creal(m->rot.m); break;
        case VH_S_POLAR_M_IM: *out = cimag(m->rot.m); break;
        case VH_S_N_SLE_STEAL: *out = 0.0; break;
        case VH_S_N_FLOOR: *out = (double)m->n_floor; break;
        case VH_S_N_COUPLE_TOT: *out = (double)m->sle.n_couple_tot; break;
        default: return 1;
    }
    return 0;
}

int orc_get_spectrum(orc_model* m, int which, double* out) {
    int nlm = m->sht.nlm;
    if (which == VH_SP_SIGMA) { memcpy(out, m->sigma_lm, nlm * sizeof(cplx)); return 0; }
    orc_resp_apply(&m->resp, &m->sht, m->sigma_lm, m->u_lm, m->n_lm);
    if (which == VH_SP_U) { memcpy(out, m->u_lm, nlm * sizeof(cplx)); return 0; }
    if (which == VH_SP_N) { memcpy(out, m->n_lm, nlm * sizeof(cplx)); return 0; }
    if (which == VH_SP_TOR) {   /* response_horizontal_toroidal (src/vilma_response.f90:884-900) */
        memset(out, 0, nlm * sizeof(cplx));
        if (m->resp.kind == RESP_VE && m->resp.nlam > 4)
            for (int k = 0; k < m->resp.nk; k++) ((cplx*)out)[m->resp.k2lm[k]] = m->resp.dWa[k];
        return 0;
    }
    return 1;
}

int orc_get_dims(orc_model* m, int which, int* out) {
    int nem = 0;
    if (m->resp.kind == RESP_VE) for (int e = 0; e < m->resp.ne; e++) nem += m->resp.MkPerDt[e] > 0.0 || !m->resp.active1d[e];
    switch (which) {
        case VH_D_LMAX: *out = m->sht.lmax; break;
        case VH_D_NLAT: *out = m->nl; break;
        case VH_D_NPHI: *out = m->np; break;
        case VH_D_NLM: *out = m->sht.nlm; break;
        case VH_D_NR: *out = m->mesh.nr; break;
        case VH_D_NE: *out = m->mesh.ne; break;
        case VH_D_NK: *out = m->resp.nk; break;
        case VH_D_NE3D: *out = m->resp.ne3d; break;
        case VH_D_NLAM: *out = m->resp.kind == RESP_VE ? m->resp.nlam : 4; break;
        case VH_D_NEM: *out = nem; break;
        case VH_D_NMN: {
            int nmn = 0;
            if (m->resp.kind == RESP_VE)
                for (int n = 0; n <= m->resp.ne; n++) {
                    int used = 0;
                    for (int e = n - 1; e <= n; e++)
                        if (e >= 0 && e < m->resp.ne && (m->resp.MkPerDt[e] > 0.0 || !m->resp.active1d[e])) used = 1;
                    nmn += used;
                }
            *out = nmn;
            break;
        }
        default: return 1;
    }
    return 0;
}

int orc_get_memory(orc_model* m, double* tau) {
    if (m->resp.kind != RESP_VE) return 1;
    size_t n = (size_t)m->resp.nlam * m->resp.ne * m->resp.nk;
    double* arr[6] = {m->resp.Are, m->resp.Aim, m->resp.Bre, m->resp.Bim, m->resp.Cre, m->resp.Cim};
    for (int q = 0; q < 6; q++) memcpy(tau + q * n, arr[q], n * 8);
    return 0;
}

int orc_set_memory(orc_model* m, const double* tau) {
    if (m->resp.kind != RESP_VE) return 1;
    size_t n = (size_t)m->resp.nlam * m->resp.ne * m->resp.nk;
    double* arr[6] = {m->resp.Are, m->resp.Aim, m->resp.Bre, m->resp.Bim, m->resp.Cre, m->resp.Cim};
    for (int q = 0; q < 6; q++) memcpy(arr[q], tau + q * n, n * 8);
    return 0;
}

const char* orc_last_error(orc_model* m) { return m->err; }

/* ------------------------------------------------------ response probes */

int orc_resp_begin_step(orc_model* m) { orc_resp_begin(&m->resp); return 0; }
int orc_resp_apply_c(orc_model* m, const double* sigma, double* u, double* n) {
    orc_resp_apply(&m->resp, &m->sht, (const cplx*)sigma, (cplx*)u, (cplx*)n);
    return 0;
}
int orc_resp_advance_c(orc_model* m, const double* sigma) {
    orc_resp_advance(&m->resp, &m->sht, (const cplx*)sigma);
    m->resp.time += m->resp.dt;
    return 0;
}
int orc_resp_set_dt_c(orc_model* m, double dt) { orc_resp_set_dt(&m->resp, dt); return 0; }
int orc_get_drift(orc_model* m, double* dsurf, double* dnodal) {
    orc_resp* r = &m->resp;
    if (r->kind != RESP_VE) return 1;
    size_t nk = r->nk, nn = (size_t)r->nr * nk;
    if (dsurf) {
        memcpy(dsurf, r->dUa, nk * 16); memcpy(dsurf + 2 * nk, r->dFa, nk * 16); memcpy(dsurf + 4 * nk, r->dVa, nk * 16);
    }
    if (dnodal) {
        memcpy(dnodal, r->dUn_re, nn * 8); memcpy(dnodal + nn, r->dUn_im, nn * 8);
        memcpy(dnodal + 2 * nn, r->dVn_re, nn * 8); memcpy(dnodal + 3 * nn, r->dVn_im, nn * 8);
    }
    return 0;
}
int orc_get_gains(orc_model* m, double* g3, double* xn) {
    orc_resp* r = &m->resp;
    if (r->kind != RESP_VE) return 1;
    int L = r->lmax;
    if (g3) { memcpy(g3, r->gu, (L + 1) * 8); memcpy(g3 + L + 1, r->gn, (L + 1) * 8); memcpy(g3 + 2 * (L + 1), r->gv, (L + 1) * 8); }
    if (xn) { memcpy(xn, r->xUn, (size_t)L * r->nr * 8); memcpy(xn + (size_t)L * r->nr, r->xVn, (size_t)L * r->nr * 8); }
    return 0;
}
int orc_sht_analysis_c(orc_model* m, const double* f, double* q) { orc_sht_analysis(&m->sht, f, (cplx*)q); return 0; }
int orc_sht_synthesis_c(orc_model* m, const double* q, double* f) { orc_sht_synthesis(&m->sht, (const cplx*)q, f); return 0; }
int orc_sph_synthesis_c(orc_model* m, const double* q, double* vt, double* vp) { orc_sph_synthesis(&m->sht, (const cplx*)q, vt, vp); return 0; }
int orc_sph_analysis_c(orc_model* m, const double* vt, const double* vp, double* q) { orc_sph_analysis(&m->sht, vt, vp, (cplx*)q); return 0; }
double orc_surface_integral_c(orc_model* m, const double* f) { return orc_surface_integral(&m->sht, f); }
int orc_grid_c(orc_model* m, double* colat, double* w) {
    memcpy(colat, m->sht.colat, m->nl * 8); memcpy(w, m->sht.w, m->nl * 8); return 0;
}
int orc_tensor_synth_c(orc_model* m, const double* c, int nch, double* dyad) {
    if (!m->resp.tsh_ready) { orc_tsh_init(&m->resp.tsh, &m->sht); m->resp.tsh_ready = 1; }
    orc_tsh_synth(&m->resp.tsh, &m->sht, (const cplx*)c, nch, dyad); return 0;
}
int orc_tensor_analysis_c(orc_model* m, const double* dyad, int nch, double* c) {
    if (!m->resp.tsh_ready) { orc_tsh_init(&m->resp.tsh, &m->sht); m->resp.tsh_ready = 1; }
    orc_tsh_analysis(&m->resp.tsh, &m->sht, dyad, nch, (cplx*)c); return 0;
}
int orc_tensor_n6_c(orc_model* m, double* n6) {
    if (!m->resp.tsh_ready) { orc_tsh_init(&m->resp.tsh, &m->sht); m->resp.tsh_ready = 1; }
    memcpy(n6, m->resp.tsh.n6, (m->sht.lmax + 1) * 8); return 0;
}
int orc_tensor_n4_c(orc_model* m, double* n4) {
    if (!m->resp.tsh_ready) { orc_tsh_init(&m->resp.tsh, &m->sht); m->resp.tsh_ready = 1; }
    memcpy(n4, m->resp.tsh.n4, (m->sht.lmax + 1) * 8); return 0;
}
int orc_tor_synthesis_c(orc_model* m, const double* q, double* vt, double* vp) { orc_tor_synthesis(&m->sht, (const cplx*)q, vt, vp); return 0; }
int orc_tor_analysis_c(orc_model* m, const double* vt, const double* vp, double* q) { orc_tor_analysis(&m->sht, vt, vp, (cplx*)q); return 0; }
int orc_rot_info_c(orc_model* m, double* out) {
    out[0] = m->rot.kTe; out[1] = m->rot.hTe; out[2] = m->rot.k_s_fluid; out[3] = m->rot.k_s_flat;
    out[4] = m->rot.dt_fe_max; out[5] = creal(m->rot.m); out[6] = cimag(m->rot.m);
    return 0;
}

/* ------------------------------------------------- component-level probes */

/* Dense A (column-major n x n) -> COO nonzeros -> pivoted band LU -> solve. */
int orc_test_band(int n, const double* A, const double* b, double* x, int* kl_ku) {
    int nz = 0;
    for (int t = 0; t < n * n; t++) nz += A[t] != 0.0;
    int* rows = malloc(nz * sizeof(int)); int* cols = malloc(nz * sizeof(int)); double* vals = malloc(nz * 8);
    int p = 0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++)
            if (A[i + j * n] != 0.0) { rows[p] = i; cols[p] = j; vals[p] = A[i + j * n]; p++; }
    orc_band bd;
    int rc = orc_band_build(&bd, n, nz, rows, cols, vals);
    if (!rc) orc_band_solve(&bd, b, x);
    if (kl_ku) { kl_ku[0] = bd.kl; kl_ku[1] = bd.ku; }
    orc_band_free(&bd); free(rows); free(cols); free(vals);
    return rc;
}

/* Loading (tidal=0) or tidal (tidal=1) Love numbers h,l,k at degree j for
 * the earth in p (src/vilma_radial_fe.f90:837-893); fluidize=1 turns every
 * Maxwell layer into a fluid (the t -> infinity limit). */
int orc_love(const vh_params* p, int j, int tidal, int fluidize, double* hlk, char* err) {
    orc_earth e; orc_mesh m; orc_radop op;
    if (orc_build_earth(p, &e, err)) return 1;
    if (fluidize)
        for (int i = 0; i < e.n; i++)
            if (e.L[i].rheology == RHEOL_MAXWELL) { e.L[i].mu = 0.0; e.L[i].rheology = RHEOL_FLUID; }
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return 1;
    if (orc_radop_assemble(&op, &e, &m, j, err)) { orc_mesh_free(&m); return 1; }
    double *b = malloc(op.ndof * 8), *x = malloc(op.ndof * 8);
    if (tidal) orc_radop_tidal_rhs(&op, 1.0, b); else orc_radop_load_rhs(&op, 1.0, b);
    orc_radop_solve(&op, b, x, NULL);
    double ua = x[idx_u(m.nr - 1)], va = x[idx_v(m.nr - 1)], fa = x[idx_f(m.nr - 1)];
    double a = e.r_earth, g = orc_gravity_at(&e, a);
    double phi = tidal ? 1.0 : 4.0 * ORC_PI * ORC_G * a * 1.0 / (2.0 * j + 1.0);
    hlk[0] = g * ua / phi; hlk[1] = g * va / phi; hlk[2] = -fa / phi - 1.0;
    free(b); free(x); orc_radop_free(&op); orc_mesh_free(&m);
    return 0;
}

/* Dense operator of degree j (with the degree-1 E_uniq penalty) for the
 * assembly tests; returns nd. */
int orc_dense_operator(const vh_params* p, int j, double* A_out, int cap, char* err) {
    orc_earth e; orc_mesh m;
    if (orc_build_earth(p, &e, err)) return -1;
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return -1;
    int nd = 4 * m.nr - 1;
    if (A_out && (long)nd * nd <= cap) {
        double* A = orc_build_dense_operator(&e, &m, j, 1);
        memcpy(A_out, A, (size_t)nd * nd * 8);
        free(A);
    }
    orc_mesh_free(&m);
    return nd;
}

/* Toroidal W operator of degree j on the mesh of p, dense nr x nr column-
 * major, before pinning and bordering (build_toroidal_operator,
 * src/vilma_radial_fe.f90:699-726); returns nr. With x and b given, also the
 * assembled (pinned, bordered) solve x = A^-1 b. */
int orc_tor_operator_c(const vh_params* p, int j, double* A_out, int cap, const double* b, double* x, char* err) {
    orc_earth e; orc_mesh m;
    if (orc_build_earth(p, &e, err)) return -1;
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return -1;
    int nr = m.nr;
    double* mu = malloc(m.ne * 8);
    for (int k = 0; k < m.ne; k++) mu[k] = e.L[m.lay[k]].mu;   /* as the response's element shear */
    if (A_out && (long)nr * nr <= cap) {
        double* A = orc_build_tor_operator(m.r, mu, m.ne, j);
        memcpy(A_out, A, (size_t)nr * nr * 8);
        free(A);
    }
    if (b && x) {
        orc_torop op;
        if (orc_torop_assemble(&op, m.r, mu, m.ne, j, err)) { free(mu); orc_mesh_free(&m); return -1; }
        orc_torop_solve(&op, b, x);
        orc_torop_free(&op);
    }
    free(mu); orc_mesh_free(&m);
    return nr;
}

/* per-element shear modulus and rheology of p's mesh (as the response sees
 * them); returns ne */
int orc_mesh_props_c(const vh_params* p, double* mu, int* rheo, int cap, char* err) {
    orc_earth e; orc_mesh m;
    if (orc_build_earth(p, &e, err)) return -1;
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return -1;
    int ne = m.ne;
    if (mu && ne <= cap)
        for (int k = 0; k < ne; k++) { mu[k] = e.L[m.lay[k]].mu; rheo[k] = e.L[m.lay[k]].rheology; }
    orc_mesh_free(&m);
    return ne;
}

/* The toroidal forcing of the memory tau = 2 mu eps(W) of a nodal W at degree
 * j, through strain_coeffs_tor and dissipative_rhs_tor (channels 4, 5 of a
 * 6-channel memory); f has nr entries. */
int orc_tor_forcing_c(const vh_params* p, int j, const double* W, double* f, char* err) {
    orc_earth e; orc_mesh m;
    if (orc_build_earth(p, &e, err)) return -1;
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return -1;
    int ne = m.ne;
    double Jr = (double)j * (j + 1.0);
    double *A = calloc((size_t)6 * ne, 8), *B = calloc((size_t)6 * ne, 8), *C = calloc((size_t)6 * ne, 8);
    for (int k = 0; k < ne; k++) {
        double a[2], b[2], c[2], mu = e.L[m.lay[k]].mu;
        orc_strain_coeffs_tor(W[k], W[k + 1], Jr, a, b, c);
        for (int q = 0; q < 2; q++) {
            A[k * 6 + 4 + q] = 2.0 * mu * a[q]; B[k * 6 + 4 + q] = 2.0 * mu * b[q]; C[k * 6 + 4 + q] = 2.0 * mu * c[q];
        }
    }
    double norm[2], sa[2][2], sb[2][2], sc[2][2];
    orc_strain_constants_tor(Jr, norm, sa, sb, sc);
    memset(f, 0, (size_t)m.nr * 8);
    orc_dissipative_rhs_tor(ne, m.r, sa, sb, sc, norm, A, B, C, f);
    free(A); free(B); free(C); orc_mesh_free(&m);
    return m.nr;
}

int orc_mesh_c(const vh_params* p, double* r, int* lay, int cap, char* err) {
    orc_earth e; orc_mesh m;
    if (orc_build_earth(p, &e, err)) return -1;
    if (orc_mesh_build(&e, p->n_elem_layer, &m, err)) return -1;
    int nr = m.nr;
    if (r && nr <= cap) { memcpy(r, m.r, nr * 8); memcpy(lay, m.lay, m.ne * sizeof(int)); }
    orc_mesh_free(&m);
    return nr;
}

void orc_ocean_function_c(int n, const double* topo, const double* ice, double* C) { orc_ocean_function(n, topo, ice, C); }

/* Degree-1 bordered band factor (test probe): copies ab (ldab x n) and ipiv
 * when the buffers are large enough; returns n, ldab, kl, ku in info. */
int orc_deg1_factor_c(orc_model* m, double* ab, int* ipiv, long cap, int* info) {
    const orc_band* b = &m->resp.ops[0].sys.band;
    info[0] = b->n; info[1] = b->ldab; info[2] = b->kl; info[3] = b->ku;
    if (ab && (long)b->ldab * b->n <= cap) {
        memcpy(ab, b->ab, (size_t)b->ldab * b->n * 8);
        memcpy(ipiv, b->ipiv, b->n * sizeof(int));
    }
    return 0;
}
