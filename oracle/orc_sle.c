/*
 * orc_sle.c -- oracle: pseudo-spectral, mass-conserving sea-level equation
 * with migrating coastlines and grounded-ice flotation.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). Restates src/vilma_sle.f90:124-405
 * (sle_solve, with the memory co-convergence passes of the implicit scheme)
 * and :407-425 (ocean_function).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* C = 1 iff topo < 0 and rho_i*I < -rho_w*topo (src/vilma_sle.f90:407-425). */
void orc_ocean_function(int n, const double* topo, const double* ice, double* C) {
    for (int i = 0; i < n; i++)
        C[i] = (topo[i] < 0.0 && ORC_RHO_ICE * ice[i] < -ORC_RHO_WATER * topo[i]) ? 1.0 : 0.0;
}

void orc_sle_solve(orc_sle* sle, const orc_sht* s, orc_resp* resp, const double* d_ice,
                   const double* ice, const double* topo0, double* rsl, double* C,
                   orc_sle_result* res, int report_only, cplx* sigma_out, const double* s_rot) {
    int ng = s->nphi * s->nlat, nlm = s->nlm;
    sle->n_solve++;
    double *load = malloc(ng * 8), *u = calloc(ng, 8), *N = calloc(ng, 8), *Sraw = malloc(ng * 8);
    double *rsl_new = malloc(ng * 8), *C0 = malloc(ng * 8), *wcorr = calloc(ng, 8);
    double *C_next = malloc(ng * 8), *d_ice_g = calloc(ng, 8), *tmp = malloc(ng * 8), *ref = malloc(ng * 8);
    cplx *load_lm = malloc(nlm * sizeof(cplx)), *u_lm = malloc(nlm * sizeof(cplx)), *N_lm = malloc(nlm * sizeof(cplx));
    double rho_ratio = ORC_RHO_ICE / ORC_RHO_WATER;
    if (!sle->warm_start) memset(rsl, 0, ng * 8);
    double dphi = 0.0, ice_int = 0.0, zeta_int = 0.0, C_int = 0.0;
    memset(res, 0, sizeof(*res));
    for (int i = 0; i < ng; i++) ref[i] = ice[i] - d_ice[i];
    orc_ocean_function(ng, topo0, ref, C0);
    orc_resp_begin(resp);
    /* the sea-level <-> memory co-convergence (src/vilma_sle.f90:218-366):
     * FE converges in one pass; TRAP re-converges the load against each new
     * tau_{n+1} estimate until the report drift settles */
    if (!report_only) orc_resp_prepare_endpoint(resp);
    int n_mem = report_only ? 1 : sle->max_mem_iter;
    for (int im = 1; im <= n_mem; im++) {
        for (int io = 1; io <= sle->n_outer; io++) {
            if (sle->fixed_ocean) {
                if (io == 1) memcpy(C, C0, ng * 8);
            } else if (io == 1) {
                for (int i = 0; i < ng; i++) tmp[i] = topo0[i] - rsl[i];
                orc_ocean_function(ng, tmp, ice, C);
            } else {
                memcpy(C, C_next, ng * 8);
            }
            for (int i = 0; i < ng; i++) d_ice_g[i] = ice[i] * (1.0 - C[i]) - (ice[i] - d_ice[i]) * (1.0 - C0[i]);
            C_int = orc_surface_integral(s, C);
            if (C_int <= 0.0) break;
            ice_int = -rho_ratio * orc_surface_integral(s, d_ice_g);
            if (sle->subgrid) {
                for (int i = 0; i < ng; i++) { wcorr[i] = -ORC_RHO_WATER * topo0[i] * (C[i] - C0[i]); tmp[i] = topo0[i] * (C[i] - C0[i]); }
                zeta_int = orc_surface_integral(s, tmp);
            } else {
                memset(wcorr, 0, ng * 8); zeta_int = 0.0;
            }
            for (int ii = 1; ii <= sle->n_inner; ii++) {
                for (int i = 0; i < ng; i++) load[i] = ORC_RHO_ICE * d_ice_g[i] + ORC_RHO_WATER * (C[i] * rsl[i]) + wcorr[i];
                orc_sht_analysis(s, load, load_lm);
                orc_resp_apply(resp, s, load_lm, u_lm, N_lm);
                orc_sht_synthesis(s, u_lm, u);
                orc_sht_synthesis(s, N_lm, N);
                for (int i = 0; i < ng; i++) Sraw[i] = N[i] - u[i];
                if (s_rot) for (int i = 0; i < ng; i++) Sraw[i] = Sraw[i] + s_rot[i];
                for (int i = 0; i < ng; i++) tmp[i] = C[i] * Sraw[i];
                double Cs_int = orc_surface_integral(s, tmp);
                dphi = (ice_int - Cs_int + zeta_int) / C_int;
                double dmax = 0.0, smax = 0.0;
                for (int i = 0; i < ng; i++) rsl_new[i] = Sraw[i] + dphi;
                for (int i = 0; i < ng; i++) dmax = fmax(dmax, fabs(C[i] * (rsl_new[i] - rsl[i])));
                memcpy(rsl, rsl_new, ng * 8);
                res->n_inner_last = ii; res->resid = dmax;
                for (int i = 0; i < ng; i++) smax = fmax(smax, fabs(C[i] * rsl[i]));
                if (dmax <= sle->tol * fmax(smax, DBL_MIN)) break;
            }
            res->n_outer_done = io;
            sle->n_outer_tot++;
            sle->n_inner_tot += res->n_inner_last;
            if (sle->fixed_ocean) break;
            for (int i = 0; i < ng; i++) tmp[i] = topo0[i] - rsl[i];
            orc_ocean_function(ng, tmp, ice, C_next);
            int flips = 0;
            for (int i = 0; i < ng; i++) flips += (C_next[i] != C[i]);
            res->n_coast_flip = flips;
            if (flips == 0) break;
        }
        for (int i = 0; i < ng; i++) load[i] = ORC_RHO_ICE * d_ice_g[i] + ORC_RHO_WATER * (C[i] * rsl[i]) + wcorr[i];
        orc_sht_analysis(s, load, load_lm);
        if (report_only) break;
        orc_resp_advance_endpoint(resp, s, load_lm);
        res->n_couple_done = im;
        if (resp->couple_done) break;
    }
    if (!report_only) { orc_resp_finalize_step(resp); sle->n_couple_tot += res->n_couple_done; }
    if (sigma_out) memcpy(sigma_out, load_lm, nlm * sizeof(cplx));
    for (int i = 0; i < ng; i++) tmp[i] = C[i] * rsl[i];
    double Cs_int = orc_surface_integral(s, tmp) - zeta_int;
    res->ocean_frac = C_int / (16.0 * atan(1.0));
    res->mass_resid = fabs(ice_int) > 0.0 ? fabs(Cs_int - ice_int) / fabs(ice_int) : fabs(Cs_int);
    res->esl = dphi;
    free(load); free(u); free(N); free(Sraw); free(rsl_new); free(C0); free(wcorr); free(C_next);
    free(d_ice_g); free(tmp); free(ref); free(load_lm); free(u_lm); free(N_lm);
}
