// vh_sle.h -- host wrappers of the sea-level-equation grid kernels (vh_sle.hip).
#pragma once
#include "vh_internal.h"

// also zeroes wcorr (n) and res[0..nres) (the solve's result slots)
void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0, double* wcorr,
             double* res, int nres);
// response_apply only: uN = [u; N] (degree-major) from sigma
void vh_apply_only(hipStream_t st, int nlm, const double2* sigma, const double* gu, const double* gn,
                   const double* dsurf, int nk, double g, int kind, double2* uN);
// out[0] = int C f
void sle_cs(hipStream_t st, const vh_sht& s, const double* C, const double* f, double* part, double* out);
// sle_solve's coastline passes and inner fixed point (src/vilma_sle.f90:231-355)
// as one cooperative launch: per pass (at most n_outer) the mask and
// grounded-ice increment with their integrals, the subgrid term, then at most
// n_inner iterations of load analysis, response, synthesis and rsl update,
// then the coastline test; every exit decision is taken on the device.
// F, FC, F2 are scratch spectra ([nlat][L+1], [nlat][L+1], [(nlat+1)/2][L+1][4]);
// part holds [2][nlat] row maxima, [nlat] flip counts, [nlat] row sums cs;
// part3 [3][nlat] the pass integrand row sums. On return C, dig, wcorr, rsl
// and Cn are those of the last pass and sigma is the spectrum of the load of
// the final rsl (the solve's closing analysis).
// res (zeroed by the caller): [0] iterations of the last pass, [1] its last
// dphi, [2] 1 if a grid barrier wait timed out, [3] C_int of the last pass,
// [4] iterations over all passes, [5] passes with ocean (C_int > 0), [6], [7]
// int d_ice_g and int topo0 (C - C0) of the last such pass, [10] int C rsl of
// the result (int C Sraw + dphi C_int of the last iterate; 0 without ocean).
// bar (VH_SLE_BAR_WORDS words, one 4 KiB page per counter): [1] timeout
// flag, [2] barriers passed (continues across launches); at
// VH_SLE_FLAG_STRIDE (1 + k), k < 8, the release words, at
// VH_SLE_FLAG_STRIDE 9 the top arrival counter, at VH_SLE_FLAG_STRIDE (10 + g)
// the arrival counter of block group g = blockIdx % 8.
#define VH_SLE_FLAG_STRIDE 1024
#define VH_SLE_BAR_WORDS (18 * VH_SLE_FLAG_STRIDE)
struct sle_solve_args {
    vh_sht s;
    const double *topo0, *ice, *dice, *C0, *srot, *gu, *gn, *dsurf;
    double *C, *Cn, *dig, *wcorr;
    int nk, kind, subgrid, n_inner, n_outer, fixed_ocean;
    double g, tol;
    double *rsl, *cs, *part, *part3, *res;
    double2 *F, *FC, *F2, *sigma;
    unsigned* bar;
};
void sle_solve_launch(hipStream_t st, const sle_solve_args& a);
int sle_solve_grid(const vh_sht& s);
// most barrier-counter steps one launch can take
long sle_solve_max_barriers(const vh_sht& s, int n_outer, int n_inner);
void sle_bar_reset(hipStream_t st, unsigned* bar);   // counter, flag and start value to 0
void grid_sub(hipStream_t st, int n, const double* a, const double* b, double* out);
void grid_ocean(hipStream_t st, int n, const double* topo, const double* ice, double* C);
void grid_bsl_integrand(hipStream_t st, int n, const double* h, const double* C, const double* heq,
                        const double* C0, double* out);
