// vh_sle.h -- host wrappers of the sea-level-equation grid kernels (vh_sle.hip).
#pragma once
#include "vh_internal.h"

// Synchronisation area of the fused solve (unsigned words; every counter on a
// 4 KiB page of its own, the claim words one per 64-byte line):
//   [1] error flag (a wait exceeded its bound), [2] items run by a block
//   other than their owner;
//   VH_SLE_FLAG_STRIDE (1 + g), g < 8: release words (stages completed);
//   VH_SLE_FLAG_STRIDE 9: the top arrival counter;
//   VH_SLE_FLAG_STRIDE (10 + g): the arrival counter of item group g = item % 8;
//   VH_SLE_CLAIM_OFF + VH_SLE_CLAIM_STRIDE i: claim word of work item i
//   (1 + the last stage it was claimed for);
//   VH_SLE_CTL_OFF: the control records of four consecutive stages (doubles).
// sle_ice resets it before every solve.
#define VH_SLE_FLAG_STRIDE 1024
#define VH_SLE_MAX_ITEMS 1024
#define VH_SLE_CLAIM_STRIDE 16
#define VH_SLE_CLAIM_OFF (18 * VH_SLE_FLAG_STRIDE)
#define VH_SLE_CTL_OFF (34 * VH_SLE_FLAG_STRIDE)
#define VH_SLE_BAR_WORDS (35 * VH_SLE_FLAG_STRIDE)

// Interpolated ice, d_ice and C0 of the sub-step; also zeroes wcorr (n),
// res[0..nres) (the solve's result slots) and resets the solve's
// synchronisation area bar for nitems work items and a first pass (n_outer >= 1).
void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0, double* wcorr,
             double* res, int nres, unsigned* bar, int nitems, int n_outer);
// response_apply only: uN = [u; N] (degree-major) from sigma
void vh_apply_only(hipStream_t st, int nlm, const double2* sigma, const double* gu, const double* gn,
                   const double* dsurf, int nk, double g, int kind, double2* uN);
// out[0] = int C f
void sle_cs(hipStream_t st, const vh_sht& s, const double* C, const double* f, double* part, double* out);
// sle_solve's coastline passes and inner fixed point (src/vilma_sle.f90:231-355)
// in one launch of ordinary blocks (no co-residency needed): per pass (at most
// n_outer) the mask and grounded-ice increment with their integrals, the
// subgrid term, then at most n_inner iterations of load analysis, response,
// synthesis and rsl update, then the coastline test; every exit decision is
// taken on the device.
// F, FC, F2 are scratch spectra ([nlat][L+1], [nlat][L+1], [(nlat+1)/2][L+1][4]);
// sraw a grid field of scratch; part holds [2][nlat] row maxima, [nlat] flip
// counts, [nlat] row sums cs; part3 [3][nlat] the pass integrand row sums. On
// return C, dig, wcorr, rsl and Cn are those of the last pass and sigma is the
// spectrum of the load of the final rsl (the solve's closing analysis).
// res (zeroed by sle_ice): [0] iterations of the last pass, [1] its last
// dphi, [2] 1 if a wait exceeded its bound (an error), [3] C_int of the last
// pass, [4] iterations over all passes, [5] passes with ocean (C_int > 0),
// [6], [7] int d_ice_g and int topo0 (C - C0) of the last such pass, [8]
// work items run by a block other than their owner, [10] int C rsl of the
// result (int C Sraw + dphi C_int of the last iterate; 0 without ocean).
struct sle_solve_args {
    vh_sht s;
    const double *topo0, *ice, *dice, *C0, *srot, *gu, *gn, *dsurf;
    double *C, *Cn, *dig, *wcorr, *sraw;
    int nk, kind, subgrid, n_inner, n_outer, fixed_ocean;
    double g, tol;
    double *rsl, *cs, *part, *part3, *res;
    double2 *F, *FC, *F2, *sigma;
    unsigned* bar;
    // diagnostic: the last late_blocks blocks start late_ticks (100 MHz) late
    int late_blocks;
    unsigned late_ticks;
    // blocks of the launch (<= 0: one per work item); block b owns items b,
    // b + blocks, ... (fewer blocks leave the device to other work, e.g. the
    // other members of an ensemble)
    int blocks;
};
// launch = false: only the kernel's LDS attribute is set (before a stream
// capture, where a function attribute cannot be changed)
void sle_solve_launch(hipStream_t st, const sle_solve_args& a, bool launch = true);
int sle_solve_grid(const vh_sht& s);   // work items of the fused solve
void grid_sub(hipStream_t st, int n, const double* a, const double* b, double* out);
void grid_absdiff(hipStream_t st, int n, const double* a, const double* b, double* out);   // |a - b|
void grid_ocean(hipStream_t st, int n, const double* topo, const double* ice, double* C);
void grid_bsl_integrand(hipStream_t st, int n, const double* h, const double* C, const double* heq,
                        const double* C0, double* out);
