// vh_sle.h -- host wrappers of the sea-level-equation grid kernels (vh_sle.hip).
#pragma once
#include "vh_internal.h"

void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0);
// C (first pass: from topo0 - rsl; later: C_next), d_ice_g, and sc[0..2] =
// (int C, int d_ice_g, int topo0 (C - C0))
void sle_outer(hipStream_t st, const vh_sht& s, int first, const double* topo0, const double* rsl,
               const double* ice, const double* dice, const double* C0, const double* Cnext, double* C,
               double* dig, double* part, double* sc);
void sle_wcorr(hipStream_t st, int n, int subgrid, const double* topo0, const double* C, const double* C0, double* w);
// load -> load_lm (degree-major)
void sle_load_analysis(hipStream_t st, const vh_sht& s, const double* dig, const double* C, const double* rsl,
                       const double* wcorr, double2* F, double2* load_lm);
// response_apply only: uN = [u; N] (degree-major) from sigma
void vh_apply_only(hipStream_t st, int nlm, const double2* sigma, const double* gu, const double* gn,
                   const double* dsurf, int nk, double g, int kind, double2* uN);
// out[0] = int C f
void sle_cs(hipStream_t st, const vh_sht& s, const double* C, const double* f, double* part, double* out);
// C_next and out[0] = number of flipped cells
void sle_flip(hipStream_t st, const vh_sht& s, const double* topo0, const double* rsl, const double* ice,
              const double* C, double* Cnext, double* part, double* out);
// The inner fixed point of the sea-level equation (src/vilma_sle.f90:296-330)
// as one cooperative launch: at most n_inner iterations of load analysis,
// response, synthesis and rsl update, with the exit test on the device.
// F, FC, F2 are scratch spectra ([nlat][L+1], [nlat][L+1], [(nlat+1)/2][L+1][4]).
// On return rsl is updated and sigma is the spectrum of the load of the final
// rsl under the current mask (the solve's closing analysis).
// res[0] = iterations run, res[1] = dphi of the last one, res[2] = 1 if a grid
// barrier wait timed out. bar: [0] barrier counter, at bar_base on entry, raised
// by the grid size at each of the launch's 2 + 2 res[0] barriers; [1] timeout flag.
struct sle_inner_args {
    vh_sht s;
    const double *dig, *C, *wcorr, *srot, *gu, *gn, *dsurf, *sc;
    int nk, kind, subgrid, n_inner;
    double g, tol;
    double *rsl, *cs, *part, *res;
    double2 *F, *FC, *F2, *sigma;
    unsigned* bar;
    unsigned bar_base;
};
void sle_inner(hipStream_t st, const sle_inner_args& a);
int sle_inner_grid(const vh_sht& s);
void sle_bar_reset(hipStream_t st, unsigned* bar);   // counter and flag to 0
void grid_sub(hipStream_t st, int n, const double* a, const double* b, double* out);
void grid_ocean(hipStream_t st, int n, const double* topo, const double* ice, double* C);
void grid_bsl_integrand(hipStream_t st, int n, const double* h, const double* C, const double* heq,
                        const double* C0, double* out);
