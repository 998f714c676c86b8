// vh_adjoint.h -- host wrappers of the reverse-mode kernels (vh_adjoint.hip,
// drift transpose in vh_drift.hip).
#pragma once
#include "vh_internal.h"

struct adv_adj_args {
    int nr, nk, nem, nlm;
    double dt;
    vh_dev v;
    const double* tauh_in;    // adjoint of tau_{n+1}
    double* tauh_out;         // adjoint of tau_n (overwritten)
    const double* tau;        // forward tau_n of the 1-D elements [e1of[em]][24][nk]
    const int* e1of;          // [nem] -> index of a 1-D element in tau, or -1
    const double* sigma;      // forward closing load transform (complex, degree-major)
    const double* dN;         // forward nodal drift [mn][4][nk]
    const int* e3of;          // [nem] -> 3-D element or -1
    const double2* Xh;        // 3-D increment adjoint, K'-scaled [i3*3 + s][lam][nlm]
    const double* n6;         // [L+1]
    double* Ue;               // [nem][8][nk] element nodal adjoints
    double* mpart;            // [nem][blocks] parts of dJ/dM_e
};
// tauh_out, dN_h, sigma_h (complex degree-major) and Kh[em] += dt dJ/dM_e
void vh_advance_adjoint(hipStream_t st, const vh_dims& d, const adv_adj_args& a, const int* mnodes,
                        const int2* node_el, double* dNh, double* sigh, double* Kh);
// blocks of k_adv_adj along the slots (the mpart row length)
inline int vh_advance_adjoint_nbx(const vh_dims& d) { return (d.nk + 63) / 64; }

// Xa = K' tau_h of the 3-D elements, tensor coefficient layout
void vh_t3_pack_adj(hipStream_t st, const vh_dims& d, const vh_dev& v, const int* e3em, int ne3d, const double* n6,
                    const double* tauh, double2* Xa);
// adjoint row pass of ne3 whole elements (3 batches each); rate_h += rate gradient
void vh_t3_rows_adj(hipStream_t st, const vh_sht& s, double2* Fa, const double2* Ff, int ne3, const double* rate,
                    double dt, double* rate_h);

// mode 2: y = x - C I/C_int; mode 4: the ice adjoints from t (and I(b)); mode 6: y += x
struct sle_adj_grid_args {
    int mode, n, nphi;
    double dlon, inv_cint, frac;
    const double* w;
    const double* C;
    const double* rh;
    const double* t;
    const double* I;   // device scalar (integral) or null
    double *x, *b, *y, *ih0, *ih1;
};
void vh_sle_adj_grid(hipStream_t st, const sle_adj_grid_args& a);
// x = rho_w C t (+ rh/W if first), b = x (first) or b += x, part[j] = row sums of x
void vh_sle_adj_row(hipStream_t st, const vh_sht& s, const double* C, const double* t, const double* rh, int first,
                    double* x, double* b, double* part);
void vh_sle_adj_spec(hipStream_t st, int nlm, int nk, int mode, const double2* src, double2* q, const double2* q1,
                     const double* gu, const double* gn, double g, double* dsurfh);

void vh_pert_grad(hipStream_t st, int ne, size_t ng, const int* e3of_mesh, const int* em_of, const double* MkPerDt,
                  const double* rate3, const double* rate_h, const double* Kh, double* out);

// drift transpose (vh_drift.hip): Fh scratch [2 nmn][2][nk]; tau_h += ...
struct vh_drift_plan;
void vh_drift_adjoint(hipStream_t st, const vh_drift_plan* p, const vh_dims& d, const vh_dev& v,
                      const double* mnorm, const double* maxnorm, double skip_tol, int deg1_cm,
                      const double* dNh, const double* dsurfh, double* Fh, double* tauh);
const int2* vh_drift_plan_node_el(const vh_drift_plan* p);
