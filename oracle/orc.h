/*
 * orc.h -- internal declarations of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is a plain-C/OpenMP restatement of the
 * reference's viscoelastic time step (VILMA v2), used as the correctness
 * checker for the HIP product path and as the CPU baseline in bench.py. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Each routine cites the reference file:line it restates.
 */
#ifndef ORC_H
#define ORC_H

#include <complex.h>
#include <stdint.h>
#include "../include/vilma_hip.h"

typedef double complex cplx;

#define ORC_PI 3.14159265358979323846264338327950288
#define ORC_G 6.67430e-11          /* src/vilma_constants.f90:18    */
#define ORC_RHO_ICE 931.0          /* src/vilma_constants.f90:27    */
#define ORC_RHO_WATER 1000.0       /* src/vilma_constants.f90:28    */
#define ORC_SEC_PER_YEAR 31556926.0 /* src/vilma_constants.f90:34   */
#define ORC_NLAM 4

enum { RHEOL_ELASTIC = 0, RHEOL_MAXWELL = 1, RHEOL_FLUID = 2 };
enum { RESP_NULL = 0, RESP_ELASTIC = 1, RESP_VE = 2 };

/* ---------------------------------------------------------------- earth */
typedef struct { double r_bot, r_top, rho, mu, eta; int rheology; } orc_layer;
typedef struct { int n; double r_earth, r_core; orc_layer L[VH_MAX_LAYER]; } orc_earth;

int    orc_build_earth(const vh_params* p, orc_earth* e, char* err);
double orc_gravity_at(const orc_earth* e, double r);

/* ----------------------------------------------------------------- mesh */
typedef struct { int nr, ne; double* r; int* lay; } orc_mesh;
int  orc_mesh_build(const orc_earth* e, const int* n_elem_layer, orc_mesh* m, char* err);
void orc_mesh_free(orc_mesh* m);

/* ------------------------------------------------------------ band LU */
typedef struct { int n, kl, ku, kv, ldab; double* ab; int* ipiv; } orc_band;
int  orc_band_build(orc_band* b, int n, int nz, const int* rows, const int* cols,
                    const double* vals);          /* 0-based COO; 0 = ok */
void orc_band_solve(const orc_band* b, const double* rhs, double* x);
void orc_band_free(orc_band* b);

typedef struct { int nd, nb, ns; orc_band band; double *dr, *dc, *w; } orc_bband;

typedef struct {
    int j, nr, ne, ndof;
    double r_earth, g_surf;
    orc_bband sys;
    double* nullmode;     /* ndof, degree 1 only */
} orc_radop;

double* orc_build_dense_operator(const orc_earth* e, const orc_mesh* m, int j, int with_uniq);
int  orc_radop_assemble(orc_radop* op, const orc_earth* e, const orc_mesh* m, int j, char* err);
void orc_radop_solve(const orc_radop* op, const double* b, double* x, const double* border);
void orc_radop_load_rhs(const orc_radop* op, double sigma, double* b);
void orc_radop_tidal_rhs(const orc_radop* op, double phi_t, double* b);
void orc_radop_free(orc_radop* op);

static inline int idx_u(int k) { return 4 * k; }      /* node k, 0-based */
static inline int idx_v(int k) { return 4 * k + 1; }
static inline int idx_f(int k) { return 4 * k + 2; }
static inline int idx_p(int e) { return 4 * e + 3; }  /* element e       */

/* ------------------------------------------------------- Maxwell kernels */
void orc_strain_coeffs(double u1, double u2, double v1, double v2, double Jr,
                       double a[4], double b[4], double c[4]);
void orc_strain_constants(double Jr, double norm[4], double sa[4][4], double sb[4][4], double sc[4][4]);
/* A/B/C per slot are [e][ld]: ld = 4, or 6 with the toroidal channels */
void orc_dissipative_rhs(int ne, const double* r, double sa[4][4], double sb[4][4], double sc[4][4],
                         const double norm[4], const double* A, const double* B, const double* C,
                         int ld, double* f);
void orc_advance_memory_trap(int ne, const double* mu, const double* Mk, const double* Un, const double* Vn,
                             const double* Up, const double* Vp, double Jr, double* A, double* B, double* C,
                             int ld, const unsigned char* active);
void orc_advance_memory_fe(int ne, const double* mu, const double* Mk, const double* Un,
                           const double* Vn, double Jr, double* A, double* B, double* C, int ld,
                           const unsigned char* active);

/* toroidal degree of freedom (src/vilma_radial_fe.f90:699-817,
 * src/vilma_viscoelastic.f90:373-439, :591-629) */
typedef struct { int j, nr; orc_bband sys; unsigned char* pinned; } orc_torop;
double* orc_build_tor_operator(const double* r, const double* mu, int ne, int j);
int  orc_torop_assemble(orc_torop* op, const double* r, const double* mu, int ne, int j, char* err);
void orc_torop_solve(const orc_torop* op, const double* b, double* x);
void orc_torop_free(orc_torop* op);
void orc_strain_coeffs_tor(double w1, double w2, double Jr, double a[2], double b[2], double c[2]);
void orc_strain_constants_tor(double Jr, double norm[2], double sa[2][2], double sb[2][2], double sc[2][2]);
void orc_dissipative_rhs_tor(int ne, const double* r, double sa[2][2], double sb[2][2], double sc[2][2],
                             const double norm[2], const double* A, const double* B, const double* C, double* f);
void orc_advance_memory_tor(int ne, const double* mu, const double* Mk, const double* Wn, const double* Wp,
                            double Jr, double* A, double* B, double* C, const unsigned char* active);

/* ------------------------------------------------------------------ SHT */
typedef struct {
    int lmax, mmax, nlat, nphi, nlm, nh;
    double *colat, *ct, *st, *lon, *w;   /* w: Gauss weights, sum 2 */
    double *plm, *dplm;                  /* [mstart(m) + (l-m)]*nh + j, j < nh */
    int* mstart;                         /* offset of (m, l=m) in LM order */
    /* FFT plan */
    int nfac, fac[32];
    cplx* tw;                            /* nphi twiddles e^{-2 pi i k/nphi} */
} orc_sht;

int  orc_sht_init(orc_sht* s, int lmax, int nlat, int nphi, char* err);
void orc_sht_free(orc_sht* s);
static inline int orc_lmidx(const orc_sht* s, int l, int m) { return s->mstart[m] + (l - m); }
double orc_surface_integral(const orc_sht* s, const double* f);
void orc_sht_analysis(const orc_sht* s, const double* f, cplx* q);
void orc_sht_synthesis(const orc_sht* s, const cplx* q, double* f);
void orc_sph_synthesis(const orc_sht* s, const cplx* q, double* vt, double* vp);
void orc_sph_analysis(const orc_sht* s, const double* vt, const double* vp, cplx* q);
/* toroidal pair (vt, vp) = e_r x grad_1 (sum q Y) = (-F, E) and its analysis */
void orc_tor_synthesis(const orc_sht* s, const cplx* q, double* vt, double* vp);
void orc_tor_analysis(const orc_sht* s, const double* vt, const double* vp, cplx* q);
void orc_fft_rows_forward(const orc_sht* s, const double* f, int nrows, cplx* F);  /* m=0..lmax, scaled 2pi/nphi */
void orc_fft_rows_inverse(const orc_sht* s, const cplx* F, int nrows, double* f);
double orc_eval_point(const orc_sht* s, const cplx* q, double colat, double lon);

/* tensor spherical harmonics (src/vilma_tensor_sh.f90) */
typedef struct {
    int lmax, nlm, nphi, nlat;
    int *ldeg, *mord;
    double *llp1, *cott, *invsin, *n6, *n4;
} orc_tsh;
int  orc_tsh_init(orc_tsh* t, const orc_sht* s);
void orc_tsh_free(orc_tsh* t);
/* nch = 4 (lambda 1,2,5,6) or 6 (+ lambda 3,4 as channels 4, 5) */
void orc_tsh_synth(const orc_tsh* t, const orc_sht* s, const cplx* c /* nch*nlm, [lam][lm] */, int nch,
                   double* dyad /* 6 planes */);
void orc_tsh_analysis(const orc_tsh* t, const orc_sht* s, const double* dyad, int nch, cplx* c);

/* ------------------------------------------------------------- response */
typedef struct {
    int kind;
    int lmax, nr, ne, ndof, nlm;
    double g, a, dt, time;
    double skip_tol;
    int deg1_cm;
    /* elastic response */
    double *ugain, *ngain, *vgain;
    /* VE */
    orc_radop* ops;          /* [lmax], ops[l-1] */
    double *r, *mu, *Mk, *MkPerDt;
    unsigned char* active1d;
    double* Jr;              /* [lmax+1] */
    double (*nrmc)[4];
    double (*sa)[4][4], (*sb)[4][4], (*sc)[4][4];
    double *gu, *gn, *gv;    /* [lmax+1] */
    double *xUn, *xVn;       /* [l-1][nr] */
    int nk;
    int *k2lm, *kdeg, *kbeg;
    double *Are, *Aim, *Bre, *Bim, *Cre, *Cim;   /* [k][e][lam] */
    cplx *dUa, *dFa, *dVa;
    double *dUn_re, *dUn_im, *dVn_re, *dVn_im;   /* [k][node] */
    double* mnorm;
    /* lateral viscosity */
    int lat_visc, ne3d, tsh_ready;
    double visc3d_tol;
    double *Mk3, *MkPerDt3;  /* [e][nlat][nphi] */
    int* e3d;
    orc_tsh tsh;
    /* controller snapshot (buffer A) */
    double *Are_s, *Aim_s, *Bre_s, *Bim_s, *Cre_s, *Cim_s;
    double time_s;
    /* implicit (trapezoidal) memory, src/vilma_response.f90:1514-1663 */
    int implicit;
    double couple_tol;
    int couple_pass, couple_done;
    cplx *sigma_n, *sigma_next, *sigma_n_s, *dUa_prev;
    int sigma_primed, sigma_primed_s, have_next;
    double *Are0, *Aim0, *Bre0, *Bim0, *Cre0, *Cim0;   /* tau_n of the step */
    double *edUn_re, *edUn_im, *edVn_re, *edVn_im;     /* endpoint drift [k][node] */
    double *Are_c, *Aim_c, *Bre_c, *Bim_c, *Cre_c, *Cim_c;   /* coarse stash (buffer B) */
    /* toroidal channels (src/vilma_response.f90:126-146, :1055-1109): once
     * carried, every memory array is [k][e][6] with lambda 3, 4 at 4, 5 */
    int nlam, toroidal;
    orc_torop* tops;                                   /* [lmax], tops[l-1] */
    double (*nrmt)[2];
    double (*sat)[2][2], (*sbt)[2][2], (*sct)[2][2];
    cplx *dWa, *dWa_prev;                              /* [nk] surface W */
    double *dWn_re, *dWn_im, *edWn_re, *edWn_im;       /* [k][node] */
} orc_resp;

int  orc_resp_init_ve(orc_resp* r, const orc_earth* e, const orc_mesh* m, const orc_sht* s, double dt, char* err);
int  orc_resp_init_elastic(orc_resp* r, const orc_earth* e, const orc_mesh* m, int lmax, char* err);
void orc_resp_free(orc_resp* r);
void orc_resp_begin(orc_resp* r);
void orc_resp_apply(const orc_resp* r, const orc_sht* s, const cplx* sigma, cplx* u, cplx* n);
void orc_resp_advance(orc_resp* r, const orc_sht* s, const cplx* sigma);
void orc_resp_set_dt(orc_resp* r, double dt);
double orc_resp_max_rate(const orc_resp* r);
double orc_resp_memory_norm(const orc_resp* r);
void orc_resp_set_implicit(orc_resp* r, int nlm);
void orc_resp_prepare_endpoint(orc_resp* r);
void orc_resp_advance_endpoint(orc_resp* r, const orc_sht* s, const cplx* sigma);
void orc_resp_finalize_step(orc_resp* r);
void orc_resp_prime_sigma(orc_resp* r, const cplx* sigma);
void orc_resp_stash_coarse(orc_resp* r);
void orc_resp_coarse_fine_error(const orc_resp* r, double* err_inf, double* tau_inf);
void orc_resp_save(orc_resp* r);
void orc_resp_restore(orc_resp* r);
int  orc_resp_enable_lateral(orc_resp* r, const orc_sht* s, const double* pert, char* err);

/* ------------------------------------------------------------------ SLE */
typedef struct {
    int n_outer, n_inner, max_mem_iter, fixed_ocean, subgrid, warm_start;
    long n_couple_tot;
    double tol;
    long n_solve, n_outer_tot, n_inner_tot;
} orc_sle;
typedef struct {
    int n_outer_done, n_couple_done, n_inner_last, n_coast_flip;
    double resid, mass_resid, ocean_frac, esl;
} orc_sle_result;

void orc_ocean_function(int n, const double* topo, const double* ice, double* C);
void orc_sle_solve(orc_sle* sle, const orc_sht* s, orc_resp* resp, const double* d_ice,
                   const double* ice, const double* topo0, double* rsl, double* C,
                   orc_sle_result* res, int report_only, cplx* sigma_out, const double* s_rot);

/* ------------------------------------------------------------- rotation */
typedef struct {
    int nr, ne, ndof, tidal;
    double Jr, dt;
    orc_radop op;
    double *r, *mu, *Mk, *MkPerDt;
    double norm[4], sa[4][4], sb[4][4], sc[4][4];
    double Fe, Ue;
    double *xUn, *xVn;
    double *Are, *Aim, *Bre, *Bim, *Cre, *Cim;   /* [e][lam] */
    cplx dF, dU;
    double *dUn_re, *dUn_im, *dVn_re, *dVn_im;
} orc_ch2;

typedef struct {
    int enabled;
    cplx m;
    double time, a, g, CminusA, Omega, k_s, k_s_fluid, k_s_flat, kTe, hTe, dt_fe_max;
    cplx cload;
    orc_ch2 load_ch, tidal_ch;
} orc_rot;

int  orc_rot_init(orc_rot* rt, const orc_earth* e, const orc_mesh* m, double dt, char* err);
void orc_rot_free(orc_rot* rt);
void orc_rot_begin(orc_rot* rt, double dt);
void orc_rot_solve_m(orc_rot* rt, const orc_sht* s, const double* load);
void orc_rot_s_rot(const orc_rot* rt, const orc_sht* s, double* srot);
void orc_rot_update(orc_rot* rt, const orc_sht* s, const double* load, double dt);
int  orc_rotation_run(const vh_params* p, int lmax, int nlat, int nphi, const double* load, double dt,
                      int nsteps, double* m_out, double* info, char* err);

/* ---------------------------------------------------------------- model */
typedef struct orc_model {
    vh_params par;
    orc_sht sht;
    orc_earth earth;
    orc_mesh mesh;
    orc_resp resp;
    orc_sle sle;
    orc_rot rot;
    /* stepper (src/vilma_timestep.f90:38-75) */
    double cfl, dt_min, dt_max, guard_growth;
    double rtol, atol, safety, grow_max, shrink_min, dt_try;
    long n_accept, n_reject, n_solve, n_floor;
    double worst_mass_resid;
    /* gauss-grid state */
    int np, nl;
    double *z_bed_eq, *h_ice_eq, *h_ice, *rsl, *z_bed, *C, *C0, *s_rot;
    cplx* sigma_lm;
    cplx *u_lm, *n_lm;
    double time, bsl;
    char err[512];
} orc_model;

#endif
