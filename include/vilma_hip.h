/*
 * vilma_hip.h -- C-ABI of the MI355X-native VILMA v2 viscoelastic time step.
 *
 * This header is the drop-in boundary. A host (the Fortran shim module
 * `vilma` in fastearth3d_amd/fortran/vilma.f90, the Python mirror in
 * fastearth3d_amd/lib.py, or any C/C++ caller) drives the model through
 * plain pointers and sizes; no torch or HIP type appears in a signature.
 *
 * Conventions (all FP64):
 *   - Grid fields are (nphi, nlat) column-major with phi fastest, latitude rows
 *     north-first on the Gauss-Legendre nodes, longitudes 2*pi*(i-1)/nphi
 *     (reference: src/vilma_sht.f90:11-15, :120-123; src/vilma_coupling.f90:678-681).
 *   - Spectra are complex, interleaved (re, im), in the SHTns "LM" order of the
 *     reference: m-major, l contiguous, 0-based index m*(2*lmax+1-m)/2 + l for
 *     mres = 1 (src/vilma_sht.f90:349-354). Orthonormal harmonics without the
 *     Condon-Shortley phase; a real field is Q_l0 Y_l0 + 2 Re sum_{m>0} Q_lm Y_lm.
 *   - Times in the API are years where the reference API uses years
 *     (solid_earth_update dt_yr), seconds inside vh_params (vilma_par_load
 *     converts dt_* from years to seconds on load, src/vilma_params.f90:269-277).
 *   - Every function returns 0 on success and a nonzero status on failure;
 *     vh_last_error() returns the message. The reference aborts with
 *     `error stop 'message'` instead (e.g. src/vilma_coupling.f90:248,279);
 *     the Fortran shim turns a nonzero status back into `error stop message`.
 *   - A handle is not thread-safe; distinct handles may be used from different
 *     threads and processes at once, each on its own stream. No kernel needs
 *     its blocks to be resident together: the sea-level solve's blocks take
 *     over the work of blocks that have not started (VH_S_N_SLE_STEAL counts
 *     it), so a device shared with other work slows a solve but does not
 *     stop it.
 */
#ifndef VILMA_HIP_H
#define VILMA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define VH_MAX_LAYER 16

/* The physics/numerics record of the reference's `&vilma` group
 * (vilma_param_class, src/vilma_params.f90:26-205), restricted to the fields
 * the viscoelastic time step reads, plus two additive knobs marked below. */
typedef struct vh_params {
    int    lmax, nlat, nphi, mmax, mres;    /* nlat/nphi <= 0: 2*lmax+2 / 4*lmax   */
    double eps_polar;                       /* accepted, ignored: exact Legendre    */
    char   earth[64];                       /* "M3-L70-V01" | "custom" | "PREM"    */
    int    n_layer;
    double r_earth, r_core;
    double r_bot[VH_MAX_LAYER], r_top[VH_MAX_LAYER];
    double rho[VH_MAX_LAYER], mu[VH_MAX_LAYER], eta[VH_MAX_LAYER];
    int    rheology[VH_MAX_LAYER];          /* 0 elastic, 1 Maxwell, 2 fluid       */
    /* additive: explicit radial element count per layer (surface-first);
     * 0 keeps the reference depth-target rule of src/vilma_radial_fe.f90:140-224 */
    int    n_elem_layer[VH_MAX_LAYER];
    char   scheme[8];                       /* "fe" forward Euler | "trap" trapezoid */
    int    max_couple_iter;
    char   earth_response[8];               /* "ve" | "elastic" | "null"           */
    int    sle_n_outer, sle_n_inner;
    double sle_tol;
    int    sle_max_mem_iter, sle_fixed_ocean, sle_subgrid;
    double dt_init, dt_min, dt_max;         /* seconds                             */
    double rtol, atol, safety, grow_max, shrink_min, cfl;
    int    rotation;
    char   deg1_frame[8];                   /* "cm" | "cf"                         */
    int    l_visc_3d;
    double visc3d_tol;
    int    l_toroidal;
    double visc_log10_min, visc_log10_max, visc3d_lid_depth, visc3d_lid_log10max;
    /* additive: drift skip gate, the reference constant skip_tol of
     * src/vilma_response.f90:93 (1e-4); 0 solves every slot */
    double skip_tol;
} vh_params;

typedef struct vh_model vh_model;

/* Field / scalar / spectrum selectors for the getters. */
enum {
    VH_F_RSL = 0,      /* se%rsl   (full relative sea-level change)          */
    VH_F_ZBED = 1,     /* se%z_bed = z_bed_eq - rsl                          */
    VH_F_C = 2,        /* se%gg%C  ocean function 0/1                        */
    VH_F_C0 = 3,       /* se%gg%C0 reference ocean function                  */
    VH_F_SROT = 4,     /* se%gg%s_rot rotational sea-level contribution      */
    VH_F_HICE = 5      /* se%gg%h_ice                                        */
};
enum {
    VH_S_TIME = 0, VH_S_BSL = 1, VH_S_WORST_MASS_RESID = 2,
    VH_S_N_ACCEPT = 3, VH_S_N_REJECT = 4, VH_S_N_SOLVE = 5,
    VH_S_N_INNER_TOT = 6, VH_S_N_OUTER_TOT = 7,
    VH_S_POLAR_M_RE = 8, VH_S_POLAR_M_IM = 9,
    VH_S_N_SLE_STEAL = 10,  /* sea-level work items a block ran for another that
                               was not resident (a device shared with other work) */
    VH_S_N_FLOOR = 11,      /* trap: steps accepted at dt_min over tolerance     */
    VH_S_N_COUPLE_TOT = 12, /* sea-level <-> memory passes over all solves       */
    VH_S_N_GRAPH_LAUNCH = 13, /* sub-steps replayed as a hipGraph (VH_T_GRAPH)   */
    VH_S_N_GRAPH_CAPTURE = 14 /* sub-step graphs captured                        */
};
enum {
    VH_SP_SIGMA = 0,   /* se%sigma_lm: converged surface load [kg m^-2]      */
    VH_SP_U = 1,       /* u_lm = gu*sigma + dUa at the last report           */
    VH_SP_N = 2,       /* N_lm = gn*sigma - dFa/g                            */
    VH_SP_TOR = 3      /* t_lm: surface toroidal displacement e_r x grad_1
                          sum t Y (response_horizontal_toroidal,
                          src/vilma_response.f90:884-900); zero unless the
                          toroidal channels are carried                      */
};
enum {
    VH_D_LMAX = 0, VH_D_NLAT = 1, VH_D_NPHI = 2, VH_D_NLM = 3, VH_D_NR = 4,
    VH_D_NE = 5, VH_D_NK = 6, VH_D_NE3D = 7, VH_D_NLAM = 8, VH_D_NEM = 9,
    VH_D_NMN = 10   /* memory nodes: nodes of Maxwell elements */
};

/* Defaults of vilma_param_class overlaid with input/vilma_defaults.nml. */
int vh_params_defaults(vh_params* p);

/* solid_earth_init (src/vilma_coupling.f90:144-232), passthrough grid
 * (:172-176): fields already on the model's Gauss grid, (nphi, nlat).
 * h_ice_init may be NULL (defaults to h_ice_eq). Calls
 * solid_earth_enable_visc_3d only through vh_enable_visc_3d* below: the
 * reference reads the lateral field from netCDF (:320-334), which the
 * boundary replaces by an in-memory array. */
int vh_create(const vh_params* p, int nphi, int nlat,
              const double* z_bed_eq, const double* h_ice_eq,
              const double* h_ice_init, vh_model** out);

/* response_enable_lateral_visc (src/vilma_response.f90:994-1053):
 * pert_elem is the log10-viscosity perturbation per element, (nphi, nlat, ne). */
int vh_enable_visc_3d(vh_model* h, const double* pert_elem);

/* response_enable_lateral_visc_from_nodes (src/vilma_response.f90:1120-1169):
 * visc_node is absolute log10(eta), (nphi*nlat, nr); lid_depth <= 0 disables
 * the lid rule. */
int vh_enable_visc_3d_nodes(vh_model* h, const double* visc_node,
                            double lid_depth, double lid_log10max, double log10_cap);

/* solid_earth_update (src/vilma_coupling.f90:336-437). */
int vh_update(vh_model* h, const double* h_ice, double dt_yr);
/* The same step with the ice thickness already in device memory of the
 * model's GPU (a caller that keeps its forcing resident in HBM); no
 * counterpart in the reference, whose fields always live on the host. */
int vh_update_device(vh_model* h, const double* h_ice_device, double dt_yr);

/* solid_earth_spinup (src/vilma_coupling.f90:453-517, relax_hold :519-571):
 * bring the model to isostatic equilibrium under h_ice_lgm ((nphi, nlat), the
 * Gauss grid) holding the reference state as the datum: a zero-length call
 * seeds the entering ice; with pre_spinup_1d the model first relaxes with
 * its lateral viscosity (the last vh_enable_visc_3d* field) replaced by the
 * lateral mean, then the real field is restored; then the full model relaxes
 * for at most equil_time_max_yr years (<= 0: skipped). Each relaxation runs
 * coupling calls of growing length (200 yr, x 1.6, at most 60) until the mean
 * bed velocity falls below equil_rate_tol [m/yr], logging one line per call
 * to stdout as the reference does. The reference reads equil_time_max,
 * equil_rate_tol and pre_spinup_1d from the parameter record; here they are
 * arguments (the Fortran shim passes se%par's). */
int vh_spinup(vh_model* h, const double* h_ice_lgm, double equil_time_max_yr, double equil_rate_tol,
              int pre_spinup_1d);

/* Outputs the host reads from the solid_earth components (:71-140). */
int vh_get_field(vh_model* h, int field, double* out);           /* (nphi,nlat) */
int vh_get_scalar(vh_model* h, int which, double* out);
int vh_get_spectrum(vh_model* h, int which, double* out);        /* 2*nlm       */
int vh_get_dims(vh_model* h, int which, int* out);

/* Maxwell memory in the reference restart layout: six arrays
 * Are,Aim,Bre,Bim,Cre,Cim, each (nlam, ne, nk) column-major, concatenated
 * (src/vilma_io.f90:495-568 reads/writes tau_* in this order). */
int vh_get_memory(vh_model* h, double* tau);
int vh_set_memory(vh_model* h, const double* tau);

/* The rest of the restart state (src/vilma_io.f90:425-491): model time
 * (years), the rsl that seeds the next sea-level solve and the ice at the
 * call boundary ((nphi,nlat) each), and with rotation on the polar motion m
 * (re, im) and both degree-2 channels' memory (load, tidal), each
 * (NLAM=4, ne, 6) column-major with the last axis [Are, Aim, Bre, Bim, Cre,
 * Cim] (rotation_get_memory, src/vilma_rotation.f90:275-307). With the
 * memory (vh_get_memory / vh_set_memory) this continues a run bit for bit.
 * Any pointer may be NULL; the rotation arguments are zero / ignored with
 * rotation off. */
int vh_get_state(vh_model* h, double* time_yr, double* rsl, double* h_ice, double* rot_m,
                 double* rot_load_mem, double* rot_tidal_mem);
int vh_set_state(vh_model* h, double time_yr, const double* rsl, const double* h_ice, const double* rot_m,
                 const double* rot_load_mem, const double* rot_tidal_mem);
/* read_ve_state's cross-resolution restart (src/vilma_io.f90:515-552): tau
 * in the vh_set_memory layout of a run at a lower truncation L' (nk_src =
 * L'(L'+3)/2 slots, degree-grouped: its slots are the model's first nk_src)
 * into the low degrees of the memory, the higher degrees zero; sigma_n of the
 * trapezoidal scheme is re-primed by the next step. nk_src above the model's
 * nk (downsampling) or not a whole number of degrees is refused with the
 * reference's messages. */
int vh_set_memory_lowres(vh_model* h, const double* tau, int nk_src);
/* The step state a restart carries besides the fields (src/vilma_io.f90:
 * 453-457, :548-551): the adaptive controller's next step size dt_try
 * (years) and, for scheme "trap", the start-of-step load sigma_n (2*nlm,
 * SHTns LM order) with its primed flag. Any pointer may be NULL; a NULL
 * sigma_n on set leaves sigma_n to be re-primed. */
int vh_get_step_state(vh_model* h, double* dt_try_yr, int* sigma_primed, double* sigma_n);
int vh_set_step_state(vh_model* h, double dt_try_yr, int sigma_primed, const double* sigma_n);

/* ---- host lon-lat grid <-> the model's Gauss grid (src/vilma_remap.f90) */
typedef struct vh_remap_handle vh_remap_handle;
/* remap_init (src/vilma_remap.f90:56-99) for host cell-centre axes
 * lon_deg[nlon], lat_deg[nlat_ll] (degrees, strictly ascending, covering the
 * sphere; longitudes within one turn) against the Gauss grid of a model with
 * nphi x nlat points (vh_get_dims; the model itself is not needed).
 * Conservative weights host -> Gauss from cell boundaries at the axis
 * midpoints (latitude ends at the poles, longitude cyclic); bilinear weights
 * Gauss -> host. The fesm-utils `coords` maps the reference uses are
 * replaced by this construction for lon-lat host grids. The map owns a
 * stream; errors of the vh_remap_* calls are in vh_last_error(NULL). */
int vh_remap_create(int nphi, int nlat, int nlon, int nlat_ll, const double* lon_deg, const double* lat_deg,
                    vh_remap_handle** out);
/* remap_to_gauss (src/vilma_remap.f90:101-138): f_ll (nlon, nlat_ll), lon
 * fastest, latitude ascending -> f_gauss (nphi, nlat) in the model layout
 * (north-first); conserve_mass != 0 rescales the result so that its Gauss
 * quadrature integral equals the host area integral (ice thickness). */
int vh_remap_to_gauss(vh_remap_handle* m, const double* f_ll, double* f_gauss, int conserve_mass);
/* remap_to_gauss as vh_remap_to_gauss (same kernels, same results), for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): no counterpart in
 * the reference. It calls no hipMalloc, hipFree or allocating library
 * routine; the host field, the Gauss field and the two mass totals are staged
 * in one fixed device array per GPU of 1048576 doubles (8 MiB, held until
 * the process ends), so nlon*nlat_ll + nphi*nlat must not exceed 1048574
 * (larger grids are refused with an error). Calls on one GPU share that
 * array: calls that overlap in time on one GPU race. Needs a map created on
 * the same device (vh_remap_create sets the array up). */
int vh_remap_to_gauss_noalloc(vh_remap_handle* m, const double* f_ll, double* f_gauss, int conserve_mass);
/* remap_to_ll (src/vilma_remap.f90:140-164): bilinear, f_gauss -> f_ll. */
int vh_remap_to_ll(vh_remap_handle* m, const double* f_gauss, double* f_ll);
/* The same with device (HBM) pointers. */
int vh_remap_to_gauss_device(vh_remap_handle* m, const double* f_ll, double* f_gauss, int conserve_mass);
int vh_remap_to_ll_device(vh_remap_handle* m, const double* f_gauss, double* f_ll);
int vh_remap_destroy(vh_remap_handle* m);

/* solid_earth_finalize (src/vilma_coupling.f90:599-626). */
int vh_destroy(vh_model* h);

/* Message of the last failure on this handle (or of the last vh_create). */
const char* vh_last_error(const vh_model* h);

/* ---- Launch-shape knobs (no counterpart in the reference; the results do
 * not depend on them, bit for bit). VH_T_SLE_BLOCKS: blocks of the fused
 * sea-level solve, 0 (the default) one per work item (latitude pairs +
 * orders, lmax + 1 + nlat / 2), else that many blocks (at most the device's
 * CU count; a solve with more items than that is refused) owning the items
 * round-robin: fewer blocks solve one model more slowly but leave the device
 * to other work, e.g. the other members of an ensemble. VH_T_GRAPH: 1 replays
 * each forward-Euler sub-step (ice interpolation, drift, sea-level solve,
 * memory advance) as one captured hipGraph, one graph per (memory buffer,
 * interpolation fraction, step), captured on first use; 0 (the default)
 * launches the kernels one by one. */
enum { VH_T_SLE_BLOCKS = 0, VH_T_GRAPH = 1 };
int vh_set_tuning(vh_model* h, int which, int value);
int vh_get_tuning(vh_model* h, int which, int* value);

/* ---- Uneven ensemble (the vh_ensemble_* set). The reference steps an
 * ensemble as separate solid_earth instances (src/vilma_coupling.f90:127-176,
 * :336-437 per instance), one after another on the host. Here the members
 * are models from vh_create (any lmax, grid, element count, earth model,
 * lateral viscosity), still owned by the caller, advanced by a pool of
 * n_workers host threads (<= 0: 16; at most one per member) that take members
 * longest first and step each on its own stream, so that several members'
 * kernels share the GPU. Results equal those of stepping the members one
 * after another, bit for bit. The handle is not thread-safe, and a member
 * must not be used otherwise while an ensemble call runs. On failure the
 * call stops taking further members and returns nonzero;
 * vh_ensemble_failed_member and vh_ensemble_last_error name the first member
 * that failed and its message; members of a vh_ensemble_run_device call may
 * then stand at different times (vh_get_scalar VH_S_TIME). */
typedef struct vh_ensemble vh_ensemble;
int vh_ensemble_create(vh_model* const* members, int n_members, int n_workers, vh_ensemble** out);
/* one solid_earth_update of every member: h_ice[i] the host grid of member i
 * ((nphi, nlat) of that member), dt_yr[i] its interval */
int vh_ensemble_update(vh_ensemble* e, const double* const* h_ice, const double* dt_yr);
/* the same with each member's ice in device memory */
int vh_ensemble_update_device(vh_ensemble* e, const double* const* h_ice_device, const double* dt_yr);
/* ncalls updates of every member without a barrier between calls: member i's
 * call k takes the grid at ice_hist_device[i] + k * nphi_i * nlat_i (device
 * memory, [ncalls][nlat_i][nphi_i]) */
int vh_ensemble_run_device(vh_ensemble* e, int ncalls, const double* const* ice_hist_device, const double* dt_yr);
/* The ensemble's autotuner: runs 3 x ncalls real updates of every member
 * (candidate k on calls k*ncalls .. (k+1)*ncalls - 1 of the histories, laid
 * out as for vh_ensemble_run_device) with the candidate launch shapes of the
 * sea-level solve (VH_T_SLE_BLOCKS 0, 64, 32) and leaves every member on the
 * fastest, returned in *chosen_sle_blocks. The members advance by 3 * ncalls
 * calls, with the same results as any other setting would give. */
int vh_ensemble_tune(vh_ensemble* e, int ncalls, const double* const* ice_hist_device, const double* dt_yr,
                     int* chosen_sle_blocks);
int vh_ensemble_failed_member(const vh_ensemble* e);   /* -1: none */
const char* vh_ensemble_last_error(const vh_ensemble* e);
int vh_ensemble_destroy(vh_ensemble* e);            /* the members stay valid */

/* ---- Measurement hook (no counterpart in the reference): per-kernel-class
 * device time from HIP events recorded on the model's stream around each
 * launch. enable = 1 starts accumulating and clears the totals, 2 resumes
 * without clearing, 0 stops (the totals stay readable). */
enum {
    VH_K_DRIFT = 0,          /* drift: forcing + restricted-operator product   */
    VH_K_ADVANCE = 1,        /* memory advance (1-D update, 3-D increments)    */
    VH_K_TSH_SYNTH = 2,      /* 3-D: tensor Legendre synthesis                 */
    VH_K_T3_ROWS = 3,        /* 3-D: IFFT, -Mk3 product, FFT per latitude row  */
    VH_K_TSH_ANALYSIS = 4,   /* 3-D: tensor Legendre analysis                  */
    VH_K_T3_GATHER = 5,      /* 3-D: memory + 2 mu strain coefficients         */
    VH_K_SLE = 6,            /* sea-level equation: the fused solve (passes)   */
    VH_K_ADJ_ROWS = 7,       /* reverse: 3-D row pass with the rate gradient   */
    VH_K_ADJ_ADVANCE = 8,    /* reverse: memory advance transpose + node sums  */
    VH_K_ADJ_SLE = 9,        /* reverse: transposed sea-level solve            */
    VH_K_ADJ_DRIFT = 10,     /* reverse: G^T product + forcing transpose       */
    VH_K_NCLASS = 12
};
int vh_profile(vh_model* h, int enable);
/* ms[c] = summed device milliseconds, n[c] = number of timed launches */
int vh_profile_read(vh_model* h, double* ms /* VH_K_NCLASS */, double* n /* VH_K_NCLASS */);

/* ---- Parity probes: the response operator API the SLE and stepper use
 * (src/vilma_response.f90:43-52, dispatch :241-319). ---------------------- */

/* response_begin_step -> solve_drift (:694-809), skip gate included. */
int vh_resp_begin_step(vh_model* h);
/* response_apply (:842-864): u = gu*sigma + dUa, N = gn*sigma - dFa/g. */
int vh_resp_apply(vh_model* h, const double* sigma_lm, double* u_lm, double* n_lm);
/* response_advance_endpoint, forward-Euler branch (:1618-1633 -> :944-992). */
int vh_resp_advance(vh_model* h, const double* sigma_lm);
/* response_set_dt (:1665-1680), seconds. */
int vh_resp_set_dt(vh_model* h, double dt_s);
/* Frozen drift of the last begin_step: dUa,dFa,dVa as complex (2*nk each,
 * slot order k2lm of :629-648) and nodal dUn_re,dUn_im,dVn_re,dVn_im as
 * (nr, nk) column-major. Either pointer group may be NULL. */
int vh_get_drift(vh_model* h, double* dsurf /* 3*2*nk */, double* dnodal /* 4*nr*nk */);
/* The toroidal W drift of the last begin_step at every mesh node, dWn_re and
 * dWn_im as (nr, nk) column-major (the tor branch of solve_drift,
 * src/vilma_response.f90:694-809; zero at nodes without memory, and
 * everywhere unless the toroidal channels are carried). */
int vh_get_drift_tor(vh_model* h, double* dwn /* 2*nr*nk */);
/* Per-degree elastic gains gu, gn, gv (0..lmax) and unit-load nodal fields
 * xUn, xVn (nr, lmax) of response_init_ve (:599-627). */
int vh_get_gains(vh_model* h, double* g3 /* 3*(lmax+1) */, double* xn /* 2*nr*lmax */);
/* Scalar transforms on the model grid (sht_grid_analysis/synthesis,
 * src/vilma_sht.f90:204-226). */
int vh_sht_analysis(vh_model* h, const double* grid, double* lm);
int vh_sht_synthesis(vh_model* h, const double* lm, double* grid);
/* Tensor transforms on the model grid (tensor_sh_synth / tensor_sh_analysis,
 * src/vilma_tensor_sh.f90:168-322): c is [4][nlm] complex in LM order for
 * lambda = 1, 2, 5, 6; dyad is [6][nlat][nphi] for rr, rt, rp, tt, tp, pp.
 * The first call calibrates n6 (tensor_sh_init, :99-133). */
int vh_tensor_synthesis(vh_model* h, const double* c, double* dyad);
int vh_tensor_analysis(vh_model* h, const double* dyad, double* c);
/* The same with nch = 4 or 6 channels: with 6, c is [6][nlm] and channels 4, 5
 * are the toroidal lambda = 3, 4 (Z3 = e_r x grad_1, Z4 its spin-2 partner,
 * src/vilma_tensor_sh.f90:168-242, :360-387), normalised by J and n4. */
int vh_tensor_synthesis_ch(vh_model* h, int nch, const double* c, double* dyad);
int vh_tensor_analysis_ch(vh_model* h, int nch, const double* dyad, double* c);
/* The spin-2 normalisations n6(l), n4(l), l = 0..lmax, of the tensor analysis
 * (tensor_sh_init's calibration at m = l, src/vilma_tensor_sh.f90:99-133). */
int vh_tensor_norms(vh_model* h, double* n6, double* n4);

/* ---- Reverse mode (no counterpart in the reference, which has no adjoint).
 * Gradients of a misfit J(rsl_1..rsl_n) of the rsl after each of n coupling
 * calls with respect to the lateral log10-viscosity perturbation (pert_elem
 * of vh_enable_visc_3d) and the ice history, through hand-written transposes
 * of the step's kernels. vh_forward_record runs ncalls solid_earth_update
 * calls of dt_yr with h_ice_hist[c] (ncalls x (nphi,nlat)), keeps the model
 * state every ckpt_every calls and writes rsl after each call to rsl_out
 * (may be NULL). vh_backward takes dJ_drsl (ncalls x (nphi,nlat)), re-runs
 * each checkpoint segment keeping its sub-steps' inputs in HBM and sweeps
 * them backwards; it returns dJ_dlogeta (ne x (nphi,nlat), per mesh element
 * as pert_elem: grid values for genuinely 3-D elements, the element-uniform
 * share elsewhere) and dJ_dice ((ncalls+1) x (nphi,nlat): the ice before
 * the first call, then each call's ice). The model is left in the state after
 * the recorded calls. The derivative is the one of the recorded branch: skip
 * gate, coastline masks, sub-step sizes and the 3-D element set are held;
 * needs earth_response = "ve" (the rotation feedback is differentiated). The _device variants take
 * device pointers. */
int vh_forward_record(vh_model* h, int ncalls, const double* h_ice_hist, double dt_yr, int ckpt_every,
                      double* rsl_out);
int vh_forward_record_device(vh_model* h, int ncalls, const double* h_ice_hist_device, double dt_yr,
                             int ckpt_every, double* rsl_out_device);
int vh_backward(vh_model* h, const double* dJ_drsl, double* dJ_dlogeta, double* dJ_dice);
/* Restores the model state at the start of the last vh_forward_record (the
 * window's initial state), e.g. to run the same window again. */
int vh_forward_rewind(vh_model* h);
/* HBM the record pass may use to keep the sub-step inputs of the trailing
 * checkpoint segments, which the backward sweep then does not re-run
 * (default 96 GiB; at least the last segment is kept; 0 keeps only it). */
int vh_adjoint_store_limit(vh_model* h, double gib);
int vh_backward_device(vh_model* h, const double* dJ_drsl_device, double* dJ_dlogeta_device,
                       double* dJ_dice_device);

#ifdef __cplusplus
}
#endif
#endif /* VILMA_HIP_H */
