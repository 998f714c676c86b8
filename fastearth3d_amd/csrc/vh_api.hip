// vh_api.hip -- the C-ABI of include/vilma_hip.h and the host side of the
// time step: model set-up, the forward-Euler sub-stepping controller, the
// sea-level fixed point and the coupling step. All state lives in device
// memory; the host only makes the discrete decisions (SLE exits, coastline
// flips, the stability guard) from a few scalars read back per iteration.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>

#include "vh_internal.h"
#include "vh_rot.h"
#include "vh_sle.h"
#include "vh_adjoint.h"
#include "vh_remap.h"
#include "../../include/vilma_hip.h"

namespace {

thread_local std::string g_create_error;

template <class T>
T* dalloc(size_t n) {
    T* p = nullptr;
    HIPCHECK(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return p;
}
template <class T>
void upload(T* d, const T* h, size_t n, hipStream_t st) {
    HIPCHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, st));
}

// ---------------------------------------------------------- earth model
struct layer { double r_bot, r_top, rho, mu, eta; int rheology; };
enum { RHEOL_ELASTIC = 0, RHEOL_MAXWELL = 1, RHEOL_FLUID = 2 };

// incompressible PREM (Dziewonski & Anderson 1981, Table I), as
// src/vilma_earth_structure.f90:118-183 uses it
void prem_rho_mu(double r, double& rho, double& mu) {
    double x = r / 6371.0e3, vs;
    if (r <= 1221.5e3) { rho = 13088.5 - 8838.1 * x * x; mu = 7.0363e10; return; }
    if (r <= 3480.0e3) { rho = 12581.5 - 1263.8 * x - 3642.6 * x * x - 5528.1 * x * x * x; mu = 7.0363e10; return; }
    if (r <= 3630.0e3) { vs = 6925.4 + 1467.2 * x - 2083.4 * x * x + 978.3 * x * x * x; rho = 7956.5 - 6476.1 * x + 5528.3 * x * x - 3080.7 * x * x * x; }
    else if (r <= 5600.0e3) { vs = 11167.1 - 13781.8 * x + 17457.5 * x * x - 9277.7 * x * x * x; rho = 7956.5 - 6476.1 * x + 5528.3 * x * x - 3080.7 * x * x * x; }
    else if (r <= 5701.0e3) { vs = 22345.9 - 17247.3 * x - 2083.4 * x * x + 978.3 * x * x * x; rho = 7956.5 - 6476.1 * x + 5528.3 * x * x - 3080.7 * x * x * x; }
    else if (r <= 5771.0e3) { vs = 9983.9 - 4932.4 * x; rho = 5319.7 - 1483.6 * x; }
    else if (r <= 5971.0e3) { vs = 22351.2 - 18585.6 * x; rho = 11249.4 - 8029.8 * x; }
    else if (r <= 6151.0e3) { vs = 8949.6 - 4459.7 * x; rho = 7108.9 - 3804.5 * x; }
    else if (r <= 6346.6e3) { vs = 2151.9 + 2348.1 * x; rho = 2691.0 + 692.4 * x; }
    else if (r <= 6356.0e3) { vs = 3900.0; rho = 2900.0; }
    else { vs = 3200.0; rho = 2600.0; }
    mu = vs * vs * rho;
}

void prem_shell_avg(double rb, double rt, double& rho, double& mu) {
    const int n = 200;
    double h = (rt - rb) / n, nrho = 0, nmu = 0, den = 0;
    for (int i = 0; i <= n; i++) {
        double w = (i == 0 || i == n) ? 1.0 : ((i % 2) ? 4.0 : 2.0);
        double r = rb + i * h, ww = w * r * r, ri, mi;
        prem_rho_mu(r, ri, mi);
        nrho += ww * ri; nmu += ww * mi; den += ww;
    }
    rho = nrho / den; mu = nmu / den;
}

// build_earth (src/vilma_earth_structure.f90:62-116, :220-238)
std::vector<layer> build_earth(const vh_params& p, double& r_earth) {
    std::vector<layer> L;
    if (!strcmp(p.earth, "M3-L70-V01")) {
        const double km = 1e3;
        r_earth = 6371.0 * km;
        L = {{6301.0 * km, 6371.0 * km, 3037.0, 5.0605e10, DBL_MAX, RHEOL_ELASTIC},
             {5951.0 * km, 6301.0 * km, 3438.0, 7.0363e10, 1.0e21, RHEOL_MAXWELL},
             {5701.0 * km, 5951.0 * km, 3871.0, 1.0549e11, 1.0e21, RHEOL_MAXWELL},
             {3480.0 * km, 5701.0 * km, 4978.0, 2.2834e11, 2.0e21, RHEOL_MAXWELL},
             {0.0, 3480.0 * km, 10750.0, 0.0, 0.0, RHEOL_FLUID}};
        return L;
    }
    bool prem = !strcmp(p.earth, "PREM");
    if (!prem && strcmp(p.earth, "custom")) throw vh_error(std::string("build_earth: unknown earth model \"") + p.earth + "\"");
    if (p.n_layer < 1 || p.n_layer > VH_MAX_LAYER) throw vh_error("build_layered: n_layer out of range for a layered earth model");
    r_earth = p.r_earth;
    for (int k = 0; k < p.n_layer; k++) {
        double rho = p.rho[k], mu = p.mu[k];
        if (prem) { prem_shell_avg(p.r_bot[k], p.r_top[k], rho, mu); if (p.rheology[k] == RHEOL_FLUID) mu = 0.0; }
        L.push_back({p.r_bot[k], p.r_top[k], rho, mu, p.eta[k], p.rheology[k]});
    }
    return L;
}

double gravity_at(const std::vector<layer>& L, double r) {
    double m = 0.0;
    for (auto& ly : L) {
        double lo = ly.r_bot, hi = std::min(ly.r_top, r);
        if (hi > lo) m += (4.0 / 3.0) * VH_PI * ly.rho * (hi * hi * hi - lo * lo * lo);
    }
    return r > 0.0 ? VH_G * m / (r * r) : 0.0;
}

// radial_mesh_build (src/vilma_radial_fe.f90:153-224) + per-layer counts
void build_mesh(const std::vector<layer>& L, double r_earth, const int* nel, std::vector<double>& r,
                std::vector<int>& lay) {
    int n = (int)L.size();
    if (std::fabs(L[0].r_top - r_earth) > 1.0) throw vh_error("radial_mesh_build: r_top(1) must equal r_earth (layers are surface-first)");
    if (L[n - 1].r_bot != 0.0) throw vh_error("radial_mesh_build: the innermost layer must reach r = 0");
    for (int i = 0; i < n; i++) {
        if (L[i].r_top <= L[i].r_bot) throw vh_error("radial_mesh_build: every layer needs r_top > r_bot (surface-first ordering)");
        if (i < n - 1 && std::fabs(L[i].r_bot - L[i + 1].r_top) > 1.0) throw vh_error("radial_mesh_build: layers must be contiguous, r_bot(k) = r_top(k+1)");
    }
    auto count = [&](int i) {
        if (nel && nel[i] > 0) return nel[i];
        double r0 = L[i].r_bot, r1 = L[i].r_top, depth = r_earth - 0.5 * (r0 + r1);
        double dr = depth <= 70.0e3 ? 5.0e3 : (depth <= 670.0e3 ? 10.0e3 : 40.0e3);
        return std::max(1, (int)std::ceil((r1 - r0) / dr));
    };
    int ne = 0;
    for (int i = 0; i < n; i++) ne += count(i);
    r.assign(ne + 1, 0.0);
    lay.assign(ne, 0);
    int off = 0;
    for (int i = n - 1; i >= 0; i--) {
        int nl = count(i);
        double h = (L[i].r_top - L[i].r_bot) / (double)nl;
        for (int k = 1; k <= nl; k++) { r[off + k] = L[i].r_bot + (double)k * h; lay[off + k - 1] = i; }
        off += nl;
    }
    r[ne] = r_earth;
}

// Gauss-Legendre nodes (descending) and weights
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0); w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = std::cos(VH_PI * (i + 0.75) / (n + 0.5)), z1, pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int k = 1; k <= n; k++) { double p3 = p2; p2 = p1; p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k; }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z; z = z1 - p1 / pp;
            if (std::fabs(z - z1) <= 1e-16) break;
        }
        double p1 = 1.0, p2 = 0.0;
        for (int k = 1; k <= n; k++) { double p3 = p2; p2 = p1; p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k; }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        x[i] = z; x[n - 1 - i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp); w[n - 1 - i] = w[i];
    }
    if (n % 2) x[n / 2] = 0.0;
}

}  // namespace

// ------------------------------------------------------------------ model
struct adj_ctx;
void adj_free(adj_ctx* a);
bool adj_recording(const adj_ctx* a);
double2* adj_fd_pending(vh_model* h);   // store for the sub-step's tensor row spectra while recording

struct vh_model {
    vh_params par{};
    std::string err;
    vh_dims d{};
    vh_sht s{};
    vh_dev v{};
    hipStream_t st = nullptr;
    int kind = 2;            // 0 null, 1 elastic, 2 ve
    int deg1_cm = 1;
    double skip_tol = 1e-4, visc3d_tol = 1e-3;
    double a = 0, g = 0, dt = 0, time_resp = 0, time_yr = 0, bsl = 0, worst_mass_resid = 0;
    // host geometry
    std::vector<double> r, mu, rho, MkPerDt, Rk, colat, lon;
    std::vector<int> lay, emap, mnodes, mnode_of, kdeg;
    double rate_max = 0.0;   // max Maxwell rate incl. lateral field
    // device: radial
    double *d_r = 0, *d_mu = 0, *d_rho = 0, *d_Rk = 0, *d_MkPerDt = 0;
    int *d_emap = 0, *d_mnodes = 0, *d_mnode = 0, *d_kdeg = 0;
    unsigned char* d_is3d = 0;
    double *d_ab = 0, *d_dr = 0, *d_dc = 0, *d_ab1 = 0, *d_null1 = 0, *d_gu = 0, *d_gn = 0, *d_gv = 0;
    double *d_xUn = 0, *d_xVn = 0;
    int *d_ipiv = 0, *d_ipiv1 = 0, *d_fail = 0;
    double *d_tau[2] = {0, 0}, *d_mnorm[2] = {0, 0}, *d_maxn = 0;   // d_maxn[2]
    double* d_advmax = 0;   // the advance's per-block maxima
    int cur = 0;
    double maxn_cur = 0.0;
    double *d_dN = 0, *d_dsurf = 0;
    vh_drift_plan* dplan = nullptr;
    int *d_emap2 = 0, *d_mnodes2 = 0, *d_mnode2 = 0;   // rotation channel memory
    double* d_G2 = 0;
    // lateral viscosity: 3-D elements (batch b = 3*i3 + shape), tensor transforms
    int *d_e3em = 0, *d_e3of = 0;
    double *d_rate3 = 0, *d_n6 = 0;
    double2 *d_X = 0, *d_Fd = 0, *d_cdel = 0;
    int nbat = 0, tchunk = 0;
    bool tsh_ready = false;
    // device: SHT
    double *d_plm_lj = 0, *d_dplm_lj = 0, *d_plm_jl = 0, *d_dplm_jl = 0, *d_w = 0, *d_st = 0, *d_ct = 0;
    int* d_mstart = 0;
    double2 *d_tw = 0, *d_twp = 0;
    int2* d_ljobs = 0;
    // device: grid state
    double *d_topo0 = 0, *d_heq = 0, *d_C0ref = 0, *d_hice = 0, *d_ice1 = 0, *d_rsl = 0, *d_rsl_n = 0;
    double *d_C = 0, *d_Cn = 0, *d_dig = 0, *d_C0s = 0, *d_ice = 0, *d_dice = 0, *d_wcorr = 0;
    double *d_sraw = 0, *d_srot = 0, *d_zbed = 0, *d_gtmp = 0;
    double2 *d_F = 0, *d_FC = 0, *d_F2 = 0, *d_sigma = 0, *d_lm = 0, *d_uN = 0;
    double *d_part = 0, *d_sc = 0;
    unsigned* d_sbar = 0;   // synchronisation area of the fused sea-level solve
    double* h_pin = nullptr;
    // rotation
    bool rot_on = false;
    rot_args ra{};
    double *d_sc2 = 0, *d_cphi = 0, *d_sphi = 0;
    double dt_fe_max = DBL_MAX, rot_kTe = 0, rot_hTe = 0, rot_ks_fluid = 0;
    // SLE knobs + counters
    int n_outer = 3, n_inner = 20, fixed_ocean = 0, subgrid = 1;
    double tol = 1e-7, cfl = 1, dt_min = 0, dt_max = DBL_MAX, guard = 1e3;
    long n_accept = 0, n_reject = 0, n_solve = 0, n_inner_tot = 0, n_outer_tot = 0;
    long n_sle_steal = 0;    // sea-level work items run for a block that was not resident
    int sle_late_blocks = 0;     // diagnostic (VH_SLE_LATE_BLOCKS): solve blocks that start late
    int sle_blocks = 0;          // blocks of the fused sea-level solve (0: one per work item)
    // hipGraph replay of the forward-Euler sub-step (VH_T_GRAPH): one
    // executable graph per (memory buffer, interpolation fraction, step,
    // rotation field), captured on first use; dropped whenever a buffer or
    // launch shape it was captured with changes
    bool use_graph = false;
    struct sub_graph { int cur; double frac, dt; const double* srot; hipGraphExec_t exec; };
    std::vector<sub_graph> graphs;
    long n_graph_launch = 0, n_graph_capture = 0;
    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
    }
    unsigned sle_late_ticks = 0;
    // trapezoidal memory scheme (scheme "trap" with earth_response "ve"):
    // controller knobs and counters, the start-of-step load sigma_n, the
    // endpoint drift, the previous report drift, and the controller's saved
    // (A) and coarse (B) memory
    bool trap = false;
    int max_mem_iter = 20;
    double rtol = 1e-4, atol = 1e-3, safety = 0.9, grow_max = 5.0, shrink_min = 0.2, dt_try = 0.0;
    long n_floor = 0, n_couple_tot = 0;
    bool sig_primed = false, sig_primed_s = false;
    double time_resp_s = 0.0;
    double *d_sig_n = 0, *d_sig_n_s = 0, *d_edN = 0, *d_dua_prev = 0, *d_tau_s = 0, *d_tau_c = 0, *d_mnorm_s = 0;
    // kernel timing (HIP events on the model's stream), per kernel class
    bool prof = false;
    struct evpair { hipEvent_t a, b; int cls; };
    std::vector<evpair> evpool;
    size_t evused = 0;
    double prof_ms[VH_K_NCLASS] = {0};
    long prof_n[VH_K_NCLASS] = {0};
    void prof_begin(int cls) {
        if (!prof) return;
        if (evused == evpool.size()) {
            evpair e{};
            HIPCHECK(hipEventCreate(&e.a));
            HIPCHECK(hipEventCreate(&e.b));
            evpool.push_back(e);
        }
        evpool[evused].cls = cls;
        HIPCHECK(hipEventRecord(evpool[evused].a, st));
    }
    void prof_end() {
        if (!prof) return;
        HIPCHECK(hipEventRecord(evpool[evused].b, st));
        evused++;
    }
    void prof_collect() {   // after a stream synchronisation
        for (size_t i = 0; i < evused; i++) {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, evpool[i].a, evpool[i].b));
            prof_ms[evpool[i].cls] += ms;
            prof_n[evpool[i].cls]++;
        }
        evused = 0;
    }
    bool lat_visc = false;
    int ne3d = 0;
    // toroidal degree of freedom (l_toroidal): requested at create, carried
    // from the first genuinely 3-D element on (six memory channels; the
    // tensor batches then have nch = 6 channels). Toroidal memory taut
    // [em][12][k] (ping-pong like tau), nodal W drift dWt [nmn][2][k] and
    // surface dWa [nk] complex; the trapezoidal scheme adds the endpoint
    // drift, the previous surface drift, the saved and the coarse memory.
    bool tor_req = false, tor_on = false;
    int nch = 4;
    vh_tor_plan* tplan = nullptr;
    double *d_taut[2] = {0, 0}, *d_dWt = 0, *d_dWa = 0, *d_n4 = 0;
    double *d_edWt = 0, *d_dwa_prev = 0, *d_taut_s = 0, *d_taut_c = 0;
    // the lateral field of the last enable through the API, for the spin-up's
    // 1-D phase: 1 = per-element perturbation (lat_field, ne x grid; lat_mk the
    // 1-D rates it was applied to), 2 = nodal absolute log10 viscosity
    // (lat_field, nr x grid; lat_lid: lid depth, lid log10 max, cap)
    int lat_src = 0;
    std::vector<double> lat_field, lat_mk;
    double lat_lid[3] = {0.0, 0.0, 0.0};
    std::vector<int> e3mesh;   // mesh element of each 3-D element
    adj_ctx* adj = nullptr;    // reverse-mode record (vh_forward_record / vh_backward)

    size_t ng() const { return (size_t)d.nphi * d.nlat; }
    void sync() { HIPCHECK(hipStreamSynchronize(st)); }
    // read n device doubles (in-order after the stream's work)
    void read(const double* dsrc, int n) {
        HIPCHECK(hipMemcpyAsync(h_pin, dsrc, n * sizeof(double), hipMemcpyDeviceToHost, st));
        sync();
    }
    void refresh_dev() {
        v.r = d_r; v.mu = d_mu; v.MkPerDt = d_MkPerDt; v.emap = d_emap; v.mnodes = d_mnodes; v.kdeg = d_kdeg;
        v.mnode = d_mnode; v.is3d = d_is3d; v.ab = d_ab; v.ipiv = d_ipiv; v.dr = d_dr; v.dc = d_dc;
        v.ab1 = d_ab1; v.ipiv1 = d_ipiv1; v.null1 = d_null1; v.gu = d_gu; v.gn = d_gn; v.gv = d_gv;
        v.xUn = d_xUn; v.xVn = d_xVn;
    }
    ~vh_model();
};

vh_model::~vh_model() {
    drop_graphs();
    if (adj) adj_free(adj);
    if (dplan) vh_drift_plan_destroy(dplan);
    if (tplan) vh_tor_plan_destroy(tplan);
    for (void* p : {(void*)d_taut[0], (void*)d_taut[1], (void*)d_dWt, (void*)d_dWa, (void*)d_n4, (void*)d_edWt,
                    (void*)d_dwa_prev, (void*)d_taut_s, (void*)d_taut_c})
        if (p) (void)hipFree(p);
    void* ptrs[] = {d_r, d_mu, d_rho, d_Rk, d_MkPerDt, d_emap, d_mnodes, d_mnode, d_kdeg, d_is3d, d_ab, d_dr, d_dc,
                    d_ab1, d_null1, d_gu, d_gn, d_gv, d_xUn, d_xVn, d_ipiv, d_ipiv1, d_fail, d_tau[0], d_tau[1],
                    d_mnorm[0], d_mnorm[1], d_maxn, d_advmax, d_dN, d_dsurf, d_plm_lj, d_dplm_lj, d_plm_jl, d_dplm_jl, d_w,
                    d_st, d_ct, d_mstart, d_tw, d_twp, d_ljobs, d_topo0, d_heq, d_C0ref, d_hice, d_ice1, d_rsl, d_rsl_n, d_C,
                    d_Cn, d_dig, d_C0s, d_ice, d_dice, d_wcorr, d_sraw, d_srot, d_zbed, d_gtmp, d_F, d_FC, d_F2,
                    d_sigma, d_lm, d_uN, d_part, d_sc, d_sbar, d_sc2, d_cphi, d_sphi,
                    ra.rs.mem, ra.rs.MkPerDt, ra.rs.xU, ra.rs.xV, ra.rs.dN, ra.rs.sc, d_emap2, d_mnodes2,
                    d_mnode2, d_G2, d_e3em, d_e3of, d_rate3, d_n6, d_X, d_Fd, d_cdel,
                    d_sig_n, d_sig_n_s, d_edN, d_dua_prev, d_tau_s, d_tau_c, d_mnorm_s};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& e : evpool) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    if (h_pin) (void)hipHostFree(h_pin);
    if (st) (void)hipStreamDestroy(st);
}

namespace {

void setup(vh_model* h, const vh_params* p, int nphi, int nlat, const double* z_bed_eq, const double* h_ice_eq,
           const double* h_ice_init) {
    h->par = *p;
    if (strcmp(p->scheme, "fe") && strcmp(p->scheme, "trap"))
        throw vh_error(std::string("scheme \"") + p->scheme + "\" not supported (fe|trap)");
    if ((p->mmax >= 0 && p->mmax != p->lmax) || p->mres != 1) throw vh_error("mmax/mres other than lmax/1 not supported");
    if (p->lmax < 2) throw vh_error("lmax must be >= 2");
    int nl = p->nlat > 0 ? p->nlat : 2 * p->lmax + 2;
    int np = p->nphi > 0 ? p->nphi : 4 * p->lmax;
    nl = std::max(nl, 32);
    if (nl != nlat || np != nphi) throw vh_error("grid does not match the model Gauss grid");
    if (np <= 2 * p->lmax) throw vh_error("sht_grid_init: nphi must exceed 2*lmax");
    if (nl <= p->lmax) throw vh_error("sht_grid_init: nlat must exceed lmax");
    if (!strcmp(p->earth_response, "ve")) h->kind = 2;
    else if (!strcmp(p->earth_response, "elastic")) h->kind = 1;
    else if (!strcmp(p->earth_response, "null")) h->kind = 0;
    else throw vh_error("solid_earth_init: unknown earth_response (use ve|elastic|null)");
    h->deg1_cm = !strcmp(p->deg1_frame, "cm");
    h->skip_tol = p->skip_tol;
    h->visc3d_tol = p->visc3d_tol;
    h->tor_req = p->l_toroidal != 0;   // src/vilma_coupling.f90:282
    HIPCHECK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
    hipStream_t st = h->st;
    HIPCHECK(hipHostMalloc((void**)&h->h_pin, (2 * (size_t)std::max(nl, 32) + 64) * sizeof(double),
                           hipHostMallocDefault));

    // ---- earth, mesh, element tables (setup, host)
    double r_earth = 0;
    std::vector<layer> L = build_earth(*p, r_earth);
    build_mesh(L, r_earth, p->n_elem_layer, h->r, h->lay);
    const int nr = (int)h->r.size(), ne = nr - 1, Lmax = p->lmax;
    h->a = r_earth;
    h->g = gravity_at(L, r_earth);
    h->mu.resize(ne); h->rho.resize(ne); h->MkPerDt.resize(ne); h->Rk.assign(ne, 0.0);
    for (int e = 0; e < ne; e++) {
        const layer& ly = L[h->lay[e]];
        h->mu[e] = ly.mu; h->rho[e] = ly.rho;
        h->MkPerDt[e] = ly.rheology == RHEOL_MAXWELL ? h->mu[e] / ly.eta : 0.0;
    }
    for (int k = 1; k < ne; k++)
        h->Rk[k] = h->Rk[k - 1] + (h->rho[k - 1] - h->rho[k]) * h->r[k] * h->r[k] * h->r[k];
    h->mnode_of.assign(nr, -1);
    for (int e = 0; e < ne; e++)
        if (h->MkPerDt[e] > 0.0) { h->emap.push_back(e); h->mnode_of[e] = 1; h->mnode_of[e + 1] = 1; }
    for (int n = 0; n < nr; n++)
        if (h->mnode_of[n] > 0) { h->mnode_of[n] = (int)h->mnodes.size(); h->mnodes.push_back(n); }
    h->rate_max = 0.0;
    for (int e = 0; e < ne; e++) h->rate_max = std::max(h->rate_max, h->MkPerDt[e]);

    vh_dims& d = h->d;
    d.L = Lmax; d.nlm = (Lmax + 1) * (Lmax + 2) / 2; d.nr = nr; d.ne = ne; d.nd = 4 * nr - 1;
    d.nk = Lmax * (Lmax + 3) / 2; d.nem = (int)h->emap.size(); d.nmn = (int)h->mnodes.size();
    d.nphi = np; d.nlat = nl; d.nh = (nl + 1) / 2;
    d.ns1 = d.nd + 1; d.kl1 = d.nd; d.ldab1 = 3 * d.nd + 1;
    h->kdeg.resize(d.nk);
    for (int l = 1, k = 0; l <= Lmax; l++)
        for (int m = 0; m <= l; m++) h->kdeg[k++] = l;
    double dt0 = p->dt_init > 0.0 ? p->dt_init : 1000.0 * VH_SEC_PER_YEAR;
    h->dt = dt0;

    // ---- device radial data
    h->d_r = dalloc<double>(nr); upload(h->d_r, h->r.data(), nr, st);
    h->d_mu = dalloc<double>(ne); upload(h->d_mu, h->mu.data(), ne, st);
    h->d_rho = dalloc<double>(ne); upload(h->d_rho, h->rho.data(), ne, st);
    h->d_Rk = dalloc<double>(ne); upload(h->d_Rk, h->Rk.data(), ne, st);
    h->d_MkPerDt = dalloc<double>(ne); upload(h->d_MkPerDt, h->MkPerDt.data(), ne, st);
    h->d_emap = dalloc<int>(d.nem); upload(h->d_emap, h->emap.data(), d.nem, st);
    h->d_mnodes = dalloc<int>(d.nmn); upload(h->d_mnodes, h->mnodes.data(), d.nmn, st);
    h->d_mnode = dalloc<int>(nr); upload(h->d_mnode, h->mnode_of.data(), nr, st);
    h->d_kdeg = dalloc<int>(d.nk); upload(h->d_kdeg, h->kdeg.data(), d.nk, st);
    std::vector<unsigned char> zero3d(std::max(d.nem, 1), 0);
    h->d_is3d = dalloc<unsigned char>(d.nem); upload(h->d_is3d, zero3d.data(), d.nem, st);
    h->d_gu = dalloc<double>(Lmax + 1); h->d_gn = dalloc<double>(Lmax + 1); h->d_gv = dalloc<double>(Lmax + 1);
    h->d_xUn = dalloc<double>((size_t)Lmax * nr); h->d_xVn = dalloc<double>((size_t)Lmax * nr);
    h->d_fail = dalloc<int>(1);
    HIPCHECK(hipMemsetAsync(h->d_fail, 0, sizeof(int), st));
    {   // degree 0: no deformation, monopole geoid (src/vilma_response.f90:599-605)
        double g0[3] = {0.0, 4.0 * VH_PI * VH_G * h->a / h->g, 0.0};
        upload(h->d_gu, &g0[0], 1, st); upload(h->d_gn, &g0[1], 1, st); upload(h->d_gv, &g0[2], 1, st);
    }

    if (h->kind >= 1) {
        // ---- per-degree operators, factors and unit-load response (GPU)
        h->d_ab = dalloc<double>((size_t)(Lmax - 1) * VH_LDAB * d.nd);
        h->d_ipiv = dalloc<int>((size_t)(Lmax - 1) * d.nd);
        h->d_dr = dalloc<double>((size_t)Lmax * d.nd);
        h->d_dc = dalloc<double>((size_t)Lmax * d.nd);
        h->d_ab1 = dalloc<double>((size_t)d.ldab1 * d.ns1);
        h->d_ipiv1 = dalloc<int>(d.ns1);
        h->d_null1 = dalloc<double>(d.nd);
        double* work1 = dalloc<double>((size_t)15 * d.nd + 2);
        vh_launch_setup_band(st, d, 2, Lmax, h->d_r, h->d_mu, h->d_rho, h->d_Rk, h->a, h->g, 0, h->d_ab, h->d_ipiv,
                             h->d_dr, h->d_dc, nullptr, h->d_xUn, h->d_xVn, h->d_gu, h->d_gn, h->d_gv, h->d_fail);
        vh_launch_setup_deg1(st, d, h->d_r, h->d_mu, h->d_rho, h->d_Rk, h->a, h->g, h->deg1_cm, h->d_ab1, h->d_ipiv1,
                             h->d_dr, h->d_dc, h->d_null1, h->d_xUn, h->d_xVn, h->d_gu, h->d_gn, h->d_gv, work1,
                             h->d_fail);
        HIPCHECK(hipStreamSynchronize(st));
        (void)hipFree(work1);
        int fail = 0;
        HIPCHECK(hipMemcpy(&fail, h->d_fail, sizeof(int), hipMemcpyDeviceToHost));
        if (fail) throw vh_error("bordered_band_factor: band LU factorization failed");
    }
    if (h->kind == 2) {
        size_t nm = (size_t)24 * std::max(d.nem, 1) * d.nk;
        for (int b = 0; b < 2; b++) {
            h->d_tau[b] = dalloc<double>(nm);
            h->d_mnorm[b] = dalloc<double>(d.nk);
            vh_fill(st, h->d_tau[b], 0.0, nm);
            vh_fill(st, h->d_mnorm[b], 0.0, d.nk);
        }
        h->d_maxn = dalloc<double>(2);
        vh_fill(st, h->d_maxn, 0.0, 2);
        h->d_advmax = dalloc<double>(vh_advance_scratch(d));
        h->d_dN = dalloc<double>((size_t)4 * std::max(d.nmn, 1) * d.nk);
        h->d_dsurf = dalloc<double>((size_t)6 * d.nk);
        vh_fill(st, h->d_dN, 0.0, (size_t)4 * std::max(d.nmn, 1) * d.nk);
        vh_fill(st, h->d_dsurf, 0.0, (size_t)6 * d.nk);
    }
    h->refresh_dev();
    if (h->kind == 2) h->dplan = vh_drift_plan_create(st, d, h->v, h->deg1_cm, h->emap, h->mnode_of);

    // ---- SHT plan
    std::vector<double> ct, w;
    gauss_legendre(nl, ct, w);
    std::vector<double> sn(nl);
    h->colat.resize(nl);
    for (int j = 0; j < nl; j++) { h->colat[j] = std::acos(ct[j]); sn[j] = std::sqrt((1.0 - ct[j]) * (1.0 + ct[j])); }
    h->lon.resize(np);
    for (int i = 0; i < np; i++) h->lon[i] = 2.0 * VH_PI * i / np;
    std::vector<int> mstart(Lmax + 1);
    for (int m = 0, off = 0; m <= Lmax; m++) { mstart[m] = off; off += Lmax - m + 1; }
    std::vector<double2> tw(np);
    for (int t = 0; t < np; t++) tw[t] = make_double2(std::cos(-2.0 * VH_PI * t / np), std::sin(-2.0 * VH_PI * t / np));
    vh_sht& s = h->s;
    s.L = Lmax; s.nlat = nl; s.nphi = np; s.nh = d.nh; s.nlm = d.nlm;
    {
        int n = np, nf = 0;
        const int pref[7] = {8, 6, 4, 2, 3, 5, 7};   // fewest passes for 2^a 3^b
        for (int q = 0; q < 7 && n > 1; q++)
            while (n % pref[q] == 0 && nf < 16) { s.fac[nf++] = pref[q]; n /= pref[q]; }
        if (n != 1) throw vh_error("sht_grid_init: nphi must factor into 2, 3, 5, 7");
        s.nfac = nf;
    }
    h->d_w = dalloc<double>(nl); upload(h->d_w, w.data(), nl, st);
    h->d_ct = dalloc<double>(nl); upload(h->d_ct, ct.data(), nl, st);
    h->d_st = dalloc<double>(nl); upload(h->d_st, sn.data(), nl, st);
    h->d_mstart = dalloc<int>(Lmax + 1); upload(h->d_mstart, mstart.data(), Lmax + 1, st);
    h->d_tw = dalloc<double2>(np); upload(h->d_tw, tw.data(), np, st);
    {   // per-pass Stockham twiddles e^{-2 pi i k r tstep / N}, [r-1][k], k < Ns
        std::vector<double2> twp;
        for (int p = 0, Ns = 1; p < s.nfac; p++) {
            const int R = s.fac[p], tstep = np / (Ns * R);
            s.twoff[p] = (int)twp.size();
            s.nsm[p] = Ns > 1 ? (unsigned)((0x100000000ull + (unsigned long long)Ns - 1) / (unsigned long long)Ns) : 0u;
            if (Ns > 1)
                for (int r = 1; r < R; r++)
                    for (int k = 0; k < Ns; k++) twp.push_back(tw[(size_t)(k * r * tstep) % np]);
            Ns *= R;
        }
        s.ntw = (int)twp.size();
        // the fixed-length device transforms derive the same plan at compile time
        const fft_cplan c = fft_make_cplan(np);
        bool same = c.nfac == s.nfac;
        for (int p = 0; same && p < s.nfac; p++) same = c.fac[p] == s.fac[p] && c.twoff[p] == s.twoff[p];
        if (!same) throw vh_error("sht_grid_init: longitude FFT plan differs from the compile-time rule");
        h->d_twp = dalloc<double2>(std::max<size_t>(twp.size(), 1)); upload(h->d_twp, twp.data(), twp.size(), st);
        h->sync();   // twp goes out of scope
    }
    {   // Legendre analysis blocks: every order, in chunks of 64 degrees
        std::vector<int2> lj;
        for (int m = 0; m <= Lmax; m++)
            for (int i0 = 0; i0 < Lmax - m + 1; i0 += 64) lj.push_back(make_int2(m, i0));
        h->d_ljobs = dalloc<int2>(lj.size()); upload(h->d_ljobs, lj.data(), lj.size(), st);
        s.ljobs = h->d_ljobs; s.nljobs = (int)lj.size();
        h->sync();
    }
    size_t ntab = (size_t)d.nlm * d.nh;
    h->d_plm_lj = dalloc<double>(ntab); h->d_dplm_lj = dalloc<double>(ntab);
    h->d_plm_jl = dalloc<double>(ntab); h->d_dplm_jl = dalloc<double>(ntab);
    s.plm_lj = h->d_plm_lj; s.dplm_lj = h->d_dplm_lj; s.plm_jl = h->d_plm_jl; s.dplm_jl = h->d_dplm_jl;
    s.mstart = h->d_mstart; s.w = h->d_w; s.st = h->d_st; s.ct = h->d_ct; s.tw = h->d_tw; s.twp = h->d_twp;
    vh_setup_legendre_tables(st, s, h->d_ct, h->d_st, h->d_plm_lj, h->d_dplm_lj, h->d_plm_jl, h->d_dplm_jl);

    // ---- grid state (solid_earth_init, src/vilma_coupling.f90:196-215)
    size_t ng = h->ng();
    double** grids[] = {&h->d_topo0, &h->d_heq, &h->d_C0ref, &h->d_hice, &h->d_ice1, &h->d_rsl, &h->d_rsl_n,
                        &h->d_C, &h->d_Cn, &h->d_dig, &h->d_C0s, &h->d_ice, &h->d_dice, &h->d_wcorr, &h->d_sraw,
                        &h->d_srot, &h->d_zbed, &h->d_gtmp};
    for (double** g : grids) { *g = dalloc<double>(ng); vh_fill(st, *g, 0.0, ng); }
    upload(h->d_topo0, z_bed_eq, ng, st);
    upload(h->d_heq, h_ice_eq, ng, st);
    upload(h->d_hice, h_ice_init ? h_ice_init : h_ice_eq, ng, st);
    upload(h->d_zbed, z_bed_eq, ng, st);
    grid_ocean(st, (int)ng, h->d_topo0, h->d_heq, h->d_C0ref);
    HIPCHECK(hipMemcpyAsync(h->d_C, h->d_C0ref, ng * sizeof(double), hipMemcpyDeviceToDevice, st));
    h->d_F = dalloc<double2>((size_t)nl * (Lmax + 1));
    h->d_FC = dalloc<double2>((size_t)nl * (Lmax + 1));
    h->d_F2 = dalloc<double2>((size_t)4 * ((nl + 1) / 2) * (Lmax + 1));
    h->d_sigma = dalloc<double2>(d.nlm); vh_fill(st, (double*)h->d_sigma, 0.0, 2 * (size_t)d.nlm);
    h->d_lm = dalloc<double2>(d.nlm);
    h->d_uN = dalloc<double2>((size_t)2 * d.nlm);
    h->d_part = dalloc<double>((size_t)7 * nl + 64);
    h->d_sc = dalloc<double>(64); vh_fill(st, h->d_sc, 0.0, 64);
    h->d_sbar = dalloc<unsigned>(VH_SLE_BAR_WORDS);   // reset before every solve (sle_ice)
    if (const char* late = getenv("VH_SLE_LATE_BLOCKS")) {
        // diagnostic "count:microseconds": that many of the solve's blocks
        // start that much later than the others (at most 1 s), so the others
        // run their work items; the results do not change
        int nb = 0;
        double us = 0.0;
        if (sscanf(late, "%d:%lf", &nb, &us) == 2 && nb > 0 && us > 0.0) {
            h->sle_late_blocks = std::min(nb, sle_solve_grid(h->s));
            h->sle_late_ticks = (unsigned)(std::min(us, 1.0e6) * 100.0);
        }
    }

    // ---- SLE / stepper knobs (build_solver, src/vilma_coupling.f90:291-308)
    h->n_outer = p->sle_n_outer; h->n_inner = p->sle_n_inner; h->tol = p->sle_tol;
    h->fixed_ocean = p->sle_fixed_ocean; h->subgrid = p->sle_subgrid;
    h->cfl = p->cfl; h->dt_min = p->dt_min; h->dt_max = p->dt_max;
    // the scheme reaches the response only for ve (src/vilma_coupling.f90:269-281)
    h->trap = h->kind == 2 && !strcmp(p->scheme, "trap");
    h->max_mem_iter = std::max(1, p->sle_max_mem_iter);
    h->rtol = p->rtol; h->atol = p->atol; h->safety = p->safety; h->grow_max = p->grow_max;
    h->shrink_min = p->shrink_min; h->dt_try = p->dt_init;

    // ---- rotation (rotation_init, src/vilma_rotation.f90:114-152)
    if (p->rotation && h->kind >= 1) {
        h->rot_on = true;
        rot_state& rs = h->ra.rs;
        rs.mem = dalloc<double>((size_t)2 * ne * 24); vh_fill(st, rs.mem, 0.0, (size_t)2 * ne * 24);
        rs.MkPerDt = dalloc<double>((size_t)2 * ne);
        rs.xU = dalloc<double>((size_t)2 * nr); rs.xV = dalloc<double>((size_t)2 * nr);
        rs.dN = dalloc<double>((size_t)8 * nr); vh_fill(st, rs.dN, 0.0, (size_t)8 * nr);
        rs.sc = dalloc<double>(RS_N); vh_fill(st, rs.sc, 0.0, RS_N);
        std::vector<double> rk2(2 * ne);
        double mx = 0.0;
        for (int e = 0; e < ne; e++) {
            double eta = L[h->lay[e]].eta;
            double v = eta > 0.0 ? h->mu[e] / eta : 0.0;
            rk2[e] = v; rk2[ne + e] = v;
            mx = std::max(mx, v);
        }
        if (mx > 0.0) h->dt_fe_max = 1.0 / mx;
        upload(rs.MkPerDt, rk2.data(), 2 * ne, st);
        // unit responses: load and tidal on the degree-2 operator, tidal on the
        // Maxwell-fluidised earth (fluid_tidal_k, :375-404)
        std::vector<double> mu_f(ne);
        for (int e = 0; e < ne; e++) mu_f[e] = L[h->lay[e]].rheology == RHEOL_MAXWELL ? 0.0 : h->mu[e];
        double* d_muf = dalloc<double>(ne); upload(d_muf, mu_f.data(), ne, st);
        double* sab = dalloc<double>((size_t)VH_LDAB * d.nd * 3);
        int* sip = dalloc<int>((size_t)d.nd * 3);
        double* sdr = dalloc<double>((size_t)3 * d.nd);   // index l-1 = 1 used
        double* sdc = dalloc<double>((size_t)3 * d.nd);
        double* xf = dalloc<double>((size_t)3 * d.nd);
        vh_launch_setup_band(st, d, 2, 2, h->d_r, h->d_mu, h->d_rho, h->d_Rk, h->a, h->g, 0, sab, sip, sdr, sdc,
                             xf, nullptr, nullptr, nullptr, nullptr, nullptr, h->d_fail);
        vh_launch_setup_band(st, d, 2, 2, h->d_r, h->d_mu, h->d_rho, h->d_Rk, h->a, h->g, 1, sab, sip, sdr, sdc,
                             xf + d.nd, nullptr, nullptr, nullptr, nullptr, nullptr, h->d_fail);
        vh_launch_setup_band(st, d, 2, 2, h->d_r, d_muf, h->d_rho, h->d_Rk, h->a, h->g, 1, sab, sip, sdr, sdc,
                             xf + 2 * d.nd, nullptr, nullptr, nullptr, nullptr, nullptr, h->d_fail);
        std::vector<double> x3((size_t)3 * d.nd);
        HIPCHECK(hipMemcpyAsync(x3.data(), xf, x3.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        h->sync();
        (void)hipFree(d_muf); (void)hipFree(sab); (void)hipFree(sip); (void)hipFree(sdr); (void)hipFree(sdc); (void)hipFree(xf);
        int fail = 0;
        HIPCHECK(hipMemcpy(&fail, h->d_fail, sizeof(int), hipMemcpyDeviceToHost));
        if (fail) throw vh_error("bordered_band_factor: band LU factorization failed");
        const double* xl = x3.data();
        const double* xt = x3.data() + d.nd;
        const double* xfl = x3.data() + 2 * d.nd;
        int sidx = nr - 1;
        std::vector<double> xu(2 * nr), xv(2 * nr);
        for (int n = 0; n < nr; n++) {
            xu[n] = xl[4 * n]; xv[n] = xl[4 * n + 1];
            xu[nr + n] = xt[4 * n]; xv[nr + n] = xt[4 * n + 1];
        }
        upload(rs.xU, xu.data(), 2 * nr, st);
        upload(rs.xV, xv.data(), 2 * nr, st);
        double Fe_tidal = xt[4 * sidx + 2];
        h->rot_kTe = -Fe_tidal - 1.0;
        h->rot_hTe = h->g * xt[4 * sidx];
        h->rot_ks_fluid = -xfl[4 * sidx + 2] / 1.0 - 1.0;
        double sc[RS_N] = {0};
        sc[RS_FE_LOAD] = xl[4 * sidx + 2];
        sc[RS_UE_TIDAL] = xt[4 * sidx];
        sc[RS_KTE] = h->rot_kTe;
        sc[RS_KS] = h->rot_ks_fluid;
        sc[RS_A] = h->a; sc[RS_G] = h->g; sc[RS_CMA] = 2.63e35; sc[RS_OMEGA] = 7.292115e-5;
        upload(rs.sc, sc, RS_N, st);
        h->ra.nr = nr; h->ra.nd = d.nd; h->ra.r = h->d_r; h->ra.mu = h->d_mu;
        {   // channel-memory elements and nodes, restricted degree-2 operator
            std::vector<int> em2, mn2, mnode2(nr, -1);
            for (int e = 0; e < ne; e++)
                if (rk2[e] > 0.0 || rk2[ne + e] > 0.0) { em2.push_back(e); mnode2[e] = 1; mnode2[e + 1] = 1; }
            for (int n = 0; n < nr; n++)
                if (mnode2[n] > 0) { mnode2[n] = (int)mn2.size(); mn2.push_back(n); }
            h->ra.nem2 = (int)em2.size(); h->ra.nmn2 = (int)mn2.size();
            h->d_emap2 = dalloc<int>(std::max<size_t>(em2.size(), 1)); upload(h->d_emap2, em2.data(), em2.size(), st);
            h->d_mnodes2 = dalloc<int>(std::max<size_t>(mn2.size(), 1)); upload(h->d_mnodes2, mn2.data(), mn2.size(), st);
            h->d_mnode2 = dalloc<int>(nr); upload(h->d_mnode2, mnode2.data(), nr, st);
            const size_t ng2 = (size_t)(2 * mn2.size() + 3) * std::max<size_t>(2 * mn2.size(), 1);
            h->d_G2 = dalloc<double>(ng2);
            vh_fill(st, h->d_G2, 0.0, ng2);
            vh_setup_restricted_ops(st, d, h->v, 2, 2, h->ra.nmn2, h->d_mnodes2, h->d_G2);
            h->ra.emap2 = h->d_emap2; h->ra.mnodes2 = h->d_mnodes2; h->ra.mnode2 = h->d_mnode2; h->ra.G2 = h->d_G2;
            h->sync();   // the host lists above go out of scope
        }
        std::vector<double> sc2(nl), cph(np), sph(np);
        for (int j = 0; j < nl; j++) sc2[j] = std::sin(h->colat[j]) * std::cos(h->colat[j]);
        for (int i = 0; i < np; i++) { cph[i] = std::cos(h->lon[i]); sph[i] = std::sin(h->lon[i]); }
        h->d_sc2 = dalloc<double>(nl); upload(h->d_sc2, sc2.data(), nl, st);
        h->d_cphi = dalloc<double>(np); upload(h->d_cphi, cph.data(), np, st);
        h->d_sphi = dalloc<double>(np); upload(h->d_sphi, sph.data(), np, st);
    }
    h->sync();
}

// The toroidal channels of one advance (none unless they are carried): memory
// cur -> nxt, nodal W drift dW (and, trapezoidal rule, the start drift dW0)
vh_adv_tor adv_tor(const vh_model* h, int cur, int nxt, const double* dW, const double* dW0) {
    vh_adv_tor t;
    if (!h->tor_on) return t;
    t.plan = h->tplan; t.taut_in = h->d_taut[cur]; t.taut_out = h->d_taut[nxt]; t.dW = dW; t.dW0 = dW0;
    return t;
}

// response_begin_step's toroidal branch (src/vilma_response.f90:694-809): the
// W drift of the memory in buffer b into dW (nodal) and d_dWa (surface)
void tor_drift(vh_model* h, int b, double* dW) {
    if (!h->tor_on) return;
    vh_tor_drift(h->st, h->tplan, h->d, h->v, h->d_taut[b], h->d_mnorm[b], h->d_maxn + b, h->skip_tol, dW, h->d_dWa);
}

// vilma_advance (src/vilma_response.f90:944-992): the 3-D elements first
// (advance_memory_3d, :1171-1225, increments into h->d_cdel), then the one
// streaming pass that applies them and advances every 1-D element.
void advance_memory(vh_model* h, const double* sigma, int nxt, double* maxn_copy = nullptr) {
    hipStream_t st = h->st;
    const vh_dims& d = h->d;
    const int nch = h->nch;
    if (h->lat_visc && h->ne3d > 0) {
        const size_t ng = (size_t)d.nphi * d.nlat;
        h->prof_begin(VH_K_T3_GATHER);
        vh_t3_gather(st, d, h->v, h->d_e3em, h->ne3d, h->d_tau[h->cur], sigma, h->d_dN, h->d_X, nch,
                     h->d_taut[h->cur], h->d_dWt);
        h->prof_end();
        for (int b0 = 0; b0 < h->nbat; b0 += h->tchunk) {
            const int nb = std::min(h->tchunk, h->nbat - b0);
            // while recording for the reverse sweep, the synthesis writes into
            // the sub-step's store (its rate gradient reads the forward rows)
            // and the row pass reads it from there
            double2* syn = h->d_Fd;
            if (h->adj && adj_recording(h->adj))
                syn = adj_fd_pending(h) + (size_t)b0 * 6 * (d.L + 1) * d.nlat;
            h->prof_begin(VH_K_TSH_SYNTH);
            vh_tsh_synth(st, h->s, h->d_X + (size_t)b0 * nch * d.nlm, nb, syn, nch);
            h->prof_end();
            h->prof_begin(VH_K_T3_ROWS);
            vh_t3_rows(st, h->s, h->d_Fd, nb, 0, h->d_rate3 + (size_t)(b0 / 3) * ng, 3, h->dt, nullptr, syn);
            h->prof_end();
            h->prof_begin(VH_K_TSH_ANALYSIS);
            vh_tsh_analysis(st, h->s, h->d_Fd, nb, h->d_n6, h->d_cdel + (size_t)b0 * nch * d.nlm, nch, h->d_n4);
            h->prof_end();
        }
    }
    h->prof_begin(VH_K_ADVANCE);
    vh_launch_advance(st, d, h->v, h->dt, h->d_tau[h->cur], h->d_tau[nxt], sigma, h->d_dN, h->d_mnorm[nxt],
                      h->d_maxn + nxt, h->d_e3of, h->d_cdel, h->d_advmax, maxn_copy, nullptr, nullptr, nch,
                      adv_tor(h, h->cur, nxt, h->d_dWt, nullptr));
    h->prof_end();
}

// ---------------------------------------------------------------- SLE solve
struct sle_res { double mass_resid = 0, err_inf = 0, n_last = 0, c_int = 0; };
void adj_stash_push(vh_model* h, double dt, double frac, const sle_res& r);

// sle_solve (src/vilma_sle.f90:124-405), forward-Euler response branch.
// The advance writes tau_{n+1} into the spare buffer; the caller commits it.
// The stream work of one solve: ice, drift, the fused solve and (unless
// report_only) the memory advance. prepare: set the kernels' launch
// attributes only, enqueue nothing.
void sle_enqueue(vh_model* h, const double* ice0, const double* ice1, double frac, int interp, bool report_only,
                 const double* srot, bool prepare = false) {
    hipStream_t st = h->st;
    const vh_sht& s = h->s;
    const int n = (int)h->ng();
    // sc[16..28): the solve's results (zeroed with wcorr by the ice kernel),
    // the new memory norm at [11], all read back once at the end of the solve
    double* r = h->d_sc + 16;
    if (!prepare)
        sle_ice(st, n, ice0, ice1, frac, interp, h->d_heq, h->d_topo0, h->d_ice, h->d_dice, h->d_C0s, h->d_wcorr, r,
                12, h->d_sbar, sle_solve_grid(s), h->n_outer);
    if (h->kind == 2 && !prepare) {
        h->prof_begin(VH_K_DRIFT);
        vh_drift_run(st, h->dplan, h->d, h->v, h->d_tau[h->cur], h->d_mnorm[h->cur], h->d_maxn + h->cur, h->skip_tol,
                     h->deg1_cm, h->d_dN, h->d_dsurf);
        tor_drift(h, h->cur, h->d_dWt);
        h->prof_end();
    }
    // the coastline passes and inner fixed point in one launch
    sle_solve_args ia{};
    ia.s = s; ia.topo0 = h->d_topo0; ia.ice = h->d_ice; ia.dice = h->d_dice; ia.C0 = h->d_C0s; ia.srot = srot;
    ia.gu = h->d_gu; ia.gn = h->d_gn; ia.dsurf = h->d_dsurf;
    ia.C = h->d_C; ia.Cn = h->d_Cn; ia.dig = h->d_dig; ia.wcorr = h->d_wcorr; ia.sraw = h->d_sraw;
    ia.nk = h->d.nk; ia.kind = h->kind; ia.subgrid = h->subgrid; ia.n_inner = h->n_inner; ia.n_outer = h->n_outer;
    ia.fixed_ocean = h->fixed_ocean; ia.g = h->g; ia.tol = h->tol; ia.rsl = h->d_rsl;
    ia.part = h->d_part; ia.cs = h->d_part + 3 * (size_t)s.nlat; ia.part3 = h->d_part + 4 * (size_t)s.nlat;
    ia.res = r; ia.F = h->d_F; ia.FC = h->d_FC; ia.F2 = h->d_F2; ia.sigma = h->d_sigma; ia.bar = h->d_sbar;
    ia.late_blocks = h->sle_late_blocks; ia.late_ticks = h->sle_late_ticks; ia.blocks = h->sle_blocks;
    if (prepare) {
        sle_solve_launch(st, ia, false);
        return;
    }
    h->prof_begin(VH_K_SLE);
    sle_solve_launch(st, ia);
    h->prof_end();
    // sigma is the closing load transform; then the memory advance
    int nxt = h->cur ^ 1;
    if (!report_only && h->kind == 2) advance_memory(h, (const double*)h->d_sigma, nxt, r + 11);
}

// sle_enqueue as one graph launch: the executable graph of this (buffer,
// fraction, step, rotation field), captured from sle_enqueue on first use
void sle_enqueue_graph(vh_model* h, double frac, const double* srot) {
    hipGraphExec_t ex = nullptr;
    for (const auto& g : h->graphs)
        if (g.cur == h->cur && g.frac == frac && g.dt == h->dt && g.srot == srot) { ex = g.exec; break; }
    if (!ex) {
        if (h->graphs.size() >= 256) h->drop_graphs();
        sle_enqueue(h, h->d_hice, h->d_ice1, frac, 1, false, srot, true);
        HIPCHECK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            sle_enqueue(h, h->d_hice, h->d_ice1, frac, 1, false, srot);
        } catch (...) {
            (void)hipStreamEndCapture(h->st, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCHECK(hipStreamEndCapture(h->st, &g));
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(e);
        h->graphs.push_back({h->cur, frac, h->dt, srot, ex});
        h->n_graph_capture++;
    }
    HIPCHECK(hipGraphLaunch(ex, h->st));
    h->n_graph_launch++;
}

sle_res sle_solve(vh_model* h, const double* ice0, const double* ice1, double frac, int interp, bool report_only,
                  const double* srot) {
    sle_res res;
    double* r = h->d_sc + 16;
    // graph replay for the stepper's sub-steps (not while timing kernels with
    // events or recording for the reverse sweep, whose store moves per sub-step)
    const bool graph = h->use_graph && !report_only && interp == 1 && ice0 == h->d_hice && ice1 == h->d_ice1 &&
                       !h->prof && !(h->adj && adj_recording(h->adj));
    if (graph) sle_enqueue_graph(h, frac, srot);
    else sle_enqueue(h, ice0, ice1, frac, interp, report_only, srot);
    hipStream_t st = h->st;
    if (!report_only) { h->time_resp += h->dt; h->n_couple_tot++; }   // forward Euler: one pass
    // r[10] = int C rsl (from the solve) for the mass diagnostic, r[11] the new memory norm
    if (report_only && h->kind == 2)
        HIPCHECK(hipMemcpyAsync(r + 11, h->d_maxn + h->cur, sizeof(double), hipMemcpyDeviceToDevice, st));
    h->read(r, 12);
    const double* hr = h->h_pin;
    if (hr[2] != 0.0) throw vh_error("sle_solve: a wait of the fused sea-level solve exceeded its bound");
    h->n_sle_steal += (long)hr[8];
    h->n_inner_tot += (long)hr[4];
    h->n_outer_tot += (long)hr[5];
    const double ice_int = -(VH_RHO_ICE / VH_RHO_WATER) * hr[6];
    const double zeta_int = h->subgrid ? hr[7] : 0.0;
    const double Cs = hr[10] - zeta_int;
    res.mass_resid = std::fabs(ice_int) > 0.0 ? std::fabs(Cs - ice_int) / std::fabs(ice_int) : std::fabs(Cs);
    res.err_inf = hr[11];
    res.n_last = hr[0];
    res.c_int = hr[3];
    return res;
}

// ------------------------------------------------------------------ TRAP
// The trapezoidal (Crank-Nicolson) memory scheme: the sea-level <->
// memory co-convergence of one solve (src/vilma_sle.f90:218-366,
// ve_response_prepare/advance_endpoint src/vilma_response.f90:1598-1646)
// and the step-doubling controller (src/vilma_timestep.f90:229-307).
// tau_n stays in d_tau[cur] for the whole solve; every pass writes its
// tau_{n+1} estimate into d_tau[cur ^ 1] and the endpoint drift of that
// estimate into d_edN / d_dsurf; the solve commits the last estimate.
// the toroidal part of the trapezoidal buffers (ensure_commit_scratch_tor,
// src/vilma_response.f90:1111-1118, and the widened save / coarse buffers)
void trap_alloc_tor(vh_model* h) {
    if (!h->tor_on || h->d_taut_s) return;
    const vh_dims& d = h->d;
    const size_t nt = (size_t)12 * std::max(d.nem, 1) * d.nk, ndw = (size_t)2 * std::max(d.nmn, 1) * d.nk;
    h->d_taut_s = dalloc<double>(nt); h->d_taut_c = dalloc<double>(nt);
    h->d_edWt = dalloc<double>(ndw);
    h->d_dwa_prev = dalloc<double>((size_t)2 * d.nk);
    vh_fill(h->st, h->d_taut_s, 0.0, nt); vh_fill(h->st, h->d_taut_c, 0.0, nt);
    vh_fill(h->st, h->d_edWt, 0.0, ndw); vh_fill(h->st, h->d_dwa_prev, 0.0, (size_t)2 * d.nk);
}

void trap_alloc(vh_model* h) {
    trap_alloc_tor(h);
    if (h->d_tau_s) return;
    const vh_dims& d = h->d;
    hipStream_t st = h->st;
    const size_t nm = (size_t)24 * std::max(d.nem, 1) * d.nk, ndn = (size_t)4 * std::max(d.nmn, 1) * d.nk;
    h->d_tau_s = dalloc<double>(nm); h->d_tau_c = dalloc<double>(nm);
    h->d_mnorm_s = dalloc<double>((size_t)d.nk + 1);
    h->d_edN = dalloc<double>(ndn);
    h->d_dua_prev = dalloc<double>((size_t)2 * d.nk);
    h->d_sig_n = dalloc<double>((size_t)2 * d.nlm); h->d_sig_n_s = dalloc<double>((size_t)2 * d.nlm);
    vh_fill(st, h->d_sig_n, 0.0, (size_t)2 * d.nlm);
}

void d2d(double* dst, const double* src, size_t n, hipStream_t st);

// trapezoid_advance_all (src/vilma_response.f90:1524-1596): 3-D elements by
// the tensor transforms of tau_n + mu (eps_n + eps_{n+1}) weighted by
// -M/(1 + M/2), then every Maxwell element's tau_{n+1} in one pass
void trap_advance(vh_model* h, const double* sigma_n) {
    hipStream_t st = h->st;
    const vh_dims& d = h->d;
    const int cur = h->cur, nxt = cur ^ 1, nch = h->nch;
    const double* sig1 = (const double*)h->d_sigma;
    if (h->lat_visc && h->ne3d > 0) {
        const size_t ng = (size_t)d.nphi * d.nlat;
        h->prof_begin(VH_K_T3_GATHER);
        vh_t3_gather_trap(st, d, h->v, h->d_e3em, h->ne3d, h->d_tau[cur], sigma_n, h->d_dN, sig1, h->d_edN, h->d_X,
                          nch, h->d_taut[cur], h->d_dWt, h->d_edWt);
        h->prof_end();
        for (int b0 = 0; b0 < h->nbat; b0 += h->tchunk) {
            const int nb = std::min(h->tchunk, h->nbat - b0);
            h->prof_begin(VH_K_TSH_SYNTH);
            vh_tsh_synth(st, h->s, h->d_X + (size_t)b0 * nch * d.nlm, nb, h->d_Fd, nch);
            h->prof_end();
            h->prof_begin(VH_K_T3_ROWS);
            vh_t3_rows(st, h->s, h->d_Fd, nb, 3, h->d_rate3 + (size_t)(b0 / 3) * ng, 3, h->dt, nullptr);
            h->prof_end();
            h->prof_begin(VH_K_TSH_ANALYSIS);
            vh_tsh_analysis(st, h->s, h->d_Fd, nb, h->d_n6, h->d_cdel + (size_t)b0 * nch * d.nlm, nch, h->d_n4);
            h->prof_end();
        }
    }
    h->prof_begin(VH_K_ADVANCE);
    vh_launch_advance(st, d, h->v, h->dt, h->d_tau[cur], h->d_tau[nxt], sig1, h->d_edN, h->d_mnorm[nxt],
                      h->d_maxn + nxt, h->d_e3of, h->d_cdel, h->d_advmax, nullptr, sigma_n, h->d_dN, nch,
                      adv_tor(h, cur, nxt, h->d_edWt, h->d_dWt));
    h->prof_end();
}

// sle_solve with the co-convergence passes: each pass one fused sea-level
// solve against the current report drift, the trapezoid advance with its
// closing load, the endpoint (and report) drift of the new estimate and the
// drift change; done once the drift settles (pass >= 2) or after
// max_mem_iter passes. Commits tau_{n+1}, sigma_n and the response time.
sle_res sle_solve_trap(vh_model* h, const double* ice0, const double* ice1, double frac, const double* srot) {
    hipStream_t st = h->st;
    const vh_sht& s = h->s;
    const vh_dims& d = h->d;
    const int n = (int)h->ng();
    double* r = h->d_sc + 16;
    const int cur = h->cur, nxt = cur ^ 1;
    sle_res res;
    sle_ice(st, n, ice0, ice1, frac, 1, h->d_heq, h->d_topo0, h->d_ice, h->d_dice, h->d_C0s, h->d_wcorr, r, 12,
            h->d_sbar, sle_solve_grid(s), h->n_outer);
    // response_begin_step: the drift of tau_n; prepare_endpoint: the endpoint
    // drift starts as it, the report drift is kept for the change test
    h->prof_begin(VH_K_DRIFT);
    vh_drift_run(st, h->dplan, d, h->v, h->d_tau[cur], h->d_mnorm[cur], h->d_maxn + cur, h->skip_tol, h->deg1_cm,
                 h->d_dN, h->d_dsurf);
    tor_drift(h, cur, h->d_dWt);
    h->prof_end();
    d2d(h->d_edN, h->d_dN, (size_t)4 * std::max(d.nmn, 1) * d.nk, st);
    d2d(h->d_dua_prev, h->d_dsurf, (size_t)2 * d.nk, st);
    if (h->tor_on) {
        d2d(h->d_edWt, h->d_dWt, (size_t)2 * std::max(d.nmn, 1) * d.nk, st);
        d2d(h->d_dwa_prev, h->d_dWa, (size_t)2 * d.nk, st);
    }
    // sigma_n: the committed start-of-step load, or (not yet primed) this
    // pass's own load (src/vilma_response.f90:1553-1558)
    const double* sigma_n = h->sig_primed ? h->d_sig_n : (const double*)h->d_sigma;
    sle_solve_args ia{};
    ia.s = s; ia.topo0 = h->d_topo0; ia.ice = h->d_ice; ia.dice = h->d_dice; ia.C0 = h->d_C0s; ia.srot = srot;
    ia.gu = h->d_gu; ia.gn = h->d_gn; ia.dsurf = h->d_dsurf;
    ia.C = h->d_C; ia.Cn = h->d_Cn; ia.dig = h->d_dig; ia.wcorr = h->d_wcorr; ia.sraw = h->d_sraw;
    ia.nk = d.nk; ia.kind = h->kind; ia.subgrid = h->subgrid; ia.n_inner = h->n_inner; ia.n_outer = h->n_outer;
    ia.fixed_ocean = h->fixed_ocean; ia.g = h->g; ia.tol = h->tol; ia.rsl = h->d_rsl;
    ia.part = h->d_part; ia.cs = h->d_part + 3 * (size_t)s.nlat; ia.part3 = h->d_part + 4 * (size_t)s.nlat;
    ia.res = r; ia.F = h->d_F; ia.FC = h->d_FC; ia.F2 = h->d_F2; ia.sigma = h->d_sigma; ia.bar = h->d_sbar;
    ia.late_blocks = h->sle_late_blocks; ia.late_ticks = h->sle_late_ticks; ia.blocks = h->sle_blocks;
    int pass = 0;
    const double* hr = h->h_pin;
    for (int im = 1; im <= h->max_mem_iter; im++) {
        // the same ice fields again and a reset of the result slots and the
        // synchronisation area for the next launch
        if (im > 1)
            sle_ice(st, n, ice0, ice1, frac, 1, h->d_heq, h->d_topo0, h->d_ice, h->d_dice, h->d_C0s, h->d_wcorr, r,
                    12, h->d_sbar, sle_solve_grid(s), h->n_outer);
        h->prof_begin(VH_K_SLE);
        sle_solve_launch(st, ia);
        h->prof_end();
        trap_advance(h, sigma_n);
        h->prof_begin(VH_K_DRIFT);
        vh_drift_run(st, h->dplan, d, h->v, h->d_tau[nxt], h->d_mnorm[nxt], h->d_maxn + nxt, h->skip_tol,
                     h->deg1_cm, h->d_edN, h->d_dsurf);
        tor_drift(h, nxt, h->d_edWt);
        h->prof_end();
        vh_drift_change(st, d.nk, h->d_dsurf, h->d_dua_prev, r + 12, h->tor_on ? h->d_dWa : nullptr,
                        h->tor_on ? h->d_dwa_prev : nullptr);
        h->read(r, 14);
        if (hr[2] != 0.0) throw vh_error("sle_solve: a wait of the fused sea-level solve exceeded its bound");
        h->n_sle_steal += (long)hr[8];
        h->n_inner_tot += (long)hr[4];
        h->n_outer_tot += (long)hr[5];
        pass++;
        // couple_tol = 1e-6 (src/vilma_response.f90:177)
        if (pass >= 2 && hr[12] <= 1.0e-6 * std::max(hr[13], DBL_MIN)) break;
    }
    h->n_couple_tot += pass;
    // ve_response_finalize_step: sigma_{n+1} becomes sigma_n; time advances
    d2d(h->d_sig_n, (const double*)h->d_sigma, (size_t)2 * d.nlm, st);
    h->sig_primed = true;
    h->time_resp += h->dt;
    h->cur = nxt;
    const double ice_int = -(VH_RHO_ICE / VH_RHO_WATER) * hr[6];
    const double zeta_int = h->subgrid ? hr[7] : 0.0;
    const double Cs = hr[10] - zeta_int;
    res.mass_resid = std::fabs(ice_int) > 0.0 ? std::fabs(Cs - ice_int) / std::fabs(ice_int) : std::fabs(Cs);
    res.n_last = hr[0];
    res.c_int = hr[3];
    return res;
}

// response_save_state / response_restore_state (buffer A: tau_n, its slot
// maxima, sigma_n, the response time)
void trap_save(vh_model* h) {
    const vh_dims& d = h->d;
    const size_t nm = (size_t)24 * std::max(d.nem, 1) * d.nk;
    d2d(h->d_tau_s, h->d_tau[h->cur], nm, h->st);
    if (h->tor_on) d2d(h->d_taut_s, h->d_taut[h->cur], nm / 2, h->st);
    d2d(h->d_mnorm_s, h->d_mnorm[h->cur], d.nk, h->st);
    d2d(h->d_mnorm_s + d.nk, h->d_maxn + h->cur, 1, h->st);
    d2d(h->d_sig_n_s, h->d_sig_n, (size_t)2 * d.nlm, h->st);
    h->sig_primed_s = h->sig_primed;
    h->time_resp_s = h->time_resp;
}
void trap_restore(vh_model* h) {
    const vh_dims& d = h->d;
    const size_t nm = (size_t)24 * std::max(d.nem, 1) * d.nk;
    d2d(h->d_tau[h->cur], h->d_tau_s, nm, h->st);
    if (h->tor_on) d2d(h->d_taut[h->cur], h->d_taut_s, nm / 2, h->st);
    d2d(h->d_mnorm[h->cur], h->d_mnorm_s, d.nk, h->st);
    d2d(h->d_maxn + h->cur, h->d_mnorm_s + d.nk, 1, h->st);
    d2d(h->d_sig_n, h->d_sig_n_s, (size_t)2 * d.nlm, h->st);
    h->sig_primed = h->sig_primed_s;
    h->time_resp = h->time_resp_s;
}

// stepper_advance, implicit branch (src/vilma_timestep.f90:229-307)
void stepper_trap(vh_model* h, double t0, double t1, const double* srot) {
    const double span = t1 - t0;
    const int n = (int)h->ng();
    const vh_dims& d = h->d;
    hipStream_t st = h->st;
    trap_alloc(h);
    if (!h->sig_primed) {   // sigma_0 from a report-only solve with the entering memory
        sle_solve(h, h->d_hice, h->d_hice, 0.0, 0, true, srot);
        d2d(h->d_sig_n, (const double*)h->d_sigma, (size_t)2 * d.nlm, st);
        h->sig_primed = true;
    }
    const int p = 2;   // scheme_order(trap)
    const double ricfac = (double)((1 << p) - 1), expo = 1.0 / (double)(p + 1);
    const size_t nm = (size_t)24 * std::max(d.nem, 1) * d.nk;
    double t = t0;
    if (h->dt_try <= 0.0) h->dt_try = span;
    auto solve_at = [&](double t_eval) {
        h->n_solve++;
        sle_res r = sle_solve_trap(h, h->d_hice, h->d_ice1, (t_eval - t0) / span, srot);
        h->worst_mass_resid = std::max(h->worst_mass_resid, r.mass_resid);
    };
    for (;;) {
        if (t >= t1 - 1.0e-9 * span) break;
        const double dt = std::min(h->dt_try, t1 - t);
        const bool at_floor = h->dt_min > 0.0 && dt <= h->dt_min * (1.0 + 1.0e-9);
        HIPCHECK(hipMemcpyAsync(h->d_rsl_n, h->d_rsl, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        trap_save(h);
        h->dt = dt;
        solve_at(t + dt);                                        // coarse: one dt
        d2d(h->d_tau_c, h->d_tau[h->cur], nm, st);               // buffer B
        if (h->tor_on) d2d(h->d_taut_c, h->d_taut[h->cur], nm / 2, st);
        trap_restore(h);
        HIPCHECK(hipMemcpyAsync(h->d_rsl, h->d_rsl_n, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        h->dt = 0.5 * dt;
        solve_at(t + 0.5 * dt);                                  // fine sub-step 1
        solve_at(t + dt);                                        // fine sub-step 2
        vh_coarse_fine(st, nm, h->d_tau[h->cur], h->d_tau_c, h->d_sc + 30);
        if (h->tor_on) vh_coarse_fine(st, nm / 2, h->d_taut[h->cur], h->d_taut_c, h->d_sc + 30, true);
        h->read(h->d_sc + 30, 2);
        const double err_inf = h->h_pin[0], tau_inf = h->h_pin[1];
        h->dt = dt;
        const double errsc = (err_inf / ricfac) / (h->atol + h->rtol * tau_inf);
        if (!(errsc <= DBL_MAX))
            throw vh_error("vilma_timestep: implicit local-error estimate went non-finite "
                           "(viscosity too stiff, or the state has already diverged)");
        const bool accept = errsc <= 1.0 || at_floor;
        if (accept) {
            t = t + dt;
            h->n_accept++;
            if (at_floor && errsc > 1.0) h->n_floor++;
        } else {
            trap_restore(h);
            HIPCHECK(hipMemcpyAsync(h->d_rsl, h->d_rsl_n, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            h->n_reject++;
        }
        double fac = errsc <= 0.0 ? h->grow_max : h->safety * std::pow(errsc, -expo);
        fac = std::min(h->grow_max, std::max(h->shrink_min, fac));
        h->dt_try = std::min(h->dt_max, std::max(h->dt_min, dt * fac));
        if (h->dt_try <= 1.0e-12 * span) throw vh_error("vilma_timestep: implicit adaptive step collapsed (tolerance unreachable)");
    }
}

// stepper_advance, forward-Euler branch (src/vilma_timestep.f90:79-227)
void stepper_advance(vh_model* h, double t0, double t1, const double* srot) {
    const double span = t1 - t0;
    const int n = (int)h->ng();
    h->worst_mass_resid = 0.0;
    if (span <= 0.0) {
        sle_res r = sle_solve(h, h->d_hice, h->d_hice, 0.0, 0, true, srot);
        h->worst_mass_resid = r.mass_resid;
        return;
    }
    if (h->trap) {
        stepper_trap(h, t0, t1, srot);
        return;
    }
    double rate = h->kind == 2 ? h->rate_max : 0.0;
    double dt_stab = rate > 0.0 ? h->cfl / rate : span;
    dt_stab = std::min(dt_stab, h->dt_max);
    if (h->dt_min > 0.0) dt_stab = std::max(dt_stab, h->dt_min);
    int n_sub = std::max(1, (int)std::ceil(span / dt_stab - 1.0e-9));
    double dt = span / (double)n_sub, t = t0;
    double tau_run = h->maxn_cur;
    for (;;) {
        if (t >= t1 - 1.0e-9 * span) break;
        dt = std::min(dt, t1 - t);
        HIPCHECK(hipMemcpyAsync(h->d_rsl_n, h->d_rsl, n * sizeof(double), hipMemcpyDeviceToDevice, h->st));
        double tau0 = h->maxn_cur, time_s = h->time_resp;
        h->dt = dt;
        double frac = ((t + dt) - t0) / span;
        h->n_solve++;
        sle_res r = sle_solve(h, h->d_hice, h->d_ice1, frac, 1, false, srot);
        h->worst_mass_resid = std::max(h->worst_mass_resid, r.mass_resid);
        double err_inf = h->kind == 2 ? r.err_inf : 0.0;
        bool finite = err_inf <= DBL_MAX;
        double scale = std::max(tau0, tau_run);
        bool accept = scale <= 0.0 ? finite : (finite && err_inf <= h->guard * scale);
        if (accept) {
            if (h->adj && adj_recording(h->adj)) adj_stash_push(h, dt, frac, r);
            if (h->kind == 2) { h->cur ^= 1; h->maxn_cur = err_inf; }
            tau_run = std::max(tau_run, err_inf);
            t = t + dt; h->n_accept++;
            dt = std::min(span / (double)n_sub, 2.0 * dt);
        } else {
            HIPCHECK(hipMemcpyAsync(h->d_rsl, h->d_rsl_n, n * sizeof(double), hipMemcpyDeviceToDevice, h->st));
            h->time_resp = time_s;
            h->n_reject++;
            dt = 0.5 * dt;
            if (dt <= 1.0e-12 * span) throw vh_error("vilma_timestep: explicit guard sub-step collapsed (viscosity too stiff for FE)");
        }
    }
}

// solid_earth_update (src/vilma_coupling.f90:336-437), passthrough grid
void update(vh_model* h, const double* h_ice, double dt_yr, bool on_device) {
    hipStream_t st = h->st;
    const int n = (int)h->ng();
    double dt = dt_yr * VH_SEC_PER_YEAR;
    double t0 = h->time_yr * VH_SEC_PER_YEAR, t1 = t0 + dt;
    if (on_device) HIPCHECK(hipMemcpyAsync(h->d_ice1, h_ice, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    else upload(h->d_ice1, h_ice, n, st);
    if (h->rot_on) {
        rot_begin(st, h->ra);
        rot_srot(st, h->ra, h->d.nphi, h->d.nlat, h->d_sc2, h->d_cphi, h->d_sphi, h->d_srot);
        stepper_advance(h, t0, t1, h->d_srot);
        vh_legendre_synthesis(st, h->s, h->d_sigma, 1, h->d_F);
        vh_fft_inverse(st, h->s, h->d_F, h->d.nlat, h->d_gtmp);
        rot_inertia(st, h->ra, h->s, h->d_gtmp, h->d_sc2, h->d_cphi, h->d_sphi, h->d_part);
        int n_sub = std::max(1, (int)std::ceil(dt / h->dt_fe_max));
        double dt_sub = dt / (double)n_sub;
        for (int k = 0; k < n_sub; k++) {
            rot_begin(st, h->ra);
            rot_solve_commit(st, h->ra, dt_sub);
        }
    } else {
        stepper_advance(h, t0, t1, nullptr);
    }
    HIPCHECK(hipMemcpyAsync(h->d_hice, h->d_ice1, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    grid_sub(st, n, h->d_topo0, h->d_rsl, h->d_zbed);
    h->time_yr += dt_yr;
    // update_bsl (src/vilma_coupling.f90:572-597)
    vh_surface_integral(st, h->s, h->d_C, h->d_part, h->d_sc + 12);
    grid_bsl_integrand(st, n, h->d_hice, h->d_C, h->d_heq, h->d_C0ref, h->d_gtmp);
    vh_surface_integral(st, h->s, h->d_gtmp, h->d_part + h->d.nlat, h->d_sc + 13);
    h->read(h->d_sc + 12, 2);
    h->prof_collect();
    double c_int = h->h_pin[0];
    h->bsl = c_int > 0.0 ? -(VH_RHO_ICE / VH_RHO_WATER) * h->h_pin[1] / c_int : 0.0;
}

}  // namespace

// ---------------------------------------------------------------- adjoint
// Reverse mode over a run of coupling calls: the forward record keeps the
// model state every `every` calls; the backward sweep restores each
// checkpoint in turn (last first), re-runs that segment's calls keeping every
// accepted FE sub-step's inputs in HBM, and walks the sub-steps backwards
// through the transposed kernels (vh_adjoint.hip). Discrete decisions (skip
// gate, coastline mask, sub-step sizes, 3-D element set) are those of the
// forward run; the sea-level solve's warm start is not differentiated (the
// converged rsl does not depend on it).
struct adj_sub {   // one accepted FE sub-step: tau_n (1-D elements), closing sigma, drift, last mask, norms
    double *tau = 0, *sigma = 0, *dN = 0, *C = 0, *mnorm = 0;
    size_t tau_n = 0;
    double2* Fd = 0;   // 3-D: the forward tensor row spectra of every batch
    size_t fd_n = 0;
    double dt = 0, frac = 0, c_int = 0;
    int call = 0, n_last = 0;
};
struct adj_ckpt {
    double *tau = 0, *mnorm = 0, *rsl = 0, *hice = 0;
    double *rmem = 0, *rsc = 0;   // rotation channels: memory and scalars
    double maxn_cur = 0, time_yr = 0, time_resp = 0;
};
struct adj_ctx {
    int ncalls = 0, every = 1;
    double dt_yr = 0;
    size_t ice_n = 0;
    double* ice = nullptr;   // [ncalls][ng]
    std::vector<adj_ckpt> ck;
    std::vector<adj_sub> sub;
    size_t nsub = 0;
    bool recording = false, recorded = false;
    // the record pass keeps the sub-steps of the last nstore segments
    // (entries seg_first[seg] .. + seg_n[seg]); the others are re-run
    bool stored_ok = false;
    int nstore = 1, ne3d_rec = -1;
    double store_gib = 96.0;
    std::vector<int> seg_first, seg_n;
    std::vector<int> e1list;   // Maxwell elements advanced by the 1-D update (their memory is stored)
    int* e1of = nullptr;       // [nem] -> index in e1list or -1
    int call = 0;
    // rotation channels (reverse): adjoint state, s_rot adjoint of the call
    // (density form), the load-inertia adjoint as a spectrum
    rot_adj ra{};
    double *srh = 0, *rtot = 0;
    double2* sig_rot = 0;
    double2* Fd_pending = nullptr;
    size_t fd_n = 0;
    // reverse sweep
    bool work = false;
    size_t rh_n = 0, ih_n = 0;
    double *tauh[2] = {0, 0};
    int ah = 0;
    double *Ue = 0, *mpart = 0, *Kh = 0, *dNh = 0, *dsurfh = 0, *sigh = 0, *Fh = 0;
    double *rate_h = 0, *iceh = 0, *rh = 0, *pgrad = 0;
    double2 *Xa = 0, *Fa = 0;
    double *gx = 0, *gb = 0, *gt = 0, *gy = 0, *part = 0, *sc = 0;
    double2 *F = 0, *q = 0, *q1 = 0;
    int *e3of_mesh = 0, *em_of = 0;
};

bool adj_recording(const adj_ctx* a) { return a->recording; }

static size_t adj_nfd(const vh_model* h) { return (size_t)h->nbat * 6 * (h->d.L + 1) * h->d.nlat; }

double2* adj_fd_pending(vh_model* h) {
    adj_ctx* A = h->adj;
    if (!A->Fd_pending || A->fd_n != adj_nfd(h)) {
        if (A->Fd_pending) HIPCHECK(hipFree(A->Fd_pending));
        A->Fd_pending = nullptr;
        A->Fd_pending = dalloc<double2>(adj_nfd(h));
        A->fd_n = adj_nfd(h);
    }
    return A->Fd_pending;
}

void adj_free(adj_ctx* a) {
    std::vector<void*> p = {a->ice, a->tauh[0], a->tauh[1], a->Ue, a->mpart, a->Kh, a->dNh, a->dsurfh, a->sigh,
                            a->Fh, a->rate_h, a->iceh, a->rh, a->pgrad, a->Xa, a->Fa, a->gx, a->gb, a->gt, a->gy,
                            a->part, a->sc, a->F, a->q, a->q1, a->e3of_mesh, a->em_of, a->Fd_pending, a->e1of, a->ra.mem, a->ra.dN, a->ra.sc, a->srh, a->rtot, a->sig_rot};
    for (auto& c : a->ck) {
        p.push_back(c.tau); p.push_back(c.mnorm); p.push_back(c.rsl); p.push_back(c.hice);
        p.push_back(c.rmem); p.push_back(c.rsc);
    }
    for (auto& u : a->sub) {
        p.push_back(u.tau); p.push_back(u.sigma); p.push_back(u.dN); p.push_back(u.C); p.push_back(u.mnorm);
        p.push_back(u.Fd);
    }
    for (void* x : p)
        if (x) (void)hipFree(x);
    delete a;
}

namespace {
size_t adj_ntau(const vh_model* h) { return (size_t)24 * std::max(h->d.nem, 1) * h->d.nk; }
size_t adj_ndN(const vh_model* h) { return (size_t)4 * std::max(h->d.nmn, 1) * h->d.nk; }
void d2d(double* dst, const double* src, size_t n, hipStream_t st) {
    HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st));
}

void adj_save(vh_model* h, adj_ckpt& c) {
    const size_t ng = h->ng();
    if (!c.tau) {
        c.tau = dalloc<double>(adj_ntau(h)); c.mnorm = dalloc<double>(h->d.nk + 1);
        c.rsl = dalloc<double>(ng); c.hice = dalloc<double>(ng);
    }
    d2d(c.tau, h->d_tau[h->cur], adj_ntau(h), h->st);
    d2d(c.mnorm, h->d_mnorm[h->cur], h->d.nk, h->st);
    d2d(c.mnorm + h->d.nk, h->d_maxn + h->cur, 1, h->st);
    d2d(c.rsl, h->d_rsl, ng, h->st);
    d2d(c.hice, h->d_hice, ng, h->st);
    if (h->rot_on) {
        const size_t nm = (size_t)2 * (h->d.nr - 1) * 24;
        if (!c.rmem) { c.rmem = dalloc<double>(nm); c.rsc = dalloc<double>(RS_N); }
        d2d(c.rmem, h->ra.rs.mem, nm, h->st);
        d2d(c.rsc, h->ra.rs.sc, RS_N, h->st);
    }
    c.maxn_cur = h->maxn_cur; c.time_yr = h->time_yr; c.time_resp = h->time_resp;
}

void adj_restore(vh_model* h, const adj_ckpt& c) {
    const size_t ng = h->ng();
    d2d(h->d_tau[h->cur], c.tau, adj_ntau(h), h->st);
    d2d(h->d_mnorm[h->cur], c.mnorm, h->d.nk, h->st);
    d2d(h->d_maxn + h->cur, c.mnorm + h->d.nk, 1, h->st);
    d2d(h->d_rsl, c.rsl, ng, h->st);
    d2d(h->d_hice, c.hice, ng, h->st);
    if (h->rot_on && c.rmem) {
        d2d(h->ra.rs.mem, c.rmem, (size_t)2 * (h->d.nr - 1) * 24, h->st);
        d2d(h->ra.rs.sc, c.rsc, RS_N, h->st);
    }
    h->maxn_cur = c.maxn_cur; h->time_yr = c.time_yr; h->time_resp = c.time_resp;
}

void adj_stash_push(vh_model* h, double dt, double frac, const sle_res& r) {
    adj_ctx* A = h->adj;
    const vh_dims& d = h->d;
    if (A->nsub == A->sub.size()) {
        adj_sub u;
        u.sigma = dalloc<double>(2 * (size_t)d.nlm);
        u.dN = dalloc<double>(adj_ndN(h)); u.C = dalloc<double>(h->ng()); u.mnorm = dalloc<double>(d.nk + 1);
        A->sub.push_back(u);
    }
    adj_sub& u = A->sub[A->nsub++];
    const size_t per_el = (size_t)24 * d.nk, ntau1 = per_el * A->e1list.size();
    if (u.tau_n != ntau1) {
        if (u.tau) HIPCHECK(hipFree(u.tau));
        u.tau = ntau1 ? dalloc<double>(ntau1) : nullptr;
        u.tau_n = ntau1;
    }
    if (A->e1list.size() == (size_t)d.nem) {
        d2d(u.tau, h->d_tau[h->cur], ntau1, h->st);
    } else {
        for (size_t i = 0; i < A->e1list.size(); i++)
            d2d(u.tau + i * per_el, h->d_tau[h->cur] + (size_t)A->e1list[i] * per_el, per_el, h->st);
    }
    d2d(u.sigma, (const double*)h->d_sigma, 2 * (size_t)d.nlm, h->st);
    d2d(u.dN, h->d_dN, adj_ndN(h), h->st);
    d2d(u.C, h->d_C, h->ng(), h->st);
    d2d(u.mnorm, h->d_mnorm[h->cur], d.nk, h->st);
    d2d(u.mnorm + d.nk, h->d_maxn + h->cur, 1, h->st);
    u.dt = dt; u.frac = frac; u.c_int = r.c_int; u.n_last = (int)r.n_last; u.call = A->call;
    if (h->lat_visc && h->ne3d > 0) {   // the sub-step's row spectra: swap the pending store in
        std::swap(u.Fd, A->Fd_pending);
        std::swap(u.fd_n, A->fd_n);
    }
}

void adj_check(const vh_model* h) {
    if (h->kind != 2) throw vh_error("adjoint: needs earth_response=ve");
}

void forward_record(vh_model* h, int ncalls, const double* ice_hist, bool dev, double dt_yr, int every,
                    double* rsl_out, bool rsl_dev) {
    adj_check(h);
    if (h->trap) throw vh_error("vh_forward_record: the reverse sweep differentiates scheme fe only");
    if (h->tor_on) throw vh_error("vh_forward_record: the reverse sweep does not carry the toroidal channels");
    if (ncalls <= 0) throw vh_error("vh_forward_record: ncalls must be positive");
    if (!(dt_yr > 0.0)) throw vh_error("vh_forward_record: dt_yr must be positive");
    if (!h->adj) h->adj = new adj_ctx();
    adj_ctx* A = h->adj;
    const size_t ng = h->ng();
    A->ncalls = ncalls; A->every = std::max(1, every); A->dt_yr = dt_yr; A->recorded = false;
    if (A->ice_n < (size_t)ncalls * ng) {
        if (A->ice) HIPCHECK(hipFree(A->ice));
        A->ice = dalloc<double>((size_t)ncalls * ng);
        A->ice_n = (size_t)ncalls * ng;
    }
    HIPCHECK(hipMemcpyAsync(A->ice, ice_hist, (size_t)ncalls * ng * sizeof(double),
                            dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, h->st));
    const int nseg = (ncalls + A->every - 1) / A->every;
    if ((int)A->ck.size() < nseg + 1) A->ck.resize(nseg + 1);
    // the 1-D Maxwell elements (their memory enters the rate adjoint)
    const vh_dims& d = h->d;
    std::vector<unsigned char> is3d(std::max(d.nem, 1), 0);
    if (h->lat_visc) HIPCHECK(hipMemcpy(is3d.data(), h->d_is3d, d.nem, hipMemcpyDeviceToHost));
    A->e1list.clear();
    std::vector<int> e1of(std::max(d.nem, 1), -1);
    for (int em = 0; em < d.nem; em++)
        if (!is3d[em]) { e1of[em] = (int)A->e1list.size(); A->e1list.push_back(em); }
    if (!A->e1of) A->e1of = dalloc<int>(std::max(d.nem, 1));
    upload(A->e1of, e1of.data(), std::max(d.nem, 1), h->st);
    A->ne3d_rec = h->lat_visc ? h->ne3d : 0;
    // trailing segments whose sub-steps fit the store limit are kept now and
    // not re-run by the backward sweep (at least the last one)
    {
        const double span = dt_yr * VH_SEC_PER_YEAR;
        double dt_stab = h->rate_max > 0.0 ? h->cfl / h->rate_max : span;
        dt_stab = std::min(dt_stab, h->dt_max);
        if (h->dt_min > 0.0) dt_stab = std::max(dt_stab, h->dt_min);
        const double nsub_call = std::max(1.0, std::ceil(span / dt_stab - 1.0e-9));
        const double per_sub = 8.0 * ((double)24 * A->e1list.size() * d.nk + 2.0 * d.nlm + (double)adj_ndN(h) +
                                      (double)ng + d.nk + 1) +
                                (h->lat_visc && h->ne3d > 0 ? 16.0 * (double)adj_nfd(h) : 0.0);
        const double per_seg = 1.1 * A->every * nsub_call * per_sub;
        A->nstore = (int)std::max(1.0, std::min((double)nseg, std::floor(A->store_gib * 1073741824.0 / per_seg)));
    }
    A->seg_first.assign(nseg, 0);
    A->seg_n.assign(nseg, 0);
    A->stored_ok = false;
    A->nsub = 0;
    for (int c = 0; c < ncalls; c++) {
        const int seg = c / A->every;
        if (c % A->every == 0) {
            adj_save(h, A->ck[seg]);
            if (seg >= nseg - A->nstore) { A->recording = true; A->seg_first[seg] = (int)A->nsub; }
        }
        A->call = c;
        try {
            update(h, A->ice + (size_t)c * ng, dt_yr, true);
        } catch (...) {
            A->recording = false;
            throw;
        }
        if (A->recording) A->seg_n[seg] = (int)A->nsub - A->seg_first[seg];
        if (rsl_out)
            HIPCHECK(hipMemcpyAsync(rsl_out + (size_t)c * ng, h->d_rsl, ng * sizeof(double),
                                    rsl_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, h->st));
    }
    A->recording = false;
    adj_save(h, A->ck[nseg]);
    h->sync();
    A->recorded = true;
    A->stored_ok = true;
}

void adj_alloc_work(vh_model* h) {
    adj_ctx* A = h->adj;
    const vh_dims& d = h->d;
    const size_t ng = h->ng();
    if (!A->work) {
        A->tauh[0] = dalloc<double>(adj_ntau(h)); A->tauh[1] = dalloc<double>(adj_ntau(h));
        A->Ue = dalloc<double>((size_t)8 * std::max(d.nem, 1) * d.nk);
        A->mpart = dalloc<double>((size_t)std::max(d.nem, 1) * vh_advance_adjoint_nbx(d));
        A->Kh = dalloc<double>(std::max(d.nem, 1));
        A->dNh = dalloc<double>(adj_ndN(h));
        A->dsurfh = dalloc<double>((size_t)6 * d.nk);
        A->sigh = dalloc<double>(2 * (size_t)d.nlm);
        A->Fh = dalloc<double>((size_t)4 * std::max(d.nmn, 1) * d.nk);
        A->gx = dalloc<double>(ng); A->gb = dalloc<double>(ng); A->gt = dalloc<double>(ng); A->gy = dalloc<double>(ng);
        A->part = dalloc<double>(2 * (size_t)d.nlat); A->sc = dalloc<double>(4);
        A->F = dalloc<double2>((size_t)d.nlat * (d.L + 1));
        A->q = dalloc<double2>(d.nlm); A->q1 = dalloc<double2>(d.nlm);
        A->pgrad = dalloc<double>((size_t)d.ne * ng);
        A->e3of_mesh = dalloc<int>(d.ne); A->em_of = dalloc<int>(d.ne);
        A->work = true;
    }
    if (h->rot_on && !A->ra.mem) {
        A->ra.mem = dalloc<double>((size_t)2 * (d.nr - 1) * 24);
        A->ra.dN = dalloc<double>((size_t)2 * 4 * d.nr);
        A->ra.sc = dalloc<double>(RA_N);
        A->srh = dalloc<double>(ng);
        A->rtot = dalloc<double>(2);
        A->sig_rot = dalloc<double2>(d.nlm);
    }
    // 3-D buffers follow the current lateral set
    if (A->rate_h) { HIPCHECK(hipFree(A->rate_h)); A->rate_h = 0; }
    if (A->Xa) { HIPCHECK(hipFree(A->Xa)); A->Xa = 0; }
    if (A->Fa) { HIPCHECK(hipFree(A->Fa)); A->Fa = 0; }
    if (h->lat_visc && h->ne3d > 0) {
        A->rate_h = dalloc<double>((size_t)h->ne3d * ng);
        A->Xa = dalloc<double2>((size_t)h->nbat * 4 * d.nlm);
        A->Fa = dalloc<double2>((size_t)std::max(h->tchunk, 1) * 6 * (d.L + 1) * d.nlat);
    }
    if (A->rh_n < (size_t)A->ncalls * ng) {
        if (A->rh) HIPCHECK(hipFree(A->rh));
        A->rh = dalloc<double>((size_t)A->ncalls * ng);
        A->rh_n = (size_t)A->ncalls * ng;
    }
    if (A->ih_n < (size_t)(A->ncalls + 1) * ng) {
        if (A->iceh) HIPCHECK(hipFree(A->iceh));
        A->iceh = dalloc<double>((size_t)(A->ncalls + 1) * ng);
        A->ih_n = (size_t)(A->ncalls + 1) * ng;
    }
    std::vector<int> e3of(d.ne, -1), emof(d.ne, -1);
    for (size_t i = 0; i < h->e3mesh.size() && h->lat_visc; i++) e3of[h->e3mesh[i]] = (int)i;
    for (int em = 0; em < d.nem; em++) emof[h->emap[em]] = em;
    upload(A->e3of_mesh, e3of.data(), d.ne, h->st);
    upload(A->em_of, emof.data(), d.ne, h->st);
}

// Transposed sea-level solve of one sub-step (src/vilma_sle.f90:296-355 for
// the forward): sigma_h (adjoint of the closing load transform) and rh (of
// the output rsl, or null) -> dsurf_h and the ice adjoints of the call's two
// ice fields. Grid adjoints are carried in density form (divided by W).
void sle_adjoint(vh_model* h, const adj_sub& S, const double* rh) {
    adj_ctx* A = h->adj;
    hipStream_t st = h->st;
    const vh_sht& s = h->s;
    const vh_dims& d = h->d;
    const size_t ng = h->ng();
    auto syn = [&](const double2* q, double* g) {
        vh_legendre_synthesis(st, s, q, 1, A->F);
        vh_fft_inverse(st, s, A->F, d.nlat, g);
    };
    auto anl = [&](const double* g, double2* q) {
        vh_fft_forward(st, s, g, d.nlat, A->F);
        vh_legendre_analysis(st, s, A->F, 1, q);
    };
    sle_adj_grid_args g{};
    g.n = (int)ng; g.nphi = d.nphi; g.dlon = 2.0 * VH_PI / d.nphi; g.w = s.w; g.C = S.C; g.frac = S.frac;
    g.ih0 = A->iceh + (size_t)S.call * ng; g.ih1 = A->iceh + (size_t)(S.call + 1) * ng;
    // q1 = sigma_h with m > 0 halved; t = Syn(q1) (so the sigma path's grid adjoint is W t)
    vh_sle_adj_spec(st, d.nlm, d.nk, 0, (const double2*)A->sigh, A->q1, nullptr, h->d_gu, h->d_gn, h->g, nullptr);
    syn(A->q1, A->gt);
    const int nl = S.c_int > 0.0 ? S.n_last : 0;
    if (nl > 0) {
        g.inv_cint = 1.0 / S.c_int;
        sle_adj_grid_args a = g;
        // x = rh/W + rho_w C t; b = x; then b = sum_{i<nl} (T~)^i x with
        // T~ x = rho_w C Syn((gn - gu) Anl(x - C I(x)/C_int))
        vh_sle_adj_row(st, s, S.C, A->gt, rh, 1, A->gx, A->gb, A->part);
        for (int i = 1; i < nl; i++) {
            vh_weighted_total(st, s, A->part, 1, A->sc);
            a = g; a.mode = 2; a.x = A->gx; a.I = A->sc; a.y = A->gy;
            vh_sle_adj_grid(st, a);
            anl(A->gy, A->q);
            vh_sle_adj_spec(st, d.nlm, d.nk, 1, nullptr, A->q, nullptr, h->d_gu, h->d_gn, h->g, nullptr);
            syn(A->q, A->gt);
            vh_sle_adj_row(st, s, S.C, A->gt, nullptr, 0, A->gx, A->gb, A->part);
        }
        // y = b - C I(b)/C_int; q = Anl(y) -> dsurf_h; q <- (gn - gu) q + q1
        vh_surface_integral(st, s, A->gb, A->part, A->sc + 1);
        a = g; a.mode = 2; a.x = A->gb; a.I = A->sc + 1; a.y = A->gy;
        vh_sle_adj_grid(st, a);
        if (h->rot_on) {   // s_rot enters Sraw: its adjoint is P^T b_h (density form y)
            a = g; a.mode = 6; a.x = A->gy; a.y = A->srh;
            vh_sle_adj_grid(st, a);
        }
        anl(A->gy, A->q);
        vh_sle_adj_spec(st, d.nlm, d.nk, 2, nullptr, A->q, A->q1, h->d_gu, h->d_gn, h->g, A->dsurfh);
        syn(A->q, A->gt);
        a = g; a.mode = 4; a.t = A->gt; a.I = A->sc + 1;
        vh_sle_adj_grid(st, a);
    } else {
        vh_fill(st, A->dsurfh, 0.0, (size_t)6 * d.nk);
        sle_adj_grid_args a = g;
        a.mode = 4; a.t = A->gt; a.I = nullptr; a.inv_cint = 0.0;
        vh_sle_adj_grid(st, a);
    }
}

void reverse_sub(vh_model* h, const adj_sub& S, const double* rh, const double2* sig_extra) {
    adj_ctx* A = h->adj;
    hipStream_t st = h->st;
    const vh_dims& d = h->d;
    const size_t ng = h->ng();
    double* th_in = A->tauh[A->ah];
    double* th_out = A->tauh[A->ah ^ 1];
    if (h->lat_visc && h->ne3d > 0) {
        const size_t per = (size_t)6 * (d.L + 1) * d.nlat;
        h->prof_begin(VH_K_T3_GATHER);
        vh_t3_pack_adj(st, d, h->v, h->d_e3em, h->ne3d, h->d_n6, th_in, A->Xa);
        h->prof_end();
        for (int b0 = 0; b0 < h->nbat; b0 += h->tchunk) {
            const int nb = std::min(h->tchunk, h->nbat - b0);
            h->prof_begin(VH_K_TSH_SYNTH);
            vh_tsh_synth(st, h->s, A->Xa + (size_t)b0 * 4 * d.nlm, nb, A->Fa);
            h->prof_end();
            h->prof_begin(VH_K_ADJ_ROWS);
            vh_t3_rows_adj(st, h->s, A->Fa, S.Fd + (size_t)b0 * per, nb / 3, h->d_rate3 + (size_t)(b0 / 3) * ng,
                           S.dt, A->rate_h + (size_t)(b0 / 3) * ng);
            h->prof_end();
            h->prof_begin(VH_K_TSH_ANALYSIS);
            vh_tsh_analysis(st, h->s, A->Fa, nb, h->d_n6, A->Xa + (size_t)b0 * 4 * d.nlm);
            h->prof_end();
        }
    }
    adv_adj_args a{};
    a.nr = d.nr; a.nk = d.nk; a.nem = d.nem; a.nlm = d.nlm; a.dt = S.dt; a.v = h->v;
    a.tauh_in = th_in; a.tauh_out = th_out; a.tau = S.tau; a.sigma = S.sigma; a.dN = S.dN;
    a.e3of = h->d_e3of; a.e1of = A->e1of; a.Xh = A->Xa; a.n6 = h->d_n6; a.Ue = A->Ue; a.mpart = A->mpart;
    h->prof_begin(VH_K_ADJ_ADVANCE);
    vh_advance_adjoint(st, d, a, h->d_mnodes, vh_drift_plan_node_el(h->dplan), A->dNh, A->sigh, A->Kh);
    if (sig_extra)   // the closing sigma also fed the rotation's load inertia
        vh_sle_adj_spec(st, d.nlm, d.nk, 4, sig_extra, (double2*)A->sigh, nullptr, h->d_gu, h->d_gn, h->g, nullptr);
    h->prof_end();
    h->prof_begin(VH_K_ADJ_SLE);
    sle_adjoint(h, S, rh);
    h->prof_end();
    h->prof_begin(VH_K_ADJ_DRIFT);
    vh_drift_adjoint(st, h->dplan, d, h->v, S.mnorm, S.mnorm + d.nk, h->skip_tol, h->deg1_cm, A->dNh, A->dsurfh,
                     A->Fh, th_out);
    h->prof_end();
    A->ah ^= 1;
}

// The rotation part of a call after its stepper (src/vilma_coupling.f90:399-420),
// transposed: the channel sub-intervals last first, then the load inertia
// of the closing sigma -> A->sig_rot (added to the last sub-step's sigma adjoint).
void rot_call_tail_adjoint(vh_model* h) {
    adj_ctx* A = h->adj;
    hipStream_t st = h->st;
    const vh_dims& d = h->d;
    const double dt = A->dt_yr * VH_SEC_PER_YEAR;
    const int n_sub = std::max(1, (int)std::ceil(dt / h->dt_fe_max));
    const double dt_sub = dt / (double)n_sub;
    for (int k = n_sub - 1; k >= 0; k--) {
        rot_commit_adj(st, h->ra, A->ra, dt_sub);
        rot_begin_adj(st, h->ra, A->ra);
    }
    // I = int load gamma (cos, sin), load = Syn(sigma): sigma_h = c_m Anl(gamma (I13_h cos + I23_h sin))
    rot_inertia_adj(st, A->ra, d.nphi, d.nlat, h->d_sc2, h->d_cphi, h->d_sphi, A->gt);
    vh_fft_forward(st, h->s, A->gt, d.nlat, A->F);
    vh_legendre_analysis(st, h->s, A->F, 1, A->q);
    vh_sle_adj_spec(st, d.nlm, d.nk, 3, A->q, A->sig_rot, nullptr, h->d_gu, h->d_gn, h->g, nullptr);
    vh_fill(st, A->ra.sc + RA_I, 0.0, 2);
}

// The rotation part of a call before its stepper (rot_begin, s_rot), transposed,
// from the call's accumulated s_rot adjoint; leaves m_h for the previous call.
void rot_call_head_adjoint(vh_model* h) {
    adj_ctx* A = h->adj;
    hipStream_t st = h->st;
    rot_srot_adj(st, h->ra, A->ra, h->s, A->srh, h->d_sc2, h->d_cphi, h->d_sphi, A->part, A->rtot);
    rot_begin_adj(st, h->ra, A->ra);
    vh_fill(st, A->srh, 0.0, h->ng());
}

void backward(vh_model* h, const double* dJ_drsl, bool dev, double* dJ_dpert, double* dJ_dice, bool out_dev) {
    adj_ctx* A = h->adj;
    if (!A || !A->recorded) throw vh_error("vh_backward: no forward record (call vh_forward_record first)");
    adj_check(h);
    hipStream_t st = h->st;
    const vh_dims& d = h->d;
    const size_t ng = h->ng();
    adj_alloc_work(h);
    HIPCHECK(hipMemcpyAsync(A->rh, dJ_drsl, (size_t)A->ncalls * ng * sizeof(double),
                            dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    A->ah = 0;
    if (h->rot_on) {
        vh_fill(st, A->ra.mem, 0.0, (size_t)2 * (d.nr - 1) * 24);
        vh_fill(st, A->ra.dN, 0.0, (size_t)2 * 4 * d.nr);
        vh_fill(st, A->ra.sc, 0.0, RA_N);
        vh_fill(st, A->srh, 0.0, ng);
    }
    vh_fill(st, A->tauh[0], 0.0, adj_ntau(h));
    vh_fill(st, A->Kh, 0.0, std::max(d.nem, 1));
    vh_fill(st, A->iceh, 0.0, (size_t)(A->ncalls + 1) * ng);
    if (A->rate_h) vh_fill(st, A->rate_h, 0.0, (size_t)h->ne3d * ng);
    const int nseg = (A->ncalls + A->every - 1) / A->every;
    if ((h->lat_visc ? h->ne3d : 0) != A->ne3d_rec)
        throw vh_error("vh_backward: the lateral viscosity changed since vh_forward_record");
    const int first_stored = A->stored_ok ? nseg - A->nstore : nseg;
    size_t top = 0;   // entries above the stored segments take the re-run ones
    for (int seg = first_stored; seg < nseg; seg++) top = std::max(top, (size_t)A->seg_first[seg] + A->seg_n[seg]);
    for (int seg = nseg - 1; seg >= 0; seg--) {
        size_t first, n;
        if (seg >= first_stored) {
            first = A->seg_first[seg];
            n = A->seg_n[seg];
        } else {   // re-run the segment from its checkpoint
            adj_restore(h, A->ck[seg]);
            A->nsub = top;
            A->recording = true;
            const int c1 = std::min(A->ncalls, (seg + 1) * A->every);
            try {
                for (int c = seg * A->every; c < c1; c++) {
                    A->call = c;
                    update(h, A->ice + (size_t)c * ng, A->dt_yr, true);
                }
            } catch (...) {
                A->recording = false;
                throw;
            }
            A->recording = false;
            first = top;
            n = A->nsub - top;
        }
        for (size_t i = first + n; i-- > first;) {
            const adj_sub& S = A->sub[i];
            const bool last = i + 1 == first + n || A->sub[i + 1].call != S.call;
            const bool head = i == first || A->sub[i - 1].call != S.call;
            if (h->rot_on && last) rot_call_tail_adjoint(h);
            reverse_sub(h, S, last ? A->rh + (size_t)S.call * ng : nullptr, h->rot_on && last ? A->sig_rot : nullptr);
            if (h->rot_on && head) rot_call_head_adjoint(h);
        }
        h->sync();
        h->prof_collect();
    }
    adj_restore(h, A->ck[nseg]);
    vh_pert_grad(st, d.ne, ng, A->e3of_mesh, A->em_of, h->d_MkPerDt, h->d_rate3, A->rate_h, A->Kh, A->pgrad);
    const hipMemcpyKind k = out_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    if (dJ_dpert) HIPCHECK(hipMemcpyAsync(dJ_dpert, A->pgrad, (size_t)d.ne * ng * sizeof(double), k, st));
    if (dJ_dice) HIPCHECK(hipMemcpyAsync(dJ_dice, A->iceh, (size_t)(A->ncalls + 1) * ng * sizeof(double), k, st));
    h->sync();
}

// degree-major (device) <-> SHTns LM order (API) for complex spectra
inline int lm_shtns(int L, int l, int m) { return m * (2 * L + 1 - m) / 2 + l; }

void to_lm(const vh_model* h, const double2* dm, double* out) {
    int L = h->d.L;
    for (int l = 0; l <= L; l++)
        for (int m = 0; m <= l; m++) {
            int a = l * (l + 1) / 2 + m, b = lm_shtns(L, l, m);
            out[2 * b] = dm[a].x; out[2 * b + 1] = dm[a].y;
        }
}
void from_lm(const vh_model* h, const double* in, double2* dm) {
    int L = h->d.L;
    for (int l = 0; l <= L; l++)
        for (int m = 0; m <= l; m++) {
            int a = l * (l + 1) / 2 + m, b = lm_shtns(L, l, m);
            dm[a] = make_double2(in[2 * b], in[2 * b + 1]);
        }
}

// the vh_remap_* calls: errors through the model-less channel of vh_last_error(NULL)
template <class F>
int remap_guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return 1;
    }
}

template <class F>
int guarded(vh_model* h, F&& f) {
    if (!h) return 2;
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        h->err = e.what();
        return 1;
    }
}

}  // namespace

// ================================================================ C-ABI
extern "C" {

int vh_params_defaults(vh_params* p) {
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
        p->r_bot[i] = rb[i]; p->r_top[i] = rt[i]; p->rho[i] = rho[i]; p->mu[i] = mu[i]; p->eta[i] = eta[i]; p->rheology[i] = rh[i];
    }
    strcpy(p->scheme, "fe"); p->max_couple_iter = 20; strcpy(p->earth_response, "ve");
    p->sle_n_outer = 3; p->sle_n_inner = 20; p->sle_tol = 1.0e-7; p->sle_max_mem_iter = 20;
    p->sle_fixed_ocean = 0; p->sle_subgrid = 1;
    p->dt_init = 0.0; p->dt_min = 0.0; p->dt_max = 1.0e30 * VH_SEC_PER_YEAR;
    p->rtol = 1.0e-4; p->atol = 1.0e-3; p->safety = 0.9; p->grow_max = 5.0; p->shrink_min = 0.2; p->cfl = 1.0;
    p->rotation = 1; strcpy(p->deg1_frame, "cm");
    p->l_visc_3d = 0; p->visc3d_tol = 1.0e-3; p->l_toroidal = 0;
    p->visc_log10_min = 19.5; p->visc_log10_max = 30.0; p->visc3d_lid_depth = 0.0; p->visc3d_lid_log10max = 22.0;
    p->skip_tol = 1.0e-4;
    return 0;
}

int vh_create(const vh_params* p, int nphi, int nlat, const double* z_bed_eq, const double* h_ice_eq,
              const double* h_ice_init, vh_model** out) {
    *out = nullptr;
    vh_model* h = nullptr;
    try {
        h = new vh_model();
        setup(h, p, nphi, nlat, z_bed_eq, h_ice_eq, h_ice_init);
        *out = h;
        return 0;
    } catch (const std::exception& e) {
        g_create_error = std::string("solid_earth_init: ") + e.what();
        delete h;
        return 1;
    }
}

int vh_update(vh_model* h, const double* h_ice, double dt_yr) {
    return guarded(h, [&] { update(h, h_ice, dt_yr, false); });
}

int vh_update_device(vh_model* h, const double* h_ice_device, double dt_yr) {
    return guarded(h, [&] { update(h, h_ice_device, dt_yr, true); });
}

int vh_get_field(vh_model* h, int field, double* out) {
    return guarded(h, [&] {
        const double* src = nullptr;
        switch (field) {
            case VH_F_RSL: src = h->d_rsl; break;
            case VH_F_ZBED: src = h->d_zbed; break;
            case VH_F_C: src = h->d_C; break;
            case VH_F_C0: src = h->d_C0ref; break;
            case VH_F_SROT: src = h->d_srot; break;
            case VH_F_HICE: src = h->d_hice; break;
            default: throw vh_error("vh_get_field: unknown field");
        }
        HIPCHECK(hipMemcpyAsync(out, src, h->ng() * sizeof(double), hipMemcpyDeviceToHost, h->st));
        h->sync();
    });
}

int vh_get_scalar(vh_model* h, int which, double* out) {
    return guarded(h, [&] {
        switch (which) {
            case VH_S_TIME: *out = h->time_yr; break;
            case VH_S_BSL: *out = h->bsl; break;
            case VH_S_WORST_MASS_RESID: *out = h->worst_mass_resid; break;
            case VH_S_N_ACCEPT: *out = (double)h->n_accept; break;
            case VH_S_N_REJECT: *out = (double)h->n_reject; break;
            case VH_S_N_SOLVE: *out = (double)h->n_solve; break;
            case VH_S_N_INNER_TOT: *out = (double)h->n_inner_tot; break;
            case VH_S_N_OUTER_TOT: *out = (double)h->n_outer_tot; break;
            case VH_S_N_SLE_STEAL: *out = (double)h->n_sle_steal; break;
            case VH_S_N_FLOOR: *out = (double)h->n_floor; break;
            case VH_S_N_COUPLE_TOT: *out = (double)h->n_couple_tot; break;
            case VH_S_N_GRAPH_LAUNCH: *out = (double)h->n_graph_launch; break;
            case VH_S_N_GRAPH_CAPTURE: *out = (double)h->n_graph_capture; break;
            case VH_S_POLAR_M_RE:
            case VH_S_POLAR_M_IM:
                if (!h->rot_on) { *out = 0.0; break; }
                h->read(h->ra.rs.sc + RS_M, 2);
                *out = h->h_pin[which == VH_S_POLAR_M_RE ? 0 : 1];
                break;
            default: throw vh_error("vh_get_scalar: unknown selector");
        }
    });
}

int vh_get_spectrum(vh_model* h, int which, double* out) {
    return guarded(h, [&] {
        std::vector<double2> tmp(h->d.nlm);
        if (which == VH_SP_SIGMA) {
            HIPCHECK(hipMemcpyAsync(tmp.data(), h->d_sigma, h->d.nlm * sizeof(double2), hipMemcpyDeviceToHost, h->st));
            h->sync();
            to_lm(h, tmp.data(), out);
            return;
        }
        if (which == VH_SP_TOR) {   // response_horizontal_toroidal (src/vilma_response.f90:884-900)
            std::fill(tmp.begin(), tmp.end(), make_double2(0.0, 0.0));
            if (h->tor_on) {
                HIPCHECK(hipMemcpyAsync(tmp.data() + 1, h->d_dWa, (size_t)h->d.nk * sizeof(double2),
                                        hipMemcpyDeviceToHost, h->st));
                h->sync();
            }
            to_lm(h, tmp.data(), out);
            return;
        }
        if (which != VH_SP_U && which != VH_SP_N) throw vh_error("vh_get_spectrum: unknown selector");
        // response_apply at the current load and frozen drift
        std::vector<double2> uN(2 * (size_t)h->d.nlm);
        vh_apply_only(h->st, h->d.nlm, h->d_sigma, h->d_gu, h->d_gn, h->d_dsurf, h->d.nk, h->g, h->kind, h->d_uN);
        HIPCHECK(hipMemcpyAsync(uN.data(), h->d_uN, uN.size() * sizeof(double2), hipMemcpyDeviceToHost, h->st));
        h->sync();
        to_lm(h, uN.data() + (which == VH_SP_U ? 0 : h->d.nlm), out);
    });
}

int vh_get_dims(vh_model* h, int which, int* out) {
    return guarded(h, [&] {
        switch (which) {
            case VH_D_LMAX: *out = h->d.L; break;
            case VH_D_NLAT: *out = h->d.nlat; break;
            case VH_D_NPHI: *out = h->d.nphi; break;
            case VH_D_NLM: *out = h->d.nlm; break;
            case VH_D_NR: *out = h->d.nr; break;
            case VH_D_NE: *out = h->d.ne; break;
            case VH_D_NK: *out = h->d.nk; break;
            case VH_D_NE3D: *out = h->ne3d; break;
            case VH_D_NLAM: *out = h->tor_on ? 6 : 4; break;
            case VH_D_NEM: *out = h->d.nem; break;
            case VH_D_NMN: *out = h->d.nmn; break;
            default: throw vh_error("vh_get_dims: unknown selector");
        }
    });
}

// reference restart layout: 6 arrays (nlam, ne, nk) column-major, nlam = 4,
// or 6 once the toroidal channels are carried (lambda 3, 4 at 5, 6)
int vh_get_memory(vh_model* h, double* tau) {
    return guarded(h, [&] {
        if (h->kind != 2) throw vh_error("vh_get_memory: no viscoelastic memory");
        const vh_dims& d = h->d;
        const int nl = h->tor_on ? 6 : 4;
        std::vector<double> dev((size_t)24 * d.nem * d.nk), devt(h->tor_on ? (size_t)12 * d.nem * d.nk : 0);
        HIPCHECK(hipMemcpyAsync(dev.data(), h->d_tau[h->cur], dev.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
        if (h->tor_on)
            HIPCHECK(hipMemcpyAsync(devt.data(), h->d_taut[h->cur], devt.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    h->st));
        h->sync();
        size_t n1 = (size_t)nl * d.ne * d.nk;
        memset(tau, 0, 6 * n1 * sizeof(double));
        // array order Are, Aim, Bre, Bim, Cre, Cim = (shape, reim)
        for (int em = 0; em < d.nem; em++) {
            int e = h->emap[em];
            for (int sh = 0; sh < 3; sh++)
                for (int ri = 0; ri < 2; ri++) {
                    double* dst = tau + (size_t)(sh * 2 + ri) * n1;
                    for (int lam = 0; lam < 4; lam++) {
                        const double* src = dev.data() + ((size_t)em * 24 + (sh * 2 + ri) * 4 + lam) * d.nk;
                        for (int k = 0; k < d.nk; k++) dst[((size_t)k * d.ne + e) * nl + lam] = src[k];
                    }
                    if (h->tor_on)
                        for (int q = 0; q < 2; q++) {
                            const double* src = devt.data() + ((size_t)em * 12 + (sh * 2 + ri) * 2 + q) * d.nk;
                            for (int k = 0; k < d.nk; k++) dst[((size_t)k * d.ne + e) * nl + 4 + q] = src[k];
                        }
                }
        }
    });
}

int vh_set_memory(vh_model* h, const double* tau) {
    return guarded(h, [&] {
        if (h->kind != 2) throw vh_error("vh_set_memory: no viscoelastic memory");
        const vh_dims& d = h->d;
        const int nl = h->tor_on ? 6 : 4;
        std::vector<double> dev((size_t)24 * d.nem * d.nk), devt(h->tor_on ? (size_t)12 * d.nem * d.nk : 0);
        size_t n1 = (size_t)nl * d.ne * d.nk;
        for (int em = 0; em < d.nem; em++) {
            int e = h->emap[em];
            for (int sh = 0; sh < 3; sh++)
                for (int ri = 0; ri < 2; ri++) {
                    const double* src = tau + (size_t)(sh * 2 + ri) * n1;
                    for (int lam = 0; lam < 4; lam++) {
                        double* dst = dev.data() + ((size_t)em * 24 + (sh * 2 + ri) * 4 + lam) * d.nk;
                        for (int k = 0; k < d.nk; k++) dst[k] = src[((size_t)k * d.ne + e) * nl + lam];
                    }
                    if (h->tor_on)
                        for (int q = 0; q < 2; q++) {
                            double* dst = devt.data() + ((size_t)em * 12 + (sh * 2 + ri) * 2 + q) * d.nk;
                            for (int k = 0; k < d.nk; k++) dst[k] = src[((size_t)k * d.ne + e) * nl + 4 + q];
                        }
                }
        }
        // elastic and fluid elements have a zero Maxwell rate, so the update
        // leaves their memory at its start value, zero (src/vilma_viscoelastic.f90:
        // 577-584 with Mk = 0); the library keeps memory for Maxwell elements
        // only, so a state that gives the others memory is refused, not dropped
        std::vector<char> maxwell(d.ne, 0);
        for (int em = 0; em < d.nem; em++) maxwell[h->emap[em]] = 1;
        for (size_t a6 = 0; a6 < 6; a6++)
            for (int k = 0; k < d.nk; k++)
                for (int e = 0; e < d.ne; e++) {
                    if (maxwell[e]) continue;
                    const double* v = tau + a6 * n1 + ((size_t)k * d.ne + e) * nl;
                    for (int lam = 0; lam < nl; lam++)
                        if (v[lam] != 0.0)
                            throw vh_error("vh_set_memory: memory of an elastic or fluid element must be zero (element " +
                                           std::to_string(e + 1) + ")");
                }
        upload(h->d_tau[h->cur], dev.data(), dev.size(), h->st);
        if (h->tor_on) upload(h->d_taut[h->cur], devt.data(), devt.size(), h->st);
        vh_launch_memnorm(h->st, d, h->d_tau[h->cur], h->d_mnorm[h->cur], h->d_maxn + h->cur,
                          h->tor_on ? h->d_taut[h->cur] : nullptr);
        h->read(h->d_maxn + h->cur, 1);
        h->maxn_cur = h->h_pin[0];
    });
}

// read_ve_state's cross-resolution branch (src/vilma_io.f90:515-552): the
// memory of a run at a lower truncation (nk_src slots, degree-grouped, so the
// shared low-degree block is a prefix of every slot index) copied into the
// low degrees, the higher ones zero; sigma_n is not carried over (the next
// trapezoidal step re-primes it)
int vh_set_memory_lowres(vh_model* h, const double* tau, int nk_src) {
    return guarded(h, [&] {
        if (h->kind != 2) throw vh_error("vh_set_memory_lowres: no viscoelastic memory");
        const vh_dims& d = h->d;
        if (nk_src > d.nk)
            throw vh_error("vilma_restart_read: cross-resolution downsampling (file lmax > model) not supported");
        bool ok = false;
        for (int L = 1; L <= d.L; L++)
            if (L * (L + 3) / 2 == nk_src) ok = true;   // slots of degrees 1..L
        if (!ok) throw vh_error("vilma_restart_read: cross-resolution restart incompatible (mmax/mres mismatch)");
        const int nl = h->tor_on ? 6 : 4;   // the source carries the model's channels
        const size_t n_src = (size_t)nl * d.ne * nk_src, n_dst = (size_t)nl * d.ne * d.nk;
        std::vector<double> full(6 * n_dst, 0.0);
        for (int a6 = 0; a6 < 6; a6++)   // (nlam, ne, nk) column-major: k slowest, so a prefix
            std::copy(tau + a6 * n_src, tau + (a6 + 1) * n_src, full.begin() + a6 * n_dst);
        if (vh_set_memory(h, full.data())) throw vh_error(h->err);
        h->sig_primed = false;
    });
}

// The step state of the trapezoidal scheme that a restart carries
// (src/vilma_io.f90:453-457, :548-551): the controller's next step size
// (years) and the start-of-step load sigma_n (interleaved complex, SHTns LM
// order, nlm values) with its primed flag.
int vh_get_step_state(vh_model* h, double* dt_try_yr, int* sigma_primed, double* sigma_n) {
    return guarded(h, [&] {
        if (dt_try_yr) *dt_try_yr = h->dt_try / VH_SEC_PER_YEAR;
        if (sigma_primed) *sigma_primed = h->sig_primed ? 1 : 0;
        if (sigma_n) {
            if (!h->d_sig_n) {
                std::fill(sigma_n, sigma_n + 2 * (size_t)h->d.nlm, 0.0);
                return;
            }
            std::vector<double2> t(h->d.nlm);
            HIPCHECK(hipMemcpyAsync(t.data(), h->d_sig_n, t.size() * sizeof(double2), hipMemcpyDeviceToHost, h->st));
            h->sync();
            to_lm(h, (const double2*)t.data(), sigma_n);
        }
    });
}

int vh_set_step_state(vh_model* h, double dt_try_yr, int sigma_primed, const double* sigma_n) {
    return guarded(h, [&] {
        h->dt_try = dt_try_yr * VH_SEC_PER_YEAR;
        if (h->kind != 2 || !h->trap) { h->sig_primed = sigma_primed != 0; return; }
        trap_alloc(h);
        if (sigma_n) {
            std::vector<double2> t(h->d.nlm);
            from_lm(h, sigma_n, t.data());
            upload((double2*)h->d_sig_n, t.data(), t.size(), h->st);
            h->sync();
        }
        h->sig_primed = sigma_primed != 0 && sigma_n != nullptr;
    });
}

// The prognostic state besides the memory (src/vilma_io.f90:425-491): the
// clock, the rsl that seeds the next sea-level solve, the ice at the call
// boundary, and the rotation state (m and both channels' memory in the
// reference packing (NLAM, ne, 6) of rotation_get_memory,
// src/vilma_rotation.f90:275-307).
// ---- host lon-lat grid <-> Gauss grid (src/vilma_remap.f90); a map owns
// its stream and reports errors through vh_last_error(NULL)
struct vh_remap_handle {
    hipStream_t st = nullptr;
    vh_remap* r = nullptr;
};

int vh_remap_create(int nphi, int nlat, int nlon, int nlat_ll, const double* lon_deg, const double* lat_deg,
                    vh_remap_handle** out) {
    return remap_guarded([&] {
        if (!out || !lon_deg || !lat_deg) throw vh_error("remap_init: null argument");
        *out = nullptr;
        if (nlat < 2 || nphi < 2) throw vh_error("remap_init: Gauss grid must be at least 2 x 2");
        std::vector<double> ct, w, colat(nlat);
        gauss_legendre(nlat, ct, w);   // north-first nodes cos(colat) and weights, as the model's grid
        for (int j = 0; j < nlat; j++) colat[j] = std::acos(ct[j]);
        auto* m = new vh_remap_handle();
        try {
            HIPCHECK(hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
            m->r = vh_remap_build(m->st, nphi, nlat, colat.data(), w.data(), nlon, nlat_ll, lon_deg, lat_deg);
            vh_remap_scratch_setup();   // this device's address of the allocation-free variant's array
        } catch (...) {
            if (m->st) (void)hipStreamDestroy(m->st);
            delete m;
            throw;
        }
        *out = m;
    });
}

int vh_remap_to_gauss(vh_remap_handle* m, const double* f_ll, double* f_gauss, int conserve_mass) {
    if (!m) return 2;
    return remap_guarded([&] { vh_remap_to_gauss_host(m->r, f_ll, f_gauss, conserve_mass); });
}

int vh_remap_to_gauss_noalloc(vh_remap_handle* m, const double* f_ll, double* f_gauss, int conserve_mass) {
    if (!m) return 2;
    return remap_guarded([&] {
        if (!f_ll || !f_gauss) throw vh_error("remap_to_gauss_noalloc: null field");
        vh_remap_to_gauss_noalloc(m->r, f_ll, f_gauss, conserve_mass);
    });
}

int vh_remap_to_ll(vh_remap_handle* m, const double* f_gauss, double* f_ll) {
    if (!m) return 2;
    return remap_guarded([&] { vh_remap_to_ll_host(m->r, f_gauss, f_ll); });
}

int vh_remap_to_gauss_device(vh_remap_handle* m, const double* f_ll, double* f_gauss, int conserve_mass) {
    if (!m) return 2;
    return remap_guarded([&] {
        vh_remap_to_gauss_dev(m->r, f_ll, f_gauss, conserve_mass);
        HIPCHECK(hipStreamSynchronize(m->st));
    });
}

int vh_remap_to_ll_device(vh_remap_handle* m, const double* f_gauss, double* f_ll) {
    if (!m) return 2;
    return remap_guarded([&] {
        vh_remap_to_ll_dev(m->r, f_gauss, f_ll);
        HIPCHECK(hipStreamSynchronize(m->st));
    });
}

int vh_remap_destroy(vh_remap_handle* m) {
    if (!m) return 0;
    vh_remap_free(m->r);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
    return 0;
}

int vh_get_state(vh_model* h, double* time_yr, double* rsl, double* h_ice, double* rot_m, double* rot_load_mem,
                 double* rot_tidal_mem) {
    return guarded(h, [&] {
        const size_t ng = h->ng();
        const int ne = h->d.ne;
        if (time_yr) *time_yr = h->time_yr;
        if (rsl) HIPCHECK(hipMemcpyAsync(rsl, h->d_rsl, ng * sizeof(double), hipMemcpyDeviceToHost, h->st));
        if (h_ice) HIPCHECK(hipMemcpyAsync(h_ice, h->d_hice, ng * sizeof(double), hipMemcpyDeviceToHost, h->st));
        std::vector<double> mem((size_t)2 * ne * 24, 0.0), sc(RS_N, 0.0);
        if (h->rot_on) {
            HIPCHECK(hipMemcpyAsync(mem.data(), h->ra.rs.mem, mem.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
            HIPCHECK(hipMemcpyAsync(sc.data(), h->ra.rs.sc, RS_N * sizeof(double), hipMemcpyDeviceToHost, h->st));
        }
        h->sync();
        if (rot_m) { rot_m[0] = sc[RS_M]; rot_m[1] = sc[RS_M + 1]; }
        double* outs[2] = {rot_load_mem, rot_tidal_mem};
        for (int ch = 0; ch < 2; ch++)
            if (outs[ch])
                for (int comp = 0; comp < 6; comp++)
                    for (int e = 0; e < ne; e++)
                        for (int lam = 0; lam < 4; lam++)
                            outs[ch][lam + 4 * (e + (size_t)ne * comp)] = mem[((size_t)ch * ne + e) * 24 + comp * 4 + lam];
    });
}

int vh_set_state(vh_model* h, double time_yr, const double* rsl, const double* h_ice, const double* rot_m,
                 const double* rot_load_mem, const double* rot_tidal_mem) {
    return guarded(h, [&] {
        const size_t ng = h->ng();
        const int ne = h->d.ne;
        h->time_yr = time_yr;
        h->time_resp = time_yr * VH_SEC_PER_YEAR;
        if (rsl) {
            upload(h->d_rsl, rsl, ng, h->st);
            grid_sub(h->st, (int)ng, h->d_topo0, h->d_rsl, h->d_zbed);
        }
        if (h_ice) upload(h->d_hice, h_ice, ng, h->st);
        if (h->rot_on) {   // ignored with rotation off (a cold start of the rotation, as the reference)
            if (rot_m) upload(h->ra.rs.sc + RS_M, rot_m, 2, h->st);
            const double* ins[2] = {rot_load_mem, rot_tidal_mem};
            std::vector<double> mem((size_t)2 * ne * 24);
            if (ins[0] || ins[1]) {
                HIPCHECK(hipMemcpyAsync(mem.data(), h->ra.rs.mem, mem.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
                h->sync();
                for (int ch = 0; ch < 2; ch++)
                    if (ins[ch])
                        for (int comp = 0; comp < 6; comp++)
                            for (int e = 0; e < ne; e++)
                                for (int lam = 0; lam < 4; lam++)
                                    mem[((size_t)ch * ne + e) * 24 + comp * 4 + lam] = ins[ch][lam + 4 * (e + (size_t)ne * comp)];
                upload(h->ra.rs.mem, mem.data(), mem.size(), h->st);
            }
        }
        h->sync();
    });
}

int vh_destroy(vh_model* h) {
    delete h;
    return 0;
}

const char* vh_last_error(const vh_model* h) {
    return h ? h->err.c_str() : g_create_error.c_str();
}

// tensor_sh_init (src/vilma_tensor_sh.f90:76-133): the n6 calibration
void tensor_init(vh_model* h) {
    if (h->tsh_ready) return;
    h->d_n6 = dalloc<double>(h->d.L + 1);
    vh_tsh_calibrate_n6(h->st, h->s, h->d_n6);
    h->tsh_ready = true;
}

// enable_toroidal (src/vilma_response.f90:1055-1109): the W operators and
// forcing tables, the toroidal memory (zero: the exact toroidal state of a run
// that has only been spheroidal), the W drift arrays, the Z4 normalisation,
// and the trapezoidal buffers if that scheme has allocated its own. The
// advance's reductions gain the toroidal rows.
void enable_toroidal(vh_model* h) {
    const vh_dims& d = h->d;
    hipStream_t st = h->st;
    h->sync();   // the reduction scratch is replaced below
    h->drop_graphs();
    h->tplan = vh_tor_plan_create(st, d, h->r, h->mu, h->emap, h->mnodes, h->mnode_of);
    const size_t nt = (size_t)12 * std::max(d.nem, 1) * d.nk, ndw = (size_t)2 * std::max(d.nmn, 1) * d.nk;
    for (int b = 0; b < 2; b++) { h->d_taut[b] = dalloc<double>(nt); vh_fill(st, h->d_taut[b], 0.0, nt); }
    h->d_dWt = dalloc<double>(ndw); vh_fill(st, h->d_dWt, 0.0, ndw);
    h->d_dWa = dalloc<double>((size_t)2 * d.nk); vh_fill(st, h->d_dWa, 0.0, (size_t)2 * d.nk);
    h->d_n4 = dalloc<double>(d.L + 1);
    vh_tsh_calibrate_n4(st, h->s, h->d_n4);
    HIPCHECK(hipFree(h->d_advmax));
    h->d_advmax = nullptr;
    h->d_advmax = dalloc<double>(vh_advance_scratch(d, true));
    h->tor_on = true;
    h->nch = 6;
    if (h->d_tau_s) trap_alloc_tor(h);
    h->sync();
}

// response_enable_lateral_visc (src/vilma_response.f90:994-1053); pert is
// [e][nlat][nphi] (phi fastest). With l_toroidal, the first genuinely 3-D
// element turns the toroidal channels on for good (:1048-1052).
void enable_lateral(vh_model* h, const double* pert) {
    if (h->kind != 2) throw vh_error("solid_earth_enable_visc_3d: lateral viscosity needs earth_response=ve");
    const vh_dims& d = h->d;
    const int ne = d.ne;
    const size_t ng = (size_t)d.nphi * d.nlat;
    tensor_init(h);
    std::vector<double> rate3(ng * ne);
    double r3max = 0.0;
    for (int e = 0; e < ne; e++)
        for (size_t i = 0; i < ng; i++) {
            double v = h->MkPerDt[e] * std::pow(10.0, -pert[(size_t)e * ng + i]);
            rate3[(size_t)e * ng + i] = v;
            r3max = std::max(r3max, v);
        }
    std::vector<int> e3;
    for (int e = 0; e < ne; e++) {
        if (h->MkPerDt[e] == 0.0) continue;
        double mx = -INFINITY, mn = INFINITY;
        for (size_t i = 0; i < ng; i++) { double v = pert[(size_t)e * ng + i]; mx = std::max(mx, v); mn = std::min(mn, v); }
        if (mx - mn > h->visc3d_tol) {
            e3.push_back(e);
        } else {
            double sum = 0.0;   // phi fastest, then latitude
            for (size_t i = 0; i < ng; i++) sum += pert[(size_t)e * ng + i];
            h->MkPerDt[e] = h->MkPerDt[e] * std::pow(10.0, -(sum / (double)ng));
        }
    }
    // Maxwell-element index of each 3-D element; per-element flags
    std::vector<int> e3em(e3.size()), e3of(std::max(d.nem, 1), -1);
    std::vector<unsigned char> is3d(std::max(d.nem, 1), 0);
    for (size_t i = 0; i < e3.size(); i++) {
        int em = (int)(std::find(h->emap.begin(), h->emap.end(), e3[i]) - h->emap.begin());
        e3em[i] = em; e3of[em] = (int)i; is3d[em] = 1;
    }
    hipStream_t st = h->st;
    h->sync();
    h->drop_graphs();   // the tensor buffers are replaced below
    for (void* p : {(void*)h->d_e3em, (void*)h->d_e3of, (void*)h->d_rate3, (void*)h->d_X, (void*)h->d_Fd,
                    (void*)h->d_cdel})
        if (p) HIPCHECK(hipFree(p));
    h->ne3d = (int)e3.size();
    h->e3mesh = e3;
    h->nbat = 3 * h->ne3d;
    h->tchunk = std::min(h->nbat, VH_TCHUNK);
    h->d_e3em = dalloc<int>(e3em.size()); upload(h->d_e3em, e3em.data(), e3em.size(), st);
    h->d_e3of = dalloc<int>(e3of.size()); upload(h->d_e3of, e3of.data(), e3of.size(), st);
    upload(h->d_is3d, is3d.data(), (size_t)d.nem, st);
    h->d_rate3 = dalloc<double>(ng * h->ne3d);
    for (int i = 0; i < h->ne3d; i++) upload(h->d_rate3 + (size_t)i * ng, rate3.data() + (size_t)e3[i] * ng, ng, st);
    if (h->tor_req && h->ne3d > 0 && !h->tor_on) enable_toroidal(h);
    h->d_X = dalloc<double2>((size_t)h->nbat * h->nch * d.nlm);
    h->d_cdel = dalloc<double2>((size_t)h->nbat * h->nch * d.nlm);
    h->d_Fd = dalloc<double2>((size_t)std::max(h->tchunk, 1) * 6 * (d.L + 1) * d.nlat);
    upload(h->d_MkPerDt, h->MkPerDt.data(), (size_t)ne, st);
    h->rate_max = r3max;
    for (int e = 0; e < ne; e++) h->rate_max = std::max(h->rate_max, h->MkPerDt[e]);
    h->lat_visc = true;
    vh_t3_rows_prepare(h->s);
    h->sync();
}

// response_enable_lateral_visc_from_nodes (src/vilma_response.f90:1120-1169):
// element value = mean of its two nodes' log10 viscosity, capped in the lid.
static void enable_lateral_nodes(vh_model* h, const double* visc_node, double lid_depth, double lid_log10max,
                          double log10_cap) {
    if (h->kind != 2) throw vh_error("solid_earth_enable_visc_3d: lateral viscosity needs earth_response=ve");
    const int ne = h->d.ne, nr = h->d.nr;
    const size_t ng = (size_t)h->d.nphi * h->d.nlat;
    double r_lid = DBL_MAX;
    if (lid_depth > 0.0) r_lid = h->r[nr - 1] - lid_depth;
    std::vector<double> pert(ng * ne, 0.0);
    for (int e = 0; e < ne; e++) {
        if (h->MkPerDt[e] == 0.0) continue;
        double logeta_ref = std::log10(h->mu[e] / h->MkPerDt[e]);
        bool in_lid = h->r[e] >= r_lid - 1.0e-6 * h->r[nr - 1];
        for (size_t sp = 0; sp < ng; sp++) {
            double ea = 0.5 * (visc_node[sp + (size_t)e * ng] + visc_node[sp + (size_t)(e + 1) * ng]);
            if (in_lid && ea > lid_log10max) ea = log10_cap;
            pert[(size_t)e * ng + sp] = ea - logeta_ref;
        }
    }
    enable_lateral(h, pert.data());
}

int vh_enable_visc_3d(vh_model* h, const double* pert_elem) {
    return guarded(h, [&] {
        if (!pert_elem) throw vh_error("enable_lateral_visc: null field");
        const size_t n = (size_t)h->d.ne * h->d.nphi * h->d.nlat;
        std::vector<double> mk = h->MkPerDt;
        enable_lateral(h, pert_elem);
        h->lat_src = 1;
        h->lat_field.assign(pert_elem, pert_elem + n);
        h->lat_mk = mk;
    });
}

int vh_enable_visc_3d_nodes(vh_model* h, const double* visc_node, double lid_depth, double lid_log10max,
                            double log10_cap) {
    return guarded(h, [&] {
        if (!visc_node) throw vh_error("enable_lateral_visc_from_nodes: null field");
        const size_t n = (size_t)h->d.nr * h->d.nphi * h->d.nlat;
        enable_lateral_nodes(h, visc_node, lid_depth, lid_log10max, log10_cap);
        h->lat_src = 2;
        h->lat_field.assign(visc_node, visc_node + n);
        h->lat_lid[0] = lid_depth; h->lat_lid[1] = lid_log10max; h->lat_lid[2] = log10_cap;
    });
}

// relax_hold (src/vilma_coupling.f90:519-571): hold ice and relax to
// bed-stationary convergence in passes of growing length (SPINUP_DT0 = 200 yr,
// x SPINUP_GROW = 1.6, at most SPINUP_MAXP = 60 passes, :65-67), each a
// coupling call; exit when the mean bed velocity |dz_bed| / pass < tol, or at
// t_cap years (t_cap < 0: no cap) with a warning. One line per pass, as the
// reference logs.
static void relax_hold(vh_model* h, const double* ice, double t_cap, const char* label, double tol) {
    const double dt0 = 200.0, grow = 1.6;
    const int maxp = 60;
    const size_t ng = h->ng();
    hipStream_t st = h->st;
    double* z_prev = dalloc<double>(ng);
    std::vector<double> diff(ng);
    double pass_dt = dt0, t_done = 0.0, rrate = DBL_MAX;
    bool converged = false;
    try {
        HIPCHECK(hipMemcpyAsync(z_prev, h->d_zbed, ng * sizeof(double), hipMemcpyDeviceToDevice, st));
        if (t_cap > 0.0) printf(" spin-up [%s]: relax LGM memory vs reference, cap %.0f yr\n", label, t_cap);
        else printf(" spin-up [%s]: relax LGM memory vs reference (1-D pre-spin)\n", label);
        for (int it = 1; it <= maxp; it++) {
            if (t_cap > 0.0) pass_dt = std::min(pass_dt, t_cap - t_done);
            if (pass_dt <= 0.0) break;
            const long nsub0 = h->n_accept;
            update(h, ice, pass_dt, false);
            t_done += pass_dt;
            const long nsub = std::max(1L, h->n_accept - nsub0);
            const double mean_dt = pass_dt / (double)nsub;
            grid_absdiff(st, (int)ng, h->d_zbed, z_prev, h->d_gtmp);
            vh_surface_integral(st, h->s, h->d_gtmp, h->d_part, h->d_sc + 14);
            HIPCHECK(hipMemcpyAsync(diff.data(), h->d_gtmp, ng * sizeof(double), hipMemcpyDeviceToHost, st));
            h->read(h->d_sc + 14, 1);
            const double rmean = h->h_pin[0] / (4.0 * VH_PI);
            double rmax = 0.0;
            for (double v : diff) rmax = std::max(rmax, v);
            rrate = rmean / pass_dt;
            printf("   pass %3d:  <dt>=%9.1f yr (%ld sub)  d|z_bed|=%9.4f m  max=%9.2f m  v=%9.2e m/yr  v/tol=%8.1e\n",
                   it, mean_dt, nsub, rmean, rmax, rrate, rrate / tol);
            if (rrate < tol) { converged = true; break; }
            HIPCHECK(hipMemcpyAsync(z_prev, h->d_zbed, ng * sizeof(double), hipMemcpyDeviceToDevice, st));
            pass_dt *= grow;
        }
        if (t_cap > 0.0 && !converged)
            printf(" WARNING: spin-up [%s] hit equil_time_max=%.1f yr before convergence (v=%9.2e m/yr, %8.1ex tol)\n",
                   label, t_done, rrate, rrate / tol);
        fflush(stdout);
        h->sync();
    } catch (...) {
        (void)hipFree(z_prev);
        throw;
    }
    (void)hipFree(z_prev);
}

// solid_earth_spinup (src/vilma_coupling.f90:453-517): seed the entering ice
// with a zero-length call, then (pre_spinup_1d) relax the model with its
// lateral viscosity replaced by the lateral mean (a 1-D model on the mean
// profile; the memory carries over) and restore the real field, then relax
// the full model up to equil_time_max years.
static void spinup(vh_model* h, const double* ice, double t_max, double tol, bool pre_1d) {
    if (!pre_1d && t_max <= 0.0) return;
    if (tol <= 0.0) throw vh_error("solid_earth_spinup: equil_rate_tol must be positive");
    update(h, ice, 0.0, false);
    if (pre_1d) {
        const size_t ng = h->ng();
        if (h->lat_src == 2) {   // lateral mean of log10(eta) per node
            const int nr = h->d.nr;
            std::vector<double> unif((size_t)nr * ng);
            for (int r = 0; r < nr; r++) {
                double sum = 0.0;
                for (size_t i = 0; i < ng; i++) sum += h->lat_field[(size_t)r * ng + i];
                std::fill(unif.begin() + (size_t)r * ng, unif.begin() + (size_t)(r + 1) * ng, sum / (double)ng);
            }
            enable_lateral_nodes(h, unif.data(), h->lat_lid[0], h->lat_lid[1], h->lat_lid[2]);
        } else if (h->lat_src == 1) {   // lateral mean of the perturbation per element
            const int ne = h->d.ne;
            std::vector<double> unif((size_t)ne * ng);
            for (int e = 0; e < ne; e++) {
                double sum = 0.0;
                for (size_t i = 0; i < ng; i++) sum += h->lat_field[(size_t)e * ng + i];
                std::fill(unif.begin() + (size_t)e * ng, unif.begin() + (size_t)(e + 1) * ng, sum / (double)ng);
            }
            h->MkPerDt = h->lat_mk;
            enable_lateral(h, unif.data());
        }
        relax_hold(h, ice, -1.0, "1-D ", tol);
        if (h->lat_src == 2) {
            enable_lateral_nodes(h, h->lat_field.data(), h->lat_lid[0], h->lat_lid[1], h->lat_lid[2]);
        } else if (h->lat_src == 1) {
            h->MkPerDt = h->lat_mk;
            enable_lateral(h, h->lat_field.data());
        }
    }
    if (t_max > 0.0) relax_hold(h, ice, t_max, "full", tol);
}

int vh_spinup(vh_model* h, const double* h_ice_lgm, double equil_time_max_yr, double equil_rate_tol,
              int pre_spinup_1d) {
    return guarded(h, [&] {
        if (!h_ice_lgm) throw vh_error("solid_earth_spinup: null ice field");
        spinup(h, h_ice_lgm, equil_time_max_yr, equil_rate_tol, pre_spinup_1d != 0);
    });
}

// tensor_sh_synth / tensor_sh_analysis (src/vilma_tensor_sh.f90:168-322):
// c is [4][nlm] complex in SHTns LM order (lambda = 1, 2, 5, 6), dyad is
// [6][nlat][nphi] (rr, rt, rp, tt, tp, pp).
// n4 on first use by the six-channel transforms (the model calibrates it
// itself when the toroidal channels are carried)
static void tensor_init_n4(vh_model* h) {
    if (h->d_n4) return;
    h->d_n4 = dalloc<double>(h->d.L + 1);
    vh_tsh_calibrate_n4(h->st, h->s, h->d_n4);
}

int vh_tensor_synthesis_ch(vh_model* h, int nch, const double* c, double* dyad) {
    return guarded(h, [&] {
        if (nch != 4 && nch != 6) throw vh_error("tensor_sh_synth: nch must be 4 or 6");
        tensor_init(h);
        const vh_dims& d = h->d;
        const size_t ng = (size_t)d.nphi * d.nlat;
        std::vector<double2> X((size_t)nch * d.nlm);
        for (int lam = 0; lam < nch; lam++) from_lm(h, c + (size_t)2 * lam * d.nlm, X.data() + (size_t)lam * d.nlm);
        double2 *dX = dalloc<double2>(X.size()), *dF = dalloc<double2>((size_t)6 * (d.L + 1) * d.nlat);
        double* dg = dalloc<double>(6 * ng);
        upload(dX, X.data(), X.size(), h->st);
        vh_tsh_synth(h->st, h->s, dX, 1, dF, nch);
        vh_t3_rows(h->st, h->s, dF, 1, 1, nullptr, 1, 0.0, dg);
        HIPCHECK(hipMemcpyAsync(dyad, dg, 6 * ng * sizeof(double), hipMemcpyDeviceToHost, h->st));
        h->sync();
        (void)hipFree(dX); (void)hipFree(dF); (void)hipFree(dg);
    });
}

int vh_tensor_analysis_ch(vh_model* h, int nch, const double* dyad, double* c) {
    return guarded(h, [&] {
        if (nch != 4 && nch != 6) throw vh_error("tensor_sh_analysis: nch must be 4 or 6");
        tensor_init(h);
        if (nch == 6) tensor_init_n4(h);
        const vh_dims& d = h->d;
        const size_t ng = (size_t)d.nphi * d.nlat;
        double2 *dC = dalloc<double2>((size_t)nch * d.nlm), *dF = dalloc<double2>((size_t)6 * (d.L + 1) * d.nlat);
        double* dg = dalloc<double>(6 * ng);
        upload(dg, dyad, 6 * ng, h->st);
        vh_t3_rows(h->st, h->s, dF, 1, 2, nullptr, 1, 0.0, dg);
        vh_tsh_analysis(h->st, h->s, dF, 1, h->d_n6, dC, nch, h->d_n4);
        std::vector<double2> C((size_t)nch * d.nlm);
        HIPCHECK(hipMemcpyAsync(C.data(), dC, C.size() * sizeof(double2), hipMemcpyDeviceToHost, h->st));
        h->sync();
        for (int lam = 0; lam < nch; lam++) to_lm(h, C.data() + (size_t)lam * d.nlm, c + (size_t)2 * lam * d.nlm);
        (void)hipFree(dC); (void)hipFree(dF); (void)hipFree(dg);
    });
}

int vh_tensor_synthesis(vh_model* h, const double* c, double* dyad) { return vh_tensor_synthesis_ch(h, 4, c, dyad); }
int vh_tensor_analysis(vh_model* h, const double* dyad, double* c) { return vh_tensor_analysis_ch(h, 4, dyad, c); }

// n6(l) and n4(l), l = 0..L (the spin-2 normalisations of the tensor analysis)
int vh_tensor_norms(vh_model* h, double* n6, double* n4) {
    return guarded(h, [&] {
        tensor_init(h);
        tensor_init_n4(h);
        if (n6) HIPCHECK(hipMemcpyAsync(n6, h->d_n6, (h->d.L + 1) * sizeof(double), hipMemcpyDeviceToHost, h->st));
        if (n4) HIPCHECK(hipMemcpyAsync(n4, h->d_n4, (h->d.L + 1) * sizeof(double), hipMemcpyDeviceToHost, h->st));
        h->sync();
    });
}

// The toroidal W drift of the last begin_step at every mesh node
// (dWn_re / dWn_im of the reference, zero at nodes that carry no memory):
// [2][nk][nr]; zero unless the toroidal channels are carried.
int vh_get_drift_tor(vh_model* h, double* dwn) {
    return guarded(h, [&] {
        const vh_dims& d = h->d;
        const size_t nn = (size_t)d.nr * d.nk;
        memset(dwn, 0, 2 * nn * sizeof(double));
        if (!h->tor_on) return;
        std::vector<double> dw((size_t)2 * d.nmn * d.nk);
        HIPCHECK(hipMemcpyAsync(dw.data(), h->d_dWt, dw.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
        h->sync();
        for (int q = 0; q < d.nmn; q++) {
            const int node = h->mnodes[q];
            for (int ri = 0; ri < 2; ri++)
                for (int k = 0; k < d.nk; k++) dwn[ri * nn + (size_t)k * d.nr + node] = dw[((size_t)q * 2 + ri) * d.nk + k];
        }
    });
}

// ---- probes
int vh_resp_begin_step(vh_model* h) {
    return guarded(h, [&] {
        if (h->kind != 2) return;
        vh_drift_run(h->st, h->dplan, h->d, h->v, h->d_tau[h->cur], h->d_mnorm[h->cur], h->d_maxn + h->cur,
                     h->skip_tol, h->deg1_cm, h->d_dN, h->d_dsurf);
        tor_drift(h, h->cur, h->d_dWt);
        h->sync();
    });
}

int vh_resp_apply(vh_model* h, const double* sigma_lm, double* u_lm, double* n_lm) {
    return guarded(h, [&] {
        std::vector<double2> sg(h->d.nlm), uN(2 * (size_t)h->d.nlm);
        from_lm(h, sigma_lm, sg.data());
        upload(h->d_lm, sg.data(), h->d.nlm, h->st);
        vh_apply_only(h->st, h->d.nlm, h->d_lm, h->d_gu, h->d_gn, h->d_dsurf, h->d.nk, h->g, h->kind, h->d_uN);
        HIPCHECK(hipMemcpyAsync(uN.data(), h->d_uN, uN.size() * sizeof(double2), hipMemcpyDeviceToHost, h->st));
        h->sync();
        to_lm(h, uN.data(), u_lm);
        to_lm(h, uN.data() + h->d.nlm, n_lm);
    });
}

int vh_resp_advance(vh_model* h, const double* sigma_lm) {
    return guarded(h, [&] {
        if (h->kind != 2) return;
        std::vector<double2> sg(h->d.nlm);
        from_lm(h, sigma_lm, sg.data());
        upload(h->d_lm, sg.data(), h->d.nlm, h->st);
        int nxt = h->cur ^ 1;
        advance_memory(h, (const double*)h->d_lm, nxt);
        h->read(h->d_maxn + nxt, 1);
        h->maxn_cur = h->h_pin[0];
        h->cur = nxt;
        h->time_resp += h->dt;
    });
}

int vh_set_tuning(vh_model* h, int which, int value) {
    return guarded(h, [&] {
        switch (which) {
            case VH_T_SLE_BLOCKS:
                if (value < 0) throw vh_error("vh_set_tuning: VH_T_SLE_BLOCKS must be >= 0");
                h->sle_blocks = value;
                h->drop_graphs();
                break;
            case VH_T_GRAPH:
                if (value != 0 && value != 1) throw vh_error("vh_set_tuning: VH_T_GRAPH must be 0 or 1");
                h->use_graph = value == 1;
                if (!h->use_graph) h->drop_graphs();
                break;
            default: throw vh_error("vh_set_tuning: unknown knob");
        }
    });
}

int vh_get_tuning(vh_model* h, int which, int* value) {
    return guarded(h, [&] {
        switch (which) {
            case VH_T_SLE_BLOCKS: *value = h->sle_blocks; break;
            case VH_T_GRAPH: *value = h->use_graph ? 1 : 0; break;
            default: throw vh_error("vh_get_tuning: unknown knob");
        }
    });
}

int vh_profile(vh_model* h, int enable) {
    return guarded(h, [&] {
        h->sync();
        h->prof_collect();
        h->prof = enable != 0;
        if (enable == 1)
            for (int c = 0; c < VH_K_NCLASS; c++) { h->prof_ms[c] = 0.0; h->prof_n[c] = 0; }
    });
}

int vh_profile_read(vh_model* h, double* ms, double* n) {
    return guarded(h, [&] {
        h->sync();
        h->prof_collect();
        for (int c = 0; c < VH_K_NCLASS; c++) { ms[c] = h->prof_ms[c]; n[c] = (double)h->prof_n[c]; }
    });
}

int vh_resp_set_dt(vh_model* h, double dt_s) {
    return guarded(h, [&] { h->dt = dt_s; });
}

int vh_get_drift(vh_model* h, double* dsurf, double* dnodal) {
    return guarded(h, [&] {
        if (h->kind != 2) throw vh_error("vh_get_drift: no viscoelastic memory");
        const vh_dims& d = h->d;
        if (dsurf) {
            HIPCHECK(hipMemcpyAsync(dsurf, h->d_dsurf, (size_t)6 * d.nk * sizeof(double), hipMemcpyDeviceToHost, h->st));
        }
        if (dnodal) {
            std::vector<double> dn((size_t)4 * d.nmn * d.nk);
            HIPCHECK(hipMemcpyAsync(dn.data(), h->d_dN, dn.size() * sizeof(double), hipMemcpyDeviceToHost, h->st));
            h->sync();
            size_t nn = (size_t)d.nr * d.nk;
            memset(dnodal, 0, 4 * nn * sizeof(double));
            for (int q = 0; q < d.nmn; q++) {
                int node = h->mnodes[q];
                for (int c = 0; c < 4; c++)
                    for (int k = 0; k < d.nk; k++) dnodal[c * nn + (size_t)k * d.nr + node] = dn[((size_t)q * 4 + c) * d.nk + k];
            }
        }
        h->sync();
    });
}

int vh_get_gains(vh_model* h, double* g3, double* xn) {
    return guarded(h, [&] {
        const vh_dims& d = h->d;
        int L = d.L;
        if (g3) {
            HIPCHECK(hipMemcpyAsync(g3, h->d_gu, (L + 1) * sizeof(double), hipMemcpyDeviceToHost, h->st));
            HIPCHECK(hipMemcpyAsync(g3 + L + 1, h->d_gn, (L + 1) * sizeof(double), hipMemcpyDeviceToHost, h->st));
            HIPCHECK(hipMemcpyAsync(g3 + 2 * (L + 1), h->d_gv, (L + 1) * sizeof(double), hipMemcpyDeviceToHost, h->st));
        }
        if (xn) {
            HIPCHECK(hipMemcpyAsync(xn, h->d_xUn, (size_t)L * d.nr * sizeof(double), hipMemcpyDeviceToHost, h->st));
            HIPCHECK(hipMemcpyAsync(xn + (size_t)L * d.nr, h->d_xVn, (size_t)L * d.nr * sizeof(double),
                                    hipMemcpyDeviceToHost, h->st));
        }
        h->sync();
    });
}

// reverse mode (no counterpart in the reference: gradients of an rsl misfit)
int vh_forward_record(vh_model* h, int ncalls, const double* h_ice_hist, double dt_yr, int ckpt_every,
                      double* rsl_out) {
    return guarded(h, [&] { forward_record(h, ncalls, h_ice_hist, false, dt_yr, ckpt_every, rsl_out, false); });
}
int vh_forward_record_device(vh_model* h, int ncalls, const double* h_ice_hist_device, double dt_yr, int ckpt_every,
                             double* rsl_out_device) {
    return guarded(h, [&] {
        forward_record(h, ncalls, h_ice_hist_device, true, dt_yr, ckpt_every, rsl_out_device, true);
    });
}
int vh_backward(vh_model* h, const double* dJ_drsl, double* dJ_dlogeta, double* dJ_dice) {
    return guarded(h, [&] { backward(h, dJ_drsl, false, dJ_dlogeta, dJ_dice, false); });
}
int vh_backward_device(vh_model* h, const double* dJ_drsl_device, double* dJ_dlogeta_device, double* dJ_dice_device) {
    return guarded(h, [&] { backward(h, dJ_drsl_device, true, dJ_dlogeta_device, dJ_dice_device, true); });
}
int vh_adjoint_store_limit(vh_model* h, double gib) {
    return guarded(h, [&] {
        if (!(gib >= 0.0)) throw vh_error("vh_adjoint_store_limit: the limit must be >= 0 GiB");
        if (!h->adj) h->adj = new adj_ctx();
        h->adj->store_gib = gib;
    });
}
int vh_forward_rewind(vh_model* h) {
    return guarded(h, [&] {
        if (!h->adj || !h->adj->recorded) throw vh_error("vh_forward_rewind: no forward record");
        adj_restore(h, h->adj->ck[0]);
        h->sync();
    });
}

int vh_sht_analysis(vh_model* h, const double* grid, double* lm) {
    return guarded(h, [&] {
        upload(h->d_gtmp, grid, h->ng(), h->st);
        vh_fft_forward(h->st, h->s, h->d_gtmp, h->d.nlat, h->d_F);
        vh_legendre_analysis(h->st, h->s, h->d_F, 1, h->d_lm);
        std::vector<double2> t(h->d.nlm);
        HIPCHECK(hipMemcpyAsync(t.data(), h->d_lm, t.size() * sizeof(double2), hipMemcpyDeviceToHost, h->st));
        h->sync();
        to_lm(h, t.data(), lm);
    });
}

int vh_sht_synthesis(vh_model* h, const double* lm, double* grid) {
    return guarded(h, [&] {
        std::vector<double2> t(h->d.nlm);
        from_lm(h, lm, t.data());
        upload(h->d_lm, t.data(), h->d.nlm, h->st);
        vh_legendre_synthesis(h->st, h->s, h->d_lm, 1, h->d_F);
        vh_fft_inverse(h->st, h->s, h->d_F, h->d.nlat, h->d_gtmp);
        HIPCHECK(hipMemcpyAsync(grid, h->d_gtmp, h->ng() * sizeof(double), hipMemcpyDeviceToHost, h->st));
        h->sync();
    });
}

}  // extern "C"
