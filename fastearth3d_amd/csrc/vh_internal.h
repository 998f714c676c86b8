// vh_internal.h -- device-side layout and kernel entry points of the
// MI355X-native viscoelastic time step (shared by the .hip translation units).
//
// Layout in HBM (all FP64):
//   * Maxwell memory tau[em][c][k]: only Maxwell elements em carry memory
//     (elastic lithosphere and fluid core never do, src/vilma_response.f90:
//     571-586); c = (shape*2 + reim)*4 + lambda with shape A,B,C and lambda
//     the four spheroidal channels 1,2,5,6; the slot index k (degree-grouped,
//     l ascending then m ascending, src/vilma_response.f90:629-648) is
//     contiguous, so every per-slot kernel reads coalesced rows of nk values.
//     Two copies (ping-pong): the advance writes tau_{n+1} into the other
//     buffer, so the step controller's rollback needs no copy.
//   * Spectra are complex in degree-major order idx(l,m) = l(l+1)/2 + m,
//     whose l >= 1 part is exactly the slot order (k = idx - 1).
//   * Grid fields are (nlat, nphi) with phi fastest, north-first rows.
//   * Per-degree band LU factors (l >= 2): kl = ku = 6 (node-interleaved P1/P0
//     dofs, src/vilma_radial_fe.f90:226-238), ldab = 19, column-major.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vh_tuned.h"
#include <stdexcept>
#include <string>
#include <vector>

struct vh_error : std::runtime_error { using std::runtime_error::runtime_error; };

// max that propagates NaN (fmax returns the non-NaN operand): the memory norm
// feeds the stability guard, whose finite test must see a NaN memory value
// (src/vilma_timestep.f90:195: "finite = err_inf <= huge", false for NaN)
__host__ __device__ __forceinline__ double vh_nanmax(double a, double b) { return (b > a || b != b) ? b : a; }

#define VH_KL 6
#define VH_KV 12
#define VH_LDAB 19
#define VH_PI 3.14159265358979323846264338327950288
#define VH_G 6.67430e-11
#define VH_RHO_ICE 931.0
#define VH_RHO_WATER 1000.0
#define VH_SEC_PER_YEAR 31556926.0

#define HIPCHECK(x)                                                             \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) throw vh_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct vh_dims {
    int L, nlm, nr, ne, nd, nk, nem, nmn, nphi, nlat, nh;
    int ns1, ldab1, kl1;   // degree-1 bordered system
};

// Everything a kernel needs about the radial discretisation and the memory.
struct vh_dev {
    // mesh / elements
    const double* r;        // [nr]
    const double* mu;       // [ne]
    const double* MkPerDt;  // [ne]   Maxwell rate mu/eta of the 1-D path (M = rate*dt)
    const int* emap;        // [nem]  Maxwell element -> mesh element
    const int* mnodes;      // [nmn]  memory node -> mesh node
    const int* kdeg;        // [nk]   degree of slot k
    const int* mnode;       // [nr]   mesh node -> memory-node index or -1
    const unsigned char* is3d;   // [nem] genuinely 3-D (advanced pseudo-spectrally)
    // operators
    const double* ab;       // [(L-1)][19*nd]  l >= 2
    const int* ipiv;        // [(L-1)][nd]
    const double* dr;       // [L][nd]  row equilibration, index l-1
    const double* dc;       // [L][nd]
    const double* ab1;      // degree 1, [ldab1*ns1]
    const int* ipiv1;       // [ns1]
    const double* null1;    // [nd]
    const double* gu;       // [L+1]
    const double* gn;
    const double* gv;
    const double* xUn;      // [L][nr], index l-1
    const double* xVn;
};

// ---------------------------------------------------------------- setup
void vh_launch_setup_band(hipStream_t st, const vh_dims& d, int lmin, int lmax, const double* r,
                          const double* mu, const double* rho, const double* Rk, double r_earth,
                          double g_surf, int rhs_kind, double* ab, int* ipiv, double* dr, double* dc,
                          double* xfull /* [njob][nd] or null */, double* xUn, double* xVn,
                          double* gu, double* gn, double* gv, int* fail);
void vh_launch_setup_deg1(hipStream_t st, const vh_dims& d, const double* r, const double* mu,
                          const double* rho, const double* Rk, double r_earth, double g_surf,
                          int deg1_cm, double* ab1, int* ipiv1, double* dr1, double* dc1,
                          double* null1, double* xUn, double* xVn, double* gu, double* gn,
                          double* gv, double* work, int* fail);
// degree-1 restricted solution operator G1 [2 nmn + 3][2 nmn] (work: ns1 doubles)
void vh_launch_setup_deg1_op(hipStream_t st, const vh_dims& d, const int* mnodes, const double* ab1,
                             const int* ipiv1, const double* dr1, const double* dc1, const double* null1,
                             int deg1_cm, double* work, double* G1);

// ---------------------------------------------------------------- (b) advance
// 3-D elements (v.is3d) take tau_in + their increment from cdel
// ([i3*3 + shape][lambda][nlm] complex, e3of[em] = i3) instead of the 1-D update.
// scratch: vh_advance_scratch(d) doubles (per-block maxima)
struct vh_tor_plan;
// The toroidal channels of one advance (all null when they are not carried):
// taut_in/out [em][12][k], the nodal W drift dW [nmn][2][k] (the endpoint
// drift under the trapezoidal rule, dW0 then the start-of-step one).
struct vh_adv_tor {
    const vh_tor_plan* plan = nullptr;
    const double* taut_in = nullptr;
    double* taut_out = nullptr;
    const double* dW = nullptr;
    const double* dW0 = nullptr;
};
void vh_launch_advance(hipStream_t st, const vh_dims& d, const vh_dev& v, double dt, const double* tau_in,
                       double* tau_out, const double* sigma, const double* dN, double* mnorm_out,
                       double* maxnorm_out, const int* e3of, const double2* cdel, double* scratch,
                       double* maxn_copy = nullptr,   // maxn_copy: a second place for the global max
                       // trapezoidal rule (sigma_n non-null): dN is the endpoint drift and
                       // (sigma_n, dN0) give the start-of-step strain
                       const double* sigma_n = nullptr, const double* dN0 = nullptr,
                       // channels per tensor batch of cdel (4, or 6 with the toroidal ones)
                       int nch = 4, const vh_adv_tor& tor = vh_adv_tor());
// scratch doubles of the advance's reductions (tor: with the toroidal rows)
size_t vh_advance_scratch(const vh_dims& d, bool tor = false);
// out2[0] = max |fine - coarse|, out2[1] = max |fine| over n values (NaN kept)
// accumulate: combine with what out2 already holds instead of starting at zero
void vh_coarse_fine(hipStream_t st, size_t n, const double* fine, const double* coarse, double* out2,
                    bool accumulate = false);
// out2[0] = max |dUa - prev|, out2[1] = max |dUa| over the nk complex dUa at
// the head of dsurf; then prev <- dUa
// (dWa, prev_t non-null: the toroidal surface drift joins both maxima and is kept likewise)
void vh_drift_change(hipStream_t st, int nk, const double* dsurf, double* prev, double* out2,
                     const double* dWa = nullptr, double* prev_t = nullptr);
// taut (toroidal memory, [em][12][k]) joins the norm when non-null
void vh_launch_memnorm(hipStream_t st, const vh_dims& d, const double* tau, double* mnorm, double* maxnorm,
                       const double* taut = nullptr);

// ---------------------------------------------------------------- SHT
// Factorisation of a longitude FFT of length N (plan of the transform)
struct fft_cplan { int nfac; int fac[8]; int twoff[8]; int ns[8]; };
// the run-time plan's rule: factors 8, 6, 4, 2, 3, 5, 7 in that order of
// preference; twiddles [r-1][k] per pass after the first
__host__ __device__ constexpr fft_cplan fft_make_cplan(int N) {
    fft_cplan c{};
    int n = N, nf = 0;
    const int pref[7] = {8, 6, 4, 2, 3, 5, 7};
    for (int q = 0; q < 7 && n > 1; q++)
        while (n % pref[q] == 0 && nf < 8) { c.fac[nf++] = pref[q]; n /= pref[q]; }
    c.nfac = n == 1 ? nf : 0;
    int off = 0, Ns = 1;
    for (int p = 0; p < nf; p++) {
        c.twoff[p] = off;
        c.ns[p] = Ns;
        if (Ns > 1) off += (c.fac[p] - 1) * Ns;
        Ns *= c.fac[p];
    }
    return c;
}

struct vh_sht {
    int L, nlat, nphi, nh, nlm;
    const double* plm_lj;  // [m][l][j], per-m offset mstart[m]*nh (synthesis)
    const double* dplm_lj;
    const double* plm_jl;  // [m][j][l], same per-m offset (analysis)
    const double* dplm_jl;
    const int* mstart;     // [L+1] (device)
    const double* w;       // [nlat]
    const double* st;      // [nlat] sin(colat)
    const double* ct;      // [nlat]
    const double2* tw;     // [nphi] e^{-2 pi i t/nphi}
    const double2* twp;    // per-pass FFT twiddles, pass p at twoff[p]: [R-1][Ns] (Ns = product of earlier radices)
    int nfac, fac[16], twoff[16], ntw;
    unsigned nsm[16];      // per pass: ceil(2^32 / Ns) (Ns > 1), so j / Ns = umulhi(j, nsm) exactly for j, Ns < 2^16
    const int2* ljobs;     // Legendre analysis blocks: (m, first degree offset), 64 degrees each
    int nljobs;
};
void vh_setup_legendre_tables(hipStream_t st, const vh_sht& s, const double* ct, const double* sn,
                              double* plm_lj, double* dplm_lj, double* plm_jl, double* dplm_jl);
// rows of real grid -> row spectra F[row][m], m = 0..L, scaled by 2 pi/nphi
void vh_fft_forward(hipStream_t st, const vh_sht& s, const double* f, int nrows, double2* F);
void vh_fft_inverse(hipStream_t st, const vh_sht& s, const double2* F, int nrows, double* f);
// batch of nb fields: F [nb][nlat][L+1] <-> q [nb][nlm] (degree-major)
void vh_legendre_analysis(hipStream_t st, const vh_sht& s, const double2* F, int nb, double2* q);
void vh_legendre_synthesis(hipStream_t st, const vh_sht& s, const double2* q, int nb, double2* F);
// synthesis into rows f*fst + j*jst (fst = 1, jst = nb interleaves the fields per latitude)
void vh_legendre_synthesis_rows(hipStream_t st, const vh_sht& s, const double2* q, int nb, double2* F,
                                int fst, int jst);
// spheroidal pair (vt, vp) row spectra <-> q, nb fields
void vh_sph_synthesis_F(hipStream_t st, const vh_sht& s, const double2* q, int nb, double2* FT, double2* FP);
void vh_sph_analysis_F(hipStream_t st, const vh_sht& s, const double2* FT, const double2* FP, int nb, double2* q);

// ---------------------------------------------------------------- reductions
// out[0] = sum_j w_j * (2 pi/nphi) * sum_i f[j][i] (deterministic two-pass)
void vh_surface_integral(hipStream_t st, const vh_sht& s, const double* f, double* part, double* out);
// out[c] = sum_j w_j * dlon * part[c*nlat + j], c < ncomp (ncomp <= 64)
void vh_weighted_total(hipStream_t st, const vh_sht& s, const double* part, int ncomp, double* out);
void vh_fill(hipStream_t st, double* p, double v, size_t n);

// ---------------------------------------------------------------- tensor (3-D)
// X [nb][nch][nlm] tensor coefficients (nch = 4: lambda 1, 2, 5, 6; nch = 6:
// also lambda 3, 4 as channels 4, 5) -> Fd [nb][6][L+1][nlat] dyad row spectra
void vh_tsh_synth(hipStream_t st, const vh_sht& s, const double2* X, int nb, double2* Fd, int nch = 4);
// Fd -> C [nb][nch][nlm]; n6 (n4) == null leaves the raw spin-2 adjoint in
// channel 3 (5)
void vh_tsh_analysis(hipStream_t st, const vh_sht& s, const double2* Fd, int nb, const double* n6, double2* C,
                     int nch = 4, const double* n4 = nullptr);
// mode 0: Fd <- FFT(-rate*dt * IFFT(Fin)), rate [nb/bpe][nlat][nphi] (Fin null: Fd, in place);
// mode 1: grid [nb][6][nlat][nphi] <- IFFT(Fin); mode 2: Fd <- FFT(grid);
// mode 3: as mode 0 with the trapezoidal weight M / (1 + M/2), M = rate*dt
void vh_t3_rows(hipStream_t st, const vh_sht& s, double2* Fd, int nb, int mode, const double* rate, int bpe,
                double dt, double* grid, const double2* Fin = nullptr);
// the row pass's launch attributes for this grid, set outside any stream capture
void vh_t3_rows_prepare(const vh_sht& s);
// X [3 ne3d][nch][nlm]; with nch = 6 the toroidal channels come from taut
// and the nodal W drift dWt ([nmn][2][nk])
void vh_t3_gather(hipStream_t st, const vh_dims& d, const vh_dev& v, const int* e3em, int ne3d, const double* tau,
                  const double* sigma, const double* dN, double2* X, int nch = 4, const double* taut = nullptr,
                  const double* dWt = nullptr);
// trapezoidal rule: X = tau_n + mu (eps_n + eps_{n+1}), eps_n from
// (sigma_n, dN0 / dWt0), eps_{n+1} from (sigma, dN1 / dWt1)
void vh_t3_gather_trap(hipStream_t st, const vh_dims& d, const vh_dev& v, const int* e3em, int ne3d,
                       const double* tau, const double* sigma_n, const double* dN0, const double* sigma,
                       const double* dN1, double2* X, int nch = 4, const double* taut = nullptr,
                       const double* dWt0 = nullptr, const double* dWt1 = nullptr);
void vh_tsh_calibrate_n6(hipStream_t st, const vh_sht& s, double* n6_dev);
// n4(l): the raw toroidal spin-2 adjoint of the unit sectoral (l, m = l)
// toroidal spin-2 synthesis (channel 5), the Z4 normalisation
void vh_tsh_calibrate_n4(hipStream_t st, const vh_sht& s, double* n4_dev);

// ---------------------------------------------------------------- toroidal
// Restricted inverses of the per-degree W operators and the forcing tables
// (host set-up, extended precision), built when the toroidal channels are
// first carried.
vh_tor_plan* vh_tor_plan_create(hipStream_t st, const vh_dims& d, const std::vector<double>& r,
                                const std::vector<double>& mu, const std::vector<int>& emap,
                                const std::vector<int>& mnodes, const std::vector<int>& mnode_of);
void vh_tor_plan_destroy(vh_tor_plan* p);
// W drift of the toroidal memory taut: nodal dWt [nmn][2][nk], surface dWa
// [nk] complex; zero where the skip gate is closed
void vh_tor_drift(hipStream_t st, const vh_tor_plan* p, const vh_dims& d, const vh_dev& v, const double* taut,
                  const double* mnorm, const double* maxnorm, double skip_tol, double* dWt, double* dWa);
struct tor_adv_args {
    int nk, nem, nlm, yoff;
    double dt;
    const int* kdeg;
    const int* emap;
    const double* mu;
    const double* MkPerDt;
    const unsigned char* is3d;
    const int* e3of;
    const int2* enode;
    const double2* cdel;   // [i3*3 + shape][6][nlm]
    const double* tau_in;
    double* tau_out;
    const double* dW;
    const double* dW0;     // trapezoidal rule only
    double* pmax;
    double* pblk;
};
void vh_launch_advance_tor(hipStream_t st, const tor_adv_args& a, bool trap);
const int2* vh_tor_plan_enode(const vh_tor_plan* p);

// ---------------------------------------------------------------- drift plan
struct vh_drift_plan;
// Restricted solution operators G_l (l = lmin..lmax, l >= 2) of the band
// factors in v: unit forcing at the U/V dofs of the nmn listed nodes ->
// solution at those dofs and at the surface U, F, V; [l-lmin][2nmn+3][2nmn].
void vh_setup_restricted_ops(hipStream_t st, const vh_dims& d, const vh_dev& v, int lmin, int lmax, int nmn,
                             const int* mnodes, double* G);
// builds the per-degree restricted solution operators from the factors in v
vh_drift_plan* vh_drift_plan_create(hipStream_t st, const vh_dims& d, const vh_dev& v, int deg1_cm,
                                    const std::vector<int>& emap, const std::vector<int>& mnode_of);
void vh_drift_plan_destroy(vh_drift_plan* p);
void vh_drift_run(hipStream_t st, const vh_drift_plan* p, const vh_dims& d, const vh_dev& v,
                  const double* tau, const double* mnorm, const double* maxnorm, double skip_tol,
                  int deg1_cm, double* dN, double* dsurf);
