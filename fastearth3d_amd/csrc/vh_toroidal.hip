// vh_toroidal.hip -- the toroidal degree of freedom (l_toroidal): the
// horizontal displacement W(r) e_r x grad_1 Y_lm that the spheroidal
// response never excites, and that lateral viscosity variations couple in
// through the two toroidal memory channels lambda = 3, 4
// (src/vilma_radial_fe.f90:699-817, src/vilma_viscoelastic.f90:373-439,
// :591-629, src/vilma_response.f90:694-809 tor branch, :966-991,
// :1055-1109).
//
// Memory layout: a second array taut[em][c][k] beside the spheroidal one,
// c = (shape*2 + reim)*2 + q (shape A, B, C; q = 0, 1 for lambda = 3, 4),
// slot k contiguous; two copies (ping-pong) like the spheroidal memory.
// W is pure drift (no load term), so per slot the step needs only
//
//   forcing  f = -int taut : delta eps_W  on the nodal W test functions
//   solve    W = A_l^{-1} f               (the per-degree W operator)
//
// with f living on the memory nodes and W read back at the memory nodes
// (for the strain of the advance) and at the surface (the reported
// toroidal displacement). As for the spheroidal drift, the solve is a
// restricted inverse formed once per degree at set-up,
//
//   Gt_l = (rows: memory nodes + surface) x (columns: memory nodes) of A_l^{-1},
//
// here on the host in extended precision: A_l is tridiagonal and symmetric
// positive definite for l >= 2 (nodes whose elements are all fluid pinned
// to zero); degree 1 carries one net-rotation constraint per solid shell
// (a bordered system, solved densely with partial pivoting after symmetric
// diagonal scaling). The forcing map of one (degree, Maxwell element) is a
// 2 x 6 table (two test nodes x three shapes x two channels), built with
// the reference's 2-point radial Gauss rule.
#include <algorithm>
#include <cmath>
#include <vector>
#include "vh_internal.h"
#include "vh_radial_dev.h"

struct vh_tor_plan {
    int L = 0, nem = 0, nmn = 0;
    double* Gt = nullptr;    // [L][nmn + 1][nmn], index l-1
    double* Kt = nullptr;    // [L][nem][2][6]
    int2* enode = nullptr;   // [nem] memory nodes (lower, upper) of the element
};

namespace {

typedef long double xd;

// W stiffness of one element, mu int [J e3_a e3_b + J (J - 2) e4_a e4_b] r^2 dr
// with e3 = psi' - psi / r, e4 = psi / r for the two linear shape functions
// (the eq-80 W block, src/vilma_radial_fe.f90:699-726); the integrand is a
// quadratic polynomial in r, so two Gauss points integrate it exactly.
void tor_element(double rk, double rk1, double mu, double J, xd K[2][2]) {
    const xd g = 0.57735026918962576450914878050195746L;
    const xd h = (xd)rk1 - (xd)rk;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) K[a][b] = 0.0L;
    for (int ig = 0; ig < 2; ig++) {
        const xd ra = 0.5L * ((xd)rk + (xd)rk1) + 0.5L * h * (ig ? g : -g);
        const xd psi[2] = {((xd)rk1 - ra) / h, (ra - (xd)rk) / h};
        const xd dpsi[2] = {-1.0L / h, 1.0L / h};
        xd e3[2], e4[2];
        for (int a = 0; a < 2; a++) { e3[a] = dpsi[a] - psi[a] / ra; e4[a] = psi[a] / ra; }
        const xd wq = 0.5L * h * ra * ra;
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                K[a][b] += wq * (xd)mu * ((xd)J * e3[a] * e3[b] + (xd)J * ((xd)J - 2.0L) * e4[a] * e4[b]);
    }
}

// int psi_a r^3 dr over one element (degree-4 integrand: three Gauss points)
void tor_k4(double rk, double rk1, xd v[2]) {
    const xd x[3] = {-0.77459666924148337703585307995647992L, 0.0L, 0.77459666924148337703585307995647992L};
    const xd w[3] = {5.0L / 9.0L, 8.0L / 9.0L, 5.0L / 9.0L};
    const xd h = (xd)rk1 - (xd)rk;
    v[0] = v[1] = 0.0L;
    for (int i = 0; i < 3; i++) {
        const xd ra = 0.5L * ((xd)rk + (xd)rk1) + 0.5L * h * x[i];
        const xd r3 = ra * ra * ra * 0.5L * h * w[i];
        v[0] += r3 * ((xd)rk1 - ra) / h;
        v[1] += r3 * (ra - (xd)rk) / h;
    }
}

// toroidal strain coefficients of nodal W (lambda = 3, 4; no lambda-4 part at
// degree 1, src/vilma_viscoelastic.f90:373-393)
void tor_strain_host(double w1, double w2, double J, double a[2], double b[2], double c[2]) {
    a[0] = -w1 + w2; b[0] = -w1; c[0] = -w2;
    a[1] = 0.0; b[1] = w1; c[1] = w2;
    if (J <= 2.0) { b[1] = 0.0; c[1] = 0.0; }
}

// x = M^{-1} b for the columns of B (n x nrhs, column-major), dense Gaussian
// elimination with partial pivoting; M is overwritten. Returns false if singular.
bool dense_solve(int n, std::vector<xd>& M, std::vector<xd>& B, int nrhs) {
    for (int k = 0; k < n; k++) {
        int p = k;
        xd best = std::fabs(M[(size_t)k * n + k]);
        for (int i = k + 1; i < n; i++) {
            const xd v = std::fabs(M[(size_t)k * n + i]);
            if (v > best) { best = v; p = i; }
        }
        if (!(best > 0.0L)) return false;
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(M[(size_t)j * n + k], M[(size_t)j * n + p]);
            for (int c = 0; c < nrhs; c++) std::swap(B[(size_t)c * n + k], B[(size_t)c * n + p]);
        }
        const xd piv = M[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) {
            const xd f = M[(size_t)k * n + i] / piv;
            if (f == 0.0L) continue;
            M[(size_t)k * n + i] = f;
            for (int j = k + 1; j < n; j++) M[(size_t)j * n + i] -= f * M[(size_t)j * n + k];
            for (int c = 0; c < nrhs; c++) B[(size_t)c * n + i] -= f * B[(size_t)c * n + k];
        }
    }
    for (int c = 0; c < nrhs; c++) {
        xd* x = B.data() + (size_t)c * n;
        for (int i = n - 1; i >= 0; i--) {
            xd s = x[i];
            for (int j = i + 1; j < n; j++) s -= M[(size_t)j * n + i] * x[j];
            x[i] = s / M[(size_t)i * n + i];
        }
    }
    return true;
}

}  // namespace

// toroidal_operator_assemble / _solve_vec (src/vilma_radial_fe.f90:728-810)
// restricted to the memory nodes, and the forcing tables.
vh_tor_plan* vh_tor_plan_create(hipStream_t st, const vh_dims& d, const std::vector<double>& r,
                                const std::vector<double>& mu, const std::vector<int>& emap,
                                const std::vector<int>& mnodes, const std::vector<int>& mnode_of) {
    const int nr = d.nr, ne = d.ne, nmn = d.nmn, nem = d.nem, L = d.L;
    if (nmn > 128) throw vh_error("toroidal drift: too many memory nodes for the LDS forcing column (nmn = " +
                                  std::to_string(nmn) + ")");
    std::vector<unsigned char> pinned(nr, 1);
    for (int e = 0; e < ne; e++)
        if (mu[e] > 0.0) { pinned[e] = 0; pinned[e + 1] = 0; }
    // solid shells for the degree-1 border (:742-771)
    std::vector<int> shell(ne, 0);
    int ns = 0;
    for (int e = 0, prev = 0; e < ne; e++) {
        const int solid = mu[e] > 0.0;
        if (solid && !prev) ns++;
        if (solid) shell[e] = ns;
        prev = solid;
    }
    if (ns > 4) throw vh_error("toroidal_operator_assemble: more than 4 solid shells");
    const int nrow = nmn + 1;
    std::vector<double> Gt((size_t)L * nrow * std::max(nmn, 1), 0.0);
    std::vector<xd> dg(nr), od(nr), X((size_t)nr * std::max(nmn, 1));
    for (int l = 1; l <= L; l++) {
        const double J = (double)l * (double)(l + 1);
        std::fill(dg.begin(), dg.end(), 0.0L);
        std::fill(od.begin(), od.end(), 0.0L);   // od[i] = A[i][i+1] = A[i+1][i]
        for (int e = 0; e < ne; e++) {
            if (!(mu[e] > 0.0)) continue;
            xd K[2][2];
            tor_element(r[e], r[e + 1], mu[e], J, K);
            dg[e] += K[0][0]; dg[e + 1] += K[1][1]; od[e] += K[0][1];
        }
        for (int n = 0; n < nr; n++)
            if (pinned[n]) { dg[n] = 1.0L; od[n] = 0.0L; if (n > 0) od[n - 1] = 0.0L; }
        std::fill(X.begin(), X.end(), 0.0L);
        for (int c = 0; c < nmn; c++) X[(size_t)c * nr + mnodes[c]] = 1.0L;
        if (l >= 2) {
            // symmetric positive definite tridiagonal: LDL^T without pivoting
            std::vector<xd> dd(nr), ll(nr, 0.0L);
            dd[0] = dg[0];
            for (int i = 1; i < nr; i++) {
                ll[i] = od[i - 1] / dd[i - 1];
                dd[i] = dg[i] - ll[i] * od[i - 1];
                if (!(dd[i] > 0.0L)) throw vh_error("toroidal_operator_assemble: W operator not positive definite");
            }
            for (int c = 0; c < nmn; c++) {
                xd* x = X.data() + (size_t)c * nr;
                for (int i = 1; i < nr; i++) x[i] -= ll[i] * x[i - 1];
                for (int i = 0; i < nr; i++) x[i] /= dd[i];
                for (int i = nr - 2; i >= 0; i--) x[i] -= ll[i + 1] * x[i + 1];
            }
        } else {
            // [A w; w^T 0] with one net-rotation row per solid shell, scaled
            // symmetrically (A part by its diagonal, border by its largest
            // scaled entry), dense elimination with partial pivoting
            const int n = nr + ns;
            std::vector<xd> M((size_t)n * n, 0.0L), s(n, 1.0L);
            for (int i = 0; i < nr; i++) {
                M[(size_t)i * n + i] = dg[i];
                if (i + 1 < nr) { M[(size_t)(i + 1) * n + i] = od[i]; M[(size_t)i * n + i + 1] = od[i]; }
            }
            for (int e = 0; e < ne; e++) {
                if (!shell[e]) continue;
                xd k4[2];
                tor_k4(r[e], r[e + 1], k4);
                const int b = nr + shell[e] - 1;
                M[(size_t)b * n + e] += k4[0]; M[(size_t)b * n + e + 1] += k4[1];
                M[(size_t)e * n + b] += k4[0]; M[(size_t)(e + 1) * n + b] += k4[1];
            }
            for (int i = 0; i < nr; i++) s[i] = 1.0L / std::sqrt(std::fabs(M[(size_t)i * n + i]));
            for (int b = nr; b < n; b++) {
                xd mx = 0.0L;
                for (int i = 0; i < nr; i++) mx = std::max(mx, std::fabs(M[(size_t)b * n + i]) * s[i]);
                s[b] = mx > 0.0L ? 1.0L / mx : 1.0L;
            }
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) M[(size_t)j * n + i] *= s[i] * s[j];
            std::vector<xd> B((size_t)n * std::max(nmn, 1), 0.0L);
            for (int c = 0; c < nmn; c++) B[(size_t)c * n + mnodes[c]] = s[mnodes[c]];
            if (!dense_solve(n, M, B, nmn)) throw vh_error("bordered_band_factor: band LU factorization failed");
            for (int c = 0; c < nmn; c++)
                for (int i = 0; i < nr; i++) X[(size_t)c * nr + i] = s[i] * B[(size_t)c * n + i];
        }
        double* G = Gt.data() + (size_t)(l - 1) * nrow * nmn;
        for (int c = 0; c < nmn; c++) {
            for (int q = 0; q < nmn; q++) G[(size_t)q * nmn + c] = (double)X[(size_t)c * nr + mnodes[q]];
            G[(size_t)nmn * nmn + c] = (double)X[(size_t)c * nr + nr - 1];
        }
    }
    // forcing tables: floc[t] = sum_j Kt[t][j] v[j], v = (A q0, A q1, B q0,
    // B q1, C q0, C q1) of one re/im block (dissipative_rhs_tor,
    // src/vilma_viscoelastic.f90:407-439: 2-point radial Gauss, norms J/2,
    // J (J - 2)/2 of lambda = 3, 4)
    std::vector<double> Kt((size_t)L * std::max(nem, 1) * 12, 0.0);
    const double xg = 0.5773502691896257;
    for (int l = 1; l <= L; l++) {
        const double J = (double)l * (double)(l + 1);
        const double norm[2] = {0.5 * J, 0.5 * J * (J - 2.0)};
        double sa[2][2], sb[2][2], sc[2][2];
        tor_strain_host(1.0, 0.0, J, sa[0], sb[0], sc[0]);
        tor_strain_host(0.0, 1.0, J, sa[1], sb[1], sc[1]);
        for (int em = 0; em < nem; em++) {
            const int e = emap[em];
            const double rk = r[e], rk1 = r[e + 1], h = rk1 - rk;
            double* K = Kt.data() + ((size_t)(l - 1) * nem + em) * 12;
            for (int ig = 0; ig < 2; ig++) {
                const double ra = 0.5 * (h * (ig ? xg : -xg) + rk + rk1);
                const double psik = (rk1 - ra) / h, psik1 = (ra - rk) / h;
                const double wq = ra * ra * h * 0.5;
                for (int t = 0; t < 2; t++)
                    for (int q = 0; q < 2; q++) {
                        const double deps = sa[t][q] / h + sb[t][q] * psik / ra + sc[t][q] * psik1 / ra;
                        const double c = -norm[q] * deps * wq;
                        K[t * 6 + 0 * 2 + q] += c / h;
                        K[t * 6 + 1 * 2 + q] += c * psik / ra;
                        K[t * 6 + 2 * 2 + q] += c * psik1 / ra;
                    }
            }
        }
    }
    std::vector<int2> en(std::max(nem, 1), make_int2(0, 0));
    for (int em = 0; em < nem; em++) en[em] = make_int2(mnode_of[emap[em]], mnode_of[emap[em] + 1]);
    vh_tor_plan* p = new vh_tor_plan();
    p->L = L; p->nem = nem; p->nmn = nmn;
    HIPCHECK(hipMalloc((void**)&p->Gt, Gt.size() * sizeof(double)));
    HIPCHECK(hipMalloc((void**)&p->Kt, Kt.size() * sizeof(double)));
    HIPCHECK(hipMalloc((void**)&p->enode, en.size() * sizeof(int2)));
    HIPCHECK(hipMemcpyAsync(p->Gt, Gt.data(), Gt.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(p->Kt, Kt.data(), Kt.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(p->enode, en.data(), en.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));   // the host tables go out of scope
    return p;
}

void vh_tor_plan_destroy(vh_tor_plan* p) {
    if (!p) return;
    (void)hipFree(p->Gt); (void)hipFree(p->Kt); (void)hipFree(p->enode);
    delete p;
}

const int2* vh_tor_plan_enode(const vh_tor_plan* p) { return p->enode; }

// ------------------------------------------------------------------ drift
// One thread per slot (64 per block); the slot's forcing column lives in
// LDS as f[node][lane] (no sharing between lanes, so no barriers). The skip
// gate is the spheroidal one (the per-slot norm covers all six channels).
__global__ void __launch_bounds__(64) k_drift_tor(int nk, int nem, int nmn, const int* __restrict__ kdeg,
        const int2* __restrict__ enode, const double* __restrict__ Gt, const double* __restrict__ Kt,
        const double* __restrict__ taut, const double* __restrict__ mnorm, const double* __restrict__ maxnorm,
        double skip_tol, double* __restrict__ dWt, double2* __restrict__ dWa) {
    extern __shared__ double fcol[];   // [nmn][64]
    const int lane = threadIdx.x, k = blockIdx.x * 64 + lane;
    if (k >= nk) return;
    const int l = kdeg[k];
    const bool act = mnorm[k] > skip_tol * maxnorm[0];
    const double* G = Gt + (size_t)(l - 1) * (nmn + 1) * nmn;
    const double* K = Kt + (size_t)(l - 1) * nem * 12;
    double* f = fcol + lane;
    double wa[2] = {0.0, 0.0};
    for (int ri = 0; ri < 2; ri++) {
        for (int q = 0; q < nmn; q++) f[q * 64] = 0.0;
        if (act)
            for (int em = 0; em < nem; em++) {
                const double* t = taut + (size_t)em * 12 * nk + k;
                double v[6];
#pragma unroll
                for (int sh = 0; sh < 3; sh++)
#pragma unroll
                    for (int q = 0; q < 2; q++) v[sh * 2 + q] = t[(size_t)((sh * 2 + ri) * 2 + q) * nk];
                const double* Ke = K + (size_t)em * 12;
                double f0 = 0.0, f1 = 0.0;
#pragma unroll
                for (int j = 0; j < 6; j++) { f0 = fma(Ke[j], v[j], f0); f1 = fma(Ke[6 + j], v[j], f1); }
                const int2 nn = enode[em];
                f[nn.x * 64] += f0;
                f[nn.y * 64] += f1;
            }
        for (int row = 0; row <= nmn; row++) {
            double w = 0.0;
            if (act) {
                const double* Gr = G + (size_t)row * nmn;
                for (int c = 0; c < nmn; c++) w = fma(Gr[c], f[c * 64], w);
            }
            if (row < nmn) dWt[((size_t)row * 2 + ri) * nk + k] = w;
            else wa[ri] = w;
        }
    }
    dWa[k] = make_double2(wa[0], wa[1]);
}

void vh_tor_drift(hipStream_t st, const vh_tor_plan* p, const vh_dims& d, const vh_dev& v, const double* taut,
                  const double* mnorm, const double* maxnorm, double skip_tol, double* dWt, double* dWa) {
    if (d.nk == 0) return;
    const size_t lds = (size_t)std::max(d.nmn, 1) * 64 * sizeof(double);
    k_drift_tor<<<(d.nk + 63) / 64, 64, lds, st>>>(d.nk, d.nem, d.nmn, v.kdeg, p->enode, p->Gt, p->Kt, taut, mnorm,
                                                   maxnorm, skip_tol, dWt, (double2*)dWa);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------- advance
// The toroidal channels of the memory update, beside k_advance: thread per
// (slot, Maxwell element, shape), four values (re/im x lambda 3, 4).
//   1-D element, forward Euler:  taut <- (1 - M) taut - 2 mu M eps(W)
//   1-D element, trapezoid:      taut <- [(1 - M/2) taut - mu M (eps(W_n) + eps(W_{n+1}))] / (1 + M/2)
//   3-D element:                 taut <- taut + increment (channels 4, 5 of cdel)
// (advance_memory_tor, src/vilma_viscoelastic.f90:591-629). Its per-slot and
// per-block maxima go to the advance's reduction rows after the spheroidal
// ones (row offset yoff), so the one per-slot norm covers all six channels.
template <bool Trap>
__global__ void __launch_bounds__(256) k_advance_tor(tor_adv_args a) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wg = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int pair = blockIdx.y * 4 + wg;
    const int nk = a.nk;
    double amax = 0.0;
    if (k < nk && pair < 3 * a.nem) {
        const int em = pair / 3, sh = pair % 3;
        const size_t base = (size_t)em * 12 * nk + k;
        if (a.is3d[em]) {
            const double2* cd = a.cdel + ((size_t)a.e3of[em] * 3 + sh) * 6 * a.nlm + k + 1;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const double2 dv = cd[(size_t)(4 + q) * a.nlm];
                const size_t o0 = base + (size_t)((sh * 2 + 0) * 2 + q) * nk;
                const size_t o1 = base + (size_t)((sh * 2 + 1) * 2 + q) * nk;
                const double t0 = a.tau_in[o0] + dv.x, t1 = a.tau_in[o1] + dv.y;
                a.tau_out[o0] = t0; a.tau_out[o1] = t1;
                amax = vh_nanmax(amax, vh_nanmax(fabs(t0), fabs(t1)));
            }
        } else {
            const int l = a.kdeg[k];
            const double J = (double)l * (double)(l + 1);
            const int2 nn = a.enode[em];
            const int e = a.emap[em];
            const double Me = a.MkPerDt[e] * a.dt;
            double om = 1.0 - Me, wgt = 2.0 * a.mu[e] * Me;
            if constexpr (Trap) {
                const double den = 1.0 + 0.5 * Me;
                om = (1.0 - 0.5 * Me) / den;
                wgt = a.mu[e] * Me / den;
            }
#pragma unroll
            for (int ri = 0; ri < 2; ri++) {
                double s[3][2];
                tor_strain(a.dW[((size_t)nn.x * 2 + ri) * nk + k], a.dW[((size_t)nn.y * 2 + ri) * nk + k], J,
                           s[0], s[1], s[2]);
                if constexpr (Trap) {
                    double s0[3][2];
                    tor_strain(a.dW0[((size_t)nn.x * 2 + ri) * nk + k], a.dW0[((size_t)nn.y * 2 + ri) * nk + k], J,
                               s0[0], s0[1], s0[2]);
#pragma unroll
                    for (int c = 0; c < 3; c++) { s[c][0] += s0[c][0]; s[c][1] += s0[c][1]; }
                }
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const size_t off = base + (size_t)((sh * 2 + ri) * 2 + q) * nk;
                    const double sv = sh == 0 ? s[0][q] : (sh == 1 ? s[1][q] : s[2][q]);
                    const double t = om * a.tau_in[off] - wgt * sv;
                    a.tau_out[off] = t;
                    amax = vh_nanmax(amax, fabs(t));
                }
            }
        }
    }
    red[wg][lane] = amax;
    __syncthreads();
    if (wg == 0) {
        double m = vh_nanmax(vh_nanmax(red[0][lane], red[1][lane]), vh_nanmax(red[2][lane], red[3][lane]));
        if (k < nk) a.pmax[(size_t)(a.yoff + blockIdx.y) * nk + k] = m;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = vh_nanmax(m, __shfl_xor(m, off));
        if (lane == 0) a.pblk[(size_t)(a.yoff + blockIdx.y) * gridDim.x + blockIdx.x] = m;
    }
}

void vh_launch_advance_tor(hipStream_t st, const tor_adv_args& a, bool trap) {
    dim3 g((a.nk + 63) / 64, (3 * a.nem + 3) / 4);
    if (trap) k_advance_tor<true><<<g, 256, 0, st>>>(a);
    else k_advance_tor<false><<<g, 256, 0, st>>>(a);
    HIPCHECK(hipGetLastError());
}
