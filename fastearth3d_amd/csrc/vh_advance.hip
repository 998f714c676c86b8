// vh_advance.hip -- (b) the forward-Euler Maxwell memory update of every
// (slot, Maxwell element): nodal strain U = sigma*xUn(l) + dUn (likewise V),
// strain coefficients (src/vilma_viscoelastic.f90:329-359) and
// tau <- (1-M) tau - 2 mu M eps for the three radial shape coefficients
// (:577-584, driven by vilma_advance, src/vilma_response.f90:944-992).
//
// One streaming pass: tau_n is read from one buffer and tau_{n+1} written to
// the other (the step controller's rollback is then free), and the same pass
// produces the per-slot max |tau_{n+1}| that the next drift solve's skip gate
// and the controller's memory-norm guard need (src/vilma_response.f90:719-728,
// :1795-1811), so no separate scan of the memory is ever made.
#include "vh_internal.h"
#include "vh_radial_dev.h"

struct advance_args {
    int nr, nk, nem;
    double dt;
    vh_dev v;
    const double* tau_in;
    double* tau_out;
    const double* sigma;   // complex degree-major
    const double* dN;      // [mn][4][nk]
    double* pmax;          // [gridDim.y][nk] per-block per-slot max |tau_{n+1}|
    double* pblk;          // [gridDim.y * gridDim.x] per-block max
    const int* e3of;       // [nem] -> 3-D element index or -1
    const double2* cdel;   // [i3*3 + shape][lambda][nlm], nch channels
    int nlm, nch;
    // trapezoidal rule: dN is the endpoint drift (eps_{n+1} with sigma), dN0
    // the start-of-step drift (eps_n with sigma_n)
    const double* sigma_n;
    const double* dN0;
};

// Grid: x = tile of 64 slots (lanes, coalesced rows of tau), y = group of
// four (Maxwell element, shape) pairs (one per wavefront), so every thread
// moves 8 memory values (re/im x 4 channels) and the launch has thousands of
// blocks. Each block stores its per-slot max |tau_{n+1}| and its overall max
// (plain stores); k_advance_max combines them (max is exact in any order).
// Trap: tau_{n+1} = [(1 - M/2) tau_n - mu M (eps_n + eps_{n+1})] / (1 + M/2)
// (src/vilma_viscoelastic.f90:551-565) in place of forward Euler.
template <bool Trap>
__global__ void __launch_bounds__(256) k_advance(advance_args a) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wg = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    const int pair = blockIdx.y * 4 + wg;             // (em, shape) = (pair / 3, pair % 3)
    const int nk = a.nk;
    double amax = 0.0;
    if (k < nk && pair < 3 * a.nem) {
        const int em = pair / 3, sh = pair % 3;
        const size_t base = (size_t)em * 24 * nk + k;
        if (a.v.is3d[em]) {   // tau_{n+1} = tau_n + 3-D increment (increment form)
            const double2* cd = a.cdel + ((size_t)a.e3of[em] * 3 + sh) * a.nch * a.nlm + k + 1;
#pragma unroll
            for (int lam = 0; lam < 4; lam++) {
                const double2 dv = cd[(size_t)lam * a.nlm];
                const size_t o0 = base + (size_t)((sh * 2 + 0) * 4 + lam) * nk;
                const size_t o1 = base + (size_t)((sh * 2 + 1) * 4 + lam) * nk;
                const double t0 = a.tau_in[o0] + dv.x, t1 = a.tau_in[o1] + dv.y;
                a.tau_out[o0] = t0; a.tau_out[o1] = t1;
                amax = vh_nanmax(amax, vh_nanmax(fabs(t0), fabs(t1)));
            }
        } else {
            const int l = a.v.kdeg[k];
            const double Jr = (double)l * (double)(l + 1);
            const double sre = a.sigma[2 * (k + 1)], sim = a.sigma[2 * (k + 1) + 1];
            const double* xU = a.v.xUn + (size_t)(l - 1) * a.nr;
            const double* xV = a.v.xVn + (size_t)(l - 1) * a.nr;
            const int e = a.v.emap[em];
            const int n0 = a.v.mnode[e], n1 = a.v.mnode[e + 1];
            const double* d0 = a.dN + (size_t)n0 * 4 * nk + k;
            const double* d1 = a.dN + (size_t)n1 * 4 * nk + k;
            double U0r = sre * xU[e] + d0[0], U0i = sim * xU[e] + d0[(size_t)nk];
            double V0r = sre * xV[e] + d0[(size_t)2 * nk], V0i = sim * xV[e] + d0[(size_t)3 * nk];
            double U1r = sre * xU[e + 1] + d1[0], U1i = sim * xU[e + 1] + d1[(size_t)nk];
            double V1r = sre * xV[e + 1] + d1[(size_t)2 * nk], V1i = sim * xV[e + 1] + d1[(size_t)3 * nk];
            double s[2][3][4];
            strain_coeffs(U0r, U1r, V0r, V1r, Jr, s[0][0], s[0][1], s[0][2]);
            strain_coeffs(U0i, U1i, V0i, V1i, Jr, s[1][0], s[1][1], s[1][2]);
            const double Me = a.v.MkPerDt[e] * a.dt;   // response_set_dt: Mk = MkPerDt*dt
            double om = 1.0 - Me, two_muM = 2.0 * a.v.mu[e] * Me;
            if constexpr (Trap) {   // eps_n from (sigma_n, dN0), summed into s; weights of the rule
                const double srn = a.sigma_n[2 * (k + 1)], sin_ = a.sigma_n[2 * (k + 1) + 1];
                const double* p0 = a.dN0 + (size_t)n0 * 4 * nk + k;
                const double* p1 = a.dN0 + (size_t)n1 * 4 * nk + k;
                double t[2][3][4];
                strain_coeffs(srn * xU[e] + p0[0], srn * xU[e + 1] + p1[0], srn * xV[e] + p0[(size_t)2 * nk],
                              srn * xV[e + 1] + p1[(size_t)2 * nk], Jr, t[0][0], t[0][1], t[0][2]);
                strain_coeffs(sin_ * xU[e] + p0[(size_t)nk], sin_ * xU[e + 1] + p1[(size_t)nk],
                              sin_ * xV[e] + p0[(size_t)3 * nk], sin_ * xV[e + 1] + p1[(size_t)3 * nk], Jr,
                              t[1][0], t[1][1], t[1][2]);
#pragma unroll
                for (int ri = 0; ri < 2; ri++)
#pragma unroll
                    for (int q = 0; q < 3; q++)
#pragma unroll
                        for (int lam = 0; lam < 4; lam++) s[ri][q][lam] += t[ri][q][lam];
                const double den = 1.0 + 0.5 * Me;
                om = (1.0 - 0.5 * Me) / den;
                two_muM = a.v.mu[e] * Me / den;   // weight of eps_n + eps_{n+1}
            }
#pragma unroll
            for (int ri = 0; ri < 2; ri++)
#pragma unroll
                for (int lam = 0; lam < 4; lam++) {
                    const size_t off = base + (size_t)((sh * 2 + ri) * 4 + lam) * nk;
                    const double sv = sh == 0 ? s[ri][0][lam] : (sh == 1 ? s[ri][1][lam] : s[ri][2][lam]);
                    const double t = om * a.tau_in[off] - two_muM * sv;
                    a.tau_out[off] = t;
                    amax = vh_nanmax(amax, fabs(t));
                }
        }
    }
    red[wg][lane] = amax;
    __syncthreads();
    if (wg == 0) {
        double m = vh_nanmax(vh_nanmax(red[0][lane], red[1][lane]), vh_nanmax(red[2][lane], red[3][lane]));
        if (k < nk) a.pmax[(size_t)blockIdx.y * nk + k] = m;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = vh_nanmax(m, __shfl_xor(m, off));
        if (lane == 0) a.pblk[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
}

// mnorm[k] = max over the advance's y-blocks of pmax[y][k]; the last block
// also forms the global max of the per-block maxima.
// Per-slot maxima over the advance's block rows, and (last block) the global
// maximum over the per-block maxima. One 64-thread block per 64 slots so the
// reduction spreads over the CUs; loads are issued 8 at a time (max is exact
// in any order).
__global__ void __launch_bounds__(256) k_advance_max(int nk, int ny, int nblk, const double* __restrict__ pmax,
                                                     const double* __restrict__ pblk, double* mnorm, double* maxn,
                                                     double* maxn_copy) {
    if ((int)blockIdx.x == (int)gridDim.x - 1) {
        __shared__ double red[256];
        double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
        int i = threadIdx.x;
        for (; i + 3 * (int)blockDim.x < nblk; i += 4 * blockDim.x) {
            m0 = vh_nanmax(m0, pblk[i]);
            m1 = vh_nanmax(m1, pblk[i + blockDim.x]);
            m2 = vh_nanmax(m2, pblk[i + 2 * blockDim.x]);
            m3 = vh_nanmax(m3, pblk[i + 3 * blockDim.x]);
        }
        for (; i < nblk; i += blockDim.x) m0 = vh_nanmax(m0, pblk[i]);
        red[threadIdx.x] = vh_nanmax(vh_nanmax(m0, m1), vh_nanmax(m2, m3));
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = vh_nanmax(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            maxn[0] = red[0];
            if (maxn_copy) maxn_copy[0] = red[0];
        }
        return;
    }
    if (threadIdx.x >= 64) return;
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= nk) return;
    double m[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int y = 0;
    for (; y + 8 <= ny; y += 8) {
#pragma unroll
        for (int u = 0; u < 8; u++) m[u] = vh_nanmax(m[u], pmax[(size_t)(y + u) * nk + k]);
    }
    for (; y < ny; y++) m[0] = vh_nanmax(m[0], pmax[(size_t)y * nk + k]);
    mnorm[k] = vh_nanmax(vh_nanmax(vh_nanmax(m[0], m[1]), vh_nanmax(m[2], m[3])), vh_nanmax(vh_nanmax(m[4], m[5]), vh_nanmax(m[6], m[7])));
}

size_t vh_advance_scratch(const vh_dims& d, bool tor) {
    const size_t gx = (d.nk + 63) / 64, gy = (3 * (size_t)d.nem + 3) / 4 * (tor ? 2 : 1);
    return gy * d.nk + gx * gy;
}

void vh_launch_advance(hipStream_t st, const vh_dims& d, const vh_dev& v, double dt, const double* tau_in,
                       double* tau_out, const double* sigma, const double* dN, double* mnorm_out,
                       double* maxnorm_out, const int* e3of, const double2* cdel, double* scratch,
                       double* maxn_copy, const double* sigma_n, const double* dN0, int nch,
                       const vh_adv_tor& tor) {
    if (d.nem == 0) {
        vh_fill(st, mnorm_out, 0.0, d.nk);
        vh_fill(st, maxnorm_out, 0.0, 1);
        if (maxn_copy) vh_fill(st, maxn_copy, 0.0, 1);
        return;
    }
    advance_args a;
    a.e3of = e3of; a.cdel = cdel; a.nlm = d.nlm; a.nch = nch;
    a.nr = d.nr; a.nk = d.nk; a.nem = d.nem; a.dt = dt; a.v = v;
    a.tau_in = tau_in; a.tau_out = tau_out; a.sigma = sigma; a.dN = dN;
    a.sigma_n = sigma_n; a.dN0 = dN0;
    dim3 g((d.nk + 63) / 64, (3 * d.nem + 3) / 4);
    const bool with_tor = tor.taut_in != nullptr;
    const unsigned ny = g.y * (with_tor ? 2 : 1);   // reduction rows: spheroidal, then toroidal
    a.pmax = scratch;
    a.pblk = scratch + (size_t)ny * d.nk;
    if (sigma_n) k_advance<true><<<g, 256, 0, st>>>(a);
    else k_advance<false><<<g, 256, 0, st>>>(a);
    HIPCHECK(hipGetLastError());
    if (with_tor) {
        tor_adv_args t;
        t.nk = d.nk; t.nem = d.nem; t.nlm = d.nlm; t.yoff = (int)g.y; t.dt = dt;
        t.kdeg = v.kdeg; t.emap = v.emap; t.mu = v.mu; t.MkPerDt = v.MkPerDt; t.is3d = v.is3d;
        t.e3of = e3of; t.enode = vh_tor_plan_enode(tor.plan); t.cdel = cdel;
        t.tau_in = tor.taut_in; t.tau_out = tor.taut_out; t.dW = tor.dW; t.dW0 = tor.dW0;
        t.pmax = a.pmax; t.pblk = a.pblk;
        vh_launch_advance_tor(st, t, sigma_n != nullptr);
    }
    // blocks 0 .. g.x-1: 64 slots each (threads 64..255 of them idle); the last: the global maximum
    k_advance_max<<<g.x + 1, 256, 0, st>>>(d.nk, (int)ny, (int)(g.x * ny), a.pmax, a.pblk, mnorm_out, maxnorm_out,
                                           maxn_copy);
    HIPCHECK(hipGetLastError());
}

// Per-slot max |tau| and the global max of a memory state that did not come
// out of an advance (a restart or a host-set memory).
__global__ void __launch_bounds__(256) k_memnorm(int nk, int nem, const double* __restrict__ tau,
                                                 const double* __restrict__ taut, double* mnorm,
                                                 unsigned long long* maxbits) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, eg = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane;
    double amax = 0.0;
    if (k < nk)
        for (int em = eg; em < nem; em += 4) {
            for (int c = 0; c < 24; c++) amax = vh_nanmax(amax, fabs(tau[((size_t)em * 24 + c) * nk + k]));
            if (taut)
                for (int c = 0; c < 12; c++) amax = vh_nanmax(amax, fabs(taut[((size_t)em * 12 + c) * nk + k]));
        }
    red[eg][lane] = amax;
    __syncthreads();
    if (eg == 0) {
        double m = vh_nanmax(vh_nanmax(red[0][lane], red[1][lane]), vh_nanmax(red[2][lane], red[3][lane]));
        if (k < nk) mnorm[k] = m;
        red[0][lane] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < 64; i++) m = vh_nanmax(m, red[0][i]);
        atomicMax(maxbits, (unsigned long long)__double_as_longlong(m));
    }
}

void vh_launch_memnorm(hipStream_t st, const vh_dims& d, const double* tau, double* mnorm, double* maxnorm,
                       const double* taut) {
    vh_fill(st, maxnorm, 0.0, 1);
    k_memnorm<<<(d.nk + 63) / 64, 256, 0, st>>>(d.nk, d.nem, tau, taut, mnorm, (unsigned long long*)maxnorm);
    HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------------ TRAP norms
// Maxima of non-negative values through atomicMax on their bit patterns
// (the order of non-negative doubles is that of their bits; a NaN from fabs
// sorts above +inf, so it is kept). out must start at zero.
__device__ __forceinline__ void max_bits(unsigned long long* out, double v) {
    atomicMax(out, (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ double block_max(double v, double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = vh_nanmax(v, __shfl_xor(v, off));
    if (lane == 0) red[w] = v;
    __syncthreads();
    double m = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) m = vh_nanmax(m, red[i]);
    return m;
}

// response_coarse_fine_error (src/vilma_response.f90:1759-1788): out[0] =
// max |fine - coarse|, out[1] = max |fine| over n values
__global__ void __launch_bounds__(256) k_coarse_fine(size_t n, const double* __restrict__ fine,
                                                     const double* __restrict__ coarse, unsigned long long* out) {
    __shared__ double red[8];
    double e = 0.0, t = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double f = fine[i];
        e = vh_nanmax(e, fabs(f - coarse[i]));
        t = vh_nanmax(t, fabs(f));
    }
    e = block_max(e, red);
    __syncthreads();
    t = block_max(t, red + 4);
    if (threadIdx.x == 0) { max_bits(out, e); max_bits(out + 1, t); }
}

// drift_change (src/vilma_response.f90:1499-1512) of the report drift dUa
// (the first nk complex values of dsurf): out[0] = max |dUa - prev|, out[1]
// = max |dUa|; then prev <- dUa (keep_drift). With the toroidal channels the
// surface W drift dWa joins both maxima (and prev_t <- dWa).
__global__ void __launch_bounds__(256) k_drift_change(int nk, const double2* __restrict__ dsurf, double2* prev,
                                                      const double2* __restrict__ dwa, double2* prev_t,
                                                      unsigned long long* out) {
    __shared__ double red[8];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double c = 0.0, s = 0.0;
    if (k < nk) {
        const double2 u = dsurf[k], p = prev[k];
        c = hypot(u.x - p.x, u.y - p.y);
        s = hypot(u.x, u.y);
        prev[k] = u;
        if (dwa) {
            const double2 w = dwa[k], q = prev_t[k];
            c = vh_nanmax(c, hypot(w.x - q.x, w.y - q.y));
            s = vh_nanmax(s, hypot(w.x, w.y));
            prev_t[k] = w;
        }
    }
    c = block_max(c, red);
    __syncthreads();
    s = block_max(s, red + 4);
    if (threadIdx.x == 0) { max_bits(out, c); max_bits(out + 1, s); }
}

void vh_coarse_fine(hipStream_t st, size_t n, const double* fine, const double* coarse, double* out2,
                    bool accumulate) {
    if (!accumulate) vh_fill(st, out2, 0.0, 2);
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
    k_coarse_fine<<<std::max(nb, 1u), 256, 0, st>>>(n, fine, coarse, (unsigned long long*)out2);
    HIPCHECK(hipGetLastError());
}

void vh_drift_change(hipStream_t st, int nk, const double* dsurf, double* prev, double* out2,
                     const double* dWa, double* prev_t) {
    vh_fill(st, out2, 0.0, 2);
    k_drift_change<<<(nk + 255) / 256, 256, 0, st>>>(nk, (const double2*)dsurf, (double2*)prev,
                                                     (const double2*)dWa, (double2*)prev_t,
                                                     (unsigned long long*)out2);
    HIPCHECK(hipGetLastError());
}
