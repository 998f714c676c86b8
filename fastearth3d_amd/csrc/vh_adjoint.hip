// vh_adjoint.hip -- reverse-mode kernels of the forward-Euler viscoelastic
// sub-step (memory advance, 3-D increment, sea-level solve), for gradients of
// a misfit of rsl with respect to the lateral log-viscosity field and the ice
// history. The forward sub-step is, for its recorded discrete decisions
// (skip-gate set, coastline mask of the last pass, sub-step sizes, the set
// of genuinely 3-D elements), linear in the memory, the ice and the drift,
// and smooth in the Maxwell rates; every kernel here is the exact transpose
// (Euclidean, on the stored real and imaginary parts) of a forward kernel:
//
//   advance   tau' = (1 - M) tau - 2 mu M eps(sigma, dN)     (1-D elements,
//             src/vilma_viscoelastic.f90:577-584)
//             tau' = tau + Delta(tau + 2 mu eps)             (3-D elements,
//             src/vilma_response.f90:1171-1225, :1274-1319)
//   3-D       Delta(c) = A(-M .* S c) with the tensor synthesis S and its
//             quadrature adjoint A = K S^T Omega W, K diagonal per (channel,
//             degree, order), Omega the dyad weights (tp counts 1/2), W the
//             Gauss weights: so Delta^T = K^{-1} Delta K and the forward
//             kernels serve the adjoint between two diagonal scalings;
//   rates     dJ/dM(grid) = -(2 pi/nphi) w_j sum_dyads Omega (S K c_h)(S c).
#include <atomic>
#include "vh_internal.h"
#include "vh_tuned.h"
#include "vh_radial_dev.h"
#include "vh_sht_dev.h"
#include "vh_adjoint.h"

// diagonal K' = 1 / (N_lam(l) c_m) of the tensor analysis (c_0 = 1, c_m = 2;
// N = 1, J, 2 J^2, n6(l) for lambda = 1, 2, 5, 6); 0 outside the transform's range
__device__ __forceinline__ double tsh_kinv(int lam, int l, int m, const double* n6) {
    const double J = (double)l * (double)(l + 1);
    double n = lam == 0 ? 1.0 : (lam == 1 ? J : (lam == 2 ? 2.0 * J * J : (l >= 2 ? n6[l] : 0.0)));
    return n * (m == 0 ? 1.0 : 2.0);
}

__device__ __forceinline__ int slot_order(int k, int l) { return k + 1 - l * (l + 1) / 2; }

// ------------------------------------------------------------ advance adjoint
// Grid: x = 64 slots, y = 4 Maxwell elements (one per wavefront); a thread
// owns the 24 memory values of its (element, slot). tauh_out = adjoint of
// tau_n (overwritten); Ue[em][8][nk] = the element's adjoints of its nodal
// (U0, U1, V0, V1) x (re, im); mpart[em][bx] = this block's part of
// dJ/dM_e (1-D elements). Xh = the 3-D increment's adjoint in K'-scaled
// tensor coefficients [i3*3 + shape][lambda][nlm].
__global__ void __launch_bounds__(256) k_adv_adj(adv_adj_args a) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane, em = blockIdx.y * 4 + wv;
    const int nk = a.nk;
    double mh = 0.0;
    if (k < nk && em < a.nem) {
        const int e = a.v.emap[em], l = a.v.kdeg[k];
        const double Jr = (double)l * (double)(l + 1);
        const size_t base = (size_t)em * 24 * nk + k;
        double Eh[2][3][4];
        if (a.v.is3d[em]) {
            const int m = slot_order(k, l);
            const double2* xh = a.Xh + (size_t)a.e3of[em] * 12 * a.nlm + k + 1;
            const double twoMu = 2.0 * a.v.mu[e];
#pragma unroll
            for (int sh = 0; sh < 3; sh++)
#pragma unroll
                for (int lam = 0; lam < 4; lam++) {
                    double2 x = xh[(size_t)(sh * 4 + lam) * a.nlm];
                    const double f = tsh_kinv(lam, l, m, a.n6);
                    x = make_double2(x.x * f, x.y * f);
                    const size_t o0 = base + (size_t)((sh * 2 + 0) * 4 + lam) * nk;
                    const size_t o1 = base + (size_t)((sh * 2 + 1) * 4 + lam) * nk;
                    a.tauh_out[o0] = a.tauh_in[o0] + x.x;
                    a.tauh_out[o1] = a.tauh_in[o1] + x.y;
                    Eh[0][sh][lam] = twoMu * x.x;
                    Eh[1][sh][lam] = twoMu * x.y;
                }
        } else {
            const double sre = a.sigma[2 * (k + 1)], sim = a.sigma[2 * (k + 1) + 1];
            const double* xU = a.v.xUn + (size_t)(l - 1) * a.nr;
            const double* xV = a.v.xVn + (size_t)(l - 1) * a.nr;
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
            const double Me = a.v.MkPerDt[e] * a.dt;
            const double om = 1.0 - Me, two_mu = 2.0 * a.v.mu[e], two_muM = two_mu * Me;
            const double* t1 = a.tau + (size_t)a.e1of[em] * 24 * nk + k;
#pragma unroll
            for (int sh = 0; sh < 3; sh++)
#pragma unroll
                for (int ri = 0; ri < 2; ri++)
#pragma unroll
                    for (int lam = 0; lam < 4; lam++) {
                        const size_t off = base + (size_t)((sh * 2 + ri) * 4 + lam) * nk;
                        const double th = a.tauh_in[off];
                        a.tauh_out[off] = om * th;
                        Eh[ri][sh][lam] = -two_muM * th;
                        mh = fma(th, -t1[(size_t)((sh * 2 + ri) * 4 + lam) * nk] - two_mu * s[ri][sh][lam], mh);
                    }
        }
        double uh[2][4];   // (U0, U1, V0, V1) per re/im
#pragma unroll
        for (int ri = 0; ri < 2; ri++) {
            uh[ri][0] = uh[ri][1] = uh[ri][2] = uh[ri][3] = 0.0;
            strain_coeffs_adj(Eh[ri][0], Eh[ri][1], Eh[ri][2], Jr, uh[ri][0], uh[ri][1], uh[ri][2], uh[ri][3]);
        }
        double* u = a.Ue + (size_t)em * 8 * nk + k;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            u[(size_t)(2 * q + 0) * nk] = uh[0][q];
            u[(size_t)(2 * q + 1) * nk] = uh[1][q];
        }
    }
    // per-block part of dJ/dM_e: a wavefront holds one element
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mh += __shfl_xor(mh, off);
    if (lane == 0 && em < a.nem) a.mpart[(size_t)em * gridDim.x + blockIdx.x] = mh;
}

// Per slot: the nodal adjoints of the memory nodes (sums over the elements
// below and above) -> dN_h, and sigma_h = sum_q xU(r_q) U_h + xV(r_q) V_h.
__global__ void __launch_bounds__(256) k_adv_nodes(int nr, int nk, int nmn, vh_dev v, const int* __restrict__ mnodes,
                                                   const int2* __restrict__ node_el, const double* __restrict__ Ue,
                                                   double* dNh, double* sigh) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) { sigh[0] = 0.0; sigh[1] = 0.0; }
    if (k >= nk) return;
    const int l = v.kdeg[k];
    const double* xU = v.xUn + (size_t)(l - 1) * nr;
    const double* xV = v.xVn + (size_t)(l - 1) * nr;
    double sr = 0.0, si = 0.0;
    for (int q = 0; q < nmn; q++) {
        const int2 ne = node_el[q];   // .x: element below (q its upper node), .y: above (q its lower node)
        double h[4] = {0.0, 0.0, 0.0, 0.0};   // U re, U im, V re, V im
        if (ne.x >= 0) {
            const double* u = Ue + (size_t)ne.x * 8 * nk + k;
            h[0] += u[(size_t)2 * nk]; h[1] += u[(size_t)3 * nk]; h[2] += u[(size_t)6 * nk]; h[3] += u[(size_t)7 * nk];
        }
        if (ne.y >= 0) {
            const double* u = Ue + (size_t)ne.y * 8 * nk + k;
            h[0] += u[0]; h[1] += u[(size_t)nk]; h[2] += u[(size_t)4 * nk]; h[3] += u[(size_t)5 * nk];
        }
        double* o = dNh + (size_t)q * 4 * nk + k;
        o[0] = h[0]; o[(size_t)nk] = h[1]; o[(size_t)2 * nk] = h[2]; o[(size_t)3 * nk] = h[3];
        const int r = mnodes[q];
        sr = fma(xU[r], h[0], fma(xV[r], h[2], sr));
        si = fma(xU[r], h[1], fma(xV[r], h[3], si));
    }
    sigh[2 * (k + 1)] = sr;
    sigh[2 * (k + 1) + 1] = si;
}

// Kh[em] += dt * sum over blocks of mpart: one 64-lane block per element,
// strided partial sums then a fixed butterfly (deterministic)
__global__ void __launch_bounds__(64) k_mreduce(int nem, int nbx, const double* __restrict__ mpart, double dt,
                                                double* Kh) {
    const int em = blockIdx.x, lane = threadIdx.x;
    double s = 0.0;
    for (int b = lane; b < nbx; b += 64) s += mpart[(size_t)em * nbx + b];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) Kh[em] += dt * s;
}

void vh_advance_adjoint(hipStream_t st, const vh_dims& d, const adv_adj_args& a0, const int* mnodes,
                        const int2* node_el, double* dNh, double* sigh, double* Kh) {
    adv_adj_args a = a0;
    if (d.nem == 0) return;
    dim3 g((d.nk + 63) / 64, (d.nem + 3) / 4);
    k_adv_adj<<<g, 256, 0, st>>>(a);
    HIPCHECK(hipGetLastError());
    k_mreduce<<<d.nem, 64, 0, st>>>(d.nem, (int)g.x, a.mpart, a.dt, Kh);
    HIPCHECK(hipGetLastError());
    k_adv_nodes<<<(d.nk + 255) / 256, 256, 0, st>>>(d.nr, d.nk, d.nmn, a.v, mnodes, node_el, a.Ue, dNh, sigh);
    HIPCHECK(hipGetLastError());
}

// --------------------------------------------------------------- 3-D adjoint
// Xa[(i3*3 + s)*4 + lam][k+1] = K'(lam, l, m) tau_h of (element e3em[i3], shape s)
__global__ void __launch_bounds__(256) k_t3_pack_adj(int nk, int nlm, vh_dev v, const int* __restrict__ e3em,
                                                     const double* __restrict__ n6, const double* __restrict__ tauh,
                                                     double2* __restrict__ Xa) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i3 = blockIdx.y;
    double2* Xb = Xa + (size_t)i3 * 12 * nlm;
    if (k == 0)
        for (int q = 0; q < 12; q++) Xb[(size_t)q * nlm] = make_double2(0.0, 0.0);
    if (k >= nk) return;
    const int em = e3em[i3], l = v.kdeg[k], m = slot_order(k, l);
    const double* t = tauh + (size_t)em * 24 * nk + k;
#pragma unroll
    for (int sh = 0; sh < 3; sh++)
#pragma unroll
        for (int lam = 0; lam < 4; lam++) {
            const double n = tsh_kinv(lam, l, m, n6);
            const double f = n != 0.0 ? 1.0 / n : 0.0;
            Xb[(size_t)(sh * 4 + lam) * nlm + k + 1] =
                make_double2(f * t[(size_t)((sh * 2 + 0) * 4 + lam) * nk], f * t[(size_t)((sh * 2 + 1) * 4 + lam) * nk]);
        }
}

__device__ __forceinline__ int xcd_block_a(int nblk) {
    const int cnt = (nblk + 7) / 8;
    return (int)(blockIdx.x % 8) * cnt + (int)(blockIdx.x / 8);
}

// Block = (3-D element i3, latitude row j); three wavefronts, one per dyad
// pair (rr, rt), (rp, tt), (tp, pp). For each of the element's three shapes:
// the forward rows Ff to real space (kept in registers), the adjoint rows Fa
// to real space, the product sum_dy Omega_dy g_f g_a accumulated per
// longitude (registers), then Fa <- FFT(-rate dt g_a) in place (the forward
// row pass, mode 0). At the end rate_h[i3][j][i] += -(2 pi/nphi) w_j dt *
// the accumulated product summed over the three wavefronts.
// RU: longitudes per lane kept in registers (nphi <= 64 RU); spectra
// loaded RA_LU orders per lane at a time (fewer registers than the row pass)
#define RA_U 8   // most longitudes per lane (nphi <= 512)
#define RA_LU 4
template <int RU>
__global__ void __launch_bounds__(192) k_t3_rows_adj(vh_sht s, double2* __restrict__ Fa, const double2* __restrict__ Ff,
                                                     int ne3, const double* __restrict__ rate, double dt,
                                                     double* __restrict__ rate_h) {
    extern __shared__ double2 lds[];
    const int L = s.L, N = s.nphi, nlat = s.nlat;
    const int lb = xcd_block_a(ne3 * nlat);
    if (lb >= ne3 * nlat) return;
    const int i3 = lb / nlat, j = lb % nlat;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double2* a = lds + (size_t)wave * 2 * N;
    double2* c = a + N;
    double* Mrow = (double*)(lds + (size_t)6 * N);
    const double* rr = rate + ((size_t)i3 * nlat + j) * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) Mrow[i] = rr[i] * dt;
    double acc[RU];
#pragma unroll
    for (int u = 0; u < RU; u++) acc[u] = 0.0;
    const double omA = wave == 2 ? 0.5 : 1.0;   // dyad weights: tp (first of wave 2) counts 1/2, the rest 1
    const int d0 = 2 * wave, d1 = 2 * wave + 1;
    __syncthreads();   // Mrow complete
    for (int sh = 0; sh < 3; sh++) {
        const int b = 3 * i3 + sh;
        double2 gf[RU];
        double2* z = nullptr;
        for (int pass = 0; pass < 2; pass++) {
            const double2* Fs = pass == 0 ? Ff : (const double2*)Fa;
            const double2* F0 = Fs + ((size_t)b * 6 + d0) * (L + 1) * nlat + j;
            const double2* F1 = Fs + ((size_t)b * 6 + d1) * (L + 1) * nlat + j;
            for (int i = lane; i < N; i += 64) a[i] = make_double2(0.0, 0.0);
            wave_lds_sync();
            for (int m0 = 0; m0 <= L; m0 += 64 * RA_LU) {
                double2 A[RA_LU], B[RA_LU];
#pragma unroll
                for (int u = 0; u < RA_LU; u++) {
                    const int m = m0 + lane + 64 * u;
                    A[u] = m <= L ? F0[(size_t)m * nlat] : make_double2(0.0, 0.0);
                    B[u] = m <= L ? F1[(size_t)m * nlat] : make_double2(0.0, 0.0);
                }
#pragma unroll
                for (int u = 0; u < RA_LU; u++) {
                    const int m = m0 + lane + 64 * u;
                    if (m > L) continue;
                    if (m == 0) { A[u].y = 0.0; B[u].y = 0.0; }
                    a[m] = make_double2(A[u].x - B[u].y, A[u].y + B[u].x);
                    if (m > 0) a[N - m] = make_double2(A[u].x + B[u].y, -A[u].y + B[u].x);
                }
            }
            wave_lds_sync();
            z = fft_wave_t<true>(a, c, s, lane);
            if (pass == 0) {
#pragma unroll
                for (int u = 0; u < RU; u++) {
                    const int i = lane + 64 * u;
                    gf[u] = i < N ? z[i] : make_double2(0.0, 0.0);
                }
                wave_lds_sync();
            }
        }
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int i = lane + 64 * u;
            if (i < N) {
                const double2 g = z[i];
                acc[u] = fma(omA * g.x, gf[u].x, fma(g.y, gf[u].y, acc[u]));
                const double M = Mrow[i];
                z[i] = make_double2(-M * g.x, -M * g.y);
            }
        }
        wave_lds_sync();
        double2* other = (z == a) ? c : a;
        double2* y = fft_wave_t<false>(z, other, s, lane);
        const double scale = 2.0 * VH_PI / N;
        double2* G0 = Fa + ((size_t)b * 6 + d0) * (L + 1) * nlat + j;
        double2* G1 = Fa + ((size_t)b * 6 + d1) * (L + 1) * nlat + j;
        for (int m = lane; m <= L; m += 64) {
            double2 zm = y[m], zc = y[(N - m) % N];
            zc.y = -zc.y;
            const double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
            const double2 Dd = csub(zm, zc);
            const double2 B = make_double2(0.5 * Dd.y, -0.5 * Dd.x);
            G0[(size_t)m * nlat] = make_double2(scale * A.x, scale * A.y);
            G1[(size_t)m * nlat] = make_double2(scale * B.x, scale * B.y);
        }
        wave_lds_sync();
    }
    // cross-wavefront sum through the (now free) FFT buffers
    __syncthreads();
    double* red = (double*)a;
#pragma unroll
    for (int u = 0; u < RU; u++) {
        const int i = lane + 64 * u;
        if (i < N) red[i] = acc[u];
    }
    __syncthreads();
    const double f = -(2.0 * VH_PI / N) * s.w[j] * dt;
    const double* r0 = (const double*)lds;
    const double* r1 = (const double*)(lds + (size_t)2 * N);
    const double* r2 = (const double*)(lds + (size_t)4 * N);
    double* out = rate_h + ((size_t)i3 * nlat + j) * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) out[i] += f * (r0[i] + r1[i] + r2[i]);
}

// The same reverse row pass over RJ consecutive latitudes per block (one
// wavefront each; every wavefront runs its latitude's three dyad pairs and
// three shapes in the order above, the pairs' products in separate
// accumulators summed at the end as above), for the grid widths with a
// compile-time transform: the spectra are read and written as whole lines
// of RJ latitudes, so the vector memory unit handles RJ times fewer lines.
// The operations of k_t3_rows_adj in the same order with the in-place
// transform (results agree to rounding). LDS: RJ transform buffers (stride
// N + 1 against bank conflicts), the twiddles, RJ rate rows.
template <int N, int RJ>
__global__ void __launch_bounds__(64 * RJ, VH_ROWS_ADJ_MINW) k_t3_rows_adj_g(vh_sht s, double2* __restrict__ Fa,
                                                           const double2* __restrict__ Ff, int ne3,
                                                           const double* __restrict__ rate, double dt,
                                                           double* __restrict__ rate_h) {
    static_assert(N % 64 == 0 && 64 % RJ == 0, "grouped reverse row pass shape");
    constexpr int RU = N / 64, MS = 64, MU = (N / 2 + MS - 1) / MS, NP = N + 1;
    extern __shared__ double2 lds[];
    const int L = s.L, nlat = s.nlat, ng = (nlat + RJ - 1) / RJ;
    const int lb = xcd_block_a(ne3 * ng);
    if (lb >= ne3 * ng) return;
    const int i3 = lb / ng, j0 = (lb % ng) * RJ, nj = min(RJ, nlat - j0);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    double2* TW = lds + (size_t)RJ * NP;
    double* Rr = (double*)(TW + s.ntw);
    for (int t0 = wave * 64; t0 < s.ntw; t0 += 64 * RJ)
        if (t0 + lane < s.ntw) lds_dma16(s.twp + t0 + lane, TW + t0);
    if (wave < nj) {
        const double2* rr = (const double2*)(rate + ((size_t)i3 * nlat + j0 + wave) * N);
        double2* rd = (double2*)(Rr + (size_t)wave * N);
#pragma unroll
        for (int c0 = 0; c0 < N / 2; c0 += 64)
            if (N / 2 % 64 == 0 || c0 + lane < N / 2) lds_dma16(rr + c0 + lane, rd + c0);
    }
    const int jj = tid % RJ, mm = tid / RJ;
    const bool jok = jj < nj;
    const size_t dstr = (size_t)(L + 1) * nlat;
    double2* z = lds + (size_t)wave * NP;         // this wavefront's latitude
    double2* aj = lds + (size_t)jj * NP;          // the latitude this thread loads and stores
    // rows (2p, 2p + 1) of batch b from F into the buffers (after a barrier)
    auto fill = [&](const double2* F, size_t off) __attribute__((always_inline)) {
        double2 A[MU], B[MU];
#pragma unroll
        for (int u = 0; u < MU; u++) {
            const int m = mm + MS * u;
            const bool ok = jok && m <= L;
            A[u] = ok ? F[off + (size_t)m * nlat] : make_double2(0.0, 0.0);
            B[u] = ok ? F[off + dstr + (size_t)m * nlat] : make_double2(0.0, 0.0);
        }
        __syncthreads();   // every wavefront is done with its buffer
        for (int i = L + 1 + lane; i < N - L; i += 64) z[i] = make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < MU; u++) {
            const int m = mm + MS * u;
            if (!jok || m > L) continue;
            double2 x = A[u], y = B[u];
            if (m == 0) { x.y = 0.0; y.y = 0.0; }
            aj[m] = make_double2(x.x - y.y, x.y + y.x);
            if (m > 0) aj[N - m] = make_double2(x.x + y.y, -x.y + y.x);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    double acc[3][RU];
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int u = 0; u < RU; u++) acc[p][u] = 0.0;
    const double* R = Rr + (size_t)wave * N;
    const double scale = 2.0 * VH_PI / N;
    static_assert(fft_last<N>::R == RU, "the last pass's outputs are the lane's longitudes");
    double2 tl[fft_last<N>::R - 1];   // the last pass's twiddles, from LDS once the copies are in
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    fft_last_twiddles<N>(TW, lane, tl);
    for (int sh = 0; sh < 3; sh++) {
        const int b = 3 * i3 + sh;
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const size_t off = ((size_t)b * 6 + 2 * p) * dstr + j0 + jj;
            const double omA = p == 2 ? 0.5 : 1.0;   // dyad weights: tp counts 1/2, the rest 1
            double2 gf[RU];
            fill(Ff, off);
            if (wave < nj)   // the forward rows to real space: kept in registers
                fft_ip_reg_last<N, true>(z, TW, tl, lane, [&](int r, int, double2 g) __attribute__((always_inline)) {
                    gf[r] = g;
                    return g;
                });
            fill(Fa, off);
            if (wave < nj) {
                // the adjoint rows to real space; the product with the forward
                // rows and with -rate dt applied as the transform stores
                fft_ip_reg_last<N, true>(z, TW, tl, lane, [&](int r, int i, double2 g) __attribute__((always_inline)) {
                    acc[p][r] = fma(omA * g.x, gf[r].x, fma(g.y, gf[r].y, acc[p][r]));
                    const double M = R[i] * dt;
                    return make_double2(-M * g.x, -M * g.y);
                });
                wave_lds_sync();
                fft_ip_reg_last<N, false>(z, TW, tl, lane, [](int, int, double2 g) __attribute__((always_inline)) { return g; });
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < MU; u++) {
                const int m = mm + MS * u;
                if (!jok || m > L) continue;
                double2 zm = aj[m], zc = aj[(N - m) % N];
                zc.y = -zc.y;
                const double2 Aq = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
                const double2 Dd = csub(zm, zc);
                const double2 Bq = make_double2(0.5 * Dd.y, -0.5 * Dd.x);
                Fa[off + (size_t)m * nlat] = make_double2(scale * Aq.x, scale * Aq.y);
                Fa[off + dstr + (size_t)m * nlat] = make_double2(scale * Bq.x, scale * Bq.y);
            }
        }
    }
    if (wave < nj) {
        const int j = j0 + wave;
        const double f = -(2.0 * VH_PI / N) * s.w[j] * dt;
        double* out = rate_h + ((size_t)i3 * nlat + j) * N;
#pragma unroll
        for (int u = 0; u < RU; u++) {
            const int i = lane + 64 * u;
            out[i] += f * (acc[0][u] + acc[1][u] + acc[2][u]);
        }
    }
}

template <int N>
static size_t rows_adj_g_lds(const vh_sht& s) {
    return ((size_t)VH_ROWS_ADJ_RJ * (N + 1) + s.ntw) * sizeof(double2) + (size_t)VH_ROWS_ADJ_RJ * N * sizeof(double);
}
template <int N>
static void rows_adj_g_limit(size_t lds) {
    static std::atomic<size_t> set{0};
    if (lds > 65536 && set.load() < lds) {
        HIPCHECK(hipFuncSetAttribute((const void*)k_t3_rows_adj_g<N, VH_ROWS_ADJ_RJ>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        set.store(lds);
    }
}

void vh_t3_pack_adj(hipStream_t st, const vh_dims& d, const vh_dev& v, const int* e3em, int ne3d, const double* n6,
                    const double* tauh, double2* Xa) {
    if (ne3d <= 0) return;
    dim3 g((d.nk + 255) / 256, ne3d);
    k_t3_pack_adj<<<g, 256, 0, st>>>(d.nk, d.nlm, v, e3em, n6, tauh, Xa);
    HIPCHECK(hipGetLastError());
}

void vh_t3_rows_adj(hipStream_t st, const vh_sht& s, double2* Fa, const double2* Ff, int ne3, const double* rate,
                    double dt, double* rate_h) {
    if (ne3 <= 0) return;
    if (2 * s.L + 1 <= s.nphi && (s.nphi == 256 || s.nphi == 384)) {   // the grouped pass
        const unsigned g = (unsigned)(((ne3 * ((s.nlat + VH_ROWS_ADJ_RJ - 1) / VH_ROWS_ADJ_RJ) + 7) / 8) * 8);
        if (s.nphi == 256) {
            const size_t l = rows_adj_g_lds<256>(s);
            rows_adj_g_limit<256>(l);
            k_t3_rows_adj_g<256, VH_ROWS_ADJ_RJ><<<g, 64 * VH_ROWS_ADJ_RJ, l, st>>>(s, Fa, Ff, ne3, rate, dt, rate_h);
        } else {
            const size_t l = rows_adj_g_lds<384>(s);
            rows_adj_g_limit<384>(l);
            k_t3_rows_adj_g<384, VH_ROWS_ADJ_RJ><<<g, 64 * VH_ROWS_ADJ_RJ, l, st>>>(s, Fa, Ff, ne3, rate, dt, rate_h);
        }
        HIPCHECK(hipGetLastError());
        return;
    }
    const size_t lds = (size_t)6 * s.nphi * sizeof(double2) + (size_t)s.nphi * sizeof(double);
    if (s.nphi > 64 * RA_U) throw vh_error("3-D adjoint: nphi above 512 is not supported by the row pass");
    const unsigned nblk = (unsigned)(((ne3 * s.nlat + 7) / 8) * 8);
    if (s.nphi <= 64 * 6) k_t3_rows_adj<6><<<nblk, 192, lds, st>>>(s, Fa, Ff, ne3, rate, dt, rate_h);
    else k_t3_rows_adj<RA_U><<<nblk, 192, lds, st>>>(s, Fa, Ff, ne3, rate, dt, rate_h);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------- sea-level adjoint
// Elementwise pieces of the transposed sea-level solve, in density form
// (grid adjoints divided by the quadrature weight W = w_j * 2 pi/nphi).
__global__ void k_sle_adj_grid(sle_adj_grid_args a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int j = i / a.nphi;
    const double W = a.w[j] * a.dlon;
    const double C = a.C[i];
    switch (a.mode) {
        case 2:     // y = x - C I(x) / C_int
            a.y[i] = a.x[i] - C * a.I[0] * a.inv_cint;
            break;
        case 6:     // y += x
            a.y[i] += a.x[i];
            break;
        case 4: {   // dig_h = rho_i W t - (rho_i/rho_w) W I(b)/C_int; ice_h = (1 - C) dig_h
            double g = VH_RHO_ICE * W * a.t[i];
            if (a.I) g -= (VH_RHO_ICE / VH_RHO_WATER) * W * a.I[0] * a.inv_cint;
            const double ih = (1.0 - C) * g;
            a.ih0[i] += (1.0 - a.frac) * ih;
            a.ih1[i] += a.frac * ih;
            break;
        }
    }
}

// One block per latitude row: x = rho_w C t (+ rh / W on the first
// iteration), b = x (first) or b += x, and the row sum of x (part[j], for
// the next iteration's integral of x).
__global__ void __launch_bounds__(256) k_sle_adj_row(int nphi, double dlon, const double* __restrict__ w,
                                                     const double* __restrict__ C, const double* __restrict__ t,
                                                     const double* __restrict__ rh, int first, double* x, double* b,
                                                     double* part) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    const double W = w[j] * dlon;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        const size_t id = (size_t)j * nphi + i;
        double v = VH_RHO_WATER * C[id] * t[id];
        if (first && rh) v = rh[id] / W + v;
        x[id] = v;
        b[id] = first ? v : b[id] + v;
        acc += v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[j] = red[0];
}

void vh_sle_adj_row(hipStream_t st, const vh_sht& s, const double* C, const double* t, const double* rh, int first,
                    double* x, double* b, double* part) {
    k_sle_adj_row<<<s.nlat, 256, 0, st>>>(s.nphi, 2.0 * VH_PI / s.nphi, s.w, C, t, rh, first, x, b, part);
    HIPCHECK(hipGetLastError());
}

__device__ __forceinline__ int idx_degree(int idx) {
    int l = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((l + 1) * (l + 2) / 2 <= idx) l++;
    while (l * (l + 1) / 2 > idx) l--;
    return l;
}

// spectral pieces: mode 0: q1 = sigma_h with m > 0 halved;
// mode 1: q *= (gn - gu)_l; mode 3: q = c_m src; mode 4: q += src;
// mode 2: dsurf_h from a = c_m q (U: -a, F: -a/g, V: 0, l >= 1), then
//         q = (gn - gu)_l q + q1
__global__ void k_sle_adj_spec(int nlm, int nk, int mode, const double2* __restrict__ src, double2* q,
                               const double2* __restrict__ q1, const double* __restrict__ gu,
                               const double* __restrict__ gn, double g, double* dsurfh) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlm) return;
    const int l = idx_degree(idx), m = idx - l * (l + 1) / 2;
    const double cm = m == 0 ? 1.0 : 2.0;
    if (mode == 0) {
        const double2 v = src[idx];
        q[idx] = make_double2(v.x / cm, v.y / cm);
    } else if (mode == 3) {   // q = c_m src
        const double2 v = src[idx];
        q[idx] = make_double2(cm * v.x, cm * v.y);
    } else if (mode == 4) {   // q += src
        const double2 v = src[idx], w = q[idx];
        q[idx] = make_double2(w.x + v.x, w.y + v.y);
    } else if (mode == 1) {
        const double f = gn[l] - gu[l];
        const double2 v = q[idx];
        q[idx] = make_double2(f * v.x, f * v.y);
    } else {
        const double2 v = q[idx];
        if (l >= 1) {
            const int k = idx - 1;
            const double2 av = make_double2(cm * v.x, cm * v.y);
            dsurfh[(size_t)(0 * nk + k) * 2 + 0] = -av.x;
            dsurfh[(size_t)(0 * nk + k) * 2 + 1] = -av.y;
            dsurfh[(size_t)(1 * nk + k) * 2 + 0] = -av.x / g;
            dsurfh[(size_t)(1 * nk + k) * 2 + 1] = -av.y / g;
            dsurfh[(size_t)(2 * nk + k) * 2 + 0] = 0.0;
            dsurfh[(size_t)(2 * nk + k) * 2 + 1] = 0.0;
        }
        const double f = gn[l] - gu[l];
        const double2 a1 = q1[idx];
        q[idx] = make_double2(f * v.x + a1.x, f * v.y + a1.y);
    }
}

void vh_sle_adj_grid(hipStream_t st, const sle_adj_grid_args& a) {
    k_sle_adj_grid<<<(a.n + 255) / 256, 256, 0, st>>>(a);
    HIPCHECK(hipGetLastError());
}

void vh_sle_adj_spec(hipStream_t st, int nlm, int nk, int mode, const double2* src, double2* q, const double2* q1,
                     const double* gu, const double* gn, double g, double* dsurfh) {
    k_sle_adj_spec<<<(nlm + 255) / 256, 256, 0, st>>>(nlm, nk, mode, src, q, q1, gu, gn, g, dsurfh);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------- gradient output
// dJ/dpert[e][i]: 3-D elements -ln10 rate3 rate_h; other Maxwell elements the
// element-uniform -ln10 MkPerDt_e Kh / ng (the rate there is MkPerDt 10^-mean(pert)).
__global__ void k_pert_grad(int ne, size_t ng, const int* __restrict__ e3of_mesh, const int* __restrict__ em_of,
                            const double* __restrict__ MkPerDt, const double* __restrict__ rate3,
                            const double* __restrict__ rate_h, const double* __restrict__ Kh, double* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    if (i >= ng || e >= ne) return;
    const double ln10 = 2.302585092994045684;
    double v = 0.0;
    const int i3 = e3of_mesh[e], em = em_of[e];
    if (i3 >= 0) v = -ln10 * rate3[(size_t)i3 * ng + i] * rate_h[(size_t)i3 * ng + i];
    else if (em >= 0) v = -ln10 * MkPerDt[e] * Kh[em] / (double)ng;
    out[(size_t)e * ng + i] = v;
}

void vh_pert_grad(hipStream_t st, int ne, size_t ng, const int* e3of_mesh, const int* em_of, const double* MkPerDt,
                  const double* rate3, const double* rate_h, const double* Kh, double* out) {
    dim3 g((unsigned)((ng + 255) / 256), ne);
    k_pert_grad<<<g, 256, 0, st>>>(ne, ng, e3of_mesh, em_of, MkPerDt, rate3, rate_h, Kh, out);
    HIPCHECK(hipGetLastError());
}
