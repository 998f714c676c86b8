// vh_sle.hip -- (d) grid-space kernels of the sea-level equation
// (src/vilma_sle.f90:124-405) and the ocean function (:407-425).
//
// The ocean/coastline mask must be bit-exact, so every expression that
// feeds it (interpolated ice, ice - d_ice, topo0 - rsl) is evaluated with
// floating-point contraction off, in the reference's operation order.
// Reductions are deterministic two-pass (per-row partials, then one ordered
// weighted sum) so a rerun gives identical bits.
#pragma clang fp contract(off)
#include <algorithm>
#include <cfloat>
#include "vh_internal.h"
#include "vh_sht_dev.h"
#include "vh_sle.h"

__device__ __forceinline__ double ocean(double topo, double ice) {
    return (topo < 0.0 && VH_RHO_ICE * ice < -VH_RHO_WATER * topo) ? 1.0 : 0.0;
}

// ice(t) = ice0 + frac*(ice1 - ice0), d_ice = ice - ice_ref, C0 = O(topo0, ice - d_ice)
// (src/vilma_timestep.f90:314-326, src/vilma_sle.f90:216)
__global__ void k_sle_ice(int n, const double* __restrict__ ice0, const double* __restrict__ ice1, double frac,
                          int interp, const double* __restrict__ ice_ref, const double* __restrict__ topo0,
                          double* ice, double* dice, double* C0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = interp ? ice0[i] + frac * (ice1[i] - ice0[i]) : ice0[i];
    double d = v - ice_ref[i];
    ice[i] = v;
    dice[i] = d;
    C0[i] = ocean(topo0[i], v - d);
}

// Coastline pass head (src/vilma_sle.f90:231-294): C, grounded-ice increment,
// and per-row partial sums of C, d_ice_g, topo0*(C - C0).
__global__ void __launch_bounds__(256) k_sle_outer(int nphi, int first, const double* __restrict__ topo0,
        const double* __restrict__ rsl, const double* __restrict__ ice, const double* __restrict__ dice,
        const double* __restrict__ C0, const double* __restrict__ Cnext, double* C, double* dig,
        double* part /* [3][nlat] */, int nlat) {
    __shared__ double red[3][256];
    const int j = blockIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        double c = first ? ocean(topo0[id] - rsl[id], ice[id]) : Cnext[id];
        double g = ice[id] * (1.0 - c) - (ice[id] - dice[id]) * (1.0 - C0[id]);
        C[id] = c;
        dig[id] = g;
        s0 += c; s1 += g; s2 += topo0[id] * (c - C0[id]);
    }
    red[0][threadIdx.x] = s0; red[1][threadIdx.x] = s1; red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[(size_t)threadIdx.x * nlat + j] = red[threadIdx.x][0];
}

// subgrid sloping-coast load (src/vilma_sle.f90:289-291)
__global__ void k_sle_wcorr(int n, int subgrid, const double* __restrict__ topo0, const double* __restrict__ C,
                            const double* __restrict__ C0, double* wcorr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    wcorr[i] = subgrid ? -VH_RHO_WATER * topo0[i] * (C[i] - C0[i]) : 0.0;
}

// surface load (assemble_load, src/vilma_sle.f90:387-403), fed straight into
// the forward longitude transform
struct InLoad {
    const double *dig, *C, *rsl, *wcorr;
    int nphi;
    __device__ double operator()(int row, int i) const {
        size_t id = (size_t)row * nphi + i;
        return VH_RHO_ICE * dig[id] + VH_RHO_WATER * (C[id] * rsl[id]) + wcorr[id];
    }
};

// response_apply (src/vilma_response.f90:842-864) into two spectra [u; N]
// (degree-major): u = gu*sigma + dUa, N = gn*sigma - dFa/g.
__global__ void k_apply(int nlm, const double2* __restrict__ sigma, const double* __restrict__ gu,
                        const double* __restrict__ gn, const double* __restrict__ dsurf, int nk, double g,
                        int kind, double2* uN) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlm) return;
    int l = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((l + 1) * (l + 2) / 2 <= idx) l++;
    while (l * (l + 1) / 2 > idx) l--;
    double2 s = sigma[idx], u = make_double2(0.0, 0.0), N = make_double2(0.0, 0.0);
    if (kind != 0) {   // 0 = null response
        u = make_double2(gu[l] * s.x, gu[l] * s.y);
        N = make_double2(gn[l] * s.x, gn[l] * s.y);
        if (kind == 2 && l >= 1) {
            int k = idx - 1;
            const double2* ds = (const double2*)dsurf;
            double2 dU = ds[k], dF = ds[nk + k];
            u = make_double2(u.x + dU.x, u.y + dU.y);
            N = make_double2(N.x - dF.x / g, N.y - dF.y / g);
        }
    }
    uN[idx] = u;
    uN[nlm + idx] = N;
}

__global__ void __launch_bounds__(256) k_rowsum_prod(int nphi, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* part) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        acc += a[id] * b[id];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[j] = red[0];
}

__global__ void __launch_bounds__(256) k_sle_flip(int nphi, const double* __restrict__ topo0,
        const double* __restrict__ rsl, const double* __restrict__ ice, const double* __restrict__ C,
        double* Cnext, double* part) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double cnt = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        double cn = ocean(topo0[id] - rsl[id], ice[id]);
        Cnext[id] = cn;
        cnt += (cn != C[id]) ? 1.0 : 0.0;
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[j] = red[0];
}

// per-row counts are whole numbers: their sum is exact in any order
__global__ void __launch_bounds__(256) k_sum_final(const double* __restrict__ part, int nlat, double* out) {
    __shared__ double red[256];
    double a = 0.0;
    for (int j = threadIdx.x; j < nlat; j += blockDim.x) a += part[j];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// z_bed = z_bed_eq - rsl ; ocean function of an arbitrary (topo, ice) pair
__global__ void k_sub(int n, const double* __restrict__ a, const double* __restrict__ b, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}
__global__ void k_ocean(int n, const double* __restrict__ topo, const double* __restrict__ ice, double* C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) C[i] = ocean(topo[i], ice[i]);
}
// bsl integrand h_ice (1 - C) - h_ice_eq (1 - C0) (src/vilma_coupling.f90:572-597)
__global__ void k_bsl_integrand(int n, const double* __restrict__ h, const double* __restrict__ C,
                                const double* __restrict__ heq, const double* __restrict__ C0, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = h[i] * (1.0 - C[i]) - heq[i] * (1.0 - C0[i]);
}

// ------------------------------------------------------------ host wrappers

// ---------------------------------------------------------------- fused inner loop
// The inner fixed point of sle_solve (src/vilma_sle.f90:296-330) in one
// cooperative launch, so no iteration waits on the host:
//   per iteration  B: order m (block m): Legendre analysis of the load's row
//                     spectra -> sigma_lm, response applied (u, N), Legendre
//                     synthesis of u and N -> row spectra F2
//                  C: row pair p (block p, waves 0/1 = rows 2p, 2p+1): inverse
//                     FFT of (u_j, N_j) -> Sraw_j kept in LDS, cs_j = sum C Sraw
//                  D: row pair p: Cs_int = sum_j w_j dlon cs_j (every block, same
//                     order), dphi, rsl update, row maxima of |C drsl|, |C rsl|;
//                     then the next load of rows 2p, 2p+1 and its forward FFT
//   and every block takes the same exit decision from the row maxima.
// Each stage's arithmetic is fixed by the sizes alone (partition of the
// Legendre sums, order of the Cs_int butterfly), so every block reaches the
// same values and a rerun gives identical bits.
// Data handed between blocks (F, F2, cs, part) moves through coherent loads and
// stores (ld_coh / st_coh), so a grid barrier needs no cache write-back or
// invalidate: each wavefront waits for its stores, then one counter per launch
// (bar[0], zeroed before the launch) is raised and awaited. The wait is bounded:
// a block that never arrives raises bar[1] and res[2] instead of a hang.

#define VH_SLE_SPIN_LIMIT 2000000u
// P-table values a thread has in flight in the order stage
#define VH_SLE_JB 16
#define VH_SLE_LB 8

__device__ __forceinline__ void sle_grid_sync(unsigned* bar, unsigned& target, double* res) {
    target += gridDim.x;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);   // this wavefront's coherent stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
            if (++spins > VH_SLE_SPIN_LIMIT) {
                __hip_atomic_fetch_or(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                res[2] = 1.0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

// LDS: [2][nphi] doubles Sraw of the block's row pair (kept from C to D);
// with TabLds the block's Legendre table [j][l] (row stride (L+1)|1 doubles,
// loaded once per launch); the FFT twiddles (s.ntw, loaded once per launch);
// then a union of the row stages' FFT buffers
// (2 wavefronts x 2 nphi complex, the update's copy of w_j and cs_j after the
// first two) and the order stage's folds, partial sums and (sigma, u, N).
__host__ __device__ static size_t sle_tab_doubles(const vh_sht& s) { return (((size_t)s.nh * ((s.L + 1) | 1)) + 1) & ~(size_t)1; }
static size_t sle_inner_lds(const vh_sht& s, bool tab) {
    const size_t rows = (size_t)2 * s.nphi + std::max((size_t)2 * s.nphi, (size_t)s.nlat);
    const size_t order = (size_t)2 * s.nh + std::max(256, s.L + 1) + 3 * (s.L + 1) + 16 * (size_t)s.nh;
    return (size_t)2 * s.nphi * sizeof(double) + (tab ? sle_tab_doubles(s) * sizeof(double) : 0) +
           ((size_t)s.ntw + std::max(rows, order)) * sizeof(double2);
}

template <bool TabLds>
__global__ void __launch_bounds__(256) k_sle_inner(sle_inner_args a) {
    extern __shared__ double2 lds[];
    const vh_sht& s = a.s;
    const int L = s.L, N = s.nphi, nlat = s.nlat, nh = s.nh;
    const int npairs = (nlat + 1) / 2, blk = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    double* srow = (double*)lds;                          // [2][N]
    double* T = srow + 2 * N;                             // [nh][tstride] (TabLds)
    double2* TW = (double2*)(T + (TabLds ? sle_tab_doubles(s) : 0));   // [s.ntw]
    double2* U = TW + s.ntw;                                            // union
    const int tstride = blk <= L ? ((L - blk + 1) | 1) : 1;
    if (TabLds && blk <= L) {
        const int nl = L - blk + 1;
        const double* Pa = s.plm_jl + (size_t)s.mstart[blk] * nh;
        for (int t = tid; t < nh * nl; t += 256) T[(t / nl) * tstride + t % nl] = Pa[t];
    }
    if (blk < npairs)
        for (int t = tid; t < s.ntw; t += 256) TW[t] = s.twp[t];
    __syncthreads();
    __shared__ double redmax[4][4];
    __shared__ double sdphi;
    unsigned target = 0;
    const double scale = 2.0 * VH_PI / N, dlon = 2.0 * VH_PI / N;
    const InLoad load{a.dig, a.C, a.rsl, a.wcorr, N};

    // initial load of rows 2p, 2p+1 -> F (k_fft_fwd<InLoad>)
    if (blk < npairs && wave == 0) {
        const int r0 = 2 * blk, r1 = r0 + 1;
        const bool has1 = r1 < nlat;
        double2* fa = U;
        double2* fb = U + N;
        for (int i0 = 0; i0 < N; i0 += 64 * VH_FFT_U) {
            double v0[VH_FFT_U], v1[VH_FFT_U];
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int i = i0 + lane + 64 * u;
                v0[u] = i < N ? load(r0, i) : 0.0;
                v1[u] = (has1 && i < N) ? load(r1, i) : 0.0;
            }
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int i = i0 + lane + 64 * u;
                if (i < N) fa[i] = make_double2(v0[u], v1[u]);
            }
        }
        wave_lds_sync();
        double2* z = fft_wave(fa, fb, s, false, lane, TW);
        lds_to_pair_spectra<true>(z, L, N, scale, a.F + (size_t)r0 * (L + 1),
                                  has1 ? a.F + (size_t)r1 * (L + 1) : nullptr, lane);
    }
    sle_grid_sync(a.bar, target, a.res);

    for (int ii = 1; ii <= a.n_inner; ii++) {
        // ---- B: order m
        if (blk <= L) {
            const int m = blk, nl = L - m + 1;
            double2* E = U;                       // [nh]
            double2* O = E + nh;                  // [nh]
            double2* red = O + nh;                // [max(256, L+1)]
            double2* Q = red + max(256, L + 1);   // [3][L+1]: sigma, u, N
            double2* red2 = Q + 3 * (L + 1);      // [4][nh][2][2]
            for (int j = tid; j < nh; j += 256) {
                const int js = nlat - 1 - j;
                const double2 fn = ld_coh(a.F + (size_t)j * (L + 1) + m);
                const double2 fs = js != j ? ld_coh(a.F + (size_t)js * (L + 1) + m) : make_double2(0.0, 0.0);
                const double w = s.w[j];
                E[j] = make_double2(w * (fn.x + fs.x), w * (fn.y + fs.y));
                O[j] = make_double2(w * (fn.x - fs.x), w * (fn.y - fs.y));
            }
            __syncthreads();
            // analysis: thread = (latitude part, degree), S parts of the north
            // rows, S = min(4, 256 / nl) so small orders still fill the block;
            // partial sums added in part order
            const double* Pa = s.plm_jl + (size_t)s.mstart[m] * nh;
            {
                const int S = max(1, min(4, 256 / nl)), jq = (nh + S - 1) / S;
                for (int t = tid; t < S * nl; t += 256) {
                    const int part = t / nl, il = t % nl;
                    const int ja = part * jq, jb = min(nh, ja + jq);
                    const double2* G = (il & 1) ? O : E;
                    double ax = 0.0, ay = 0.0;
                    for (int j0 = ja; j0 < jb; j0 += VH_SLE_JB) {
                        double pv[VH_SLE_JB];
#pragma unroll
                        for (int u = 0; u < VH_SLE_JB; u++) {
                            if constexpr (TabLds) pv[u] = j0 + u < jb ? T[(j0 + u) * tstride + il] : 0.0;
                            else pv[u] = j0 + u < jb ? Pa[(size_t)(j0 + u) * nl + il] : 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < VH_SLE_JB; u++)
                            if (j0 + u < jb) { ax = fma(pv[u], G[j0 + u].x, ax); ay = fma(pv[u], G[j0 + u].y, ay); }
                    }
                    red[t] = make_double2(ax, ay);
                }
                __syncthreads();
                for (int il = tid; il < nl; il += 256) {
                    const int l = m + il;
                    double2 r = red[il];
                    for (int k = 1; k < S; k++) r = make_double2(r.x + red[k * nl + il].x, r.y + red[k * nl + il].y);
                    a.sigma[(size_t)l * (l + 1) / 2 + m] = r;
                    Q[il] = r;
                }
                __syncthreads();
            }
            // response applied to the load (response_apply, src/vilma_response.f90:842-864):
            // u = gu sigma + dUa, N = gn sigma - dFa/g
            for (int il = tid; il < nl; il += 256) {
                const int l = m + il, idx = l * (l + 1) / 2 + m;
                double2 u = make_double2(0.0, 0.0), Nv = u;
                if (a.kind != 0) {
                    const double2 sg = Q[il];
                    u = make_double2(a.gu[l] * sg.x, a.gu[l] * sg.y);
                    Nv = make_double2(a.gn[l] * sg.x, a.gn[l] * sg.y);
                    if (a.kind == 2 && l >= 1) {
                        const double2 dU = ((const double2*)a.dsurf)[idx - 1];
                        const double2 dF = ((const double2*)a.dsurf)[a.nk + idx - 1];
                        u = make_double2(u.x + dU.x, u.y + dU.y);
                        Nv = make_double2(Nv.x - dF.x / a.g, Nv.y - dF.y / a.g);
                    }
                }
                Q[(L + 1) + il] = u;
                Q[2 * (L + 1) + il] = Nv;
            }
            __syncthreads();
            // synthesis of u, N (two fields, rows interleaved as the inverse
            // transform pairs them): thread = (degree part, north latitude),
            // S parts of even length, S = min(4, 256 / nh); parity sums kept
            // apart for the south rows, parts added in order
            const double* Ps = s.plm_lj + (size_t)s.mstart[m] * nh;
            const int S = max(1, min(4, 256 / nh));
            const int lq = ((nl + 2 * S - 1) / (2 * S)) * 2;
            for (int t = tid; t < S * nh; t += 256) {
                const int gq = t / nh, j = t % nh;
                const int la = min(nl, gq * lq), lb = min(nl, la + lq);
                double2 e[2], o[2];
                e[0] = e[1] = o[0] = o[1] = make_double2(0.0, 0.0);
                for (int l0 = la; l0 < lb; l0 += 2 * VH_SLE_LB) {
                    double pe[VH_SLE_LB], po[VH_SLE_LB];
#pragma unroll
                    for (int u = 0; u < VH_SLE_LB; u++) {
                        const int il = l0 + 2 * u;
                        if constexpr (TabLds) {
                            pe[u] = il < lb ? T[j * tstride + il] : 0.0;
                            po[u] = il + 1 < lb ? T[j * tstride + il + 1] : 0.0;
                        } else {
                            pe[u] = il < lb ? Ps[(size_t)il * nh + j] : 0.0;
                            po[u] = il + 1 < lb ? Ps[(size_t)(il + 1) * nh + j] : 0.0;
                        }
                    }
#pragma unroll
                    for (int u = 0; u < VH_SLE_LB; u++) {
                        const int il = l0 + 2 * u;
                        if (il >= lb) break;
                        const bool two = il + 1 < lb;
#pragma unroll
                        for (int f = 0; f < 2; f++) {
                            const double2 c0 = Q[(f + 1) * (L + 1) + il];
                            e[f].x = fma(c0.x, pe[u], e[f].x); e[f].y = fma(c0.y, pe[u], e[f].y);
                            if (two) {
                                const double2 c1 = Q[(f + 1) * (L + 1) + il + 1];
                                o[f].x = fma(c1.x, po[u], o[f].x); o[f].y = fma(c1.y, po[u], o[f].y);
                            }
                        }
                    }
                }
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    double2* r = red2 + (((size_t)gq * nh + j) * 2 + f) * 2;
                    r[0] = e[f];
                    r[1] = o[f];
                }
            }
            __syncthreads();
            for (int t = tid; t < 2 * nh; t += 256) {
                const int jj = t / 2, f = t % 2, js = nlat - 1 - jj;
                double2 ev = make_double2(0.0, 0.0), od = ev;
                for (int k = 0; k < S; k++) {
                    const double2* r = red2 + (((size_t)k * nh + jj) * 2 + f) * 2;
                    ev = make_double2(ev.x + r[0].x, ev.y + r[0].y);
                    od = make_double2(od.x + r[1].x, od.y + r[1].y);
                }
                st_coh(a.F2 + ((size_t)jj * 2 + f) * (L + 1) + m, make_double2(ev.x + od.x, ev.y + od.y));
                if (js != jj) st_coh(a.F2 + ((size_t)js * 2 + f) * (L + 1) + m, make_double2(ev.x - od.x, ev.y - od.y));
            }
        }
        sle_grid_sync(a.bar, target, a.res);

        // ---- C: inverse FFT of (u_j, N_j) -> Sraw_j in LDS, cs_j
        if (blk < npairs && wave < 2) {
            const int j = 2 * blk + wave;
            if (j < nlat) {
                double2* fa = U + (size_t)wave * 2 * N;
                double2* fb = fa + N;
                // the row's mask and rotation term in flight over the transform
                double cpre[VH_FFT_U], spre[VH_FFT_U];
#pragma unroll
                for (int u = 0; u < VH_FFT_U; u++) {
                    const int i = lane + 64 * u;
                    cpre[u] = i < N ? a.C[(size_t)j * N + i] : 0.0;
                    spre[u] = (a.srot && i < N) ? a.srot[(size_t)j * N + i] : 0.0;
                }
                pair_spectra_to_lds<true>(a.F2 + (size_t)(2 * j) * (L + 1), a.F2 + (size_t)(2 * j + 1) * (L + 1), L, N,
                                          fa, lane);
                double2* z = fft_wave(fa, fb, s, true, lane, TW);
                double rs = 0.0;
                for (int i0 = 0; i0 < N; i0 += 64 * VH_FFT_U) {
#pragma unroll
                    for (int u = 0; u < VH_FFT_U; u++) {
                        const int i = i0 + lane + 64 * u;
                        if (i >= N) break;
                        const size_t id = (size_t)j * N + i;
                        double v = z[i].y - z[i].x;
                        if (a.srot) v = v + (i0 == 0 ? spre[u] : a.srot[id]);
                        srow[wave * N + i] = v;
                        rs += (i0 == 0 ? cpre[u] : a.C[id]) * v;
                    }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) rs += __shfl_xor(rs, off);
                if (lane == 0) st_coh(a.cs + j, rs);
            }
        }
        sle_grid_sync(a.bar, target, a.res);

        // ---- D: dphi, rsl update, row maxima; the next load and its FFT
        if (blk < npairs) {
            const int r0 = 2 * blk;
            double* wcs = (double*)(U + 2 * N);   // [nlat]: w_j cs_j dlon
            for (int q = tid; q < nlat; q += 256) wcs[q] = s.w[q] * ld_coh(a.cs + q) * dlon;
            __syncthreads();
            if (wave == 0) {
                // Cs_int: per-lane sums in latitude order, then a butterfly
                // (every lane, every block: the same bits)
                double total = 0.0;
                for (int q = lane; q < nlat; q += 64) total += wcs[q];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off);
                const double rho_ratio = VH_RHO_ICE / VH_RHO_WATER;
                const double ice_int = -rho_ratio * a.sc[1];
                const double zeta = a.subgrid ? a.sc[2] : 0.0;
                if (lane == 0) {
                    sdphi = (ice_int - total + zeta) / a.sc[0];
                    if (blk == 0) { a.res[0] = (double)ii; a.res[1] = sdphi; }
                }
            }
            __syncthreads();
            const double dphi = sdphi;
            const bool has1 = r0 + 1 < nlat;
            double2* fa = U;
            double2* fb = U + N;
            double dm[2] = {0.0, 0.0}, sm[2] = {0.0, 0.0};
#pragma unroll
            for (int w = 0; w < 2; w++) {
                const int j = r0 + w;
                if (j >= nlat) {
                    for (int i = tid; i < N; i += 256) fa[i].y = 0.0;
                    continue;
                }
                for (int i = tid; i < N; i += 256) {
                    const size_t id = (size_t)j * N + i;
                    const double rn = srow[w * N + i] + dphi;
                    const double c = a.C[id];
                    dm[w] = fmax(dm[w], fabs(c * (rn - a.rsl[id])));
                    a.rsl[id] = rn;
                    sm[w] = fmax(sm[w], fabs(c * rn));
                    const double ld = VH_RHO_ICE * a.dig[id] + VH_RHO_WATER * (c * rn) + a.wcorr[id];
                    if (w == 0) fa[i].x = ld; else fa[i].y = ld;
                }
            }
#pragma unroll
            for (int q = 0; q < 2; q++)
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    dm[q] = fmax(dm[q], __shfl_xor(dm[q], off));
                    sm[q] = fmax(sm[q], __shfl_xor(sm[q], off));
                }
            if (lane == 0) { redmax[wave][0] = dm[0]; redmax[wave][1] = dm[1]; redmax[wave][2] = sm[0]; redmax[wave][3] = sm[1]; }
            __syncthreads();
            if (tid < 4) {
                const int w = tid & 1;
                if (r0 + w < nlat) {
                    const double v = fmax(fmax(redmax[0][tid], redmax[1][tid]), fmax(redmax[2][tid], redmax[3][tid]));
                    st_coh(a.part + (tid < 2 ? 0 : nlat) + r0 + w, v);
                }
            }
            if (wave == 0) {
                double2* z = fft_wave(fa, fb, s, false, lane, TW);
                lds_to_pair_spectra<true>(z, L, N, scale, a.F + (size_t)r0 * (L + 1),
                                          has1 ? a.F + (size_t)(r0 + 1) * (L + 1) : nullptr, lane);
            }
        }
        sle_grid_sync(a.bar, target, a.res);

        // ---- exit test, the same in every block (max is exact in any order)
        double dmx = 0.0, smx = 0.0;
        for (int q = tid; q < nlat; q += 256) {
            dmx = fmax(dmx, ld_coh(a.part + q));
            smx = fmax(smx, ld_coh(a.part + nlat + q));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) { dmx = fmax(dmx, __shfl_xor(dmx, off)); smx = fmax(smx, __shfl_xor(smx, off)); }
        __syncthreads();
        if (lane == 0) { redmax[wave][0] = dmx; redmax[wave][1] = smx; }
        __syncthreads();
        dmx = fmax(fmax(redmax[0][0], redmax[1][0]), fmax(redmax[2][0], redmax[3][0]));
        smx = fmax(fmax(redmax[0][1], redmax[1][1]), fmax(redmax[2][1], redmax[3][1]));
        if (dmx <= a.tol * fmax(smx, DBL_MIN)) break;
    }
}

__global__ void k_sle_inner_prep(unsigned* bar, double* res) {
    if (threadIdx.x < 2) bar[threadIdx.x] = 0u;
    if (threadIdx.x < 3) res[threadIdx.x] = 0.0;
}

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0) {
    k_sle_ice<<<nb256(n), 256, 0, st>>>(n, ice0, ice1, frac, interp, ice_ref, topo0, ice, dice, C0);
}
void sle_outer(hipStream_t st, const vh_sht& s, int first, const double* topo0, const double* rsl,
               const double* ice, const double* dice, const double* C0, const double* Cnext, double* C,
               double* dig, double* part, double* sc) {
    k_sle_outer<<<s.nlat, 256, 0, st>>>(s.nphi, first, topo0, rsl, ice, dice, C0, Cnext, C, dig, part, s.nlat);
    vh_weighted_total(st, s, part, 3, sc);
}
void sle_wcorr(hipStream_t st, int n, int subgrid, const double* topo0, const double* C, const double* C0, double* w) {
    k_sle_wcorr<<<nb256(n), 256, 0, st>>>(n, subgrid, topo0, C, C0, w);
}
void sle_load_analysis(hipStream_t st, const vh_sht& s, const double* dig, const double* C, const double* rsl,
                       const double* wcorr, double2* F, double2* load_lm) {
    vh_fft_forward_t(st, s, InLoad{dig, C, rsl, wcorr, s.nphi}, s.nlat, F);
    vh_legendre_analysis(st, s, F, 1, load_lm);
}
void vh_apply_only(hipStream_t st, int nlm, const double2* sigma, const double* gu, const double* gn,
                   const double* dsurf, int nk, double g, int kind, double2* uN) {
    k_apply<<<nb256(nlm), 256, 0, st>>>(nlm, sigma, gu, gn, dsurf, nk, g, kind, uN);
}
void sle_cs(hipStream_t st, const vh_sht& s, const double* C, const double* f, double* part, double* out) {
    k_rowsum_prod<<<s.nlat, 256, 0, st>>>(s.nphi, C, f, part);
    vh_weighted_total(st, s, part, 1, out);
}
void sle_flip(hipStream_t st, const vh_sht& s, const double* topo0, const double* rsl, const double* ice,
              const double* C, double* Cnext, double* part, double* out) {
    k_sle_flip<<<s.nlat, 256, 0, st>>>(s.nphi, topo0, rsl, ice, C, Cnext, part);
    k_sum_final<<<1, 256, 0, st>>>(part, s.nlat, out);
}
void grid_sub(hipStream_t st, int n, const double* a, const double* b, double* out) {
    k_sub<<<nb256(n), 256, 0, st>>>(n, a, b, out);
}
void grid_ocean(hipStream_t st, int n, const double* topo, const double* ice, double* C) {
    k_ocean<<<nb256(n), 256, 0, st>>>(n, topo, ice, C);
}
void grid_bsl_integrand(hipStream_t st, int n, const double* h, const double* C, const double* heq,
                        const double* C0, double* out) {
    k_bsl_integrand<<<nb256(n), 256, 0, st>>>(n, h, C, heq, C0, out);
}

void sle_inner(hipStream_t st, const sle_inner_args& a) {
    const vh_sht& s = a.s;
    const int grid = std::max(s.L + 1, (s.nlat + 1) / 2);
    // the Legendre table of each order stays in LDS when it fits beside the rest
    const bool tab = sle_inner_lds(s, true) <= (size_t)150 * 1024;
    const size_t lds = sle_inner_lds(s, tab);
    const void* fn = tab ? (const void*)k_sle_inner<true> : (const void*)k_sle_inner<false>;
    // resident-block capacity, checked once per (device, LDS size)
    static thread_local int cap_dev = -1;
    static thread_local size_t cap_lds = 0;
    static thread_local long cap = 0;
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (cap_dev != dev || cap_lds != lds) {
        int coop = 0, per_cu = 0, ncu = 0;
        HIPCHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        if (!coop) throw vh_error("sle: the device does not support cooperative launches");
        if (lds > 65536)
            HIPCHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds));
        HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        cap_dev = dev; cap_lds = lds; cap = (long)per_cu * ncu;
    }
    if (cap < grid)
        throw vh_error("sle: the fused inner loop needs " + std::to_string(grid) + " resident blocks, the device holds " +
                       std::to_string(cap));
    k_sle_inner_prep<<<1, 64, 0, st>>>(a.bar, a.res);
    HIPCHECK(hipGetLastError());
    sle_inner_args arg = a;
    void* params[] = {&arg};
    HIPCHECK(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), params, (unsigned)lds, st));
}
