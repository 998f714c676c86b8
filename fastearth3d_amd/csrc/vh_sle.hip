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
#include <mutex>
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
                          double* ice, double* dice, double* C0, double* wcorr, double* res, int nres) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nres) res[i] = 0.0;
    if (i >= n) return;
    wcorr[i] = 0.0;
    double v = interp ? ice0[i] + frac * (ice1[i] - ice0[i]) : ice0[i];
    double d = v - ice_ref[i];
    ice[i] = v;
    dice[i] = d;
    C0[i] = ocean(topo0[i], v - d);
}

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
// cooperative launch, so no iteration waits on the host. Row blocks own a
// latitude pair p (rows 2p, 2p+1), order blocks an order m.
//   start  rows:   FC = FFT(C), F = FFT(load(rsl))            [barrier]
//          orders: F -> sigma -> (u, N) -> F2                  [barrier]
//   per iteration
//          rows:   inverse FFT of (u_j, N_j) -> Sraw_j (kept in LDS),
//                  cs_j = sum C Sraw; base = rho_i dI + rho_w C Sraw + wcorr,
//                  Fb = FFT(base)                              [barrier]
//          all:    Cs_int = sum_j w_j dlon cs_j, dphi
//          rows:   rsl = Sraw + dphi, row maxima of |C drsl|, |C rsl|
//          orders: the next load's spectrum Fb + rho_w dphi FC (the transform
//                  of rho_i dI + rho_w C rsl + wcorr, by linearity) ->
//                  sigma -> (u, N) -> F2                        [barrier]
//          all:    the same exit decision from the row maxima.
// On exit sigma is the spectrum of the load of the final rsl: the closing
// analysis of the solve. Each stage's arithmetic is fixed by the sizes alone
// (partition of the Legendre sums, order of the Cs_int butterfly), so every
// block reaches the same values and a rerun gives identical bits.
// Data handed between blocks (F, FC, F2, cs, part, part3) moves through
// agent-scope loads and stores (ld_coh / st_coh), which bypass the XCDs'
// non-coherent caches, so a grid barrier needs no cache write-back or
// invalidate: each wavefront waits for its stores, then the block arrives in
// the counter of its group blockIdx % 8; the last block of a group arrives in
// the top counter, and the block that completes the top count writes the
// barrier number into eight release words, which the blocks of each group poll
// (all counters and words on pages of their own, so the arrivals of a hundred
// blocks that finish together do not queue on one address). The counts run on
// across launches from the barrier number kept in bar[2]. Data from the
// kernels before the launch is acquired once at entry. The wait is bounded: a
// block that never arrives raises bar[1] and res[2] instead of a hang.

#define VH_SLE_SPIN_LIMIT 2000000u
// P-table values a thread has in flight in the order stage
#define VH_SLE_JB 8
#define VH_SLE_LB 4
// grid values per lane a row wavefront loads ahead of its transform
#define VH_SLE_PU 6
// longitudes per thread and row whose inputs the pass head loads together
#define VH_SLE_PH 2

__device__ __forceinline__ void sle_grid_sync(unsigned* bar, unsigned& nbar, double* res) {
    nbar += 1u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);   // this wavefront's stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = blockIdx.x & 7u, nb = gridDim.x;
        const unsigned ng = (nb - g + 7u) / 8u, ngroups = nb < 8u ? nb : 8u;
        unsigned* flag = bar + VH_SLE_FLAG_STRIDE * (1 + g);
        // arrival in the group counter; the group's last block arrives in the
        // top counter, and the last group's block raises the release words
        if (__hip_atomic_fetch_add(bar + VH_SLE_FLAG_STRIDE * (10 + g), 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) + 1u == ng * nbar &&
            __hip_atomic_fetch_add(bar + VH_SLE_FLAG_STRIDE * 9, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u ==
                ngroups * nbar) {
#pragma unroll
            for (int k = 0; k < 8; k++)
                __hip_atomic_store(bar + VH_SLE_FLAG_STRIDE * (1 + k), nbar, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned spins = 0;
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nbar) {
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

// LDS of a row block: Sraw [2][nphi] doubles (kept across a barrier), the FFT
// twiddles (s.ntw complex) and two FFT buffer pairs (2 x 2 nphi complex; the
// pass head's row sums use the same space).
// Of an order block: with TabLds its Legendre
// table [j][l] (row stride (L+1)|1 doubles, loaded once per launch), then the
// folds, partial sums and (sigma, u, N).
__host__ __device__ static size_t sle_tab_doubles(const vh_sht& s) {
    return (((size_t)s.nh * ((s.L + 1) | 1)) + 1) & ~(size_t)1;
}
static size_t sle_solve_lds(const vh_sht& s, bool tab) {
    const size_t rows = (size_t)2 * s.nphi * sizeof(double) + (size_t)s.ntw * sizeof(double2) +
                        std::max((size_t)4 * s.nphi * sizeof(double2), (size_t)6 * 256 * sizeof(double));
    const size_t order = (tab ? sle_tab_doubles(s) * sizeof(double) : 0) +
                         ((size_t)2 * s.nh + std::max(256, s.L + 1) + 3 * (s.L + 1) + 16 * (size_t)s.nh) * sizeof(double2);
    return std::max(rows, order);
}

// wavefront sum of v over the block's lanes 0..63 (butterfly: every lane the same bits)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// dphi = (ice_int - Cs_int + zeta) / C_int with Cs_int from the row sums cs
// (wavefront 0; result in *sd for the block)
__device__ __forceinline__ void sle_dphi(const sle_solve_args& a, const double* tot, int nlat, double dlon, int wave,
                                         int lane, double* sd, double* scs) {
    if (wave == 0) {
        double total = 0.0;
        for (int q = lane; q < nlat; q += 64) total += a.s.w[q] * ld_coh(a.cs + q) * dlon;
        total = wave_sum(total);
        if (lane == 0) {
            *scs = total;
            const double rho_ratio = VH_RHO_ICE / VH_RHO_WATER;
            const double ice_int = -rho_ratio * tot[1];
            const double zeta = a.subgrid ? tot[2] : 0.0;
            *sd = (ice_int - total + zeta) / tot[0];
        }
    }
}

// Order stage of block m: fold of the load spectrum (F, plus kdp * FC when
// kdp != 0) -> Legendre analysis -> sigma -> response -> Legendre synthesis of
// (u, N) -> F2. fn/fs: this thread's north / south spectrum values (j = tid < nh)
// already loaded by the caller.
template <bool TabLds>
__device__ __forceinline__ void sle_order_stage(const sle_solve_args& a, int m, const double* T, int tstride,
                                                double2* U, double2 fn, double2 fs) {
    const vh_sht& s = a.s;
    const int L = s.L, nlat = s.nlat, nh = s.nh, nl = L - m + 1, tid = threadIdx.x;
    (void)nlat;
    double2* E = U;                       // [nh]
    double2* O = E + nh;                  // [nh]
    double2* red = O + nh;                // [max(256, L+1)]
    double2* Q = red + max(256, L + 1);   // [3][L+1]: sigma, u, N
    double2* red2 = Q + 3 * (L + 1);      // [4][nh][2][2]
    if (tid < nh) {   // nh <= 256 (checked at launch)
        const double w = s.w[tid];
        E[tid] = make_double2(w * (fn.x + fs.x), w * (fn.y + fs.y));
        O[tid] = make_double2(w * (fn.x - fs.x), w * (fn.y - fs.y));
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
            int j = ja;
            for (; j + VH_SLE_JB <= jb; j += VH_SLE_JB) {   // whole batches: loads without predicates
                double pv[VH_SLE_JB];
                double2 gv[VH_SLE_JB];
#pragma unroll
                for (int u = 0; u < VH_SLE_JB; u++) {
                    if constexpr (TabLds) pv[u] = T[(j + u) * tstride + il];
                    else pv[u] = Pa[(size_t)(j + u) * nl + il];
                    gv[u] = G[j + u];
                }
#pragma unroll
                for (int u = 0; u < VH_SLE_JB; u++) { ax = fma(pv[u], gv[u].x, ax); ay = fma(pv[u], gv[u].y, ay); }
            }
            for (; j < jb; j++) {
                double pj;
                if constexpr (TabLds) pj = T[j * tstride + il];
                else pj = Pa[(size_t)j * nl + il];
                ax = fma(pj, G[j].x, ax); ay = fma(pj, G[j].y, ay);
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
        int il = la;
        for (; il + 2 * VH_SLE_LB <= lb; il += 2 * VH_SLE_LB) {   // whole batches: loads without predicates
            double pe[VH_SLE_LB], po[VH_SLE_LB];
            double2 qe[2][VH_SLE_LB], qo[2][VH_SLE_LB];
#pragma unroll
            for (int u = 0; u < VH_SLE_LB; u++) {
                if constexpr (TabLds) {
                    pe[u] = T[j * tstride + il + 2 * u];
                    po[u] = T[j * tstride + il + 2 * u + 1];
                } else {
                    pe[u] = Ps[(size_t)(il + 2 * u) * nh + j];
                    po[u] = Ps[(size_t)(il + 2 * u + 1) * nh + j];
                }
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    qe[f][u] = Q[(f + 1) * (L + 1) + il + 2 * u];
                    qo[f][u] = Q[(f + 1) * (L + 1) + il + 2 * u + 1];
                }
            }
#pragma unroll
            for (int u = 0; u < VH_SLE_LB; u++)
#pragma unroll
                for (int f = 0; f < 2; f++) {
                    e[f].x = fma(qe[f][u].x, pe[u], e[f].x); e[f].y = fma(qe[f][u].y, pe[u], e[f].y);
                    o[f].x = fma(qo[f][u].x, po[u], o[f].x); o[f].y = fma(qo[f][u].y, po[u], o[f].y);
                }
        }
        for (; il < lb; il += 2) {
            const bool two = il + 1 < lb;
            double p0, p1 = 0.0;
            if constexpr (TabLds) {
                p0 = T[j * tstride + il];
                if (two) p1 = T[j * tstride + il + 1];
            } else {
                p0 = Ps[(size_t)il * nh + j];
                if (two) p1 = Ps[(size_t)(il + 1) * nh + j];
            }
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const double2 c0 = Q[(f + 1) * (L + 1) + il];
                e[f].x = fma(c0.x, p0, e[f].x); e[f].y = fma(c0.y, p0, e[f].y);
                if (two) {
                    const double2 c1 = Q[(f + 1) * (L + 1) + il + 1];
                    o[f].x = fma(c1.x, p1, o[f].x); o[f].y = fma(c1.y, p1, o[f].y);
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
        // F2 rows q = 2j + field in groups of 4 (one latitude pair), order-major
        // inside a group: [q/4][m][q%4] (64 contiguous bytes per order written,
        // one contiguous 64 (L+1) bytes per latitude pair read)
        const int qn = 2 * jj + f, qs = 2 * js + f;
        st_coh(a.F2 + ((size_t)(qn >> 2) * (L + 1) + m) * 4 + (qn & 3), make_double2(ev.x + od.x, ev.y + od.y));
        if (js != jj)
            st_coh(a.F2 + ((size_t)(qs >> 2) * (L + 1) + m) * 4 + (qs & 3), make_double2(ev.x - od.x, ev.y - od.y));
    }
}

// The kernel's longitude FFT: the compile-time transforms where registers
// allow (MinW 1); with two wavefronts per SIMD (256 registers) the run-time
// one, whose smaller register need avoids spills.
template <bool INV, int MinW>
__device__ __forceinline__ double2* sle_fft(double2* a, double2* b, const vh_sht& s, int lane, const double2* tw) {
    if constexpr (MinW == 1) return fft_wave_t<INV>(a, b, s, lane, tw);
    else return fft_wave(a, b, s, INV, lane, tw);
}

// MinW: wavefronts per SIMD the build must allow; 2 halves the registers
// (with spills) so that two blocks share a CU when the grid (latitude pairs +
// orders, 386 blocks at lmax 192) exceeds the CU count.
template <bool TabLds, int MinW>
__global__ void __launch_bounds__(256, MinW) k_sle_solve(sle_solve_args a) {
    extern __shared__ double2 lds[];
    const vh_sht& s = a.s;
    const int L = s.L, N = s.nphi, nlat = s.nlat, nh = s.nh;
    const int npairs = (nlat + 1) / 2, blk = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const bool is_row = blk < npairs;
    const int p = blk, m = blk - npairs;
    __shared__ double redmax[4][4];
    __shared__ double sdphi, scs;
    // Inputs written by the kernels before this launch (ice, drift, rsl, ...)
    // were written by other XCDs: an agent-scope acquire invalidates this
    // CU's and XCD's cached copies before anything is read, whatever fences
    // the launch itself carries; a release at the end publishes the outputs.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    unsigned target = a.bar[2];   // barriers passed, counted on across launches
    const double scale = 2.0 * VH_PI / N, dlon = 2.0 * VH_PI / N;
    __shared__ double stot[4];    // pass totals: C_int, int d_ice_g, int topo0 (C - C0); flips
    double n_iters = 0.0, n_passes = 0.0, last_ice = 0.0, last_zeta = 0.0, last_cs = 0.0;
    // row-block LDS
    double* srow = (double*)lds;                          // [2][N]
    double2* TW = (double2*)(srow + 2 * N);               // [s.ntw]
    double2* RB = TW + s.ntw;                             // [2][2N] FFT buffer pairs
    // order-block LDS
    double* T = (double*)lds;                             // [nh][tstride] (TabLds)
    double2* U = (double2*)(T + (TabLds ? sle_tab_doubles(s) : 0));
    const int tstride = is_row ? 1 : ((L - m + 1) | 1);
    if (is_row) {
        for (int t = tid; t < s.ntw; t += 256) TW[t] = s.twp[t];
    }
    // the order blocks' Legendre table is loaded after the first barrier of the
    // first pass, while the row blocks run the pass's start transforms
    auto load_table = [&]() {
        // 8 table values per thread in flight
        const int nl = L - m + 1, nt = nh * nl;
        const double* Pa = s.plm_jl + (size_t)s.mstart[m] * nh;
        for (int t0 = tid; t0 < nt; t0 += 256 * 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = t0 + 256 * u < nt ? Pa[t0 + 256 * u] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int t = t0 + 256 * u;
                if (t < nt) T[(t / nl) * tstride + t % nl] = v[u];
            }
        }
    };
    __syncthreads();

  for (int io = 1; io <= a.n_outer; io++) {
    // ---- coastline pass head (src/vilma_sle.f90:231-294), rows: the mask C
    // (first pass: from topo0 - rsl; later: the previous pass's C_next; fixed
    // ocean: C0), the grounded-ice increment d_ice_g and, per row, the sums of
    // C, d_ice_g and topo0 (C - C0) (a 256-thread tree per row)
    if (is_row) {
        double* red = (double*)RB;   // [6][256]: (quantity, row of the pair)
        double sp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        // the inputs of a thread's points of both rows (VH_SLE_PH per row) are
        // loaded before any is used; sums per thread in the same (row, longitude) order
        for (int i0 = 0; i0 < N; i0 += 256 * VH_SLE_PH) {
            double vC0[2][VH_SLE_PH], vice[2][VH_SLE_PH], vdice[2][VH_SLE_PH], vtopo[2][VH_SLE_PH],
                vx[2][VH_SLE_PH];
#pragma unroll
            for (int w = 0; w < 2; w++)
#pragma unroll
                for (int u = 0; u < VH_SLE_PH; u++) {
                    const int j = 2 * p + w, i = i0 + tid + 256 * u;
                    const bool ok = j < nlat && i < N;
                    const size_t id = ok ? (size_t)j * N + i : 0;
                    vC0[w][u] = ok ? a.C0[id] : 0.0;
                    vice[w][u] = ok ? a.ice[id] : 0.0;
                    vdice[w][u] = ok ? a.dice[id] : 0.0;
                    vtopo[w][u] = ok ? a.topo0[id] : 0.0;
                    vx[w][u] = (!ok || a.fixed_ocean) ? 0.0 : (io == 1 ? a.rsl[id] : a.Cn[id]);
                }
#pragma unroll
            for (int w = 0; w < 2; w++)
#pragma unroll
                for (int u = 0; u < VH_SLE_PH; u++) {
                    const int j = 2 * p + w, i = i0 + tid + 256 * u;
                    if (j >= nlat || i >= N) continue;
                    const size_t id = (size_t)j * N + i;
                    const double c = a.fixed_ocean ? vC0[w][u]
                                                   : (io == 1 ? ocean(vtopo[w][u] - vx[w][u], vice[w][u]) : vx[w][u]);
                    const double g = vice[w][u] * (1.0 - c) - (vice[w][u] - vdice[w][u]) * (1.0 - vC0[w][u]);
                    a.C[id] = c;
                    a.dig[id] = g;
                    sp[3 * w + 0] += c; sp[3 * w + 1] += g; sp[3 * w + 2] += vtopo[w][u] * (c - vC0[w][u]);
                }
        }
#pragma unroll
        for (int q = 0; q < 6; q++) red[q * 256 + tid] = sp[q];
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (tid < st)
#pragma unroll
                for (int q = 0; q < 6; q++) red[q * 256 + tid] += red[q * 256 + tid + st];
            __syncthreads();
        }
        if (tid < 6) {
            const int w = tid / 3, q = tid % 3, j = 2 * p + w;
            if (j < nlat) st_coh(a.part3 + (size_t)q * nlat + j, red[tid * 256]);
        }
    }
    sle_grid_sync(a.bar, target, a.res);
    // every block: the pass totals sum_j w_j dlon part_j (latitude order, as
    // k_weighted_total): the terms are loaded by all threads into the FFT /
    // order-stage scratch, then one lane per quantity adds them in order
    {
        double* tw3 = is_row ? (double*)RB : (double*)U;   // [3][nlat], free here
        for (int t = tid; t < 3 * nlat; t += 256) {
            const int q = t % nlat;
            tw3[t] = s.w[q] * ld_coh(a.part3 + t) * dlon;
        }
        __syncthreads();
        if (tid < 3) {   // in latitude order, 8 terms read ahead of their sums
            double total = 0.0;
            int q = 0;
            for (; q + 8 <= nlat; q += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = tw3[tid * nlat + q + u];
#pragma unroll
                for (int u = 0; u < 8; u++) total += v[u];
            }
            for (; q < nlat; q++) total += tw3[tid * nlat + q];
            stot[tid] = total;
        }
        __syncthreads();
    }
    if constexpr (TabLds) {
        if (io == 1 && !is_row) load_table();   // read by the order stage after the next barrier
    }
    const double C_int = stot[0];
    if (blk == 0 && tid == 0) a.res[3] = C_int;
    last_cs = 0.0;   // a pass without ocean leaves int C rsl = 0
    if (C_int > 0.0) {
        // subgrid sloping-coast load (src/vilma_sle.f90:289-291), own rows
        if (is_row)
            for (int w = 0; w < 2; w++) {
                const int j = 2 * p + w;
                if (j >= nlat) continue;
                for (int i = tid; i < N; i += 256) {
                    const size_t id = (size_t)j * N + i;
                    a.wcorr[id] = a.subgrid ? -VH_RHO_WATER * a.topo0[id] * (a.C[id] - a.C0[id]) : 0.0;
                }
            }
        n_passes += 1.0;
        last_ice = stot[1];
        last_zeta = stot[2];
    }
    __syncthreads();   // wcorr of the block's rows, written by all its threads, is read below by wavefronts 0, 1
    // ---- start: FC = FFT(C) (wave 0), F = FFT(load(rsl)) (wave 1)
    if (is_row && wave < 2) {
        const int r0 = 2 * p, r1 = r0 + 1;
        const bool has1 = r1 < nlat;
        double2* fa = RB + (size_t)wave * 2 * N;
        double2* fb = fa + N;
        const InLoad load{a.dig, a.C, a.rsl, a.wcorr, N};
        for (int i0 = 0; i0 < N; i0 += 64 * VH_FFT_U) {
            double v0[VH_FFT_U], v1[VH_FFT_U];
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int i = i0 + lane + 64 * u;
                if (wave == 0) {
                    v0[u] = i < N ? a.C[(size_t)r0 * N + i] : 0.0;
                    v1[u] = (has1 && i < N) ? a.C[(size_t)r1 * N + i] : 0.0;
                } else {
                    v0[u] = i < N ? load(r0, i) : 0.0;
                    v1[u] = (has1 && i < N) ? load(r1, i) : 0.0;
                }
            }
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int i = i0 + lane + 64 * u;
                if (i < N) fa[i] = make_double2(v0[u], v1[u]);
            }
        }
        wave_lds_sync();
        double2* z = sle_fft<false, MinW>(fa, fb, s, lane, TW);
        double2* dst = wave == 0 ? a.FC : a.F;
        lds_to_pair_spectra<true>(z, L, N, scale, dst + (size_t)r0 * (L + 1),
                                  has1 ? dst + (size_t)r1 * (L + 1) : nullptr, lane);
    }
    sle_grid_sync(a.bar, target, a.res);
    if (!is_row) {
        double2 fn = make_double2(0.0, 0.0), fs = fn;
        if (tid < nh) {
            const int js = nlat - 1 - tid;
            fn = ld_coh(a.F + (size_t)tid * (L + 1) + m);
            fs = js != tid ? ld_coh(a.F + (size_t)js * (L + 1) + m) : make_double2(0.0, 0.0);
        }
        sle_order_stage<TabLds>(a, m, T, tstride, U, fn, fs);
    }
    sle_grid_sync(a.bar, target, a.res);
    if (C_int <= 0.0) break;   // no ocean: load only; sigma above is the closing analysis

    for (int ii = 1; ii <= a.n_inner; ii++) {
        n_iters += 1.0;
        // ---- rows: inverse FFT of (u_j, N_j) -> Sraw_j, cs_j, the base load and its FFT
        if (is_row) {
            // after a wavefront's transform its result sits in buffer a when the
            // pass count is even, else in b; the other one takes the row's base load
            const size_t zoff = (s.nfac & 1) ? (size_t)N : 0, soff = (size_t)N - zoff;
            if (wave < 2) {
                const int j = 2 * p + wave;
                if (j < nlat) {
                    double2* fa = RB + (size_t)wave * 2 * N;
                    double2* fb = fa + N;
                    // the row's mask, rotation term, grounded ice and subgrid term
                    // in flight over the transform (longitudes below 64 VH_SLE_PU)
                    double cpre[VH_SLE_PU], spre[VH_SLE_PU], gpre[VH_SLE_PU], wpre[VH_SLE_PU];
#pragma unroll
                    for (int u = 0; u < VH_SLE_PU; u++) {
                        const int i = lane + 64 * u;
                        const size_t id = (size_t)j * N + i;
                        cpre[u] = i < N ? a.C[id] : 0.0;
                        spre[u] = (a.srot && i < N) ? a.srot[id] : 0.0;
                        gpre[u] = i < N ? a.dig[id] : 0.0;
                        wpre[u] = i < N ? a.wcorr[id] : 0.0;
                    }
                    const double2* F2p = a.F2 + (size_t)p * (L + 1) * 4 + 2 * wave;   // rows q = 2j, 2j + 1
                    pair_spectra_to_lds<true>(F2p, F2p + 1, L, N, fa, lane, 4);
                    double2* z = sle_fft<true, MinW>(fa, fb, s, lane, TW);
                    double rs = 0.0;
                    for (int i0 = 0; i0 < N; i0 += 64 * VH_SLE_PU) {
#pragma unroll
                        for (int u = 0; u < VH_SLE_PU; u++) {
                            const int i = i0 + lane + 64 * u;
                            if (i >= N) break;
                            const size_t id = (size_t)j * N + i;
                            const bool pre = i0 == 0;
                            double v = z[i].y - z[i].x;
                            if (a.srot) v = v + (pre ? spre[u] : a.srot[id]);
                            srow[wave * N + i] = v;
                            const double c = pre ? cpre[u] : a.C[id];
                            rs += c * v;
                            const double base = VH_RHO_ICE * (pre ? gpre[u] : a.dig[id]) + VH_RHO_WATER * (c * v) +
                                                (pre ? wpre[u] : a.wcorr[id]);
                            ((double*)(RB + (size_t)wave * 2 * N + soff))[i] = base;
                        }
                    }
                    rs = wave_sum(rs);
                    if (lane == 0) st_coh(a.cs + j, rs);
                }
            }
            __syncthreads();
            if (wave == 0) {
                // pack the pair's base loads (rows 2p, 2p+1) into wave 0's result
                // buffer, transform with its other buffer as scratch
                const int r0 = 2 * p;
                const bool has1 = r0 + 1 < nlat;
                double2* P = RB + zoff;
                const double* b0 = (const double*)(RB + soff);
                const double* b1 = (const double*)(RB + 2 * (size_t)N + soff);
                for (int i = lane; i < N; i += 64) P[i] = make_double2(b0[i], has1 ? b1[i] : 0.0);
                wave_lds_sync();
                double2* z = sle_fft<false, MinW>(P, RB + soff, s, lane, TW);
                lds_to_pair_spectra<true>(z, L, N, scale, a.F + (size_t)r0 * (L + 1),
                                          r0 + 1 < nlat ? a.F + (size_t)(r0 + 1) * (L + 1) : nullptr, lane);
            }
        }
        sle_grid_sync(a.bar, target, a.res);

        // ---- dphi (every block); rows: rsl update and row maxima; orders: the next sigma and (u, N)
        if (is_row) {
            sle_dphi(a, stot, nlat, dlon, wave, lane, &sdphi, &scs);
            __syncthreads();
            const double dphi = sdphi;
            last_cs = scs + dphi * C_int;   // int C rsl of this iterate: int C Sraw + dphi int C
            if (blk == 0 && tid == 0) { a.res[0] = (double)ii; a.res[1] = dphi; }
            const int r0 = 2 * p;
            double dm[2] = {0.0, 0.0}, sm[2] = {0.0, 0.0};
#pragma unroll
            for (int w = 0; w < 2; w++) {
                const int j = r0 + w;
                if (j >= nlat) continue;
                for (int i = tid; i < N; i += 256) {
                    const size_t id = (size_t)j * N + i;
                    const double rn = srow[w * N + i] + dphi;
                    const double c = a.C[id];
                    dm[w] = fmax(dm[w], fabs(c * (rn - a.rsl[id])));
                    a.rsl[id] = rn;
                    sm[w] = fmax(sm[w], fabs(c * rn));
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
        } else {
            // this thread's spectrum values in flight while dphi is formed
            double2 bn = make_double2(0.0, 0.0), bs = bn, cn = bn, cs = bn;
            const int js = nlat - 1 - tid;
            if (tid < nh) {
                bn = ld_coh(a.F + (size_t)tid * (L + 1) + m);
                cn = ld_coh(a.FC + (size_t)tid * (L + 1) + m);
                if (js != tid) {
                    bs = ld_coh(a.F + (size_t)js * (L + 1) + m);
                    cs = ld_coh(a.FC + (size_t)js * (L + 1) + m);
                }
            }
            sle_dphi(a, stot, nlat, dlon, wave, lane, &sdphi, &scs);
            __syncthreads();
            const double k = VH_RHO_WATER * sdphi;
            last_cs = scs + sdphi * C_int;
            const double2 fn = make_double2(bn.x + k * cn.x, bn.y + k * cn.y);
            const double2 fs = make_double2(bs.x + k * cs.x, bs.y + k * cs.y);
            sle_order_stage<TabLds>(a, m, T, tstride, U, fn, fs);
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
    if (a.fixed_ocean) break;
    // ---- coastline fixed point (src/vilma_sle.f90:336-344): C_next from topo0 - rsl,
    // flipped cells per row; every block stops when none flipped
    if (is_row) {
        double cnt[2] = {0.0, 0.0};
#pragma unroll
        for (int w = 0; w < 2; w++) {
            const int j = 2 * p + w;
            if (j >= nlat) continue;
            for (int i = tid; i < N; i += 256) {
                const size_t id = (size_t)j * N + i;
                const double cn = ocean(a.topo0[id] - a.rsl[id], a.ice[id]);
                a.Cn[id] = cn;
                cnt[w] += (cn != a.C[id]) ? 1.0 : 0.0;
            }
        }
#pragma unroll
        for (int w = 0; w < 2; w++) cnt[w] = wave_sum(cnt[w]);
        if (lane == 0) { redmax[wave][0] = cnt[0]; redmax[wave][1] = cnt[1]; }
        __syncthreads();
        if (tid < 2 && 2 * p + tid < nlat)
            st_coh(a.part + 2 * nlat + 2 * p + tid, redmax[0][tid] + redmax[1][tid] + redmax[2][tid] + redmax[3][tid]);
    }
    sle_grid_sync(a.bar, target, a.res);
    double nf = 0.0;   // whole numbers: exact in any order
    for (int q = tid; q < nlat; q += 256) nf += ld_coh(a.part + 2 * nlat + q);
    nf = wave_sum(nf);
    __syncthreads();
    if (lane == 0) redmax[wave][0] = nf;
    __syncthreads();
    nf = redmax[0][0] + redmax[1][0] + redmax[2][0] + redmax[3][0];
    __syncthreads();
    if (nf == 0.0) break;
  }
    if (blk == 0 && tid == 0) {
        a.res[4] = n_iters; a.res[5] = n_passes; a.res[6] = last_ice; a.res[7] = last_zeta;
        a.res[10] = last_cs;
        a.bar[2] = target;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

__global__ void k_sle_bar_reset(unsigned* bar) {
    if (threadIdx.x < 3) bar[threadIdx.x] = 0u;
    if (threadIdx.x < 17) bar[VH_SLE_FLAG_STRIDE * (1 + threadIdx.x)] = 0u;
}

void sle_bar_reset(hipStream_t st, unsigned* bar) {
    k_sle_bar_reset<<<1, 64, 0, st>>>(bar);
    HIPCHECK(hipGetLastError());
}

int sle_solve_grid(const vh_sht& s) { return (s.nlat + 1) / 2 + s.L + 1; }



long sle_solve_max_barriers(const vh_sht& s, int n_outer, int n_inner) {
    return (long)sle_solve_grid(s) * n_outer * (4 + 2 * (long)n_inner);
}

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0, double* wcorr,
             double* res, int nres) {
    k_sle_ice<<<nb256(std::max(n, nres)), 256, 0, st>>>(n, ice0, ice1, frac, interp, ice_ref, topo0, ice, dice, C0,
                                                         wcorr, res, nres);
}
void vh_apply_only(hipStream_t st, int nlm, const double2* sigma, const double* gu, const double* gn,
                   const double* dsurf, int nk, double g, int kind, double2* uN) {
    k_apply<<<nb256(nlm), 256, 0, st>>>(nlm, sigma, gu, gn, dsurf, nk, g, kind, uN);
}
void sle_cs(hipStream_t st, const vh_sht& s, const double* C, const double* f, double* part, double* out) {
    k_rowsum_prod<<<s.nlat, 256, 0, st>>>(s.nphi, C, f, part);
    vh_weighted_total(st, s, part, 1, out);
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

void sle_solve_launch(hipStream_t st, const sle_solve_args& a) {
    const vh_sht& s = a.s;
    if (s.nh > 256) throw vh_error("sle: the fused sea-level loop supports nlat <= 512");
    const int grid = sle_solve_grid(s);   // row blocks, then order blocks
    // the Legendre table of each order stays in LDS when it fits beside the rest
    const bool tab = sle_solve_lds(s, true) <= (size_t)156 * 1024;
    const size_t lds = sle_solve_lds(s, tab);
    int dev = 0, ncu = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const bool two = grid > ncu;   // one block per CU does not hold the grid
    const void* fn = tab ? (two ? (const void*)k_sle_solve<true, 2> : (const void*)k_sle_solve<true, 1>)
                         : (two ? (const void*)k_sle_solve<false, 2> : (const void*)k_sle_solve<false, 1>);
    // resident-block capacity, checked once per (device, kernel, LDS size)
    static thread_local int cap_dev = -1;
    static thread_local const void* cap_fn = nullptr;
    static thread_local size_t cap_lds = 0;
    static thread_local long cap = 0;
    if (cap_dev != dev || cap_fn != fn || cap_lds != lds) {
        int coop = 0, per_cu = 0;
        HIPCHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
        if (!coop) throw vh_error("sle: the device does not support cooperative launches");
        if (lds > 65536) HIPCHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds));
        cap_dev = dev; cap_fn = fn; cap_lds = lds; cap = (long)per_cu * ncu;
    }
    if (cap < grid)
        throw vh_error("sle: the fused sea-level loop needs " + std::to_string(grid) +
                       " resident blocks, the device holds " + std::to_string(cap));
    sle_solve_args arg = a;
    void* params[] = {&arg};
    // Cooperative launches of different models (streams) run one after the
    // other on the device: two grids that each hold part of the CUs while
    // waiting for the rest would never finish (the stream waits on the
    // previous launch's completion event).
    static std::mutex mu;
    static hipEvent_t last_ev[64] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (dev < 0 || dev >= 64) throw vh_error("sle: device index out of range");
    if (!last_ev[dev]) HIPCHECK(hipEventCreateWithFlags(&last_ev[dev], hipEventDisableTiming));
    else HIPCHECK(hipStreamWaitEvent(st, last_ev[dev], 0));
    HIPCHECK(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(256), params, (unsigned)lds, st));
    HIPCHECK(hipEventRecord(last_ev[dev], st));
}
