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
#include <atomic>
#include <cfloat>
#include "vh_internal.h"
#include "vh_sht_dev.h"
#include "vh_sle.h"

__device__ __forceinline__ double ocean(double topo, double ice) {
    return (topo < 0.0 && VH_RHO_ICE * ice < -VH_RHO_WATER * topo) ? 1.0 : 0.0;
}

// ice(t) = ice0 + frac*(ice1 - ice0), d_ice = ice - ice_ref, C0 = O(topo0, ice - d_ice)
// (src/vilma_timestep.f90:314-326, src/vilma_sle.f90:216); also the reset of
// the fused solve's synchronisation area (the solve that follows on the same
// stream starts from stage 0 with the record of a first pass)
__global__ void k_sle_ice(int n, const double* __restrict__ ice0, const double* __restrict__ ice1, double frac,
                          int interp, const double* __restrict__ ice_ref, const double* __restrict__ topo0,
                          double* ice, double* dice, double* C0, double* wcorr, double* res, int nres, unsigned* bar,
                          int nitems, int n_outer) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nres) res[i] = 0.0;
    if (bar) {
        if (i < nitems) bar[VH_SLE_CLAIM_OFF + VH_SLE_CLAIM_STRIDE * i] = 0u;
        if (i < 8) { bar[VH_SLE_FLAG_STRIDE * (1 + i)] = 0u; bar[VH_SLE_FLAG_STRIDE * (10 + i)] = 0u; }
        if (i < 4) bar[16 * (1 + i)] = i == 0 ? 1u : 0u;   // record slot 0 holds stage 0
        if (i == 0) { bar[VH_SLE_FLAG_STRIDE * 9] = 0u; bar[1] = 0u; bar[2] = 0u; }
        if (i < 16) {
            double* ctl = (double*)(bar + VH_SLE_CTL_OFF);
            ctl[i] = i == 0 ? (n_outer >= 1 ? 0.0 : 6.0) : (i == 1 ? 1.0 : 0.0);   // kind HEAD (DONE), pass 1
        }
    }
    if (i >= n) return;
    wcorr[i] = 0.0;
    double v = interp ? ice0[i] + frac * (ice1[i] - ice0[i]) : ice0[i];
    double d = v - ice_ref[i];
    ice[i] = v;
    dice[i] = d;
    C0[i] = ocean(topo0[i], v - d);
}

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
__global__ void k_absdiff(int n, const double* __restrict__ a, const double* __restrict__ b, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fabs(a[i] - b[i]);
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


// ---------------------------------------------------------------- fused solve
// The coastline passes and inner fixed point of sle_solve
// (src/vilma_sle.f90:231-355) in one launch of ordinary blocks, so no
// iteration waits on the host. The solve is a sequence of stages; each stage
// has one work item per block of the grid: latitude pair p (rows 2p, 2p+1)
// for items p < npairs, order m = item - npairs for the others (an item may
// have nothing to do in a stage):
//   HEAD   rows: mask C (first pass from topo0 - rsl, later the previous
//          pass's C_next; fixed ocean: C0), grounded-ice increment dig and
//          the row sums of C, dig and topo0 (C - C0)
//   START  rows: subgrid term wcorr, FC = FFT(C), F = FFT(load(rsl));
//          orders: their Legendre table into LDS
//   ORD0   orders: F -> sigma -> (u, N) -> F2
//   ROWS   rows: inverse FFT of (u_j, N_j) -> Sraw_j, cs_j = sum C Sraw,
//          base = rho_i dig + rho_w C Sraw + wcorr, F = FFT(base)
//   UPD    rows: rsl = Sraw + dphi, row maxima of |C drsl|, |C rsl|;
//          orders: the next load's spectrum F + rho_w dphi FC (the transform
//          of rho_i dig + rho_w C rsl + wcorr, by linearity) -> sigma -> F2
//   COAST  rows: C_next from topo0 - rsl, flipped cells per row
// in the order HEAD START ORD0 (ROWS UPD)* COAST per pass, then DONE. On
// exit sigma is the spectrum of the load of the final rsl: the closing
// analysis of the solve.
// After a stage every block in step (one that holds its item of the next
// stage, so the stage's outputs stay in place) takes the decision the next
// stage needs (pass totals and C_int; dphi; the inner exit test; the
// coastline fixed point) from the stage's outputs; the block whose arrival
// completed the stage also publishes it, with the running counts, in a
// control record that blocks which fell behind read instead.
// Each stage's arithmetic is fixed by the sizes alone (partition of the
// Legendre sums, order of the totals), so a rerun gives identical bits.
// No block waits for a block that has not started: of a launch of B blocks,
// block b owns items b, b + B, ... of every stage (B = the item count unless
// the caller asks for fewer blocks), claims each as it reaches it, and claims
// its first item of the next stage before it arrives at the current stage's
// end; a block that has waited VH_SLE_STEAL_POLLS polls on a stage
// claims and runs the stage's unclaimed items (whose owners are not resident:
// a device shared with other work or another process). A stage therefore
// completes as long as one block runs, and its result does not depend on
// which block ran an item. Arrivals go to the counter of the item's group
// (item % 8); the last item of a group arrives in the top counter (counters
// on pages of their own, so arrivals that come together do not queue on one
// address). Grid data written during the launch moves through agent-scope
// loads and stores (ld_coh / st_coh, coherent across the XCDs' L2s), so no
// cache write-back or invalidate is needed between stages; inputs written by
// earlier kernels are acquired once at entry. A block keeps LDS copies (its
// row pair's Sraw, the FFT twiddles, its order's Legendre table) and uses
// them only while they are valid. Every wait is bounded: past
// VH_SLE_WAIT_POLLS the error flag bar[1] and res[2] are raised and every
// block leaves.

// polls of a release word (each an agent-scope load, about a microsecond,
// plus a short sleep) before a waiting block takes over unclaimed work, and
// before a wait counts as failed
#define VH_SLE_STEAL_POLLS 48u
#define VH_SLE_WAIT_POLLS 4000000u
// P-table values a thread has in flight in the order stage
#define VH_SLE_JB 8
#define VH_SLE_LB 4
// grid values per lane a row wavefront loads ahead of its transform
#define VH_SLE_PU 6
// longitudes per thread and row whose inputs the pass head loads together
#define VH_SLE_PH 2
// longitudes per thread and row loaded together by the later row stages
#define VH_SLE_UB 4
// static LDS of the solve kernel (prefetch stage, records, flags), bytes
#define VH_SLE_STATIC_LDS (4 * 256 * 16 + 1024)

enum { SK_HEAD = 0, SK_START, SK_ORD0, SK_ROWS, SK_UPD, SK_COAST, SK_DONE };
// control record of a stage (doubles; small integers are exact)
enum { CT_KIND = 0, CT_IO, CT_II, CT_CINT, CT_TOT1, CT_TOT2, CT_DPHI, CT_NITER, CT_NPASS, CT_LICE, CT_LZETA,
       CT_LCS, CT_ITLAST, CT_DPLAST, CT_STEALS, CT_N = 16 };

__device__ __forceinline__ unsigned* sle_rel(unsigned* bar, int g) { return bar + VH_SLE_FLAG_STRIDE * (1 + g); }
__device__ __forceinline__ unsigned* sle_top(unsigned* bar) { return bar + VH_SLE_FLAG_STRIDE * 9; }
__device__ __forceinline__ unsigned* sle_grp(unsigned* bar, int g) { return bar + VH_SLE_FLAG_STRIDE * (10 + g); }
__device__ __forceinline__ unsigned* sle_claim(unsigned* bar, int i) {
    return bar + VH_SLE_CLAIM_OFF + VH_SLE_CLAIM_STRIDE * i;
}
__host__ __device__ __forceinline__ double* sle_ctl(unsigned* bar, unsigned sn) {
    return (double*)(bar + VH_SLE_CTL_OFF) + (sn & 3u) * CT_N;
}
__device__ __forceinline__ unsigned ld_u(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS of the row work: Sraw [2][nphi] doubles, the FFT twiddles (s.ntw
// complex) and two FFT buffer pairs (2 x 2 nphi complex; the pass head's row
// sums and the decisions use the same space). Of the order work: with TabLds
// the order's Legendre table [j][l] (row stride (L+1)|1 doubles), then the
// folds, partial sums and (sigma, u, N). The two layouts share the block's LDS.
__host__ __device__ static size_t sle_tab_doubles(const vh_sht& s) {
    return (((size_t)s.nh * ((s.L + 1) | 1)) + 1) & ~(size_t)1;
}
__host__ __device__ static size_t sle_rows_bytes(const vh_sht& s) {
    return (size_t)2 * s.nphi * sizeof(double) + (size_t)s.ntw * sizeof(double2) +
           std::max((size_t)4 * s.nphi * sizeof(double2), (size_t)6 * 256 * sizeof(double)) +
           (s.L < 64 * 3 ? (size_t)2 * 2 * 3 * 64 * sizeof(double2) : 0);   // prefetched pair spectra (L < 192)
}
static size_t sle_solve_lds(const vh_sht& s, bool tab) {
    const size_t order = (tab ? sle_tab_doubles(s) * sizeof(double) : 0) +
                         ((size_t)2 * s.nh + std::max(256, s.L + 1) + 3 * (s.L + 1) + 16 * (size_t)s.nh) * sizeof(double2);
    return std::max(sle_rows_bytes(s), order);
}

// wavefront sum of v over the block's lanes 0..63 (butterfly: every lane the same bits)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
// Order work of order m: fold of the load spectrum -> Legendre analysis ->
// sigma -> response -> Legendre synthesis of (u, N) -> F2. fn/fs: this
// thread's north / south spectrum values (j = tid < nh) already loaded by the
// caller. TabLds: the order's table is in T (LDS); otherwise it is read from
// the plm tables (the same values, the same sums).
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
            st_coh(a.sigma + (size_t)l * (l + 1) / 2 + m, r);
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

// ---- row work of latitude pair p (rows 2p, 2p+1)

// HEAD: the coastline pass head (src/vilma_sle.f90:231-294): the mask C (first
// pass: from topo0 - rsl; later: the previous pass's C_next; fixed ocean: C0),
// the grounded-ice increment d_ice_g and, per row, the sums of C, d_ice_g and
// topo0 (C - C0) (a 256-thread tree per row, red [6][256] in LDS)
__device__ __forceinline__ void sle_row_head(const sle_solve_args& a, int p, int io, double* red) {
    const int N = a.s.nphi, nlat = a.s.nlat, tid = threadIdx.x;
    double sp[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // the inputs of a thread's points of both rows (VH_SLE_PH per row) are
    // loaded before any is used; sums per thread in the same (row, longitude) order
    for (int i0 = 0; i0 < N; i0 += 256 * VH_SLE_PH) {
        double vC0[2][VH_SLE_PH], vice[2][VH_SLE_PH], vdice[2][VH_SLE_PH], vtopo[2][VH_SLE_PH], vx[2][VH_SLE_PH];
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
                vx[w][u] = (!ok || a.fixed_ocean) ? 0.0 : ld_coh(io == 1 ? a.rsl + id : a.Cn + id);
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
                st_coh(a.C + id, c);
                st_coh(a.dig + id, g);
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

// START: the subgrid sloping-coast load of the own rows (src/vilma_sle.f90:
// 289-291) when the pass has ocean, then FC = FFT(C) (wave 0) and
// F = FFT(load(rsl)) (wave 1). The load forms wcorr with the same expression
// instead of reading the value just stored.
template <int MinW>
__device__ __forceinline__ void sle_row_start(const sle_solve_args& a, int p, double C_int, double2* RB,
                                              const double2* TW) {
    const vh_sht& s = a.s;
    const int L = s.L, N = s.nphi, nlat = s.nlat, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double scale = 2.0 * VH_PI / N;
    const bool sea = C_int > 0.0;
    if (sea)
        for (int i0 = 0; i0 < N; i0 += 256 * VH_SLE_UB) {   // loads batched as in UPD
            double vc[2][VH_SLE_UB], vt[2][VH_SLE_UB], v0[2][VH_SLE_UB];
#pragma unroll
            for (int w = 0; w < 2; w++)
#pragma unroll
                for (int u = 0; u < VH_SLE_UB; u++) {
                    const int j = 2 * p + w, i = i0 + tid + 256 * u;
                    const bool ok = j < nlat && i < N;
                    const size_t id = ok ? (size_t)j * N + i : 0;
                    vc[w][u] = ok ? ld_coh(a.C + id) : 0.0;
                    vt[w][u] = ok ? a.topo0[id] : 0.0;
                    v0[w][u] = ok ? a.C0[id] : 0.0;
                }
#pragma unroll
            for (int w = 0; w < 2; w++)
#pragma unroll
                for (int u = 0; u < VH_SLE_UB; u++) {
                    const int j = 2 * p + w, i = i0 + tid + 256 * u;
                    if (j >= nlat || i >= N) continue;
                    st_coh(a.wcorr + (size_t)j * N + i,
                           a.subgrid ? -VH_RHO_WATER * vt[w][u] * (vc[w][u] - v0[w][u]) : 0.0);
                }
        }
    if (wave < 2) {
        const int r0 = 2 * p, r1 = r0 + 1;
        const bool has1 = r1 < nlat;
        double2* fa = RB + (size_t)wave * 2 * N;
        double2* fb = fa + N;
        auto load = [&](int row, int i) {
            const size_t id = (size_t)row * N + i;
            const double c = ld_coh(a.C + id);
            const double wc = sea ? (a.subgrid ? -VH_RHO_WATER * a.topo0[id] * (c - a.C0[id]) : 0.0) : ld_coh(a.wcorr + id);
            return VH_RHO_ICE * ld_coh(a.dig + id) + VH_RHO_WATER * (c * ld_coh(a.rsl + id)) + wc;
        };
        for (int i0 = 0; i0 < N; i0 += 64 * VH_FFT_U) {
            double v0[VH_FFT_U], v1[VH_FFT_U];
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int i = i0 + lane + 64 * u;
                if (wave == 0) {
                    v0[u] = i < N ? ld_coh(a.C + (size_t)r0 * N + i) : 0.0;
                    v1[u] = (has1 && i < N) ? ld_coh(a.C + (size_t)r1 * N + i) : 0.0;
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
}

// A row block's pair spectra for a ROWS item, loaded over a stage decision:
// orders m = lane + 64 u of wavefront w's row (F2 entries 2w, 2w + 1 of the
// pair) into stg[2u][lane], stg[2u + 1][lane] (stg: 2 SLE_RPF x 64 complex
// per wavefront), then packed as pair_spectra_to_lds packs them.
constexpr int SLE_RPF = 3;   // lmax < 64 SLE_RPF
__device__ __forceinline__ void row_spectra_issue(const double2* F2p, int L, int lane, double2* stg) {
#pragma unroll
    for (int u = 0; u < SLE_RPF; u++) {
        if (lane + 64 * u <= L) {
            const double2* f = F2p + (size_t)(lane + 64 * u) * 4;
            lds_dma16<true>(f, stg + 2 * u * 64);
            lds_dma16<true>(f + 1, stg + (2 * u + 1) * 64);
        }
    }
}
__device__ __forceinline__ void row_spectra_pack(const double2* stg, int L, int N, double2* a, int lane) {
    for (int i = lane; i < N; i += 64) a[i] = make_double2(0.0, 0.0);
    wave_lds_sync();
#pragma unroll 1   // (unrolled, the kernel spills)
    for (int u = 0; u < SLE_RPF; u++) {
        const int m = lane + 64 * u;
        if (m > L) continue;
        double2 x = stg[2 * u * 64 + lane], y = stg[(2 * u + 1) * 64 + lane];
        if (m == 0) { x.y = 0.0; y.y = 0.0; }
        a[m] = make_double2(x.x - y.y, x.y + y.x);
        if (m > 0) a[N - m] = make_double2(x.x + y.y, -x.y + y.x);
    }
    wave_lds_sync();
}

// ROWS: inverse FFT of (u_j, N_j) -> Sraw_j (into srow, and into sraw for a
// later stage run elsewhere), cs_j = sum C Sraw, the base load
// rho_i dig + rho_w C Sraw + wcorr and its FFT (the pair's rows packed) -> F
// stg: null, or the pair's spectra already in LDS (row_spectra_issue over the
// previous stage's decision, complete)
template <int MinW>
__device__ __forceinline__ void sle_row_rows(const sle_solve_args& a, int p, double* srow, double2* RB,
                                             const double2* TW, const double2* stg) {
    const vh_sht& s = a.s;
    const int L = s.L, N = s.nphi, nlat = s.nlat, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double scale = 2.0 * VH_PI / N;
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
                cpre[u] = i < N ? ld_coh(a.C + id) : 0.0;
                spre[u] = (a.srot && i < N) ? a.srot[id] : 0.0;
                gpre[u] = i < N ? ld_coh(a.dig + id) : 0.0;
                wpre[u] = i < N ? ld_coh(a.wcorr + id) : 0.0;
            }
            const double2* F2p = a.F2 + (size_t)p * (L + 1) * 4 + 2 * wave;   // rows q = 2j, 2j + 1
            if (stg) row_spectra_pack(stg + (size_t)wave * 2 * SLE_RPF * 64, L, N, fa, lane);
            else pair_spectra_to_lds<true>(F2p, F2p + 1, L, N, fa, lane, 4);
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
                    st_coh(a.sraw + id, v);
                    const double c = pre ? cpre[u] : ld_coh(a.C + id);
                    rs += c * v;
                    const double base = VH_RHO_ICE * (pre ? gpre[u] : ld_coh(a.dig + id)) + VH_RHO_WATER * (c * v) +
                                        (pre ? wpre[u] : ld_coh(a.wcorr + id));
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
                                  has1 ? a.F + (size_t)(r0 + 1) * (L + 1) : nullptr, lane);
    }
}

// UPD: rsl = Sraw + dphi (Sraw from srow when this block made it in the
// previous stage, else from sraw) and the row maxima of |C drsl|, |C rsl|.
// A thread's points (VH_SLE_UB per row and chunk) are loaded before any is
// used: coherent loads and stores keep their program order, so a
// load-use-store per point would wait one memory round trip per point.
__device__ __forceinline__ void sle_row_upd(const sle_solve_args& a, int p, double dphi, const double* srow,
                                            bool sraw_lds, double (*redmax)[4]) {
    const int N = a.s.nphi, nlat = a.s.nlat, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int r0 = 2 * p;
    double dm[2] = {0.0, 0.0}, sm[2] = {0.0, 0.0};
    for (int i0 = 0; i0 < N; i0 += 256 * VH_SLE_UB) {
        double vc[2][VH_SLE_UB], vr[2][VH_SLE_UB], vs[2][VH_SLE_UB];
#pragma unroll
        for (int w = 0; w < 2; w++)
#pragma unroll
            for (int u = 0; u < VH_SLE_UB; u++) {
                const int j = r0 + w, i = i0 + tid + 256 * u;
                const bool ok = j < nlat && i < N;
                const size_t id = ok ? (size_t)j * N + i : 0;
                vc[w][u] = ok ? ld_coh(a.C + id) : 0.0;
                vr[w][u] = ok ? ld_coh(a.rsl + id) : 0.0;
                vs[w][u] = (!ok || sraw_lds) ? 0.0 : ld_coh(a.sraw + id);
            }
#pragma unroll
        for (int w = 0; w < 2; w++)
#pragma unroll
            for (int u = 0; u < VH_SLE_UB; u++) {
                const int j = r0 + w, i = i0 + tid + 256 * u;
                if (j >= nlat || i >= N) continue;
                const size_t id = (size_t)j * N + i;
                const double rn = (sraw_lds ? srow[w * N + i] : vs[w][u]) + dphi;
                const double c = vc[w][u];
                dm[w] = fmax(dm[w], fabs(c * (rn - vr[w][u])));
                st_coh(a.rsl + id, rn);
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
}

// COAST: the coastline fixed point (src/vilma_sle.f90:336-344): C_next from
// topo0 - rsl and the flipped cells per row (loads batched as in UPD)
__device__ __forceinline__ void sle_row_coast(const sle_solve_args& a, int p, double (*redmax)[4]) {
    const int N = a.s.nphi, nlat = a.s.nlat, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    double cnt[2] = {0.0, 0.0};
    for (int i0 = 0; i0 < N; i0 += 256 * VH_SLE_UB) {
        double vt[2][VH_SLE_UB], vr[2][VH_SLE_UB], vi[2][VH_SLE_UB], vc[2][VH_SLE_UB];
#pragma unroll
        for (int w = 0; w < 2; w++)
#pragma unroll
            for (int u = 0; u < VH_SLE_UB; u++) {
                const int j = 2 * p + w, i = i0 + tid + 256 * u;
                const bool ok = j < nlat && i < N;
                const size_t id = ok ? (size_t)j * N + i : 0;
                vt[w][u] = ok ? a.topo0[id] : 0.0;
                vi[w][u] = ok ? a.ice[id] : 0.0;
                vr[w][u] = ok ? ld_coh(a.rsl + id) : 0.0;
                vc[w][u] = ok ? ld_coh(a.C + id) : 0.0;
            }
#pragma unroll
        for (int w = 0; w < 2; w++)
#pragma unroll
            for (int u = 0; u < VH_SLE_UB; u++) {
                const int j = 2 * p + w, i = i0 + tid + 256 * u;
                if (j >= nlat || i >= N) continue;
                const double cn = ocean(vt[w][u] - vr[w][u], vi[w][u]);
                st_coh(a.Cn + (size_t)j * N + i, cn);
                cnt[w] += (cn != vc[w][u]) ? 1.0 : 0.0;
            }
    }
#pragma unroll
    for (int w = 0; w < 2; w++) cnt[w] = wave_sum(cnt[w]);
    if (lane == 0) { redmax[wave][0] = cnt[0]; redmax[wave][1] = cnt[1]; }
    __syncthreads();
    if (tid < 2 && 2 * p + tid < nlat)
        st_coh(a.part + 2 * nlat + 2 * p + tid, redmax[0][tid] + redmax[1][tid] + redmax[2][tid] + redmax[3][tid]);
}

// The decision after stage `cur` (its control record) from the stage's
// outputs: the next stage's record into nx (LDS). scr: >= 3 nlat doubles of LDS.
__device__ __forceinline__ void sle_decide(const sle_solve_args& a, const double* cur, double* nx, double* scr,
                                           double (*redmax)[4]) {
    const vh_sht& s = a.s;
    const int nlat = s.nlat, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double dlon = 2.0 * VH_PI / s.nphi;
    if (tid < CT_N) nx[tid] = cur[tid];
    __syncthreads();
    const int kind = (int)cur[CT_KIND];
    if (kind == SK_HEAD) {
        // the pass totals sum_j w_j dlon part_j (latitude order, as
        // k_weighted_total): the terms are loaded by all threads, then one lane
        // per quantity adds them in order, 8 terms read ahead of their sums
        for (int t = tid; t < 3 * nlat; t += 256) {
            const int q = t % nlat;
            scr[t] = s.w[q] * ld_coh(a.part3 + t) * dlon;
        }
        __syncthreads();
        if (tid < 3) {
            double total = 0.0;
            int q = 0;
            for (; q + 8 <= nlat; q += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = scr[tid * nlat + q + u];
#pragma unroll
                for (int u = 0; u < 8; u++) total += v[u];
            }
            for (; q < nlat; q++) total += scr[tid * nlat + q];
            nx[CT_CINT + tid] = total;   // C_int, int d_ice_g, int topo0 (C - C0)
        }
        __syncthreads();
        if (tid == 0) {
            nx[CT_KIND] = SK_START;
            nx[CT_LCS] = 0.0;            // a pass without ocean leaves int C rsl = 0
            if (nx[CT_CINT] > 0.0) { nx[CT_NPASS] += 1.0; nx[CT_LICE] = nx[CT_TOT1]; nx[CT_LZETA] = nx[CT_TOT2]; }
        }
    } else if (kind == SK_START) {
        if (tid == 0) nx[CT_KIND] = SK_ORD0;
    } else if (kind == SK_ORD0) {
        // no ocean: load only; sigma of ORD0 is the closing analysis
        if (tid == 0) {
            if (cur[CT_CINT] <= 0.0) nx[CT_KIND] = SK_DONE;
            else { nx[CT_KIND] = SK_ROWS; nx[CT_II] = 1.0; nx[CT_NITER] += 1.0; }
        }
    } else if (kind == SK_ROWS) {
        // dphi = (ice_int - Cs_int + zeta) / C_int, Cs_int from the row sums cs
        if (wave == 0) {
            double total = 0.0;
            for (int q = lane; q < nlat; q += 64) total += s.w[q] * ld_coh(a.cs + q) * dlon;
            total = wave_sum(total);
            if (lane == 0) {
                const double ice_int = -(VH_RHO_ICE / VH_RHO_WATER) * cur[CT_TOT1];
                const double zeta = a.subgrid ? cur[CT_TOT2] : 0.0;
                const double dphi = (ice_int - total + zeta) / cur[CT_CINT];
                nx[CT_KIND] = SK_UPD;
                nx[CT_DPHI] = dphi;
                nx[CT_LCS] = total + dphi * cur[CT_CINT];   // int C rsl of this iterate
                nx[CT_ITLAST] = cur[CT_II];
                nx[CT_DPLAST] = dphi;
            }
        }
    } else if (kind == SK_UPD) {
        // the inner exit test (max is exact in any order)
        double dmx = 0.0, smx = 0.0;
        for (int q = tid; q < nlat; q += 256) {
            dmx = fmax(dmx, ld_coh(a.part + q));
            smx = fmax(smx, ld_coh(a.part + nlat + q));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) { dmx = fmax(dmx, __shfl_xor(dmx, off)); smx = fmax(smx, __shfl_xor(smx, off)); }
        if (lane == 0) { redmax[wave][0] = dmx; redmax[wave][1] = smx; }
        __syncthreads();
        dmx = fmax(fmax(redmax[0][0], redmax[1][0]), fmax(redmax[2][0], redmax[3][0]));
        smx = fmax(fmax(redmax[0][1], redmax[1][1]), fmax(redmax[2][1], redmax[3][1]));
        if (tid == 0) {
            const bool conv = dmx <= a.tol * fmax(smx, DBL_MIN);
            if (!conv && cur[CT_II] < (double)a.n_inner) {
                nx[CT_KIND] = SK_ROWS; nx[CT_II] = cur[CT_II] + 1.0; nx[CT_NITER] += 1.0;
            } else {
                nx[CT_KIND] = a.fixed_ocean ? SK_DONE : SK_COAST;
            }
        }
    } else if (kind == SK_COAST) {
        double nf = 0.0;   // whole numbers: exact in any order
        for (int q = tid; q < nlat; q += 256) nf += ld_coh(a.part + 2 * nlat + q);
        nf = wave_sum(nf);
        if (lane == 0) redmax[wave][0] = nf;
        __syncthreads();
        nf = redmax[0][0] + redmax[1][0] + redmax[2][0] + redmax[3][0];
        if (tid == 0) {
            if (nf == 0.0 || cur[CT_IO] >= (double)a.n_outer) nx[CT_KIND] = SK_DONE;
            else { nx[CT_KIND] = SK_HEAD; nx[CT_IO] = cur[CT_IO] + 1.0; nx[CT_II] = 0.0; }
        }
    }
    __syncthreads();
}

// page-0 words of the synchronisation area: per control-record slot k, the
// stage number + 1 of the record it holds (written after the record)
__host__ __device__ __forceinline__ unsigned* sle_slot_seq(unsigned* bar, unsigned sn) { return bar + 16 * (1 + (sn & 3u)); }

// MinW: wavefronts per SIMD the build must allow; 2 halves the registers
// (with spills) so that two blocks share a CU when the grid (latitude pairs +
// orders, 386 blocks at lmax 192) exceeds the CU count.
// Multi: the launch has fewer blocks than items (a block owns several).
template <bool TabLds, int MinW, bool Multi>
__global__ void __launch_bounds__(256, MinW) k_sle_solve(sle_solve_args a) {
    extern __shared__ double2 lds[];
    const vh_sht& s = a.s;
    const int L = s.L, N = s.nphi, nlat = s.nlat, nh = s.nh;
    // block own owns items own, own + B, own + 2B, ... of every stage
    const int npairs = (nlat + 1) / 2, G = npairs + L + 1, own = blockIdx.x, B = gridDim.x;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __shared__ double redmax[4][4];
    __shared__ double sct[CT_N], snx[CT_N];   // this stage's control record, the next one
    __shared__ unsigned su[2];
    __shared__ int slater;                     // the block's next owned item of the stage (item count B..)
    __shared__ unsigned long long smask;
    __shared__ double2 spf[4][256];            // an order's prefetched row spectra (nh <= 256)
    // Inputs written by the kernels before this launch (ice, drift, rsl, ...)
    // were written by other XCDs: an agent-scope acquire invalidates this
    // CU's and XCD's cached copies before anything is read, whatever fences
    // the launch itself carries; a release at the end publishes the outputs.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (a.late_blocks > 0 && own >= B - a.late_blocks) {   // diagnostic: a block that starts late
        if (tid == 0) {
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            while (__builtin_amdgcn_s_memrealtime() - t0 < (unsigned long long)a.late_ticks) __builtin_amdgcn_s_sleep(10);
        }
        __syncthreads();
    }
    // row layout of the LDS
    double* srow = (double*)lds;                          // [2][N]
    double2* TW = (double2*)(srow + 2 * N);               // [s.ntw]
    double2* RB = TW + s.ntw;                             // [2][2N] FFT buffer pairs
    // order layout
    double* T = (double*)lds;                             // [nh][tstride] (TabLds)
    double2* U = (double2*)(T + (TabLds ? sle_tab_doubles(s) : 0));
    const bool order_over_rows = !TabLds || sle_tab_doubles(s) * sizeof(double) < sle_rows_bytes(s);
    // what the LDS holds: the twiddles, the Sraw of a pair (and the stage it
    // was made in), the Legendre table of an order
    bool tw_ok = false, rows_pf = false;   // rows_pf: the own pair's spectra are packed for its next ROWS item
    int srow_item = -1, tab_m = -1;
    unsigned srow_stage = 0xffffffffu;

    auto bcast = [&](unsigned v) __attribute__((always_inline)) -> unsigned {   // thread 0's v to every thread
        if (tid == 0) su[0] = v;
        __syncthreads();
        const unsigned r = su[0];
        __syncthreads();
        return r;
    };
    auto claim = [&](int i, unsigned sn) __attribute__((always_inline)) -> bool {
        unsigned ok = 0;
        if (tid == 0)
            ok = __hip_atomic_fetch_max(sle_claim(a.bar, i), sn + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < sn + 1u;
        return bcast(ok) != 0u;
    };
    auto stage_now = [&]() __attribute__((always_inline)) -> unsigned { return bcast(tid == 0 ? ld_u(sle_rel(a.bar, own & 7)) : 0u); };
    // the published control record of stage sn into sct (waits for it; false
    // if the stage has moved on meanwhile or the launch failed)
    auto read_ctl = [&](unsigned sn) __attribute__((always_inline)) -> bool {
        unsigned ok = 0;
        if (tid == 0) {
            for (unsigned polls = 0;; polls++) {
                const unsigned q = ld_u(sle_slot_seq(a.bar, sn));
                if (q == sn + 1u) { ok = 1; break; }
                if (q > sn + 1u || ld_u(a.bar + 1) != 0u) break;
                if (polls > VH_SLE_WAIT_POLLS) {
                    __hip_atomic_fetch_or(a.bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    st_coh(a.res + 2, 1.0);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!bcast(ok)) return false;
        if (tid < CT_N) sct[tid] = ld_coh(sle_ctl(a.bar, sn) + tid);
        __syncthreads();
        return true;
    };
    auto publish = [&](unsigned sn1, const double* rec) __attribute__((always_inline)) {
        if (tid < CT_N) st_coh(sle_ctl(a.bar, sn1) + tid, rec[tid]);
        if (tid == 0) {
            const double steals = (double)ld_u(a.bar + 2);
            if ((int)rec[CT_KIND] == SK_DONE) {
                st_coh(a.res + 0, rec[CT_ITLAST]); st_coh(a.res + 1, rec[CT_DPLAST]); st_coh(a.res + 3, rec[CT_CINT]);
                st_coh(a.res + 4, rec[CT_NITER]); st_coh(a.res + 5, rec[CT_NPASS]); st_coh(a.res + 6, rec[CT_LICE]);
                st_coh(a.res + 7, rec[CT_LZETA]); st_coh(a.res + 8, steals); st_coh(a.res + 10, rec[CT_LCS]);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(sle_slot_seq(a.bar, sn1), sn1 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto scratch = [&]() __attribute__((always_inline)) -> double* { return (TabLds && tab_m >= 0) ? (double*)U : (double*)RB; };
    auto load_table = [&](int m) __attribute__((always_inline)) {
        // 8 table values per thread in flight
        const int nl = L - m + 1, nt = nh * nl, tstride = nl | 1;
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
        __syncthreads();
    };
    // order m's row spectra at the north / south rows of j = tid < nh: F
    // (and FC) into spf[0..1][tid] (spf[2..3][tid]), staged in LDS so that no
    // registers stay live across the row work's transforms
    bool pf_ok = false;
    auto fetch_order = [&](int m, bool withc) __attribute__((always_inline)) {
        const double2 z = make_double2(0.0, 0.0);
        if (tid < nh) {
            const int js = nlat - 1 - tid;
            const double2 f0 = ld_coh(a.F + (size_t)tid * (L + 1) + m);
            const double2 f1 = js != tid ? ld_coh(a.F + (size_t)js * (L + 1) + m) : z;
            double2 c0 = z, c1 = z;
            if (withc) {
                c0 = ld_coh(a.FC + (size_t)tid * (L + 1) + m);
                c1 = js != tid ? ld_coh(a.FC + (size_t)js * (L + 1) + m) : z;
            }
            spf[0][tid] = f0; spf[1][tid] = f1; spf[2][tid] = c0; spf[3][tid] = c1;
        }
    };
    // the same entries copied straight into spf (complete once the vector
    // memory counter has drained), in flight over a stage decision
    auto fetch_order_dma = [&](int m, bool withc) __attribute__((always_inline)) {
        if (tid < nh) {
            const int js = nlat - 1 - tid, w0 = 64 * wave;
            lds_dma16<true>(a.F + (size_t)tid * (L + 1) + m, &spf[0][w0]);
            if (withc) lds_dma16<true>(a.FC + (size_t)tid * (L + 1) + m, &spf[2][w0]);
            if (js != tid) {
                lds_dma16<true>(a.F + (size_t)js * (L + 1) + m, &spf[1][w0]);
                if (withc) lds_dma16<true>(a.FC + (size_t)js * (L + 1) + m, &spf[3][w0]);
            } else {
                spf[1][tid] = make_double2(0.0, 0.0); spf[3][tid] = make_double2(0.0, 0.0);
            }
        }
    };
    // the work of item i in stage sn (record sct); stolen: run for an owner
    // that is not resident
    auto run_item = [&](int i, unsigned sn, bool stolen) __attribute__((always_inline)) {
        const int kind = (int)sct[CT_KIND];
        if (i < npairs) {
            if (kind == SK_ORD0) return;
            if (TabLds) tab_m = -1;   // the row layout covers the table
            if (!tw_ok) {
                for (int t = tid; t < s.ntw; t += 256) TW[t] = s.twp[t];
                __syncthreads();
                tw_ok = true;
            }
            switch (kind) {
                case SK_HEAD: sle_row_head(a, i, (int)sct[CT_IO], (double*)RB); break;
                case SK_START: sle_row_start<MinW>(a, i, sct[CT_CINT], RB, TW); break;
                case SK_ROWS:
                    sle_row_rows<MinW>(a, i, srow, RB, TW, rows_pf && i == own ? RB + (size_t)4 * N : nullptr);
                    srow_item = i; srow_stage = sn;
                    break;
                case SK_UPD: sle_row_upd(a, i, sct[CT_DPHI], srow, srow_item == i && srow_stage + 1u == sn, redmax); break;
                case SK_COAST: sle_row_coast(a, i, redmax); break;
                default: break;
            }
            return;
        }
        const int m = i - npairs;
        if (kind != SK_START && kind != SK_ORD0 && kind != SK_UPD) return;
        const bool tab = TabLds && !stolen;   // a block runs another order's work from the global tables
        if (tab && tab_m != m) {
            tw_ok = false; srow_item = -1;    // the table covers the row layout
            load_table(m);
            tab_m = m;
        }
        if (kind == SK_START) return;          // the owner's table load overlaps the rows' transforms
        if (order_over_rows) { tw_ok = false; srow_item = -1; }
        if (!(pf_ok && i == own)) fetch_order(m, kind == SK_UPD);
        pf_ok = false;
        double2 fn = make_double2(0.0, 0.0), fs = fn;
        if (tid < nh) {   // each thread reads back its own entries
            fn = spf[0][tid]; fs = spf[1][tid];
            if (kind == SK_UPD) {
                const double k = VH_RHO_WATER * sct[CT_DPHI];
                const double2 c0 = spf[2][tid], c1 = spf[3][tid];
                fn = make_double2(fn.x + k * c0.x, fn.y + k * c0.y);
                fs = make_double2(fs.x + k * c1.x, fs.y + k * c1.y);
            }
        }
        if constexpr (TabLds) {
            if (tab) sle_order_stage<true>(a, m, T, (L - m + 1) | 1, U, fn, fs);
            else sle_order_stage<false>(a, m, T, 1, U, fn, fs);
        } else {
            sle_order_stage<false>(a, m, T, 1, U, fn, fs);
        }
    };
    // arrival of item i at the end of stage sn, with (pre) the claim of the
    // block's own item of stage sn + 1 issued beside it. The block whose
    // arrival completes the stage releases it: at once when it holds its item
    // of stage sn + 1 (sync: the stage's outputs stay until that item is
    // done; it publishes the decision afterwards), else after taking and
    // publishing the decision. Returns {completed, own item of sn + 1 claimed}.
    bool pending_publish = false;   // this block completed stage pending_sn - 1 and owes record pending_sn
    unsigned pending_sn = 0;
    auto arrive = [&](int i, unsigned sn, bool pre, bool sync) __attribute__((always_inline)) -> bool {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);   // this wavefront's stores are complete
        __syncthreads();
        if (tid == 0) {
            unsigned got = 0, last = 0;
            unsigned cl = 0xffffffffu;
            if (pre)
                cl = __hip_atomic_fetch_max(sle_claim(a.bar, own), sn + 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int g = i & 7;
            const unsigned ng = (unsigned)(G - g + 7) / 8u;
            if (__hip_atomic_fetch_add(sle_grp(a.bar, g), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == ng * (sn + 1u))
                last = __hip_atomic_fetch_add(sle_top(a.bar), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u ==
                       8u * (sn + 1u);
            got = pre ? (cl < sn + 2u) : 0u;
            su[0] = last; su[1] = got;
        }
        __syncthreads();
        const bool last = su[0] != 0u, got = su[1] != 0u;
        __syncthreads();
        if (last) {
            const bool after = pre ? got : sync;
            if (!after) {
                sle_decide(a, sct, snx, scratch(), redmax);
                publish(sn + 1u, snx);
            }
            if (tid == 0)
#pragma unroll
                for (int k = 0; k < 8; k++)
                    __hip_atomic_store(sle_rel(a.bar, k), sn + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pending_publish = after;
            pending_sn = sn + 1u;
        }
        return pre ? got : last;
    };
    // wait until stage sn is released. A block that has waited
    // VH_SLE_STEAL_POLLS polls looks for the stage's unclaimed items and claims
    // one: returns its index (>= 0; the caller runs it), -1 once the stage is
    // released, -2 on the error bound.
    auto wait_or_take = [&](unsigned sn, unsigned& polls) __attribute__((always_inline)) -> int {
        for (;;) {
            unsigned st = 0;
            if (tid == 0) {
                for (;; polls++) {
                    if (ld_u(sle_rel(a.bar, own & 7)) >= sn + 1u) { st = 0; break; }
                    if ((polls & 7u) == 7u && ld_u(a.bar + 1) != 0u) { st = 2; break; }
                    if (polls > VH_SLE_WAIT_POLLS) {
                        __hip_atomic_fetch_or(a.bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        st_coh(a.res + 2, 1.0);
                        st = 2;
                        break;
                    }
                    if (polls % VH_SLE_STEAL_POLLS == VH_SLE_STEAL_POLLS - 1u) { st = 1; polls++; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            st = bcast(st);
            if (st == 0) return -1;
            if (st == 2) return -2;
            for (int base = 0; base < G; base += 64) {
                if (wave == 0) {
                    const int j = base + lane;
                    const bool cand = j < G && ld_u(sle_claim(a.bar, j)) < sn + 1u;
                    const unsigned long long mk = __ballot(cand);
                    if (lane == 0) smask = mk;
                }
                __syncthreads();
                unsigned long long mk = smask;
                __syncthreads();
                while (mk != 0ull) {
                    const int j = base + __builtin_ctzll(mk);
                    mk &= mk - 1ull;
                    if (claim(j, sn)) return j;
                }
            }
        }
    };

    unsigned sn = stage_now();
    bool have = read_ctl(sn);                  // sct holds the record of stage sn
    bool claimed = have && claim(own, sn);
    for (;;) {
        if (have) {
            if ((int)sct[CT_KIND] == SK_DONE) {
                if (pending_publish && pending_sn == sn) publish(sn, sct);   // the result slots
                break;
            }
            bool sync = false;                 // this block holds its item of stage sn + 1
            int item = claimed ? own : -1;
            bool pre = claimed;                // the own item's arrival claims the next one
            unsigned polls = 0;                // of this stage's release word (thread 0)
            int w = 0;
            if (Multi && tid == 0) slater = own + B;   // read after the next barrier
            for (;;) {
                if (item >= 0) {
                    // the one place where work items run (keeps the transforms inlined once)
                    if (item != own) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    run_item(item, sn, item != own);
                    rows_pf = false;
                    if (pending_publish && pending_sn == sn) { publish(sn, sct); pending_publish = false; }
                    if ((Multi ? item % B : item) != own && tid == 0)   // an item of another block
                        __hip_atomic_fetch_add(a.bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (pre) sync = arrive(own, sn, true, false);
                    else arrive(item, sn, false, sync);
                    pre = false;
                }
                // the block's further owned items, each claimed as the block reaches it
                item = -1;
                if constexpr (Multi) {
                    __syncthreads();
                    for (int j = slater; j < G; j += B) {
                        if (claim(j, sn)) { item = j; break; }   // claim ends with barriers
                    }
                    if (tid == 0) slater = item >= 0 ? item + B : G;
                    if (item >= 0) continue;
                }
                w = wait_or_take(sn, polls);
                if (w < 0) break;
                item = w;
            }
            if (w == -2) break;
            if (sync) {
                // in step: the decision from the stage's outputs (which stay
                // until this block's item of stage sn + 1 is done). The
                // spectra of the block's next work are loaded first, in flight
                // over the decision, and staged in LDS after it: an order
                // block's after START (F) and ROWS (F and FC), a row block's
                // pair after ORD0 and UPD (F2, used when ROWS follows; not in
                // the instances with a register cap or several items per
                // block, where its addressing would spill)
                const int k0 = (int)sct[CT_KIND];
                const bool opf = own >= npairs && (k0 == SK_START || k0 == SK_ROWS);
                const bool rpf = MinW == 1 && !Multi && own < npairs && (k0 == SK_ORD0 || k0 == SK_UPD) && L < 64 * SLE_RPF;
                const bool rw = rpf && wave < 2 && 2 * own + wave < nlat;   // this wavefront's row
                if (opf) fetch_order_dma(own - npairs, k0 == SK_ROWS);
                if (rw)   // staged beyond the FFT buffers, unused by the decision
                    row_spectra_issue(a.F2 + (size_t)own * (L + 1) * 4 + 2 * wave, L, lane,
                                      RB + (size_t)4 * N + (size_t)wave * 2 * SLE_RPF * 64);
                sle_decide(a, sct, snx, scratch(), redmax);
                if (tid < CT_N) sct[tid] = snx[tid];
                __syncthreads();
                // the copies are complete before anything reads or reuses their LDS
                // (each thread of an order block reads back its own entries,
                // each row wavefront its own row's)
                if (opf || rw) __builtin_amdgcn_s_waitcnt(0);
                pf_ok = opf;
                rows_pf = rpf && (int)sct[CT_KIND] == SK_ROWS;
                sn += 1u;
                // the record of stage sn is published (pending_publish) after this
                // block's own work of the stage, off the stage's critical path
                continue;
            }
        } else if (bcast(tid == 0 ? ld_u(a.bar + 1) : 0u) != 0u) {
            break;                             // the launch failed
        }
        // behind: the published record of the current stage
        pending_publish = false;
        sn = stage_now();
        have = read_ctl(sn);
        claimed = have && claim(own, sn);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

int sle_solve_grid(const vh_sht& s) { return (s.nlat + 1) / 2 + s.L + 1; }

void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0, double* wcorr,
             double* res, int nres, unsigned* bar, int nitems, int n_outer) {
    const size_t nt = std::max<size_t>(std::max(n, nres), std::max(nitems, 64));
    k_sle_ice<<<nb256(nt), 256, 0, st>>>(n, ice0, ice1, frac, interp, ice_ref, topo0, ice, dice, C0, wcorr, res,
                                          nres, bar, nitems, n_outer);
    HIPCHECK(hipGetLastError());
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
void grid_absdiff(hipStream_t st, int n, const double* a, const double* b, double* out) {
    k_absdiff<<<nb256(n), 256, 0, st>>>(n, a, b, out);
}
void grid_ocean(hipStream_t st, int n, const double* topo, const double* ice, double* C) {
    k_ocean<<<nb256(n), 256, 0, st>>>(n, topo, ice, C);
}
void grid_bsl_integrand(hipStream_t st, int n, const double* h, const double* C, const double* heq,
                        const double* C0, double* out) {
    k_bsl_integrand<<<nb256(n), 256, 0, st>>>(n, h, C, heq, C0, out);
}

// Raises the dynamic-LDS limit of one instance of the solve kernel once per
// process (the attribute belongs to the function, not to a launch).
template <bool TabLds, int MinW, bool Multi>
static void sle_lds_limit(size_t lds) {
    static std::atomic<size_t> set{0};
    if (lds > 65536 && set.load() < lds) {
        HIPCHECK(hipFuncSetAttribute((const void*)k_sle_solve<TabLds, MinW, Multi>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        set.store(lds);
    }
}
template <bool TabLds, int MinW, bool Multi = false>
static void sle_launch_one(hipStream_t st, int grid, size_t lds, const sle_solve_args& a, bool launch) {
    sle_lds_limit<TabLds, MinW, Multi>(lds);
    if (!launch) return;
    k_sle_solve<TabLds, MinW, Multi><<<grid, 256, lds, st>>>(a);
    HIPCHECK(hipGetLastError());
}

void sle_solve_launch(hipStream_t st, const sle_solve_args& a, bool launch) {
    const vh_sht& s = a.s;
    if (s.nh > 256) throw vh_error("sle: the fused sea-level loop supports nlat <= 512");
    const int items = sle_solve_grid(s);   // latitude-pair items, then order items
    if (items > VH_SLE_MAX_ITEMS) throw vh_error("sle: the fused sea-level loop supports at most 1024 work items");
    // one block per item, or a.blocks blocks owning items round-robin
    const int grid = a.blocks > 0 ? std::min(items, a.blocks) : items;
    // the Legendre table of each order stays in LDS when it fits beside the
    // rest (160 KiB per block less the kernel's static LDS)
    const size_t lds_max = (size_t)160 * 1024 - VH_SLE_STATIC_LDS;
    const bool tab = sle_solve_lds(s, true) <= lds_max;
    const size_t lds = sle_solve_lds(s, tab);
    if (lds > lds_max) throw vh_error("sle: the fused sea-level loop needs more than 160 KiB of LDS");
    int dev = 0, ncu = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const bool two = grid > ncu;   // one block per CU does not hold the grid
    if (grid < items) {            // blocks owning several items: at most one per CU asked for
        if (two) throw vh_error("sle: VH_T_SLE_BLOCKS above the CU count");
        if (tab) sle_launch_one<true, 1, true>(st, grid, lds, a, launch);
        else sle_launch_one<false, 1, true>(st, grid, lds, a, launch);
        return;
    }
    if (tab) {
        if (two) sle_launch_one<true, 2>(st, grid, lds, a, launch);
        else sle_launch_one<true, 1>(st, grid, lds, a, launch);
    } else {
        if (two) sle_launch_one<false, 2>(st, grid, lds, a, launch);
        else sle_launch_one<false, 1>(st, grid, lds, a, launch);
    }
}
