// vh_drift.hip -- (a) the frozen relaxation drift of every (l,m) slot: the
// dissipative memory forcing -int tau:delta eps (2-point radial Gauss,
// src/vilma_viscoelastic.f90:441-480) followed by the equilibrated band
// solve (src/vilma_band.f90:111-142 via src/vilma_radial_fe.f90:578-605),
// with the reference's skip gate (mnorm <= skip_tol*max mnorm => zero drift,
// src/vilma_response.f90:711-760) and the degree-1 centre-of-mass projection
// (:386-420, :770-779).
//
// The forcing of a slot lives only on the U/V dofs of the memory nodes (the
// nodes of Maxwell elements), and the step only ever reads the solution back
// at those dofs and at the surface U, F, V. So at set-up, from the same
// pivoted band LU (bordered LU for degree 1), the solution operator restricted
// to exactly those rows and columns is formed once per degree:
//
//     G_l = S diag(dc) A_l^{-1} diag(dr) R,   (2 nmn + 3) x (2 nmn),
//
// with the centre-of-mass projection folded into G_1. Each step's drift is
// then a batched product Y_l = G_l F_l, one workgroup per (degree, tile of
// orders): the tile's forcing F_l is built in LDS straight from the memory
// (one coalesced pass over tau) and multiplied by G_l (L2-resident) on FP64
// MFMA, without a sequential triangular sweep. The result equals the band
// solve up to rounding (the same factor produced G).
#include <vector>
#include "vh_internal.h"
#include "vh_radial_dev.h"

// ------------------------------------------------------------ set-up (G_l)
// blockIdx.x = degree offset (l = lmin + bx), blockIdx.y = chunk of 64
// columns; lane = column; x[j*64 + lane] in global scratch. G holds one
// (2 nmn + 3) x (2 nmn) operator per block row, starting at degree lmin.
__global__ void __launch_bounds__(64) k_setup_gop(int lmin, int nr, int nd, int nmn, const int* __restrict__ mnodes,
        const double* __restrict__ ab, const int* __restrict__ ipiv, const double* __restrict__ drv,
        const double* __restrict__ dcv, double* work, double* G) {
    const int l = lmin + blockIdx.x, lane = threadIdx.x;
    const int ncol = 2 * nmn, nrow = ncol + 3;
    const int col = blockIdx.y * 64 + lane;
    double* x = work + ((size_t)blockIdx.x * gridDim.y + blockIdx.y) * nd * 64 + lane;
    const double* dr = drv + (size_t)(l - 1) * nd;
    const double* dc = dcv + (size_t)(l - 1) * nd;
    for (int j = 0; j < nd; j++) x[j * 64] = 0.0;
    if (col >= ncol) return;
    int dof = 4 * mnodes[col >> 1] + (col & 1);
    x[dof * 64] = dr[dof];
    lane_band_solve(nd, ab + (size_t)(l - 2) * VH_LDAB * nd, ipiv + (size_t)(l - 2) * nd, x, 64);
    double* Gl = G + (size_t)blockIdx.x * nrow * ncol;
    for (int row = 0; row < nrow; row++) {
        int src;
        if (row < ncol) src = 4 * mnodes[row >> 1] + (row & 1);
        else src = 4 * (nr - 1) + (row == ncol ? 0 : (row == ncol + 1 ? 2 : 1));
        Gl[(size_t)row * ncol + col] = dc[src] * x[src * 64];
    }
}

void vh_setup_restricted_ops(hipStream_t st, const vh_dims& d, const vh_dev& v, int lmin, int lmax, int nmn,
                             const int* mnodes, double* G) {
    const int ncol = 2 * nmn, nrow = ncol + 3;
    if (ncol == 0 || lmax < lmin) return;
    if (lmin < 2) throw vh_error("restricted operator: band factors start at degree 2");
    const int nch = (ncol + 63) / 64;
    const size_t per = (size_t)nch * d.nd * 64 * sizeof(double);
    int nb = (int)std::max<size_t>(1, ((size_t)256 << 20) / per);   // scratch under 256 MiB per batch
    nb = std::min(nb, lmax - lmin + 1);
    double* w = nullptr;
    HIPCHECK(hipMalloc((void**)&w, per * nb));
    for (int l0 = lmin; l0 <= lmax; l0 += nb) {
        int n = std::min(nb, lmax - l0 + 1);
        k_setup_gop<<<dim3(n, nch), 64, 0, st>>>(l0, d.nr, d.nd, nmn, mnodes, v.ab, v.ipiv, v.dr, v.dc, w,
                                                 G + (size_t)(l0 - lmin) * nrow * ncol);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(st));   // set-up only: release the scratch right away
    HIPCHECK(hipFree(w));
}

// ------------------------------------------------------------------ drift
struct drift_args {
    int nk, nem, nmn, ts, rw;
    const int2* jobs;     // (l, first order m0)
    const double* G;      // [l-1][nrow][ncol]
    const double* r;
    const int* emap;
    const int* mnode;
    const int* kdeg;
    const int2* node_el;  // [nmn] (Maxwell element below, above the node) or -1
    const double* Kt;     // [l-1][em][4][12] forcing map of (element, degree): floc = Kt (A, B, C)
    const double* tau;
    const double* mnorm;
    const double* maxnorm;
    double skip_tol;
    int deg1_cm;
    double* dN;
    double* dsurf;
};

#define DRIFT_THREADS 512
typedef double vh_d4 __attribute__((ext_vector_type(4)));

// Forcing of the block's right-hand sides (rhs = reim*ts + (m - m0)) at every
// memory-node dof, built in LDS straight from the memory: per (Maxwell
// element, rhs) the dissipative forcing -int tau:delta eps on the element's
// test dofs U_e, U_e+1, V_e, V_e+1, added at its two nodes. Elements of one
// parity share no node, so a pass per parity has no conflicting updates; a
// node then sums its element below and above (in either order: two terms).
template <int RW>
__device__ __forceinline__ void drift_forcing(const drift_args& a, int l, int kb, int nsl, double thr, double* F) {
    constexpr int rw = RW, ts = RW / 2;
    const int ncol = 2 * a.nmn, nk = a.nk;
    for (int t = threadIdx.x; t < ncol * rw; t += DRIFT_THREADS) F[t] = 0.0;
    __syncthreads();
    const int rhs = threadIdx.x % rw, grp = threadIdx.x / rw, ngrp = DRIFT_THREADS / rw;
    const int ml = rhs % ts, reim = rhs / ts;
    const bool act = ml < nsl && a.mnorm[kb + ml] > thr;
    const int k = kb + ml;
    const double Jr = (double)l * (double)(l + 1);
    constexpr int U = 3;   // elements whose memory a thread has in flight at once
    // VH_DRIFT_ONEPASS: every element in one pass, the two contributions a
    // node receives (element below, element above) added with LDS atomics,
    // which commute exactly for two terms on a zero start; else one pass per
    // element parity (elements of one parity share no node)
    constexpr int NPAR = VH_DRIFT_ONEPASS ? 1 : 2, EST = VH_DRIFT_ONEPASS ? 1 : 2;
    for (int par = 0; par < NPAR; par++) {
        if (act)
            for (int em0 = par + EST * grp; em0 < a.nem; em0 += EST * ngrp * U) {
                double v[U][12];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int em = em0 + EST * ngrp * u;
                    const double* t = a.tau + (size_t)min(em, a.nem - 1) * 24 * nk + k;
#pragma unroll
                    for (int q = 0; q < 12; q++) {   // (shape, lambda) of this re/im half
                        const int sh = q >> 2, lam = q & 3;
                        v[u][q] = em < a.nem ? t[(size_t)((sh * 2 + reim) * 4 + lam) * nk] : 0.0;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int em = em0 + EST * ngrp * u;
                    if (em >= a.nem) break;
                    const int e = a.emap[em];
                    typedef const __attribute__((address_space(4))) double cdouble;
                    cdouble* K = (cdouble*)(a.Kt + ((size_t)(l - 1) * a.nem + em) * 48);
                    double floc[4];
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        double f = 0.0;
#pragma unroll
                        for (int j = 0; j < 12; j++) f = fma(K[t * 12 + j], v[u][j], f);
                        floc[t] = f;
                    }
                    const int qlo = a.mnode[e], qup = a.mnode[e + 1];
                    if constexpr (VH_DRIFT_ONEPASS) {
                        atomicAdd(&F[(2 * qlo + 0) * rw + rhs], floc[0]);
                        atomicAdd(&F[(2 * qup + 0) * rw + rhs], floc[1]);
                        atomicAdd(&F[(2 * qlo + 1) * rw + rhs], floc[2]);
                        atomicAdd(&F[(2 * qup + 1) * rw + rhs], floc[3]);
                    } else {
                        F[(2 * qlo + 0) * rw + rhs] += floc[0];
                        F[(2 * qup + 0) * rw + rhs] += floc[1];
                        F[(2 * qlo + 1) * rw + rhs] += floc[2];
                        F[(2 * qup + 1) * rw + rhs] += floc[3];
                    }
                }
            }
        __syncthreads();
    }
}

// one solution row of one rhs -> nodal drift dN or surface dU, dF, dV
__device__ __forceinline__ void drift_store(const drift_args& a, int l, int k, int reim, int row, double y) {
    const int ncol = 2 * a.nmn, nk = a.nk;
    if (row < ncol) {
        const int q = row >> 1, uv = row & 1;
        a.dN[(size_t)(q * 4 + uv * 2 + reim) * nk + k] = y;
    } else {
        const int s = row - ncol;                      // 0: U, 1: F, 2: V
        const double val = (s == 1 && l == 1 && !a.deg1_cm) ? 0.0 : y;
        a.dsurf[(size_t)(s * nk + k) * 2 + reim] = val;
    }
}

// Workgroup = (degree l, ts consecutive orders): the forcing in LDS, then
// Y = G_l F. With RW >= 16 the product runs on FP64 MFMA (16x16x4 tiles;
// lane l: A[l%16][l/16], B[l/16][l%16], D[4r + l/16][l%16]): a wavefront per
// 16-row tile of G_l, all RW/16 16-column tiles of the right-hand sides,
// the G_l fragment read once per k-step from L2. Otherwise four rows per
// thread with scalar loads of G.
template <int RW>
__global__ void __launch_bounds__(DRIFT_THREADS, VH_DRIFT_MINW) k_drift(drift_args a) {
    extern __shared__ double F[];   // [ncol][rw]
    const int2 job = a.jobs[blockIdx.x];
    const int l = job.x, m0 = job.y;
    if (l <= 0) return;   // padding of the XCD-grouped job list
    constexpr int rw = RW, ts = RW / 2;
    const int ncol = 2 * a.nmn, nrow = ncol + 3;
    const int nsl = min(ts, l + 1 - m0);
    const int kb = l * (l + 1) / 2 - 1 + m0;
    const double thr = a.skip_tol * a.maxnorm[0];
    drift_forcing<RW>(a, l, kb, nsl, thr, F);
    const double* __restrict__ G = a.G + (size_t)(l - 1) * nrow * ncol;
    if (RW >= 16 && ncol <= 128) {   // the G fragments of a wavefront fit its registers
        constexpr int CT = RW >= 16 ? RW / 16 : 1;
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nrt = (nrow + 15) / 16;
        for (int rt = wave; rt < nrt; rt += DRIFT_THREADS / 64) {
            vh_d4 acc[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) acc[ct] = vh_d4{0.0, 0.0, 0.0, 0.0};
            const int arow = rt * 16 + (lane & 15);
            const double* Grow = G + (size_t)min(arow, nrow - 1) * ncol;
            // the wavefront's G_l fragment in flight VH_DRIFT_AK k-steps at a
            // time (ncol <= 128 with RW = 64: 32 k-steps in all)
            constexpr int AK = VH_DRIFT_AK;
#pragma unroll
            for (int k0 = 0; k0 < 32; k0 += AK) {
                if (4 * k0 >= ncol) break;
                double av[AK];
#pragma unroll
                for (int q = 0; q < AK; q++) {
                    const int kk = 4 * (k0 + q) + (lane >> 4);
                    av[q] = (arow < nrow && kk < ncol) ? Grow[kk] : 0.0;
                }
#pragma unroll
                for (int q = 0; q < AK; q++) {
                    const int ks = k0 + q;
                    if (4 * ks >= ncol) break;
                    const int kk = 4 * ks + (lane >> 4);
                    const bool kok = kk < ncol;
#pragma unroll
                    for (int ct = 0; ct < CT; ct++) {
                        const double bv = kok ? F[kk * RW + ct * 16 + (lane & 15)] : 0.0;
                        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], bv, acc[ct], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int ct = 0; ct < CT; ct++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = rt * 16 + 4 * r + (lane >> 4), rhs = ct * 16 + (lane & 15);
                    const int ml = rhs % ts, reim = rhs / ts;
                    if (row < nrow && ml < nsl) drift_store(a, l, kb + ml, reim, row, acc[ct][r]);
                }
        }
    } else {
        const int rhs = threadIdx.x % rw, ngrp = DRIFT_THREADS / rw, grp = threadIdx.x / rw;
        const int ml = rhs % ts, reim = rhs / ts;
        const bool valid = ml < nsl;
        const int nq = (nrow + 3) / 4;
        for (int qi = grp; qi < nq; qi += ngrp) {
            const int r0 = 4 * qi;
            const double* g0 = G + (size_t)min(r0 + 0, nrow - 1) * ncol;
            const double* g1 = G + (size_t)min(r0 + 1, nrow - 1) * ncol;
            const double* g2 = G + (size_t)min(r0 + 2, nrow - 1) * ncol;
            const double* g3 = G + (size_t)min(r0 + 3, nrow - 1) * ncol;
            double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
            for (int c = 0; c < ncol; c++) {
                const double f = F[c * rw + rhs];
                y0 = fma(g0[c], f, y0);
                y1 = fma(g1[c], f, y1);
                y2 = fma(g2[c], f, y2);
                y3 = fma(g3[c], f, y3);
            }
            if (!valid) continue;
            const double y[4] = {y0, y1, y2, y3};
            for (int i = 0; i < 4; i++)
                if (r0 + i < nrow) drift_store(a, l, kb + ml, reim, r0 + i, y[i]);
        }
    }
}

// ------------------------------------------------------------------ adjoint
// Transpose of k_drift for the same skip-gate set (slots the gate drops have
// a zero drift, so they pass no adjoint back): Y = [dN_h; dsurf_h] of the
// block's right-hand sides in LDS, Fh = G_l^T Y (FP64 MFMA for RW >= 16), then
// per (Maxwell element, rhs) the forcing transpose onto the element's memory,
// tau_h += (forcing)^T of its four test dofs (the block owns its slots).
template <int RW>
__global__ void __launch_bounds__(DRIFT_THREADS) k_drift_adj(drift_args a, const double* __restrict__ dNh,
                                                             const double* __restrict__ dsurfh, double* tauh) {
    extern __shared__ double Y[];   // [nrow][rw], then Fh [ncol][rw]
    const int2 job = a.jobs[blockIdx.x];
    const int l = job.x, m0 = job.y;
    if (l <= 0) return;   // padding of the XCD-grouped job list
    constexpr int rw = RW, ts = RW / 2;
    const int ncol = 2 * a.nmn, nrow = ncol + 3, nk = a.nk;
    double* Fh = Y + (size_t)nrow * rw;
    const int nsl = min(ts, l + 1 - m0);
    const int kb = l * (l + 1) / 2 - 1 + m0;
    const double thr = a.skip_tol * a.maxnorm[0];
    const int rhs = threadIdx.x % rw, grp = threadIdx.x / rw, ngrp = DRIFT_THREADS / rw;
    const int ml = rhs % ts, reim = rhs / ts;
    const bool act = ml < nsl && a.mnorm[kb + ml] > thr;
    const int k = kb + ml;
    for (int row = grp; row < nrow; row += ngrp) {
        double y = 0.0;
        if (act) {
            if (row < ncol) {
                const int q = row >> 1, uv = row & 1;
                y = dNh[(size_t)(q * 4 + uv * 2 + reim) * nk + k];
            } else {
                const int sidx = row - ncol;   // 0: U, 1: F, 2: V
                y = (sidx == 1 && l == 1 && !a.deg1_cm) ? 0.0 : dsurfh[(size_t)(sidx * nk + k) * 2 + reim];
            }
        }
        Y[row * rw + rhs] = y;
    }
    __syncthreads();
    const double* __restrict__ G = a.G + (size_t)(l - 1) * nrow * ncol;
    if (RW >= 16 && ncol <= 128) {   // the G fragments of a wavefront fit its registers
        constexpr int CT = RW >= 16 ? RW / 16 : 1;
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int nct = (ncol + 15) / 16;
        for (int rt = wave; rt < nct; rt += DRIFT_THREADS / 64) {
            vh_d4 acc[CT];
#pragma unroll
            for (int ct = 0; ct < CT; ct++) acc[ct] = vh_d4{0.0, 0.0, 0.0, 0.0};
            const int ac = rt * 16 + (lane & 15);   // A row = a column of G_l
            double av[33];                          // A column = a row of G_l (nrow <= 131)
#pragma unroll
            for (int ks = 0; ks < 33; ks++) {
                const int kk = 4 * ks + (lane >> 4);
                av[ks] = (ac < ncol && kk < nrow) ? G[(size_t)kk * ncol + ac] : 0.0;
            }
#pragma unroll
            for (int ks = 0; ks < 33; ks++) {
                if (4 * ks >= nrow) break;
                const int kk = 4 * ks + (lane >> 4);
                const bool kok = kk < nrow;
#pragma unroll
                for (int ct = 0; ct < CT; ct++) {
                    const double bv = kok ? Y[kk * RW + ct * 16 + (lane & 15)] : 0.0;
                    acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], bv, acc[ct], 0, 0, 0);
                }
            }
#pragma unroll
            for (int ct = 0; ct < CT; ct++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int c = rt * 16 + 4 * r + (lane >> 4);
                    if (c < ncol) Fh[c * RW + ct * 16 + (lane & 15)] = acc[ct][r];
                }
        }
    } else {
        for (int c = grp; c < ncol; c += ngrp) {
            double f = 0.0;
            for (int row = 0; row < nrow; row++) f = fma(G[(size_t)row * ncol + c], Y[row * rw + rhs], f);
            Fh[c * rw + rhs] = f;
        }
    }
    __syncthreads();
    if (!act) return;
    for (int em = grp; em < a.nem; em += ngrp) {
        const int e = a.emap[em];
        const int qlo = a.mnode[e], qup = a.mnode[e + 1];
        // test dofs U_e, U_e+1, V_e, V_e+1 <- (lower, U), (upper, U), (lower, V), (upper, V)
        const double fl[4] = {Fh[(2 * qlo + 0) * rw + rhs], Fh[(2 * qup + 0) * rw + rhs],
                              Fh[(2 * qlo + 1) * rw + rhs], Fh[(2 * qup + 1) * rw + rhs]};
        typedef const __attribute__((address_space(4))) double cdouble;
        cdouble* K = (cdouble*)(a.Kt + ((size_t)(l - 1) * a.nem + em) * 48);
        double* t = tauh + (size_t)em * 24 * nk + k;
#pragma unroll
        for (int j = 0; j < 12; j++) {   // (shape, lambda) of this re/im half
            const double g = fma(K[j], fl[0], fma(K[12 + j], fl[1], fma(K[24 + j], fl[2], K[36 + j] * fl[3])));
            const int sh = j >> 2, lam = j & 3;
            t[(size_t)((sh * 2 + reim) * 4 + lam) * nk] += g;
        }
    }
}

// The dissipative forcing of (element, degree) is linear in the element's
// 12 memory values of one re/im half: its 4 x 12 matrix, column j the forcing
// of the unit memory e_j (the same element integrals, evaluated once here).
__global__ void k_drift_ktab(int L, int nem, const double* __restrict__ r, const int* __restrict__ emap, double* Kt) {
    const int em = blockIdx.x, j = threadIdx.x % 12, lo = threadIdx.x / 12;
    const int l = blockIdx.y * (blockDim.x / 12) + lo + 1;
    if (threadIdx.x >= (blockDim.x / 12) * 12 || l > L || em >= nem) return;
    const int e = emap[em];
    const double Jr = (double)l * (double)(l + 1);
    double A[4] = {0.0, 0.0, 0.0, 0.0}, B[4] = {0.0, 0.0, 0.0, 0.0}, Cc[4] = {0.0, 0.0, 0.0, 0.0}, f[4];
    if (j < 4) A[j] = 1.0; else if (j < 8) B[j - 4] = 1.0; else Cc[j - 8] = 1.0;
    dissipative_element(r[e], r[e + 1], Jr, A, B, Cc, f);
    double* K = Kt + ((size_t)(l - 1) * nem + em) * 48;
    for (int t = 0; t < 4; t++) K[t * 12 + j] = f[t];
}

// ------------------------------------------------------------------ plan
struct vh_drift_plan {
    int2* jobs = nullptr;
    int2* node_el = nullptr;
    double* Kt = nullptr;
    double* G = nullptr;
    int njobs = 0, ts = 32, rw = 64;
    size_t lds = 0, lds_t = 0;
};

const int2* vh_drift_plan_node_el(const vh_drift_plan* p) { return p->node_el; }

vh_drift_plan* vh_drift_plan_create(hipStream_t st, const vh_dims& d, const vh_dev& v, int deg1_cm,
                                    const std::vector<int>& emap, const std::vector<int>& mnode_of) {
    vh_drift_plan* p = new vh_drift_plan();
    {   // per memory node: the Maxwell element below (upper node = q) and above (lower node = q)
        std::vector<int2> ne(std::max(d.nmn, 1), make_int2(-1, -1));
        for (int em = 0; em < d.nem; em++) {
            const int qlo = mnode_of[emap[em]], qup = mnode_of[emap[em] + 1];
            ne[qlo].y = em;
            ne[qup].x = em;
        }
        HIPCHECK(hipMalloc((void**)&p->node_el, ne.size() * sizeof(int2)));
        HIPCHECK(hipMemcpy(p->node_el, ne.data(), ne.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    if (d.nem > 0) {
        HIPCHECK(hipMalloc((void**)&p->Kt, (size_t)d.L * d.nem * 48 * sizeof(double)));
        dim3 g(d.nem, (d.L + 20) / 21);
        k_drift_ktab<<<g, 252, 0, st>>>(d.L, d.nem, v.r, v.emap, p->Kt);
        HIPCHECK(hipGetLastError());
    }
    const int ncol = 2 * d.nmn, nrow = ncol + 3;
    // right-hand sides per workgroup: as many as fit a 64 KiB forcing tile
    int rw = VH_DRIFT_RW;
    while (rw > 2 && (size_t)std::max(ncol, 1) * rw * 8 > 64 * 1024) rw /= 2;
    if ((size_t)std::max(ncol, 1) * rw * 8 > 160 * 1024)
        throw vh_error("drift: too many memory nodes for the LDS forcing tile (nmn = " + std::to_string(d.nmn) + ")");
    p->rw = rw; p->ts = rw / 2;   // the MFMA paths (rw >= 16) also need ncol <= 128
    p->lds = (size_t)std::max(ncol, 1) * rw * 8;
    // jobs (degree, first order): blocks are dealt round-robin to the 8 XCDs
    // (block b on XCD b % 8), so every job of one degree is placed on one
    // XCD, whose L2 then fetches that degree's G_l once; degrees go to the
    // least loaded XCD, largest first; (0, 0) pads the list
    std::vector<int2> per[8];
    for (int l = d.L; l >= 1; l--) {
        int x = 0;
        for (int q = 1; q < 8; q++)
            if (per[q].size() < per[x].size()) x = q;
        for (int m0 = 0; m0 <= l; m0 += p->ts) per[x].push_back(make_int2(l, m0));
    }
    size_t nmax = 0;
    for (const auto& v : per) nmax = std::max(nmax, v.size());
    std::vector<int2> jobs;
    for (size_t k = 0; k < nmax; k++)
        for (int x = 0; x < 8; x++) jobs.push_back(k < per[x].size() ? per[x][k] : make_int2(0, 0));
    p->njobs = (int)jobs.size();
    HIPCHECK(hipMalloc((void**)&p->jobs, jobs.size() * sizeof(int2)));
    HIPCHECK(hipMemcpy(p->jobs, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc((void**)&p->G, (size_t)d.L * nrow * std::max(ncol, 1) * sizeof(double)));
    HIPCHECK(hipMemsetAsync(p->G, 0, (size_t)d.L * nrow * std::max(ncol, 1) * sizeof(double), st));
    if (ncol > 0) {
        // degree 1 through the bordered factor (centre-of-mass projection folded in)
        double* w1 = nullptr;
        HIPCHECK(hipMalloc((void**)&w1, (size_t)(d.ns1 + 1) * sizeof(double)));
        vh_launch_setup_deg1_op(st, d, v.mnodes, v.ab1, v.ipiv1, v.dr, v.dc, v.null1, deg1_cm, w1, p->G);
        vh_setup_restricted_ops(st, d, v, 2, d.L, d.nmn, v.mnodes, p->G + (size_t)nrow * ncol);
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(hipFree(w1));
    }
    p->lds_t = (size_t)(2 * std::max(ncol, 1) + 3) * rw * 8;
    if (p->lds_t > 160 * 1024)
        throw vh_error("drift adjoint: too many memory nodes for the LDS tile (nmn = " + std::to_string(d.nmn) + ")");
    switch (p->rw) {
        case 64: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_adj<64>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 32: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_adj<32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 16: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_adj<16>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 8: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_adj<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 4: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_adj<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        default: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_adj<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
    }
    switch (p->rw) {
        case 64: HIPCHECK(hipFuncSetAttribute((const void*)k_drift<64>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds)); break;
        case 32: HIPCHECK(hipFuncSetAttribute((const void*)k_drift<32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds)); break;
        case 16: HIPCHECK(hipFuncSetAttribute((const void*)k_drift<16>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds)); break;
        case 8: HIPCHECK(hipFuncSetAttribute((const void*)k_drift<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds)); break;
        case 4: HIPCHECK(hipFuncSetAttribute((const void*)k_drift<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds)); break;
        default: HIPCHECK(hipFuncSetAttribute((const void*)k_drift<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds)); break;
    }
    return p;
}

void vh_drift_plan_destroy(vh_drift_plan* p) {
    if (!p) return;
    (void)hipFree(p->jobs);
    (void)hipFree(p->G);
    (void)hipFree(p->node_el);
    (void)hipFree(p->Kt);
    delete p;
}

void vh_drift_run(hipStream_t st, const vh_drift_plan* p, const vh_dims& d, const vh_dev& v,
                  const double* tau, const double* mnorm, const double* maxnorm, double skip_tol,
                  int deg1_cm, double* dN, double* dsurf) {
    drift_args a;
    a.nk = d.nk; a.nem = d.nem; a.nmn = d.nmn; a.ts = p->ts; a.rw = p->rw;
    a.jobs = p->jobs; a.G = p->G; a.r = v.r; a.emap = v.emap; a.mnode = v.mnode; a.kdeg = v.kdeg;
    a.node_el = p->node_el; a.Kt = p->Kt;
    a.tau = tau; a.mnorm = mnorm; a.maxnorm = maxnorm;
    a.skip_tol = skip_tol; a.deg1_cm = deg1_cm; a.dN = dN; a.dsurf = dsurf;
    switch (p->rw) {
        case 64: k_drift<64><<<p->njobs, DRIFT_THREADS, p->lds, st>>>(a); break;
        case 32: k_drift<32><<<p->njobs, DRIFT_THREADS, p->lds, st>>>(a); break;
        case 16: k_drift<16><<<p->njobs, DRIFT_THREADS, p->lds, st>>>(a); break;
        case 8: k_drift<8><<<p->njobs, DRIFT_THREADS, p->lds, st>>>(a); break;
        case 4: k_drift<4><<<p->njobs, DRIFT_THREADS, p->lds, st>>>(a); break;
        default: k_drift<2><<<p->njobs, DRIFT_THREADS, p->lds, st>>>(a); break;
    }
    HIPCHECK(hipGetLastError());
}

void vh_drift_adjoint(hipStream_t st, const vh_drift_plan* p, const vh_dims& d, const vh_dev& v,
                      const double* mnorm, const double* maxnorm, double skip_tol, int deg1_cm,
                      const double* dNh, const double* dsurfh, double* Fh, double* tauh) {
    (void)Fh;
    if (d.nem == 0) return;
    drift_args a;
    a.nk = d.nk; a.nem = d.nem; a.nmn = d.nmn; a.ts = p->ts; a.rw = p->rw;
    a.jobs = p->jobs; a.G = p->G; a.r = v.r; a.emap = v.emap; a.mnode = v.mnode; a.kdeg = v.kdeg;
    a.node_el = p->node_el; a.Kt = p->Kt;
    a.tau = nullptr; a.mnorm = mnorm; a.maxnorm = maxnorm;
    a.skip_tol = skip_tol; a.deg1_cm = deg1_cm; a.dN = nullptr; a.dsurf = nullptr;
    switch (p->rw) {
        case 64: k_drift_adj<64><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, tauh); break;
        case 32: k_drift_adj<32><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, tauh); break;
        case 16: k_drift_adj<16><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, tauh); break;
        case 8: k_drift_adj<8><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, tauh); break;
        case 4: k_drift_adj<4><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, tauh); break;
        default: k_drift_adj<2><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, tauh); break;
    }
    HIPCHECK(hipGetLastError());
}
