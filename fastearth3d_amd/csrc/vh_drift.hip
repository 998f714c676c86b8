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
// (one coalesced pass over tau) and multiplied by G_l (L2-resident) without
// a sequential triangular sweep. The result equals the band solve up to
// rounding (the same factor produced G).
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
    double* Fe;           // [em][4][2][nk] element forcing (floc, re/im)
    const double* tau;
    const double* mnorm;
    const double* maxnorm;
    double skip_tol;
    int deg1_cm;
    double* dN;
    double* dsurf;
};

#define DRIFT_THREADS 512

// Dissipative forcing of every (slot, Maxwell element): Fe[em][c][reim][k] =
// floc_c (test dofs U_e, U_e+1, V_e, V_e+1), for the slots the skip gate keeps.
// Grid: x = 64 slots (coalesced), y = 4 elements.
__global__ void __launch_bounds__(256) k_drift_force(drift_args a) {
    const int lane = threadIdx.x & 63, em = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.x * 64 + lane, nk = a.nk;
    if (k >= nk || em >= a.nem) return;
    if (!(a.mnorm[k] > a.skip_tol * a.maxnorm[0])) return;
    const int l = a.kdeg[k], e = a.emap[em];
    const double Jr = (double)l * (double)(l + 1);
    const double* t = a.tau + (size_t)em * 24 * nk + k;
    double* o = a.Fe + (size_t)em * 8 * nk + k;
#pragma unroll
    for (int reim = 0; reim < 2; reim++) {
        double A[4], B[4], Cc[4], floc[4];
#pragma unroll
        for (int lam = 0; lam < 4; lam++) {
            A[lam] = t[(size_t)((0 * 2 + reim) * 4 + lam) * nk];
            B[lam] = t[(size_t)((1 * 2 + reim) * 4 + lam) * nk];
            Cc[lam] = t[(size_t)((2 * 2 + reim) * 4 + lam) * nk];
        }
        dissipative_element(a.r[e], a.r[e + 1], Jr, A, B, Cc, floc);
#pragma unroll
        for (int c = 0; c < 4; c++) o[(size_t)(c * 2 + reim) * nk] = floc[c];
    }
}

// Workgroup = (degree l, ts consecutive orders): rhs = reim*ts + (m - m0);
// RW = 2 ts right-hand sides (64: one wavefront per row group).
template <int RW>
__global__ void __launch_bounds__(DRIFT_THREADS) k_drift(drift_args a) {
    extern __shared__ double F[];   // [ncol][rw]
    const int2 job = a.jobs[blockIdx.x];
    const int l = job.x, m0 = job.y;
    constexpr int rw = RW, ts = RW / 2;
    const int ncol = 2 * a.nmn, nrow = ncol + 3, nk = a.nk;
    const int nsl = min(ts, l + 1 - m0);
    const int kb = l * (l + 1) / 2 - 1 + m0;
    const double thr = a.skip_tol * a.maxnorm[0];
    const int rhs = threadIdx.x % rw, ngrp = DRIFT_THREADS / rw;
    // with RW = 64 the row group is the wavefront index: uniform, so the G rows
    // below stream through scalar loads instead of wave-wide vector loads
    const int grp = RW == 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : threadIdx.x / rw;
    const int ml = rhs % ts, reim = rhs / ts;
    const bool valid = ml < nsl;
    const bool act = valid && a.mnorm[kb + ml] > thr;
    // forcing at node q, dof uv: the upper-node term of the element below it
    // plus the lower-node term of the element above it (assembly order of
    // the reference: the element below comes first, src/vilma_viscoelastic.f90:441-480)
    for (int c = grp; c < ncol; c += ngrp) {
        double f = 0.0;
        if (act) {
            const int q = c >> 1, uv = c & 1;
            const int2 ne = a.node_el[q];
            const size_t kk = (size_t)reim * nk + kb + ml;
            if (ne.x >= 0) f = f + a.Fe[((size_t)ne.x * 4 + (uv ? 3 : 1)) * 2 * nk + kk];
            if (ne.y >= 0) f = f + a.Fe[((size_t)ne.y * 4 + (uv ? 2 : 0)) * 2 * nk + kk];
        }
        F[c * rw + rhs] = f;
    }
    __syncthreads();
    // Y = G_l F, four rows per thread per pass
    const double* __restrict__ G = a.G + (size_t)(l - 1) * nrow * ncol;
    const int nq = (nrow + 3) / 4;
    const int k = kb + ml;
    // G is read-only for the whole launch: read it through the constant
    // address space, so the wave-uniform rows become scalar loads
    typedef const __attribute__((address_space(4))) double cdouble;
    for (int qi = grp; qi < nq; qi += ngrp) {
        const int r0 = 4 * qi;
        cdouble* g0 = (cdouble*)(G + (size_t)min(r0 + 0, nrow - 1) * ncol);
        cdouble* g1 = (cdouble*)(G + (size_t)min(r0 + 1, nrow - 1) * ncol);
        cdouble* g2 = (cdouble*)(G + (size_t)min(r0 + 2, nrow - 1) * ncol);
        cdouble* g3 = (cdouble*)(G + (size_t)min(r0 + 3, nrow - 1) * ncol);
        double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
#pragma unroll 4
        for (int c = 0; c < ncol; c++) {
            const double f = F[c * rw + rhs];
            y0 = fma(g0[c], f, y0);
            y1 = fma(g1[c], f, y1);
            y2 = fma(g2[c], f, y2);
            y3 = fma(g3[c], f, y3);
        }
        if (!valid) continue;
        const double y[4] = {y0, y1, y2, y3};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = r0 + i;
            if (row < ncol) {
                const int q = row >> 1, uv = row & 1;
                a.dN[(size_t)(q * 4 + uv * 2 + reim) * nk + k] = y[i];
            } else if (row < nrow) {
                const int s = row - ncol;                      // 0: U, 1: F, 2: V
                double val = (s == 1 && l == 1 && !a.deg1_cm) ? 0.0 : y[i];
                a.dsurf[(size_t)(s * nk + k) * 2 + reim] = val;
            }
        }
    }
}

// ------------------------------------------------------------------ adjoint
// Transpose of the two kernels above for the same skip-gate set (slots the
// gate drops have a zero drift, so they pass no adjoint back).
//   k_drift_gT:        Fh = G_l^T [dN_h; dsurf_h] per (slot, re/im),
//                      Fh[(c*2 + reim)][nk], c = 2 q + uv the node dof;
//   k_drift_force_adj: tau_h += (forcing)^T of the element's four test dofs.
template <int RW>
__global__ void __launch_bounds__(DRIFT_THREADS) k_drift_gT(drift_args a, const double* __restrict__ dNh,
                                                           const double* __restrict__ dsurfh, double* Fh) {
    extern __shared__ double Y[];   // [nrow][rw]
    const int2 job = a.jobs[blockIdx.x];
    const int l = job.x, m0 = job.y;
    constexpr int rw = RW, ts = RW / 2;
    const int ncol = 2 * a.nmn, nrow = ncol + 3, nk = a.nk;
    const int nsl = min(ts, l + 1 - m0);
    const int kb = l * (l + 1) / 2 - 1 + m0;
    const double thr = a.skip_tol * a.maxnorm[0];
    const int rhs = threadIdx.x % rw, ngrp = DRIFT_THREADS / rw;
    const int grp = RW == 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : threadIdx.x / rw;
    const int ml = rhs % ts, reim = rhs / ts;
    const bool valid = ml < nsl;
    const bool act = valid && a.mnorm[kb + ml] > thr;
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
    typedef const __attribute__((address_space(4))) double cdouble;
    const int nq = (ncol + 3) / 4;
    for (int qi = grp; qi < nq; qi += ngrp) {
        const int c0 = 4 * qi;
        const int i0 = min(c0 + 0, ncol - 1), i1 = min(c0 + 1, ncol - 1), i2 = min(c0 + 2, ncol - 1),
                  i3 = min(c0 + 3, ncol - 1);
        double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
#pragma unroll 4
        for (int row = 0; row < nrow; row++) {
            cdouble* g = (cdouble*)(G + (size_t)row * ncol);
            const double y = Y[row * rw + rhs];
            f0 = fma(g[i0], y, f0);
            f1 = fma(g[i1], y, f1);
            f2 = fma(g[i2], y, f2);
            f3 = fma(g[i3], y, f3);
        }
        if (!valid) continue;
        const double f[4] = {f0, f1, f2, f3};
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (c0 + i < ncol) Fh[((size_t)(c0 + i) * 2 + reim) * nk + k] = f[i];
    }
}

__global__ void __launch_bounds__(256) k_drift_force_adj(drift_args a, const double* __restrict__ Fh,
                                                         double* tauh) {
    const int lane = threadIdx.x & 63, em = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.x * 64 + lane, nk = a.nk;
    if (k >= nk || em >= a.nem) return;
    if (!(a.mnorm[k] > a.skip_tol * a.maxnorm[0])) return;
    const int l = a.kdeg[k], e = a.emap[em];
    const double Jr = (double)l * (double)(l + 1);
    const int qlo = a.mnode[e], qup = a.mnode[e + 1];
    double* t = tauh + (size_t)em * 24 * nk + k;
#pragma unroll
    for (int reim = 0; reim < 2; reim++) {
        // test dofs U_e, U_e+1, V_e, V_e+1 <- (lower, U), (upper, U), (lower, V), (upper, V)
        const double fl[4] = {Fh[((size_t)(2 * qlo + 0) * 2 + reim) * nk + k], Fh[((size_t)(2 * qup + 0) * 2 + reim) * nk + k],
                              Fh[((size_t)(2 * qlo + 1) * 2 + reim) * nk + k], Fh[((size_t)(2 * qup + 1) * 2 + reim) * nk + k]};
        double A[4], B[4], Cc[4];
        dissipative_element_adj(a.r[e], a.r[e + 1], Jr, fl, A, B, Cc);
#pragma unroll
        for (int lam = 0; lam < 4; lam++) {
            t[(size_t)((0 * 2 + reim) * 4 + lam) * nk] += A[lam];
            t[(size_t)((1 * 2 + reim) * 4 + lam) * nk] += B[lam];
            t[(size_t)((2 * 2 + reim) * 4 + lam) * nk] += Cc[lam];
        }
    }
}

// ------------------------------------------------------------------ plan
struct vh_drift_plan {
    int2* jobs = nullptr;
    int2* node_el = nullptr;
    double* Fe = nullptr;
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
        HIPCHECK(hipMalloc((void**)&p->Fe, (size_t)std::max(d.nem, 1) * 8 * d.nk * sizeof(double)));
    }
    const int ncol = 2 * d.nmn, nrow = ncol + 3;
    // right-hand sides per workgroup: as many as fit a 64 KiB forcing tile
    int rw = 64;
    while (rw > 2 && (size_t)std::max(ncol, 1) * rw * 8 > 64 * 1024) rw /= 2;
    if ((size_t)std::max(ncol, 1) * rw * 8 > 160 * 1024)
        throw vh_error("drift: too many memory nodes for the LDS forcing tile (nmn = " + std::to_string(d.nmn) + ")");
    p->rw = rw; p->ts = rw / 2;
    p->lds = (size_t)std::max(ncol, 1) * rw * 8;
    std::vector<int2> jobs;
    for (int l = 1; l <= d.L; l++)
        for (int m0 = 0; m0 <= l; m0 += p->ts) jobs.push_back(make_int2(l, m0));
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
    p->lds_t = (size_t)(std::max(ncol, 1) + 3) * rw * 8;
    if (p->lds_t > 160 * 1024)
        throw vh_error("drift adjoint: too many memory nodes for the LDS tile (nmn = " + std::to_string(d.nmn) + ")");
    switch (p->rw) {
        case 64: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_gT<64>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 32: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_gT<32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 16: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_gT<16>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 8: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_gT<8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        case 4: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_gT<4>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
        default: HIPCHECK(hipFuncSetAttribute((const void*)k_drift_gT<2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds_t)); break;
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
    (void)hipFree(p->Fe);
    delete p;
}

void vh_drift_run(hipStream_t st, const vh_drift_plan* p, const vh_dims& d, const vh_dev& v,
                  const double* tau, const double* mnorm, const double* maxnorm, double skip_tol,
                  int deg1_cm, double* dN, double* dsurf) {
    drift_args a;
    a.nk = d.nk; a.nem = d.nem; a.nmn = d.nmn; a.ts = p->ts; a.rw = p->rw;
    a.jobs = p->jobs; a.G = p->G; a.r = v.r; a.emap = v.emap; a.mnode = v.mnode; a.kdeg = v.kdeg;
    a.node_el = p->node_el; a.Fe = p->Fe;
    a.tau = tau; a.mnorm = mnorm; a.maxnorm = maxnorm;
    a.skip_tol = skip_tol; a.deg1_cm = deg1_cm; a.dN = dN; a.dsurf = dsurf;
    if (d.nem > 0) {
        dim3 g((d.nk + 63) / 64, (d.nem + 3) / 4);
        k_drift_force<<<g, 256, 0, st>>>(a);
        HIPCHECK(hipGetLastError());
    }
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
    if (d.nem == 0) return;
    drift_args a;
    a.nk = d.nk; a.nem = d.nem; a.nmn = d.nmn; a.ts = p->ts; a.rw = p->rw;
    a.jobs = p->jobs; a.G = p->G; a.r = v.r; a.emap = v.emap; a.mnode = v.mnode; a.kdeg = v.kdeg;
    a.node_el = p->node_el; a.Fe = p->Fe;
    a.tau = nullptr; a.mnorm = mnorm; a.maxnorm = maxnorm;
    a.skip_tol = skip_tol; a.deg1_cm = deg1_cm; a.dN = nullptr; a.dsurf = nullptr;
    switch (p->rw) {
        case 64: k_drift_gT<64><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, Fh); break;
        case 32: k_drift_gT<32><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, Fh); break;
        case 16: k_drift_gT<16><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, Fh); break;
        case 8: k_drift_gT<8><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, Fh); break;
        case 4: k_drift_gT<4><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, Fh); break;
        default: k_drift_gT<2><<<p->njobs, DRIFT_THREADS, p->lds_t, st>>>(a, dNh, dsurfh, Fh); break;
    }
    HIPCHECK(hipGetLastError());
    dim3 g((d.nk + 63) / 64, (d.nem + 3) / 4);
    k_drift_force_adj<<<g, 256, 0, st>>>(a, Fh, tauh);
    HIPCHECK(hipGetLastError());
}
