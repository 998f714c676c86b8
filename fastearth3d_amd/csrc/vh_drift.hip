// vh_drift.hip -- (a) the frozen relaxation drift of every (l,m) slot: the
// dissipative memory forcing -int tau:delta eps (2-point radial Gauss,
// src/vilma_viscoelastic.f90:441-480) followed by the equilibrated band
// solve (src/vilma_band.f90:111-142 via src/vilma_radial_fe.f90:578-605),
// with the reference's skip gate (mnorm <= skip_tol*max mnorm => zero drift,
// src/vilma_response.f90:711-760) and the degree-1 centre-of-mass projection
// (:386-420, :770-779).
//
// Mapping: one 64-lane wavefront per (degree l, 64 right-hand sides); lane =
// (order m, re/im) of that degree, so all lanes share one band factor, which
// is read through the scalar cache (uniform addresses), while each lane's
// solution vector lives in LDS as x[j*64 + lane] (bank-conflict free).
// Degree 1 (the KKT-bordered, effectively dense system) runs as one extra
// wavefront of the same launch that solves its four right-hand sides
// cooperatively.
#include "vh_internal.h"
#include "vh_radial_dev.h"

struct drift_args {
    int nr, nd, nk, nem, nmn, lpb;
    const int2* jobs;
    vh_dev v;
    const double* tau;
    const double* mnorm;
    const double* maxnorm;
    double skip_tol;
    int deg1_cm, ns1, ldab1;
    double* dN;
    double* dsurf;
};

// RHS of one lane: accumulate f (physical) into x (stride 64), then scale by dr.
__device__ __forceinline__ void build_rhs(const drift_args& a, int l, int k, int reim, double* x, int stride,
                                          const double* dr) {
    const double Jr = (double)l * (double)(l + 1);
    for (int em = 0; em < a.nem; em++) {
        int e = a.v.emap[em];
        const double* t = a.tau + (size_t)em * 24 * a.nk + k;
        double A[4], B[4], Cc[4], floc[4];
#pragma unroll
        for (int lam = 0; lam < 4; lam++) {
            A[lam] = t[(size_t)((0 * 2 + reim) * 4 + lam) * a.nk];
            B[lam] = t[(size_t)((1 * 2 + reim) * 4 + lam) * a.nk];
            Cc[lam] = t[(size_t)((2 * 2 + reim) * 4 + lam) * a.nk];
        }
        dissipative_element(a.v.r[e], a.v.r[e + 1], Jr, A, B, Cc, floc);
        x[(4 * e) * stride] += floc[0];
        x[(4 * e + 4) * stride] += floc[1];
        x[(4 * e + 1) * stride] += floc[2];
        x[(4 * e + 5) * stride] += floc[3];
    }
    for (int j = 0; j < a.nd; j++) x[j * stride] = dr[j] * x[j * stride];
}

// writes the outputs of one solved lane (x = equilibrated solution y, stride)
__device__ __forceinline__ void store_lane(const drift_args& a, int l, int k, int reim, bool act,
                                           const double* x, int stride, const double* dc) {
    const int nk = a.nk, nr = a.nr;
    for (int q = 0; q < a.nmn; q++) {
        int n = a.v.mnodes[q];
        double U = act ? dc[4 * n] * x[(4 * n) * stride] : 0.0;
        double V = act ? dc[4 * n + 1] * x[(4 * n + 1) * stride] : 0.0;
        a.dN[(size_t)(q * 4 + reim) * nk + k] = U;
        a.dN[(size_t)(q * 4 + 2 + reim) * nk + k] = V;
    }
    int s = nr - 1;
    double Ua = act ? dc[4 * s] * x[(4 * s) * stride] : 0.0;
    double Fa = act ? dc[4 * s + 2] * x[(4 * s + 2) * stride] : 0.0;
    double Va = act ? dc[4 * s + 1] * x[(4 * s + 1) * stride] : 0.0;
    if (l == 1 && !a.deg1_cm) Fa = 0.0;
    a.dsurf[(size_t)(0 * nk + k) * 2 + reim] = Ua;
    a.dsurf[(size_t)(1 * nk + k) * 2 + reim] = Fa;
    a.dsurf[(size_t)(2 * nk + k) * 2 + reim] = Va;
}

__global__ void __launch_bounds__(64) k_drift(drift_args a) {
    extern __shared__ double xs[];
    const int2 job = a.jobs[blockIdx.x];
    const int l = job.x, lane = threadIdx.x;
    const double thr = a.skip_tol * a.maxnorm[0];
    const int nd = a.nd;
    if (l == 1) {
        // ---- degree 1: 4 RHS (m = 0,1 x re/im), bordered system of size ns1,
        // solved cooperatively by the wavefront (x[r*ns1 + j]).
        const int ns = a.ns1, kl = nd, kv = 2 * nd, ldab = a.ldab1;
        const double* dr = a.v.dr;   // index l-1 = 0
        const double* dc = a.v.dc;
        bool act4[4];
        for (int r = 0; r < 4; r++) act4[r] = a.mnorm[r % 2] > thr;
        for (int t = lane; t < 4 * ns; t += 64) xs[t] = 0.0;
        __syncthreads();
        if (lane < 4) {
            int m = lane % 2, reim = lane / 2;
            if (act4[lane]) build_rhs(a, 1, m, reim, xs + lane * ns, 1, dr);
        }
        __syncthreads();
        for (int j = 0; j < ns - 1; j++) {
            int lm = min(kl, ns - 1 - j);
            int p = a.v.ipiv1[j];
            if (lane < 4 && p != j) {
                double t = xs[lane * ns + p]; xs[lane * ns + p] = xs[lane * ns + j]; xs[lane * ns + j] = t;
            }
            __syncthreads();
            double t0 = xs[0 * ns + j], t1 = xs[1 * ns + j], t2 = xs[2 * ns + j], t3 = xs[3 * ns + j];
            const double* col = a.v.ab1 + (size_t)j * ldab + kv;
            for (int i = 1 + lane; i <= lm; i += 64) {
                double Lij = col[i];
                xs[0 * ns + j + i] -= Lij * t0;
                xs[1 * ns + j + i] -= Lij * t1;
                xs[2 * ns + j + i] -= Lij * t2;
                xs[3 * ns + j + i] -= Lij * t3;
            }
            __syncthreads();
        }
        for (int j = ns - 1; j >= 0; j--) {
            const double* col = a.v.ab1 + (size_t)j * ldab + kv;
            if (lane < 4) xs[lane * ns + j] = xs[lane * ns + j] / col[0];
            __syncthreads();
            double t0 = xs[0 * ns + j], t1 = xs[1 * ns + j], t2 = xs[2 * ns + j], t3 = xs[3 * ns + j];
            int lm = min(kv, j);
            for (int i = 1 + lane; i <= lm; i += 64) {
                double Uij = col[-i];
                xs[0 * ns + j - i] -= Uij * t0;
                xs[1 * ns + j - i] -= Uij * t1;
                xs[2 * ns + j - i] -= Uij * t2;
                xs[3 * ns + j - i] -= Uij * t3;
            }
            __syncthreads();
        }
        // physical solution, CM projection, outputs (one RHS per lane < 4)
        for (int t = lane; t < 4 * nd; t += 64) { int r = t / nd, i = t % nd; xs[r * ns + i] = dc[i] * xs[r * ns + i]; }
        __syncthreads();
        if (lane < 4) {
            double* x = xs + lane * ns;
            if (a.deg1_cm) {
                int f = 4 * (a.nr - 1) + 2;
                double fa_n = a.v.null1[f];
                if (fa_n != 0.0) {
                    double c = -x[f] / fa_n;
                    for (int i = 0; i < nd; i++) x[i] = x[i] + c * a.v.null1[i];
                }
            }
            // dc already applied: store with unit scaling
            const int nk = a.nk, nr = a.nr, m = lane % 2, reim = lane / 2;
            bool act = act4[lane];
            for (int q = 0; q < a.nmn; q++) {
                int n = a.v.mnodes[q];
                a.dN[(size_t)(q * 4 + reim) * nk + m] = act ? x[4 * n] : 0.0;
                a.dN[(size_t)(q * 4 + 2 + reim) * nk + m] = act ? x[4 * n + 1] : 0.0;
            }
            int s = nr - 1;
            a.dsurf[(size_t)(0 * nk + m) * 2 + reim] = act ? x[4 * s] : 0.0;
            a.dsurf[(size_t)(1 * nk + m) * 2 + reim] = (act && a.deg1_cm) ? x[4 * s + 2] : 0.0;
            a.dsurf[(size_t)(2 * nk + m) * 2 + reim] = act ? x[4 * s + 1] : 0.0;
        }
        return;
    }
    // ---- degrees >= 2: one lane per RHS
    const int nrhs = 2 * (l + 1);
    const int rr = job.y + lane;
    const bool valid = lane < a.lpb && rr < nrhs;
    const int m = valid ? rr % (l + 1) : 0, reim = valid ? rr / (l + 1) : 0;
    const int k = l * (l + 1) / 2 - 1 + m;
    const bool act = valid && a.mnorm[k] > thr;
    double* x = xs + lane;
    const double* dr = a.v.dr + (size_t)(l - 1) * nd;
    const double* dc = a.v.dc + (size_t)(l - 1) * nd;
    const double* ab = a.v.ab + (size_t)(l - 2) * VH_LDAB * nd;
    const int* ip = a.v.ipiv + (size_t)(l - 2) * nd;
    for (int j = 0; j < nd; j++) x[j * 64] = 0.0;
    if (act) build_rhs(a, l, k, reim, x, 64, dr);
    // forward: P and L (dgbtrs), kl = 6
    for (int j = 0; j < nd - 1; j++) {
        int p = ip[j];
        double xj = x[j * 64];
        if (p != j) { double xp = x[p * 64]; x[p * 64] = xj; xj = xp; x[j * 64] = xj; }
        const double* col = ab + (size_t)j * VH_LDAB + VH_KV;
        int lm = min(VH_KL, nd - 1 - j);
#pragma unroll
        for (int i = 1; i <= VH_KL; i++)
            if (i <= lm) x[(j + i) * 64] -= col[i] * xj;
    }
    // back: U x = y, kv = 12
    for (int j = nd - 1; j >= 0; j--) {
        const double* col = ab + (size_t)j * VH_LDAB + VH_KV;
        double xj = x[j * 64] / col[0];
        x[j * 64] = xj;
        int lm = min(VH_KV, j);
#pragma unroll
        for (int i = 1; i <= VH_KV; i++)
            if (i <= lm) x[(j - i) * 64] -= col[-i] * xj;
    }
    if (valid) store_lane(a, l, k, reim, act, x, 64, dc);
}

// ------------------------------------------------------------------ plan
struct vh_drift_plan {
    int2* jobs = nullptr;
    int njobs = 0, lpb = 64;
    size_t lds = 0;
};

vh_drift_plan* vh_drift_plan_create(const vh_dims& d) {
    vh_drift_plan* p = new vh_drift_plan();
    // lanes per wavefront limited by LDS (x: lanes*nd doubles)
    int lpb = 64;
    while (lpb > 1 && (size_t)64 * d.nd * 8 > 160 * 1024 && (size_t)lpb * d.nd * 8 > 150 * 1024) lpb /= 2;
    p->lpb = lpb;
    std::vector<int2> jobs;
    if (d.L >= 1) jobs.push_back(make_int2(1, 0));
    for (int l = 2; l <= d.L; l++)
        for (int r0 = 0; r0 < 2 * (l + 1); r0 += lpb) jobs.push_back(make_int2(l, r0));
    p->njobs = (int)jobs.size();
    HIPCHECK(hipMalloc((void**)&p->jobs, jobs.size() * sizeof(int2)));
    HIPCHECK(hipMemcpy(p->jobs, jobs.data(), jobs.size() * sizeof(int2), hipMemcpyHostToDevice));
    size_t lds_band = (size_t)64 * d.nd * 8;
    size_t lds_deg1 = (size_t)4 * d.ns1 * 8;
    p->lds = std::max(lds_band, lds_deg1);
    if (p->lds > 160 * 1024) throw vh_error("drift: radial system too large for LDS (nd = " + std::to_string(d.nd) + ")");
    HIPCHECK(hipFuncSetAttribute((const void*)k_drift, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds));
    return p;
}

void vh_drift_plan_destroy(vh_drift_plan* p) {
    if (!p) return;
    (void)hipFree(p->jobs);
    delete p;
}

void vh_drift_run(hipStream_t st, const vh_drift_plan* p, const vh_dims& d, const vh_dev& v,
                  const double* tau, const double* mnorm, const double* maxnorm, double skip_tol,
                  int deg1_cm, double* dN, double* dsurf) {
    drift_args a;
    a.nr = d.nr; a.nd = d.nd; a.nk = d.nk; a.nem = d.nem; a.nmn = d.nmn; a.lpb = p->lpb;
    a.jobs = p->jobs; a.v = v; a.tau = tau; a.mnorm = mnorm; a.maxnorm = maxnorm;
    a.skip_tol = skip_tol; a.deg1_cm = deg1_cm; a.ns1 = d.ns1; a.ldab1 = d.ldab1;
    a.dN = dN; a.dsurf = dsurf;
    k_drift<<<p->njobs, 64, p->lds, st>>>(a);
    HIPCHECK(hipGetLastError());
}
