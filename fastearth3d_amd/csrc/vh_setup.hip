// vh_setup.hip -- per-degree radial operator assembly, geometric-mean
// equilibration and pivoted band LU factorisation, all degrees in one launch.
//
// One workgroup per degree l >= 2 builds the node-interleaved saddle-point
// operator of Martinec (2000) eqs 80-84 (src/vilma_radial_fe.f90:274-423)
// directly in band form (kl = ku = 6), equilibrates it
// (src/vilma_radial_fe.f90:497-576), factors it with partial pivoting in
// LAPACK dgbtf2 order (src/vilma_band.f90:61-109) and solves the unit surface
// load (src/vilma_response.f90:607-621). Degree 1 is the KKT-bordered system
// (rigid-translation constraint, src/vilma_radial_fe.f90:475-480) whose border
// makes the band full; it gets one workgroup of its own.
#include "vh_internal.h"

// ------------------------------------------------------ element integrals
// Martinec (2000) App. C (src/vilma_radial_integrals.f90:28-126).
struct elint {
    double i1[2][2], i2[2][2], i3[2][2], i4[2][2], i5[2][2], i6[2][2], i7[2][2], k1[2], k2[2];
};

__device__ static void elem_integrals(double rk, double rk1, bool with_i7, elint& E) {
    double h = rk1 - rk;
    double d = (rk1 * rk1 + rk1 * rk + rk * rk) / (3.0 * h);
    E.i1[0][0] = d; E.i1[0][1] = -d; E.i1[1][0] = -d; E.i1[1][1] = d;
    E.i2[0][0] = -(rk1 * rk1 + 2.0 * rk1 * rk + 3.0 * rk * rk) / 12.0;
    E.i2[0][1] = -E.i2[0][0];
    E.i2[1][1] = (3.0 * rk1 * rk1 + 2.0 * rk1 * rk + rk * rk) / 12.0;
    E.i2[1][0] = -E.i2[1][1];
    E.i3[0][0] = -(rk1 + 2.0 * rk) / 6.0;
    E.i3[0][1] = -E.i3[0][0];
    E.i3[1][1] = (2.0 * rk1 + rk) / 6.0;
    E.i3[1][0] = -E.i3[1][1];
    E.i4[0][0] = h / 30.0 * (rk1 * rk1 + 3.0 * rk1 * rk + 6.0 * rk * rk);
    E.i4[0][1] = h / 60.0 * (3.0 * rk1 * rk1 + 4.0 * rk1 * rk + 3.0 * rk * rk);
    E.i4[1][0] = E.i4[0][1];
    E.i4[1][1] = h / 30.0 * (6.0 * rk1 * rk1 + 3.0 * rk1 * rk + rk * rk);
    E.i5[0][0] = h / 12.0 * (rk1 + 3.0 * rk);
    E.i5[0][1] = h / 12.0 * (rk1 + rk);
    E.i5[1][0] = E.i5[0][1];
    E.i5[1][1] = h / 12.0 * (3.0 * rk1 + rk);
    E.i6[0][0] = h / 3.0; E.i6[0][1] = h / 6.0; E.i6[1][0] = h / 6.0; E.i6[1][1] = h / 3.0;
    if (with_i7) {
        double lg = log(rk1 / rk);
        E.i7[0][0] = (rk1 * rk1 / h * lg - 2.0 * rk1 + 0.5 * (rk1 + rk)) / h;
        E.i7[0][1] = (-rk1 * rk / h * lg + 0.5 * (rk1 + rk)) / h;
        E.i7[1][0] = E.i7[0][1];
        E.i7[1][1] = (rk * rk / h * lg - 2.0 * rk + 0.5 * (rk1 + rk)) / h;
    } else {
        E.i7[0][0] = E.i7[0][1] = E.i7[1][0] = E.i7[1][1] = 0.0;
    }
    E.k1[0] = -(rk1 * rk1 + rk1 * rk + rk * rk) / 3.0; E.k1[1] = -E.k1[0];
    E.k2[0] = h / 6.0 * (rk1 + 2.0 * rk); E.k2[1] = h / 6.0 * (2.0 * rk1 + rk);
}

// Local 7x7 block [U1 V1 F1 U2 V2 F2 P] of element e, degree j (eqs 80-82).
__device__ static void element_block(int j, double rlo, double rhi, double mu_k, double rho_k,
                                     double Rk, double Al[7][7]) {
    elint E;
    elem_integrals(rlo, rhi, Rk != 0.0, E);
    const double Jr = (double)j * (double)(j + 1);
    const double fourpiG = 4.0 * VH_PI * VH_G;
    const int lU[2] = {0, 3}, lV[2] = {1, 4}, lF[2] = {2, 5}, lP = 6;
    for (int a = 0; a < 7; a++)
        for (int b = 0; b < 7; b++) Al[a][b] = 0.0;
    for (int ia = 0; ia < 2; ia++) {
        for (int ib = 0; ib < 2; ib++) {
            Al[lU[ib]][lU[ia]] += mu_k * (2.0 * E.i1[ia][ib]);
            Al[lV[ib]][lV[ia]] += mu_k * (Jr * E.i1[ia][ib]);
            Al[lV[ib]][lU[ia]] += mu_k * (Jr * E.i3[ia][ib]);
            Al[lV[ib]][lV[ia]] -= mu_k * (Jr * E.i3[ia][ib]);
            Al[lV[ib]][lV[ia]] -= mu_k * (Jr * E.i3[ib][ia]);
            Al[lU[ib]][lV[ia]] += mu_k * (Jr * E.i3[ib][ia]);
            Al[lU[ib]][lU[ia]] += mu_k * (Jr * E.i6[ia][ib]);
            Al[lU[ib]][lV[ia]] -= mu_k * (Jr * E.i6[ia][ib]);
            Al[lV[ib]][lU[ia]] -= mu_k * (Jr * E.i6[ia][ib]);
            Al[lV[ib]][lV[ia]] += mu_k * (Jr * E.i6[ia][ib]);
            Al[lU[ib]][lU[ia]] += mu_k * (4.0 * E.i6[ia][ib]);
            Al[lV[ib]][lU[ia]] -= mu_k * (2.0 * Jr * E.i6[ia][ib]);
            Al[lU[ib]][lV[ia]] -= mu_k * (2.0 * Jr * E.i6[ia][ib]);
            Al[lV[ib]][lV[ia]] += mu_k * (Jr * Jr * E.i6[ia][ib]);
            Al[lV[ib]][lV[ia]] += mu_k * (Jr * (Jr - 2.0) * E.i6[ia][ib]);
            double gg = (fourpiG / 3.0) * (rho_k * E.i4[ia][ib] + Rk * E.i7[ia][ib]);
            Al[lU[ib]][lU[ia]] += rho_k * (-4.0 * gg + fourpiG * rho_k * E.i4[ia][ib]);
            Al[lU[ib]][lV[ia]] += rho_k * (Jr * gg);
            Al[lU[ib]][lF[ia]] += rho_k * (E.i2[ib][ia]);
            Al[lV[ib]][lU[ia]] += rho_k * (Jr * gg);
            Al[lV[ib]][lF[ia]] += rho_k * (Jr * E.i5[ia][ib]);
            Al[lF[ib]][lF[ia]] += (E.i1[ia][ib] + Jr * E.i6[ia][ib]) / fourpiG;
            Al[lF[ib]][lU[ia]] += rho_k * (E.i2[ia][ib]);
            Al[lF[ib]][lV[ia]] += rho_k * (Jr * E.i5[ia][ib]);
        }
    }
    for (int ia = 0; ia < 2; ia++) {
        Al[lP][lU[ia]] += (E.k1[ia] + 2.0 * E.k2[ia]);
        Al[lU[ia]][lP] += (E.k1[ia] + 2.0 * E.k2[ia]);
        Al[lP][lV[ia]] -= (Jr * E.k2[ia]);
        Al[lV[ia]][lP] -= (Jr * E.k2[ia]);
    }
}

// A(i,j) (|i-j| <= 6) lives at A13[(6 + i - j) + 13*j].
__device__ static void assemble_band13(int j, int nr, const double* r, const double* mu,
                                       const double* rho, const double* Rk, double r_earth,
                                       double* A13) {
    const int ne = nr - 1, nd = 4 * nr - 1;
    for (int t = threadIdx.x; t < 13 * nd; t += blockDim.x) A13[t] = 0.0;
    __syncthreads();
    for (int parity = 0; parity < 2; parity++) {
        for (int e = 2 * threadIdx.x + parity; e < ne; e += 2 * blockDim.x) {
            double Al[7][7];
            element_block(j, r[e], r[e + 1], mu[e], rho[e], Rk[e], Al);
            int g[7] = {4 * e, 4 * e + 1, 4 * e + 2, 4 * e + 4, 4 * e + 5, 4 * e + 6, 4 * e + 3};
            for (int c = 0; c < 7; c++)
                for (int rr = 0; rr < 7; rr++) A13[(6 + g[rr] - g[c]) + 13 * g[c]] += Al[rr][c];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int f = 4 * (nr - 1) + 2;
        A13[6 + 13 * f] += r_earth / (4.0 * VH_PI * VH_G) * (double)(j + 1);
    }
    __syncthreads();
}

// Block-parallel pivoted band LU in dgbtf2 order; ab column-major (ldab, n).
__device__ static bool band_factor_block(int n, int kl, int ku, double* ab, int ldab, int* ipiv,
                                         double* sred, int* ired) {
    const int kv = kl + ku;
    int ju = 0;
    for (int j = 0; j < n; j++) {
        if (j + kv < n)
            for (int i = threadIdx.x; i < kl; i += blockDim.x) ab[i + (size_t)(j + kv) * ldab] = 0.0;
        int km = min(kl, n - 1 - j);
        // arg max |ab(kv+p, j)|, first index on ties (strict >)
        double best = -1.0; int bi = 0;
        for (int p = threadIdx.x; p <= km; p += blockDim.x) {
            double a = fabs(ab[kv + p + (size_t)j * ldab]);
            if (a > best) { best = a; bi = p; }
        }
        sred[threadIdx.x] = best; ired[threadIdx.x] = bi;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                double o = sred[threadIdx.x + s]; int oi = ired[threadIdx.x + s];
                if (o > sred[threadIdx.x] || (o == sred[threadIdx.x] && oi < ired[threadIdx.x])) {
                    sred[threadIdx.x] = o; ired[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        int jp = ired[0];
        __syncthreads();
        if (threadIdx.x == 0) ipiv[j] = j + jp;
        if (ab[kv + jp + (size_t)j * ldab] == 0.0) return false;
        ju = max(ju, min(j + ku + jp, n - 1));
        if (jp != 0) {
            for (int jj = j + threadIdx.x; jj <= ju; jj += blockDim.x) {
                double t = ab[kv + j - jj + (size_t)jj * ldab];
                ab[kv + j - jj + (size_t)jj * ldab] = ab[kv + j + jp - jj + (size_t)jj * ldab];
                ab[kv + j + jp - jj + (size_t)jj * ldab] = t;
            }
        }
        __syncthreads();
        if (km > 0) {
            double rp = 1.0 / ab[kv + (size_t)j * ldab];
            for (int p = 1 + threadIdx.x; p <= km; p += blockDim.x) ab[kv + p + (size_t)j * ldab] *= rp;
            __syncthreads();
            int ncol = ju - j;
            for (int t = threadIdx.x; t < ncol * km; t += blockDim.x) {
                int jj = j + 1 + t / km, p = 1 + t % km;
                double u = ab[kv + j - jj + (size_t)jj * ldab];
                if (u != 0.0) ab[kv + j + p - jj + (size_t)jj * ldab] -= ab[kv + p + (size_t)j * ldab] * u;
            }
        }
        __syncthreads();
    }
    return true;
}

// One RHS x (length n, in place), block-parallel dgbtrs.
__device__ static void band_solve_block(int n, int kl, int kv, const double* ab, int ldab,
                                        const int* ipiv, double* x) {
    for (int j = 0; j < n - 1; j++) {
        int lm = min(kl, n - 1 - j);
        int l = ipiv[j];
        __syncthreads();
        if (threadIdx.x == 0 && l != j) { double t = x[l]; x[l] = x[j]; x[j] = t; }
        __syncthreads();
        double t = x[j];
        for (int i = 1 + threadIdx.x; i <= lm; i += blockDim.x) x[j + i] -= ab[kv + i + (size_t)j * ldab] * t;
    }
    for (int j = n - 1; j >= 0; j--) {
        __syncthreads();
        if (threadIdx.x == 0) x[j] = x[j] / ab[kv + (size_t)j * ldab];
        __syncthreads();
        double t = x[j];
        int lm = min(kv, j);
        for (int i = 1 + threadIdx.x; i <= lm; i += blockDim.x) x[j - i] -= ab[kv - i + (size_t)j * ldab] * t;
    }
    __syncthreads();
}

// rhs_kind 0: unit surface load (src/vilma_radial_fe.f90:616-625);
//          1: unit external potential (tidal, :627-645).
__device__ static void unit_rhs(int rhs_kind, int j, int nr, double r_earth, double g_surf, double* b, int nd) {
    for (int i = threadIdx.x; i < nd; i += blockDim.x) b[i] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        if (rhs_kind == 0) {
            b[4 * (nr - 1)] = -r_earth * r_earth * 1.0 * g_surf;
            b[4 * (nr - 1) + 2] = -r_earth * r_earth * 1.0;
        } else {
            b[4 * (nr - 1) + 2] = -r_earth / (4.0 * VH_PI * VH_G) * (double)(2 * j + 1) * 1.0;
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------- degrees >= 2
// blockIdx.x = l - lmin. work: per block 13*nd (A13) + nd (x).
__global__ void __launch_bounds__(256) k_setup_band(int lmin, int nr, const double* __restrict__ r,
        const double* __restrict__ mu, const double* __restrict__ rho, const double* __restrict__ Rk,
        double r_earth, double g_surf, int rhs_kind, double* work, double* ab, int* ipiv, double* drv,
        double* dcv, double* xfull, double* xUn, double* xVn, double* gu, double* gn, double* gv, int* fail) {
    __shared__ double sred[256];
    __shared__ int ired[256];
    const int l = lmin + blockIdx.x;
    const int nd = 4 * nr - 1;
    double* A13 = work + (size_t)blockIdx.x * 14 * nd;
    double* x = A13 + 13 * (size_t)nd;
    double* abl = ab + (size_t)(l - 2) * VH_LDAB * nd;
    int* ipl = ipiv + (size_t)(l - 2) * nd;
    double* dr = drv + (size_t)(l - 1) * nd;
    double* dc = dcv + (size_t)(l - 1) * nd;
    assemble_band13(l, nr, r, mu, rho, Rk, r_earth, A13);
    for (int k = threadIdx.x; k < nd; k += blockDim.x) {
        double m = 0.0;
        for (int i = max(0, k - 6); i <= min(nd - 1, k + 6); i++) m = fmax(m, fabs(A13[(6 + i - k) + 13 * k]));
        dc[k] = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
        double m = 0.0;
        for (int k = max(0, i - 6); k <= min(nd - 1, i + 6); k++) m = fmax(m, fabs(A13[(6 + i - k) + 13 * k] * dc[k]));
        dr[i] = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < VH_LDAB * nd; t += blockDim.x) abl[t] = 0.0;
    __syncthreads();
    for (int t = threadIdx.x; t < 13 * nd; t += blockDim.x) {
        int k = t / 13, i = k + (t % 13) - 6;
        if (i < 0 || i >= nd) continue;
        double a = A13[t];
        if (a != 0.0) abl[(VH_KV + i - k) + (size_t)VH_LDAB * k] = dr[i] * a * dc[k];
    }
    __syncthreads();
    if (!band_factor_block(nd, VH_KL, VH_KL, abl, VH_LDAB, ipl, sred, ired)) { if (threadIdx.x == 0) atomicAdd(fail, 1); return; }
    __syncthreads();
    unit_rhs(rhs_kind, l, nr, r_earth, g_surf, x, nd);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = dr[i] * x[i];
    band_solve_block(nd, VH_KL, VH_KV, abl, VH_LDAB, ipl, x);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = dc[i] * x[i];
    __syncthreads();
    if (xfull)
        for (int i = threadIdx.x; i < nd; i += blockDim.x) xfull[(size_t)blockIdx.x * nd + i] = x[i];
    if (xUn) {
        double g = g_surf;
        for (int k = threadIdx.x; k < nr; k += blockDim.x) {
            xUn[(size_t)(l - 1) * nr + k] = x[4 * k];
            xVn[(size_t)(l - 1) * nr + k] = x[4 * k + 1];
        }
        if (threadIdx.x == 0) {
            gu[l] = x[4 * (nr - 1)];
            gn[l] = -x[4 * (nr - 1) + 2] / g;
            gv[l] = x[4 * (nr - 1) + 1];
        }
    }
}

void vh_launch_setup_band(hipStream_t st, const vh_dims& d, int lmin, int lmax, const double* r,
                          const double* mu, const double* rho, const double* Rk, double r_earth,
                          double g_surf, int rhs_kind, double* ab, int* ipiv, double* dr, double* dc,
                          double* xfull, double* xUn, double* xVn, double* gu, double* gn, double* gv,
                          int* fail) {
    int nb = lmax - lmin + 1;
    if (nb <= 0) return;
    double* work = nullptr;
    HIPCHECK(hipMalloc((void**)&work, (size_t)nb * 14 * d.nd * sizeof(double)));
    k_setup_band<<<nb, 256, 0, st>>>(lmin, d.nr, r, mu, rho, Rk, r_earth, g_surf, rhs_kind, work, ab,
                                     ipiv, dr, dc, xfull, xUn, xVn, gu, gn, gv, fail);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));   // set-up only: release the scratch right away
    HIPCHECK(hipFree(work));
}

// ------------------------------------------------------------- degree 1
__global__ void __launch_bounds__(256) k_setup_deg1(int nr, const double* __restrict__ r,
        const double* __restrict__ mu, const double* __restrict__ rho, const double* __restrict__ Rk,
        double r_earth, double g_surf, int deg1_cm, double* ab1, int ldab1, int* ipiv1, double* dr,
        double* dc, double* null1, double* xUn, double* xVn, double* gu, double* gn, double* gv,
        double* work, int* fail) {
    __shared__ double sred[256];
    __shared__ int ired[256];
    const int nd = 4 * nr - 1, ns = nd + 1, kl = nd, kv = 2 * nd;
    double* A13 = work;
    double* w = A13 + 13 * (size_t)nd;
    double* x = w + nd;
    assemble_band13(1, nr, r, mu, rho, Rk, r_earth, A13);
    // uniq_weight (src/vilma_radial_fe.f90:425-443), node-parity passes
    for (int i = threadIdx.x; i < nd; i += blockDim.x) w[i] = 0.0;
    __syncthreads();
    for (int parity = 0; parity < 2; parity++) {
        for (int e = 2 * threadIdx.x + parity; e < nr - 1; e += 2 * blockDim.x) {
            double rk = r[e], rk1 = r[e + 1], h = rk1 - rk;
            double k3a = h / 12.0 * (rk1 * rk1 + 2.0 * rk1 * rk + 3.0 * rk * rk);
            double k3b = h / 12.0 * (3.0 * rk1 * rk1 + 2.0 * rk1 * rk + rk * rk);
            w[4 * e] += k3a; w[4 * e + 4] += k3b;
            w[4 * e + 1] += 2.0 * k3a; w[4 * e + 5] += 2.0 * k3b;
        }
        __syncthreads();
    }
    // dc with the border folded in; dc_b = colnorm(w)
    double wmax = 0.0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) wmax = fmax(wmax, fabs(w[i]));
    sred[threadIdx.x] = wmax;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sred[threadIdx.x] = fmax(sred[threadIdx.x], sred[threadIdx.x + s]);
        __syncthreads();
    }
    double dc_b = sred[0] > 0.0 ? 1.0 / sqrt(sred[0]) : 1.0;
    __syncthreads();
    for (int k = threadIdx.x; k < nd; k += blockDim.x) {
        double m = 0.0;
        for (int i = max(0, k - 6); i <= min(nd - 1, k + 6); i++) m = fmax(m, fabs(A13[(6 + i - k) + 13 * k]));
        double v = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
        if (fabs(w[k]) > 1.0 / (v * v)) v = 1.0 / sqrt(fabs(w[k]));
        dc[k] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
        double m = 0.0;
        for (int k = max(0, i - 6); k <= min(nd - 1, i + 6); k++) m = fmax(m, fabs(A13[(6 + i - k) + 13 * k] * dc[k]));
        double v = m > 0.0 ? 1.0 / sqrt(m) : 1.0;
        if (fabs(w[i]) * dc_b > 1.0 / (v * v)) v = 1.0 / sqrt(fabs(w[i]) * dc_b);
        dr[i] = v;
    }
    double rb = 0.0;
    for (int k = threadIdx.x; k < nd; k += blockDim.x) rb = fmax(rb, fabs(w[k] * dc[k]));
    __syncthreads();
    sred[threadIdx.x] = rb;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sred[threadIdx.x] = fmax(sred[threadIdx.x], sred[threadIdx.x + s]);
        __syncthreads();
    }
    double dr_b = sred[0] > 0.0 ? 1.0 / sqrt(sred[0]) : 1.0;
    __syncthreads();
    for (size_t t = threadIdx.x; t < (size_t)ldab1 * ns; t += blockDim.x) ab1[t] = 0.0;
    __syncthreads();
    for (int t = threadIdx.x; t < 13 * nd; t += blockDim.x) {
        int k = t / 13, i = k + (t % 13) - 6;
        if (i < 0 || i >= nd) continue;
        double a = A13[t];
        if (a != 0.0) ab1[(kv + i - k) + (size_t)ldab1 * k] = dr[i] * a * dc[k];
    }
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
        if (w[i] == 0.0) continue;
        ab1[(kv + i - nd) + (size_t)ldab1 * nd] = dr[i] * w[i] * dc_b;      // border column
        ab1[(kv + nd - i) + (size_t)ldab1 * i] = dr_b * w[i] * dc[i];       // border row
    }
    __syncthreads();
    if (!band_factor_block(ns, kl, kl, ab1, ldab1, ipiv1, sred, ired)) { if (threadIdx.x == 0) atomicAdd(fail, 1); return; }
    __syncthreads();
    // null mode: zero physical RHS, border value 1
    for (int i = threadIdx.x; i < ns; i += blockDim.x) x[i] = (i == nd) ? 1.0 : 0.0;
    __syncthreads();
    band_solve_block(ns, kl, kv, ab1, ldab1, ipiv1, x);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) null1[i] = dc[i] * x[i];
    __syncthreads();
    // unit surface load
    unit_rhs(0, 1, nr, r_earth, g_surf, x, nd);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = dr[i] * x[i];
    if (threadIdx.x == 0) x[nd] = 0.0;
    __syncthreads();
    band_solve_block(ns, kl, kv, ab1, ldab1, ipiv1, x);
    for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = dc[i] * x[i];
    __syncthreads();
    if (deg1_cm) {
        double fa_n = null1[4 * (nr - 1) + 2];
        if (fa_n != 0.0) {
            double c = -x[4 * (nr - 1) + 2] / fa_n;
            __syncthreads();
            for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = x[i] + c * null1[i];
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < nr; k += blockDim.x) { xUn[k] = x[4 * k]; xVn[k] = x[4 * k + 1]; }
    if (threadIdx.x == 0) {
        gu[1] = x[4 * (nr - 1)];
        gn[1] = deg1_cm ? -x[4 * (nr - 1) + 2] / g_surf : 0.0;
        gv[1] = x[4 * (nr - 1) + 1];
    }
}

// Degree-1 drift operator: the bordered system's response, through its LU
// factor, to a unit forcing at each memory-node U/V dof, restricted to the
// rows the drift keeps (memory-node U,V, then surface U, F, V) and with the
// centre-of-mass projection folded in. G1[row][col], col = 2*q + (0:U, 1:V).
__global__ void __launch_bounds__(256) k_setup_deg1_op(int nr, int nmn, const int* __restrict__ mnodes,
        const double* __restrict__ ab1, int ldab1, const int* __restrict__ ipiv1, const double* __restrict__ dr,
        const double* __restrict__ dc, const double* __restrict__ null1, int deg1_cm, double* work, double* G1) {
    const int nd = 4 * nr - 1, ns = nd + 1, kl = nd, kv = 2 * nd;
    const int ncol = 2 * nmn, nrow = 2 * nmn + 3;
    double* x = work;
    for (int col = 0; col < ncol; col++) {
        int dof = 4 * mnodes[col >> 1] + (col & 1);
        for (int i = threadIdx.x; i < ns; i += blockDim.x) x[i] = (i == dof) ? dr[i] : 0.0;
        __syncthreads();
        band_solve_block(ns, kl, kv, ab1, ldab1, ipiv1, x);
        for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = dc[i] * x[i];
        __syncthreads();
        if (deg1_cm) {
            int f = 4 * (nr - 1) + 2;
            double fa_n = null1[f];
            if (fa_n != 0.0) {
                double c = -x[f] / fa_n;
                __syncthreads();
                for (int i = threadIdx.x; i < nd; i += blockDim.x) x[i] = x[i] + c * null1[i];
            }
            __syncthreads();
        }
        for (int row = threadIdx.x; row < nrow; row += blockDim.x) {
            int src;
            if (row < 2 * nmn) src = 4 * mnodes[row >> 1] + (row & 1);
            else src = 4 * (nr - 1) + (row - 2 * nmn == 0 ? 0 : (row - 2 * nmn == 1 ? 2 : 1));
            G1[(size_t)row * ncol + col] = x[src];
        }
        __syncthreads();
    }
}

void vh_launch_setup_deg1_op(hipStream_t st, const vh_dims& d, const int* mnodes, const double* ab1,
                             const int* ipiv1, const double* dr1, const double* dc1, const double* null1,
                             int deg1_cm, double* work, double* G1) {
    k_setup_deg1_op<<<1, 256, 0, st>>>(d.nr, d.nmn, mnodes, ab1, d.ldab1, ipiv1, dr1, dc1, null1, deg1_cm, work, G1);
    HIPCHECK(hipGetLastError());
}

void vh_launch_setup_deg1(hipStream_t st, const vh_dims& d, const double* r, const double* mu,
                          const double* rho, const double* Rk, double r_earth, double g_surf,
                          int deg1_cm, double* ab1, int* ipiv1, double* dr1, double* dc1,
                          double* null1, double* xUn, double* xVn, double* gu, double* gn,
                          double* gv, double* work, int* fail) {
    k_setup_deg1<<<1, 256, 0, st>>>(d.nr, r, mu, rho, Rk, r_earth, g_surf, deg1_cm, ab1, d.ldab1,
                                    ipiv1, dr1, dc1, null1, xUn, xVn, gu, gn, gv, work, fail);
    HIPCHECK(hipGetLastError());
}
