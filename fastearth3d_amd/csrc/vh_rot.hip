// vh_rot.hip -- rotational feedback on the GPU: the degree-2 Liouville polar
// motion with two compact Maxwell channels (loading and tidal)
// (src/vilma_rotation.f90:154-249, :335-363, :408-487).
//
// The channels use the same degree-2 operator as the field driver; their
// frozen drift goes through that operator's solution restricted to the nodes
// of the elements that carry channel memory (formed once from its band LU). Device state (rot_state below) holds the
// channel memory [ch][e][c] (c = (shape*2 + reim)*4 + lambda, every mesh
// element, as the reference carries memory wherever eta > 0), the frozen
// drift and the complex scalars, so a coupling interval needs no host round
// trip for the rotation.
#include <algorithm>
#include "vh_internal.h"
#include "vh_radial_dev.h"
#include "vh_rot.h"

// k_rot_begin: dissipative forcing of the two channel memories (rhs =
// channel*2 + re/im) built in LDS, then the frozen nodal and surface drift
// through the restricted degree-2 operator G2.
__global__ void __launch_bounds__(256) k_rot_begin(rot_args a) {
    extern __shared__ double F[];   // [2 nmn2][4]
    const int nr = a.nr, ne = nr - 1, ncol = 2 * a.nmn2, nrow = ncol + 3;
    for (int t = threadIdx.x; t < ncol * 4; t += blockDim.x) F[t] = 0.0;
    __syncthreads();
    const int rhs = threadIdx.x & 3, grp = threadIdx.x >> 2, ngrp = blockDim.x >> 2;
    const int ch = rhs >> 1, reim = rhs & 1;
    const double Jr = 6.0;
    for (int par = 0; par < 2; par++) {   // consecutive elements share a node: one parity per pass
        for (int em = par + 2 * grp; em < a.nem2; em += 2 * ngrp) {
            const int e = a.emap2[em], q = a.mnode2[e];
            const double* mem = a.rs.mem + ((size_t)ch * ne + e) * 24;
            double A[4], B[4], Cc[4], floc[4];
#pragma unroll
            for (int lam = 0; lam < 4; lam++) {
                A[lam] = mem[(0 * 2 + reim) * 4 + lam];
                B[lam] = mem[(1 * 2 + reim) * 4 + lam];
                Cc[lam] = mem[(2 * 2 + reim) * 4 + lam];
            }
            dissipative_element(a.r[e], a.r[e + 1], Jr, A, B, Cc, floc);
            F[(2 * q) * 4 + rhs] += floc[0];
            F[(2 * q + 2) * 4 + rhs] += floc[1];
            F[(2 * q + 1) * 4 + rhs] += floc[2];
            F[(2 * q + 3) * 4 + rhs] += floc[3];
        }
        __syncthreads();
    }
    for (int row = grp; row < nrow - 1; row += ngrp) {   // the surface V row is not needed
        const double* g = a.G2 + (size_t)row * ncol;
        double y = 0.0;
        for (int c = 0; c < ncol; c++) y = fma(g[c], F[c * 4 + rhs], y);
        if (row < ncol) {
            const int n = a.mnodes2[row >> 1], uv = row & 1;
            a.rs.dN[((size_t)ch * 4 + uv * 2 + reim) * nr + n] = y;
        } else if (row == ncol) {
            a.rs.sc[RS_DU + 2 * ch + reim] = y;
        } else {
            a.rs.sc[RS_DF + 2 * ch + reim] = y;
        }
    }
}

// rotation_s_rot (src/vilma_rotation.f90:192-222)
__global__ void k_rot_srot(rot_args a, int nphi, int nlat, const double* __restrict__ sc2,
                           const double* __restrict__ cphi, const double* __restrict__ sphi, double* srot) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nphi * nlat) return;
    const double* S = a.rs.sc;
    double mr = S[RS_M], mi = S[RS_M + 1];
    double kTe = S[RS_KTE];
    double cNr = mr + (kTe * mr - S[RS_DF + 2]), cNi = mi + (kTe * mi - S[RS_DF + 3]);
    double Ue = S[RS_UE_TIDAL];
    double cUr = Ue * mr + S[RS_DU + 2], cUi = Ue * mi + S[RS_DU + 3];
    double Om = S[RS_OMEGA], ar = S[RS_A], g = S[RS_G];
    double kN = Om * Om * ar * ar / g, ku = Om * Om * ar * ar;
    int j = i / nphi, ip = i % nphi;
    double gam = sc2[j];
    srot[i] = gam * (kN * (cNr * cphi[ip] + cNi * sphi[ip]) - ku * (cUr * cphi[ip] + cUi * sphi[ip]));
}

// per-row partial sums of load * sin cos * (cos phi, sin phi)
__global__ void __launch_bounds__(256) k_rot_inertia_rows(int nphi, int nlat, const double* __restrict__ load,
        const double* __restrict__ sc2, const double* __restrict__ cphi, const double* __restrict__ sphi,
        double* part) {
    __shared__ double r0[256], r1[256];
    int j = blockIdx.x;
    double a0 = 0.0, a1 = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        double v = load[(size_t)j * nphi + i];
        a0 += v * sc2[j] * cphi[i];
        a1 += v * sc2[j] * sphi[i];
    }
    r0[threadIdx.x] = a0; r1[threadIdx.x] = a1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) { r0[threadIdx.x] += r0[threadIdx.x + s]; r1[threadIdx.x] += r1[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[j] = r0[0]; part[nlat + j] = r1[0]; }
}

// solve_m + commit (src/vilma_rotation.f90:170-190, :224-233, :466-487)
__global__ void __launch_bounds__(256) k_rot_solve_commit(rot_args a, double dt) {
    __shared__ double coef[4];
    double* S = a.rs.sc;
    if (threadIdx.x == 0) {
        double aa = S[RS_A];
        double a4 = aa * aa * aa * aa;
        double Ir = -a4 * S[RS_I13], Ii = -a4 * S[RS_I23];
        double scl = 5.0 / (4.0 * VH_PI * VH_G * aa);
        double clr = Ir * scl, cli = Ii * scl;
        double Fe = S[RS_FE_LOAD];
        double Itr = -(Fe * clr + S[RS_DF + 0]), Iti = -(Fe * cli + S[RS_DF + 1]);
        double CA = S[RS_CMA];
        double pr = Itr / CA, pi = Iti / CA;
        double ks = S[RS_KS], kTe = S[RS_KTE];
        double den = 1.0 - kTe / ks;
        double mr = (pr - S[RS_DF + 2] / ks) / den, mi = (pi - S[RS_DF + 3] / ks) / den;
        S[RS_CLOAD] = clr; S[RS_CLOAD + 1] = cli;
        S[RS_M] = mr; S[RS_M + 1] = mi;
        coef[0] = clr; coef[1] = cli; coef[2] = mr; coef[3] = mi;
    }
    __syncthreads();
    const int ne = a.nr - 1, nr = a.nr;
    for (int t = threadIdx.x; t < 2 * ne; t += blockDim.x) {
        int ch = t / ne, e = t % ne;
        double cr = coef[2 * ch], ci = coef[2 * ch + 1];
        const double* xU = a.rs.xU + (size_t)ch * nr;
        const double* xV = a.rs.xV + (size_t)ch * nr;
        const double* dN = a.rs.dN + (size_t)ch * 4 * nr;
        double U0r = cr * xU[e] + dN[0 * nr + e], U1r = cr * xU[e + 1] + dN[0 * nr + e + 1];
        double U0i = ci * xU[e] + dN[1 * nr + e], U1i = ci * xU[e + 1] + dN[1 * nr + e + 1];
        double V0r = cr * xV[e] + dN[2 * nr + e], V1r = cr * xV[e + 1] + dN[2 * nr + e + 1];
        double V0i = ci * xV[e] + dN[3 * nr + e], V1i = ci * xV[e + 1] + dN[3 * nr + e + 1];
        double s[2][3][4];
        strain_coeffs(U0r, U1r, V0r, V1r, 6.0, s[0][0], s[0][1], s[0][2]);
        strain_coeffs(U0i, U1i, V0i, V1i, 6.0, s[1][0], s[1][1], s[1][2]);
        double Me = a.rs.MkPerDt[(size_t)ch * ne + e] * dt;
        double om = 1.0 - Me, two_muM = 2.0 * a.mu[e] * Me;
        double* mem = a.rs.mem + ((size_t)ch * ne + e) * 24;
        for (int sh = 0; sh < 3; sh++)
            for (int ri = 0; ri < 2; ri++)
                for (int lam = 0; lam < 4; lam++) {
                    int c = (sh * 2 + ri) * 4 + lam;
                    mem[c] = om * mem[c] - two_muM * s[ri][sh][lam];
                }
    }
}

void rot_begin(hipStream_t st, const rot_args& a) {
    size_t lds = (size_t)4 * std::max(2 * a.nmn2, 1) * sizeof(double);
    k_rot_begin<<<1, 256, lds, st>>>(a);
    HIPCHECK(hipGetLastError());
}

void rot_srot(hipStream_t st, const rot_args& a, int nphi, int nlat, const double* sc2, const double* cphi,
              const double* sphi, double* srot) {
    int n = nphi * nlat;
    k_rot_srot<<<(n + 255) / 256, 256, 0, st>>>(a, nphi, nlat, sc2, cphi, sphi, srot);
    HIPCHECK(hipGetLastError());
}

void rot_inertia(hipStream_t st, const rot_args& a, const vh_sht& s, const double* load, const double* sc2,
                 const double* cphi, const double* sphi, double* part) {
    k_rot_inertia_rows<<<s.nlat, 256, 0, st>>>(s.nphi, s.nlat, load, sc2, cphi, sphi, part);
    vh_weighted_total(st, s, part, 2, a.rs.sc + RS_I13);
}

void rot_solve_commit(hipStream_t st, const rot_args& a, double dt) {
    k_rot_solve_commit<<<1, 256, 0, st>>>(a, dt);
    HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------------ reverse mode
// Transposes of the channel kernels above (the rotation path is linear in
// the channel memory, m, the drift scalars and the load inertia). Adjoint
// state: rot_adj (vh_rot.h). One block each: the channel problem is tiny.

// solve_m + commit transposed, one sub-interval: mem_h <- (1 - M) mem_h,
// nodal/drift adjoints of the strain term, then the scalar chain
// m = (I_t/C-A - dF_tidal/ks)/den, I_t = -(Fe c_load + dF_load), c_load = -a^4 I scl.
__global__ void __launch_bounds__(256) k_rot_commit_adj(rot_args a, rot_adj h, double dt) {
    __shared__ double red[4][256];
    const int ne = a.nr - 1, nr = a.nr;
    double cacc[4] = {0.0, 0.0, 0.0, 0.0};   // (ch 0 re, ch 0 im, ch 1 re, ch 1 im): adjoints of c_load, m
    for (int par = 0; par < 2; par++) {      // neighbouring elements share a node: one parity per pass
        for (int t = threadIdx.x; t < 2 * ne; t += blockDim.x) {
            const int ch = t / ne, e = t % ne;
            if ((e & 1) != par) continue;
            const double Me = a.rs.MkPerDt[(size_t)ch * ne + e] * dt;
            const double om = 1.0 - Me, two_muM = 2.0 * a.mu[e] * Me;
            double* mh = h.mem + ((size_t)ch * ne + e) * 24;
            double Eh[2][3][4];
            for (int sh = 0; sh < 3; sh++)
                for (int ri = 0; ri < 2; ri++)
                    for (int lam = 0; lam < 4; lam++) {
                        const int c = (sh * 2 + ri) * 4 + lam;
                        const double v = mh[c];
                        mh[c] = om * v;
                        Eh[ri][sh][lam] = -two_muM * v;
                    }
            const double* xU = a.rs.xU + (size_t)ch * nr;
            const double* xV = a.rs.xV + (size_t)ch * nr;
            double* dNh = h.dN + (size_t)ch * 4 * nr;
            for (int ri = 0; ri < 2; ri++) {
                double u1 = 0.0, u2 = 0.0, v1 = 0.0, v2 = 0.0;
                strain_coeffs_adj(Eh[ri][0], Eh[ri][1], Eh[ri][2], 6.0, u1, u2, v1, v2);
                // U = c xU + dN (re: row 0/2, im: row 1/3)
                dNh[(0 + ri) * nr + e] += u1; dNh[(0 + ri) * nr + e + 1] += u2;
                dNh[(2 + ri) * nr + e] += v1; dNh[(2 + ri) * nr + e + 1] += v2;
                cacc[2 * ch + ri] += xU[e] * u1 + xU[e + 1] * u2 + xV[e] * v1 + xV[e + 1] * v2;
            }
        }
        __syncthreads();
    }
    for (int q = 0; q < 4; q++) red[q][threadIdx.x] = cacc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double* S = a.rs.sc;
        const double aa = S[RS_A], a4 = aa * aa * aa * aa, scl = 5.0 / (4.0 * VH_PI * VH_G * aa);
        const double Fe = S[RS_FE_LOAD], CA = S[RS_CMA], ks = S[RS_KS], kTe = S[RS_KTE];
        const double den = 1.0 - kTe / ks;
        for (int ri = 0; ri < 2; ri++) {
            const double mh = red[2 + ri][0] + h.sc[RA_M + ri];           // this sub-interval's m
            const double clh = red[0 + ri][0] - Fe / (CA * den) * mh;     // its c_load
            h.sc[RA_I + ri] += -a4 * scl * clh;
            h.sc[RA_DF + 0 + ri] += -mh / (CA * den);
            h.sc[RA_DF + 2 + ri] += -mh / (ks * den);
            h.sc[RA_M + ri] = 0.0;                                        // m before it does not enter
        }
    }
}

// rotation begin transposed: drift adjoints (dN_h at the channel nodes,
// dU_h, dF_h) -> G2^T -> forcing transpose -> mem_h; the drift adjoints are cleared.
__global__ void __launch_bounds__(256) k_rot_begin_adj(rot_args a, rot_adj h) {
    extern __shared__ double F[];   // [2 nmn2][4] forcing adjoints, then nothing else
    const int nr = a.nr, ne = nr - 1, ncol = 2 * a.nmn2, nrow = ncol + 3;
    const int rhs = threadIdx.x & 3, grp = threadIdx.x >> 2, ngrp = blockDim.x >> 2;
    const int ch = rhs >> 1, reim = rhs & 1;
    for (int c = grp; c < ncol; c += ngrp) {
        double f = 0.0;
        for (int row = 0; row < nrow - 1; row++) {
            double y;
            if (row < ncol) {
                const int n = a.mnodes2[row >> 1], uv = row & 1;
                y = h.dN[((size_t)ch * 4 + uv * 2 + reim) * nr + n];
            } else if (row == ncol) {
                y = h.sc[RA_DU + 2 * ch + reim];
            } else {
                y = h.sc[RA_DF + 2 * ch + reim];
            }
            f = fma(a.G2[(size_t)row * ncol + c], y, f);
        }
        F[c * 4 + rhs] = f;
    }
    __syncthreads();
    for (int em = grp; em < a.nem2; em += ngrp) {
        const int e = a.emap2[em], q = a.mnode2[e];
        const double fl[4] = {F[(2 * q) * 4 + rhs], F[(2 * q + 2) * 4 + rhs], F[(2 * q + 1) * 4 + rhs],
                              F[(2 * q + 3) * 4 + rhs]};
        double A[4], B[4], Cc[4];
        dissipative_element_adj(a.r[e], a.r[e + 1], 6.0, fl, A, B, Cc);
        double* mh = h.mem + ((size_t)ch * ne + e) * 24;
        for (int lam = 0; lam < 4; lam++) {
            mh[(0 * 2 + reim) * 4 + lam] += A[lam];
            mh[(1 * 2 + reim) * 4 + lam] += B[lam];
            mh[(2 * 2 + reim) * 4 + lam] += Cc[lam];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 2 * 4 * nr; t += blockDim.x) h.dN[t] = 0.0;
    if (threadIdx.x < 4) { h.sc[RA_DU + threadIdx.x] = 0.0; h.sc[RA_DF + threadIdx.x] = 0.0; }
}

// rotation_s_rot transposed: tot = sum_ij srot_h gamma_j (cos, sin) -> m_h,
// dF_h (tidal), dU_h (tidal)
__global__ void k_rot_srot_adj(rot_args a, rot_adj h, const double* __restrict__ tot) {
    if (threadIdx.x != 0) return;
    const double* S = a.rs.sc;
    const double kTe = S[RS_KTE], Ue = S[RS_UE_TIDAL], Om = S[RS_OMEGA], ar = S[RS_A], g = S[RS_G];
    const double kN = Om * Om * ar * ar / g, ku = Om * Om * ar * ar;
    for (int ri = 0; ri < 2; ri++) {
        const double cNh = kN * tot[ri], cUh = -ku * tot[ri];
        h.sc[RA_M + ri] += (1.0 + kTe) * cNh + Ue * cUh;
        h.sc[RA_DF + 2 + ri] += -cNh;
        h.sc[RA_DU + 2 + ri] += cUh;
    }
}

// load-inertia transposed: g_ij = gamma_j (I13_h cos phi_i + I23_h sin phi_i)
__global__ void k_rot_inertia_adj(rot_adj h, int nphi, int nlat, const double* __restrict__ sc2,
                                  const double* __restrict__ cphi, const double* __restrict__ sphi, double* g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nphi * nlat) return;
    const int j = i / nphi, ip = i % nphi;
    g[i] = sc2[j] * (h.sc[RA_I] * cphi[ip] + h.sc[RA_I + 1] * sphi[ip]);
}

void rot_commit_adj(hipStream_t st, const rot_args& a, const rot_adj& h, double dt) {
    k_rot_commit_adj<<<1, 256, 0, st>>>(a, h, dt);
    HIPCHECK(hipGetLastError());
}
void rot_begin_adj(hipStream_t st, const rot_args& a, const rot_adj& h) {
    const size_t lds = (size_t)4 * std::max(2 * a.nmn2, 1) * sizeof(double);
    k_rot_begin_adj<<<1, 256, lds, st>>>(a, h);
    HIPCHECK(hipGetLastError());
}
void rot_srot_adj(hipStream_t st, const rot_args& a, const rot_adj& h, const vh_sht& s, const double* srot_h,
                  const double* sc2, const double* cphi, const double* sphi, double* part, double* tot) {
    // sum_ij srot_h gamma (cos, sin) = weighted totals of the density-form rows
    k_rot_inertia_rows<<<s.nlat, 256, 0, st>>>(s.nphi, s.nlat, srot_h, sc2, cphi, sphi, part);
    vh_weighted_total(st, s, part, 2, tot);
    k_rot_srot_adj<<<1, 64, 0, st>>>(a, h, tot);
    HIPCHECK(hipGetLastError());
}
void rot_inertia_adj(hipStream_t st, const rot_adj& h, int nphi, int nlat, const double* sc2, const double* cphi,
                     const double* sphi, double* g) {
    const int n = nphi * nlat;
    k_rot_inertia_adj<<<(n + 255) / 256, 256, 0, st>>>(h, nphi, nlat, sc2, cphi, sphi, g);
    HIPCHECK(hipGetLastError());
}
