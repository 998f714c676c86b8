// vh_radial_dev.h -- per-element Maxwell kernels shared by the drift,
// advance and rotation translation units.
#pragma once
#include "vh_internal.h"

// spheroidal strain coefficients of nodal (U,V) (src/vilma_viscoelastic.f90:329-359)
__device__ __forceinline__ void strain_coeffs(double u1, double u2, double v1, double v2, double Jr,
                                              double a[4], double b[4], double c[4]) {
    a[0] = -u1 + u2; b[0] = 0.0; c[0] = 0.0;
    a[1] = -v1 + v2; b[1] = u1 - v1; c[1] = u2 - v2;
    a[2] = 0.0; b[2] = -u1 / Jr + 0.5 * v1; c[2] = -u2 / Jr + 0.5 * v2;
    a[3] = 0.0; b[3] = 0.5 * v1; c[3] = 0.5 * v2;
    if (Jr <= 2.0) { b[3] = 0.0; c[3] = 0.0; }
}

// toroidal strain coefficients (lambda = 3, 4) of nodal W; no lambda-4 part at
// degree 1 (src/vilma_viscoelastic.f90:373-393)
__device__ __forceinline__ void tor_strain(double w1, double w2, double Jr, double a[2], double b[2], double c[2]) {
    a[0] = -w1 + w2; b[0] = -w1; c[0] = -w2;
    a[1] = 0.0; b[1] = w1; c[1] = w2;
    if (Jr <= 2.0) { b[1] = 0.0; c[1] = 0.0; }
}

// Dissipative forcing of one element for one re/im memory block
// (A,B,C)[lambda] -> floc[4] on the test dofs U_e, U_e+1, V_e, V_e+1.
__device__ __forceinline__ void dissipative_element(double rk, double rk1, double Jr,
                                                    const double A[4], const double B[4],
                                                    const double Cc[4], double floc[4]) {
    const double xg = 0.5773502691896257;
    double norm[4] = {1.0, 0.5 * Jr, 2.0 * Jr * Jr, 2.0 * Jr * (Jr - 2.0)};
    double sa[4][4], sb[4][4], sc[4][4];
    strain_coeffs(1.0, 0.0, 0.0, 0.0, Jr, sa[0], sb[0], sc[0]);
    strain_coeffs(0.0, 1.0, 0.0, 0.0, Jr, sa[1], sb[1], sc[1]);
    strain_coeffs(0.0, 0.0, 1.0, 0.0, Jr, sa[2], sb[2], sc[2]);
    strain_coeffs(0.0, 0.0, 0.0, 1.0, Jr, sa[3], sb[3], sc[3]);
    double h = rk1 - rk;
    floc[0] = floc[1] = floc[2] = floc[3] = 0.0;
#pragma unroll
    for (int ig = 0; ig < 2; ig++) {
        double gp = ig == 0 ? -xg : xg;
        double ra = 0.5 * (h * gp + rk + rk1);
        double psik = (rk1 - ra) / h, psik1 = (ra - rk) / h;
        double tauV[4];
#pragma unroll
        for (int m = 0; m < 4; m++) tauV[m] = A[m] / h + B[m] * psik / ra + Cc[m] * psik1 / ra;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            double D = 0.0;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                double deps = sa[t][m] / h + sb[t][m] * psik / ra + sc[t][m] * psik1 / ra;
                D = D + norm[m] * tauV[m] * deps;
            }
            floc[t] = floc[t] - D * ra * ra * h * 0.5;
        }
    }
}


// Transpose of strain_coeffs: the adjoints of the three shape coefficient
// vectors (ea, eb, ec) -> adjoints of (u1, u2, v1, v2), accumulated.
__device__ __forceinline__ void strain_coeffs_adj(const double ea[4], const double eb[4], const double ec[4],
                                                  double Jr, double& u1, double& u2, double& v1, double& v2) {
    const double b3 = Jr <= 2.0 ? 0.0 : eb[3], c3 = Jr <= 2.0 ? 0.0 : ec[3];
    u1 += -ea[0] + eb[1] - eb[2] / Jr;
    u2 += ea[0] + ec[1] - ec[2] / Jr;
    v1 += -ea[1] - eb[1] + 0.5 * eb[2] + 0.5 * b3;
    v2 += ea[1] - ec[1] + 0.5 * ec[2] + 0.5 * c3;
}

// Transpose of dissipative_element: the adjoint of floc[4] -> adjoints of
// (A, B, C)[lambda] (overwritten).
__device__ __forceinline__ void dissipative_element_adj(double rk, double rk1, double Jr, const double fl[4],
                                                        double A[4], double B[4], double Cc[4]) {
    const double xg = 0.5773502691896257;
    double norm[4] = {1.0, 0.5 * Jr, 2.0 * Jr * Jr, 2.0 * Jr * (Jr - 2.0)};
    double sa[4][4], sb[4][4], sc[4][4];
    strain_coeffs(1.0, 0.0, 0.0, 0.0, Jr, sa[0], sb[0], sc[0]);
    strain_coeffs(0.0, 1.0, 0.0, 0.0, Jr, sa[1], sb[1], sc[1]);
    strain_coeffs(0.0, 0.0, 1.0, 0.0, Jr, sa[2], sb[2], sc[2]);
    strain_coeffs(0.0, 0.0, 0.0, 1.0, Jr, sa[3], sb[3], sc[3]);
    double h = rk1 - rk;
#pragma unroll
    for (int m = 0; m < 4; m++) { A[m] = 0.0; B[m] = 0.0; Cc[m] = 0.0; }
#pragma unroll
    for (int ig = 0; ig < 2; ig++) {
        double gp = ig == 0 ? -xg : xg;
        double ra = 0.5 * (h * gp + rk + rk1);
        double psik = (rk1 - ra) / h, psik1 = (ra - rk) / h;
        const double coef = -(ra * ra * h * 0.5);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            double sv = 0.0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                double deps = sa[t][m] / h + sb[t][m] * psik / ra + sc[t][m] * psik1 / ra;
                sv = sv + fl[t] * deps;
            }
            sv = sv * coef * norm[m];
            A[m] = A[m] + sv / h;
            B[m] = B[m] + sv * psik / ra;
            Cc[m] = Cc[m] + sv * psik1 / ra;
        }
    }
}

// In-place dgbtrs of one lane's right-hand side against the kl = ku = 6 band
// factor of a degree l >= 2 (x[j*stride], equilibrated space).
__device__ __forceinline__ void lane_band_solve(int nd, const double* __restrict__ ab,
                                                const int* __restrict__ ip, double* x, int stride) {
    for (int j = 0; j < nd - 1; j++) {
        int p = ip[j];
        double xj = x[j * stride];
        if (p != j) { double xp = x[p * stride]; x[p * stride] = xj; xj = xp; x[j * stride] = xj; }
        const double* col = ab + (size_t)j * VH_LDAB + VH_KV;
        int lm = min(VH_KL, nd - 1 - j);
#pragma unroll
        for (int i = 1; i <= VH_KL; i++)
            if (i <= lm) x[(j + i) * stride] -= col[i] * xj;
    }
    for (int j = nd - 1; j >= 0; j--) {
        const double* col = ab + (size_t)j * VH_LDAB + VH_KV;
        double xj = x[j * stride] / col[0];
        x[j * stride] = xj;
        int lm = min(VH_KV, j);
#pragma unroll
        for (int i = 1; i <= VH_KV; i++)
            if (i <= lm) x[(j - i) * stride] -= col[-i] * xj;
    }
}
