/*
 * orc_sht.c -- oracle: Gauss-grid spherical-harmonic transforms and the
 * dyadic tensor-SH transforms.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h). The reference calls SHTns 3.7.5
 * (not vendored; src/vilma_sht.f90:23, :135-153) with orthonormal harmonics,
 * no Condon-Shortley phase, a Gauss grid, phi-contiguous layout and m >= 0
 * spectra. This file restates the transforms SHTns computes for that
 * configuration from their definitions: exact Gauss quadrature in latitude,
 * DFT in longitude (real field = Q_l0 Y_l0 + 2 Re sum_{m>0} Q_lm Y_lm), the
 * spheroidal vector pair (d_theta, d_phi/sin) and its analysis as the
 * quadrature adjoint divided by l(l+1). SHTns' polar optimisation (eps) is
 * not restated: tables are exact. The tensor-SH part restates
 * src/vilma_tensor_sh.f90:76-133 (init + n6 calibration), :168-266 (synth),
 * :270-358 (analysis, spin-2 adjoint).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* Gauss-Legendre nodes x (descending, north first) and weights (sum 2). */
static void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(ORC_PI * (i + 0.75) / (n + 0.5)), z1, pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int k = 1; k <= n; k++) {
                double p3 = p2; p2 = p1;
                p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z; z = z1 - p1 / pp;
            if (fabs(z - z1) <= 1e-16) break;
        }
        /* one more evaluation at the converged node for the weight */
        double p1 = 1.0, p2 = 0.0;
        for (int k = 1; k <= n; k++) {
            double p3 = p2; p2 = p1;
            p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        x[i] = z; x[n - 1 - i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp); w[n - 1 - i] = w[i];
    }
    if (n % 2) x[n / 2] = 0.0;
}

/* Orthonormal associated Legendre functions without CS phase, and their
 * theta-derivatives, at one colatitude (cos t = x, sin t = s), for order m,
 * degrees m..lmax. */
static void legendre_m(int lmax, int m, double x, double s, double* p, double* dp) {
    double pmm = 1.0 / sqrt(4.0 * ORC_PI);
    for (int k = 1; k <= m; k++) pmm *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    p[0] = pmm;
    if (m < lmax) p[1] = sqrt(2.0 * m + 3.0) * x * pmm;
    for (int l = m + 2; l <= lmax; l++) {
        double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        double b = sqrt((((double)l - 1) * (l - 1) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
        p[l - m] = a * (x * p[l - m - 1] - b * p[l - m - 2]);
    }
    for (int l = m; l <= lmax; l++) {
        double pl1 = (l > m) ? p[l - m - 1] : 0.0;
        double c = sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) * ((double)l * l - (double)m * m));
        dp[l - m] = (l * x * p[l - m] - (l > m ? c * pl1 : 0.0)) / s;
    }
}

/* --------------------------------------------------------------- FFT */

static void fft_rec(const cplx* in, cplx* out, int n, int stride, const int* fac,
                    const cplx* tw, int twstride, int N, int inv) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = fac[0], m = n / p;
    for (int q = 0; q < p; q++)
        fft_rec(in + q * stride, out + q * m, m, stride * p, fac + 1, tw, twstride * p, N, inv);
    cplx y[16], z[16];
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            int t = (q * k) % n;
            cplx wv = tw[(size_t)t * twstride];
            if (inv) wv = conj(wv);
            y[q] = out[k + q * m] * wv;
        }
        for (int u = 0; u < p; u++) {
            cplx acc = 0.0;
            for (int q = 0; q < p; q++) {
                int t = (q * u * m) % n;
                cplx wv = tw[(size_t)t * twstride];
                if (inv) wv = conj(wv);
                acc += y[q] * wv;
            }
            z[u] = acc;
        }
        for (int u = 0; u < p; u++) out[k + u * m] = z[u];
    }
    (void)N;
}

static void fft(const orc_sht* s, const cplx* in, cplx* out, int inv) {
    fft_rec(in, out, s->nphi, 1, s->fac, s->tw, 1, s->nphi, inv);
}

/* Forward DFT of real rows: F[row*(lmax+1)+m] = (2 pi/nphi) sum_i f_i e^{-i m phi_i}. */
void orc_fft_rows_forward(const orc_sht* s, const double* f, int nrows, cplx* F) {
    int N = s->nphi, L = s->lmax;
    double sc = 2.0 * ORC_PI / N;
#pragma omp parallel for schedule(static) if (nrows > 8)
    for (int rp = 0; rp < (nrows + 1) / 2; rp++) {
        int r0 = 2 * rp, r1 = 2 * rp + 1;
        cplx* z = (cplx*)malloc(N * sizeof(cplx));
        cplx* Z = (cplx*)malloc(N * sizeof(cplx));
        for (int i = 0; i < N; i++) z[i] = f[(size_t)r0 * N + i] + I * (r1 < nrows ? f[(size_t)r1 * N + i] : 0.0);
        fft(s, z, Z, 0);
        for (int m = 0; m <= L; m++) {
            cplx a = Z[m], b = conj(Z[(N - m) % N]);
            F[(size_t)r0 * (L + 1) + m] = sc * 0.5 * (a + b);
            if (r1 < nrows) F[(size_t)r1 * (L + 1) + m] = sc * (a - b) / (2.0 * I);
        }
        free(z); free(Z);
    }
}

/* Inverse: f_i = Re F_0 + 2 Re sum_{m>=1} F_m e^{i m phi_i}. */
void orc_fft_rows_inverse(const orc_sht* s, const cplx* F, int nrows, double* f) {
    int N = s->nphi, L = s->lmax;
#pragma omp parallel for schedule(static) if (nrows > 8)
    for (int rp = 0; rp < (nrows + 1) / 2; rp++) {
        int r0 = 2 * rp, r1 = 2 * rp + 1;
        cplx* z = (cplx*)calloc(N, sizeof(cplx));
        cplx* Z = (cplx*)malloc(N * sizeof(cplx));
        for (int m = 0; m <= L; m++) {
            cplx a = F[(size_t)r0 * (L + 1) + m];
            cplx b = (r1 < nrows) ? F[(size_t)r1 * (L + 1) + m] : 0.0;
            if (m == 0) { a = creal(a); b = creal(b); }
            z[m] += a + I * b;
            if (m > 0) z[N - m] += conj(a) + I * conj(b);
        }
        fft(s, z, Z, 1);
        for (int i = 0; i < N; i++) {
            f[(size_t)r0 * N + i] = creal(Z[i]);
            if (r1 < nrows) f[(size_t)r1 * N + i] = cimag(Z[i]);
        }
        free(z); free(Z);
    }
}

/* --------------------------------------------------------------- init */

int orc_sht_init(orc_sht* s, int lmax, int nlat, int nphi, char* err) {
    memset(s, 0, sizeof(*s));
    if (nphi <= 2 * lmax) { snprintf(err, 256, "sht_grid_init: nphi must exceed 2*lmax"); return 1; }
    if (nlat <= lmax) { snprintf(err, 256, "sht_grid_init: nlat must exceed lmax"); return 1; }
    s->lmax = lmax; s->mmax = lmax; s->nlat = nlat; s->nphi = nphi;
    s->nlm = (lmax + 1) * (lmax + 2) / 2;
    s->nh = (nlat + 1) / 2;
    s->colat = malloc(nlat * sizeof(double)); s->ct = malloc(nlat * sizeof(double));
    s->st = malloc(nlat * sizeof(double)); s->w = malloc(nlat * sizeof(double));
    s->lon = malloc(nphi * sizeof(double));
    gauss_legendre(nlat, s->ct, s->w);
    for (int j = 0; j < nlat; j++) { s->colat[j] = acos(s->ct[j]); s->st[j] = sqrt((1.0 - s->ct[j]) * (1.0 + s->ct[j])); }
    for (int i = 0; i < nphi; i++) s->lon[i] = 2.0 * ORC_PI * i / nphi;
    s->mstart = malloc((lmax + 1) * sizeof(int));
    int off = 0;
    for (int m = 0; m <= lmax; m++) { s->mstart[m] = off; off += lmax - m + 1; }
    s->plm = malloc((size_t)s->nlm * s->nh * sizeof(double));
    s->dplm = malloc((size_t)s->nlm * s->nh * sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= lmax; m++) {
        double* p = malloc((lmax + 1) * sizeof(double));
        double* dp = malloc((lmax + 1) * sizeof(double));
        for (int j = 0; j < s->nh; j++) {
            legendre_m(lmax, m, s->ct[j], s->st[j], p, dp);
            for (int l = m; l <= lmax; l++) {
                s->plm[(size_t)(s->mstart[m] + l - m) * s->nh + j] = p[l - m];
                s->dplm[(size_t)(s->mstart[m] + l - m) * s->nh + j] = dp[l - m];
            }
        }
        free(p); free(dp);
    }
    /* FFT factors + twiddles */
    int n = nphi, nf = 0;
    const int pref[5] = {4, 2, 3, 5, 7};
    for (int q = 0; q < 5 && n > 1; q++)
        while (n % pref[q] == 0) { s->fac[nf++] = pref[q]; n /= pref[q]; }
    for (int p = 11; n > 1; p += 2) {
        while (n % p == 0) {
            if (p > 16) { snprintf(err, 256, "sht_grid_init: nphi has a prime factor > 16"); return 1; }
            s->fac[nf++] = p; n /= p;
        }
    }
    s->nfac = nf;
    s->tw = malloc(nphi * sizeof(cplx));
    for (int t = 0; t < nphi; t++) s->tw[t] = cexp(-2.0 * I * ORC_PI * t / nphi);
    return 0;
}

void orc_sht_free(orc_sht* s) {
    free(s->colat); free(s->ct); free(s->st); free(s->w); free(s->lon);
    free(s->mstart); free(s->plm); free(s->dplm); free(s->tw);
    memset(s, 0, sizeof(*s));
}

/* src/vilma_sht.f90:184-202 */
double orc_surface_integral(const orc_sht* s, const double* f) {
    double dlon = 2.0 * ORC_PI / s->nphi, total = 0.0;
    for (int j = 0; j < s->nlat; j++) {
        double row = 0.0;
        for (int i = 0; i < s->nphi; i++) row += f[(size_t)j * s->nphi + i];
        total += s->w[j] * row * dlon;
    }
    return total;
}

/* south row of north row j (-1 for the equator of an odd grid) */
static inline int south(const orc_sht* s, int j) {
    int js = s->nlat - 1 - j;
    return js == j ? -1 : js;
}

/* Legendre part of a scalar analysis from row spectra F (nlat x (L+1)). */
static void legendre_analysis(const orc_sht* s, const cplx* F, cplx* q) {
    int L = s->lmax, nh = s->nh;
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= L; m++) {
        cplx* E = malloc(nh * sizeof(cplx));
        cplx* O = malloc(nh * sizeof(cplx));
        for (int j = 0; j < nh; j++) {
            int js = south(s, j);
            cplx fn = F[(size_t)j * (L + 1) + m], fs = js >= 0 ? F[(size_t)js * (L + 1) + m] : 0.0;
            E[j] = s->w[j] * (fn + fs);
            O[j] = s->w[j] * (fn - fs);
        }
        for (int l = m; l <= L; l++) {
            const double* P = s->plm + (size_t)(s->mstart[m] + l - m) * nh;
            const cplx* G = ((l + m) % 2 == 0) ? E : O;
            cplx acc = 0.0;
            for (int j = 0; j < nh; j++) acc += P[j] * G[j];
            q[s->mstart[m] + l - m] = acc;
        }
        free(E); free(O);
    }
}

/* Legendre part of a scalar synthesis: F (nlat x (L+1)) from q. */
static void legendre_synthesis(const orc_sht* s, const cplx* q, cplx* F) {
    int L = s->lmax, nh = s->nh;
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= L; m++) {
        for (int j = 0; j < nh; j++) {
            cplx e = 0.0, o = 0.0;
            for (int l = m; l <= L; l++) {
                double P = s->plm[(size_t)(s->mstart[m] + l - m) * nh + j];
                cplx v = q[s->mstart[m] + l - m] * P;
                if ((l + m) % 2 == 0) e += v; else o += v;
            }
            int js = south(s, j);
            F[(size_t)j * (L + 1) + m] = e + o;
            if (js >= 0) F[(size_t)js * (L + 1) + m] = e - o;
        }
    }
}

void orc_sht_analysis(const orc_sht* s, const double* f, cplx* q) {
    cplx* F = malloc((size_t)s->nlat * (s->lmax + 1) * sizeof(cplx));
    orc_fft_rows_forward(s, f, s->nlat, F);
    legendre_analysis(s, F, q);
    free(F);
}

void orc_sht_synthesis(const orc_sht* s, const cplx* q, double* f) {
    cplx* F = malloc((size_t)s->nlat * (s->lmax + 1) * sizeof(cplx));
    legendre_synthesis(s, q, F);
    orc_fft_rows_inverse(s, F, s->nlat, f);
    free(F);
}

/* Spheroidal synthesis (vt, vp) = (d_theta, (1/sin) d_phi) of sum q Y
 * (src/vilma_sht.f90:228-240). */
void orc_sph_synthesis(const orc_sht* s, const cplx* q, double* vt, double* vp) {
    int L = s->lmax, nh = s->nh, nl = s->nlat;
    cplx* FT = malloc((size_t)nl * (L + 1) * sizeof(cplx));
    cplx* FP = malloc((size_t)nl * (L + 1) * sizeof(cplx));
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= L; m++) {
        for (int j = 0; j < nh; j++) {
            cplx e = 0.0, o = 0.0, de = 0.0, dO = 0.0;
            for (int l = m; l <= L; l++) {
                size_t id = (size_t)(s->mstart[m] + l - m);
                cplx c = q[id];
                double P = s->plm[id * nh + j], D = s->dplm[id * nh + j];
                if ((l + m) % 2 == 0) { e += c * P; de += c * D; }
                else { o += c * P; dO += c * D; }
            }
            cplx im_s = I * (double)m / s->st[j];
            int js = south(s, j);
            FT[(size_t)j * (L + 1) + m] = de + dO;
            FP[(size_t)j * (L + 1) + m] = im_s * (e + o);
            if (js >= 0) {
                FT[(size_t)js * (L + 1) + m] = dO - de;
                FP[(size_t)js * (L + 1) + m] = im_s * (e - o);
            }
        }
    }
    orc_fft_rows_inverse(s, FT, nl, vt);
    orc_fft_rows_inverse(s, FP, nl, vp);
    free(FT); free(FP);
}

/* Spheroidal part of the vector analysis (spat_to_SHsphtor's S, the
 * inverse of orc_sph_synthesis; src/vilma_sht.f90:242-256). */
void orc_sph_analysis(const orc_sht* s, const double* vt, const double* vp, cplx* q) {
    int L = s->lmax, nh = s->nh, nl = s->nlat;
    cplx* FT = malloc((size_t)nl * (L + 1) * sizeof(cplx));
    cplx* FP = malloc((size_t)nl * (L + 1) * sizeof(cplx));
    orc_fft_rows_forward(s, vt, nl, FT);
    orc_fft_rows_forward(s, vp, nl, FP);
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= L; m++) {
        cplx* Te = malloc(nh * sizeof(cplx)); cplx* To = malloc(nh * sizeof(cplx));
        cplx* Pe = malloc(nh * sizeof(cplx)); cplx* Po = malloc(nh * sizeof(cplx));
        for (int j = 0; j < nh; j++) {
            int js = south(s, j);
            cplx tn = FT[(size_t)j * (L + 1) + m], ts = js >= 0 ? FT[(size_t)js * (L + 1) + m] : 0.0;
            cplx pn = FP[(size_t)j * (L + 1) + m], ps = js >= 0 ? FP[(size_t)js * (L + 1) + m] : 0.0;
            cplx mim_s = -I * (double)m / s->st[j];
            /* even l+m: dP(south) = -dP(north), P(south) = P(north) */
            Te[j] = s->w[j] * (tn - ts); To[j] = s->w[j] * (tn + ts);
            Pe[j] = s->w[j] * mim_s * (pn + ps); Po[j] = s->w[j] * mim_s * (pn - ps);
        }
        for (int l = m; l <= L; l++) {
            size_t id = (size_t)(s->mstart[m] + l - m);
            if (l == 0) { q[id] = 0.0; continue; }
            const double* P = s->plm + id * nh;
            const double* D = s->dplm + id * nh;
            int ev = (l + m) % 2 == 0;
            const cplx* T = ev ? Te : To;
            const cplx* PP = ev ? Pe : Po;
            cplx acc = 0.0;
            for (int j = 0; j < nh; j++) acc += D[j] * T[j] + P[j] * PP[j];
            q[id] = acc / ((double)l * (l + 1.0));
        }
        free(Te); free(To); free(Pe); free(Po);
    }
    free(FT); free(FP);
}

/* Toroidal synthesis (vt, vp) = e_r x grad_1 (sum q Y) = (-F, E), the
 * module's orientation (src/vilma_sht.f90:258-277: the negated SHTns field). */
void orc_tor_synthesis(const orc_sht* s, const cplx* q, double* vt, double* vp) {
    int L = s->lmax, nh = s->nh, nl = s->nlat;
    cplx* FT = malloc((size_t)nl * (L + 1) * sizeof(cplx));
    cplx* FP = malloc((size_t)nl * (L + 1) * sizeof(cplx));
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= L; m++) {
        for (int j = 0; j < nh; j++) {
            cplx e = 0.0, o = 0.0, de = 0.0, dO = 0.0;
            for (int l = m; l <= L; l++) {
                size_t id = (size_t)(s->mstart[m] + l - m);
                cplx c = q[id];
                double P = s->plm[id * nh + j], D = s->dplm[id * nh + j];
                if ((l + m) % 2 == 0) { e += c * P; de += c * D; }
                else { o += c * P; dO += c * D; }
            }
            cplx mim_s = -I * (double)m / s->st[j];
            int js = south(s, j);
            FT[(size_t)j * (L + 1) + m] = mim_s * (e + o);
            FP[(size_t)j * (L + 1) + m] = de + dO;
            if (js >= 0) {
                FT[(size_t)js * (L + 1) + m] = mim_s * (e - o);
                FP[(size_t)js * (L + 1) + m] = dO - de;
            }
        }
    }
    orc_fft_rows_inverse(s, FT, nl, vt);
    orc_fft_rows_inverse(s, FP, nl, vp);
    free(FT); free(FP);
}

/* Toroidal part of the vector analysis in the same orientation
 * (src/vilma_sht.f90:293-308): q_lm = (1/J) int v . conj(e_r x grad_1 Y_lm). */
void orc_tor_analysis(const orc_sht* s, const double* vt, const double* vp, cplx* q) {
    int L = s->lmax, nh = s->nh, nl = s->nlat;
    cplx* FT = malloc((size_t)nl * (L + 1) * sizeof(cplx));
    cplx* FP = malloc((size_t)nl * (L + 1) * sizeof(cplx));
    orc_fft_rows_forward(s, vt, nl, FT);
    orc_fft_rows_forward(s, vp, nl, FP);
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= L; m++) {
        cplx* Te = malloc(nh * sizeof(cplx)); cplx* To = malloc(nh * sizeof(cplx));
        cplx* Pe = malloc(nh * sizeof(cplx)); cplx* Po = malloc(nh * sizeof(cplx));
        for (int j = 0; j < nh; j++) {
            int js = south(s, j);
            cplx tn = FT[(size_t)j * (L + 1) + m], ts = js >= 0 ? FT[(size_t)js * (L + 1) + m] : 0.0;
            cplx pn = FP[(size_t)j * (L + 1) + m], ps = js >= 0 ? FP[(size_t)js * (L + 1) + m] : 0.0;
            cplx im_s = I * (double)m / s->st[j];
            /* vt pairs with (i m/s) P, vp with dP; even l+m: P symmetric, dP antisymmetric */
            Te[j] = s->w[j] * im_s * (tn + ts); To[j] = s->w[j] * im_s * (tn - ts);
            Pe[j] = s->w[j] * (pn - ps); Po[j] = s->w[j] * (pn + ps);
        }
        for (int l = m; l <= L; l++) {
            size_t id = (size_t)(s->mstart[m] + l - m);
            if (l == 0) { q[id] = 0.0; continue; }
            const double* P = s->plm + id * nh;
            const double* D = s->dplm + id * nh;
            int ev = (l + m) % 2 == 0;
            const cplx* T = ev ? Te : To;
            const cplx* PP = ev ? Pe : Po;
            cplx acc = 0.0;
            for (int j = 0; j < nh; j++) acc += P[j] * T[j] + D[j] * PP[j];
            q[id] = acc / ((double)l * (l + 1.0));
        }
        free(Te); free(To); free(Pe); free(Po);
    }
    free(FT); free(FP);
}

/* Point evaluation of a scalar field (sht_grid_eval_point, src/vilma_sht.f90:310-325). */
double orc_eval_point(const orc_sht* s, const cplx* q, double colat, double lon) {
    int L = s->lmax;
    double x = cos(colat), st = sin(colat), val = 0.0;
    double* p = malloc((L + 1) * sizeof(double));
    double* dp = malloc((L + 1) * sizeof(double));
    for (int m = 0; m <= L; m++) {
        legendre_m(L, m, x, st > 0 ? st : 1e-300, p, dp);
        cplx acc = 0.0;
        for (int l = m; l <= L; l++) acc += q[s->mstart[m] + l - m] * p[l - m];
        cplx e = cexp(I * m * lon);
        val += (m == 0) ? creal(q[s->mstart[0] + 0] * 0.0 + acc) : 2.0 * creal(acc * e);
    }
    free(p); free(dp);
    return val;
}

/* ---------------------------------------------------------- tensor SH */

static void byprof(const orc_sht* s, const double* f, const double* prof, double* out) {
    for (int j = 0; j < s->nlat; j++)
        for (int i = 0; i < s->nphi; i++) out[(size_t)j * s->nphi + i] = f[(size_t)j * s->nphi + i] * prof[j];
}

/* spin-2 grid fields Sg = sum c G, Sh = sum c H (src/vilma_tensor_sh.f90:244-266). */
static void spin2_fields(const orc_tsh* t, const orc_sht* s, const cplx* c6, double* sg, double* sh) {
    size_t ng = (size_t)s->nphi * s->nlat;
    int nlm = s->nlm;
    cplx* imc = malloc(nlm * sizeof(cplx));
    cplx* lc = malloc(nlm * sizeof(cplx));
    double *gt = malloc(ng * 8), *gp = malloc(ng * 8), *gtf = malloc(ng * 8), *gpf = malloc(ng * 8), *lap = malloc(ng * 8);
    double* tmp = malloc(ng * 8);
    for (int lm = 0; lm < nlm; lm++) { imc[lm] = I * (double)t->mord[lm] * c6[lm]; lc[lm] = -t->llp1[lm] * c6[lm]; }
    orc_sph_synthesis(s, c6, gt, gp);
    orc_sph_synthesis(s, imc, gtf, gpf);
    orc_sht_synthesis(s, lc, lap);
    for (int j = 0; j < s->nlat; j++)
        for (int i = 0; i < s->nphi; i++) {
            size_t id = (size_t)j * s->nphi + i;
            sg[id] = lap[id] - 2.0 * (gt[id] * t->cott[j]) - 2.0 * (gpf[id] * t->invsin[j]);
            sh[id] = (gtf[id] * t->invsin[j]) - (gp[id] * t->cott[j]);
        }
    free(imc); free(lc); free(gt); free(gp); free(gtf); free(gpf); free(lap); free(tmp);
}

/* unnormalised adjoint of the spin-2 synthesis (src/vilma_tensor_sh.f90:324-358). */
static void spin2_adjoint(const orc_tsh* t, const orc_sht* s, const double* dtt, const double* dpp,
                          const double* dtp, cplx* craw) {
    size_t ng = (size_t)s->nphi * s->nlat;
    int nlm = s->nlm;
    double *D = malloc(ng * 8), *vt = malloc(ng * 8), *vp = malloc(ng * 8);
    cplx* q = malloc(nlm * sizeof(cplx));
    cplx* sq = malloc(nlm * sizeof(cplx));
    for (size_t id = 0; id < ng; id++) D[id] = dtt[id] - dpp[id];
    orc_sht_analysis(s, D, q);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = -t->llp1[lm] * q[lm];
    byprof(s, D, t->cott, vt); memset(vp, 0, ng * 8);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] - 2.0 * t->llp1[lm] * sq[lm];
    memset(vt, 0, ng * 8); byprof(s, D, t->invsin, vp);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] + 2.0 * (I * (double)t->mord[lm]) * t->llp1[lm] * sq[lm];
    byprof(s, dtp, t->invsin, vt); memset(vp, 0, ng * 8);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] - 2.0 * (I * (double)t->mord[lm]) * t->llp1[lm] * sq[lm];
    memset(vt, 0, ng * 8); byprof(s, dtp, t->cott, vp);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] - 2.0 * t->llp1[lm] * sq[lm];
    free(D); free(vt); free(vp); free(q); free(sq);
}

/* unnormalised adjoint of the toroidal spin-2 synthesis (src/vilma_tensor_sh.f90:
 * 360-387): 1/2 Sg*(tp) - Sh*(tt - pp) */
static void spin2_adjoint_tor(const orc_tsh* t, const orc_sht* s, const double* dtt, const double* dpp,
                              const double* dtp, cplx* craw) {
    size_t ng = (size_t)s->nphi * s->nlat;
    int nlm = s->nlm;
    double *D = malloc(ng * 8), *vt = malloc(ng * 8), *vp = malloc(ng * 8);
    cplx* q = malloc(nlm * sizeof(cplx));
    cplx* sq = malloc(nlm * sizeof(cplx));
    orc_sht_analysis(s, dtp, q);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = -0.5 * t->llp1[lm] * q[lm];
    byprof(s, dtp, t->cott, vt); memset(vp, 0, ng * 8);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] - t->llp1[lm] * sq[lm];
    memset(vt, 0, ng * 8); byprof(s, dtp, t->invsin, vp);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] + (I * (double)t->mord[lm]) * t->llp1[lm] * sq[lm];
    for (size_t id = 0; id < ng; id++) D[id] = dtt[id] - dpp[id];
    byprof(s, D, t->invsin, vt); memset(vp, 0, ng * 8);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] + (I * (double)t->mord[lm]) * t->llp1[lm] * sq[lm];
    memset(vt, 0, ng * 8); byprof(s, D, t->cott, vp);
    orc_sph_analysis(s, vt, vp, sq);
    for (int lm = 0; lm < nlm; lm++) craw[lm] = craw[lm] + t->llp1[lm] * sq[lm];
    free(D); free(vt); free(vp); free(q); free(sq);
}

/* src/vilma_tensor_sh.f90:76-133 */
int orc_tsh_init(orc_tsh* t, const orc_sht* s) {
    memset(t, 0, sizeof(*t));
    t->lmax = s->lmax; t->nlm = s->nlm; t->nphi = s->nphi; t->nlat = s->nlat;
    t->ldeg = malloc(s->nlm * sizeof(int)); t->mord = malloc(s->nlm * sizeof(int));
    t->llp1 = malloc(s->nlm * sizeof(double));
    for (int m = 0; m <= s->lmax; m++)
        for (int l = m; l <= s->lmax; l++) {
            int lm = orc_lmidx(s, l, m);
            t->ldeg[lm] = l; t->mord[lm] = m; t->llp1[lm] = (double)l * (double)(l + 1);
        }
    t->cott = malloc(s->nlat * sizeof(double)); t->invsin = malloc(s->nlat * sizeof(double));
    for (int j = 0; j < s->nlat; j++) {
        t->cott[j] = cos(s->colat[j]) / sin(s->colat[j]);
        t->invsin[j] = 1.0 / sin(s->colat[j]);
    }
    t->n6 = calloc(s->lmax + 1, sizeof(double));
    t->n4 = calloc(s->lmax + 1, sizeof(double));
    size_t ng = (size_t)s->nphi * s->nlat;
    cplx* c6 = calloc(s->nlm, sizeof(cplx));
    cplx* craw = malloc(s->nlm * sizeof(cplx));
    double *sg = malloc(ng * 8), *sh = malloc(ng * 8), *tt = malloc(ng * 8), *pp = malloc(ng * 8), *tp = malloc(ng * 8);
    for (int l = 2; l <= s->lmax; l++) {
        int m = l;                                    /* sectoral, mmax = lmax */
        int lm = orc_lmidx(s, l, m);
        memset(c6, 0, s->nlm * sizeof(cplx));
        c6[lm] = 1.0;
        spin2_fields(t, s, c6, sg, sh);
        for (size_t id = 0; id < ng; id++) { tt[id] = sg[id]; pp[id] = -sg[id]; tp[id] = 4.0 * sh[id]; }
        spin2_adjoint(t, s, tt, pp, tp, craw);
        t->n6[l] = creal(craw[lm]);
        /* the same coefficient as Z4: tt = -Sh, pp = +Sh, tp = Sg */
        for (size_t id = 0; id < ng; id++) { tt[id] = -sh[id]; pp[id] = sh[id]; tp[id] = sg[id]; }
        spin2_adjoint_tor(t, s, tt, pp, tp, craw);
        t->n4[l] = creal(craw[lm]);
    }
    free(c6); free(craw); free(sg); free(sh); free(tt); free(pp); free(tp);
    return 0;
}

void orc_tsh_free(orc_tsh* t) {
    free(t->ldeg); free(t->mord); free(t->llp1); free(t->cott); free(t->invsin); free(t->n6); free(t->n4);
    memset(t, 0, sizeof(*t));
}

/* Tensor coefficients c[lam*nlm + lm] (lam = 1,2,5,6 -> 0..3; with nch = 6
 * also lam = 3,4 -> 4,5) -> six dyadic planes rr, rt, rp, tt, tp, pp
 * (src/vilma_tensor_sh.f90:168-242). */
void orc_tsh_synth(const orc_tsh* t, const orc_sht* s, const cplx* c, int nch, double* dyad) {
    size_t ng = (size_t)s->nphi * s->nlat;
    int nlm = s->nlm;
    double *tr = malloc(ng * 8), *sg = malloc(ng * 8), *sh = malloc(ng * 8);
    cplx* sc = malloc(nlm * sizeof(cplx));
    orc_sht_synthesis(s, c, dyad + 0 * ng);
    orc_sph_synthesis(s, c + nlm, dyad + 1 * ng, dyad + 2 * ng);
    for (int lm = 0; lm < nlm; lm++) sc[lm] = -t->llp1[lm] * c[2 * nlm + lm];
    orc_sht_synthesis(s, sc, tr);
    spin2_fields(t, s, c + 3 * nlm, sg, sh);
    for (size_t id = 0; id < ng; id++) {
        dyad[3 * ng + id] = tr[id] + sg[id];          /* tt */
        dyad[5 * ng + id] = tr[id] - sg[id];          /* pp */
        dyad[4 * ng + id] = 4.0 * sh[id];             /* tp */
    }
    if (nch == 6) {
        /* Z3: (rt, rp) += e_r x grad_1; Z4: (tt, pp, tp) += (-Sh, +Sh, Sg) */
        double *vt = malloc(ng * 8), *vp = malloc(ng * 8);
        orc_tor_synthesis(s, c + 4 * nlm, vt, vp);
        spin2_fields(t, s, c + 5 * nlm, sg, sh);
        for (size_t id = 0; id < ng; id++) {
            dyad[1 * ng + id] += vt[id];
            dyad[2 * ng + id] += vp[id];
            dyad[3 * ng + id] += -sh[id];
            dyad[5 * ng + id] += sh[id];
            dyad[4 * ng + id] += sg[id];
        }
        free(vt); free(vp);
    }
    free(tr); free(sg); free(sh); free(sc);
}

/* Six dyadic planes -> tensor coefficients, nch = 4 or 6 channels
 * (src/vilma_tensor_sh.f90:270-322). */
void orc_tsh_analysis(const orc_tsh* t, const orc_sht* s, const double* dyad, int nch, cplx* c) {
    size_t ng = (size_t)s->nphi * s->nlat;
    int nlm = s->nlm;
    double* tmp = malloc(ng * 8);
    cplx* craw = malloc(nlm * sizeof(cplx));
    orc_sht_analysis(s, dyad + 0 * ng, c);
    orc_sph_analysis(s, dyad + 1 * ng, dyad + 2 * ng, c + nlm);
    for (size_t id = 0; id < ng; id++) tmp[id] = dyad[3 * ng + id] + dyad[5 * ng + id];
    orc_sht_analysis(s, tmp, craw);
    for (int lm = 0; lm < nlm; lm++)
        c[2 * nlm + lm] = t->llp1[lm] > 0.0 ? craw[lm] / (-2.0 * t->llp1[lm]) : 0.0;
    spin2_adjoint(t, s, dyad + 3 * ng, dyad + 5 * ng, dyad + 4 * ng, craw);
    for (int lm = 0; lm < nlm; lm++) {
        double n6 = t->n6[t->ldeg[lm]];
        c[3 * nlm + lm] = n6 != 0.0 ? craw[lm] / n6 : 0.0;
    }
    if (nch == 6) {
        orc_tor_analysis(s, dyad + 1 * ng, dyad + 2 * ng, c + 4 * nlm);
        spin2_adjoint_tor(t, s, dyad + 3 * ng, dyad + 5 * ng, dyad + 4 * ng, craw);
        for (int lm = 0; lm < nlm; lm++) {
            double n4 = t->n4[t->ldeg[lm]];
            c[5 * nlm + lm] = n4 != 0.0 ? craw[lm] / n4 : 0.0;
        }
    }
    free(tmp); free(craw);
}
