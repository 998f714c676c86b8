// vh_sht.hip -- Gauss-grid spherical-harmonic transforms on the GPU.
//
// Conventions restate the reference's SHTns configuration
// (src/vilma_sht.f90:11-15, :135-153): orthonormal harmonics without the
// Condon-Shortley phase, Gauss-Legendre latitudes north-first, longitudes
// 2 pi i/nphi, a real field = Q_l0 Y_l0 + 2 Re sum_{m>0} Q_lm Y_lm, spheroidal
// vector pair (d_theta, d_phi/sin theta) with its analysis the quadrature
// adjoint divided by l(l+1). Tables are exact (no polar truncation).
//
// Kernels:
//   * k_legendre_tables: P_lm and dP_lm/dtheta for the northern half grid, in
//     two layouts: [m][l][j] (synthesis reads rows of j) and [m][j][l]
//     (analysis reads rows of l), so both directions read coalesced.
//   * longitude transforms: a Stockham mixed-radix complex FFT per wavefront
//     in LDS; two real rows travel as one complex row.
//   * Legendre sums use the equatorial parity fold P_lm(pi-t) = (-1)^{l+m} P_lm(t).
#include "vh_internal.h"
#include "vh_sht_dev.h"

// ------------------------------------------------------------- tables
__global__ void k_legendre_tables(int L, int nh, const double* __restrict__ ct, const double* __restrict__ st,
                                  const int* __restrict__ mstart, double* plm_lj, double* dplm_lj,
                                  double* plm_jl, double* dplm_jl) {
    int m = blockIdx.x;
    int j = blockIdx.y * blockDim.x + threadIdx.x;
    if (j >= nh || m > L) return;
    double x = ct[j], s = st[j];
    const int nl = L - m + 1;
    const size_t base = (size_t)mstart[m] * nh;
    double pmm = 1.0 / sqrt(4.0 * VH_PI);
    for (int k = 1; k <= m; k++) pmm *= sqrt((2.0 * k + 1.0) / (2.0 * k)) * s;
    double pm2 = 0.0, pm1 = 0.0;   // P_{l-2,m}, P_{l-1,m}
    for (int l = m; l <= L; l++) {
        double p;
        if (l == m) p = pmm;
        else if (l == m + 1) p = sqrt(2.0 * m + 3.0) * x * pmm;
        else {
            double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            double b = sqrt((((double)l - 1) * (l - 1) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
            p = a * (x * pm1 - b * pm2);
        }
        // dP/dtheta = (l cos P_l - sqrt((2l+1)/(2l-1) (l^2-m^2)) P_{l-1}) / sin
        double c = sqrt((2.0 * l + 1.0) / (2.0 * l - 1.0) * ((double)l * l - (double)m * m));
        double dp = (l * x * p - (l > m ? c * pm1 : 0.0)) / s;
        plm_lj[base + (size_t)(l - m) * nh + j] = p;
        dplm_lj[base + (size_t)(l - m) * nh + j] = dp;
        plm_jl[base + (size_t)j * nl + (l - m)] = p;
        dplm_jl[base + (size_t)j * nl + (l - m)] = dp;
        pm2 = pm1;
        pm1 = p;
    }
}

void vh_setup_legendre_tables(hipStream_t st, const vh_sht& s, const double* ct, const double* sn,
                              double* plm_lj, double* dplm_lj, double* plm_jl, double* dplm_jl) {
    dim3 g(s.L + 1, (s.nh + 63) / 64);
    k_legendre_tables<<<g, 64, 0, st>>>(s.L, s.nh, ct, sn, s.mstart, plm_lj, dplm_lj, plm_jl, dplm_jl);
    HIPCHECK(hipGetLastError());
}

// ------------------------------------------------------ Legendre sums
// F rows: field f, latitude j at row f*fs + j*js (fs = nlat, js = 1 for
// stacked fields; fs = 1, js = nb for fields interleaved per latitude);
// q layout [f][nlm] degree-major.

// Analysis: block = (order m, chunk of 64 degrees) from s.ljobs; thread =
// (degree in chunk, quarter of the north latitudes). The latitude pair is
// folded into even/odd l+m partial fields first; the four quarter sums are
// added in a fixed order (deterministic).
__global__ void __launch_bounds__(256) k_leg_analysis(vh_sht s, const double2* __restrict__ F, int nb,
                                                      double2* __restrict__ q, int fst, int jst) {
    extern __shared__ double2 sh[];
    const int2 job = s.ljobs[blockIdx.x];
    const int m = job.x, il0 = job.y, L = s.L, nh = s.nh, nl = L - m + 1;
    double2* E = sh;                        // [nb][nh]
    double2* O = sh + nb * nh;              // [nb][nh]
    double2* red = sh + 2 * nb * nh;        // [4][64][nb]
    for (int t = threadIdx.x; t < nb * nh; t += blockDim.x) {
        int f = t / nh, j = t % nh, js = s.nlat - 1 - j;
        const double2* Ff = F + (size_t)f * fst * (L + 1);
        double2 fn = Ff[(size_t)j * jst * (L + 1) + m];
        double2 fs = js != j ? Ff[(size_t)js * jst * (L + 1) + m] : make_double2(0.0, 0.0);
        double w = s.w[j];
        E[t] = make_double2(w * (fn.x + fs.x), w * (fn.y + fs.y));
        O[t] = make_double2(w * (fn.x - fs.x), w * (fn.y - fs.y));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int il = il0 + lane;
    const int ilc = il < nl ? il : nl - 1;
    const int jq = (nh + 3) / 4, ja = g * jq, jb = min(nh, ja + jq);
    const double* P = s.plm_jl + (size_t)s.mstart[m] * nh;
    const double2* G = (ilc & 1) ? O : E;   // parity of l + m = 2m + il
    for (int f = 0; f < nb; f++) {
        double ax = 0.0, ay = 0.0;
        const double2* Gf = G + f * nh;
        for (int j = ja; j < jb; j++) {
            const double p = P[(size_t)j * nl + ilc];
            ax = fma(p, Gf[j].x, ax); ay = fma(p, Gf[j].y, ay);
        }
        red[(g * 64 + lane) * nb + f] = make_double2(ax, ay);
    }
    __syncthreads();
    if (g == 0 && il < nl) {
        const int l = m + il;
        for (int f = 0; f < nb; f++) {
            double2 r = red[lane * nb + f];
            for (int k = 1; k < 4; k++) r = make_double2(r.x + red[(k * 64 + lane) * nb + f].x,
                                                         r.y + red[(k * 64 + lane) * nb + f].y);
            q[(size_t)f * s.nlm + l * (l + 1) / 2 + m] = r;
        }
    }
}

static size_t leg_anal_lds(const vh_sht& s, int nb) { return ((size_t)2 * nb * s.nh + 4 * 64 * nb) * sizeof(double2); }


void vh_legendre_analysis(hipStream_t st, const vh_sht& s, const double2* F, int nb, double2* q) {
    k_leg_analysis<<<s.nljobs, 256, leg_anal_lds(s, nb), st>>>(s, F, nb, q, s.nlat, 1);
    HIPCHECK(hipGetLastError());
}


void vh_legendre_synthesis(hipStream_t st, const vh_sht& s, const double2* q, int nb, double2* F) {
    vh_legendre_synthesis_t(st, s, SpecSrc{q, s.nlm}, nb, F, s.nlat, 1);
    HIPCHECK(hipGetLastError());
}

void vh_legendre_synthesis_rows(hipStream_t st, const vh_sht& s, const double2* q, int nb, double2* F,
                                int fst, int jst) {
    vh_legendre_synthesis_t(st, s, SpecSrc{q, s.nlm}, nb, F, fst, jst);
    HIPCHECK(hipGetLastError());
}

// spheroidal synthesis: FT = sum q dP (theta comp.), FP = (i m/sin) sum q P
__global__ void __launch_bounds__(256) k_sph_synthesis(vh_sht s, const double2* __restrict__ q, int nb,
                                                       double2* __restrict__ FT, double2* __restrict__ FP) {
    extern __shared__ double2 sh[];
    const int m = blockIdx.x, L = s.L, nh = s.nh, nl = L - m + 1;
    for (int t = threadIdx.x; t < nb * nl; t += blockDim.x) {
        int f = t / nl, il = t % nl, l = m + il;
        sh[t] = q[(size_t)f * s.nlm + l * (l + 1) / 2 + m];
    }
    __syncthreads();
    const double* P = s.plm_lj + (size_t)s.mstart[m] * nh;
    const double* D = s.dplm_lj + (size_t)s.mstart[m] * nh;
    for (int t = threadIdx.x; t < nb * nh; t += blockDim.x) {
        int f = t / nh, j = t % nh, js = s.nlat - 1 - j;
        const double2* Q = sh + f * nl;
        double ex = 0, ey = 0, ox = 0, oy = 0, dex = 0, dey = 0, dox = 0, doy = 0;
        for (int il = 0; il < nl; il++) {
            double p = P[(size_t)il * nh + j], d = D[(size_t)il * nh + j];
            if (il & 1) { ox += Q[il].x * p; oy += Q[il].y * p; dox += Q[il].x * d; doy += Q[il].y * d; }
            else { ex += Q[il].x * p; ey += Q[il].y * p; dex += Q[il].x * d; dey += Q[il].y * d; }
        }
        double ms = (double)m / s.st[j];
        size_t o = (size_t)f * s.nlat * (L + 1);
        FT[o + (size_t)j * (L + 1) + m] = make_double2(dex + dox, dey + doy);
        FP[o + (size_t)j * (L + 1) + m] = make_double2(-ms * (ey + oy), ms * (ex + ox));
        if (js != j) {
            FT[o + (size_t)js * (L + 1) + m] = make_double2(dox - dex, doy - dey);
            FP[o + (size_t)js * (L + 1) + m] = make_double2(-ms * (ey - oy), ms * (ex - ox));
        }
    }
}

// spheroidal part of the vector analysis (inverse of k_sph_synthesis)
__global__ void __launch_bounds__(256) k_sph_analysis(vh_sht s, const double2* __restrict__ FT,
                                                      const double2* __restrict__ FP, int nb,
                                                      double2* __restrict__ q) {
    extern __shared__ double2 sh[];
    const int m = blockIdx.x, L = s.L, nh = s.nh, nl = L - m + 1;
    double2* Te = sh;
    double2* To = sh + nb * nh;
    double2* Pe = sh + 2 * nb * nh;
    double2* Po = sh + 3 * nb * nh;
    for (int t = threadIdx.x; t < nb * nh; t += blockDim.x) {
        int f = t / nh, j = t % nh, js = s.nlat - 1 - j;
        size_t o = (size_t)f * s.nlat * (L + 1);
        double2 tn = FT[o + (size_t)j * (L + 1) + m], pn = FP[o + (size_t)j * (L + 1) + m];
        double2 ts = make_double2(0, 0), ps = make_double2(0, 0);
        if (js != j) { ts = FT[o + (size_t)js * (L + 1) + m]; ps = FP[o + (size_t)js * (L + 1) + m]; }
        double w = s.w[j], ms = (double)m / s.st[j];
        Te[t] = make_double2(w * (tn.x - ts.x), w * (tn.y - ts.y));
        To[t] = make_double2(w * (tn.x + ts.x), w * (tn.y + ts.y));
        // (-i m/s) * z = (ms*z.y, -ms*z.x)
        double2 a = make_double2(pn.x + ps.x, pn.y + ps.y), b = make_double2(pn.x - ps.x, pn.y - ps.y);
        Pe[t] = make_double2(w * (ms * a.y), w * (-ms * a.x));
        Po[t] = make_double2(w * (ms * b.y), w * (-ms * b.x));
    }
    __syncthreads();
    const double* P = s.plm_jl + (size_t)s.mstart[m] * nh;
    const double* D = s.dplm_jl + (size_t)s.mstart[m] * nh;
    for (int il = threadIdx.x; il < nl; il += blockDim.x) {
        int l = m + il;
        bool ev = ((l + m) & 1) == 0;
        for (int f = 0; f < nb; f++) {
            const double2* T = (ev ? Te : To) + f * nh;
            const double2* PP = (ev ? Pe : Po) + f * nh;
            double ax = 0.0, ay = 0.0;
            for (int j = 0; j < nh; j++) {
                double d = D[(size_t)j * nl + il], p = P[(size_t)j * nl + il];
                ax += d * T[j].x + p * PP[j].x;
                ay += d * T[j].y + p * PP[j].y;
            }
            double2 r = make_double2(0.0, 0.0);
            if (l > 0) { double inv = (double)l * (l + 1.0); r = make_double2(ax / inv, ay / inv); }
            q[(size_t)f * s.nlm + l * (l + 1) / 2 + m] = r;
        }
    }
}

void vh_sph_synthesis_F(hipStream_t st, const vh_sht& s, const double2* q, int nb, double2* FT, double2* FP) {
    size_t lds = (size_t)nb * (s.L + 1) * sizeof(double2);
    k_sph_synthesis<<<s.L + 1, 256, lds, st>>>(s, q, nb, FT, FP);
    HIPCHECK(hipGetLastError());
}

void vh_sph_analysis_F(hipStream_t st, const vh_sht& s, const double2* FT, const double2* FP, int nb, double2* q) {
    size_t lds = (size_t)4 * nb * s.nh * sizeof(double2);
    k_sph_analysis<<<s.L + 1, 256, lds, st>>>(s, FT, FP, nb, q);
    HIPCHECK(hipGetLastError());
}

// ---------------------------------------------------------------- FFT
void vh_fft_forward(hipStream_t st, const vh_sht& s, const double* f, int nrows, double2* F) {
    vh_fft_forward_t(st, s, InGrid{f, s.nphi}, nrows, F);
}

void vh_fft_inverse(hipStream_t st, const vh_sht& s, const double2* F, int nrows, double* f) {
    vh_fft_inverse_pair_t(st, s, F, nrows, OutGrid{f, s.nphi});
}

// ----------------------------------------------------------- reductions
__global__ void __launch_bounds__(256) k_rowsum(const double* __restrict__ f, int nphi, double* part) {
    __shared__ double red[256];
    const double* row = f + (size_t)blockIdx.x * nphi;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) acc += row[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// One block per component: the terms are staged in LDS by the whole block,
// then summed in latitude order by one lane (the order fixes the rounding).
__global__ void __launch_bounds__(256) k_weighted_total(const double* __restrict__ part,
        const double* __restrict__ w, int nlat, double dlon, double* out) {
    extern __shared__ double sv[];   // [2][nlat]
    const int c = blockIdx.x;
    for (int j = threadIdx.x; j < nlat; j += blockDim.x) {
        sv[j] = w[j];
        sv[nlat + j] = part[(size_t)c * nlat + j];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double total = 0.0;
    for (int j = 0; j < nlat; j++) total += sv[j] * sv[nlat + j] * dlon;
    out[c] = total;
}

void vh_surface_integral(hipStream_t st, const vh_sht& s, const double* f, double* part, double* out) {
    k_rowsum<<<s.nlat, 256, 0, st>>>(f, s.nphi, part);
    k_weighted_total<<<1, 256, 2 * s.nlat * sizeof(double), st>>>(part, s.w, s.nlat, 2.0 * VH_PI / s.nphi, out);
    HIPCHECK(hipGetLastError());
}

void vh_weighted_total(hipStream_t st, const vh_sht& s, const double* part, int ncomp, double* out) {
    if (ncomp <= 0) return;
    k_weighted_total<<<ncomp, 256, 2 * s.nlat * sizeof(double), st>>>(part, s.w, s.nlat, 2.0 * VH_PI / s.nphi, out);
    HIPCHECK(hipGetLastError());
}

__global__ void k_fill(double* p, double v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void vh_fill(hipStream_t st, double* p, double v, size_t n) {
    if (n == 0) return;
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, v, (int)n);
    HIPCHECK(hipGetLastError());
}
