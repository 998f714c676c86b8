// vh_sht_dev.h -- longitude transforms: a Stockham mixed-radix complex FFT
// run by one wavefront in LDS, and the real-row wrappers. Templated on a
// producer (forward) / consumer (inverse) functor so grid-space work (the sea
// level load assembly, the rsl epilogue) fuses into the transform instead of
// making another pass over HBM.
#pragma once
#include "vh_internal.h"

struct InGrid {
    const double* f;
    int nphi;
    __device__ double operator()(int row, int i) const { return f[(size_t)row * nphi + i]; }
};
// consumers of the inverse transform receive the two real rows of a pair
// (r0, r0 + 1) at longitude i
struct OutGrid {
    double* f;
    int nphi;
    __device__ void operator()(int r0, int i, double v0, double v1, bool has1) const {
        f[(size_t)r0 * nphi + i] = v0;
        if (has1) f[(size_t)(r0 + 1) * nphi + i] = v1;
    }
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// twiddle e^{-+2 pi i t/N}, 0 <= t < N
__device__ __forceinline__ double2 twid(const double2* tw, int t, bool inv) {
    double2 w = tw[t];
    return inv ? make_double2(w.x, -w.y) : w;
}

// Size-R DFT of v (in registers; compile-time R so nothing spills).
// w[q] = e^{-+2 pi i q/R}, q = 1..R-1.
template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], double2 (&V)[R], const double2 (&w)[R], bool inv) {
    if constexpr (R == 2) {
        V[0] = cadd(v[0], v[1]); V[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        // forward: W4 = -i ; -i*z = (z.y, -z.x); inverse: +i*z = (-z.y, z.x)
        double2 jd13 = inv ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
        V[0] = cadd(s02, s13); V[2] = csub(s02, s13);
        V[1] = cadd(d02, jd13); V[3] = csub(d02, jd13);
    } else {
#pragma unroll
        for (int u = 0; u < R; u++) {
            double2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; q++) {
                const int e = (q * u) % R;
                acc = cadd(acc, e == 0 ? v[q] : cmul(v[q], w[e]));
            }
            V[u] = acc;
        }
    }
}

// One Stockham radix-R pass over N points by nthr threads (thread index t).
template <int R>
__device__ __forceinline__ void fft_pass(const double2* src, double2* dst, int N, int Ns, const double2* tw,
                                         bool inv, int t, int nthr) {
    const int NR = N / R, tstep = N / (Ns * R);
    double2 w[R];
#pragma unroll
    for (int q = 1; q < R; q++) w[q] = twid(tw, q * NR, inv);
    w[0] = make_double2(1.0, 0.0);
    for (int j = t; j < NR; j += nthr) {
        const int k = j % Ns;
        double2 v[R], V[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[j + r * NR];
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = cmul(v[r], twid(tw, k * r * tstep, inv));
        dft_small<R>(v, V, w, inv);
        const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[idxD + r * Ns] = V[r];
    }
}

// One complex FFT of length N by the calling wavefront (lane 0..63). a, b:
// LDS buffers of N; returns the buffer that holds the result. Natural order
// in and out (Stockham autosort). fac: radices (2, 3, 4, 5, 7), product N.
// Every wavefront of the block must call it with the same N (the passes are
// separated by block barriers).
__device__ double2* fft_wave(double2* a, double2* b, int N, const int* fac, int nfac, const double2* tw,
                             bool inv, int lane) {
    int Ns = 1;
    double2 *src = a, *dst = b;
    for (int s = 0; s < nfac; s++) {
        const int R = fac[s];
        switch (R) {
            case 2: fft_pass<2>(src, dst, N, Ns, tw, inv, lane, 64); break;
            case 3: fft_pass<3>(src, dst, N, Ns, tw, inv, lane, 64); break;
            case 4: fft_pass<4>(src, dst, N, Ns, tw, inv, lane, 64); break;
            case 5: fft_pass<5>(src, dst, N, Ns, tw, inv, lane, 64); break;
            default: fft_pass<7>(src, dst, N, Ns, tw, inv, lane, 64); break;
        }
        __syncthreads();
        double2* t = src; src = dst; dst = t;
        Ns *= R;
    }
    return src;
}

// Real rows -> F[row][m] (m = 0..L) scaled by `scale`; one row pair per
// wavefront, LDS 2 buffers x N per wavefront.
template <class In>
__global__ void __launch_bounds__(256) k_fft_fwd(vh_sht s, In in, int nrows, double2* __restrict__ F, double scale) {
    extern __shared__ double2 lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, N = s.nphi, L = s.L;
    double2* a = lds + (size_t)wave * 2 * N;
    double2* b = a + N;
    const int pair = blockIdx.x * (blockDim.x >> 6) + wave;
    const int r0 = 2 * pair, r1 = 2 * pair + 1;
    const bool has0 = r0 < nrows, has1 = r1 < nrows;
    for (int i = lane; i < N; i += 64)
        a[i] = make_double2(has0 ? in(r0, i) : 0.0, has1 ? in(r1, i) : 0.0);
    __syncthreads();
    double2* z = fft_wave(a, b, N, s.fac, s.nfac, s.tw, false, lane);
    if (has0) {
        for (int m = lane; m <= L; m += 64) {
            double2 zm = z[m], zc = z[(N - m) % N];
            zc.y = -zc.y;
            double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
            double2 D = csub(zm, zc);
            double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
            F[(size_t)r0 * (L + 1) + m] = make_double2(scale * A.x, scale * A.y);
            if (has1) F[(size_t)r1 * (L + 1) + m] = make_double2(scale * B.x, scale * B.y);
        }
    }
}

// F[row][m] -> real rows f = Re F0 + 2 Re sum_{m>0} F_m e^{i m phi}.
template <class Out>
__global__ void __launch_bounds__(256) k_fft_inv(vh_sht s, const double2* __restrict__ F, int nrows, Out out) {
    extern __shared__ double2 lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, N = s.nphi, L = s.L;
    double2* a = lds + (size_t)wave * 2 * N;
    double2* b = a + N;
    const int pair = blockIdx.x * (blockDim.x >> 6) + wave;
    const int r0 = 2 * pair, r1 = 2 * pair + 1;
    const bool has0 = r0 < nrows, has1 = r1 < nrows;
    for (int i = lane; i < N; i += 64) a[i] = make_double2(0.0, 0.0);
    __syncthreads();
    if (has0) {
        for (int m = lane; m <= L; m += 64) {
            double2 A = F[(size_t)r0 * (L + 1) + m];
            double2 B = has1 ? F[(size_t)r1 * (L + 1) + m] : make_double2(0.0, 0.0);
            if (m == 0) { A.y = 0.0; B.y = 0.0; }
            a[m] = make_double2(A.x - B.y, A.y + B.x);
            if (m > 0) a[N - m] = make_double2(A.x + B.y, -A.y + B.x);
        }
    }
    __syncthreads();
    double2* z = fft_wave(a, b, N, s.fac, s.nfac, s.tw, true, lane);
    if (has0)
        for (int i = lane; i < N; i += 64) out(r0, i, z[i].x, z[i].y, has1);
}

template <class In>
void vh_fft_forward_t(hipStream_t st, const vh_sht& s, In in, int nrows, double2* F) {
    int npairs = (nrows + 1) / 2;
    size_t lds = (size_t)2 * s.nphi * sizeof(double2);
    k_fft_fwd<In><<<npairs, 64, lds, st>>>(s, in, nrows, F, 2.0 * VH_PI / s.nphi);
    HIPCHECK(hipGetLastError());
}

template <class Out>
void vh_fft_inverse_pair_t(hipStream_t st, const vh_sht& s, const double2* F, int nrows, Out out) {
    int npairs = (nrows + 1) / 2;
    size_t lds = (size_t)2 * s.nphi * sizeof(double2);
    k_fft_inv<Out><<<npairs, 64, lds, st>>>(s, F, nrows, out);
    HIPCHECK(hipGetLastError());
}
