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
// (r0, r0 + 1) at longitude i; with kRowSum the values they return are summed
// over the row (wavefront reduction) and handed to row_total(r0, sum)
struct OutGrid {
    static constexpr bool kRowSum = false;
    double* f;
    int nphi;
    __device__ double operator()(int r0, int i, double v0, double v1, bool has1) const {
        f[(size_t)r0 * nphi + i] = v0;
        if (has1) f[(size_t)(r0 + 1) * nphi + i] = v1;
        return 0.0;
    }
    __device__ void row_total(int, double) const {}
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// grid points per lane whose loads are issued together in the row kernels
#define VH_FFT_U 8

__device__ __forceinline__ double2 cconj_if(double2 w, bool inv) { return inv ? make_double2(w.x, -w.y) : w; }
// -i z (forward) / +i z (inverse)
__device__ __forceinline__ double2 mul_mi(double2 z, bool inv) {
    return inv ? make_double2(-z.y, z.x) : make_double2(z.y, -z.x);
}

// Size-R DFT of v in registers (compile-time R: nothing spills), sign -+ for
// forward / inverse.
template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], double2 (&V)[R], bool inv) {
    if constexpr (R == 2) {
        V[0] = cadd(v[0], v[1]); V[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
        const double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const double2 s13 = cadd(v[1], v[3]), d13 = mul_mi(csub(v[1], v[3]), inv);
        V[0] = cadd(s02, s13); V[2] = csub(s02, s13);
        V[1] = cadd(d02, d13); V[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        const double c = -0.5, sn = 0.86602540378443864676;   // cos, sin of 2 pi/3
        const double2 t1 = cadd(v[1], v[2]);
        const double2 t2 = make_double2(v[0].x + c * t1.x, v[0].y + c * t1.y);
        const double2 d = csub(v[1], v[2]);
        const double2 t3 = mul_mi(make_double2(sn * d.x, sn * d.y), inv);
        V[0] = cadd(v[0], t1); V[1] = cadd(t2, t3); V[2] = csub(t2, t3);
    } else if constexpr (R == 6) {   // 2 x 3: even/odd halves, W6^k on the odd one
        double2 e[3] = {v[0], v[2], v[4]}, o[3] = {v[1], v[3], v[5]}, E[3], O[3];
        dft_small<3>(e, E, inv);
        dft_small<3>(o, O, inv);
        const double h = 0.5, sn = 0.86602540378443864676;
        const double2 w1 = cconj_if(make_double2(h, -sn), inv), w2 = cconj_if(make_double2(-h, -sn), inv);
        const double2 O1 = cmul(O[1], w1), O2 = cmul(O[2], w2);
        V[0] = cadd(E[0], O[0]); V[3] = csub(E[0], O[0]);
        V[1] = cadd(E[1], O1); V[4] = csub(E[1], O1);
        V[2] = cadd(E[2], O2); V[5] = csub(E[2], O2);
    } else if constexpr (R == 8) {   // 2 x 4
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]}, E[4], O[4];
        dft_small<4>(e, E, inv);
        dft_small<4>(o, O, inv);
        const double r2 = 0.70710678118654752440;
        const double2 w1 = cconj_if(make_double2(r2, -r2), inv), w3 = cconj_if(make_double2(-r2, -r2), inv);
        const double2 O1 = cmul(O[1], w1), O2 = mul_mi(O[2], inv), O3 = cmul(O[3], w3);
        V[0] = cadd(E[0], O[0]); V[4] = csub(E[0], O[0]);
        V[1] = cadd(E[1], O1); V[5] = csub(E[1], O1);
        V[2] = cadd(E[2], O2); V[6] = csub(E[2], O2);
        V[3] = cadd(E[3], O3); V[7] = csub(E[3], O3);
    } else {   // 5, 7: direct sum with exact-index twiddles
#pragma unroll
        for (int u = 0; u < R; u++) {
            double2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; q++) {
                const int e = (q * u) % R;
                if (e == 0) { acc = cadd(acc, v[q]); continue; }
                double sv, cv;
                sincospi(2.0 * e / R, &sv, &cv);
                acc = cadd(acc, cmul(v[q], cconj_if(make_double2(cv, -sv), inv)));
            }
            V[u] = acc;
        }
    }
}

// LDS exchange between the passes of one wavefront's FFT: LDS operations of
// a wavefront complete in order, so only compiler reordering must be fenced.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One Stockham radix-R pass over N points. tp: this pass's twiddles
// [r-1][k] = e^{-2 pi i k r tstep/N}, k = j mod Ns < Ns (none when Ns = 1: the
// first pass multiplies by 1).
// swz(i): i with its low 3 bits xor-ed by bits 3..5, a permutation inside each
// aligned group of 8 elements. The first pass stores with stride R across
// lanes (all in one LDS bank group); swizzling that buffer spreads the lanes of
// a group over 8 distinct 16-byte slots, and the second pass reads it back the
// same way.
__device__ __forceinline__ int swz(int i) { return i ^ ((i >> 3) & 7); }

// SwS / SwD are run-time flags (one instantiation per radix keeps the
// instruction footprint of a kernel with several transforms small).
template <int R>
__device__ __forceinline__ void fft_pass(const double2* src, double2* dst, int N, int Ns, unsigned nsm,
                                         const double2* tp, bool inv, int t, int nthr, bool SwS = false,
                                         bool SwD = false) {
    const int NR = N / R;
    for (int j = t; j < NR; j += nthr) {
        // j = q Ns + k by the pass's reciprocal (no integer division)
        const int q = Ns > 1 ? (int)__umulhi((unsigned)j, nsm) : j;
        const int k = j - q * Ns;
        double2 v[R], V[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[SwS ? swz(j + r * NR) : j + r * NR];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], cconj_if(tp[(r - 1) * Ns + k], inv));
        }
        dft_small<R>(v, V, inv);
        const int idxD = q * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[SwD ? swz(idxD + r * Ns) : idxD + r * Ns] = V[r];
    }
}

// One complex FFT of length s.nphi by the calling wavefront alone (lane
// 0..63): a, b are this wavefront's LDS buffers of N points; returns the one
// holding the result. Natural order in and out (Stockham autosort). The
// caller makes a's contents visible to the whole wavefront first. tw: a copy
// of the s.ntw per-pass twiddles (e.g. in LDS), or null for s.twp.
__device__ __forceinline__ double2* fft_wave(double2* a, double2* b, const vh_sht& s, bool inv, int lane,
                                             const double2* tw = nullptr) {
    const int N = s.nphi;
    int Ns = 1;
    double2 *src = a, *dst = b;
    // the buffer between the first and second pass is swizzled (swz)
    const bool sw = s.nfac > 1 && (N & 7) == 0;
    for (int p = 0; p < s.nfac; p++) {
        const int R = s.fac[p];
        const double2* tp = (tw ? tw : s.twp) + s.twoff[p];
        const int mode = !sw ? 0 : (p == 0 ? 1 : (p == 1 ? 2 : 0));   // 1: store swizzled, 2: load swizzled
        const unsigned M = s.nsm[p];
        const bool sws = mode == 2, swd = mode == 1;
        switch (R) {
            case 8: fft_pass<8>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
            case 6: fft_pass<6>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
            case 4: fft_pass<4>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
            case 2: fft_pass<2>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
            case 3: fft_pass<3>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
            case 5: fft_pass<5>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
            default: fft_pass<7>(src, dst, N, Ns, M, tp, inv, lane, 64, sws, swd); break;
        }
        wave_lds_sync();
        double2* t = src; src = dst; dst = t;
        Ns *= R;
    }
    return src;
}

// ---- the same transform with the length fixed at compile time
// For the grid widths of the usual truncations the factorisation, strides and
// twiddle offsets become constants: index arithmetic folds into the LDS
// instructions' immediate offsets and shifts, and the direction into the
// butterflies' signs. Factors (fft_make_cplan, which the run-time plan is
// checked against), twiddle table and the order of every floating-point
// operation are those of fft_wave, so both give the same bits.
template <int N, int P, bool INV>
__device__ __forceinline__ void fft_pass_n(const double2* __restrict__ src, double2* __restrict__ dst,
                                           const double2* __restrict__ tw, int lane) {
    constexpr fft_cplan c = fft_make_cplan(N);
    constexpr int R = c.fac[P], Ns = c.ns[P], NR = N / R;
    constexpr bool sw = c.nfac > 1 && (N & 7) == 0, SwS = sw && P == 1, SwD = sw && P == 0;
    const double2* tp = tw + c.twoff[P];
#pragma unroll 1
    for (int j0 = 0; j0 < NR; j0 += 64) {
        const int j = j0 + lane;
        if (NR % 64 == 0 || j < NR) {
            const int q = j / Ns, k = j - q * Ns;
            double2 v[R], V[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = src[SwS ? swz(j + r * NR) : j + r * NR];
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[r] = cmul(v[r], cconj_if(tp[(r - 1) * Ns + k], INV));
            }
            dft_small<R>(v, V, INV);
            const int idxD = q * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) dst[SwD ? swz(idxD + r * Ns) : idxD + r * Ns] = V[r];
        }
    }
}

template <int N, int P, bool INV>
__device__ __forceinline__ double2* fft_passes_n(double2* src, double2* dst, const double2* tw, int lane) {
    constexpr fft_cplan c = fft_make_cplan(N);
    if constexpr (P >= c.nfac) {
        return src;
    } else {
        fft_pass_n<N, P, INV>(src, dst, tw, lane);
        wave_lds_sync();
        return fft_passes_n<N, P + 1, INV>(dst, src, tw, lane);
    }
}

// In-place form of a compile-time pass: every lane reads the inputs of its
// butterflies (NR / 64 rounded up) into registers, the wavefront's LDS reads
// are ordered before its writes, then the outputs go back into the same
// buffer. Same arithmetic as fft_pass_n; half the LDS.
template <int N, int P, bool INV>
__device__ __forceinline__ void fft_pass_ip(double2* buf, const double2* __restrict__ tw, int lane) {
    constexpr fft_cplan c = fft_make_cplan(N);
    constexpr int R = c.fac[P], Ns = c.ns[P], NR = N / R, NB = (NR + 63) / 64;
    constexpr bool sw = c.nfac > 1 && (N & 7) == 0, SwS = sw && P == 1, SwD = sw && P == 0;
    const double2* tp = tw + c.twoff[P];
    double2 v[NB][R];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = lane + 64 * b;
        if (NR % 64 == 0 || j < NR) {
#pragma unroll
            for (int r = 0; r < R; r++) v[b][r] = buf[SwS ? swz(j + r * NR) : j + r * NR];
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = lane + 64 * b;
        if (NR % 64 == 0 || j < NR) {
            const int q = j / Ns, k = j - q * Ns;
            double2 V[R];
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[b][r] = cmul(v[b][r], cconj_if(tp[(r - 1) * Ns + k], INV));
            }
            dft_small<R>(v[b], V, INV);
            const int idxD = q * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[SwD ? swz(idxD + r * Ns) : idxD + r * Ns] = V[r];
        }
    }
}

template <int N, int P, bool INV>
__device__ __forceinline__ void fft_passes_ip(double2* buf, const double2* tw, int lane) {
    constexpr fft_cplan c = fft_make_cplan(N);
    if constexpr (P < c.nfac) {
        fft_pass_ip<N, P, INV>(buf, tw, lane);
        wave_lds_sync();
        fft_passes_ip<N, P + 1, INV>(buf, tw, lane);
    }
}

// The in-place transform of length N with its last pass taken apart: passes
// 0 .. nfac-2 as fft_passes_ip, then the last one (64 butterflies, one per
// lane, stride 64) with its twiddles held in registers (tl: the lane's
// R - 1 values, loaded once) and epi(r, i, v) applied to the lane's output r
// (element i = lane + 64 r) before it is stored. The operations of
// fft_passes_ip in the same order.
template <int N, int P, bool INV>
__device__ __forceinline__ void fft_passes_ip_head(double2* buf, const double2* tw, int lane) {
    constexpr fft_cplan c = fft_make_cplan(N);
    if constexpr (P < c.nfac - 1) {
        fft_pass_ip<N, P, INV>(buf, tw, lane);
        wave_lds_sync();
        fft_passes_ip_head<N, P + 1, INV>(buf, tw, lane);
    }
}
template <int N>
struct fft_last {
    static constexpr fft_cplan c = fft_make_cplan(N);
    static constexpr int P = c.nfac - 1, R = c.fac[P], Ns = c.ns[P];
    static_assert(c.nfac >= 3 && N / R == 64 && Ns == 64, "last pass: one butterfly per lane, no swizzle");
};
template <int N>
__device__ __forceinline__ void fft_last_twiddles(const double2* tw, int lane, double2 (&tl)[fft_last<N>::R - 1]) {
    constexpr int R = fft_last<N>::R;
    const double2* tp = tw + fft_last<N>::c.twoff[fft_last<N>::P];
#pragma unroll
    for (int r = 1; r < R; r++) tl[r - 1] = tp[(r - 1) * 64 + lane];
}
template <int N, bool INV, class Epi>
__device__ __forceinline__ void fft_ip_reg_last(double2* buf, const double2* tw,
                                                const double2 (&tl)[fft_last<N>::R - 1], int lane, Epi epi) {
    constexpr int R = fft_last<N>::R;
    fft_passes_ip_head<N, 0, INV>(buf, tw, lane);
    double2 v[R], V[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = buf[lane + r * 64];
    wave_lds_sync();
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], cconj_if(tl[r - 1], INV));
    dft_small<R>(v, V, INV);
#pragma unroll
    for (int r = 0; r < R; r++) buf[lane + r * 64] = epi(r, lane + r * 64, V[r]);
}

// grid widths with a compile-time transform: the default 4 lmax of lmax 64
// and 192, and 384 (lmax 96 by default, lmax 128 on its 192 x 384 grid)
#define VH_FFT_FIXED_SIZES(X) X(256) X(384) X(768)

// fft_wave with the direction as a template argument: the fixed-size
// transform where the grid width has one, else the run-time one.
template <bool INV>
__device__ __forceinline__ double2* fft_wave_t(double2* a, double2* b, const vh_sht& s, int lane,
                                               const double2* tw = nullptr) {
    const double2* t = tw ? tw : s.twp;
    switch (s.nphi) {
#define VH_FFT_FIXED_CASE(NN) \
        case NN: return fft_passes_n<NN, 0, INV>(a, b, t, lane);
        VH_FFT_FIXED_SIZES(VH_FFT_FIXED_CASE)
#undef VH_FFT_FIXED_CASE
        default: return fft_wave(a, b, s, INV, lane, tw);
    }
}

// True when nphi has a compile-time transform (fft_wave_ip then works in place).
__host__ __device__ inline bool fft_fixed_size(int n) {
#define VH_FFT_FIXED_TEST(NN) if (n == NN) return true;
    VH_FFT_FIXED_SIZES(VH_FFT_FIXED_TEST)
#undef VH_FFT_FIXED_TEST
    return false;
}

// fft_wave_t for callers that budget LDS by fft_fixed_size: in place in a
// (returned) for the compile-time lengths, else the run-time transform with
// b as its second buffer.
template <bool INV>
__device__ __forceinline__ double2* fft_wave_ip(double2* a, double2* b, const vh_sht& s, int lane,
                                                const double2* tw = nullptr) {
    const double2* t = tw ? tw : s.twp;
    switch (s.nphi) {
#define VH_FFT_IP_CASE(NN) \
        case NN: fft_passes_ip<NN, 0, INV>(a, t, lane); return a;
        VH_FFT_FIXED_SIZES(VH_FFT_IP_CASE)
#undef VH_FFT_IP_CASE
        default: return fft_wave(a, b, s, INV, lane, tw);
    }
}

// Loads and stores of data that other blocks of a running cooperative launch
// read or wrote: relaxed atomics at agent scope, which the compiler issues
// coherent across the XCDs' L2 caches.
__device__ __forceinline__ double ld_coh(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double2 ld_coh(const double2* p) {
    return make_double2(ld_coh((const double*)p), ld_coh((const double*)p + 1));
}
__device__ __forceinline__ void st_coh(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(double2* p, double2 v) {
    st_coh((double*)p, v.x);
    st_coh((double*)p + 1, v.y);
}
// 16 bytes per lane from global memory straight into LDS: lane k's bytes land
// at dst + k, dst the same for the whole wavefront. No registers hold them;
// they are in LDS once the wavefront's vector memory counter has drained.
// Coh: coherent at agent scope, as ld_coh (data other running blocks wrote).
template <bool Coh = false>
__device__ __forceinline__ void lds_dma16(const double2* src, double2* dst) {
    __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)dst, 16, 0,
                                     Coh ? 16 /* sc1 */ : 0);
}

template <bool Coh>
__device__ __forceinline__ double2 ld_g(const double2* p) {
    if constexpr (Coh) return ld_coh(p); else return *p;
}
template <bool Coh>
__device__ __forceinline__ void st_g(double2* p, double2 v) {
    if constexpr (Coh) st_coh(p, v); else *p = v;
}

// a <- the packed complex row Z = A + iB of two real rows given by their half
// spectra A = F0[m], B = F1[m] (m = 0..L; imaginary parts at m = 0 dropped):
// Z_m = A_m + i B_m, Z_{N-m} = conj(A_m) + i conj(B_m). A null F0 / F1 reads
// zeros. Called by the whole wavefront; a is complete and visible on return.
// Coh: F0/F1 are read with coherent loads (written by other running blocks).
// mstride: distance between consecutive orders in F0/F1.
template <bool Coh = false>
__device__ __forceinline__ void pair_spectra_to_lds(const double2* F0, const double2* F1, int L, int N, double2* a,
                                                    int lane, int mstride = 1) {
    for (int i = lane; i < N; i += 64) a[i] = make_double2(0.0, 0.0);
    wave_lds_sync();
    if (F0) {
        for (int m0 = 0; m0 <= L; m0 += 64 * VH_FFT_U) {
            double2 A[VH_FFT_U], B[VH_FFT_U];
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int m = m0 + lane + 64 * u;
                A[u] = m <= L ? ld_g<Coh>(F0 + (size_t)m * mstride) : make_double2(0.0, 0.0);
                B[u] = (F1 && m <= L) ? ld_g<Coh>(F1 + (size_t)m * mstride) : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int m = m0 + lane + 64 * u;
                if (m > L) continue;
                if (m == 0) { A[u].y = 0.0; B[u].y = 0.0; }
                a[m] = make_double2(A[u].x - B[u].y, A[u].y + B[u].x);
                if (m > 0) a[N - m] = make_double2(A[u].x + B[u].y, -A[u].y + B[u].x);
            }
        }
    }
    wave_lds_sync();
}

// The FFT z of a packed pair of real rows split into their half spectra
// (m = 0..L), scaled: F0[m] = scale (Z_m + conj Z_{N-m})/2 and
// F1[m] = scale (Z_m - conj Z_{N-m})/(2i); F1 may be null. Coh: coherent stores.
template <bool Coh = false>
__device__ __forceinline__ void lds_to_pair_spectra(const double2* z, int L, int N, double scale, double2* F0,
                                                    double2* F1, int lane) {
    for (int m = lane; m <= L; m += 64) {
        double2 zm = z[m], zc = z[(N - m) % N];
        zc.y = -zc.y;
        const double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        const double2 D = csub(zm, zc);
        const double2 B = make_double2(0.5 * D.y, -0.5 * D.x);
        st_g<Coh>(F0 + m, make_double2(scale * A.x, scale * A.y));
        if (F1) st_g<Coh>(F1 + m, make_double2(scale * B.x, scale * B.y));
    }
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
    // all of a lane's grid loads in flight before the first LDS store
    for (int i0 = 0; i0 < N; i0 += 64 * VH_FFT_U) {
        double v0[VH_FFT_U], v1[VH_FFT_U];
#pragma unroll
        for (int u = 0; u < VH_FFT_U; u++) {
            const int i = i0 + lane + 64 * u;
            v0[u] = (has0 && i < N) ? in(r0, i) : 0.0;
            v1[u] = (has1 && i < N) ? in(r1, i) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < VH_FFT_U; u++) {
            const int i = i0 + lane + 64 * u;
            if (i < N) a[i] = make_double2(v0[u], v1[u]);
        }
    }
    wave_lds_sync();
    double2* z = fft_wave_t<false>(a, b, s, lane);
    if (has0) lds_to_pair_spectra(z, L, N, scale, F + (size_t)r0 * (L + 1), has1 ? F + (size_t)r1 * (L + 1) : nullptr, lane);
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
    pair_spectra_to_lds(has0 ? F + (size_t)r0 * (L + 1) : nullptr, has1 ? F + (size_t)r1 * (L + 1) : nullptr, L, N,
                        a, lane);
    double2* z = fft_wave_t<true>(a, b, s, lane);
    double rs = 0.0;   // consumers may return a per-row term to be summed
    if (has0)
        for (int i0 = 0; i0 < N; i0 += 64 * VH_FFT_U) {
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int i = i0 + lane + 64 * u;
                if (i < N) rs += out(r0, i, z[i].x, z[i].y, has1);
            }
        }
    if constexpr (Out::kRowSum) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) rs += __shfl_xor(rs, off);
        if (lane == 0 && has0) out.row_total(r0, rs);
    }
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

// Legendre synthesis: block = order m; thread = (north latitude j, quarter
// of the degrees), every field per table load; quarter sums (even / odd l+m)
// added in a fixed order. nb <= 4 fields; src(f, l, m, idx) supplies the
// coefficient of field f at degree-major index idx (a spectrum, or one
// computed on the fly such as the response applied to a load).
struct SpecSrc {
    const double2* q;
    int nlm;
    __device__ double2 operator()(int f, int l, int m, int idx) const { (void)l; (void)m; return q[(size_t)f * nlm + idx]; }
};

template <class Src>
__global__ void __launch_bounds__(512) k_leg_synthesis(vh_sht s, Src src, int nb, double2* __restrict__ F, int fst,
                                                       int jst) {
    extern __shared__ double2 sh[];
    const int m = blockIdx.x, L = s.L, nh = s.nh, nl = L - m + 1;
    double2* Q = sh;                        // [nb][nl]
    double2* red = sh + nb * (L + 1);       // [4][nh][nb][2]
    for (int t = threadIdx.x; t < nb * nl; t += blockDim.x) {
        int f = t / nl, il = t % nl, l = m + il;
        Q[f * nl + il] = src(f, l, m, l * (l + 1) / 2 + m);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 4 * nh; t += blockDim.x) {
        const int g = t / nh, j = t % nh;
        const int lq = ((nl + 7) / 8) * 2;      // even count per quarter keeps parity aligned
        const int la = min(nl, g * lq), lb = min(nl, la + lq);
        const double* P = s.plm_lj + (size_t)s.mstart[m] * nh;
        double2 e[4], o[4];
#pragma unroll
        for (int f = 0; f < 4; f++) e[f] = o[f] = make_double2(0.0, 0.0);
        for (int il = la; il < lb; il += 2) {
            const double p = P[(size_t)il * nh + j];
            const bool two = il + 1 < lb;
            const double p1 = two ? P[(size_t)(il + 1) * nh + j] : 0.0;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                if (f >= nb) break;
                const double2 c0 = Q[f * nl + il];
                e[f].x = fma(c0.x, p, e[f].x); e[f].y = fma(c0.y, p, e[f].y);
                if (two) {
                    const double2 c1 = Q[f * nl + il + 1];
                    o[f].x = fma(c1.x, p1, o[f].x); o[f].y = fma(c1.y, p1, o[f].y);
                }
            }
        }
#pragma unroll
        for (int f = 0; f < 4; f++) {
            if (f >= nb) break;
            double2* r = red + (((size_t)g * nh + j) * nb + f) * 2;
            r[0] = e[f];
            r[1] = o[f];
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nh * nb; t += blockDim.x) {
        const int jj = t / nb, f = t % nb, js = s.nlat - 1 - jj;
        double2 ev = make_double2(0.0, 0.0), od = ev;
        for (int k = 0; k < 4; k++) {
            const double2* r = red + (((size_t)k * nh + jj) * nb + f) * 2;
            ev = make_double2(ev.x + r[0].x, ev.y + r[0].y);
            od = make_double2(od.x + r[1].x, od.y + r[1].y);
        }
        double2* Ff = F + (size_t)f * fst * (L + 1);
        Ff[(size_t)jj * jst * (L + 1) + m] = make_double2(ev.x + od.x, ev.y + od.y);
        if (js != jj) Ff[(size_t)js * jst * (L + 1) + m] = make_double2(ev.x - od.x, ev.y - od.y);
    }
}

template <class Src>
void vh_legendre_synthesis_t(hipStream_t st, const vh_sht& s, Src src, int nb, double2* F, int fst, int jst) {
    if (nb > 4) throw vh_error("legendre synthesis: at most 4 fields per call");
    const int nthr = std::min(512, ((4 * s.nh + 63) / 64) * 64);   // (quarter, latitude) tasks, looped
    const size_t lds = ((size_t)nb * (s.L + 1) + (size_t)4 * s.nh * nb * 2) * sizeof(double2);
    k_leg_synthesis<Src><<<s.L + 1, nthr, lds, st>>>(s, src, nb, F, fst, jst);
    HIPCHECK(hipGetLastError());
}

