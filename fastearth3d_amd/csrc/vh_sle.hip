// vh_sle.hip -- (d) grid-space kernels of the sea-level equation
// (src/vilma_sle.f90:124-405) and the ocean function (:407-425).
//
// The ocean/coastline mask must be bit-exact, so every expression that
// feeds it (interpolated ice, ice - d_ice, topo0 - rsl) is evaluated with
// floating-point contraction off, in the reference's operation order.
// Reductions are deterministic two-pass (per-row partials, then one ordered
// weighted sum) so a rerun gives identical bits.
#pragma clang fp contract(off)
#include "vh_internal.h"
#include "vh_sht_dev.h"

__device__ __forceinline__ double ocean(double topo, double ice) {
    return (topo < 0.0 && VH_RHO_ICE * ice < -VH_RHO_WATER * topo) ? 1.0 : 0.0;
}

// ice(t) = ice0 + frac*(ice1 - ice0), d_ice = ice - ice_ref, C0 = O(topo0, ice - d_ice)
// (src/vilma_timestep.f90:314-326, src/vilma_sle.f90:216)
__global__ void k_sle_ice(int n, const double* __restrict__ ice0, const double* __restrict__ ice1, double frac,
                          int interp, const double* __restrict__ ice_ref, const double* __restrict__ topo0,
                          double* ice, double* dice, double* C0) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = interp ? ice0[i] + frac * (ice1[i] - ice0[i]) : ice0[i];
    double d = v - ice_ref[i];
    ice[i] = v;
    dice[i] = d;
    C0[i] = ocean(topo0[i], v - d);
}

// Coastline pass head (src/vilma_sle.f90:231-294): C, grounded-ice increment,
// and per-row partial sums of C, d_ice_g, topo0*(C - C0).
__global__ void __launch_bounds__(256) k_sle_outer(int nphi, int first, const double* __restrict__ topo0,
        const double* __restrict__ rsl, const double* __restrict__ ice, const double* __restrict__ dice,
        const double* __restrict__ C0, const double* __restrict__ Cnext, double* C, double* dig,
        double* part /* [3][nlat] */, int nlat) {
    __shared__ double red[3][256];
    const int j = blockIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        double c = first ? ocean(topo0[id] - rsl[id], ice[id]) : Cnext[id];
        double g = ice[id] * (1.0 - c) - (ice[id] - dice[id]) * (1.0 - C0[id]);
        C[id] = c;
        dig[id] = g;
        s0 += c; s1 += g; s2 += topo0[id] * (c - C0[id]);
    }
    red[0][threadIdx.x] = s0; red[1][threadIdx.x] = s1; red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[(size_t)threadIdx.x * nlat + j] = red[threadIdx.x][0];
}

// subgrid sloping-coast load (src/vilma_sle.f90:289-291)
__global__ void k_sle_wcorr(int n, int subgrid, const double* __restrict__ topo0, const double* __restrict__ C,
                            const double* __restrict__ C0, double* wcorr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    wcorr[i] = subgrid ? -VH_RHO_WATER * topo0[i] * (C[i] - C0[i]) : 0.0;
}

// surface load (assemble_load, src/vilma_sle.f90:387-403), fed straight into
// the forward longitude transform
struct InLoad {
    const double *dig, *C, *rsl, *wcorr;
    int nphi;
    __device__ double operator()(int row, int i) const {
        size_t id = (size_t)row * nphi + i;
        return VH_RHO_ICE * dig[id] + VH_RHO_WATER * (C[id] * rsl[id]) + wcorr[id];
    }
};

// response_apply (src/vilma_response.f90:842-864) into two spectra [u; N]
// (degree-major): u = gu*sigma + dUa, N = gn*sigma - dFa/g.
__global__ void k_apply(int nlm, const double2* __restrict__ sigma, const double* __restrict__ gu,
                        const double* __restrict__ gn, const double* __restrict__ dsurf, int nk, double g,
                        int kind, double2* uN) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlm) return;
    int l = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
    while ((l + 1) * (l + 2) / 2 <= idx) l++;
    while (l * (l + 1) / 2 > idx) l--;
    double2 s = sigma[idx], u = make_double2(0.0, 0.0), N = make_double2(0.0, 0.0);
    if (kind != 0) {   // 0 = null response
        u = make_double2(gu[l] * s.x, gu[l] * s.y);
        N = make_double2(gn[l] * s.x, gn[l] * s.y);
        if (kind == 2 && l >= 1) {
            int k = idx - 1;
            const double2* ds = (const double2*)dsurf;
            double2 dU = ds[k], dF = ds[nk + k];
            u = make_double2(u.x + dU.x, u.y + dU.y);
            N = make_double2(N.x - dF.x / g, N.y - dF.y / g);
        }
    }
    uN[idx] = u;
    uN[nlm + idx] = N;
}

// the response applied to the load as the synthesis reads it: field 0 is
// u = gu*sigma + dUa, field 1 is N = gn*sigma - dFa/g (response_apply,
// src/vilma_response.f90:842-864)
struct ApplySrc {
    const double2* sigma;
    const double* gu;
    const double* gn;
    const double* dsurf;
    int nk, kind;
    double g;
    __device__ double2 operator()(int f, int l, int m, int idx) const {
        (void)m;
        if (kind == 0) return make_double2(0.0, 0.0);   // null response
        const double2 s = sigma[idx];
        if (f == 0) {
            double2 u = make_double2(gu[l] * s.x, gu[l] * s.y);
            if (kind == 2 && l >= 1) {
                const double2 dU = ((const double2*)dsurf)[idx - 1];
                u = make_double2(u.x + dU.x, u.y + dU.y);
            }
            return u;
        }
        double2 N = make_double2(gn[l] * s.x, gn[l] * s.y);
        if (kind == 2 && l >= 1) {
            const double2 dF = ((const double2*)dsurf)[nk + idx - 1];
            N = make_double2(N.x - dF.x / g, N.y - dF.y / g);
        }
        return N;
    }
};

// inverse transform epilogue: rows arrive as interleaved pairs (u_j, N_j);
// Sraw = N - u (+ s_rot); the row's sum of C * Sraw goes to cs[j]
struct OutSraw {
    static constexpr bool kRowSum = true;
    double* sraw;
    const double* srot;
    const double* C;
    double* cs;
    int nphi;
    __device__ double operator()(int r0, int i, double v0, double v1, bool has1) const {
        (void)has1;
        int j = r0 >> 1;
        size_t id = (size_t)j * nphi + i;
        double s = v1 - v0;
        if (srot) s = s + srot[id];
        sraw[id] = s;
        return C[id] * s;
    }
    __device__ void row_total(int r0, double v) const { cs[r0 >> 1] = v; }
};

__global__ void __launch_bounds__(256) k_rowsum_prod(int nphi, const double* __restrict__ a,
                                                     const double* __restrict__ b, double* part) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        acc += a[id] * b[id];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[j] = red[0];
}

// Cs_int = sum_j w_j dlon cs_j (latitude order, every block the same);
// dphi = (ice_int - Cs_int + zeta_int)/C_int ; rsl_new = Sraw + dphi ;
// per row: dmax = max|C (rsl_new - rsl)|, smax = max|C rsl_new| ; rsl = rsl_new.
// sc: [0] C_int [1] int d_ice_g [2] zeta_int; writes sc[3] = Cs_int, sc[4] = dphi.
// out: [0, nlat) row dmax, [nlat, 2 nlat) row smax, [2 nlat] dphi.
__global__ void __launch_bounds__(256) k_sle_update(int nphi, const double* __restrict__ sraw,
        const double* __restrict__ C, double* rsl, double* sc, int subgrid, const double* __restrict__ cs,
        const double* __restrict__ w, double dlon, double* out, int nlat) {
    __shared__ double r0[256], r1[256];
    __shared__ double sv[2][512];
    __shared__ double dphi_s;
    const int j = blockIdx.x;
    for (int t = threadIdx.x; t < nlat; t += blockDim.x) { sv[0][t] = w[t]; sv[1][t] = cs[t]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = 0.0;
        for (int q = 0; q < nlat; q++) total += sv[0][q] * sv[1][q] * dlon;
        const double rho_ratio = VH_RHO_ICE / VH_RHO_WATER;
        const double ice_int = -rho_ratio * sc[1];
        const double zeta = subgrid ? sc[2] : 0.0;
        dphi_s = (ice_int - total + zeta) / sc[0];
        if (j == 0) { sc[3] = total; sc[4] = dphi_s; out[2 * nlat] = dphi_s; }
    }
    __syncthreads();
    const double dphi = dphi_s;
    double dmax = 0.0, smax = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        double rn = sraw[id] + dphi;
        dmax = fmax(dmax, fabs(C[id] * (rn - rsl[id])));
        rsl[id] = rn;
        smax = fmax(smax, fabs(C[id] * rn));
    }
    r0[threadIdx.x] = dmax; r1[threadIdx.x] = smax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + s]);
            r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[j] = r0[0];
        out[nlat + j] = r1[0];
    }
}

// one block per component; max is exact in any order
__global__ void __launch_bounds__(256) k_max_final(const double* __restrict__ part, int nlat, double* out) {
    __shared__ double red[256];
    const int c = blockIdx.x;
    double m = 0.0;
    for (int j = threadIdx.x; j < nlat; j += blockDim.x) m = fmax(m, part[(size_t)c * nlat + j]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// C_next = O(topo0 - rsl, ice); per-row count of flipped cells
__global__ void __launch_bounds__(256) k_sle_flip(int nphi, const double* __restrict__ topo0,
        const double* __restrict__ rsl, const double* __restrict__ ice, const double* __restrict__ C,
        double* Cnext, double* part) {
    __shared__ double red[256];
    const int j = blockIdx.x;
    double cnt = 0.0;
    for (int i = threadIdx.x; i < nphi; i += blockDim.x) {
        size_t id = (size_t)j * nphi + i;
        double cn = ocean(topo0[id] - rsl[id], ice[id]);
        Cnext[id] = cn;
        cnt += (cn != C[id]) ? 1.0 : 0.0;
    }
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[j] = red[0];
}

// per-row counts are whole numbers: their sum is exact in any order
__global__ void __launch_bounds__(256) k_sum_final(const double* __restrict__ part, int nlat, double* out) {
    __shared__ double red[256];
    double a = 0.0;
    for (int j = threadIdx.x; j < nlat; j += blockDim.x) a += part[j];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// z_bed = z_bed_eq - rsl ; ocean function of an arbitrary (topo, ice) pair
__global__ void k_sub(int n, const double* __restrict__ a, const double* __restrict__ b, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] - b[i];
}
__global__ void k_ocean(int n, const double* __restrict__ topo, const double* __restrict__ ice, double* C) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) C[i] = ocean(topo[i], ice[i]);
}
// bsl integrand h_ice (1 - C) - h_ice_eq (1 - C0) (src/vilma_coupling.f90:572-597)
__global__ void k_bsl_integrand(int n, const double* __restrict__ h, const double* __restrict__ C,
                                const double* __restrict__ heq, const double* __restrict__ C0, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = h[i] * (1.0 - C[i]) - heq[i] * (1.0 - C0[i]);
}

// ------------------------------------------------------------ host wrappers
#include "vh_sle.h"

static inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

void sle_ice(hipStream_t st, int n, const double* ice0, const double* ice1, double frac, int interp,
             const double* ice_ref, const double* topo0, double* ice, double* dice, double* C0) {
    k_sle_ice<<<nb256(n), 256, 0, st>>>(n, ice0, ice1, frac, interp, ice_ref, topo0, ice, dice, C0);
}
void sle_outer(hipStream_t st, const vh_sht& s, int first, const double* topo0, const double* rsl,
               const double* ice, const double* dice, const double* C0, const double* Cnext, double* C,
               double* dig, double* part, double* sc) {
    k_sle_outer<<<s.nlat, 256, 0, st>>>(s.nphi, first, topo0, rsl, ice, dice, C0, Cnext, C, dig, part, s.nlat);
    vh_weighted_total(st, s, part, 3, sc);
}
void sle_wcorr(hipStream_t st, int n, int subgrid, const double* topo0, const double* C, const double* C0, double* w) {
    k_sle_wcorr<<<nb256(n), 256, 0, st>>>(n, subgrid, topo0, C, C0, w);
}
void sle_load_analysis(hipStream_t st, const vh_sht& s, const double* dig, const double* C, const double* rsl,
                       const double* wcorr, double2* F, double2* load_lm) {
    vh_fft_forward_t(st, s, InLoad{dig, C, rsl, wcorr, s.nphi}, s.nlat, F);
    vh_legendre_analysis(st, s, F, 1, load_lm);
}
void sle_apply_synth(hipStream_t st, const vh_sht& s, const double2* sigma, const double* gu, const double* gn,
                     const double* dsurf, int nk, double g, int kind, double2* F2, const double* srot,
                     const double* C, double* sraw, double* cs) {
    vh_legendre_synthesis_t(st, s, ApplySrc{sigma, gu, gn, dsurf, nk, kind, g}, 2, F2, 1, 2);
    vh_fft_inverse_pair_t(st, s, F2, 2 * s.nlat, OutSraw{sraw, srot, C, cs, s.nphi});
}
void vh_apply_only(hipStream_t st, int nlm, const double2* sigma, const double* gu, const double* gn,
                   const double* dsurf, int nk, double g, int kind, double2* uN) {
    k_apply<<<nb256(nlm), 256, 0, st>>>(nlm, sigma, gu, gn, dsurf, nk, g, kind, uN);
}
void sle_cs(hipStream_t st, const vh_sht& s, const double* C, const double* f, double* part, double* out) {
    k_rowsum_prod<<<s.nlat, 256, 0, st>>>(s.nphi, C, f, part);
    vh_weighted_total(st, s, part, 1, out);
}
void sle_update(hipStream_t st, const vh_sht& s, const double* sraw, const double* C, double* rsl, double* sc,
                int subgrid, const double* cs, double* out) {
    if (s.nlat > 512) throw vh_error("sle_update: nlat above 512 is not supported");
    k_sle_update<<<s.nlat, 256, 0, st>>>(s.nphi, sraw, C, rsl, sc, subgrid, cs, s.w, 2.0 * VH_PI / s.nphi, out,
                                         s.nlat);
}
void sle_flip(hipStream_t st, const vh_sht& s, const double* topo0, const double* rsl, const double* ice,
              const double* C, double* Cnext, double* part, double* out) {
    k_sle_flip<<<s.nlat, 256, 0, st>>>(s.nphi, topo0, rsl, ice, C, Cnext, part);
    k_sum_final<<<1, 256, 0, st>>>(part, s.nlat, out);
}
void grid_sub(hipStream_t st, int n, const double* a, const double* b, double* out) {
    k_sub<<<nb256(n), 256, 0, st>>>(n, a, b, out);
}
void grid_ocean(hipStream_t st, int n, const double* topo, const double* ice, double* C) {
    k_ocean<<<nb256(n), 256, 0, st>>>(n, topo, ice, C);
}
void grid_bsl_integrand(hipStream_t st, int n, const double* h, const double* C, const double* heq,
                        const double* C0, double* out) {
    k_bsl_integrand<<<nb256(n), 256, 0, st>>>(n, h, C, heq, C0, out);
}
