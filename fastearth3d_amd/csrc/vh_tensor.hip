// vh_tensor.hip -- (c) the lateral-viscosity memory update of the genuinely
// 3-D elements: for each such element e and radial shape s in {A, B, C}
//
//     c_s <- c_s + analysis( -Mk3(:,:,e) * synth(c_s + 2 mu_e eps_s) )
//
// (increment form, src/vilma_response.f90:1171-1225, :1227-1319), where the
// tensor transforms map the four spheroidal channels lambda = 1, 2, 5, 6 --
// and, when the toroidal degree of freedom is carried, lambda = 3, 4 as
// channels 4, 5 -- to the six dyads rr, rt, rp, tt, tp, pp
// (src/vilma_tensor_sh.f90:168-387).
//
// The reference composes each tensor transform from 3 + 3 (synthesis) and
// 3 + 5 (analysis) scalar/vector SHTs plus grid-space products with cot and
// 1/sin. All of those factors depend on latitude only and commute with the
// longitude FFT, so here a tensor transform is one Legendre pass in (m, j)
// space producing/consuming all six dyads at once, from the P and dP tables:
//
//   rr = sum c1 P            rt = sum c2 dP          rp = (i m/s) sum c2 P
//   tr = sum (-J c5) P       J = l(l+1)
//   Sg = sum c6 [-J P - 2 cot dP + 2 m^2/s^2 P]      Sh = (i m/s) sum c6 (dP - cot P)
//   tt = tr + Sg,  pp = tr - Sg,  tp = 4 Sh
//
// and with the toroidal channels (Z3 = e_r x grad_1, Z4 its spin-2 partner)
//
//   rt += -(i m/s) sum c3 P     rp += sum c3 dP
//   tt += -Sh4,  pp += Sh4,  tp += Sg4   (Sg4, Sh4: the spin-2 sums of c4)
//
// and the analysis is its quadrature adjoint with the reference's
// normalisations (c2 / J, c5 / (-2J), c6 / n6(l), c3 / J, c4 / n4(l); n6 and
// n4 calibrated at m = l as in src/vilma_tensor_sh.f90:99-133). The pointwise product with the grid
// field -Mk3 needs real space: one row kernel per (batch, latitude) does
// inverse FFT -> multiply -> forward FFT for the six dyads in LDS, so the
// grid never reaches HBM.
//
// Layouts: coefficients X[b][lam][idx] (nch = 4 or 6 channels; complex, degree-major idx =
// l(l+1)/2 + m); dyad row spectra Fd[b][dy][m][j] (complex, latitude
// contiguous, so the Legendre kernels read and write coalesced rows and the
// row kernel's column access is shared through L2 by an XCD-aware block
// order); batch b = 3*i3 + s.
#include <atomic>
#include <vector>
#include <algorithm>
#include "vh_internal.h"
#include "vh_radial_dev.h"
#include "vh_sht_dev.h"

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// range of logical block ids so that concurrently resident blocks on one XCD
// work on neighbouring rows (shared L2 lines). Grid padded to a multiple of 8.
__device__ __forceinline__ int xcd_block(int nblk) {
    const int cnt = (nblk + 7) / 8;
    return (int)(blockIdx.x % 8) * cnt + (int)(blockIdx.x / 8);
}
static inline unsigned xcd_grid(int nblk) { return (unsigned)(((nblk + 7) / 8) * 8); }

// (order m, batch group g) blocks of the tensor Legendre kernels: XCD x
// takes the orders m = x, x + 8, ... (balanced work, since a block costs
// ~ L - m + 1) and runs all batch groups of one order back to back, so the
// order's P/dP table rows are fetched into that XCD's L2 once.
__device__ __forceinline__ bool xcd_order_block(int ngrp, int L, int& m, int& g) {
    const int x = (int)(blockIdx.x % 8), r = (int)(blockIdx.x / 8);
    m = x + 8 * (r / ngrp);
    g = r % ngrp;
    return m <= L;
}
static inline unsigned xcd_order_grid(int ngrp, int L) { return (unsigned)(8 * ngrp * ((L + 1 + 7) / 8)); }

// ---------------------------------------------------------- synthesis
// Block = (order m, TB consecutive batches); lane pair (2j, 2j+1) = north
// latitude j, the even / odd l+m half of the degree sum. The P/dP table rows
// of the order are loaded once per block for all TB batches; the pair swaps
// its partial sums and the even lane writes the north row, the odd lane the
// south row.
template <int TB, bool JCHUNK, int NCH>
__global__ void __launch_bounds__(256) k_tsh_synth(vh_sht s, const double2* __restrict__ X, int nb,
                                                   double2* __restrict__ Fd) {
    // coefficient slots per batch and Legendre sums per batch
    constexpr int NS = NCH == 4 ? 5 : 8, NQ = NCH == 4 ? 7 : 12;
    extern __shared__ double2 cs[];   // [TB][NS][nl]: c1, c2, -J c5, c6, -J c6 (, c3, c4, -J c4)
    const int L = s.L, nh = s.nh, nlat = s.nlat, nlm = s.nlm;
    const int ngrp = (nb + TB - 1) / TB;
    // orders fastest within a batch group: neighbouring orders share the
    // degree-major coefficient lines in L2
    const int lb = xcd_block(ngrp * (L + 1));
    if (lb >= ngrp * (L + 1)) return;
    const int b0 = (lb / (L + 1)) * TB, m = lb % (L + 1), nl = L - m + 1;
    const int nbb = min(TB, nb - b0);
    // coefficients into LDS, SQ loads per thread issued before their stores
    constexpr int SQ = 4;
    for (int t0 = threadIdx.x; t0 < TB * NCH * nl; t0 += SQ * blockDim.x) {
        double2 v[SQ];
#pragma unroll
        for (int q = 0; q < SQ; q++) {
            const int t = t0 + q * blockDim.x;
            const int bb = t / (NCH * nl), r = t % (NCH * nl), lam = r / nl, il = r % nl, l = m + il;
            v[q] = make_double2(0.0, 0.0);
            if (t < TB * NCH * nl && bb < nbb) v[q] = X[((size_t)(b0 + bb) * NCH + lam) * nlm + l * (l + 1) / 2 + m];
        }
#pragma unroll
        for (int q = 0; q < SQ; q++) {
            const int t = t0 + q * blockDim.x;
            if (t >= TB * NCH * nl) break;
            const int bb = t / (NCH * nl), r = t % (NCH * nl), lam = r / nl, il = r % nl, l = m + il;
            const double J = (double)l * (double)(l + 1);
            double2* cb = cs + (size_t)bb * NS * nl;
            if (lam < 4) {
                cb[lam * nl + il] = lam == 2 ? make_double2(-J * v[q].x, -J * v[q].y) : v[q];
                if (lam == 3) cb[4 * nl + il] = make_double2(-J * v[q].x, -J * v[q].y);
            } else {   // c3 -> slot 5, c4 -> slot 6 and -J c4 -> slot 7
                cb[(lam + 1) * nl + il] = v[q];
                if (lam == 5) cb[7 * nl + il] = make_double2(-J * v[q].x, -J * v[q].y);
            }
        }
    }
    __syncthreads();
    const double* P = s.plm_lj + (size_t)s.mstart[m] * nh;
    const double* D = s.dplm_lj + (size_t)s.mstart[m] * nh;
    // lane pairs take north latitudes j0 + threadIdx.x / 2, 128 at a time
    // (JCHUNK: more than 128 north latitudes)
    for (int j0 = 0; j0 < (JCHUNK ? nh : 1); j0 += 128) {
    const int j = j0 + (threadIdx.x >> 1), par = threadIdx.x & 1;
    const bool live = j < nh;   // both lanes of a pair agree
    const int jj = live ? j : 0;
    // [batch][sum]: A1, A2, D2, A5, A6, L6, D6 (, A3, D3, A4, L4, D4) over
    // this lane's parity
    double2 acc[TB][NQ];
#pragma unroll
    for (int bb = 0; bb < TB; bb++)
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[bb][q] = make_double2(0.0, 0.0);
    auto term = [&](int il, double pv, double dv) __attribute__((always_inline)) {
#pragma unroll
        for (int bb = 0; bb < TB; bb++) {
            const double2* cb = cs + (size_t)bb * NS * nl;
            const double2 c1 = cb[il], c2 = cb[nl + il], c5 = cb[2 * nl + il], c6 = cb[3 * nl + il],
                          l6 = cb[4 * nl + il];
            double2* a = acc[bb];
            a[0].x = fma(c1.x, pv, a[0].x); a[0].y = fma(c1.y, pv, a[0].y);
            a[1].x = fma(c2.x, pv, a[1].x); a[1].y = fma(c2.y, pv, a[1].y);
            a[2].x = fma(c2.x, dv, a[2].x); a[2].y = fma(c2.y, dv, a[2].y);
            a[3].x = fma(c5.x, pv, a[3].x); a[3].y = fma(c5.y, pv, a[3].y);
            a[4].x = fma(c6.x, pv, a[4].x); a[4].y = fma(c6.y, pv, a[4].y);
            a[5].x = fma(l6.x, pv, a[5].x); a[5].y = fma(l6.y, pv, a[5].y);
            a[6].x = fma(c6.x, dv, a[6].x); a[6].y = fma(c6.y, dv, a[6].y);
            if constexpr (NCH == 6) {
                const double2 c3 = cb[5 * nl + il], c4 = cb[6 * nl + il], l4 = cb[7 * nl + il];
                a[7].x = fma(c3.x, pv, a[7].x); a[7].y = fma(c3.y, pv, a[7].y);
                a[8].x = fma(c3.x, dv, a[8].x); a[8].y = fma(c3.y, dv, a[8].y);
                a[9].x = fma(c4.x, pv, a[9].x); a[9].y = fma(c4.y, pv, a[9].y);
                a[10].x = fma(l4.x, pv, a[10].x); a[10].y = fma(l4.y, pv, a[10].y);
                a[11].x = fma(c4.x, dv, a[11].x); a[11].y = fma(c4.y, dv, a[11].y);
            }
        }
    };
    // VH_TSH_SU degrees of this parity per step, their table values loaded
    // together (latency of the table rows overlapped)
    int il = par;
    for (; il + 2 * (VH_TSH_SU - 1) < nl; il += 2 * VH_TSH_SU) {
        double pv[VH_TSH_SU], dv[VH_TSH_SU];
#pragma unroll
        for (int u = 0; u < VH_TSH_SU; u++) {
            pv[u] = P[(size_t)(il + 2 * u) * nh + jj];
            dv[u] = D[(size_t)(il + 2 * u) * nh + jj];
        }
#pragma unroll
        for (int u = 0; u < VH_TSH_SU; u++) term(il + 2 * u, pv[u], dv[u]);
    }
    for (; il < nl; il += 2) term(il, P[(size_t)il * nh + jj], D[(size_t)il * nh + jj]);
    // even lane: north row (E + O; dP-type E + O), odd lane: south row
    // (P-type E - O; dP-type O - E); own = this lane's parity sum
    const int row = par == 0 ? jj : nlat - 1 - jj;
    const bool write = live && !(par == 1 && row == jj);   // no south twin on the equator of an odd grid
    const double st = s.st[row], cot = s.ct[row] / st, is = 1.0 / st;
    const double mis = (double)m * is, g2 = 2.0 * mis * mis;
#pragma unroll
    for (int bb = 0; bb < TB; bb++) {
        double2 S[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const double2 own = acc[bb][q];
            const double2 oth = make_double2(__shfl_xor(own.x, 1), __shfl_xor(own.y, 1));
            const double2 e = par == 0 ? own : oth, o = par == 0 ? oth : own;
            const bool dtype = (q == 2 || q == 6 || q == 8 || q == 11);
            if (par == 0) S[q] = make_double2(e.x + o.x, e.y + o.y);
            else S[q] = dtype ? make_double2(o.x - e.x, o.y - e.y) : make_double2(e.x - o.x, e.y - o.y);
        }
        if (!write || bb >= nbb) continue;
        double2 out[6];
        out[0] = S[0];                                             // rr
        out[1] = S[2];                                             // rt
        out[2] = make_double2(-mis * S[1].y, mis * S[1].x);        // rp = i (m/s) A2
        const double2 tr = S[3];
        const double2 Sg = make_double2(S[5].x - 2.0 * cot * S[6].x + g2 * S[4].x,
                                        S[5].y - 2.0 * cot * S[6].y + g2 * S[4].y);
        const double2 hd = make_double2(S[6].x - cot * S[4].x, S[6].y - cot * S[4].y);
        const double2 Sh = make_double2(-mis * hd.y, mis * hd.x); // i (m/s) (D6 - cot A6)
        out[3] = make_double2(tr.x + Sg.x, tr.y + Sg.y);           // tt
        out[4] = make_double2(4.0 * Sh.x, 4.0 * Sh.y);             // tp
        out[5] = make_double2(tr.x - Sg.x, tr.y - Sg.y);           // pp
        if constexpr (NCH == 6) {
            out[1].x += mis * S[7].y; out[1].y -= mis * S[7].x;       // rt += -(i m/s) A3
            out[2].x += S[8].x; out[2].y += S[8].y;                   // rp += D3
            const double2 Sg4 = make_double2(S[10].x - 2.0 * cot * S[11].x + g2 * S[9].x,
                                             S[10].y - 2.0 * cot * S[11].y + g2 * S[9].y);
            const double2 h4 = make_double2(S[11].x - cot * S[9].x, S[11].y - cot * S[9].y);
            const double2 Sh4 = make_double2(-mis * h4.y, mis * h4.x);   // i (m/s) (D4 - cot A4)
            out[3].x -= Sh4.x; out[3].y -= Sh4.y;                     // tt -= Sh4
            out[5].x += Sh4.x; out[5].y += Sh4.y;                     // pp += Sh4
            out[4].x += Sg4.x; out[4].y += Sg4.y;                     // tp += Sg4
        }
#pragma unroll
        for (int dy = 0; dy < 6; dy++)
            Fd[(((size_t)(b0 + bb) * 6 + dy) * (L + 1) + m) * nlat + row] = out[dy];
    }
    }
}

// ----------------------------------------------------------- analysis
// FP64 MFMA form. For one order m and degree parity p the analysis is the
// product  S[l][col] = sum_j T[l][j] B[j][col]  with T = P or dP over the north
// latitudes and B the latitude pair's folded, weighted fields. Four channels:
// per batch five P-type complex fields (f1, f2P, f5, fD, fb: 10 real
// columns) and two dP-type (f2D, fd: 4 columns); block = (order m, 4
// batches): P-type columns 0..39 (three 16-wide tiles, 40..47 zero), dP-type
// 48..63 (one tile). Six channels add per batch the P-type f3P, fD4, fb4 and
// the dP-type f3D, fd4 (16 + 8 real columns); block = (order m, 2 batches):
// P-type 0..31, dP-type 32..47 (three tiles). Latitudes are taken TA_JC at a
// time: the block folds them into LDS (B), then every wavefront runs its
// (parity, 16-degree) row tiles through v_mfma_f64_16x16x4 (lane l:
// A[l%16][l/16], B[l/16][l%16], D[4r+l/16][l%16]) with the P/dP fragments
// read straight from the [m][j][l] tables. The per-(degree, batch)
// normalisations follow from the accumulated tiles.
typedef double vh_d4 __attribute__((ext_vector_type(4)));
#define TA_JC VH_TA_JC   // latitudes per LDS chunk
// column layout of one channel count
template <int NCH> struct ta_cols {
    static constexpr int TB = NCH == 4 ? 4 : 2;        // batches per block
    static constexpr int NQP = NCH == 4 ? 5 : 8;       // P-type complex fields per batch
    static constexpr int NQD = NCH == 4 ? 2 : 4;       // dP-type complex fields per batch
    static constexpr int CP = 2 * NQP, CD = 2 * NQD;   // real columns per batch
    static constexpr int PC = NCH == 4 ? 48 : 32;      // first dP-type column
    static constexpr int NT = NCH == 4 ? 4 : 3;        // 16-column tiles
    static_assert(TB * CP <= PC && PC + TB * CD == 16 * NT && 16 * NT <= 64, "analysis column layout");
};
// TA_ITEMS: row tiles per wavefront (4 wavefronts): 3 holds the 16-degree
// tiles of both parities up to lmax 159 (ceil(ne/16) + ceil(no/16) <= 12),
// 4 up to lmax 255
template <int TA_ITEMS, int NCH>
__global__ void __launch_bounds__(256) k_tsh_analysis(vh_sht s, const double2* __restrict__ Fd, int nb,
                                                      const double* __restrict__ n6, const double* __restrict__ n4,
                                                      double2* __restrict__ C) {
    using K = ta_cols<NCH>;
    constexpr int TB = K::TB, NQP = K::NQP, NQD = K::NQD, CP = K::CP, CD = K::CD, PC = K::PC, NT = K::NT;
    // one LDS array, two views: the folded latitude chunk Bs [parity][latitude
    // in chunk][column] (rows padded: the fold stores stride over banks) while
    // the products run, then the per-wavefront staging Ds of one row tile in
    // the epilogue (the same 4160 doubles)
    __shared__ double Sh[2 * TA_JC * 65];
    static_assert(2 * TA_JC * 65 >= 4 * 16 * 65, "epilogue staging must fit the chunk buffer");
    double (*Bs)[TA_JC][65] = (double (*)[TA_JC][65])Sh;
    double (*Ds)[16][65] = (double (*)[16][65])Sh;
    const int L = s.L, nh = s.nh, nlat = s.nlat, nlm = s.nlm;
    const int ngrp = (nb + TB - 1) / TB;
    int m, g;
    if (!xcd_order_block(ngrp, L, m, g)) return;
    const int b0 = g * TB, nl = L - m + 1, nbb = min(TB, nb - b0);
    const int ne = (nl + 1) / 2, no = nl / 2;
    const int nte = (ne + 15) / 16, nto = (no + 15) / 16, nitems = nte + nto;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const double ms = (double)m;
    const double* P = s.plm_jl + (size_t)s.mstart[m] * nh;
    const double* D = s.dplm_jl + (size_t)s.mstart[m] * nh;
    vh_d4 acc[TA_ITEMS][NT];
#pragma unroll
    for (int it = 0; it < TA_ITEMS; it++)
#pragma unroll
        for (int t = 0; t < NT; t++) acc[it][t] = vh_d4{0.0, 0.0, 0.0, 0.0};
    // the wavefront's row tiles: parity, degree offset, live rows
    int il_[TA_ITEMS], par_[TA_ITEMS];
    bool row_[TA_ITEMS];
#pragma unroll
    for (int it = 0; it < TA_ITEMS; it++) {
        const int item = wave + 4 * it;
        const int p = item < nte ? 0 : 1, rt = item < nte ? item : item - nte;
        const int i = rt * 16 + (lane & 15);
        par_[it] = p;
        il_[it] = 2 * i + p;
        row_[it] = item < nitems && i < (p ? no : ne);
    }
    // P/dP fragments of k-step ks of the chunk at j0 (the A operand: rows of
    // the [m][j][l] tables, read straight from global memory)
    auto frag = [&](int j0c, int ks, double (&fP)[TA_ITEMS], double (&fD)[TA_ITEMS]) {
        const int j = j0c + ks * 4 + (lane >> 4);
#pragma unroll
        for (int it = 0; it < TA_ITEMS; it++) {
            const bool ok = row_[it] && j < nh;
            fP[it] = ok ? P[(size_t)j * nl + il_[it]] : 0.0;
            fD[it] = ok ? D[(size_t)j * nl + il_[it]] : 0.0;
        }
    };
    constexpr int PF = VH_TA_PF, KS = TA_JC / 4;   // fragments in flight ahead of the products
    static_assert(PF >= 1 && PF < KS, "prefetch distance");
    for (int j0 = 0; j0 < nh; j0 += TA_JC) {
        // the chunk's first fragments are in flight while the block folds
        double aP[PF + 1][TA_ITEMS], aD[PF + 1][TA_ITEMS];
#pragma unroll
        for (int q = 0; q < PF; q++) frag(j0, q, aP[q], aD[q]);
        __syncthreads();   // previous chunk's fragments consumed
        // ---- fold the latitude pairs (j, nlat-1-j), j = j0 + jl, of the block's batches
        for (int t = threadIdx.x; t < TB * TA_JC; t += blockDim.x) {
            const int bb = t / TA_JC, jl = t % TA_JC, j = j0 + jl;
            double2 f[2][NQP + NQD];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int q = 0; q < NQP + NQD; q++) f[h][q] = make_double2(0.0, 0.0);
            if (bb < nbb && j < nh) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int row = h == 0 ? j : nlat - 1 - j;
                    if (h == 1 && row == j) break;
                    double2 d[6];
#pragma unroll
                    for (int dy = 0; dy < 6; dy++) d[dy] = Fd[(((size_t)(b0 + bb) * 6 + dy) * (L + 1) + m) * nlat + row];
                    const double w = s.w[row], st = s.st[row], cot = s.ct[row] / st, is = 1.0 / st;
                    const double mis = ms * is;
                    const double2 Dv = make_double2(d[3].x - d[5].x, d[3].y - d[5].y);     // tt - pp
                    const double2 tp = d[4];
                    const double g2 = 2.0 * mis * mis, hc = 2.0 * mis * cot;
                    f[h][0] = make_double2(w * d[0].x, w * d[0].y);                      // rr
                    f[h][1] = make_double2(w * (mis * d[2].y), w * (-mis * d[2].x));      // (-i m/s) rp
                    f[h][2] = make_double2(w * (d[3].x + d[5].x), w * (d[3].y + d[5].y)); // tt + pp
                    f[h][3] = make_double2(w * Dv.x, w * Dv.y);
                    // 2 m^2/s^2 D + 2 i m cot/s tp
                    f[h][4] = make_double2(w * (g2 * Dv.x - hc * tp.y), w * (g2 * Dv.y + hc * tp.x));
                    f[h][NQP] = make_double2(w * d[1].x, w * d[1].y);                     // rt
                    f[h][NQP + 1] = make_double2(w * (-2.0 * cot * Dv.x + 2.0 * mis * tp.y),   // -2 cot D - 2 i m/s tp
                                                 w * (-2.0 * cot * Dv.y - 2.0 * mis * tp.x));
                    if constexpr (NCH == 6) {
                        f[h][5] = make_double2(w * (-mis * d[1].y), w * (mis * d[1].x));    // (i m/s) rt
                        f[h][6] = make_double2(w * tp.x, w * tp.y);                          // tp
                        // 2 m^2/s^2 tp - 2 i m cot/s D
                        f[h][7] = make_double2(w * (g2 * tp.x + hc * Dv.y), w * (g2 * tp.y - hc * Dv.x));
                        f[h][NQP + 2] = make_double2(w * d[2].x, w * d[2].y);                 // rp
                        f[h][NQP + 3] = make_double2(w * (-2.0 * cot * tp.x - 2.0 * mis * Dv.y),   // -2 cot tp + 2 i m/s D
                                                     w * (-2.0 * cot * tp.y + 2.0 * mis * Dv.x));
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NQP + NQD; q++) {
                const bool dtype = q >= NQP;
                const double2 a = f[0][q], c = f[1][q];
                const double2 plus = make_double2(a.x + c.x, a.y + c.y), minus = make_double2(a.x - c.x, a.y - c.y);
                const double2 ev = dtype ? minus : plus, od = dtype ? plus : minus;   // even / odd l+m
                const int col = dtype ? PC + bb * CD + 2 * (q - NQP) : bb * CP + 2 * q;
                Bs[0][jl][col] = ev.x; Bs[0][jl][col + 1] = ev.y;
                Bs[1][jl][col] = od.x; Bs[1][jl][col + 1] = od.y;
            }
            if (bb == 0)
                for (int c = TB * CP; c < PC; c++) { Bs[0][jl][c] = 0.0; Bs[1][jl][c] = 0.0; }
        }
        __syncthreads();
        // ---- MFMA over this chunk's latitudes for the wavefront's row tiles;
        // the fragments of the next PF k-steps are in flight while one runs
        // (a ring of PF + 1 register sets, indices fixed by the full unroll)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            if (ks + PF < KS) frag(j0, ks + PF, aP[(ks + PF) % (PF + 1)], aD[(ks + PF) % (PF + 1)]);
            const int cur = ks % (PF + 1);
            const int jl = ks * 4 + (lane >> 4);
#pragma unroll
            for (int it = 0; it < TA_ITEMS; it++) {
                if (wave + 4 * it >= nitems) break;
                const double* Bp = &Bs[par_[it]][jl][0];
#pragma unroll
                for (int t = 0; t < NT; t++)
                    acc[it][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(16 * t < PC ? aP[cur][it] : aD[cur][it],
                                                                      Bp[16 * t + (lane & 15)], acc[it][t], 0, 0, 0);
            }
        }
    }
    // ---- epilogue: stage each row tile, then lane = (degree row, batch)
    __syncthreads();   // every wavefront is done with the last chunk before its staging overwrites it
#pragma unroll
    for (int it = 0; it < TA_ITEMS; it++) {
        const int item = wave + 4 * it;
        if (item >= nitems) break;
        const int p = item < nte ? 0 : 1, rt = item < nte ? item : item - nte;
        const int np_ = p ? no : ne;
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) Ds[wave][4 * r + (lane >> 4)][t * 16 + (lane & 15)] = acc[it][t][r];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int rr = lane & 15, bb = lane >> 4, i = rt * 16 + rr;
        if (i < np_ && bb < nbb) {
            const int l = m + 2 * i + p;
            const double* row = &Ds[wave][rr][0];
            const double* pr = row + bb * CP;        // P-type columns of this batch
            const double* dr = row + PC + bb * CD;   // dP-type columns
            const double2 s1 = make_double2(pr[0], pr[1]);
            const double2 s2 = make_double2(pr[2] + dr[0], pr[3] + dr[1]);
            const double2 s5 = make_double2(pr[4], pr[5]);
            const double2 sD = make_double2(pr[6], pr[7]);
            const double2 sb = make_double2(pr[8], pr[9]);
            const double2 sd = make_double2(dr[2], dr[3]);
            const double J = (double)l * (double)(l + 1);
            const size_t idx = (size_t)l * (l + 1) / 2 + m;
            double2* Cb = C + (size_t)(b0 + bb) * NCH * nlm;
            Cb[idx] = s1;
            Cb[nlm + idx] = J > 0.0 ? make_double2(s2.x / J, s2.y / J) : make_double2(0.0, 0.0);
            Cb[2 * nlm + idx] = J > 0.0 ? make_double2(s5.x / (-2.0 * J), s5.y / (-2.0 * J)) : make_double2(0.0, 0.0);
            const double2 raw = make_double2(-J * sD.x + sb.x + sd.x, -J * sD.y + sb.y + sd.y);
            double2 c6 = raw;
            if (n6) {
                const double n = n6[l];
                c6 = n != 0.0 ? make_double2(raw.x / n, raw.y / n) : make_double2(0.0, 0.0);
            }
            Cb[3 * nlm + idx] = c6;
            if constexpr (NCH == 6) {
                // c3 = (1/J) int v . conj(e_r x grad_1 Y); c4: the toroidal
                // spin-2 adjoint 1/2 Sg*(tp) - Sh*(tt - pp) over n4
                const double2 s3 = make_double2(pr[10] + dr[4], pr[11] + dr[5]);
                const double2 sD4 = make_double2(pr[12], pr[13]);
                const double2 sb4 = make_double2(pr[14], pr[15]);
                const double2 sd4 = make_double2(dr[6], dr[7]);
                Cb[4 * nlm + idx] = J > 0.0 ? make_double2(s3.x / J, s3.y / J) : make_double2(0.0, 0.0);
                const double2 raw4 = make_double2(0.5 * (-J * sD4.x + sb4.x + sd4.x), 0.5 * (-J * sD4.y + sb4.y + sd4.y));
                double2 c4 = raw4;
                if (n4) {
                    const double n = n4[l];
                    c4 = n != 0.0 ? make_double2(raw4.x / n, raw4.y / n) : make_double2(0.0, 0.0);
                }
                Cb[5 * nlm + idx] = c4;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ------------------------------------------------------------- row pass
// Block = (batch b, latitude row j), three wavefronts, one per dyad pair
// (rr, rt), (rp, tt), (tp, pp) packed as one complex FFT of two real rows.
//   mode 0: inverse FFT, multiply by -rate*dt (rate row of the batch's
//           element), forward FFT, in place in Fd
//   mode 1: inverse FFT to grid[b][dy][j][i]
//   mode 2: forward FFT of grid[b][dy][j][i] into Fd
//   mode 3: as mode 0 with -M/(1 + M/2), M = rate*dt (trapezoidal rule,
//           src/vilma_response.f90:1447-1478, in increment form)
// twl: the twiddles are copied into the block's LDS (else read from s.twp).
__global__ void __launch_bounds__(192) k_t3_rows(vh_sht s, const double2* Fin, double2* Fd, int nb, int mode,
                                                 const double* __restrict__ rate, int bpe, double dt,
                                                 double* __restrict__ grid, int twl) {
    extern __shared__ double2 lds[];
    const int L = s.L, N = s.nphi, nlat = s.nlat;
    const int lb = xcd_block(nb * nlat);
    if (lb >= nb * nlat) return;
    const int b = lb / nlat, j = lb % nlat;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // one LDS buffer per wavefront when the transform runs in place, else two
    const int nbuf = fft_fixed_size(N) ? 1 : 2;
    double2* a = lds + (size_t)wave * nbuf * N;
    double2* c = nbuf == 2 ? a + N : nullptr;
    double2* TW = lds + (size_t)3 * nbuf * N;                    // [s.ntw] when twl
    double* Rrow = (double*)(TW + (twl ? s.ntw : 0));            // the batch element's rate row [N]
    const double2* tw = twl ? TW : nullptr;
    const int d0 = 2 * wave, d1 = 2 * wave + 1;
    const double2* I0 = Fin + ((size_t)b * 6 + d0) * (L + 1) * nlat + j;
    const double2* I1 = Fin + ((size_t)b * 6 + d1) * (L + 1) * nlat + j;
    double2* F0 = Fd + ((size_t)b * 6 + d0) * (L + 1) * nlat + j;
    double2* F1 = Fd + ((size_t)b * 6 + d1) * (L + 1) * nlat + j;
    // the twiddles and the rate row go into LDS first, copied by the memory
    // unit (no registers held) while the spectra's loads are in flight; the
    // block waits for them once, before its first transform
    if (twl)
        for (int t0 = wave * 64; t0 < s.ntw; t0 += 192)
            if (t0 + lane < s.ntw) lds_dma16(s.twp + t0 + lane, TW + t0);
    const bool prod = mode == 0 || mode == 3;
    if (prod) {
        const double* rr = rate + ((size_t)(b / bpe) * nlat + j) * N;
        if ((N & 1) == 0) {   // 16-byte aligned pairs of the row
            for (int c0 = wave * 64; c0 < N / 2; c0 += 192)
                if (c0 + lane < N / 2) lds_dma16((const double2*)rr + c0 + lane, (double2*)Rrow + c0);
        } else {
            for (int i = threadIdx.x; i < N; i += blockDim.x) Rrow[i] = rr[i];
        }
    }
    double2* z;
    if (mode != 2) {
        for (int i = lane; i < N; i += 64) a[i] = make_double2(0.0, 0.0);
        wave_lds_sync();
        for (int m0 = 0; m0 <= L; m0 += 64 * VH_FFT_U) {
            double2 A[VH_FFT_U], B[VH_FFT_U];
#pragma unroll
            for (int u = 0; u < VH_FFT_U; u++) {
                const int m = m0 + lane + 64 * u;
                A[u] = m <= L ? I0[(size_t)m * nlat] : make_double2(0.0, 0.0);
                B[u] = m <= L ? I1[(size_t)m * nlat] : make_double2(0.0, 0.0);
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
        __builtin_amdgcn_s_waitcnt(0);   // the copies into LDS are complete
        __syncthreads();                 // and visible to the whole block
        z = fft_wave_ip<true>(a, c, s, lane, tw);
        if (mode == 1) {
            double* g0 = grid + (((size_t)b * 6 + d0) * nlat + j) * N;
            double* g1 = grid + (((size_t)b * 6 + d1) * nlat + j) * N;
            for (int i = lane; i < N; i += 64) { g0[i] = z[i].x; g1[i] = z[i].y; }
            return;
        }
        for (int i = lane; i < N; i += 64) {
            const double M0 = Rrow[i] * dt;
            const double M = mode == 3 ? M0 / (1.0 + 0.5 * M0) : M0;   // trapezoidal weight
            z[i] = make_double2(-M * z[i].x, -M * z[i].y);
        }
        wave_lds_sync();
    } else {
        const double* g0 = grid + (((size_t)b * 6 + d0) * nlat + j) * N;
        const double* g1 = grid + (((size_t)b * 6 + d1) * nlat + j) * N;
        for (int i = lane; i < N; i += 64) a[i] = make_double2(g0[i], g1[i]);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        z = a;
    }
    double2* other = (z == a) ? c : a;
    double2* y = fft_wave_ip<false>(z, other, s, lane, tw);
    const double scale = 2.0 * VH_PI / N;
    for (int m = lane; m <= L; m += 64) {
        double2 zm = y[m], zc = y[(N - m) % N];
        zc.y = -zc.y;
        const double2 A = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        const double2 Dd = csub(zm, zc);
        const double2 B = make_double2(0.5 * Dd.y, -0.5 * Dd.x);
        F0[(size_t)m * nlat] = make_double2(scale * A.x, scale * A.y);
        F1[(size_t)m * nlat] = make_double2(scale * B.x, scale * B.y);
    }
}

// Row pass of modes 0 and 3 over RJ consecutive latitudes and one dyad pair
// per block, for the grid widths with a compile-time transform: wavefront w
// transforms latitude j0 + w; the spectra are read and written as whole
// lines (the RJ latitudes of one order and dyad are RJ x 16 contiguous
// bytes), so the vector memory unit handles RJ times fewer lines than with
// one latitude per block. The operations of k_t3_rows in the same order.
// LDS: RJ transform buffers of N points (stride N + 1), the twiddles, RJ rate rows.
template <int N, int RJ>
__global__ void __launch_bounds__(64 * RJ) k_t3_rows_g(vh_sht s, const double2* Fin, double2* Fd, int nb, int mode,
                                                       const double* __restrict__ rate, int bpe, double dt) {
    static_assert(64 % RJ == 0, "RJ divides the wavefront");
    constexpr int MS = 64;                     // orders per load round (64 RJ threads over RJ latitudes)
    constexpr int MU = (N / 2 + MS - 1) / MS;  // load rounds (2L + 1 <= N)
    constexpr int NP = N + 1;                  // buffer stride: the RJ buffers' entry m in distinct LDS banks
    extern __shared__ double2 lds[];
    const int L = s.L, nlat = s.nlat, ng = (nlat + RJ - 1) / RJ;
    const int lb = xcd_block(nb * ng * 3);
    if (lb >= nb * ng * 3) return;
    // the three pairs of one (batch, latitude group) are neighbours on one XCD (shared rate rows)
    const int p = lb % 3, g = (lb / 3) % ng, b = lb / (3 * ng);
    const int j0 = g * RJ, nj = min(RJ, nlat - j0);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    double2* TW = lds + (size_t)RJ * NP;             // [s.ntw]
    double* Rr = (double*)(TW + s.ntw);              // [RJ][N]
    // twiddles and this wavefront's rate row copied into LDS by the memory unit
    for (int t0 = wave * 64; t0 < s.ntw; t0 += 64 * RJ)
        if (t0 + lane < s.ntw) lds_dma16(s.twp + t0 + lane, TW + t0);
    if (wave < nj) {
        const double2* rr = (const double2*)(rate + ((size_t)(b / bpe) * nlat + j0 + wave) * N);
        double2* rd = (double2*)(Rr + (size_t)wave * N);
#pragma unroll
        for (int c0 = 0; c0 < N / 2; c0 += 64)
            if (N / 2 % 64 == 0 || c0 + lane < N / 2) lds_dma16(rr + c0 + lane, rd + c0);
    }
    // spectra: thread (order mm + 64u, latitude j0 + jj)
    const int jj = tid % RJ, mm = tid / RJ;
    const bool jok = jj < nj;
    const size_t dstr = (size_t)(L + 1) * nlat;      // between dyads
    const size_t off = ((size_t)b * 6 + 2 * p) * dstr + j0 + jj;
    double2 A[MU], B[MU];
#pragma unroll
    for (int u = 0; u < MU; u++) {
        const int m = mm + MS * u;
        const bool ok = jok && m <= L;
        A[u] = ok ? Fin[off + (size_t)m * nlat] : make_double2(0.0, 0.0);
        B[u] = ok ? Fin[off + dstr + (size_t)m * nlat] : make_double2(0.0, 0.0);
    }
    {
        double2* a = lds + (size_t)wave * NP;          // the orders without spectrum
        for (int i = L + 1 + lane; i < N - L; i += 64) a[i] = make_double2(0.0, 0.0);
        double2* aj = lds + (size_t)jj * NP;
#pragma unroll
        for (int u = 0; u < MU; u++) {
            const int m = mm + MS * u;
            if (!jok || m > L) continue;
            double2 x = A[u], y = B[u];
            if (m == 0) { x.y = 0.0; y.y = 0.0; }
            aj[m] = make_double2(x.x - y.y, x.y + y.x);
            if (m > 0) aj[N - m] = make_double2(x.x + y.y, -x.y + y.x);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);   // the copies into LDS are complete
    __syncthreads();
    if (wave < nj) {
        double2* z = lds + (size_t)wave * NP;
        const double* R = Rr + (size_t)wave * N;
        // the last pass's twiddles stay in registers for both transforms; the
        // product with -rate dt is applied as the inverse transform stores
        double2 tl[fft_last<N>::R - 1];
        fft_last_twiddles<N>(TW, lane, tl);
        fft_ip_reg_last<N, true>(z, TW, tl, lane, [&](int, int i, double2 g) __attribute__((always_inline)) {
            const double M0 = R[i] * dt;
            const double M = mode == 3 ? M0 / (1.0 + 0.5 * M0) : M0;   // trapezoidal weight
            return make_double2(-M * g.x, -M * g.y);
        });
        wave_lds_sync();
        fft_ip_reg_last<N, false>(z, TW, tl, lane, [](int, int, double2 g) __attribute__((always_inline)) { return g; });
    }
    __syncthreads();
    const double scale = 2.0 * VH_PI / N;
    const double2* y = lds + (size_t)jj * NP;
#pragma unroll
    for (int u = 0; u < MU; u++) {
        const int m = mm + MS * u;
        if (!jok || m > L) continue;
        double2 zm = y[m], zc = y[(N - m) % N];
        zc.y = -zc.y;
        const double2 Aq = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y + zc.y));
        const double2 Dd = csub(zm, zc);
        const double2 Bq = make_double2(0.5 * Dd.y, -0.5 * Dd.x);
        Fd[off + (size_t)m * nlat] = make_double2(scale * Aq.x, scale * Aq.y);
        Fd[off + dstr + (size_t)m * nlat] = make_double2(scale * Bq.x, scale * Bq.y);
    }
}

// ------------------------------------------------------------- gather
// Toroidal channels 4, 5 of X: taut_s + wgt * (eps_s(W1) [+ eps_s(W0)]) of
// slot k, element em (memory nodes n0, n1), written after the spheroidal four.
__device__ __forceinline__ void gather_tor(int nk, int nlm, int k, int em, int n0, int n1, double Jr, double wgt,
                                           const double* __restrict__ taut, const double* __restrict__ W1,
                                           const double* __restrict__ W0, double2* Xb) {
    double st[2][3][2];
#pragma unroll
    for (int ri = 0; ri < 2; ri++) {
        tor_strain(W1[((size_t)n0 * 2 + ri) * nk + k], W1[((size_t)n1 * 2 + ri) * nk + k], Jr, st[ri][0], st[ri][1],
                   st[ri][2]);
        if (W0) {
            double s0[3][2];
            tor_strain(W0[((size_t)n0 * 2 + ri) * nk + k], W0[((size_t)n1 * 2 + ri) * nk + k], Jr, s0[0], s0[1], s0[2]);
#pragma unroll
            for (int c = 0; c < 3; c++) { st[ri][c][0] += s0[c][0]; st[ri][c][1] += s0[c][1]; }
        }
    }
    const double* t = taut + (size_t)em * 12 * nk + k;
#pragma unroll
    for (int sh = 0; sh < 3; sh++)
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double re = t[(size_t)((sh * 2 + 0) * 2 + q) * nk] + wgt * st[0][sh][q];
            const double im = t[(size_t)((sh * 2 + 1) * 2 + q) * nk] + wgt * st[1][sh][q];
            Xb[(size_t)(sh * 6 + 4 + q) * nlm + k + 1] = make_double2(re, im);
        }
}

// X[(i3*3 + s)*nch + lam][k+1] = tau_s + 2 mu eps_s of slot k, element e3em[i3].
__global__ void __launch_bounds__(256) k_t3_gather(int nr, int nk, int nlm, vh_dev v, const int* __restrict__ e3em,
        const double* __restrict__ tau, const double* __restrict__ sigma, const double* __restrict__ dN,
        double2* __restrict__ X, int nch, const double* __restrict__ taut, const double* __restrict__ dWt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i3 = blockIdx.y;
    double2* Xb = X + (size_t)i3 * 3 * nch * nlm;
    if (k == 0)
        for (int q = 0; q < 3 * nch; q++) Xb[(size_t)q * nlm] = make_double2(0.0, 0.0);
    if (k >= nk) return;
    const int em = e3em[i3], e = v.emap[em];
    const int l = v.kdeg[k];
    const double Jr = (double)l * (double)(l + 1);
    const double sre = sigma[2 * (k + 1)], sim = sigma[2 * (k + 1) + 1];
    const double* xU = v.xUn + (size_t)(l - 1) * nr;
    const double* xV = v.xVn + (size_t)(l - 1) * nr;
    const int n0 = v.mnode[e], n1 = v.mnode[e + 1];
    const double* q0 = dN + (size_t)n0 * 4 * nk + k;
    const double* q1 = dN + (size_t)n1 * 4 * nk + k;
    double U0r = sre * xU[e] + q0[0], U0i = sim * xU[e] + q0[(size_t)nk];
    double V0r = sre * xV[e] + q0[(size_t)2 * nk], V0i = sim * xV[e] + q0[(size_t)3 * nk];
    double U1r = sre * xU[e + 1] + q1[0], U1i = sim * xU[e + 1] + q1[(size_t)nk];
    double V1r = sre * xV[e + 1] + q1[(size_t)2 * nk], V1i = sim * xV[e + 1] + q1[(size_t)3 * nk];
    double sc[2][3][4];
    strain_coeffs(U0r, U1r, V0r, V1r, Jr, sc[0][0], sc[0][1], sc[0][2]);
    strain_coeffs(U0i, U1i, V0i, V1i, Jr, sc[1][0], sc[1][1], sc[1][2]);
    const double twoMu = 2.0 * v.mu[e];
    const double* t = tau + (size_t)em * 24 * nk + k;
#pragma unroll
    for (int sh = 0; sh < 3; sh++)
#pragma unroll
        for (int lam = 0; lam < 4; lam++) {
            double re = t[(size_t)((sh * 2 + 0) * 4 + lam) * nk] + twoMu * sc[0][sh][lam];
            double im = t[(size_t)((sh * 2 + 1) * 4 + lam) * nk] + twoMu * sc[1][sh][lam];
            Xb[(size_t)(sh * nch + lam) * nlm + k + 1] = make_double2(re, im);
        }
    if (nch == 6) gather_tor(nk, nlm, k, em, n0, n1, Jr, twoMu, taut, dWt, nullptr, Xb);
}

// X[(i3*3 + s)*4 + lam][k+1] = tau_s + mu (eps_n,s + eps_{n+1},s) for the
// trapezoidal rule (gather_tensor_coeffs_trap, src/vilma_response.f90:1371-1445;
// the rule's increment -M/(1+M/2) [tau_n + mu (eps_n + eps_{n+1})] is
// linear in this one tensor, so the three syntheses of the reference become one)
__global__ void __launch_bounds__(256) k_t3_gather_trap(int nr, int nk, int nlm, vh_dev v, const int* __restrict__ e3em,
        const double* __restrict__ tau, const double* __restrict__ sigma_n, const double* __restrict__ dN0,
        const double* __restrict__ sigma, const double* __restrict__ dN1, double2* __restrict__ X, int nch,
        const double* __restrict__ taut, const double* __restrict__ dWt0, const double* __restrict__ dWt1) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i3 = blockIdx.y;
    double2* Xb = X + (size_t)i3 * 3 * nch * nlm;
    if (k == 0)
        for (int q = 0; q < 3 * nch; q++) Xb[(size_t)q * nlm] = make_double2(0.0, 0.0);
    if (k >= nk) return;
    const int em = e3em[i3], e = v.emap[em];
    const int l = v.kdeg[k];
    const double Jr = (double)l * (double)(l + 1);
    const double* xU = v.xUn + (size_t)(l - 1) * nr;
    const double* xV = v.xVn + (size_t)(l - 1) * nr;
    const int n0 = v.mnode[e], n1 = v.mnode[e + 1];
    double sc[2][3][4];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int sh = 0; sh < 3; sh++)
#pragma unroll
            for (int lam = 0; lam < 4; lam++) sc[q][sh][lam] = 0.0;
#pragma unroll
    for (int w = 0; w < 2; w++) {   // eps_n, then eps_{n+1}
        const double* sg = w == 0 ? sigma_n : sigma;
        const double* dN = w == 0 ? dN0 : dN1;
        const double sre = sg[2 * (k + 1)], sim = sg[2 * (k + 1) + 1];
        const double* q0 = dN + (size_t)n0 * 4 * nk + k;
        const double* q1 = dN + (size_t)n1 * 4 * nk + k;
        double t[2][3][4];
        strain_coeffs(sre * xU[e] + q0[0], sre * xU[e + 1] + q1[0], sre * xV[e] + q0[(size_t)2 * nk],
                      sre * xV[e + 1] + q1[(size_t)2 * nk], Jr, t[0][0], t[0][1], t[0][2]);
        strain_coeffs(sim * xU[e] + q0[(size_t)nk], sim * xU[e + 1] + q1[(size_t)nk], sim * xV[e] + q0[(size_t)3 * nk],
                      sim * xV[e + 1] + q1[(size_t)3 * nk], Jr, t[1][0], t[1][1], t[1][2]);
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int sh = 0; sh < 3; sh++)
#pragma unroll
                for (int lam = 0; lam < 4; lam++) sc[q][sh][lam] += t[q][sh][lam];
    }
    const double mu = v.mu[e];
    const double* t = tau + (size_t)em * 24 * nk + k;
#pragma unroll
    for (int sh = 0; sh < 3; sh++)
#pragma unroll
        for (int lam = 0; lam < 4; lam++) {
            double re = t[(size_t)((sh * 2 + 0) * 4 + lam) * nk] + mu * sc[0][sh][lam];
            double im = t[(size_t)((sh * 2 + 1) * 4 + lam) * nk] + mu * sc[1][sh][lam];
            Xb[(size_t)(sh * nch + lam) * nlm + k + 1] = make_double2(re, im);
        }
    if (nch == 6) gather_tor(nk, nlm, k, em, n0, n1, Jr, mu, taut, dWt1, dWt0, Xb);
}

// ------------------------------------------------------------- host side
// LDS of the synthesis (VH_TSH_TB_SYNTH batches per block with four
// channels, two batches with six) and of the row pass
#define TSH_TB6 2
static size_t synth_lds(const vh_sht& s, int nch) {
    return nch == 4 ? (size_t)VH_TSH_TB_SYNTH * 5 * (s.L + 1) * sizeof(double2)
                    : (size_t)TSH_TB6 * 8 * (s.L + 1) * sizeof(double2);
}
// LDS of the row pass: the wavefronts' FFT buffers, the twiddles (twl), the rate row
static size_t rows_lds(const vh_sht& s, bool twl) {
    return (size_t)3 * (fft_fixed_size(s.nphi) ? 1 : 2) * s.nphi * sizeof(double2) +
           (twl ? (size_t)s.ntw * sizeof(double2) : 0) + (size_t)s.nphi * sizeof(double);
}

void vh_tsh_synth(hipStream_t st, const vh_sht& s, const double2* X, int nb, double2* Fd, int nch) {
    if (nb <= 0) return;
    if (nch != 4 && nch != 6) throw vh_error("tensor synthesis: 4 or 6 channels");
    const int tb = nch == 4 ? VH_TSH_TB_SYNTH : TSH_TB6;
    const int ngrp = (nb + tb - 1) / tb;
    const int nthr = std::min(256, ((2 * s.nh + 63) / 64) * 64);
    const dim3 g(xcd_grid(ngrp * (s.L + 1)));
    const size_t lds = synth_lds(s, nch);
    if (nch == 4) {
        if (s.nh > 128) k_tsh_synth<VH_TSH_TB_SYNTH, true, 4><<<g, nthr, lds, st>>>(s, X, nb, Fd);
        else k_tsh_synth<VH_TSH_TB_SYNTH, false, 4><<<g, nthr, lds, st>>>(s, X, nb, Fd);
    } else {
        if (s.nh > 128) k_tsh_synth<TSH_TB6, true, 6><<<g, nthr, lds, st>>>(s, X, nb, Fd);
        else k_tsh_synth<TSH_TB6, false, 6><<<g, nthr, lds, st>>>(s, X, nb, Fd);
    }
    HIPCHECK(hipGetLastError());
}

void vh_tsh_analysis(hipStream_t st, const vh_sht& s, const double2* Fd, int nb, const double* n6, double2* C,
                     int nch, const double* n4) {
    if (nb <= 0) return;
    if (nch != 4 && nch != 6) throw vh_error("tensor analysis: 4 or 6 channels");
    // row tiles of one order: ceil(ne/16) + ceil(no/16) <= 4 wavefronts x TA_ITEMS
    const int tiles = ((s.L + 2) / 2 + 15) / 16 + ((s.L + 1) / 2 + 15) / 16;
    if (tiles > 16) throw vh_error("tensor analysis: lmax above 255 is not supported by the MFMA tiling");
    const int tb = nch == 4 ? ta_cols<4>::TB : ta_cols<6>::TB;
    const int ngrp = (nb + tb - 1) / tb;
    const dim3 g(xcd_order_grid(ngrp, s.L));
    if (nch == 4) {
        if (tiles <= 12) k_tsh_analysis<3, 4><<<g, 256, 0, st>>>(s, Fd, nb, n6, nullptr, C);
        else k_tsh_analysis<4, 4><<<g, 256, 0, st>>>(s, Fd, nb, n6, nullptr, C);
    } else {
        if (tiles <= 12) k_tsh_analysis<3, 6><<<g, 256, 0, st>>>(s, Fd, nb, n6, n4, C);
        else k_tsh_analysis<4, 6><<<g, 256, 0, st>>>(s, Fd, nb, n6, n4, C);
    }
    HIPCHECK(hipGetLastError());
}

// The grouped row pass (k_t3_rows_g) of one grid width: its LDS bytes, and
// the dynamic-LDS limit raised once per process (a function attribute; set
// before any graph capture through vh_t3_rows_prepare).
template <int N>
static size_t rows_g_lds(const vh_sht& s) {
    return ((size_t)VH_ROWS_RJ * (N + 1) + s.ntw) * sizeof(double2) + (size_t)VH_ROWS_RJ * N * sizeof(double);
}
template <int N>
static void rows_g_limit(size_t lds) {
    static std::atomic<size_t> set{0};
    if (lds > 65536 && set.load() < lds) {
        HIPCHECK(hipFuncSetAttribute((const void*)k_t3_rows_g<N, VH_ROWS_RJ>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        set.store(lds);
    }
}
// true (and the attribute set) when the grouped pass serves this grid
static bool rows_g_ready(const vh_sht& s) {
    if (2 * s.L + 1 > s.nphi || (s.nphi & 1)) return false;
    switch (s.nphi) {
        case 256: { const size_t l = rows_g_lds<256>(s); if (l > 160 * 1024) return false; rows_g_limit<256>(l); return true; }
        case 384: { const size_t l = rows_g_lds<384>(s); if (l > 160 * 1024) return false; rows_g_limit<384>(l); return true; }
        default: return false;
    }
}
void vh_t3_rows_prepare(const vh_sht& s) { (void)rows_g_ready(s); }

void vh_t3_rows(hipStream_t st, const vh_sht& s, double2* Fd, int nb, int mode, const double* rate, int bpe,
                double dt, double* grid, const double2* Fin) {
    if (nb <= 0) return;
    if ((mode == 0 || mode == 3) && rows_g_ready(s)) {
        const int ng = (s.nlat + VH_ROWS_RJ - 1) / VH_ROWS_RJ;
        const dim3 g(xcd_grid(nb * ng * 3));
        const double2* in = Fin ? Fin : Fd;
        if (s.nphi == 256)
            k_t3_rows_g<256, VH_ROWS_RJ><<<g, 64 * VH_ROWS_RJ, rows_g_lds<256>(s), st>>>(s, in, Fd, nb, mode, rate, bpe, dt);
        else
            k_t3_rows_g<384, VH_ROWS_RJ><<<g, 64 * VH_ROWS_RJ, rows_g_lds<384>(s), st>>>(s, in, Fd, nb, mode, rate, bpe, dt);
        HIPCHECK(hipGetLastError());
        return;
    }
    const bool twl = rows_lds(s, true) <= (size_t)64 * 1024;   // within the default dynamic LDS limit
    k_t3_rows<<<xcd_grid(nb * s.nlat), 192, rows_lds(s, twl), st>>>(s, Fin ? Fin : Fd, Fd, nb, mode, rate, bpe, dt,
                                                                   grid, twl ? 1 : 0);
    HIPCHECK(hipGetLastError());
}

void vh_t3_gather_trap(hipStream_t st, const vh_dims& d, const vh_dev& v, const int* e3em, int ne3d,
                       const double* tau, const double* sigma_n, const double* dN0, const double* sigma,
                       const double* dN1, double2* X, int nch, const double* taut, const double* dWt0,
                       const double* dWt1) {
    if (ne3d <= 0) return;
    if (nch == 6 && !(taut && dWt0 && dWt1)) throw vh_error("tensor gather: toroidal channels without their memory");
    dim3 g((d.nk + 255) / 256, ne3d);
    k_t3_gather_trap<<<g, 256, 0, st>>>(d.nr, d.nk, d.nlm, v, e3em, tau, sigma_n, dN0, sigma, dN1, X, nch, taut,
                                        dWt0, dWt1);
    HIPCHECK(hipGetLastError());
}

void vh_t3_gather(hipStream_t st, const vh_dims& d, const vh_dev& v, const int* e3em, int ne3d, const double* tau,
                  const double* sigma, const double* dN, double2* X, int nch, const double* taut, const double* dWt) {
    if (ne3d <= 0) return;
    if (nch == 6 && !(taut && dWt)) throw vh_error("tensor gather: toroidal channels without their memory");
    dim3 g((d.nk + 255) / 256, ne3d);
    k_t3_gather<<<g, 256, 0, st>>>(d.nr, d.nk, d.nlm, v, e3em, tau, sigma, dN, X, nch, taut, dWt);
    HIPCHECK(hipGetLastError());
}

// n6(l), l = 2..L: the raw spin-2 adjoint of the unit sectoral (l, m = l)
// spin-2 synthesis (tt = Sg, pp = -Sg, tp = 4 Sh), one batch holding every
// sectoral unit at once (each order only sees its own).
void vh_tsh_calibrate_n6(hipStream_t st, const vh_sht& s, double* n6_dev) {
    const int L = s.L, nlm = s.nlm;
    std::vector<double2> X((size_t)4 * nlm, make_double2(0.0, 0.0));
    for (int l = 2; l <= L; l++) X[(size_t)3 * nlm + l * (l + 1) / 2 + l] = make_double2(1.0, 0.0);
    double2 *dX = nullptr, *dF = nullptr, *dC = nullptr;
    double* dg = nullptr;
    HIPCHECK(hipMalloc((void**)&dX, X.size() * sizeof(double2)));
    HIPCHECK(hipMalloc((void**)&dF, (size_t)6 * (L + 1) * s.nlat * sizeof(double2)));
    HIPCHECK(hipMalloc((void**)&dC, X.size() * sizeof(double2)));
    HIPCHECK(hipMalloc((void**)&dg, (size_t)6 * s.nlat * s.nphi * sizeof(double)));
    HIPCHECK(hipMemcpyAsync(dX, X.data(), X.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    // through the grid, as the transforms are used: synthesis -> dyads -> analysis
    vh_tsh_synth(st, s, dX, 1, dF);
    vh_t3_rows(st, s, dF, 1, 1, nullptr, 1, 0.0, dg);
    vh_t3_rows(st, s, dF, 1, 2, nullptr, 1, 0.0, dg);
    vh_tsh_analysis(st, s, dF, 1, nullptr, dC);
    std::vector<double2> C(X.size());
    HIPCHECK(hipMemcpyAsync(C.data(), dC, C.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    (void)hipFree(dX); (void)hipFree(dF); (void)hipFree(dC); (void)hipFree(dg);
    std::vector<double> n6(L + 1, 0.0);
    for (int l = 2; l <= L; l++) n6[l] = C[(size_t)3 * nlm + l * (l + 1) / 2 + l].x;
    HIPCHECK(hipMemcpyAsync(n6_dev, n6.data(), (L + 1) * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
}

// n4(l), l = 2..L: as n6 for the toroidal spin-2 channel (unit sectoral c4
// synthesises tt = -Sh4, pp = Sh4, tp = Sg4; its raw adjoint at (l, l)).
void vh_tsh_calibrate_n4(hipStream_t st, const vh_sht& s, double* n4_dev) {
    const int L = s.L, nlm = s.nlm;
    std::vector<double2> X((size_t)6 * nlm, make_double2(0.0, 0.0));
    for (int l = 2; l <= L; l++) X[(size_t)5 * nlm + l * (l + 1) / 2 + l] = make_double2(1.0, 0.0);
    double2 *dX = nullptr, *dF = nullptr, *dC = nullptr;
    double* dg = nullptr;
    HIPCHECK(hipMalloc((void**)&dX, X.size() * sizeof(double2)));
    HIPCHECK(hipMalloc((void**)&dF, (size_t)6 * (L + 1) * s.nlat * sizeof(double2)));
    HIPCHECK(hipMalloc((void**)&dC, X.size() * sizeof(double2)));
    HIPCHECK(hipMalloc((void**)&dg, (size_t)6 * s.nlat * s.nphi * sizeof(double)));
    HIPCHECK(hipMemcpyAsync(dX, X.data(), X.size() * sizeof(double2), hipMemcpyHostToDevice, st));
    vh_tsh_synth(st, s, dX, 1, dF, 6);
    vh_t3_rows(st, s, dF, 1, 1, nullptr, 1, 0.0, dg);
    vh_t3_rows(st, s, dF, 1, 2, nullptr, 1, 0.0, dg);
    vh_tsh_analysis(st, s, dF, 1, nullptr, dC, 6, nullptr);
    std::vector<double2> C(X.size());
    HIPCHECK(hipMemcpyAsync(C.data(), dC, C.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    (void)hipFree(dX); (void)hipFree(dF); (void)hipFree(dC); (void)hipFree(dg);
    std::vector<double> n4(L + 1, 0.0);
    for (int l = 2; l <= L; l++) n4[l] = C[(size_t)5 * nlm + l * (l + 1) / 2 + l].x;
    HIPCHECK(hipMemcpyAsync(n4_dev, n4.data(), (L + 1) * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
}
