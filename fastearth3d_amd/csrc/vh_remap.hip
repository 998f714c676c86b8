// vh_remap.hip -- host lon-lat grid <-> model Gauss grid (src/vilma_remap.f90).
//
// Conservative in (host -> Gauss, remap_to_gauss, :101-138) and bilinear out
// (Gauss -> host, remap_to_ll, :140-164). The reference builds both maps with
// the fesm-utils `coords` library (SCRIP-style weights from cell boundaries at
// the axis midpoints). For two lon-lat grids the conservative weights are
// separable: the spherical area of a cell is dlon * d(sin lat), so the overlap
// of a target and a source cell is (longitude overlap) x (sin-latitude
// overlap), and the area-weighted mean over the overlapping source cells is
//   f_t(j, i) = sum_q sum_p Wlat[j][q] Wlon[i][p] f_s(q, p),
// with each row of Wlat / Wlon normalised to 1 (full coverage: the target
// cell area). A target cell overlaps a contiguous run of source cells in each
// direction (cyclic in longitude), stored as (first index, count, weights).
// Cell boundaries: midpoints between neighbouring centres; latitude runs from
// -90 to +90 at the ends, longitude is cyclic with period 360 degrees.
// Bilinear: a host point between two Gauss longitudes (cyclic) and two Gauss
// latitudes takes the bilinear weights of its four neighbours; beyond the
// outermost Gauss latitude the nearest row is taken (the reference leaves such
// choices to `coords`).
#include <cfloat>
#include <mutex>
#include <cmath>
#include <mutex>
#include <vector>

#include "vh_internal.h"
#include "vh_remap.h"

namespace {

// cell boundaries of an ascending axis (degrees): latitude clamped to the poles
std::vector<double> lat_bounds(const std::vector<double>& c) {
    const size_t n = c.size();
    std::vector<double> b(n + 1);
    b[0] = -90.0;
    b[n] = 90.0;
    for (size_t k = 1; k < n; k++) b[k] = 0.5 * (c[k - 1] + c[k]);
    return b;
}
std::vector<double> lon_bounds(const std::vector<double>& c) {
    const size_t n = c.size();
    std::vector<double> b(n + 1);
    for (size_t k = 1; k < n; k++) b[k] = 0.5 * (c[k - 1] + c[k]);
    b[0] = 0.5 * (c[n - 1] - 360.0 + c[0]);
    b[n] = b[0] + 360.0;
    return b;
}

// overlap runs: for each target interval [tb[i], tb[i+1]) the source
// intervals it meets, measured by g(bound) (sin of latitude, or longitude in
// degrees), normalised to sum 1. cyclic: the source axis repeats every 360.
struct runs {
    std::vector<int> first, count;
    std::vector<double> w;   // [target][kmax]
    int kmax = 0;
};

runs make_runs(const std::vector<double>& tb, const std::vector<double>& sb, bool cyclic) {
    const int nt = (int)tb.size() - 1, ns = (int)sb.size() - 1;
    std::vector<std::vector<std::pair<int, double>>> hit(nt);
    auto meas = [&](double x) { return cyclic ? x : std::sin(x * VH_PI / 180.0); };
    for (int i = 0; i < nt; i++) {
        const double a = tb[i], b = tb[i + 1];
        if (!cyclic) {
            for (int q = 0; q < ns; q++) {
                const double lo = std::max(a, sb[q]), hi = std::min(b, sb[q + 1]);
                if (hi > lo) hit[i].push_back({q, meas(hi) - meas(lo)});
            }
        } else {
            // shift the source axis by whole turns so that it covers [a, b)
            for (int turn = -1; turn <= 1; turn++) {
                const double sh = 360.0 * turn;
                for (int q = 0; q < ns; q++) {
                    const double lo = std::max(a, sb[q] + sh), hi = std::min(b, sb[q + 1] + sh);
                    if (hi > lo) hit[i].push_back({q, hi - lo});
                }
            }
        }
    }
    runs r;
    r.first.resize(nt);
    r.count.resize(nt);
    for (int i = 0; i < nt; i++) {
        if (hit[i].empty()) throw vh_error("remap: a Gauss cell overlaps no host cell (host axes must cover the sphere)");
        r.kmax = std::max(r.kmax, (int)hit[i].size());
    }
    r.w.assign((size_t)nt * r.kmax, 0.0);
    for (int i = 0; i < nt; i++) {
        double tot = 0.0;
        for (auto& h : hit[i]) tot += h.second;
        r.first[i] = hit[i][0].first;
        r.count[i] = (int)hit[i].size();
        for (size_t k = 0; k < hit[i].size(); k++) {
            // runs are contiguous (cyclic): entry k is source first + k
            if ((hit[i][0].first + (int)k) % ns != hit[i][k].first)
                throw vh_error("remap: host axis cells must be ordered and contiguous");
            r.w[(size_t)i * r.kmax + k] = hit[i][k].second / tot;
        }
    }
    return r;
}

template <class T>
T* dup(const std::vector<T>& h, hipStream_t st) {
    T* d = nullptr;
    HIPCHECK(hipMalloc((void**)&d, std::max<size_t>(h.size(), 1) * sizeof(T)));
    if (!h.empty()) HIPCHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return d;
}

}   // namespace

// host (nlon, nlat_ll) -> Gauss (nphi, nlat), one thread per Gauss point;
// Gauss row j is north-first, its latitude runs index the ascending host rows
__global__ void __launch_bounds__(256) k_remap_to_gauss(int nphi, int nlat, int nlon, int kla, int klo,
                                                        const int* __restrict__ lat_first,
                                                        const int* __restrict__ lat_count,
                                                        const double* __restrict__ lat_w,
                                                        const int* __restrict__ lon_first,
                                                        const int* __restrict__ lon_count,
                                                        const double* __restrict__ lon_w,
                                                        const double* __restrict__ f, double* __restrict__ g) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nphi * nlat) return;
    const int i = t % nphi, j = t / nphi;
    const int q0 = lat_first[j], nq = lat_count[j], p0 = lon_first[i], np_ = lon_count[i];
    double acc = 0.0;
    for (int a = 0; a < nq; a++) {
        const double* row = f + (size_t)(q0 + a) * nlon;
        double s = 0.0;
        for (int b = 0; b < np_; b++) {
            int p = p0 + b;
            if (p >= nlon) p -= nlon;
            s += lon_w[(size_t)i * klo + b] * row[p];
        }
        acc += lat_w[(size_t)j * kla + a] * s;
    }
    g[t] = acc;
}

// Gauss (nphi, nlat) -> host (nlon, nlat_ll), one thread per host point
__global__ void __launch_bounds__(256) k_remap_to_ll(int nphi, int nlon, int nlat_ll, const int* __restrict__ li,
                                                     const double* __restrict__ lt, const int* __restrict__ pj,
                                                     const double* __restrict__ pt, const double* __restrict__ g,
                                                     double* __restrict__ f) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlon * nlat_ll) return;
    const int p = t % nlon, q = t / nlon;
    const int i0 = li[p], i1 = i0 + 1 == nphi ? 0 : i0 + 1;
    const int j0 = pj[2 * q], j1 = pj[2 * q + 1];
    const double tx = lt[p], ty = pt[q];
    const double a = (1.0 - tx) * g[(size_t)j0 * nphi + i0] + tx * g[(size_t)j0 * nphi + i1];
    const double b = (1.0 - tx) * g[(size_t)j1 * nphi + i0] + tx * g[(size_t)j1 * nphi + i1];
    f[t] = (1.0 - ty) * a + ty * b;
}

// sum of w[k] * x[k] over n values (one block; the remap's mass integrals)
__global__ void __launch_bounds__(256) k_remap_dot(int n, const double* __restrict__ w, const double* __restrict__ x,
                                                   int wmod, double* out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s += w[wmod > 0 ? k / wmod : k] * x[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// g *= src / gauss when |gauss| > tiny (remap_to_gauss conserve_mass, :124-134)
__global__ void __launch_bounds__(256) k_remap_scale(int n, const double* __restrict__ tot, double* g) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double src = tot[0], gauss = tot[1];
    if (fabs(gauss) > DBL_MIN) g[t] *= src / gauss;
}

struct vh_remap {
    int nlon = 0, nlat_ll = 0, nphi = 0, nlat = 0, kla = 0, klo = 0;
    hipStream_t st = nullptr;
    int *lat_first = nullptr, *lat_count = nullptr, *lon_first = nullptr, *lon_count = nullptr;
    double *lat_w = nullptr, *lon_w = nullptr;
    int *li = nullptr, *pj = nullptr;
    double *lt = nullptr, *pt = nullptr;
    double *src_wt = nullptr;   // [nlat_ll][nlon] host cell areas scaled to a 4 pi total
    double *gw = nullptr;       // [nlat] Gauss weight * 2 pi / nphi
    double *f_ll = nullptr, *f_g = nullptr, *tot = nullptr;
};

vh_remap* vh_remap_build(hipStream_t st, int nphi, int nlat, const double* colat, const double* gw_host, int nlon,
                         int nlat_ll, const double* lon_deg, const double* lat_deg) {
    if (nlon < 2 || nlat_ll < 2) throw vh_error("remap: the host grid needs at least 2 x 2 cells");
    std::vector<double> lon(lon_deg, lon_deg + nlon), lat(lat_deg, lat_deg + nlat_ll);
    for (int p = 1; p < nlon; p++)
        if (!(lon[p] > lon[p - 1])) throw vh_error("remap: host longitudes must be strictly ascending");
    if (lon[nlon - 1] - lon[0] >= 360.0) throw vh_error("remap: host longitudes must span less than 360 degrees");
    for (int q = 1; q < nlat_ll; q++)
        if (!(lat[q] > lat[q - 1])) throw vh_error("remap: host latitudes must be strictly ascending");
    if (lat[0] < -90.0 || lat[nlat_ll - 1] > 90.0) throw vh_error("remap: host latitudes must lie in [-90, 90]");
    // Gauss axes: longitudes 360 i / nphi; latitudes ascending (south first) for the runs
    std::vector<double> glon(nphi), glat_asc(nlat);
    for (int i = 0; i < nphi; i++) glon[i] = 360.0 * i / nphi;
    for (int j = 0; j < nlat; j++) glat_asc[j] = 90.0 - colat[nlat - 1 - j] * 180.0 / VH_PI;
    runs rlat = make_runs(lat_bounds(glat_asc), lat_bounds(lat), false);
    runs rlon = make_runs(lon_bounds(glon), lon_bounds(lon), true);
    // Gauss rows north-first: row j takes the ascending run nlat-1-j
    std::vector<int> lf(nlat), lc(nlat);
    std::vector<double> lw((size_t)nlat * rlat.kmax);
    for (int j = 0; j < nlat; j++) {
        const int ja = nlat - 1 - j;
        lf[j] = rlat.first[ja];
        lc[j] = rlat.count[ja];
        for (int k = 0; k < rlat.kmax; k++) lw[(size_t)j * rlat.kmax + k] = rlat.w[(size_t)ja * rlat.kmax + k];
    }
    // bilinear: host longitude between Gauss columns i0, i0 + 1 (cyclic)
    std::vector<int> li(nlon);
    std::vector<double> lt(nlon);
    for (int p = 0; p < nlon; p++) {
        double x = std::fmod(lon[p], 360.0);
        if (x < 0.0) x += 360.0;
        const double u = x * nphi / 360.0;
        int i0 = (int)std::floor(u);
        double t = u - i0;
        if (i0 >= nphi) { i0 -= nphi; }
        li[p] = i0;
        lt[p] = t;
    }
    // host latitude between Gauss rows (north-first indices j0, j1), nearest row beyond the outermost
    std::vector<int> pj(2 * (size_t)nlat_ll);
    std::vector<double> pt(nlat_ll);
    for (int q = 0; q < nlat_ll; q++) {
        const double y = lat[q];
        if (y <= glat_asc[0]) { pj[2 * q] = nlat - 1; pj[2 * q + 1] = nlat - 1; pt[q] = 0.0; continue; }
        if (y >= glat_asc[nlat - 1]) { pj[2 * q] = 0; pj[2 * q + 1] = 0; pt[q] = 0.0; continue; }
        int ja = 0;   // ascending index with glat_asc[ja] <= y < glat_asc[ja + 1]
        while (ja + 1 < nlat && glat_asc[ja + 1] <= y) ja++;
        pj[2 * q] = nlat - 1 - ja;             // north-first row of the lower latitude
        pj[2 * q + 1] = nlat - 1 - (ja + 1);   // and of the upper one
        pt[q] = (y - glat_asc[ja]) / (glat_asc[ja + 1] - glat_asc[ja]);
    }
    // host cell areas (steradians) scaled by 4 pi / their total: the conserved
    // source integral is sum(f area) / sum(area) * 4 pi (remap_to_gauss, :126-128)
    std::vector<double> lb = lat_bounds(lat), ob = lon_bounds(lon);
    std::vector<double> area((size_t)nlon * nlat_ll);
    double atot = 0.0;
    for (int q = 0; q < nlat_ll; q++)
        for (int p = 0; p < nlon; p++) {
            area[(size_t)q * nlon + p] = (ob[p + 1] - ob[p]) * VH_PI / 180.0 *
                                         (std::sin(lb[q + 1] * VH_PI / 180.0) - std::sin(lb[q] * VH_PI / 180.0));
            atot += area[(size_t)q * nlon + p];
        }
    for (double& a : area) a *= 4.0 * VH_PI / atot;
    std::vector<double> gw(nlat);
    for (int j = 0; j < nlat; j++) gw[j] = gw_host[j] * 2.0 * VH_PI / nphi;

    vh_remap* r = new vh_remap();
    r->nlon = nlon; r->nlat_ll = nlat_ll; r->nphi = nphi; r->nlat = nlat;
    r->kla = rlat.kmax; r->klo = rlon.kmax; r->st = st;
    r->lat_first = dup(lf, st); r->lat_count = dup(lc, st); r->lat_w = dup(lw, st);
    r->lon_first = dup(rlon.first, st); r->lon_count = dup(rlon.count, st); r->lon_w = dup(rlon.w, st);
    r->li = dup(li, st); r->lt = dup(lt, st); r->pj = dup(pj, st); r->pt = dup(pt, st);
    r->src_wt = dup(area, st);
    r->gw = dup(gw, st);
    HIPCHECK(hipMalloc((void**)&r->f_ll, (size_t)nlon * nlat_ll * sizeof(double)));
    HIPCHECK(hipMalloc((void**)&r->f_g, (size_t)nphi * nlat * sizeof(double)));
    HIPCHECK(hipMalloc((void**)&r->tot, 2 * sizeof(double)));
    HIPCHECK(hipStreamSynchronize(st));
    return r;
}

static inline unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// the conservative map, then (conserve_mass) the two mass totals into tot[0..1]
// and the rescale; f_ll, f_g and tot in device memory
static void remap_to_gauss_tot(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass, double* tot) {
    const size_t ng = (size_t)r->nphi * r->nlat, nh = (size_t)r->nlon * r->nlat_ll;
    k_remap_to_gauss<<<nblk(ng), 256, 0, r->st>>>(r->nphi, r->nlat, r->nlon, r->kla, r->klo, r->lat_first,
                                                   r->lat_count, r->lat_w, r->lon_first, r->lon_count, r->lon_w,
                                                   f_ll, f_g);
    HIPCHECK(hipGetLastError());
    if (conserve_mass) {
        // source integral in steradians (the host cells' areas sum to 4 pi) and
        // the Gauss quadrature integral of the result
        k_remap_dot<<<1, 256, 0, r->st>>>((int)nh, r->src_wt, f_ll, 0, tot);
        k_remap_dot<<<1, 256, 0, r->st>>>((int)ng, r->gw, f_g, r->nphi, tot + 1);
        k_remap_scale<<<nblk(ng), 256, 0, r->st>>>((int)ng, tot, f_g);
        HIPCHECK(hipGetLastError());
    }
}

void vh_remap_to_gauss_dev(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass) {
    remap_to_gauss_tot(r, f_ll, f_g, conserve_mass, r->tot);
}

// ---- the allocation-free remap_to_gauss (vh_remap_to_gauss_noalloc)
// Its device memory beyond the handle's tables is this file-scope array:
// the staged host field, the staged Gauss field and the two mass totals
// (nlon nlat_ll + nphi nlat + 2 doubles at most VH_REMAP_SCRATCH_DOUBLES).
// Each GPU holds its own copy for the life of the process; its address on a
// device comes from hipGetSymbolAddress once per device (vh_remap_scratch_setup,
// called by vh_remap_create). Calls on one GPU share it, so calls that overlap
// in time on one GPU race; the array keeps the last call's contents.
#define VH_REMAP_SCRATCH_DOUBLES (1 << 20)
#define VH_REMAP_MAX_DEVICES 64
static __device__ __align__(256) double scratch[VH_REMAP_SCRATCH_DOUBLES];
static double* g_scratch_ptr[VH_REMAP_MAX_DEVICES];
static std::mutex g_scratch_mu;

void vh_remap_scratch_setup() {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= VH_REMAP_MAX_DEVICES) throw vh_error("remap: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return;
    double* p = nullptr;
    HIPCHECK(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
}

size_t vh_remap_scratch_doubles() { return VH_REMAP_SCRATCH_DOUBLES; }

void vh_remap_to_gauss_noalloc(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass) {
    const size_t ng = (size_t)r->nphi * r->nlat, nh = (size_t)r->nlon * r->nlat_ll;
    if (nh + ng + 2 > (size_t)VH_REMAP_SCRATCH_DOUBLES)
        throw vh_error("remap_to_gauss_noalloc: host grid (" + std::to_string(nh) + ") plus Gauss grid (" +
                       std::to_string(ng) + ") points exceed the fixed scratch of " +
                       std::to_string(VH_REMAP_SCRATCH_DOUBLES - 2) + " values");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    double* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = (dev >= 0 && dev < VH_REMAP_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
    }
    if (!base) throw vh_error("remap_to_gauss_noalloc: no remap map was created on this device");
    double* s_ll = base;
    double* s_g = base + nh;
    double* tot = base + nh + ng;
    HIPCHECK(hipMemcpyAsync(s_ll, f_ll, nh * sizeof(double), hipMemcpyHostToDevice, r->st));
    remap_to_gauss_tot(r, s_ll, s_g, conserve_mass, tot);
    HIPCHECK(hipMemcpyAsync(f_g, s_g, ng * sizeof(double), hipMemcpyDeviceToHost, r->st));
    HIPCHECK(hipStreamSynchronize(r->st));
}

void vh_remap_to_ll_dev(vh_remap* r, const double* f_g, double* f_ll) {
    const size_t nh = (size_t)r->nlon * r->nlat_ll;
    k_remap_to_ll<<<nblk(nh), 256, 0, r->st>>>(r->nphi, r->nlon, r->nlat_ll, r->li, r->lt, r->pj, r->pt, f_g, f_ll);
    HIPCHECK(hipGetLastError());
}

void vh_remap_to_gauss_host(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass) {
    const size_t ng = (size_t)r->nphi * r->nlat, nh = (size_t)r->nlon * r->nlat_ll;
    HIPCHECK(hipMemcpyAsync(r->f_ll, f_ll, nh * sizeof(double), hipMemcpyHostToDevice, r->st));
    vh_remap_to_gauss_dev(r, r->f_ll, r->f_g, conserve_mass);
    HIPCHECK(hipMemcpyAsync(f_g, r->f_g, ng * sizeof(double), hipMemcpyDeviceToHost, r->st));
    HIPCHECK(hipStreamSynchronize(r->st));
}

void vh_remap_to_ll_host(vh_remap* r, const double* f_g, double* f_ll) {
    const size_t ng = (size_t)r->nphi * r->nlat, nh = (size_t)r->nlon * r->nlat_ll;
    HIPCHECK(hipMemcpyAsync(r->f_g, f_g, ng * sizeof(double), hipMemcpyHostToDevice, r->st));
    vh_remap_to_ll_dev(r, r->f_g, r->f_ll);
    HIPCHECK(hipMemcpyAsync(f_ll, r->f_ll, nh * sizeof(double), hipMemcpyDeviceToHost, r->st));
    HIPCHECK(hipStreamSynchronize(r->st));
}

void vh_remap_free(vh_remap* r) {
    if (!r) return;
    for (void* p : {(void*)r->lat_first, (void*)r->lat_count, (void*)r->lat_w, (void*)r->lon_first,
                    (void*)r->lon_count, (void*)r->lon_w, (void*)r->li, (void*)r->lt, (void*)r->pj, (void*)r->pt,
                    (void*)r->src_wt, (void*)r->gw, (void*)r->f_ll, (void*)r->f_g, (void*)r->tot})
        if (p) (void)hipFree(p);
    delete r;
}
