// vh_remap.h -- host lon-lat grid <-> Gauss grid maps (vh_remap.hip).
#pragma once
#include <string>
#include <hip/hip_runtime.h>

struct vh_remap;
// Build both maps for host cell-centre axes lon_deg[nlon], lat_deg[nlat_ll]
// (degrees, ascending) against a Gauss grid (colatitudes north-first, Gauss
// weights w); throws vh_error on a host grid it cannot map.
vh_remap* vh_remap_build(hipStream_t st, int nphi, int nlat, const double* colat, const double* w, int nlon,
                         int nlat_ll, const double* lon_deg, const double* lat_deg);
void vh_remap_to_gauss_host(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass);
void vh_remap_to_ll_host(vh_remap* r, const double* f_g, double* f_ll);
void vh_remap_to_gauss_dev(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass);
void vh_remap_to_ll_dev(vh_remap* r, const double* f_g, double* f_ll);
void vh_remap_free(vh_remap* r);
// the allocation-free remap_to_gauss: the device memory it uses beyond the
// map's tables is one fixed array of vh_remap_scratch_doubles() values per
// GPU, set up once per device by vh_remap_scratch_setup (under vh_remap_create)
void vh_remap_scratch_setup();
size_t vh_remap_scratch_doubles();
void vh_remap_to_gauss_noalloc(vh_remap* r, const double* f_ll, double* f_g, int conserve_mass);
