"""CPU restatement of the host lon-lat <-> Gauss grid maps (src/vilma_remap.f90).

TEST INFRASTRUCTURE ONLY: the checker of the library's vh_remap_* entry
points. Only tests/ may import it.

The reference builds its maps with the fesm-utils `coords` library
(`map_init(..., method="con", gen="coords")` for host -> Gauss,
`method="bilinear", max_neighbors=8` for Gauss -> host,
src/vilma_remap.f90:89-97), which is not vendored in the reference tree
(the `fesm-utils:dev` dependency, no pinned version). Its published
description (src/vilma_remap.f90:3-4, 22-24: SCRIP-style conservative
weights, cell boundaries at the axis midpoints) is restated here for
lon-lat host grids, written independently of the device code as dense
overlap matrices:
  * conservative (remap_to_gauss, :101-138): the target value is the
    area-weighted mean of the source cells it overlaps, the spherical area of
    a lon-lat cell being dlon * d(sin lat); boundaries at the midpoints of
    neighbouring centres, latitude closed at the poles, longitude cyclic;
    conserve_mass rescales by (source integral in sr) / (Gauss quadrature
    integral), the source integral being sum(f area) / sum(area) * 4 pi
    (:124-134);
  * bilinear (remap_to_ll, :140-164): weights of the four Gauss neighbours,
    cyclic in longitude, the nearest Gauss row beyond the outermost latitude.
Parity unpinned against `coords` itself (absent): these functions are pinned
by the invariants in tests/test_remap.py (constants kept, exact area
conservation, the quadrature-integral rescale, exactness of the bilinear map
for fields linear in longitude or latitude, nested cells averaging to their
parent).
"""
import numpy as np


def lat_bounds(c):
    c = np.asarray(c, dtype=np.float64)
    b = np.empty(c.size + 1)
    b[0], b[-1] = -90.0, 90.0
    b[1:-1] = 0.5 * (c[:-1] + c[1:])
    return b


def lon_bounds(c):
    c = np.asarray(c, dtype=np.float64)
    b = np.empty(c.size + 1)
    b[1:-1] = 0.5 * (c[:-1] + c[1:])
    b[0] = 0.5 * (c[-1] - 360.0 + c[0])
    b[-1] = b[0] + 360.0
    return b


def gauss_axes(nlat, nphi):
    """Gauss latitudes north-first (degrees), weights, longitudes (degrees)."""
    x, w = np.polynomial.legendre.leggauss(nlat)   # x ascending = cos(colat) ascending: south first
    lat_nf = np.degrees(np.arcsin(x[::-1]))
    return lat_nf, w[::-1].copy(), 360.0 * np.arange(nphi) / nphi


def _overlap(tb, sb):
    """Overlap lengths of target intervals [tb[i], tb[i+1]) with source ones."""
    lo = np.maximum(tb[:-1, None], sb[None, :-1])
    hi = np.minimum(tb[1:, None], sb[None, 1:])
    return np.clip(hi - lo, 0.0, None)


def weights(lon, lat, nlat, nphi):
    """Dense overlap matrices: Wlat (nlat Gauss rows north-first, nlat_ll) in
    sin-latitude, Wlon (nphi, nlon) in degrees (cyclic)."""
    glat_nf, _, glon = gauss_axes(nlat, nphi)
    tb_lat = np.sin(np.radians(lat_bounds(glat_nf[::-1])))
    sb_lat = np.sin(np.radians(lat_bounds(lat)))
    wlat = _overlap(tb_lat, sb_lat)[::-1]          # rows north-first
    tb_lon, sb_lon = lon_bounds(glon), lon_bounds(lon)
    wlon = sum(_overlap(tb_lon, sb_lon + 360.0 * k) for k in (-1, 0, 1))
    return wlat, wlon


def to_gauss(f_ll, lon, lat, nlat, nphi, conserve_mass=False):
    """f_ll (nlat_ll, nlon), latitude ascending -> (nlat, nphi) north-first."""
    wlat, wlon = weights(lon, lat, nlat, nphi)
    g = (wlat @ f_ll @ wlon.T) / np.outer(wlat.sum(axis=1), wlon.sum(axis=1))
    if conserve_mass:
        area = cell_areas(lon, lat)
        src = np.sum(f_ll * area) / np.sum(area) * 4.0 * np.pi
        _, w, _ = gauss_axes(nlat, nphi)
        gint = np.sum(w[:, None] * g) * 2.0 * np.pi / nphi
        if abs(gint) > np.finfo(float).tiny:
            g = g * (src / gint)
    return g


def cell_areas(lon, lat):
    """Spherical areas (sr) of the host cells, (nlat_ll, nlon)."""
    dl = np.radians(np.diff(lon_bounds(lon)))
    ds = np.diff(np.sin(np.radians(lat_bounds(lat))))
    return ds[:, None] * dl[None, :]


def to_ll(g, lon, lat, nlat, nphi):
    """Bilinear: g (nlat, nphi) north-first -> (nlat_ll, nlon)."""
    glat_nf, _, _ = gauss_axes(nlat, nphi)
    glat = glat_nf[::-1]                             # ascending
    ga = g[::-1]
    out = np.empty((len(lat), len(lon)))
    for q, y in enumerate(lat):
        if y <= glat[0]:
            row = ga[0]
        elif y >= glat[-1]:
            row = ga[-1]
        else:
            ja = np.searchsorted(glat, y, side="right") - 1
            t = (y - glat[ja]) / (glat[ja + 1] - glat[ja])
            row = (1.0 - t) * ga[ja] + t * ga[ja + 1]
        for p, x in enumerate(lon):
            u = (x % 360.0) * nphi / 360.0
            i0 = int(np.floor(u)) % nphi
            t = u - np.floor(u)
            out[q, p] = (1.0 - t) * row[i0] + t * row[(i0 + 1) % nphi]
    return out
