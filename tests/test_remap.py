"""The host-grid <-> Gauss-grid maps (src/vilma_remap.f90): the oracle pinned by
invariants on the CPU; the library's vh_remap_* entry points against the
oracle on the GPU. The reference's own maps come from the absent fesm-utils
`coords` library, so these tests pin the restated algorithm (conservative:
area-weighted mean over overlapping cells, midpoint boundaries; bilinear:
four Gauss neighbours), not `coords` bit for bit."""
import numpy as np
import pytest

import remap as R

NLAT, NPHI = 66, 128          # the lmax 32 default Gauss grid


def _grid(dlon, dlat, lon0=0.0):
    lon = lon0 + dlon * (np.arange(int(round(360.0 / dlon))) + 0.5)
    lat = -90.0 + dlat * (np.arange(int(round(180.0 / dlat))) + 0.5)
    return lon, lat


@pytest.mark.parametrize("dlon,dlat,lon0", [(1.5, 1.5, 0.0), (5.0, 4.0, -180.0), (2.8125, 2.8125, 0.3)])
def test_conservative_invariants(dlon, dlat, lon0):
    lon, lat = _grid(dlon, dlat, lon0)
    rng = np.random.default_rng(0)
    f = rng.standard_normal((lat.size, lon.size))
    # constants are kept
    g1 = R.to_gauss(np.full_like(f, 3.25), lon, lat, NLAT, NPHI)
    assert np.abs(g1 - 3.25).max() <= 1e-13
    # the area integral is conserved with the Gauss cells' own areas
    g = R.to_gauss(f, lon, lat, NLAT, NPHI)
    glat_nf, _, glon = R.gauss_axes(NLAT, NPHI)
    gareas = R.cell_areas(glon, glat_nf[::-1])[::-1]
    src = np.sum(f * R.cell_areas(lon, lat))
    assert abs(np.sum(g * gareas) - src) <= 1e-12 * np.sum(np.abs(f) * R.cell_areas(lon, lat))
    # conserve_mass: the Gauss quadrature integral equals the source integral
    gm = R.to_gauss(f + 2.0, lon, lat, NLAT, NPHI, conserve_mass=True)
    _, w, _ = R.gauss_axes(NLAT, NPHI)
    assert abs(np.sum(w[:, None] * gm) * 2.0 * np.pi / NPHI - np.sum((f + 2.0) * R.cell_areas(lon, lat))) <= 1e-12 * 4 * np.pi


def test_nested_cells_average_to_their_parent():
    """A host grid whose cells tile each Gauss cell exactly (4 x 3 per cell):
    the conservative map is the area-weighted mean of the 12 children."""
    glat_nf, _, glon = R.gauss_axes(NLAT, NPHI)
    tb_lon = R.lon_bounds(glon)
    tb_lat = R.lat_bounds(glat_nf[::-1])
    lonb = np.concatenate([np.linspace(tb_lon[i], tb_lon[i + 1], 5)[:-1] for i in range(NPHI)] + [[tb_lon[-1]]])
    latb = np.concatenate([np.linspace(tb_lat[j], tb_lat[j + 1], 4)[:-1] for j in range(NLAT)] + [[tb_lat[-1]]])
    # centres whose midpoints give back exactly these boundaries: solve the midpoint recursion
    lon = np.empty(lonb.size - 1)
    lon[0] = 0.5 * (lonb[0] + lonb[1])
    for k in range(1, lon.size):
        lon[k] = 2.0 * lonb[k] - lon[k - 1]
    lat = np.empty(latb.size - 1)
    lat[0] = 0.5 * (latb[0] + latb[1])
    for k in range(1, lat.size):
        lat[k] = 2.0 * latb[k] - lat[k - 1]
    if not (np.allclose(R.lon_bounds(lon)[1:-1], lonb[1:-1]) and np.allclose(R.lat_bounds(lat)[1:-1], latb[1:-1])):
        pytest.skip("the midpoint recursion does not reproduce these boundaries")
    rng = np.random.default_rng(1)
    f = rng.standard_normal((lat.size, lon.size))
    a = R.cell_areas(lon, lat)
    g = R.to_gauss(f, lon, lat, NLAT, NPHI)[::-1]          # ascending rows
    par = (f * a).reshape(NLAT, 3, NPHI, 4).sum(axis=(1, 3)) / a.reshape(NLAT, 3, NPHI, 4).sum(axis=(1, 3))
    assert np.abs(g - par).max() <= 1e-12


def test_bilinear_invariants():
    lon, lat = _grid(1.0, 1.0, 0.25)
    glat_nf, _, glon = R.gauss_axes(NLAT, NPHI)
    # constant
    assert np.abs(R.to_ll(np.full((NLAT, NPHI), -1.5), lon, lat, NLAT, NPHI) + 1.5).max() <= 1e-14
    # linear in latitude between the Gauss rows
    g = np.repeat(glat_nf[:, None], NPHI, axis=1)
    f = R.to_ll(g, lon, lat, NLAT, NPHI)
    inside = (lat > glat_nf.min()) & (lat < glat_nf.max())
    assert np.abs(f[inside] - lat[inside, None]).max() <= 1e-12
    # beyond the outermost rows: the nearest row
    assert np.abs(f[~inside] - np.where(lat[~inside] > 0, glat_nf[0], glat_nf[-1])[:, None]).max() <= 1e-12
    # linear in longitude away from the cyclic seam
    g = np.repeat(glon[None, :], NLAT, axis=0)
    f = R.to_ll(g, lon, lat, NLAT, NPHI)
    away = lon < glon[-1]
    assert np.abs(f[:, away] - lon[None, away]).max() <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("dlon,dlat,lon0", [(1.5, 1.5, 0.0), (5.0, 4.0, -180.0)])
def test_gpu_remap_matches_oracle(dlon, dlat, lon0):
    from fastearth3d_amd import configs, lib
    c = configs.config(1, lmax=32)
    m = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"])
    assert (m.nlat, m.nphi) == (NLAT, NPHI)
    lon, lat = _grid(dlon, dlat, lon0)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((lat.size, lon.size)) + 1.0
    rm = m.remap(lon, lat)
    for cm in (False, True):
        go = R.to_gauss(f, lon, lat, NLAT, NPHI, conserve_mass=cm)
        gg = rm.to_gauss(f, conserve_mass=cm)
        assert np.abs(gg - go).max() <= 1e-12 * np.abs(go).max()
    g = rng.standard_normal((NLAT, NPHI))
    fo = R.to_ll(g, lon, lat, NLAT, NPHI)
    fg = rm.to_ll(g)
    assert np.abs(fg - fo).max() <= 1e-13 * np.abs(fo).max()
    # the coupling's round trip (ice in conservatively, rsl out bilinearly) keeps a constant
    assert np.abs(rm.to_ll(rm.to_gauss(np.full_like(f, 7.0))) - 7.0).max() <= 1e-12
    rm.close()
    m.finalize()


@pytest.mark.gpu
def test_gpu_remap_errors():
    from fastearth3d_amd import configs, lib
    from fastearth3d_amd.model import VilmaError
    c = configs.config(1, lmax=16)
    m = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"])
    lon, lat = _grid(5.0, 5.0)
    with pytest.raises(VilmaError, match="ascending"):
        m.remap(lon[::-1], lat)
    with pytest.raises(VilmaError, match="ascending"):
        m.remap(lon, lat[::-1])
    with pytest.raises(VilmaError, match="360"):
        m.remap(np.linspace(0.0, 400.0, 50), lat)
    m.finalize()
