"""The allocation-free remap entry point vh_remap_to_gauss_noalloc against
vh_remap_to_gauss (the entry point it varies), through the C-ABI: the same
kernels, so the same bits; two calls in a row, the second with a smaller
host grid (the fixed array keeps the first call's contents, which must not
leak into the second), with and without the mass rescale; a grid too large
for the fixed array is refused with an error."""
import numpy as np
import pytest

from fastearth3d_amd import configs

pytestmark = pytest.mark.gpu


def _model(lmax):
    from fastearth3d_amd import lib
    c = configs.config(1, lmax=lmax)
    return lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"])


def _axes(dlon, dlat):
    lon = np.arange(0.0, 360.0, dlon) + 0.5 * dlon
    lat = np.arange(-90.0, 90.0, dlat) + 0.5 * dlat
    return lon, lat


def test_noalloc_matches_remap_to_gauss():
    from fastearth3d_amd import lib
    m = _model(48)
    rng = np.random.default_rng(5)
    for dlon, dlat in ((1.0, 1.0), (2.5, 2.0)):     # the second call has the smaller host grid
        lon, lat = _axes(dlon, dlat)
        r = lib.Remap(m, lon, lat)
        f = 1000.0 + 300.0 * rng.standard_normal((lat.size, lon.size))
        for cm in (True, False):
            want = r.to_gauss(f, conserve_mass=cm)
            got = r.to_gauss_noalloc(f, conserve_mass=cm)
            np.testing.assert_array_equal(got, want)
        r.close()


def test_noalloc_refuses_grid_beyond_its_array():
    from fastearth3d_amd import lib
    from fastearth3d_amd.model import VilmaError
    m = _model(16)
    lon, lat = _axes(0.2, 0.2)                      # 1800 x 900 host cells > 2^20
    r = lib.Remap(m, lon, lat)
    with pytest.raises(VilmaError, match="exceed the fixed scratch"):
        r.to_gauss_noalloc(np.zeros((lat.size, lon.size)))
    r.close()
