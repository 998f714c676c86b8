"""The benchmark workloads' shapes and generators (CPU)."""
import numpy as np

from fastearth3d_amd import configs


def test_ensemble_members():
    m = configs.ensemble_members(64)
    assert len(m) == 64
    assert set(x[0] for x in m) == set(configs.ENSEMBLE_LMAX)
    assert all(32 <= x[1] <= 64 for x in m)
    assert sum(x[2] for x in m) == 16
    assert set(x[0] for x in m if x[2]) == set(configs.ENSEMBLE_LMAX)   # 3-D members at every lmax, 192 too
    assert all(60.0 <= x[3] <= 120.0 for x in m)
    assert all(1e20 <= x[4] <= 1e22 and 1e20 <= x[5] <= 1e22 for x in m)
    assert len(set(x[3] for x in m)) == 64 and len(set(x[4] for x in m)) == 64   # every member its own earth
    assert configs.ensemble_members(64) == m                                      # seeded
    assert configs.ensemble_members(64, seed=1) != m


def test_ensemble_member_workload():
    mem = configs.ensemble_members(64)
    i = next(k for k, x in enumerate(mem) if x[0] == 96 and not x[2])
    c = configs.ensemble_member(i, 64)
    p = c["params"]
    L, ne, _, litho, eu, el = mem[i]
    assert p.lmax == 96 and p.rotation == 1 and p.l_visc_3d == 0 and c["dt_yr"] == 100.0 and c["steps"] == 200
    nphi, nlat = p.grid()
    assert c["z_bed_eq"].shape == (nlat, nphi) and c["ice"](0).shape == (nlat, nphi)
    assert sum(p.n_elem_layer[i] for i in range(p.n_layer)) == ne
    assert abs((p.r_earth - p.r_bot[0]) / 1e3 - litho) < 1e-9 and p.r_top[1] == p.r_bot[0]
    assert p.eta[1] == p.eta[2] == eu and p.eta[3] == el and p.rheology[0] == 0
    # 70 % ocean, +-5 km, the member's own topography and ice
    assert abs((c["z_bed_eq"] < 0).mean() - 0.70) < 0.01
    assert abs(np.abs(c["z_bed_eq"]).max() - 5000.0) < 1e-6
    j = next(k for k, x in enumerate(mem) if x[0] == 96 and k != i)
    c2 = configs.ensemble_member(j, 64)
    assert not np.array_equal(c2["z_bed_eq"], c["z_bed_eq"])
    assert not np.array_equal(c2["ice"](100), c["ice"](100))
    k3 = next(k for k, x in enumerate(mem) if x[2])
    c3 = configs.ensemble_member(k3, 64)
    assert c3["params"].l_visc_3d == 1 and c3["pert"].shape == (mem[k3][1],) + c3["z_bed_eq"].shape


def test_fast_random_field_equals_direct_sum():
    """random_sh_field_fast (per-order Legendre sums, then one product over
    longitudes) draws the same coefficients as the direct summation of
    random_sh_field and so gives the same field, on a Gauss grid and on an
    odd grid."""
    for nlat, nphi, L in ((24, 48, 11), (17, 40, 16)):
        colat, lon = configs.gauss_grid(nlat, nphi)
        a = configs.random_sh_field(colat, lon, L, -2.0, 5)
        b = configs.random_sh_field_fast(colat, lon, L, -2.0, 5)
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
