"""Lateral-viscosity (3-D) path on the GPU vs the oracle, through the C-ABI.

Tensor transforms (tensor_sh_synth/analysis, src/vilma_tensor_sh.f90:168-322),
the 3-D memory advance (src/vilma_response.f90:1171-1319) and whole coupling
calls of the 3-D viscosity workload at oracle-sized lmax; the laterally
uniform field forced through the 3-D path reproduces the 1-D run
(tests/test_response_3d.f90:224-246 of the reference states the same
property at 1e-9).
"""
import numpy as np
import pytest

from conftest import rel
from fastearth3d_amd import configs

pytestmark = pytest.mark.gpu

SEC_PER_YEAR = 31556926.0


def _lm_index(L, l, m):
    return m * (2 * L + 1 - m) // 2 + l


def _random_tensor(L, nlm, seed):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((4, nlm)) + 1j * rng.standard_normal((4, nlm))
    for mm in range(L + 1):
        for l in range(mm, L + 1):
            i = _lm_index(L, l, mm)
            if mm == 0:
                c[:, i] = c[:, i].real
            if l < 2:
                c[3, i] = 0.0
            if l < 1:
                c[1:, i] = 0.0
    return c


def _pair(oracle, idx, lmax, **over):
    from fastearth3d_amd import lib
    c = configs.config(idx, lmax=lmax)
    p = c["params"]
    for k, v in over.items():
        setattr(p, k, v)
    mo = oracle.OracleModel(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    mg = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    return c, mo, mg


@pytest.mark.parametrize("L", [16, 48, 64, 96, 192])
def test_tensor_transforms(oracle, L):
    """Tensor synthesis / analysis against the oracle; nphi 64 and 192 take the run-time
    longitude FFT, 256 (lmax 64) the compile-time in-place one. lmax 192 (386 latitudes)
    runs the synthesis over two chunks of north latitudes and the analysis with four
    row tiles per wave."""
    c, mo, mg = _pair(oracle, 1, L)
    nlm = mo.dims["nlm"]
    x = _random_tensor(L, nlm, 3)
    do = mo.tensor_synth(x)
    dg = mg.tensor_synth(x)
    for p in range(6):
        assert rel(dg[p], do[p]) <= 1e-11, p
    rng = np.random.default_rng(4)
    d = rng.standard_normal(do.shape)
    co = mo.tensor_analysis(d)
    cg = mg.tensor_analysis(d)
    for lam in range(4):
        assert rel(cg[lam], co[lam]) <= 1e-11, lam
    # round trip of a band-limited tensor field
    assert rel(mg.tensor_analysis(dg), x) <= 1e-10


@pytest.mark.parametrize("L", [16, 64, 192])
def test_advance_3d(oracle, L):
    """One 3-D advance from a random memory state (config-2 shaped pert); at lmax 64 the row
    pass runs the compile-time in-place FFT (nphi 256); lmax 192 is the largest 3-D member
    of the ensemble workload."""
    c, mo, mg = _pair(oracle, 2, L)
    for m in (mo, mg):
        m.enable_visc_3d(c["pert"])
    assert mg.dims["ne3d"] == mo.dims["ne3d"] > 0
    rng = np.random.default_rng(8)
    tau = rng.standard_normal(mo.memory().shape) * 1.0e3
    r, lay = oracle.mesh(c["params"])
    p = c["params"]
    maxwell = np.array([p.rheology[lay[e]] == 1 for e in range(len(lay))])
    tau[:, :, ~maxwell, :] = 0.0
    tau[:, :2, :, 3] = 0.0
    mo.set_memory(tau)
    mg.set_memory(tau)
    mo.begin_step()
    mg.begin_step()
    nlm = mo.dims["nlm"]
    sig = rng.standard_normal(nlm) + 1j * rng.standard_normal(nlm)
    sig[: p.lmax + 1] = sig[: p.lmax + 1].real
    for m in (mo, mg):
        m.set_dt(20.0 * SEC_PER_YEAR)
        m.advance(sig)
    assert rel(mg.memory(), mo.memory()) <= 1e-10


def test_3d_uniform_equals_1d():
    """Every Maxwell element forced 3-D under a uniform field == the 1-D run."""
    from fastearth3d_amd import lib
    c = configs.config(0, lmax=16)
    a = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    p3 = configs.config(0, lmax=16)["params"]
    p3.visc3d_tol = -1.0
    b = lib.SolidEarth(p3, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    b.enable_visc_3d(np.zeros((b.dims["ne"], b.nlat, b.nphi)))
    assert b.dims["ne3d"] == b.dims["nem"]
    for k in range(4):
        a.update(c["ice"](k), c["dt_yr"])
        b.update(c["ice"](k), c["dt_yr"])
    assert rel(b.memory(), a.memory()) <= 1e-9
    assert rel(b.spectrum(1), a.spectrum(1)) <= 1e-9


def test_config2_run(oracle):
    """3-D viscosity + SLE + rotation at lmax 24: several coupling calls."""
    c, mo, mg = _pair(oracle, 2, 24)
    for m in (mo, mg):
        m.enable_visc_3d(c["pert"])
    assert mg.dims["ne3d"] == mo.dims["ne3d"]
    for k in range(4):
        h = c["ice"](k + 60)
        mo.update(h, c["dt_yr"])
        mg.update(h, c["dt_yr"])
        assert np.array_equal(mg.C, mo.C)
        assert rel(mg.rsl, mo.rsl) <= 1e-8
    assert mg.scalar(3) == mo.scalar(3)              # same number of accepted sub-steps
    assert rel(mg.memory(), mo.memory()) <= 1e-8
    assert rel(mg.sigma_lm, mo.sigma_lm) <= 1e-8
    assert mg.worst_mass_resid <= 1e-10


def test_config2_full_size(oracle):
    """The benchmarked workload itself: lmax 128 on the 192 x 384 Gauss grid, 64 elements
    (36 genuinely 3-D), sea level with migrating coastlines and rotation, six 50-yr coupling
    calls (30 FE sub-steps) from the same state; the GPU takes its compile-time FFTs here.
    Per call: the same accepted sub-steps and sea-level iterations, the mask bit-exact,
    rsl and sigma within 1e-10, the memory within 1e-9 (measured: 1e-13 and 1e-11)."""
    c, mo, mg = _pair(oracle, 2, 128)
    for m in (mo, mg):
        m.enable_visc_3d(c["pert"])
    assert mg.dims["ne3d"] == mo.dims["ne3d"] > 0
    for k in range(6):
        h = c["ice"](60 + k)
        mo.update(h, c["dt_yr"])
        mg.update(h, c["dt_yr"])
        assert mg.scalar(3) == mo.scalar(3)
        assert mg.scalar(6) == mo.scalar(6)   # sea-level inner iterations over the run
        assert np.array_equal(mg.C, mo.C)
        assert rel(mg.rsl, mo.rsl) <= 1e-10
        assert rel(mg.sigma_lm, mo.sigma_lm) <= 1e-10
        assert rel(mg.memory(), mo.memory()) <= 1e-9
    assert mg.worst_mass_resid <= 1e-10


def test_enable_3d_nodes(oracle):
    """solid_earth_enable_visc_3d from nodal log10 viscosity (lid cap)."""
    c, mo, mg = _pair(oracle, 2, 12)
    nr = mo.dims["nr"]
    rng = np.random.default_rng(12)
    visc = 21.0 + 0.3 * rng.standard_normal((nr, mo.nlat, mo.nphi))
    for m in (mo, mg):
        m.enable_visc_3d_nodes(visc, lid_depth=80.0e3, lid_log10max=22.0, log10_cap=30.0)
    assert mg.dims["ne3d"] == mo.dims["ne3d"]
    for k in range(2):
        mo.update(c["ice"](k + 60), c["dt_yr"])
        mg.update(c["ice"](k + 60), c["dt_yr"])
    assert rel(mg.memory(), mo.memory()) <= 1e-8
