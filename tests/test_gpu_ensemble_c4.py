"""The uneven ensemble through the vh_ensemble_* entry points.

A reduced ensemble of the uneven-ensemble workload: eight members of their
own earth models (lithosphere 62-115 km, upper / lower mantle viscosity
2e20-8e21 Pa s, 32-48 elements), lmax 64, 96, 128 and 192, four of them with
lateral viscosity, each with its own topography and ice domes, sea level and
rotation on, dt 100 yr. The viscosities keep the oracle's sub-step count small
(2-6 per call) so that it finishes in seconds per member.

* every member, stepped by the ensemble's worker pool, against the oracle
  after each of two calls: the ocean function bit for bit, rsl, z_bed, the
  load spectrum and the memory within 1e-10, the same sub-step count;
* the autotuner (vh_ensemble_tune) and vh_ensemble_run_device (no barrier
  between calls, ice histories in HBM) against the same members stepped one
  after another: bit for bit;
* a member that fails stops the call with its index and message.
"""
import numpy as np
import pytest

from conftest import rel
from fastearth3d_amd import configs

pytestmark = pytest.mark.gpu

MEMBERS = [  # (lmax, ne, 3-D, lithosphere km, eta_um, eta_lm)
    (64, 32, True, 70.0, 3e21, 5e21), (64, 40, False, 110.0, 2e20, 1e21), (96, 36, True, 90.0, 5e21, 2e21),
    (96, 48, False, 62.0, 4e20, 8e21), (128, 32, False, 100.0, 1e21, 3e21), (128, 34, True, 85.0, 6e21, 4e21),
    (192, 32, False, 75.0, 8e20, 2e21), (192, 32, True, 115.0, 8e21, 6e21)]
FIRST = 40   # ice of calls 40, 41 of the history (a fifth of the glacial maximum)
TOL = 1e-10


def _member(i):
    return configs.ensemble_member(i, members=MEMBERS)


def _gpu(c):
    from fastearth3d_amd import lib
    m = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        m.enable_visc_3d(c["pert"])
    return m


def _snap(m):
    return dict(C=m.C.copy(), rsl=m.rsl.copy(), z_bed=m.z_bed.copy(), sigma=m.sigma_lm.copy(),
                nsub=m.scalar(3), bsl=m.bsl)


@pytest.fixture(scope="module")
def ensemble_calls():
    """Two calls of the eight members through vh_ensemble_update (host ice,
    four workers); per member the state after each call and the final
    memory."""
    from fastearth3d_amd import lib
    cs = [_member(i) for i in range(len(MEMBERS))]
    ms = [_gpu(c) for c in cs]
    ens = lib.Ensemble(ms, n_workers=4)
    snaps = [[] for _ in ms]
    for k in range(2):
        ens.update([c["ice"](FIRST + k) for c in cs], [c["dt_yr"] for c in cs])
        for i, m in enumerate(ms):
            snaps[i].append(_snap(m))
    mem = [m.memory().copy() for m in ms]
    ens.close()
    for m in ms:
        m.finalize()
    return cs, snaps, mem


@pytest.mark.parametrize("i", range(len(MEMBERS)))
def test_ensemble_member_matches_oracle(oracle, ensemble_calls, i):
    cs, snaps, mem = ensemble_calls
    c = cs[i]
    mo = oracle.OracleModel(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        mo.enable_visc_3d(c["pert"])
    for k in range(2):
        mo.update(c["ice"](FIRST + k), c["dt_yr"])
        g = snaps[i][k]
        assert np.array_equal(g["C"], mo.C), (i, k)
        assert rel(g["rsl"], mo.rsl) <= TOL, (i, k)
        assert rel(g["z_bed"] - c["z_bed_eq"], mo.z_bed - c["z_bed_eq"]) <= TOL, (i, k)
        assert rel(g["sigma"], mo.sigma_lm) <= TOL, (i, k)
        assert g["nsub"] == mo.scalar(3), (i, k)
        assert abs(g["bsl"] - mo.bsl) <= TOL * max(1.0, abs(mo.bsl))
    assert rel(mem[i], mo.memory()) <= TOL
    mo.finalize()


def test_ensemble_run_equals_sequential():
    """Five calls of every member with the ice histories resident in HBM:
    the autotuner (three calls, one per sea-level launch shape) and then
    vh_ensemble_run_device (eight workers) against the members stepped one
    after another through vh_update_device; bit for bit."""
    from devmem import DeviceArray
    from fastearth3d_amd import lib
    cs = [_member(i) for i in range(len(MEMBERS))]
    nc = 5
    hist = [DeviceArray(np.stack([c["ice"](FIRST + k) for k in range(nc)])) for c in cs]
    seq = [_gpu(c) for c in cs]
    for m, h, c in zip(seq, hist, cs):
        ng = h.n // nc
        for k in range(nc):
            m.update_device(h.ptr(k * ng), c["dt_yr"])
    par = [_gpu(c) for c in cs]
    ens = lib.Ensemble(par, n_workers=8)
    chosen = ens.tune(1, [h.ptr() for h in hist], [c["dt_yr"] for c in cs])
    assert chosen in (0, 32, 64)
    ens.run_device(nc - 3, [h.ptr(3 * (h.n // nc)) for h in hist], [c["dt_yr"] for c in cs])
    ens.close()
    for h in hist:
        h.free()
    for a, b in zip(seq, par):
        np.testing.assert_array_equal(a.rsl, b.rsl)
        np.testing.assert_array_equal(a.C, b.C)
        np.testing.assert_array_equal(a.memory(), b.memory())
        assert a.scalar(0) == b.scalar(0) and a.scalar(3) == b.scalar(3)
    for m in seq + par:
        m.finalize()


def test_ensemble_failure_names_member(oracle):
    """A NaN in one member's Maxwell memory makes its guard collapse; the
    ensemble call fails naming that member and its message, the other
    members are unaffected."""
    from fastearth3d_amd import lib
    from fastearth3d_amd.model import VilmaError
    cs = [_member(i) for i in (1, 3, 4)]
    ms = [_gpu(c) for c in cs]
    ens = lib.Ensemble(ms, n_workers=3)
    ens.update([c["ice"](FIRST) for c in cs], 100.0)
    tau = ms[1].memory()
    r, lay = oracle.mesh(cs[1]["params"])
    e = [k for k in range(len(lay)) if cs[1]["params"].rheology[lay[k]] == 1][2]
    tau[0, 5, e, 1] = np.nan
    ms[1].set_memory(tau)
    with pytest.raises(VilmaError, match="ensemble member 1: .*explicit guard sub-step collapsed"):
        ens.update([c["ice"](FIRST + 1) for c in cs], 100.0)
    assert ens.failed_member() == 1
    assert ms[0].scalar(0) == ms[2].scalar(0) == 200.0   # the others made their call
    ens.close()
    for m in ms:
        m.finalize()


@pytest.mark.parametrize("blocks", [7, 32])
def test_sle_block_budget_bit_identical(blocks):
    """The sea-level solve on fewer blocks than work items (each block owning
    several, VH_T_SLE_BLOCKS) gives the one-block-per-item result bit for bit;
    a budget above the CU count is refused for a solve with more items."""
    from fastearth3d_amd import _abi
    from fastearth3d_amd.model import VilmaError
    c = _member(2)                          # lmax 96, 3-D: 97 + 49 items
    ref, got = _gpu(c), _gpu(c)
    got.set_tuning(_abi.T_SLE_BLOCKS, blocks)
    for k in range(2):
        for m in (ref, got):
            m.update(c["ice"](FIRST + k), c["dt_yr"])
    np.testing.assert_array_equal(got.rsl, ref.rsl)
    np.testing.assert_array_equal(got.C, ref.C)
    np.testing.assert_array_equal(got.memory(), ref.memory())
    assert got.scalar(6) == ref.scalar(6)
    big = _gpu(_member(7))                  # lmax 192: 386 items
    big.set_tuning(_abi.T_SLE_BLOCKS, 300)
    with pytest.raises(VilmaError, match="VH_T_SLE_BLOCKS above the CU count"):
        big.update(_member(7)["ice"](FIRST), 100.0)
    for m in (ref, got, big):
        m.finalize()
