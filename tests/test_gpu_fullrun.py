"""Whole runs of the benchmark configurations 0, 1 and 2 on the GPU against
the oracle's trajectories (tests/golden/fullrun_c{0,1,2}.npz, written by
tests/golden/fullrun.py with the same run protocol): every coupling call of
the run (200 calls at lmax 64; 350 and 400 calls at lmax 128 on the 192 x 384
grid), compared after each call: the same accepted and rejected sub-steps,
sea-level solves, inner iterations and coastline passes, the ocean mask bit
for bit (CRC-32 per call, the packed mask at four checkpoints), model time,
bsl, polar motion and max |rsl| within 1e-8; at the checkpoints sigma and
the six memory maxima within 1e-8, and the rsl grid at half time and at the
end within 1e-8 (the parity contract's bound after a full run).

Also, through the library, the guard and frame pins of the oracle tests
(test_oracle_timedomain.py): NaN memory rejected 40 times with rsl restored
and the reference's collapse error; the rejections of a too-large cfl equal
the oracle's; the degree-1 drift's F(a) = 0 in the cm frame; the closed-form
sub-step count.
"""
import os
import sys
import zlib

import numpy as np
import pytest

from conftest import rel
from fastearth3d_amd import configs

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import fullrun  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-8
SEC_PER_YEAR = 31556926.0


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_full_run_matches_oracle(idx):
    from fastearth3d_amd import lib
    if not os.path.exists(fullrun.fixture_path(idx)):
        pytest.skip("fixture %s not generated" % fullrun.fixture_path(idx))
    ref = fullrun.load(idx)
    cfg = configs.config(idx)
    p = cfg["params"]
    m = lib.SolidEarth(p, cfg["z_bed_eq"], cfg["h_ice_eq"], cfg["h_ice_init"])
    cols = {c: i for i, c in enumerate(fullrun.COLS)}
    exact = ("n_accept", "n_reject", "n_solve", "n_inner", "n_outer", "mask_crc")
    close = ("time", "bsl", "m_re", "m_im", "rsl_max")
    seen = {}

    def check(k, model):
        row = ref["rows"][k]
        for c in exact:
            got = model.scalar(fullrun.SCALARS[c]) if c in fullrun.SCALARS else None
            if c == "mask_crc":
                got = float(zlib.crc32(fullrun.mask_bits(model.C).tobytes()))
            assert got == row[cols[c]], (idx, k, c, got, row[cols[c]])
        for c in close:
            if c == "rsl_max":
                got = float(np.abs(model.rsl).max())
            else:
                got = model.scalar(fullrun.SCALARS[c])
            want = row[cols[c]]
            scale = max(abs(want), np.abs(ref["rows"][:, cols[c]]).max() * 1e-6, 1e-300)
            assert abs(got - want) <= TOL * scale, (idx, k, c, got, want)
        seen[k] = True

    rec = fullrun.run_record(m, cfg, on_call=check)
    assert len(seen) == cfg["steps"]
    for n, s in ref["sigma"].items():
        assert rel(rec["sigma"][n], s) <= TOL, (idx, n)
        assert np.array_equal(rec["mask"][n], ref["mask"][n]), (idx, n)
        t = ref["taumax"][n]
        for a in range(6):
            if t[a] > 0:
                assert abs(rec["taumax"][n][a] - t[a]) <= TOL * t[a], (idx, n, a)
            else:
                assert rec["taumax"][n][a] == 0.0
    for n, r in ref["rsl"].items():
        assert rel(rec["rsl"][n], r) <= TOL, (idx, n)
    assert m.worst_mass_resid <= 1e-10 or idx == 0


def _pair(oracle, idx, lmax, **over):
    from fastearth3d_amd import lib
    c = configs.config(idx, lmax=lmax)
    p = c["params"]
    for k, v in over.items():
        setattr(p, k, v)
    mo = oracle.OracleModel(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    mg = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    return c, mo, mg


def test_guard_rejects_nan_memory(oracle):
    """A NaN in the memory: 40 rejections with rsl restored bit for bit, then
    the reference's collapse error, the same on the GPU as on the oracle."""
    from fastearth3d_amd.model import VilmaError
    c, mo, mg = _pair(oracle, 1, 12)
    for m in (mo, mg):
        m.update(c["ice"](50), c["dt_yr"])
    assert rel(mg.rsl, mo.rsl) <= 1e-10
    rsl0 = mg.rsl
    tau = mg.memory()
    r, lay = oracle.mesh(c["params"])
    e = [k for k in range(len(lay)) if c["params"].rheology[lay[k]] == 1][2]   # a Maxwell element
    tau[0, 5, e, 1] = np.nan
    for m in (mo, mg):
        m.set_memory(tau)
        with pytest.raises(VilmaError, match="explicit guard sub-step collapsed"):
            m.update(c["ice"](51), c["dt_yr"])
    assert mg.scalar(4) == mo.scalar(4) == 40
    assert mg.scalar(3) == mo.scalar(3)
    assert np.array_equal(mg.rsl, rsl0)


def test_guard_rollback_matches_oracle(oracle):
    """cfl = 3000 with 300-kyr calls: rejections in every call; the GPU makes
    the same accept / reject sequence counts and the same rsl as the oracle."""
    c, mo, mg = _pair(oracle, 1, 12, cfl=3000.0)
    for m in (mo, mg):
        m.update(c["ice"](40), 1.0)
    for k in range(3):
        for m in (mo, mg):
            m.update(c["ice"](41 + k), 3.0e5)
        assert mg.scalar(4) == mo.scalar(4) and mg.scalar(3) == mo.scalar(3), k
        assert mg.scalar(4) > 0
        assert np.array_equal(mg.C, mo.C)
        assert rel(mg.rsl, mo.rsl) <= 1e-8, k


def test_deg1_cm_drift(oracle):
    """The GPU's degree-1 drift in the cm frame has F(a) = 0, like the oracle."""
    c, mo, mg = _pair(oracle, 1, 16)
    rng = np.random.default_rng(31)
    tau = rng.standard_normal(mo.memory().shape) * 1.0e4
    r, lay = oracle.mesh(c["params"])
    p = c["params"]
    maxwell = np.array([p.rheology[lay[e]] == 1 for e in range(len(lay))])
    tau[:, :, ~maxwell, :] = 0.0
    tau[:, :2, :, 3] = 0.0
    for m in (mo, mg):
        m.set_memory(tau)
        m.begin_step()
    sg, _ = mg.drift()
    so, _ = mo.drift()
    assert np.abs(sg[1, :2]).max() <= 1e-12 * np.abs(sg[1]).max()
    assert rel(sg[0, :2], so[0, :2]) <= 1e-10


@pytest.mark.parametrize("idx", [0, 2])
def test_nsub_closed_form(idx):
    """Accepted sub-steps per call = max(1, ceil(span max(mu/eta 10^-pert) /
    cfl - 1e-9)) (src/vilma_timestep.f90:177-186)."""
    from fastearth3d_amd import lib
    from test_oracle_timedomain import _rate_max
    c = configs.config(idx, lmax=16)
    p = c["params"]
    m = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        m.enable_visc_3d(c["pert"])
    for dt_yr in (c["dt_yr"], 7.0 * c["dt_yr"]):
        n0 = m.scalar(3)
        m.update(c["ice"](3), dt_yr)
        expect = max(1, int(np.ceil(dt_yr * SEC_PER_YEAR * _rate_max(p, c["pert"]) / p.cfl - 1e-9)))
        assert m.scalar(3) - n0 == expect
