"""The oracle's implicit (trapezoidal) memory scheme and its step-doubling
controller (scheme = "trap": src/vilma_response.f90:1321-1663,
src/vilma_timestep.f90:229-307, src/vilma_sle.f90:218-366), pinned by what
the mathematics fixes, CPU only:

* second order in time: with the controller held to fixed steps (tolerances
  out of reach, dt_max = h) the surface displacement of the load-only
  workload converges as h^2 (successive differences shrink 4x as h halves;
  a forward-Euler or lagged-strain rule would give 2x);
* the same solution as the forward-Euler path (pinned against the
  normal-mode Love numbers elsewhere): small steps of both agree, and steps
  beyond the explicit stability limit (M about 4.4) stay on the fine-step
  solution (A-stability);
* a laterally uniform field run through the 3-D (pseudo-spectral) trapezoid
  equals the 1-D spectral trapezoid;
* the sea-level <-> memory co-convergence takes at least two passes per
  solve (the convergence test needs a previous drift) and conserves mass;
* the controller's bookkeeping: three solves per attempt, rejections under a
  tight tolerance, more accepted steps for a tighter tolerance, and the
  reference's error on a non-finite estimate.
"""
import numpy as np
import pytest

from conftest import rel
from fastearth3d_amd import configs
from fastearth3d_amd.model import VilmaError

YR = 31556926.0


def _model(oracle, idx, lmax, scheme="trap", **over):
    c = configs.config(idx, lmax=lmax)
    p = c["params"]
    p.scheme = scheme.encode()
    p.skip_tol = 0.0
    for k, v in over.items():
        setattr(p, k, v)
    m = oracle.OracleModel(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    return c, m


def _load_only_u(oracle, h=None, span=400.0, ncalls=5, scheme="trap", **over):
    if scheme == "trap" and h is not None:
        over.update(rtol=1e30, atol=1e30, dt_max=h * YR)
    c, m = _model(oracle, 0, 16, scheme, **over)
    for k in range(ncalls):
        m.update(c["ice"](k), span)
    u = m.spectrum(1).copy()
    info = dict(acc=m.scalar(3), rej=m.scalar(4), solves=m.scalar(5), couple=m.scalar(12))
    m.finalize()
    return u, info


def test_trap_second_order(oracle):
    us = {h: _load_only_u(oracle, h)[0] for h in (400.0, 200.0, 100.0, 50.0)}
    for a, b, c in ((400.0, 200.0, 100.0), (200.0, 100.0, 50.0)):
        order = np.log2(np.abs(us[a] - us[b]).max() / np.abs(us[b] - us[c]).max())
        assert 1.9 <= order <= 2.1, (a, order)


def test_trap_matches_forward_euler(oracle):
    """Small steps: TRAP (h = 25 yr) and forward Euler (cfl 0.02, about 9 yr
    sub-steps) agree to the forward-Euler error. Steps of 2 kyr (fine
    sub-steps of 1 kyr: M = mu dt / eta about 4.4 in the upper mantle, past
    forward Euler's stability bound M < 2) over 10 kyr stay within 2e-3 of
    fine-step forward Euler (A-stability; the rule is not L-stable, so much
    larger M leaves stiff components ringing with |(1 - M/2) / (1 + M/2)| -> 1)."""
    ut, _ = _load_only_u(oracle, 25.0)
    uf, _ = _load_only_u(oracle, None, scheme="fe", cfl=0.02)
    assert rel(ut, uf) <= 3e-3
    ub, info = _load_only_u(oracle, 2000.0, span=2000.0, ncalls=5)
    assert info["acc"] == 5
    ue, _ = _load_only_u(oracle, None, span=2000.0, ncalls=5, scheme="fe", cfl=0.05)
    assert rel(ub, ue) <= 2e-3


def test_trap_uniform_3d_equals_1d(oracle):
    """Lateral field zero with the 3-D threshold below zero: every Maxwell
    element takes the pseudo-spectral trapezoid (tensor transforms), which
    must reproduce the 1-D spectral trapezoid."""
    c = configs.config(2, lmax=12)
    ref = None
    for tol in (1e-3, -1.0):
        p = configs.config(2, lmax=12)["params"]
        p.scheme, p.skip_tol, p.visc3d_tol = b"trap", 0.0, tol
        m = oracle.OracleModel(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
        m.enable_visc_3d(np.zeros_like(c["pert"]))
        assert (m.dims["ne3d"] > 0) == (tol < 0)
        for k in range(2):
            m.update(c["ice"](k + 60), 200.0)
        out = (m.rsl.copy(), m.memory().copy(), m.scalar(3))
        m.finalize()
        if ref is None:
            ref = out
    assert out[2] == ref[2]
    assert rel(out[0], ref[0]) <= 1e-9
    assert rel(out[1], ref[1]) <= 1e-9


def test_trap_sea_level_coconvergence(oracle):
    """Sea level with migrating coastlines and rotation: every solve runs at
    least two sea-level <-> memory passes, mass is conserved, and the result
    follows fine-step forward Euler (cfl 0.02, 34 sub-steps a call) within
    that run's own first-order error (3e-3; the default forward-Euler run,
    one sub-step a call, is 4.6 % away from both)."""
    c, m = _model(oracle, 1, 16)
    cf, mf = _model(oracle, 1, 16, scheme="fe", cfl=0.02)
    for k in range(3):
        for x in (m, mf):
            x.update(c["ice"](k + 60), c["dt_yr"])
    assert m.scalar(12) >= 2 * m.scalar(5)
    assert mf.scalar(12) == mf.scalar(5)      # forward Euler: one pass per solve
    assert m.worst_mass_resid <= 1e-10
    assert np.array_equal(m.C, mf.C)
    assert rel(m.rsl, mf.rsl) <= 3e-3
    m.finalize()
    mf.finalize()


def test_trap_controller(oracle):
    c, m = _model(oracle, 0, 16, rtol=1e-7, atol=1e-6)
    for k in range(3):
        m.update(c["ice"](k), 2000.0)
    acc, rej, sol = m.scalar(3), m.scalar(4), m.scalar(5)
    assert sol == 3 * (acc + rej)             # coarse + two fine solves per attempt
    assert rej > 0                            # the first attempt spans a whole call
    c2, m2 = _model(oracle, 0, 16, rtol=1e-3, atol=1e-2)
    for k in range(3):
        m2.update(c2["ice"](k), 2000.0)
    assert m2.scalar(3) < acc
    assert rel(m2.spectrum(1), m.spectrum(1)) <= 1e-2
    m.finalize()
    m2.finalize()


def test_trap_nonfinite_estimate(oracle):
    c, m = _model(oracle, 1, 8)
    m.update(c["ice"](50), c["dt_yr"])
    tau = m.memory()
    r, lay = oracle.mesh(c["params"])
    e = [k for k in range(len(lay)) if c["params"].rheology[lay[k]] == 1][2]
    tau[0, 5, e, 1] = np.nan
    m.set_memory(tau)
    with pytest.raises(VilmaError, match="implicit local-error estimate went non-finite"):
        m.update(c["ice"](51), c["dt_yr"])
    m.finalize()
