"""Restart pieces beyond the fields (src/vilma_io.f90:414-568), through the
C-ABI:

* the trapezoidal scheme continues bit for bit across a restart that carries
  the memory, the fields, and the step state (the controller's dt_try and the
  start-of-step load sigma_n);
* a cross-resolution restart: the memory of an lmax 16 run placed into an
  lmax 24 model (the low-degree block of the degree-grouped slots, the rest
  zero), then a zero-length seeding call and a coupling call, against the
  oracle given the same memory padded independently here; downsampling and a
  slot count that is not a whole number of degrees are refused with the
  reference's messages.
"""
import numpy as np
import pytest

from conftest import rel
from fastearth3d_amd import configs

pytestmark = pytest.mark.gpu


def _gpu(c, **over):
    from fastearth3d_amd import lib
    p = c["params"]
    for k, v in over.items():
        setattr(p, k, v)
    return lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])


def test_trap_restart_continuation():
    c = configs.config(1, lmax=16)
    ref = _gpu(c, scheme=b"trap")
    for k in range(4):
        ref.update(c["ice"](k + 60), c["dt_yr"])
    a = _gpu(configs.config(1, lmax=16), scheme=b"trap")
    for k in range(2):
        a.update(c["ice"](k + 60), c["dt_yr"])
    st, mem, step = a.state(), a.memory(), a.step_state()
    assert step["sigma_primed"] and step["dt_try"] > 0.0
    a.finalize()
    b = _gpu(configs.config(1, lmax=16), scheme=b"trap")
    b.set_memory(mem)
    b.set_state(st)
    b.set_step_state(step)
    for k in range(2, 4):
        b.update(c["ice"](k + 60), c["dt_yr"])
    np.testing.assert_array_equal(b.rsl, ref.rsl)
    np.testing.assert_array_equal(b.memory(), ref.memory())
    np.testing.assert_array_equal(b.C, ref.C)
    for m in (ref, b):
        m.finalize()


def test_lowres_restart_matches_oracle(oracle):
    from fastearth3d_amd.model import VilmaError
    lo = configs.config(1, lmax=16)
    m16 = _gpu(lo)
    for k in range(3):
        m16.update(lo["ice"](k + 60), lo["dt_yr"])
    tau16 = m16.memory()
    m16.finalize()
    hi = configs.config(1, lmax=24)
    g = _gpu(hi)
    g.set_memory_lowres(tau16)
    tau = g.memory()
    nk16 = tau16.shape[1]
    assert nk16 == 16 * 19 // 2
    np.testing.assert_array_equal(tau[:, :nk16], tau16)
    assert not np.any(tau[:, nk16:])
    padded = np.zeros_like(tau)   # the reference's rule, restated here
    padded[:, :nk16] = tau16
    o = oracle.OracleModel(hi["params"], hi["z_bed_eq"], hi["h_ice_eq"], hi["h_ice_init"])
    o.set_memory(padded)
    for m in (g, o):
        m.update(hi["ice"](63), 0.0)            # seed the entering ice (zero-length call)
        m.update(hi["ice"](63), hi["dt_yr"])
    assert np.array_equal(g.C, o.C)
    assert rel(g.rsl, o.rsl) <= 1e-10
    assert rel(g.memory(), o.memory()) <= 1e-10
    with pytest.raises(VilmaError, match="downsampling"):
        m8 = _gpu(configs.config(1, lmax=8))
        m8.set_memory_lowres(tau16)
    with pytest.raises(VilmaError, match="incompatible"):
        g.set_memory_lowres(tau16[:, :nk16 - 1])
    m8.finalize()
    for m in (g, o):
        m.finalize()
