"""The toroidal degree of freedom (l_toroidal) on the GPU against the oracle
(src/vilma_radial_fe.f90:699-817 W operator, src/vilma_viscoelastic.f90:
373-439, :591-629 toroidal strain / forcing / advance,
src/vilma_tensor_sh.f90:168-387 six-channel tensor transforms,
src/vilma_response.f90:694-809 tor branch, :1055-1109 enabling):

* six-channel tensor transforms at lmax 10, 48 and 192 (the long-latitude
  and four-tile paths): synthesis and analysis within 1e-11 of the oracle's,
  the Z4 normalisation n4 within 1e-12, the round trip;
* the response probes of tests/test_toroidal.f90 (a held Y20 load over a
  sin^2 cos(2 phi) log-viscosity perturbation): the drifted uplift, the
  toroidal surface coefficients t_lm and the six-channel memory within 1e-10
  after every step, the reflection selection rule, and an axisymmetric
  perturbation driving no W;
* whole coupling calls of the lateral-viscosity workload at lmax 16 with the
  channels carried, forward Euler and the trapezoidal scheme: ocean function
  bit for bit, rsl, sigma, the memory and t_lm within 1e-10, equal step
  counts; twenty calls at lmax 12 within 1e-8 with equal sub-step and
  sea-level iteration counts;
* a restart of the six-channel memory continuing bit for bit, and the
  reverse sweep refusing the toroidal channels.
"""
import numpy as np
import pytest

from conftest import rel
from fastearth3d_amd import _abi, configs

pytestmark = pytest.mark.gpu
YR = 31556926.0
TOL = 1e-10


def _lm(L, l, m):
    return m * (2 * L + 1 - m) // 2 + l


def _coeffs(L, nch, seed):
    rng = np.random.default_rng(seed)
    nlm = (L + 1) * (L + 2) // 2
    c = rng.standard_normal((nch, nlm)) + 1j * rng.standard_normal((nch, nlm))
    for mm in range(L + 1):
        for l in range(mm, L + 1):
            i = _lm(L, l, mm)
            if mm == 0:
                c[:, i] = c[:, i].real
            for lam, lmin in zip(range(nch), (0, 1, 1, 2, 1, 2)):
                if l < lmin:
                    c[lam, i] = 0.0
    return c


def _m3(oracle, lmax, nlat=0, nphi=0):
    p = oracle.params_defaults()
    p.earth = b"M3-L70-V01"
    p.lmax, p.nlat, p.nphi = lmax, nlat, nphi
    p.rotation = 0
    return p


@pytest.mark.parametrize("L,fac", [(10, 4), (48, 0), (192, 0)])
def test_tensor_six_channels(oracle, L, fac):
    from fastearth3d_amd import lib
    p = _m3(oracle, L, nlat=fac * L, nphi=fac * L)
    nphi, nlat = p.grid()
    z, hz = np.full((nlat, nphi), 1000.0), np.zeros((nlat, nphi))
    mo = oracle.OracleModel(p, z, hz)
    mg = lib.SolidEarth(p, z, hz)
    c = _coeffs(L, 6, 3)
    dg, do = mg.tensor_synth(c), mo.tensor_synth(c)
    assert rel(dg, do) <= 1e-11
    cg, co = mg.tensor_analysis(do, nch=6), mo.tensor_analysis(do, nch=6)
    assert rel(cg, co) <= 1e-10
    assert rel(cg, c) <= 1e-9
    n6, n4 = mg.tensor_norms()
    assert rel(n4, mo.tensor_n4()) <= 1e-12 and rel(n6, mo.tensor_n6()) <= 1e-12
    # four channels unchanged by the six-channel path
    assert rel(mg.tensor_analysis(do, nch=4), co[:4]) <= 1e-10
    for m in (mo, mg):
        m.finalize()


def _tor_pair(oracle, delta, axisym, L=12):
    from fastearth3d_amd import lib
    p = _m3(oracle, L, nlat=3 * L, nphi=3 * L)
    p.l_visc_3d = 1
    p.l_toroidal = 1
    nphi, nlat = p.grid()
    z, hz = np.full((nlat, nphi), 1000.0), np.zeros((nlat, nphi))
    mo = oracle.OracleModel(p, z, hz)
    mg = lib.SolidEarth(p, z, hz)
    colat, lon, _ = mo.grid_geometry()
    r, mu, rh = oracle.mesh_props(_m3(oracle, L))
    st, ct = np.sin(colat)[:, None], np.cos(colat)[:, None]
    pat = delta * ct * ct * np.ones((1, nphi)) if axisym else delta * st * st * np.cos(2.0 * lon)[None, :]
    pert = np.zeros((mo.dims["ne"], nlat, nphi))
    pert[rh == 1] = pat
    for m in (mo, mg):
        m.set_dt(40.0 * YR)
        m.enable_visc_3d(pert)
    return mo, mg


def test_toroidal_response_matches_oracle(oracle):
    L = 12
    mo, mg = _tor_pair(oracle, 0.2, False, L)
    assert mg.dims["nlam"] == mo.dims["nlam"] == 6
    sig = np.zeros(mg.dims["nlm"], complex)
    sig[_lm(L, 2, 0)] = 1.0e3
    for step in range(40):
        for m in (mo, mg):
            m.begin_step()
        ug, _ = mg.apply(sig)
        uo, _ = mo.apply(sig)
        assert rel(ug, uo) <= TOL, step
        if step:
            assert rel(mg.spectrum(_abi.SP_TOR), mo.spectrum(_abi.SP_TOR)) <= TOL, step
        for m in (mo, mg):
            m.advance(sig)
        assert rel(mg.memory(), mo.memory()) <= TOL, step
    for m in (mo, mg):
        m.begin_step()
    u, _ = mg.apply(sig)
    t = mg.spectrum(_abi.SP_TOR)
    assert rel(t, mo.spectrum(_abi.SP_TOR)) <= TOL
    wmax, umax = np.abs(t).max(), np.abs(u).max()
    assert wmax >= 1e-6 * umax
    err = 0.0   # reflection selection rule: U even / real, W odd / imaginary, even m only
    for l in range(1, L + 1):
        for mm in range(l + 1):
            i = _lm(L, l, mm)
            if mm % 2:
                err = max(err, abs(u[i]) / umax, abs(t[i]) / wmax)
            else:
                err = max(err, abs(u[i].imag) / umax, abs(t[i].real) / wmax)
    assert err <= 1e-10
    for m in (mo, mg):
        m.finalize()


def test_toroidal_axisymmetric_null_on_gpu(oracle):
    mo, mg = _tor_pair(oracle, 0.2, True)
    L = 12
    sig = np.zeros(mg.dims["nlm"], complex)
    sig[_lm(L, 2, 0)] = 1.0e3
    for _ in range(20):
        mg.begin_step()
        mg.apply(sig)
        mg.advance(sig)
    mg.begin_step()
    u, _ = mg.apply(sig)
    assert np.abs(mg.spectrum(_abi.SP_TOR)).max() <= 1e-12 * np.abs(u).max()
    for m in (mo, mg):
        m.finalize()


def _run_pair(oracle, scheme, lmax=16):
    from fastearth3d_amd import lib
    c = configs.config(2, lmax=lmax)
    p = c["params"]
    p.l_toroidal = 1
    p.scheme = scheme
    mo = oracle.OracleModel(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    mg = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    for m in (mo, mg):
        m.enable_visc_3d(c["pert"])
    assert mg.dims["nlam"] == mo.dims["nlam"] == 6
    return c, mo, mg


@pytest.mark.parametrize("scheme", [b"fe", b"trap"])
def test_toroidal_model_run_matches_oracle(oracle, scheme):
    c, mo, mg = _run_pair(oracle, scheme)
    for k in range(3):
        for m in (mo, mg):
            m.update(c["ice"](k + 60), c["dt_yr"])
        assert np.array_equal(mg.C, mo.C), k
        assert rel(mg.rsl, mo.rsl) <= TOL, k
        assert rel(mg.sigma_lm, mo.sigma_lm) <= TOL, k
        assert rel(mg.spectrum(_abi.SP_TOR), mo.spectrum(_abi.SP_TOR)) <= TOL, k
        assert tuple(mg.scalar(i) for i in (3, 4, 5)) == tuple(mo.scalar(i) for i in (3, 4, 5)), k
    assert rel(mg.memory(), mo.memory()) <= TOL
    assert np.abs(mg.memory()[:, :, :, 4:]).max() > 0.0   # the toroidal channels are excited
    for m in (mo, mg):
        m.finalize()


def test_toroidal_restart_continuation():
    from fastearth3d_amd import lib
    c = configs.config(2, lmax=16)
    p = c["params"]
    p.l_toroidal = 1

    def model():
        m = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
        m.enable_visc_3d(c["pert"])
        return m

    a = model()
    for k in range(2):
        a.update(c["ice"](k + 60), c["dt_yr"])
    tau, st = a.memory(), a.state()
    a.update(c["ice"](62), c["dt_yr"])
    b = model()
    b.set_memory(tau)
    b.set_state(st)
    b.update(c["ice"](62), c["dt_yr"])
    assert np.array_equal(a.rsl, b.rsl)
    assert np.array_equal(a.memory(), b.memory())
    assert np.array_equal(a.spectrum(_abi.SP_TOR), b.spectrum(_abi.SP_TOR))
    with pytest.raises(lib.VilmaError, match="toroidal"):
        b.forward_record([c["ice"](63)], c["dt_yr"])
    for m in (a, b):
        m.finalize()


def test_toroidal_long_run_matches_oracle(oracle):
    """Twenty coupling calls with the channels carried (lmax 12): rsl, sigma,
    t_lm and the memory within 1e-8, the mask bit for bit and the same
    sub-step counts after every call."""
    c, mo, mg = _run_pair(oracle, b"fe", lmax=12)
    for k in range(20):
        for m in (mo, mg):
            m.update(c["ice"](k + 60), c["dt_yr"])
        assert np.array_equal(mg.C, mo.C), k
        assert rel(mg.rsl, mo.rsl) <= 1e-8, k
        assert rel(mg.sigma_lm, mo.sigma_lm) <= 1e-8, k
        assert rel(mg.spectrum(_abi.SP_TOR), mo.spectrum(_abi.SP_TOR)) <= 1e-8, k
        assert tuple(mg.scalar(i) for i in (3, 4, 5, 6)) == tuple(mo.scalar(i) for i in (3, 4, 5, 6)), k
    assert rel(mg.memory(), mo.memory()) <= 1e-8
    for m in (mo, mg):
        m.finalize()
