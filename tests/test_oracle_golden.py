"""The oracle pinned against the reference's own golden data and the
mathematics: band LU test matrices (tests/test_band.f90), P1 element
integrals (tests/test_integrals.f90), operator symmetry (tests/test_assembly.f90),
analytic Love limits (tests/test_love.f90), the M3-L70-V01 benchmark Love
table (tests/test_benchmark_love.f90, data/benchmarks/love_M3-L70-V01),
SHT conventions (tests/test_sht.f90), tensor-SH identities
(tests/test_tensor_sh.f90), flotation (tests/test_flotation.f90), SLE mass
conservation (tests/test_sle.f90:118), the 3-D == 1-D reduction
(tests/test_response_3d.f90:224-246) and the rotation channels against the
Spada et al. (2011) Test 3/2 polar-motion benchmark (tests/test_rotation.f90,
data/benchmarks/rotation_spada2011). CPU only."""
import json
import os

import numpy as np
import pytest
from scipy.special import lpmv, factorial

from conftest import rel
from fastearth3d_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ band LU
def test_band_zero_diagonal(oracle):
    A = np.array([[0, 2, 0, 0], [3, 1, 4, 0], [0, 5, 1, 6], [0, 0, 7, 2]], float)
    b = np.array([1, 2, 3, 4.0])
    x, _ = oracle.band_solve_dense(A, b)
    assert rel(x, np.linalg.solve(A, b)) <= 1e-12


def test_band_random_band(oracle):
    n, kl, ku = 50, 4, 3
    A = np.zeros((n, n))
    for j in range(1, n + 1):
        for i in range(max(1, j - ku), min(n, j + kl) + 1):
            A[i - 1, j - 1] = (7 * i + 3 * j) % 11 - 5
        A[j - 1, j - 1] += 2 * (kl + ku) + 5
    b = np.array([(13 * i) % 17 - 8 for i in range(1, n + 1)], float)
    x, klku = oracle.band_solve_dense(A, b)
    assert klku == (kl, ku)
    assert rel(x, np.linalg.solve(A, b)) <= 1e-12


def test_band_saddle(oracle):
    A = np.zeros((6, 6))
    for i in range(4):
        A[i, i] = 4.0
    for (i, j) in [(0, 1), (2, 3), (0, 4), (1, 4), (2, 5), (3, 5)]:
        A[i, j] = A[j, i] = 1.0
    b = np.arange(1, 7, dtype=float)
    x, _ = oracle.band_solve_dense(A, b)
    assert rel(x, np.linalg.solve(A, b)) <= 1e-12


# ------------------------------------------------------- element integrals
def test_element_integrals(oracle):
    rk, rk1 = 3.4e6, 3.6e6
    xg, wg = np.polynomial.legendre.leggauss(20)
    r = 0.5 * (rk1 - rk) * xg + 0.5 * (rk1 + rk)
    w = 0.5 * (rk1 - rk) * wg
    h = rk1 - rk
    psi = [(rk1 - r) / h, (r - rk) / h]
    dpsi = [-np.ones_like(r) / h, np.ones_like(r) / h]
    got = oracle.elem_integrals(rk, rk1)
    defs = [lambda a, b: dpsi[a] * dpsi[b] * r**2, lambda a, b: psi[a] * dpsi[b] * r**2,
            lambda a, b: psi[a] * dpsi[b] * r, lambda a, b: psi[a] * psi[b] * r**2,
            lambda a, b: psi[a] * psi[b] * r, lambda a, b: psi[a] * psi[b], lambda a, b: psi[a] * psi[b] / r]
    for q, f in enumerate(defs):
        for a in range(2):
            for b in range(2):
                ref = np.sum(w * f(a, b))
                assert abs(got[4 * q + 2 * a + b] - ref) <= 1e-11 * abs(ref), (q, a, b)
    for q, f in enumerate([lambda a: dpsi[a] * r**2, lambda a: psi[a] * r, lambda a: psi[a] * r**2]):
        for a in range(2):
            ref = np.sum(w * f(a))
            assert abs(got[28 + 2 * q + a] - ref) <= 1e-11 * abs(ref)


# -------------------------------------------------------------- operator
@pytest.mark.parametrize("j", [1, 2, 7])
def test_operator_symmetric(oracle, j):
    p = configs.five_layer_params(8, 40)
    A = oracle.dense_operator(p, j)
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()


def _homogeneous(mu):
    """Homogeneous incompressible sphere (tests/test_love.f90 solve_homogeneous)."""
    p = configs.five_layer_params(8, 40)
    p.earth = b"custom"
    p.n_layer = 1
    p.r_bot[0], p.r_top[0], p.rho[0], p.mu[0], p.eta[0], p.rheology[0] = 0.0, 6371e3, 5511.0, mu, 1e40, 0
    p.n_elem_layer[0] = 200
    return p


@pytest.mark.parametrize("j", [2, 3, 4, 5, 6])
def test_love_fluid_and_rigid_limits(oracle, j):
    h, l, k = oracle.love(_homogeneous(1.0), j)
    assert abs(h + (2 * j + 1) / 3.0) <= 1e-3
    assert abs(k + 1.0) <= 1e-3
    h, l, k = oracle.love(_homogeneous(1.0e18), j)
    assert max(abs(h), abs(l), abs(k)) <= 1e-4


def test_love_table_m3l70v01(oracle):
    """Elastic Love numbers at 1% for j = 2..48, fluid limit at 1% for j <= 8
    (tests/test_benchmark_love.f90:71-101), reference depth-target mesh."""
    tab = json.load(open(os.path.join(GOLDEN, "love_M3-L70-V01.json")))["degrees"]
    p = oracle.params_defaults()
    p.earth = b"M3-L70-V01"
    for j in list(range(2, 9)) + [16, 32, 48]:
        e = tab[str(j)]["elastic"]
        h, l, k = oracle.love(p, j)
        assert abs(h - e["h"]) <= 1e-2 * abs(e["h"]), (j, h, e)
        assert abs(k - e["k"]) <= 1e-2 * abs(e["k"]), (j, k, e)
        if j <= 8:
            f = tab[str(j)]["fluid"]
            hf, lf, kf = oracle.love(p, j, fluidize=True)
            for got, ref in ((hf, f["h"]), (lf, f["l"]), (kf, f["k"])):
                assert abs(got - ref) <= 1e-2 * abs(ref), (j, got, ref)


# -------------------------------------------------------------- rotation
def _spada_load(name, h, colat, lon):
    """Heaviside ice loads of Spada et al. (2011) Table 4, centroid colatitude
    25 deg, longitude 75 deg, half-width 10 deg, rho_ice 931: the parabolic cap
    sigma = rho h sqrt((cos g - cos a)/(1 - cos a)) and the uniform disc
    sigma = rho h (as tests/test_rotation.f90:129-155 builds them)."""
    deg = np.pi / 180.0
    thc, lac, al = 25.0 * deg, 75.0 * deg, 10.0 * deg
    cg = np.cos(thc) * np.cos(colat)[:, None] + np.sin(thc) * np.sin(colat)[:, None] * np.cos(lon[None, :] - lac)
    out = np.zeros((colat.size, lon.size))
    inside = cg >= np.cos(al)
    if name == "cap":
        out[inside] = 931.0 * h * np.sqrt(np.maximum((cg[inside] - np.cos(al)) / (1.0 - np.cos(al)), 0.0))
    else:
        out[inside] = 931.0 * h
    return out


@pytest.mark.parametrize("name,h", [("cap", 1500.0), ("disc", 1000.0)])
def test_rotation_spada2011(oracle, name, h):
    """Rotation channels (orc_rotation.c) vs the Spada et al. (2011) Test 3/2
    benchmark, the checks of tests/test_rotation.f90: rigid (2,1) inertia of
    the load / (C - A) vs the published G_sigma (2 %), 0 < k_T,elastic < k_s <
    1.5 and 0.5 < k_s, and |m(t)| under the Heaviside load vs Table 14 with the
    Chandler wobble excluded (the quasi-static column; 3 %) at t = 0, 1, 2, 5,
    10, 20 kyr; M3-L70-V01 on the reference depth-target mesh, lmax 128 on a
    256 x 512 Gauss grid, dt = 25 yr (Spada's year, 3.15576e7 s)."""
    ref = json.load(open(os.path.join(GOLDEN, "spada2011_test3_2.json")))
    p = oracle.params_defaults()
    p.earth = b"M3-L70-V01"
    lmax, nlat, nphi = 128, 256, 512
    x, _ = np.polynomial.legendre.leggauss(nlat)
    colat = np.arccos(x[::-1])                      # north first
    lon = 2.0 * np.pi * np.arange(nphi) / nphi
    yr = 3.15576e7
    dt = 25.0 * yr
    m, kTe, ks, I21 = oracle.rotation_run(p, lmax, nlat, nphi, _spada_load(name, h, colat, lon), dt, 800)
    assert 0.5 < ks < 1.5 and 0.0 < kTe < ks, (kTe, ks)
    G = complex(*ref["G_sigma"][name])
    assert abs(I21 / 2.63e35 - G) <= 0.02 * abs(G), (I21 / 2.63e35, G)
    mdeg = np.abs(m) * 180.0 / np.pi
    for row in ref["table"][name]["excl"]:
        i = int(round(row["t_kyr"] * 1000.0 * yr / dt))
        assert abs(mdeg[i] - row["m_Gs_deg"]) <= 0.03 * row["m_Gs_deg"], (row, mdeg[i])


# ------------------------------------------------------------------- SHT
@pytest.fixture(scope="module")
def small_model(oracle):
    c = configs.config(0, lmax=12)
    return oracle.OracleModel(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])


def _lm_index(L, l, m):
    return m * (2 * L + 1 - m) // 2 + l


def test_sht_closed_forms(small_model):
    m = small_model
    colat, lon, w = m.grid_geometry()
    L = m.dims["lmax"]
    assert abs(w.sum() - 2.0) <= 1e-14
    assert abs(m.surface_integral(np.ones((m.nlat, m.nphi))) - 4 * np.pi) <= 1e-12
    # Y10 = sqrt(3/4pi) cos(theta): its coefficient is 1 and synthesis reproduces it
    f = np.sqrt(3.0 / (4 * np.pi)) * np.cos(colat)[:, None] * np.ones(m.nphi)
    q = m.sht_analysis(f)
    e = np.zeros_like(q); e[_lm_index(L, 1, 0)] = 1.0
    assert np.abs(q - e).max() <= 1e-12
    assert np.abs(m.sht_synthesis(e) - f).max() <= 1e-12
    # orthonormal Y_lm without Condon-Shortley phase against scipy (lpmv has the CS phase)
    for (l, mm) in [(3, 2), (7, 5), (12, 12)]:
        norm = np.sqrt((2 * l + 1) / (4 * np.pi) * factorial(l - mm) / factorial(l + mm))
        P = (-1) ** mm * lpmv(mm, l, np.cos(colat)) * norm
        g = 2.0 * P[:, None] * np.cos(mm * lon[None, :])          # Re of Q Y with Q = 1, counted twice
        e = np.zeros_like(q); e[_lm_index(L, l, mm)] = 1.0
        assert np.abs(m.sht_synthesis(e) - g).max() <= 1e-12 * max(1, np.abs(g).max())


def test_sht_roundtrip_and_vector(small_model):
    m = small_model
    L = m.dims["lmax"]
    rng = np.random.default_rng(3)
    q = rng.standard_normal(m.dims["nlm"]) + 1j * rng.standard_normal(m.dims["nlm"])
    q[: L + 1] = q[: L + 1].real
    assert np.abs(m.sht_analysis(m.sht_synthesis(q)) - q).max() <= 1e-11
    vt, vp = m.sph_synthesis(q)
    q2 = q.copy(); q2[0] = 0.0
    assert np.abs(m.sph_analysis(vt, vp) - q2).max() <= 1e-11
    # gradient identity: d/dtheta of Y10 = -sqrt(3/4pi) sin(theta)
    colat, lon, w = m.grid_geometry()
    e = np.zeros_like(q); e[_lm_index(L, 1, 0)] = 1.0
    vt, vp = m.sph_synthesis(e)
    assert np.abs(vt + np.sqrt(3 / (4 * np.pi)) * np.sin(colat)[:, None]).max() <= 1e-12
    assert np.abs(vp).max() <= 1e-12


def test_tensor_sh_roundtrip_and_norm(small_model):
    m = small_model
    L = m.dims["lmax"]
    nlm = m.dims["nlm"]
    rng = np.random.default_rng(5)
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
    d = m.tensor_synth(c)
    c2 = m.tensor_analysis(d)
    assert np.abs(c2 - c).max() <= 1e-9 * np.abs(c).max()
    # n6(l) = B13 norm 2 J (J-2) of Z6 for exact quadrature (sectoral calibration)
    n6 = m.tensor_n6()
    for l in range(2, L + 1):
        J = l * (l + 1.0)
        assert abs(n6[l] - 2.0 * J * (J - 2.0)) <= 1e-9 * 2.0 * J * (J - 2.0)


# -------------------------------------------------------------- flotation
def test_flotation_truth_table(oracle):
    topo = np.array([-100.0, -100.0, -100.0, 100.0, -1.0, 0.0])
    ice = np.array([0.0, 100.0, 200.0, 0.0, 1.074, 0.0])
    # rho_i I < -rho_w topo: 0 < 1e5 ocean; 93100 < 1e5 ocean; 186200 >= 1e5 land; land; 999.9 < 1000 ocean; topo=0 land
    expect = np.array([1, 1, 0, 0, 1, 0], float)
    assert np.array_equal(oracle.ocean_function(topo, ice), expect)


# --------------------------------------------------------------- SLE + 3-D
def test_sle_mass_conservation(oracle):
    c = configs.config(1, lmax=16)
    m = oracle.OracleModel(c["params"], c["z_bed_eq"], c["h_ice_eq"])
    for k in range(3):
        m.update(c["ice"](k + 40), c["dt_yr"])
        assert m.worst_mass_resid <= 1e-10
    assert m.scalar(6) > 0   # the SLE actually iterated


def test_3d_uniform_equals_1d(oracle):
    """A laterally uniform perturbation forced through the pseudo-spectral
    3-D advance (every Maxwell element 3-D) reproduces the 1-D run."""
    c = configs.config(0, lmax=8)
    p = c["params"]
    a = oracle.OracleModel(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    p3 = configs.config(0, lmax=8)["params"]
    p3.visc3d_tol = -1.0
    b = oracle.OracleModel(p3, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    ne = b.dims["ne"]
    b.enable_visc_3d(np.zeros((ne, b.nlat, b.nphi)))
    assert b.dims["ne3d"] == b.dims["nem"]
    for k in range(4):
        a.update(c["ice"](k), c["dt_yr"])
        b.update(c["ice"](k), c["dt_yr"])
    assert rel(b.memory(), a.memory()) <= 1e-9
    assert rel(b.spectrum(1), a.spectrum(1)) <= 1e-9
