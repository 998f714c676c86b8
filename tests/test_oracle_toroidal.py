"""The oracle's toroidal degree of freedom (l_toroidal; src/vilma_radial_fe.f90:
699-817, src/vilma_viscoelastic.f90:373-439, :591-629, src/vilma_tensor_sh.f90:
168-387, src/vilma_response.f90:694-809, :1055-1109) pinned against what the
mathematics fixes, CPU only:

* the W stiffness re-derived from the strain representation: 2 int mu
  [|Z3|^2 e3 de3 + |Z4|^2 e4 de4] r^2 dr with e3 = W' - W/r, e4 = W/r and the
  B13 norms J/2, J(J-2)/2, by Gauss quadrature here, against the oracle's
  closed-form assembly (machine precision); its degree-1 null mode W = r
  (a rigid rotation of the solid shell);
* the strain route: the dissipative forcing of the memory tau = 2 mu eps(W)
  through the oracle's toroidal strain coefficients equals -A W (the
  2-point Gauss rule is exact on that integrand);
* the toroidal vector transform e_r x grad_1: orientation (-F, E) against the
  spheroidal synthesis of the same coefficients, and its round trip;
* the six-channel tensor transforms: round trip, the physical norm int tau:tau
  against the B13 norms of all six channels, and completeness on an arbitrary
  smooth symmetric tensor field built here from a Cartesian polynomial
  (tests/test_tensor_sh.f90), which four channels cannot represent;
* the coupled response (tests/test_toroidal.f90): a held Y20 load over a
  sin^2(theta) cos(2 phi) log-viscosity perturbation drives W; the reflection
  selection rule (U even / real, W odd / imaginary, even m only) holds to
  round-off; W scales as delta^1 and its uplift feedback as delta^2; an
  axisymmetric perturbation drives no W; with the channels off no W is
  reported;
* rotational invariance with the channels carried (tests/test_rotinv.f90):
  a tapered cap over a low-viscosity column on the pole and off the pole give
  the same cap-centre uplift.
"""
import numpy as np
import pytest

from fastearth3d_amd import _abi

SEC_PER_YEAR = 31556926.0


def _lm(L, l, m):
    return m * (2 * L + 1 - m) // 2 + l


def _m3(oracle, lmax, nlat=0, nphi=0):
    p = oracle.params_defaults()
    p.earth = b"M3-L70-V01"
    p.lmax, p.nlat, p.nphi = lmax, nlat, nphi
    p.rotation = 0
    return p


def _elements(oracle, p):
    """node radii and per-element shear modulus / Maxwell flag of p's mesh"""
    r, mu, rh = oracle.mesh_props(p)
    return r, mu, rh == 1


def _stiffness_by_strain(r, mu, j):
    """sum_e mu_e int [J e3_a e3_b + J (J-2) e4_a e4_b] r^2 dr, e3 = psi' - psi/r,
    e4 = psi/r, by 4-point Gauss-Legendre per element (exact: polynomial)."""
    J = j * (j + 1.0)
    x, w = np.polynomial.legendre.leggauss(4)
    nr = r.size
    A = np.zeros((nr, nr))
    for e in range(nr - 1):
        if mu[e] == 0.0:
            continue
        rk, rk1 = r[e], r[e + 1]
        h = rk1 - rk
        for xg, wg in zip(x, w):
            ra = 0.5 * (rk + rk1) + 0.5 * h * xg
            psi = np.array([(rk1 - ra) / h, (ra - rk) / h])
            dpsi = np.array([-1.0 / h, 1.0 / h])
            e3 = dpsi - psi / ra
            e4 = psi / ra
            blk = J * np.outer(e3, e3) + J * (J - 2.0) * np.outer(e4, e4)
            A[e:e + 2, e:e + 2] += mu[e] * 0.5 * h * wg * ra * ra * blk
    return A


@pytest.mark.parametrize("j", [1, 2, 7, 30])
def test_w_stiffness_by_strain_route(oracle, j):
    p = _m3(oracle, 8)
    r, mu, _ = _elements(oracle, p)
    A = oracle.tor_operator(p, j)
    B = _stiffness_by_strain(r, mu, j)
    assert np.abs(A - A.T).max() == 0.0
    assert np.abs(A - B).max() <= 1e-12 * np.abs(B).max()
    if j == 1:
        # W = r on the solid nodes (zero in the fluid core): a rigid rotation
        solid = np.zeros(r.size, bool)
        for e in range(r.size - 1):
            if mu[e] > 0:
                solid[e] = solid[e + 1] = True
        W = np.where(solid, r, 0.0)
        assert np.abs(A @ W).max() <= 1e-12 * np.abs(A).max() * r.max()


@pytest.mark.parametrize("j", [1, 3, 12])
def test_w_solve_and_strain_route(oracle, j):
    """The assembled solve inverts the stiffness on the live nodes (and at
    degree 1 returns the solution with no net rotation of the solid shell);
    the strain route (memory 2 mu eps(W) through the oracle's strain
    coefficients and forcing) reproduces -A W."""
    p = _m3(oracle, 8)
    r, mu, _ = _elements(oracle, p)
    rng = np.random.default_rng(j)
    live = np.zeros(r.size, bool)
    for e in range(r.size - 1):
        if mu[e] > 0:
            live[e] = live[e + 1] = True
    b = np.where(live, rng.standard_normal(r.size), 0.0)
    if j == 1:
        # no net torque (b orthogonal to the null mode W = r of the solid
        # shell), so the bordered solve is A x = b with int W r^3 dr = 0
        n = np.where(live, r, 0.0)
        b = b - (b @ n) / (n @ n) * n
        k4 = np.zeros(r.size)   # int psi_k r^3 dr over the solid elements, by 4-point Gauss here
        xg, wg = np.polynomial.legendre.leggauss(4)
        for e in range(r.size - 1):
            if mu[e] > 0:
                rk, rk1 = r[e], r[e + 1]
                h = rk1 - rk
                ra = 0.5 * (rk + rk1) + 0.5 * h * xg
                k4[e] += np.sum(0.5 * h * wg * (rk1 - ra) / h * ra ** 3)
                k4[e + 1] += np.sum(0.5 * h * wg * (ra - rk) / h * ra ** 3)
    A, x = oracle.tor_operator(p, j, b)
    assert not np.any(x[~live])
    res = A @ x - b
    assert np.abs(res[live]).max() <= 1e-9 * np.abs(A).max() * np.abs(x).max()
    if j == 1:
        assert abs(k4 @ x) <= 1e-10 * np.abs(k4).max() * np.abs(x).max() * r.size
    # strain route: f = -int tau:deps with tau = 2 mu eps(x) equals -A x
    f = oracle.tor_forcing_of_w(p, j, x)
    assert np.abs(f + A @ x).max() <= 1e-10 * np.abs(A @ x).max()


def test_tor_vector_transform(oracle):
    L = 16
    p = _m3(oracle, L)
    m = oracle.OracleModel(p, np.full(p.grid()[::-1], 1000.0), np.zeros(p.grid()[::-1]))
    rng = np.random.default_rng(3)
    q = rng.standard_normal(m.dims["nlm"]) + 1j * rng.standard_normal(m.dims["nlm"])
    for l in range(L + 1):
        q[_lm(L, l, 0)] = q[_lm(L, l, 0)].real   # a real field
    q[_lm(L, 0, 0)] = 0.0
    vt, vp = m.tor_synthesis(q)
    st, sp = m.sph_synthesis(q)     # (E, F)
    assert np.abs(vt + sp).max() <= 1e-12 * np.abs(sp).max()   # -F
    assert np.abs(vp - st).max() <= 1e-12 * np.abs(st).max()   # +E
    q2 = m.tor_analysis(vt, vp)
    assert np.abs(q2 - q).max() <= 1e-11 * np.abs(q).max()
    assert np.abs(m.sph_analysis(vt, vp)).max() <= 1e-11 * np.abs(q).max()   # no spheroidal leak
    m.finalize()


def _random_coeffs(L, nch, seed):
    rng = np.random.default_rng(seed)
    nlm = (L + 1) * (L + 2) // 2
    c = rng.standard_normal((nch, nlm)) + 1j * rng.standard_normal((nch, nlm))
    for mm in range(L + 1):
        for l in range(mm, L + 1):
            i = _lm(L, l, mm)
            if mm == 0:
                c[:, i] = c[:, i].real
            # channels without a harmonic: Z2, Z5, Z3 need l >= 1; Z6, Z4 need l >= 2
            for lam, lmin in zip(range(nch), (0, 1, 1, 2, 1, 2)):
                if l < lmin:
                    c[lam, i] = 0.0
    return c


def _tensor_norm(d, w, nphi):
    """int tau:tau dOmega from the dyad planes rr, rt, rp, tt, tp, pp (the
    off-diagonal planes hold 2 tau_ab, e_ab:e_ab = 1/2)."""
    rr, rt, rp, tt, tp, pp = d
    f = rr ** 2 + tt ** 2 + pp ** 2 + 0.5 * (rt ** 2 + rp ** 2 + tp ** 2)
    return float(np.sum(w[:, None] * f) * 2.0 * np.pi / nphi)


def _spectral_norm(c, L):
    tot = 0.0
    for mm in range(L + 1):
        kap = 1.0 if mm == 0 else 2.0
        for l in range(mm, L + 1):
            J = l * (l + 1.0)
            nrm = [1.0, 0.5 * J, 2.0 * J * J, 2.0 * J * (J - 2.0), 0.5 * J, 0.5 * J * (J - 2.0)]
            for lam in range(c.shape[0]):
                tot += kap * nrm[lam] * abs(c[lam, _lm(L, l, mm)]) ** 2
    return tot


def _cartesian_tensor_field(colat, lon, seed=12345):
    """T_ij(x) = A_ij + B_ijk x_k + C_ijkl x_k x_l, symmetric in ij, on the
    unit sphere, resolved on (e_r, e_t, e_p); degree <= 4; planes hold the
    coefficient of (e_a e_b + e_b e_a)/2, i.e. 2 T_ab off the diagonal. The
    uniform tangential trace (no Z carries it) is removed."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (3, 3)); A = A + A.T
    B = rng.uniform(-1, 1, (3, 3, 3)); B = B + B.transpose(1, 0, 2)
    Cq = rng.uniform(-1, 1, (3, 3, 3, 3)); Cq = Cq + Cq.transpose(1, 0, 2, 3)
    th, ph = np.meshgrid(colat, lon, indexing="ij")
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    er = np.stack([st * cp, st * sp, ct])
    et = np.stack([ct * cp, ct * sp, -st])
    ep = np.stack([-sp, cp, np.zeros_like(st)])
    x = er
    T = A[:, :, None, None] + np.einsum("ijk,k...->ij...", B, x) + np.einsum("ijkl,k...,l...->ij...", Cq, x, x)

    def q(a, b):
        return np.einsum("i...,ij...,j...->...", a, T, b)

    d = np.stack([q(er, er), 2 * q(er, et), 2 * q(er, ep), q(et, et), 2 * q(et, ep), q(ep, ep)])
    return d


def test_tensor_six_channels(oracle):
    L = 10
    p = _m3(oracle, L, nlat=4 * L, nphi=4 * L)
    nphi, nlat = p.grid()
    m = oracle.OracleModel(p, np.full((nlat, nphi), 1000.0), np.zeros((nlat, nphi)))
    colat, lon, w = m.grid_geometry()
    c = _random_coeffs(L, 6, 1)
    d = m.tensor_synth(c)
    c2 = m.tensor_analysis(d, nch=6)
    assert np.abs(c2 - c).max() <= 1e-9 * np.abs(c).max()
    # physical norm of the synthesised field = B13 norms of all six channels
    assert abs(_tensor_norm(d, w, nphi) - _spectral_norm(c, L)) <= 1e-10 * _spectral_norm(c, L)
    # each toroidal channel alone: no leak into the others
    for lam in (4, 5):
        cc = np.zeros_like(c); cc[lam] = c[lam]
        back = m.tensor_analysis(m.tensor_synth(cc), nch=6)
        others = np.delete(back, lam, axis=0)
        assert np.abs(others).max() <= 1e-11 * np.abs(c[lam]).max()
    # completeness: an arbitrary smooth symmetric tensor survives with six channels, not with four
    d0 = _cartesian_tensor_field(colat, lon)
    tr = d0[3] + d0[5]
    mean = np.sum(w[:, None] * tr) / (2.0 * nphi)
    d0[3] -= 0.5 * mean; d0[5] -= 0.5 * mean
    d6 = m.tensor_synth(m.tensor_analysis(d0, nch=6))
    d4 = m.tensor_synth(m.tensor_analysis(d0, nch=4))
    assert np.abs(d6 - d0).max() <= 1e-11 * np.abs(d0).max()
    assert np.abs(d4 - d0).max() >= 1e-2 * np.abs(d0).max()
    n4 = m.tensor_n4()
    assert n4[0] == 0.0 and n4[1] == 0.0 and np.all(n4[2:] > 0)
    m.finalize()


def _tor_run(oracle, delta, axisym, tor, L=12, nstep=40, dt_yr=40.0):
    """tests/test_toroidal.f90 run(): a 1e3 Y20 load held nstep steps of
    dt_yr over M3-L70-V01 with a log10-viscosity perturbation in the Maxwell
    layers; the surface uplift and toroidal W coefficients after the last
    step's drift."""
    p = _m3(oracle, L, nlat=3 * L, nphi=3 * L)
    p.l_visc_3d = 1
    p.l_toroidal = 1 if tor else 0
    nphi, nlat = p.grid()
    m = oracle.OracleModel(p, np.full((nlat, nphi), 1000.0), np.zeros((nlat, nphi)))
    m.set_dt(dt_yr * SEC_PER_YEAR)
    colat, lon, _ = m.grid_geometry()
    r, mu, maxwell = _elements(oracle, _m3(oracle, L))
    st, ct = np.sin(colat)[:, None], np.cos(colat)[:, None]
    pat = delta * ct * ct * np.ones((1, nphi)) if axisym else delta * st * st * np.cos(2.0 * lon)[None, :]
    pert = np.zeros((m.dims["ne"], nlat, nphi))
    pert[maxwell] = pat
    m.enable_visc_3d(pert)
    assert m.dims["nlam"] == (6 if tor else 4)
    sig = np.zeros(m.dims["nlm"], complex)
    sig[_lm(L, 2, 0)] = 1.0e3
    for _ in range(nstep):
        m.begin_step()
        m.apply(sig)
        m.advance(sig)
    m.begin_step()
    u, _n = m.apply(sig)
    t = m.spectrum(_abi.SP_TOR)
    m.finalize()
    return u, t


def test_toroidal_selection_rule_and_scaling(oracle):
    L, delta = 12, 0.2
    u_on, t_on = _tor_run(oracle, delta, False, True)
    wmax, umax = np.abs(t_on).max(), np.abs(u_on).max()
    assert wmax >= 1e-6 * umax                    # (1) W is driven
    err = 0.0
    for l in range(1, L + 1):
        for mm in range(l + 1):
            i = _lm(L, l, mm)
            if mm % 2:
                err = max(err, abs(u_on[i]) / umax, abs(t_on[i]) / wmax)
            else:
                err = max(err, abs(u_on[i].imag) / umax, abs(t_on[i].real) / wmax)
    assert err <= 1e-10                           # (2) reflection selection rule
    u_off, t_off = _tor_run(oracle, delta, False, False)
    assert not np.any(t_off)
    u_h, t_h = _tor_run(oracle, 0.5 * delta, False, True)
    u_hoff, _ = _tor_run(oracle, 0.5 * delta, False, False)
    slope_w = np.log(np.abs(t_on).max() / np.abs(t_h).max()) / np.log(2.0)
    slope_du = np.log(np.abs(u_on - u_off).max() / np.abs(u_h - u_hoff).max()) / np.log(2.0)
    assert abs(slope_w - 1.0) <= 0.1, slope_w     # (3) W first order in delta
    assert abs(slope_du - 2.0) <= 0.2, slope_du   #     its feedback second order
    # the reference's own figures for this configuration: W ~ delta^0.987, its
    # uplift feedback ~ delta^1.986 (doc/design-toroidal.md:56-57); measured
    # 0.9866 and 1.9858
    assert abs(slope_w - 0.987) <= 0.001, slope_w
    assert abs(slope_du - 1.986) <= 0.001, slope_du


def test_toroidal_axisymmetric_null(oracle):
    u, t = _tor_run(oracle, 0.2, True, True)
    assert np.abs(t).max() <= 1e-12 * np.abs(u).max()


def _taper(g, cin=0.20, cout=0.40):
    return np.where(g <= cin, 1.0, np.where(g >= cout, 0.0, 0.5 * (1.0 + np.cos(np.pi * (g - cin) / (cout - cin)))))


def _rotinv_case(oracle, beta, lon0, L=16):
    """tests/test_rotinv.f90 run_case with the toroidal channels carried: a
    200 m tapered cap (ramped over 100 yr, held to 200 yr) over a column of
    log10-viscosity -2 at 70-170 km depth under it, axis (beta, lon0); forward
    Euler at 2 yr (stable for the 1e19 Pa s column). Returns the cap-centre
    uplift and max|W(a)| / max|U(a)|."""
    p = _m3(oracle, L, nlat=3 * L, nphi=3 * L)
    p.l_visc_3d = 1
    p.l_toroidal = 1
    nphi, nlat = p.grid()
    m = oracle.OracleModel(p, np.full((nlat, nphi), 1000.0), np.zeros((nlat, nphi)))
    dt = 2.0
    m.set_dt(dt * SEC_PER_YEAR)
    colat, lon, _ = m.grid_geometry()
    ct = np.cos(colat)[:, None] * np.cos(beta) + np.sin(colat)[:, None] * np.sin(beta) * np.cos(lon[None, :] - lon0)
    g = np.arccos(np.clip(ct, -1.0, 1.0))
    cap_lm = m.sht_analysis(_taper(g))
    r, _ = oracle.mesh(p)
    pert = np.zeros((m.dims["ne"], nlat, nphi))
    for e in range(m.dims["ne"]):
        depth = 6371.0e3 - 0.5 * (r[e] + r[e + 1])
        if 70.0e3 < depth < 170.0e3:
            pert[e] = -2.0 * _taper(g)
    m.enable_visc_3d(pert)
    nstep = int(round(200.0 / dt))
    for i in range(1, nstep + 1):
        sig = 931.0 * min(i * dt / 100.0, 1.0) * 200.0 * cap_lm
        m.begin_step()
        u, _n = m.apply(sig)
        m.advance(sig)
    t = m.spectrum(_abi.SP_TOR)
    import martinec
    uc = martinec.eval_points(u, L, np.array([beta]), np.array([lon0]))[0]
    m.finalize()
    return uc, np.abs(t).max() / np.abs(u).max()


def test_toroidal_rotational_invariance(oracle):
    u_pole, w_pole = _rotinv_case(oracle, 0.0, 0.0)
    u_off, w_off = _rotinv_case(oracle, np.radians(60.0), np.radians(40.0))
    tol = 0.03   # max(1 %, 48 / lmax %) at lmax 16, tests/test_rotinv.f90:52
    assert abs(u_off - u_pole) <= tol * abs(u_pole), (u_pole, u_off)
    assert w_pole <= 1e-12
    assert w_off <= tol
