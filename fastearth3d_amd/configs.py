"""Seeded synthetic workloads of the benchmark configurations.

Each builder returns a dict with the `&vilma` record (`params`), the
reference state (`z_bed_eq`, `h_ice_eq`, optional `h_ice_init`), the ice
history as a function `ice(step)` of the call index, the interval `dt_yr`,
the number of calls `steps`, and for 3-D runs the per-element log10
viscosity perturbation `pert` (ne, nlat, nphi). Grid fields are (nlat, nphi)
C-order arrays (phi fastest), latitude rows north-first on the Gauss nodes.

The radial element counts of these workloads (40 / 64 elements) are not
expressible with the reference's depth-target mesh rule
(src/vilma_radial_fe.f90:140-224, which gives 217-218 elements for these
earths); they are set through the additive `n_elem_layer` knob, which the
oracle honours identically.
"""
import numpy as np

from .namelist import params_defaults

KM = 1.0e3
# surface-first layers: elastic lithosphere 0-80 km, upper mantle 80-420 and
# 420-670 km (5e20 Pa s), lower mantle 670-2891 km (2e21 Pa s), fluid core;
# M3-L70-V01 densities and rigidities (PREM-like, piecewise constant).
_LAYERS = [
    # r_bot, r_top, rho, mu, eta, rheology
    (6291.0 * KM, 6371.0 * KM, 3037.0, 5.0605e10, 1.0e40, 0),
    (5951.0 * KM, 6291.0 * KM, 3438.0, 7.0363e10, 5.0e20, 1),
    (5701.0 * KM, 5951.0 * KM, 3871.0, 1.0549e11, 5.0e20, 1),
    (3480.0 * KM, 5701.0 * KM, 4978.0, 2.2834e11, 2.0e21, 1),
    (0.0, 3480.0 * KM, 10750.0, 0.0, 0.0, 2),
]
_ELEMS = {40: (6, 8, 6, 12, 8), 64: (8, 12, 8, 20, 16)}


def five_layer_params(lmax, ne, nlat=0, nphi=0):
    p = params_defaults()
    p.lmax, p.nlat, p.nphi = lmax, nlat, nphi
    p.earth = b"custom"
    p.n_layer = 5
    p.r_earth, p.r_core = 6371.0 * KM, 3480.0 * KM
    for i, (rb, rt, rho, mu, eta, rh) in enumerate(_LAYERS):
        p.r_bot[i], p.r_top[i], p.rho[i], p.mu[i], p.eta[i], p.rheology[i] = rb, rt, rho, mu, eta, rh
    if ne in _ELEMS:
        counts = _ELEMS[ne]
    else:  # proportional split of an arbitrary total (ensemble members)
        w = np.array([8, 12, 8, 20, 16], float)
        c = np.maximum(1, np.floor(w * ne / w.sum())).astype(int)
        while c.sum() < ne:
            c[np.argmax(w * ne / w.sum() - c)] += 1
        counts = tuple(int(x) for x in c)
    for i, n in enumerate(counts):
        p.n_elem_layer[i] = n
    return p


def gauss_grid(nlat, nphi):
    """Colatitudes (north first) and longitudes of the model Gauss grid."""
    x, _ = np.polynomial.legendre.leggauss(nlat)
    colat = np.arccos(x[::-1])
    lon = 2.0 * np.pi * np.arange(nphi) / nphi
    return colat, lon


def _angdist(colat, lon, c0, l0):
    ct = np.cos(colat)[:, None] * np.cos(c0) + np.sin(colat)[:, None] * np.sin(c0) * np.cos(lon[None, :] - l0)
    return np.arccos(np.clip(ct, -1.0, 1.0))


def spherical_cap_sqrt(colat, lon, colat0_deg, lon0_deg, radius_deg, h0):
    """h0*sqrt(1-(d/d0)^2) inside angular radius d0 of the centre."""
    d = _angdist(colat, lon, np.radians(colat0_deg), np.radians(lon0_deg))
    d0 = np.radians(radius_deg)
    return np.where(d < d0, h0 * np.sqrt(np.clip(1.0 - (d / d0) ** 2, 0.0, None)), 0.0)


def random_sh_field(colat, lon, lmax, slope, seed):
    """Real random field sum_{l<=lmax} with per-degree power ~ l^slope
    (l >= 1), evaluated on the grid by direct Legendre summation."""
    rng = np.random.default_rng(seed)
    x = np.cos(colat)
    st = np.sin(colat)
    out = np.zeros((colat.size, lon.size))
    # orthonormal P_l^m via the standard recurrences (no CS phase)
    for m in range(0, lmax + 1):
        pmm = np.full_like(x, 1.0 / np.sqrt(4 * np.pi))
        for k in range(1, m + 1):
            pmm = pmm * np.sqrt((2 * k + 1) / (2 * k)) * st
        p_prev, p = None, pmm
        cosm, sinm = np.cos(m * lon), np.sin(m * lon)
        for l in range(m, lmax + 1):
            if l == m + 1:
                p_prev, p = p, np.sqrt(2 * m + 3) * x * pmm
            elif l > m + 1:
                a = np.sqrt((4 * l * l - 1) / (l * l - m * m))
                b = np.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
                p_prev, p = p, a * (x * p - b * p_prev)
            if l == 0:
                continue
            amp = np.sqrt(float(l) ** slope / (2 * l + 1))
            ca, sa = rng.standard_normal(2) * amp
            if m == 0:
                out += ca * p[:, None]
            else:
                out += np.sqrt(2.0) * p[:, None] * (ca * cosm[None, :] + sa * sinm[None, :])
    return out


def random_sh_field_fast(colat, lon, lmax, slope, seed):
    """The field of random_sh_field (same coefficients, same draw order),
    evaluated per order m as a Legendre column sum over the latitudes and
    then one matrix product over the longitudes; O(lmax^2 nlat + lmax nlat
    nphi) instead of O(lmax^2 nlat nphi)."""
    rng = np.random.default_rng(seed)
    x = np.cos(colat)
    st = np.sin(colat)
    ac = np.zeros((colat.size, lmax + 1))
    asn = np.zeros((colat.size, lmax + 1))
    for m in range(0, lmax + 1):
        pmm = np.full_like(x, 1.0 / np.sqrt(4 * np.pi))
        for k in range(1, m + 1):
            pmm = pmm * np.sqrt((2 * k + 1) / (2 * k)) * st
        p_prev, p = None, pmm
        for l in range(m, lmax + 1):
            if l == m + 1:
                p_prev, p = p, np.sqrt(2 * m + 3) * x * pmm
            elif l > m + 1:
                a = np.sqrt((4 * l * l - 1) / (l * l - m * m))
                b = np.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
                p_prev, p = p, a * (x * p - b * p_prev)
            if l == 0:
                continue
            amp = np.sqrt(float(l) ** slope / (2 * l + 1))
            ca, sa = rng.standard_normal(2) * amp
            w = 1.0 if m == 0 else np.sqrt(2.0)
            ac[:, m] += w * ca * p
            asn[:, m] += w * sa * p
    mm = np.arange(lmax + 1)[:, None] * lon[None, :]
    return ac @ np.cos(mm) + asn @ np.sin(mm)


def ocean_topography(colat, lon, lmax, seed, ocean_frac=0.70, amp=5000.0, fast=False):
    """Random topography (power ~ l^-2 up to lmax) offset so `ocean_frac` of
    the grid is below sea level and scaled to +-amp metres."""
    f = (random_sh_field_fast if fast else random_sh_field)(colat, lon, lmax, -2.0, seed)
    f = f - np.quantile(f, ocean_frac)
    return f * (amp / np.abs(f).max())


def three_domes(colat, lon, seed):
    """Three Gaussian ice domes, sigma 8-15 deg, peaks 2000-3500 m, centred at
    random mid/high latitudes (45-75 deg, either hemisphere)."""
    rng = np.random.default_rng(seed)
    h = np.zeros((colat.size, lon.size))
    for _ in range(3):
        lat = rng.uniform(45.0, 75.0) * rng.choice([-1.0, 1.0])
        lo = rng.uniform(0.0, 360.0)
        sig = np.radians(rng.uniform(8.0, 15.0))
        peak = rng.uniform(2000.0, 3500.0)
        d = _angdist(colat, lon, np.radians(90.0 - lat), np.radians(lo))
        h += peak * np.exp(-0.5 * (d / sig) ** 2)
    return h


def ice_history_factor(t_kyr):
    """0 -> 1 linearly over 20 kyr, hold 5 kyr, back to 0 over 10 kyr."""
    t = np.asarray(t_kyr, float)
    return np.where(t < 20.0, t / 20.0, np.where(t < 25.0, 1.0, np.clip(1.0 - (t - 25.0) / 10.0, 0.0, 1.0)))


def element_midpoint_depths(params):
    """Depth [m] of each radial element's midpoint for the mesh of params
    (whole-sphere mesh, innermost element first)."""
    n = params.n_layer
    depths = []
    for i in range(n - 1, -1, -1):
        rb, rt, ne = params.r_bot[i], params.r_top[i], params.n_elem_layer[i]
        h = (rt - rb) / ne
        for k in range(ne):
            depths.append(params.r_earth - (rb + (k + 0.5) * h))
    return np.array(depths)


def delta_visc(params, colat, lon, seed, lmax_delta=32, std=0.5, fast=False):
    """log10 viscosity perturbation per element (ne, nlat, nphi): a random SH
    field (l <= lmax_delta) of standard deviation `std`, full strength from
    the lithosphere base to 100 km depth, decaying as exp(-(z-100 km)/300 km)
    below, zero in the lithosphere."""
    f = (random_sh_field_fast if fast else random_sh_field)(colat, lon, lmax_delta, -2.0, seed)
    f = (f - f.mean()) * (std / f.std())
    z = element_midpoint_depths(params)
    litho = params.r_earth - params.r_bot[0]
    s = np.where(z < litho, 0.0, np.where(z <= 100.0 * KM, 1.0, np.exp(-(z - 100.0 * KM) / (300.0 * KM))))
    return s[:, None, None] * f[None, :, :]


ENSEMBLE_LMAX = (64, 96, 128, 192)


def member_params(lmax, ne, litho_km, eta_um, eta_lm, nlat=0, nphi=0):
    """The five-layer earth with a lithosphere of litho_km, upper mantle
    (lithosphere base to 670 km, split at 420 km) of viscosity eta_um and
    lower mantle (670 km to the core) of eta_lm; densities and rigidities of
    the five-layer earth; ne elements split over the layers as there."""
    p = five_layer_params(lmax, ne, nlat, nphi)
    p.r_bot[0] = p.r_top[1] = p.r_earth - litho_km * KM
    p.eta[1] = p.eta[2] = eta_um
    p.eta[3] = eta_lm
    return p


def ensemble_members(n=64, seed=0):
    """The members of the uneven ensemble: per member (lmax, ne, is_3d,
    litho_km, eta_um, eta_lm) from its own seeded draws: lmax from
    {64, 96, 128, 192}, ne uniform in 32..64, lithosphere uniform in
    60-120 km, upper- and lower-mantle viscosity log-uniform in
    1e20-1e22 Pa s. A quarter of the members (16 of 64), picked by a seeded
    permutation, are laterally heterogeneous (3-D)."""
    three = set(np.random.default_rng(seed + 4242).permutation(n)[: n // 4].tolist())
    out = []
    for i in range(n):
        r = np.random.default_rng([seed, 7, i])
        L = int(r.choice(ENSEMBLE_LMAX))
        ne = int(r.integers(32, 65))
        litho = float(r.uniform(60.0, 120.0))
        eu, el = (float(10.0 ** r.uniform(20.0, 22.0)) for _ in range(2))
        out.append((L, ne, i in three, litho, eu, el))
    return out


def ensemble_member(i, n=64, seed=0, members=None):
    """Workload of ensemble member i: its own earth (ensemble_members), the
    sea-level + rotation setting of the 1-D sea-level workload (migrating
    coastlines, subgrid term, dt 100 yr, 200 calls) at the member's lmax
    (Gauss grid 192 x 384 at lmax 128, the default grid otherwise) with its
    own topography (power ~ l^-2 up to lmax, 70 % ocean, +-5 km) and three
    ice domes, and for 3-D members a lateral viscosity field of its own
    (std 0.5 up to degree 32). `members` overrides the drawn member list."""
    L, ne, is3d, litho, eu, el = (members or ensemble_members(n, seed))[i]
    nlat, nphi = (192, 384) if L == 128 else (0, 0)
    p = member_params(L, ne, litho, eu, el, nlat, nphi)
    p.rotation = 1
    nphi, nlat = p.grid()
    colat, lon = gauss_grid(nlat, nphi)
    z = ocean_topography(colat, lon, L, seed + 1000 + i, fast=True)
    domes = three_domes(colat, lon, seed + 101 + i)
    pert = None
    if is3d:
        p.l_visc_3d = 1
        pert = delta_visc(p, colat, lon, seed + 211 + i, fast=True)
    dt = 100.0
    return dict(name="c4m%d" % i, params=p, z_bed_eq=z, h_ice_eq=np.zeros_like(z), h_ice_init=None,
                ice=lambda k, dt=dt: domes * float(ice_history_factor((k + 1) * dt / 1000.0)),
                domes=domes, dt_yr=dt, steps=200, pert=pert)


def config(idx, lmax=None, ne=None, seed=0):
    """Workload of benchmark configuration idx (0, 1 or 2); lmax / ne may be
    lowered for the oracle-sized parity cases. The ensemble (configuration 4)
    is built member by member (ensemble_member)."""
    if idx == 0:
        L = lmax or 64
        p = five_layer_params(L, ne or 40)
        p.rotation = 0
        p.deg1_frame = b"cf"
        nphi, nlat = p.grid()
        colat, lon = gauss_grid(nlat, nphi)
        cap = spherical_cap_sqrt(colat, lon, 30.0, 280.0, 12.0, 1500.0)
        z = np.full((nlat, nphi), 1000.0)
        return dict(name="c0", params=p, z_bed_eq=z, h_ice_eq=np.zeros_like(z), h_ice_init=cap,
                    ice=lambda k: cap, dt_yr=50.0, steps=200, pert=None)
    if idx in (1, 2):
        L = lmax or 128
        nlat, nphi = (192, 384) if L == 128 else (0, 0)
        p = five_layer_params(L, ne or 64, nlat, nphi)
        p.rotation = 1
        nphi, nlat = p.grid()
        colat, lon = gauss_grid(nlat, nphi)
        z = ocean_topography(colat, lon, L, seed + 11)
        domes = three_domes(colat, lon, seed + 23)
        dt = 100.0 if idx == 1 else 50.0
        steps = 350 if idx == 1 else 400
        pert = delta_visc(p, colat, lon, seed + 37) if idx == 2 else None
        if idx == 2:
            p.l_visc_3d = 1
        return dict(name="c%d" % idx, params=p, z_bed_eq=z, h_ice_eq=np.zeros_like(z), h_ice_init=None,
                    ice=lambda k, dt=dt: domes * float(ice_history_factor((k + 1) * dt / 1000.0)),
                    dt_yr=dt, steps=steps, pert=pert)
    raise ValueError("config %d" % idx)


# Martinec et al. (2018) sea-level benchmark cases, data-file letters
# (tests/test_benchmark_sle.f90:210-247): cap height [m], basin centre
# (colat, lon) [deg], basin bmax, b0 [m], fixed ocean (SLE1), Heaviside (T0),
# paleotopography spin-up.
MARTINEC_CASES = {
    "C2": (1500.0, (100.0, 320.0), (760.0, 1200.0), True, True, False),
    "D3": (500.0, (35.0, 25.0), (3800.0, 6000.0), True, False, False),
    "E2": (500.0, (35.0, 25.0), (3800.0, 6000.0), False, False, False),
    "F1": (500.0, (35.0, 25.0), (3800.0, 6000.0), False, False, True),
}


def martinec_cap(colat, lon, colat0_deg, lon0_deg, alpha_deg, h0):
    """h0 sqrt[(cos d - cos a)/(1 - cos a)] within angular radius a of the
    centre (src/vilma_field.f90:27-50)."""
    d = _angdist(colat, lon, np.radians(colat0_deg), np.radians(lon0_deg))
    ca = np.cos(np.radians(alpha_deg))
    return np.where(d <= np.radians(alpha_deg), h0 * np.sqrt(np.clip((np.cos(d) - ca) / (1.0 - ca), 0.0, None)), 0.0)


def martinec_basin(colat, lon, colat0_deg, lon0_deg, bmax, b0, sigma_deg=26.0):
    """bmax - b0 exp(-d^2 / 2 sigma^2) (src/vilma_field.f90:52-70)."""
    d = _angdist(colat, lon, np.radians(colat0_deg), np.radians(lon0_deg))
    s = np.radians(sigma_deg)
    return bmax - b0 * np.exp(-d * d / (2.0 * s * s))


def martinec2018(case, lmax=128):
    """Workload of a Martinec (2018) sea-level benchmark case
    (tests/test_benchmark_sle.f90:100-135, :255-290): the M3-L70-V01 earth on
    the (2 lmax, 4 lmax) Gauss grid, a sqrt cap of radius 10 deg centred at
    colat 25 / lon 75 on an exponential basin, dt 20 yr from a relaxed,
    ice-free state. T0: full ice for 500 calls; T1: the cap (radius and
    height) grows linearly over 500 calls, then is held 250. No rotational
    feedback; 20 inner sea-level iterations. `spinup` asks the caller for the
    paleotopography iteration of case F1."""
    h0, (bc, bl), (bmax, b0), fixed, heaviside, spinup = MARTINEC_CASES[case]
    p = params_defaults()
    p.earth = b"M3-L70-V01"
    p.lmax, p.nlat, p.nphi = lmax, 2 * lmax, 4 * lmax
    p.rotation = 0
    p.sle_n_inner = 20
    p.sle_fixed_ocean = 1 if fixed else 0
    p.sle_subgrid = 1
    nphi, nlat = p.grid()
    colat, lon = gauss_grid(nlat, nphi)
    basin = martinec_basin(colat, lon, bc, bl, bmax, b0)
    steps = 500 if heaviside else 750

    def ice(k):
        frac = 1.0 if heaviside else min(1.0, (k + 1) / 500.0)
        return martinec_cap(colat, lon, 25.0, 75.0, 10.0 * frac, h0 * frac)

    return dict(name="martinec_" + case, params=p, z_bed_eq=basin, h_ice_eq=np.zeros_like(basin),
                h_ice_init=None, ice=ice, dt_yr=20.0, steps=steps, pert=None, basin=basin, spinup=spinup,
                colat=colat, lon=lon)
