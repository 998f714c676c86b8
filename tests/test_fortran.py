"""The Fortran host module `vilma` (fastearth3d_amd/fortran/vilma.f90).

CPU: vilma_par_load through the intrinsic NAMELIST gives the same record as
the Python loader (overlay of a run file on a defaults file, years -> s).
GPU: a Fortran host running solid_earth_init / solid_earth_update through
the module matches the Python host on the same library bit for bit, on the
Gauss grid and on a lon-lat host grid (through the remap).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from fastearth3d_amd import build, configs, namelist

HERE = os.path.dirname(os.path.abspath(__file__))

DEFAULTS = """&vilma
  solver = "v2"
  lmax = 16, nlat = 0, nphi = 0, mmax = -1, mres = 1, eps_polar = -1.0
  earth = "custom", n_layer = 3, r_earth = 6371.0e3, r_core = 3480.0e3
  r_bot = 6291.0e3, 5701.0e3, 3480.0e3
  r_top = 6371.0e3, 6291.0e3, 5701.0e3
  rho = 3300.0, 4000.0, 5000.0
  mu = 6.0e10, 1.0e11, 2.0e11
  eta = 1.0e40, 5.0e20, 2.0e21
  rheology = 0, 1, 1
  scheme = "fe", max_couple_iter = 20, earth_response = "ve"
  sle_n_outer = 3, sle_n_inner = 20, sle_tol = 1.0e-7, sle_max_mem_iter = 20
  sle_fixed_ocean = .false., sle_subgrid = .true.
  dt_init = 0.0, dt_min = 0.0, dt_max = 1.0e30
  rtol = 1.0e-4, atol = 1.0e-3, safety = 0.9, grow_max = 5.0, shrink_min = 0.2, cfl = 1.0
  rotation = .true., equil_time_max = 0.0, equil_rate_tol = 1.0e-3, pre_spinup_1d = .false.
  deg1_frame = "cm", l_visc_3d = .false., visc_3d_file = "visc.nc"
  visc3d_tol = 1.0e-3, l_toroidal = .false., visc_log10_min = 19.5, visc_log10_max = 30.0
  visc3d_lid_depth = 0.0, visc3d_lid_log10max = 22.0
  vilma_v1_jmax = 170, vilma_v1_nsub = 1
/
"""
RUN = """&vilma
  lmax = 12
  eta = 1.0e40, 3.0e20, 1.0e21
  dt_max = 200.0
  rotation = .false.
  deg1_frame = "cf"
/
"""


def _driver():
    if not os.path.exists(build.DRIVER):
        if not os.path.exists(build.FLANG):
            pytest.skip("no flang in this image")
        build.build()
    return build.DRIVER


def test_fortran_par_load_matches_python(tmp_path):
    drv = _driver()
    d = tmp_path / "defaults.nml"; d.write_text(DEFAULTS)
    r = tmp_path / "run.nml"; r.write_text(RUN)
    out = subprocess.run([drv, str(r), "--print-params", str(d)], capture_output=True, text=True, check=True).stdout
    rec = {line.split()[0]: line.split()[1:] for line in out.splitlines() if line.strip()}
    p = namelist.vilma_par_load(str(r), defaults_file=str(d))
    assert [int(v) for v in rec["grid"]] == [p.lmax, p.nlat, p.nphi, p.mmax, p.mres]
    assert rec["earth"] == [p.earth.decode(), str(p.n_layer)]
    # layers beyond n_layer are unused (the two loaders start from different
    # built-in records there: the type defaults vs the overlaid defaults file)
    n = p.n_layer
    assert [float(v) for v in rec["r_bot"]][:n] == [p.r_bot[i] for i in range(n)]
    assert [float(v) for v in rec["eta"]][:n] == [p.eta[i] for i in range(n)]
    assert [int(v) for v in rec["rheology"]][:3] == [p.rheology[i] for i in range(3)]
    assert [float(v) for v in rec["dt"]] == [p.dt_init, p.dt_min, p.dt_max]
    assert float(rec["sle"][0]) == p.sle_tol
    assert rec["flags"][:3] == ["F", "T", "cf"]
    assert p.rotation == 0 and p.deg1_frame == b"cf"
    assert float(rec["dt"][2]) == 200.0 * 31556926.0


def test_fortran_rejects_bad_frame(tmp_path):
    drv = _driver()
    r = tmp_path / "bad.nml"
    r.write_text(DEFAULTS.replace('deg1_frame = "cm"', 'deg1_frame = "xx"'))
    res = subprocess.run([drv, str(r), "--print-params"], capture_output=True, text=True)
    assert res.returncode != 0
    assert "deg1_frame" in res.stdout + res.stderr


def _namelist(tmp_path, p, extra=()):
    nml = tmp_path / "c.nml"
    lines = ["&vilma", "  lmax = %d" % p.lmax, "  earth = \"custom\"", "  n_layer = %d" % p.n_layer,
             "  r_earth = %r, r_core = %r" % (p.r_earth, p.r_core)]
    for key in ("r_bot", "r_top", "rho", "mu", "eta"):
        lines.append("  %s = %s" % (key, ", ".join(repr(getattr(p, key)[i]) for i in range(p.n_layer))))
    for key in ("rheology", "n_elem_layer"):
        lines.append("  %s = %s" % (key, ", ".join(str(getattr(p, key)[i]) for i in range(p.n_layer))))
    lines += ["  rotation = .false.", "  deg1_frame = \"%s\"" % p.deg1_frame.decode()] + list(extra) + ["/"]
    nml.write_text("\n".join(lines) + "\n")
    return nml


def _run_both(tmp_path, nml, fin, extra=(), py_extra=()):
    """The Fortran driver and the Python host (tests/fortran/py_host.py), each
    in a process of its own. Returns both output files' bytes."""
    ff, fp = tmp_path / "out_f.bin", tmp_path / "out_p.bin"
    subprocess.run([_driver(), str(nml), str(fin), str(ff)] + list(extra), check=True, timeout=300)
    subprocess.run([sys.executable, os.path.join(HERE, "fortran", "py_host.py"), str(nml), str(fin), str(fp)]
                   + list(extra) + list(py_extra), check=True, timeout=300)
    return np.fromfile(ff, np.float64), np.fromfile(fp, np.float64)


@pytest.mark.gpu
def test_fortran_host_matches_python_host(tmp_path):
    """Fortran module -> C-ABI vs Python ctypes -> C-ABI: identical state."""
    c = configs.config(1, lmax=16)
    p = c["params"]
    p.rotation = 0
    nphi, nlat = p.grid()
    nml = _namelist(tmp_path, p)
    ice = np.stack([c["ice"](k + 60) for k in range(3)])
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array([nphi, nlat, 3], np.int32).tofile(f)
        np.array([c["dt_yr"]], np.float64).tofile(f)
        c["z_bed_eq"].astype(np.float64).tofile(f)   # (nlat, nphi) C order == (nphi, nlat) Fortran order
        ice.astype(np.float64).tofile(f)
    of, op = _run_both(tmp_path, nml, fin)
    assert of.size == op.size == 3 * nphi * nlat + 2
    assert np.array_equal(of, op)                   # rsl, z_bed, C, bsl, time
    assert np.abs(of[:nphi * nlat]).max() > 0.0


@pytest.mark.gpu
def test_fortran_host_grid_coupling(tmp_path):
    """solid_earth_init with a lon-lat host grid (src/vilma_coupling.f90:172-184,
    355-359, 439-449): ice and bed mapped in conservatively (the ice with its
    mass kept), rsl and z_bed returned bilinearly on the host grid. The Fortran
    host equals a Python host doing the same maps around the same model."""
    import remap as R
    c = configs.config(1, lmax=16)
    p = c["params"]
    p.rotation = 0
    nml = _namelist(tmp_path, p)
    nphi, nlat = p.grid()
    # a 4 x 3 degree host grid; the host fields sampled (bilinearly, on the CPU)
    # from the Gauss-grid ones
    lon = 2.0 + 4.0 * np.arange(90)
    lat = -88.5 + 3.0 * np.arange(60)
    z_ll = R.to_ll(c["z_bed_eq"], lon, lat, nlat, nphi)
    ice_ll = np.stack([np.maximum(R.to_ll(c["ice"](k + 60), lon, lat, nlat, nphi), 0.0) for k in range(3)])
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array([lon.size, lat.size, 3], np.int32).tofile(f)
        np.array([c["dt_yr"]], np.float64).tofile(f)
        lon.astype(np.float64).tofile(f)
        lat.astype(np.float64).tofile(f)
        z_ll.astype(np.float64).tofile(f)
        ice_ll.astype(np.float64).tofile(f)
    of, op = _run_both(tmp_path, nml, fin, ["--host-grid"])
    nh = lon.size * lat.size
    assert of.size == op.size == 2 * nh + nphi * nlat + 2
    assert np.array_equal(of, op)                   # host rsl, host z_bed, Gauss C, bsl, time
    rsl = of[:nh].reshape(lat.size, lon.size)
    assert np.abs(rsl).max() > 0.0
    assert np.array_equal(of[nh:2 * nh].reshape(lat.size, lon.size), z_ll - rsl)


@pytest.mark.gpu
def test_fortran_spinup_matches_python_host(tmp_path):
    """solid_earth_spinup through the Fortran module (the namelist's
    equil_time_max = 3000 yr, equil_rate_tol) then three coupling calls equal
    the Python host calling vh_spinup with the same values, bit for bit."""
    c = configs.config(1, lmax=16)
    p = c["params"]
    p.rotation = 0
    nphi, nlat = p.grid()
    nml = _namelist(tmp_path, p, ["  equil_time_max = 3000.0, equil_rate_tol = 1.0e-4, pre_spinup_1d = .false."])
    ice = np.stack([c["ice"](k + 60) for k in range(3)])
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as f:
        np.array([nphi, nlat, 3], np.int32).tofile(f)
        np.array([c["dt_yr"]], np.float64).tofile(f)
        c["z_bed_eq"].astype(np.float64).tofile(f)
        ice.astype(np.float64).tofile(f)
    of, op = _run_both(tmp_path, nml, fin, ["--spinup"], ["3000.0", "1.0e-4", "0"])
    assert of.size == op.size == 3 * nphi * nlat + 2
    assert np.array_equal(of, op)
    assert of[-1] > 3000.0          # the spin-up advanced the clock before the three calls
