"""The Python host of the Fortran driver's workload, run as its own process
(tests/test_fortran.py): same namelist, same input.bin / output.bin formats as
tests/fortran/drive.f90, through the ctypes front end of the same library.
Usage: python py_host.py <namelist> <input.bin> <output.bin> [--host-grid]
       [--spinup equil_time_max_yr equil_rate_tol pre_spinup_1d]"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from fastearth3d_amd import lib, namelist  # noqa: E402


def main():
    nml, fin, fout = sys.argv[1:4]
    host = "--host-grid" in sys.argv[4:]
    spin = None
    if "--spinup" in sys.argv[4:]:
        k = sys.argv.index("--spinup")
        spin = (float(sys.argv[k + 1]), float(sys.argv[k + 2]), int(sys.argv[k + 3]))
    q = namelist.vilma_par_load(nml, defaults_file=None)
    with open(fin, "rb") as f:
        nx, ny, nsteps = np.fromfile(f, np.int32, 3)
        dt = float(np.fromfile(f, np.float64, 1)[0])
        if host:
            lon = np.fromfile(f, np.float64, nx)
            lat = np.fromfile(f, np.float64, ny)
        z = np.fromfile(f, np.float64, nx * ny).reshape(ny, nx)
        ice = np.fromfile(f, np.float64, nx * ny * nsteps).reshape(nsteps, ny, nx)
    if host:
        nphi, nlat = q.grid()

        class _G:   # the model's Gauss grid dimensions, for the map
            pass
        g = _G()
        g.nphi, g.nlat = nphi, nlat
        rm = lib.Remap(g, lon, lat)
        se = lib.SolidEarth(q, rm.to_gauss(z), rm.to_gauss(np.zeros_like(z), conserve_mass=True))
        if spin:
            se.spinup(rm.to_gauss(ice[0], conserve_mass=True), *spin)
        for k in range(nsteps):
            se.update(rm.to_gauss(ice[k], conserve_mass=True), dt)
        rsl = rm.to_ll(se.rsl)
        zb = z - rsl
        rm.close()
    else:
        se = lib.SolidEarth(q, z, np.zeros_like(z))
        if spin:
            se.spinup(ice[0], *spin)
        for k in range(nsteps):
            se.update(ice[k], dt)
        rsl, zb = se.rsl, se.z_bed
    with open(fout, "wb") as f:
        rsl.astype(np.float64).tofile(f)
        zb.astype(np.float64).tofile(f)
        se.C.astype(np.float64).tofile(f)
        np.array([se.bsl, se.time], np.float64).tofile(f)
    se.finalize()


if __name__ == "__main__":
    main()
