"""Host-side logic on the CPU: the &vilma namelist reader (vilma_par_load,
src/vilma_params.f90:209-331), the workload generators, and that the HIP
library loads and exports every entry point include/vilma_hip.h declares
(no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from fastearth3d_amd import configs, namelist
from fastearth3d_amd._abi import SEC_PER_YEAR, VhParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_namelist_overlay(tmp_path):
    f = tmp_path / "run.nml"
    f.write_text("""
&ctl
   file_forcing = "x.nc"
/
&vilma
   lmax = 96        ! comment
   earth = "custom", n_layer = 2
   r_bot = 3480.0e3, 0.0
   r_top = 6371.0e3, 3480.0e3
   rho = 4000.0, 10000.0 ; mu = 1.0d11, 0.0
   eta(1) = 1.0e21
   rheology = 1, 2
   n_elem_layer = 30, 10
   dt_max = 500.0     ! years
   rotation = .false.
   sle_subgrid = F
   deg1_frame = 'cf'
   visc_3d_file = "ignored.nc"
/
""".replace(";", ","))
    p = namelist.vilma_par_load(str(f))
    assert p.lmax == 96 and p.earth == b"custom" and p.n_layer == 2
    assert list(p.r_bot[:2]) == [3480.0e3, 0.0] and p.mu[0] == 1.0e11 and p.eta[0] == 1.0e21
    assert list(p.rheology[:2]) == [1, 2] and list(p.n_elem_layer[:2]) == [30, 10]
    assert p.dt_max == 500.0 * SEC_PER_YEAR
    assert p.rotation == 0 and p.sle_subgrid == 0 and p.deg1_frame == b"cf"
    assert p.sle_n_inner == 20 and p.scheme == b"fe"      # defaults kept


def test_namelist_defaults_file_and_validation(tmp_path):
    d = tmp_path / "defaults.nml"
    d.write_text("&vilma\n lmax = 32\n cfl = 0.5\n/\n")
    f = tmp_path / "run.nml"
    f.write_text("&vilma\n lmax = 48\n/\n")
    p = namelist.vilma_par_load(str(f), defaults_file=str(d))
    assert p.lmax == 48 and p.cfl == 0.5
    f.write_text("&vilma\n deg1_frame = 'xx'\n/\n")
    with pytest.raises(namelist.NamelistError):
        namelist.vilma_par_load(str(f))
    with pytest.raises(namelist.NamelistError):
        namelist.vilma_par_load(str(d), group="nothere")


def test_params_struct_matches_header():
    """ctypes layout == C layout of struct vh_params (size check through the
    oracle, which is compiled against include/vilma_hip.h)."""
    import oracle as O
    p = O.params_defaults()
    q = namelist.params_defaults()
    assert ctypes.sizeof(VhParams) > 0
    for name, _ in VhParams._fields_:
        a, b = getattr(p, name), getattr(q, name)
        if hasattr(a, "__len__") and not isinstance(a, bytes):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def test_library_exports_header_symbols():
    so = os.path.join(ROOT, "fastearth3d_amd", "libvilma_hip.so")
    if not os.path.exists(so):
        from fastearth3d_amd import build
        build.build()
    txt = open(os.path.join(ROOT, "include", "vilma_hip.h")).read()
    names = re.findall(r"^\s*(?:int|const char\*)\s+(vh_\w+)\s*\(", txt, re.M)
    assert len(names) >= 20
    lib = ctypes.CDLL(so)
    for n in names:
        assert hasattr(lib, n), n
    # the Python binding gives every entry point a ctypes signature (an
    # undeclared one would truncate the 64-bit handle to a C int)
    from fastearth3d_amd import lib as vlib
    L = vlib.load()
    for n in names:
        assert getattr(L, n).argtypes is not None, n


def test_workload_generators():
    c = configs.config(1, lmax=16)
    z = c["z_bed_eq"]
    assert abs((z < 0).mean() - 0.70) < 0.02
    assert np.abs(z).max() == pytest.approx(5000.0)
    assert configs.ice_history_factor(10.0) == pytest.approx(0.5)
    assert configs.ice_history_factor(22.0) == 1.0
    assert configs.ice_history_factor(40.0) == 0.0
    c2 = configs.config(2, lmax=16)
    pert = c2["pert"]
    depth = configs.element_midpoint_depths(c2["params"])
    assert pert.shape[0] == depth.size
    assert np.all(pert[depth < 80e3] == 0.0)
    top = np.argmin(np.where(depth > 80e3, depth, np.inf))
    assert pert[top].std() == pytest.approx(0.5, rel=1e-6)


def test_wrong_array_shapes_are_refused(oracle):
    """Arrays handed to entry points that read a fixed number of values through
    a pointer are size-checked before the call (the lateral field per element
    and per node, the restart memory and state)."""
    from fastearth3d_amd import configs
    from fastearth3d_amd.model import VilmaError
    c = configs.config(2, lmax=8)
    m = oracle.OracleModel(c["params"], c["z_bed_eq"], c["h_ice_eq"])
    ne, nr = m.dims["ne"], m.dims["nr"]
    with pytest.raises(VilmaError, match="pert_elem"):
        m.enable_visc_3d(np.zeros((ne - 1, m.nlat, m.nphi)))
    with pytest.raises(VilmaError, match="visc_node"):
        m.enable_visc_3d_nodes(np.zeros((ne, m.nlat, m.nphi)))
    with pytest.raises(VilmaError, match="tau"):
        m.set_memory(np.zeros(m.memory().size - 4))
    with pytest.raises(VilmaError, match="grid field"):
        m.update(np.zeros(m.nlat * m.nphi + 1), 10.0)
