"""ctypes front end of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: the oracle is the correctness checker for the HIP
product path and the CPU baseline of bench.py. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it; the
product package never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from fastearth3d_amd import _abi
from fastearth3d_amd._abi import VhParams, dptr, iptr
from fastearth3d_amd.model import ModelHandle, VilmaError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        _abi.declare_api(L, "orc_")
        L.orc_create.restype = C.c_int
        L.orc_create.argtypes = [C.POINTER(VhParams), C.c_int, C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_void_p), C.c_char_p]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_params_defaults.argtypes = [C.POINTER(VhParams)]
        for n in ("orc_sht_analysis_c", "orc_sht_synthesis_c"):
            getattr(L, n).argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        for n in ("orc_tensor_synth_c", "orc_tensor_analysis_c"):
            getattr(L, n).argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
        for n in ("orc_tensor_n6_c", "orc_tensor_n4_c", "orc_rot_info_c"):
            getattr(L, n).argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_tor_synthesis_c.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_tor_analysis_c.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_mesh_props_c.argtypes = [C.POINTER(VhParams), C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int,
                                       C.c_char_p]
        L.orc_tor_forcing_c.argtypes = [C.POINTER(VhParams), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.c_char_p]
        L.orc_tor_operator_c.argtypes = [C.POINTER(VhParams), C.c_int, C.POINTER(C.c_double), C.c_long,
                                         C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p]
        L.orc_sph_synthesis_c.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_sph_analysis_c.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_surface_integral_c.restype = C.c_double
        L.orc_surface_integral_c.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_grid_c.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_resp_apply_c.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_resp_advance_c.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.orc_resp_set_dt_c.argtypes = [C.c_void_p, C.c_double]
        L.orc_test_band.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_love.argtypes = [C.POINTER(VhParams), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_char_p]
        L.orc_dense_operator.argtypes = [C.POINTER(VhParams), C.c_int, C.POINTER(C.c_double), C.c_long, C.c_char_p]
        L.orc_mesh_c.argtypes = [C.POINTER(VhParams), C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_char_p]
        L.orc_elem_integrals.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.orc_ocean_function_c.argtypes = [C.c_int] + [C.POINTER(C.c_double)] * 3
        L.orc_rotation_run.argtypes = [C.POINTER(VhParams), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                       C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.c_char_p]
        _lib = L
    return _lib


def _cp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleModel(ModelHandle):
    """The oracle's solid_earth (same methods as the HIP model handle)."""

    def __init__(self, params, z_bed_eq, h_ice_eq, h_ice_init=None):
        L = lib()
        nphi, nlat = params.grid()
        z = np.ascontiguousarray(z_bed_eq, dtype=np.float64)
        h = np.ascontiguousarray(h_ice_eq, dtype=np.float64)
        hi = None if h_ice_init is None else np.ascontiguousarray(h_ice_init, dtype=np.float64)
        out = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = L.orc_create(C.byref(params), nphi, nlat, dptr(z), dptr(h), dptr(hi), C.byref(out), err)
        if rc != 0:
            L.orc_destroy(out)
            raise VilmaError("solid_earth_init: " + err.value.decode())
        super().__init__(L, "orc_", out.value, params)

    # transforms on the model grid
    def sht_analysis(self, f):
        q = np.empty(self.dims["nlm"], dtype=np.complex128)
        self._lib.orc_sht_analysis_c(self._h, _cp(np.ascontiguousarray(f, dtype=np.float64)), _cp(q))
        return q

    def sht_synthesis(self, q):
        f = np.empty((self.nlat, self.nphi))
        self._lib.orc_sht_synthesis_c(self._h, _cp(np.ascontiguousarray(q, dtype=np.complex128)), _cp(f))
        return f

    def sph_synthesis(self, q):
        vt = np.empty((self.nlat, self.nphi)); vp = np.empty((self.nlat, self.nphi))
        self._lib.orc_sph_synthesis_c(self._h, _cp(np.ascontiguousarray(q, dtype=np.complex128)), _cp(vt), _cp(vp))
        return vt, vp

    def sph_analysis(self, vt, vp):
        q = np.empty(self.dims["nlm"], dtype=np.complex128)
        self._lib.orc_sph_analysis_c(self._h, _cp(np.ascontiguousarray(vt, dtype=np.float64)),
                                     _cp(np.ascontiguousarray(vp, dtype=np.float64)), _cp(q))
        return q

    def surface_integral(self, f):
        return self._lib.orc_surface_integral_c(self._h, _cp(np.ascontiguousarray(f, dtype=np.float64)))

    def grid_geometry(self):
        colat = np.empty(self.nlat); w = np.empty(self.nlat)
        self._lib.orc_grid_c(self._h, _cp(colat), _cp(w))
        lon = 2 * np.pi * np.arange(self.nphi) / self.nphi
        return colat, lon, w

    def tensor_synth(self, c):
        """c (4 or 6, nlm): channels lambda 1, 2, 5, 6 (and 3, 4) -> dyads (6, nlat, nphi)."""
        c = np.ascontiguousarray(c, dtype=np.complex128)
        d = np.empty((6, self.nlat, self.nphi))
        self._lib.orc_tensor_synth_c(self._h, _cp(c), c.shape[0], _cp(d))
        return d

    def tensor_analysis(self, d, nch=4):
        c = np.empty((nch, self.dims["nlm"]), dtype=np.complex128)
        self._lib.orc_tensor_analysis_c(self._h, _cp(np.ascontiguousarray(d, dtype=np.float64)), nch, _cp(c))
        return c

    def tensor_n6(self):
        n6 = np.empty(self.dims["lmax"] + 1)
        self._lib.orc_tensor_n6_c(self._h, _cp(n6))
        return n6

    def tensor_n4(self):
        n4 = np.empty(self.dims["lmax"] + 1)
        self._lib.orc_tensor_n4_c(self._h, _cp(n4))
        return n4

    def tor_synthesis(self, q):
        vt = np.empty((self.nlat, self.nphi)); vp = np.empty((self.nlat, self.nphi))
        self._lib.orc_tor_synthesis_c(self._h, _cp(np.ascontiguousarray(q, dtype=np.complex128)), _cp(vt), _cp(vp))
        return vt, vp

    def tor_analysis(self, vt, vp):
        q = np.empty(self.dims["nlm"], dtype=np.complex128)
        self._lib.orc_tor_analysis_c(self._h, _cp(np.ascontiguousarray(vt, dtype=np.float64)),
                                     _cp(np.ascontiguousarray(vp, dtype=np.float64)), _cp(q))
        return q

    def rot_info(self):
        o = np.empty(7)
        self._lib.orc_rot_info_c(self._h, _cp(o))
        return dict(kTe=o[0], hTe=o[1], k_s_fluid=o[2], k_s_flat=o[3], dt_fe_max=o[4], m=o[5] + 1j * o[6])

    def apply(self, sigma):
        u = np.empty(self.dims["nlm"], dtype=np.complex128); n = np.empty_like(u)
        self._lib.orc_resp_apply_c(self._h, _cp(np.ascontiguousarray(sigma, dtype=np.complex128)), _cp(u), _cp(n))
        return u, n

    def advance(self, sigma):
        self._lib.orc_resp_advance_c(self._h, _cp(np.ascontiguousarray(sigma, dtype=np.complex128)))

    def set_dt(self, dt_s):
        self._lib.orc_resp_set_dt_c(self._h, float(dt_s))


def params_defaults():
    p = VhParams()
    lib().orc_params_defaults(C.byref(p))
    return p


def band_solve_dense(A, b):
    """Pivoted band LU of the nonzero pattern of dense A, solve A x = b."""
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    x = np.empty(n)
    klku = np.zeros(2, dtype=np.int32)
    rc = lib().orc_test_band(n, _cp(A), _cp(np.ascontiguousarray(b, dtype=np.float64)), _cp(x), iptr(klku))
    if rc:
        raise VilmaError("band LU factorization failed")
    return x, tuple(klku)


def love(params, j, tidal=False, fluidize=False):
    out = np.empty(3)
    err = C.create_string_buffer(512)
    if lib().orc_love(C.byref(params), j, int(tidal), int(fluidize), _cp(out), err):
        raise VilmaError(err.value.decode())
    return out


def rotation_run(params, lmax, nlat, nphi, load, dt_s, nsteps):
    """The rotation channels alone under a fixed load (the reference's
    tests/test_rotation.f90:73-117 driving). Returns (m[nsteps+1] complex,
    kTe, k_s, rigid (2,1) inertia complex)."""
    load = np.ascontiguousarray(load, dtype=np.float64)
    assert load.size == nlat * nphi
    mo = np.empty(2 * (nsteps + 1))
    info = np.empty(4)
    err = C.create_string_buffer(512)
    if lib().orc_rotation_run(C.byref(params), lmax, nlat, nphi, _cp(load), dt_s, nsteps, _cp(mo), _cp(info), err):
        raise VilmaError(err.value.decode())
    return mo[0::2] + 1j * mo[1::2], info[0], info[1], info[2] + 1j * info[3]


def dense_operator(params, j):
    err = C.create_string_buffer(512)
    nd = lib().orc_dense_operator(C.byref(params), j, None, 0, err)
    A = np.zeros((nd, nd), order="F")
    lib().orc_dense_operator(C.byref(params), j, _cp(A), nd * nd, err)
    return A


def tor_operator(params, j, b=None):
    """The degree-j toroidal W stiffness (nr x nr, before pinning and
    bordering) and, with b, the assembled solve A^-1 b."""
    err = C.create_string_buffer(512)
    nr = lib().orc_tor_operator_c(C.byref(params), j, None, 0, None, None, err)
    if nr < 0:
        raise VilmaError(err.value.decode())
    A = np.zeros((nr, nr), order="F")
    x = None
    if b is None:
        lib().orc_tor_operator_c(C.byref(params), j, _cp(A), nr * nr, None, None, err)
        return A
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.empty(nr)
    if lib().orc_tor_operator_c(C.byref(params), j, _cp(A), nr * nr, _cp(bb), _cp(x), err) < 0:
        raise VilmaError(err.value.decode())
    return A, x


def mesh_props(params):
    """Node radii, per-element shear modulus and rheology (0 elastic,
    1 Maxwell, 2 fluid) of the params' mesh."""
    r, _ = mesh(params)
    err = C.create_string_buffer(512)
    mu = np.empty(r.size - 1); rh = np.empty(r.size - 1, dtype=np.int32)
    lib().orc_mesh_props_c(C.byref(params), _cp(mu), iptr(rh), r.size - 1, err)
    return r, mu, rh


def tor_forcing_of_w(params, j, W):
    """The toroidal dissipative forcing of the memory 2 mu eps(W)."""
    err = C.create_string_buffer(512)
    Wc = np.ascontiguousarray(W, dtype=np.float64)
    f = np.empty(Wc.size)
    if lib().orc_tor_forcing_c(C.byref(params), j, _cp(Wc), _cp(f), err) < 0:
        raise VilmaError(err.value.decode())
    return f


def mesh(params):
    err = C.create_string_buffer(512)
    nr = lib().orc_mesh_c(C.byref(params), None, None, 0, err)
    if nr < 0:
        raise VilmaError(err.value.decode())
    r = np.empty(nr); lay = np.empty(nr - 1, dtype=np.int32)
    lib().orc_mesh_c(C.byref(params), _cp(r), iptr(lay), nr, err)
    return r, lay


def elem_integrals(rk, rk1):
    out = np.empty(34)
    lib().orc_elem_integrals(rk, rk1, _cp(out))
    return out


def ocean_function(topo, ice):
    t = np.ascontiguousarray(topo, dtype=np.float64); i = np.ascontiguousarray(ice, dtype=np.float64)
    C_ = np.empty_like(t)
    lib().orc_ocean_function_c(t.size, _cp(t), _cp(i), _cp(C_))
    return C_
