"""Python mirror of the reference's `solid_earth` host API over the C-ABI.

`SolidEarth` keeps the reference's names and argument meaning:
`solid_earth_init(se, z_bed_eq, h_ice_eq[, h_ice_init])`
(src/vilma_coupling.f90:144-232), `solid_earth_update(se, h_ice, dt_yr)`
(:336-437), `solid_earth_finalize(se)` (:599-626), and exposes the outputs a
host reads (`se%rsl`, `se%z_bed`, `se%bsl`, `se%time`, `se%worst_mass_resid`,
`se%sigma_lm`, `se%gg%C` ...) as properties. Grid fields are numpy arrays of
shape (nlat, nphi) in C order, i.e. the reference's (nphi, nlat) column-major
layout with phi fastest. Failures raise VilmaError with the library's message
(the reference aborts with `error stop` instead).
"""
import ctypes as C

import numpy as np

from . import _abi
from ._abi import dptr


class VilmaError(RuntimeError):
    pass


class ModelHandle:
    """Thin wrapper of one model handle of a library exposing the model
    entry points under `prefix` (vh_ for the HIP library)."""

    def __init__(self, lib, prefix, handle, params):
        self._lib, self._pfx, self._h, self.par = lib, prefix, handle, params
        self.nphi, self.nlat = params.grid()
        self.dims = {k: self._dim(v) for k, v in (
            ("lmax", _abi.D_LMAX), ("nlm", _abi.D_NLM), ("nr", _abi.D_NR), ("ne", _abi.D_NE),
            ("nk", _abi.D_NK), ("nlam", _abi.D_NLAM), ("nem", _abi.D_NEM), ("nmn", _abi.D_NMN))}

    # ---------------------------------------------------------------- plumbing
    def _fn(self, name):
        return getattr(self._lib, self._pfx + name)

    def _err(self):
        fn = getattr(self._lib, self._pfx + "last_error")
        fn.restype = C.c_char_p
        fn.argtypes = [C.c_void_p]
        msg = fn(self._h)
        return msg.decode() if msg else "unknown error"

    def _check(self, rc, what):
        if rc != 0:
            raise VilmaError("%s: %s" % (what, self._err()))

    def _dim(self, which):
        v = C.c_int(0)
        self._check(self._fn("get_dims")(self._h, which, C.byref(v)), "get_dims")
        return v.value

    def _grid_in(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != self.nphi * self.nlat:
            raise VilmaError("grid field has %d values, expected nphi*nlat = %d" % (a.size, self.nphi * self.nlat))
        return a

    # -------------------------------------------------------------- host API
    def update(self, h_ice, dt_yr):
        h = self._grid_in(h_ice)
        self._check(self._fn("update")(self._h, dptr(h), float(dt_yr)), "solid_earth_update")

    def set_tuning(self, which, value):
        """A launch-shape knob (vh_set_tuning; results do not depend on it)."""
        self._check(self._fn("set_tuning")(self._h, int(which), int(value)), "set_tuning")

    def update_device(self, ptr, dt_yr):
        """solid_earth_update with the ice grid at device address ptr (the
        library's vh_update_device; not offered by the oracle)."""
        self._check(self._fn("update_device")(self._h, C.c_void_p(int(ptr)), float(dt_yr)), "solid_earth_update")

    def _sized(self, a, n, what):
        """a as a C-contiguous float64 array of exactly n values (the library
        reads n values through the pointer)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != n:
            raise VilmaError("%s: array has %d values, expected %d" % (what, a.size, n))
        return a

    def enable_visc_3d(self, pert_elem):
        p = self._sized(pert_elem, self.dims["ne"] * self.nphi * self.nlat,
                        "enable_lateral_visc: pert_elem must be (ne, nlat, nphi)")
        self._check(self._fn("enable_visc_3d")(self._h, dptr(p)), "enable_lateral_visc")
        self.dims["ne3d"] = self._dim(_abi.D_NE3D)
        self.dims["nlam"] = self._dim(_abi.D_NLAM)   # 6 once the toroidal channels are carried

    def enable_visc_3d_nodes(self, visc_node, lid_depth=0.0, lid_log10max=22.0, log10_cap=30.0):
        v = self._sized(visc_node, self.dims["nr"] * self.nphi * self.nlat,
                        "enable_lateral_visc_from_nodes: visc_node must be (nr, nlat, nphi)")
        self._check(self._fn("enable_visc_3d_nodes")(self._h, dptr(v), lid_depth, lid_log10max, log10_cap),
                    "enable_lateral_visc_from_nodes")
        self.dims["ne3d"] = self._dim(_abi.D_NE3D)
        self.dims["nlam"] = self._dim(_abi.D_NLAM)

    def field(self, which):
        out = np.empty((self.nlat, self.nphi))
        self._check(self._fn("get_field")(self._h, which, dptr(out)), "get_field")
        return out

    def scalar(self, which):
        v = C.c_double(0.0)
        self._check(self._fn("get_scalar")(self._h, which, C.byref(v)), "get_scalar")
        return v.value

    def spectrum(self, which):
        out = np.empty(self.dims["nlm"], dtype=np.complex128)
        self._check(self._fn("get_spectrum")(self._h, which, out.ctypes.data_as(C.POINTER(C.c_double))), "get_spectrum")
        return out

    def memory(self):
        d = self.dims
        out = np.empty((6, d["nk"], d["ne"], d["nlam"]))
        self._check(self._fn("get_memory")(self._h, dptr(out)), "get_memory")
        return out

    def set_memory(self, tau):
        d = self.dims
        t = self._sized(tau, 6 * d["nk"] * d["ne"] * d["nlam"], "set_memory: tau must be (6, nk, ne, nlam)")
        self._check(self._fn("set_memory")(self._h, dptr(t)), "set_memory")

    # ------------------------------------------------------- response probes
    def begin_step(self):
        self._check(self._fn("resp_begin_step")(self._h), "response_begin_step")

    def drift(self):
        d = self.dims
        surf = np.empty(3 * d["nk"], dtype=np.complex128)
        nodal = np.empty((4, d["nk"], d["nr"]))
        self._check(self._fn("get_drift")(self._h, surf.ctypes.data_as(C.POINTER(C.c_double)), dptr(nodal)), "get_drift")
        return surf.reshape(3, d["nk"]), nodal

    def gains(self):
        d = self.dims
        g3 = np.empty((3, d["lmax"] + 1))
        xn = np.empty((2, d["lmax"], d["nr"]))
        self._check(self._fn("get_gains")(self._h, dptr(g3), dptr(xn)), "get_gains")
        return g3, xn

    # ------------------------------------------------ reference-named views
    @property
    def rsl(self):
        return self.field(_abi.F_RSL)

    @property
    def z_bed(self):
        return self.field(_abi.F_ZBED)

    @property
    def C(self):
        return self.field(_abi.F_C)

    @property
    def C0(self):
        return self.field(_abi.F_C0)

    @property
    def s_rot(self):
        return self.field(_abi.F_SROT)

    @property
    def time(self):
        return self.scalar(_abi.S_TIME)

    @property
    def bsl(self):
        return self.scalar(_abi.S_BSL)

    @property
    def worst_mass_resid(self):
        return self.scalar(_abi.S_WORST_MASS_RESID)

    @property
    def sigma_lm(self):
        return self.spectrum(_abi.SP_SIGMA)

    def finalize(self):
        if self._h:
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass
