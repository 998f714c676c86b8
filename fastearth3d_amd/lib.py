"""Loader of the HIP library and the Python `solid_earth` over it.

The library is the product: if libvilma_hip.so is missing the import of this
module's `load()` raises (there is no CPU fallback).
"""
import ctypes as C
import os

import numpy as np

from . import _abi
from ._abi import VhParams, dptr
from .model import ModelHandle, VilmaError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvilma_hip.so")
HEADER = os.path.join(HERE, "..", "include", "vilma_hip.h")
_lib = None


def load():
    """ctypes handle of libvilma_hip.so (built by fastearth3d_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise VilmaError("libvilma_hip.so is not built (run python -m fastearth3d_amd.build)")
        L = C.CDLL(LIB)
        _abi.declare_api(L, "vh_")
        L.vh_create.restype = C.c_int
        L.vh_create.argtypes = [C.POINTER(VhParams), C.c_int, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_void_p)]
        L.vh_destroy.argtypes = [C.c_void_p]
        L.vh_last_error.restype = C.c_char_p
        L.vh_last_error.argtypes = [C.c_void_p]
        L.vh_params_defaults.argtypes = [C.POINTER(VhParams)]
        L.vh_resp_apply.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.vh_resp_advance.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.vh_resp_set_dt.argtypes = [C.c_void_p, C.c_double]
        L.vh_sht_analysis.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_sht_synthesis.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_tensor_synthesis.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_tensor_analysis.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_tensor_synthesis_ch.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_tensor_analysis_ch.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_tensor_norms.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_get_drift_tor.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.vh_update_device.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        L.vh_profile.argtypes = [C.c_void_p, C.c_int]
        L.vh_profile_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_forward_record.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_double, C.c_int,
                                        C.POINTER(C.c_double)]
        L.vh_forward_record_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_void_p]
        L.vh_backward.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.vh_backward_device.argtypes = [C.c_void_p] + [C.c_void_p] * 3
        L.vh_forward_rewind.argtypes = [C.c_void_p]
        L.vh_adjoint_store_limit.argtypes = [C.c_void_p, C.c_double]
        L.vh_get_state.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 6
        L.vh_set_state.argtypes = [C.c_void_p, C.c_double] + [C.POINTER(C.c_double)] * 5
        L.vh_remap_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.c_void_p)]
        L.vh_remap_to_gauss.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.vh_remap_to_ll.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vh_remap_to_gauss_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        L.vh_remap_to_gauss_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.vh_remap_to_ll_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.vh_remap_destroy.argtypes = [C.c_void_p]
        L.vh_spinup.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double, C.c_double, C.c_int]
        L.vh_set_memory_lowres.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.vh_get_step_state.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.vh_set_step_state.argtypes = [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.vh_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.vh_get_tuning.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        L.vh_ensemble_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.vh_ensemble_update.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_double)]
        L.vh_ensemble_update_device.argtypes = [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_double)]
        L.vh_ensemble_run_device.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_double)]
        L.vh_ensemble_tune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int)]
        L.vh_ensemble_failed_member.argtypes = [C.c_void_p]
        L.vh_ensemble_last_error.argtypes = [C.c_void_p]
        L.vh_ensemble_last_error.restype = C.c_char_p
        L.vh_ensemble_destroy.argtypes = [C.c_void_p]
        for name in exported_symbols():
            fn = getattr(L, name)
            if fn.argtypes is None:   # every entry point takes a handle or pointer: never call undeclared
                raise VilmaError("libvilma_hip.so: %s has no ctypes signature" % name)
            if fn.restype is C.c_int or name in ("vh_last_error", "vh_ensemble_last_error"):
                continue
            fn.restype = C.c_int
        _lib = L
    return _lib


def exported_symbols():
    """Names of the entry points declared in include/vilma_hip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(vh_\w+)\s*\(", txt, re.M)))


def _cp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class SolidEarth(ModelHandle):
    """solid_earth on the GPU: SolidEarth(par, z_bed_eq, h_ice_eq[, h_ice_init])
    is solid_earth_init; .update(h_ice, dt_yr) is solid_earth_update;
    .finalize() is solid_earth_finalize."""

    def __init__(self, params, z_bed_eq, h_ice_eq, h_ice_init=None):
        L = load()
        nphi, nlat = params.grid()
        z = np.ascontiguousarray(z_bed_eq, dtype=np.float64)
        h = np.ascontiguousarray(h_ice_eq, dtype=np.float64)
        hi = None if h_ice_init is None else np.ascontiguousarray(h_ice_init, dtype=np.float64)
        for a in (z, h) + ((hi,) if hi is not None else ()):
            if a.size != nphi * nlat:
                raise VilmaError("solid_earth_init: grid fields must have nphi*nlat = %d values" % (nphi * nlat))
        out = C.c_void_p()
        rc = L.vh_create(C.byref(params), nphi, nlat, dptr(z), dptr(h), dptr(hi), C.byref(out))
        if rc != 0:
            raise VilmaError(L.vh_last_error(None).decode())
        super().__init__(L, "vh_", out.value, params)

    def sht_analysis(self, f):
        q = np.empty(self.dims["nlm"], dtype=np.complex128)
        self._check(self._lib.vh_sht_analysis(self._h, _cp(np.ascontiguousarray(f, dtype=np.float64)), _cp(q)),
                    "sht_analysis")
        return q

    def sht_synthesis(self, q):
        f = np.empty((self.nlat, self.nphi))
        self._check(self._lib.vh_sht_synthesis(self._h, _cp(np.ascontiguousarray(q, dtype=np.complex128)), _cp(f)),
                    "sht_synthesis")
        return f

    def tensor_synth(self, c):
        """tensor_sh_synth: c (4 or 6, nlm) complex LM order (lambda 1, 2, 5, 6
        and, with six rows, 3, 4) -> dyads (6, nlat, nphi)."""
        c = np.ascontiguousarray(c, dtype=np.complex128)
        if c.ndim != 2 or c.shape[0] not in (4, 6) or c.shape[1] != self.dims["nlm"]:
            raise VilmaError("tensor_sh_synth: c must be (4 or 6, nlm)")
        d = np.empty((6, self.nlat, self.nphi))
        self._check(self._lib.vh_tensor_synthesis_ch(self._h, c.shape[0], _cp(c), _cp(d)), "tensor_sh_synth")
        return d

    def tensor_analysis(self, d, nch=4):
        """tensor_sh_analysis: dyads (6, nlat, nphi) -> c (nch, nlm) complex."""
        d = self._sized(d, 6 * self.nlat * self.nphi, "tensor_sh_analysis: dyads must be (6, nlat, nphi)")
        c = np.empty((nch, self.dims["nlm"]), dtype=np.complex128)
        self._check(self._lib.vh_tensor_analysis_ch(self._h, int(nch), _cp(d), _cp(c)), "tensor_sh_analysis")
        return c

    def tensor_norms(self):
        """(n6, n4) per degree 0..lmax: the spin-2 normalisations."""
        n6 = np.empty(self.dims["lmax"] + 1); n4 = np.empty_like(n6)
        self._check(self._lib.vh_tensor_norms(self._h, _cp(n6), _cp(n4)), "tensor_sh_init")
        return n6, n4

    def drift_tor(self):
        """Nodal toroidal W drift (2, nk, nr): re, im (zero unless carried)."""
        d = self.dims
        out = np.empty((2, d["nk"], d["nr"]))
        self._check(self._lib.vh_get_drift_tor(self._h, _cp(out)), "get_drift_tor")
        return out

    def apply(self, sigma):
        u = np.empty(self.dims["nlm"], dtype=np.complex128)
        n = np.empty_like(u)
        self._check(self._lib.vh_resp_apply(self._h, _cp(np.ascontiguousarray(sigma, dtype=np.complex128)),
                                            _cp(u), _cp(n)), "response_apply")
        return u, n

    def advance(self, sigma):
        self._check(self._lib.vh_resp_advance(self._h, _cp(np.ascontiguousarray(sigma, dtype=np.complex128))),
                    "response_advance_endpoint")

    def set_dt(self, dt_s):
        self._check(self._lib.vh_resp_set_dt(self._h, float(dt_s)), "response_set_dt")

    def state(self):
        """The restart state besides the memory: time (yr), rsl and h_ice
        (nlat, nphi), rotation m (re, im) and the load / tidal channel memory
        (6, ne, 4) (C order of the reference's (NLAM, ne, 6))."""
        ne = self.dims["ne"]
        t = C.c_double()
        rsl = np.empty((self.nlat, self.nphi)); hi = np.empty_like(rsl)
        m = np.empty(2); lm = np.empty((6, ne, 4)); tm = np.empty((6, ne, 4))
        self._check(self._lib.vh_get_state(self._h, C.byref(t), _cp(rsl), _cp(hi), _cp(m), _cp(lm), _cp(tm)),
                    "get_state")
        return dict(time=t.value, rsl=rsl, h_ice=hi, rot_m=m, rot_load_mem=lm, rot_tidal_mem=tm)

    def set_state(self, st):
        ng, ne = self.nphi * self.nlat, self.dims["ne"]
        n = dict(rsl=ng, h_ice=ng, rot_m=2, rot_load_mem=24 * ne, rot_tidal_mem=24 * ne)
        arr = {k: self._sized(st[k], n[k], "set_state: %s" % k) for k in n}
        self._check(self._lib.vh_set_state(self._h, float(st["time"]), _cp(arr["rsl"]), _cp(arr["h_ice"]),
                                           _cp(arr["rot_m"]), _cp(arr["rot_load_mem"]), _cp(arr["rot_tidal_mem"])),
                    "set_state")

    def set_memory_lowres(self, tau):
        """The memory of a run at a lower truncation, (6, nk_src, ne, nlam) as
        memory() returns it, into the low degrees (the rest zero; the
        cross-resolution restart of src/vilma_io.f90:515-552)."""
        t = np.ascontiguousarray(tau, dtype=np.float64)
        d = self.dims
        if t.ndim != 4 or t.shape[0] != 6 or t.shape[2] != d["ne"] or t.shape[3] != d["nlam"]:
            raise VilmaError("set_memory_lowres: tau must be (6, nk_src, ne, nlam)")
        self._check(self._lib.vh_set_memory_lowres(self._h, _cp(t), int(t.shape[1])), "set_memory_lowres")

    def step_state(self):
        """dt_try (yr), sigma_primed and sigma_n (LM order) of the adaptive
        controller / trapezoidal scheme."""
        dt, pr = C.c_double(), C.c_int()
        sig = np.empty(self.dims["nlm"], dtype=np.complex128)
        self._check(self._lib.vh_get_step_state(self._h, C.byref(dt), C.byref(pr),
                                                sig.ctypes.data_as(C.POINTER(C.c_double))), "get_step_state")
        return dict(dt_try=dt.value, sigma_primed=bool(pr.value), sigma_n=sig)

    def set_step_state(self, st):
        sig = None if st.get("sigma_n") is None else np.ascontiguousarray(st["sigma_n"], dtype=np.complex128)
        self._check(self._lib.vh_set_step_state(self._h, float(st["dt_try"]), int(bool(st["sigma_primed"])),
                                                None if sig is None else sig.ctypes.data_as(C.POINTER(C.c_double))),
                    "set_step_state")

    def spinup(self, h_ice_lgm, equil_time_max_yr, equil_rate_tol=1.0e-3, pre_spinup_1d=False):
        """solid_earth_spinup (src/vilma_coupling.f90:453-517): relax under the
        held start-slice ice h_ice_lgm (model grid), optionally first with the
        lateral viscosity replaced by its lateral mean, then for at most
        equil_time_max_yr years, until the bed velocity < equil_rate_tol m/yr."""
        h = self._grid_in(h_ice_lgm)
        self._check(self._lib.vh_spinup(self._h, dptr(h), float(equil_time_max_yr), float(equil_rate_tol),
                                        int(bool(pre_spinup_1d))), "solid_earth_spinup")

    def remap(self, lon_deg, lat_deg):
        """Map pair between a host lon-lat grid (cell-centre axes in degrees,
        ascending) and this model's Gauss grid (remap_init,
        src/vilma_remap.f90:56-99)."""
        return Remap(self, lon_deg, lat_deg)

    def forward_record(self, ice_hist, dt_yr, ckpt_every=10):
        """Run len(ice_hist) coupling calls keeping checkpoints; returns the
        rsl after each call, (ncalls, nlat, nphi)."""
        ice = np.ascontiguousarray(ice_hist, dtype=np.float64)
        n = ice.shape[0]
        if ice.size != n * self.nphi * self.nlat:
            raise VilmaError("forward_record: ice history must be (ncalls, nlat, nphi)")
        rsl = np.empty((n, self.nlat, self.nphi))
        self._check(self._lib.vh_forward_record(self._h, n, _cp(ice), float(dt_yr), int(ckpt_every), _cp(rsl)),
                    "forward_record")
        self._ncalls = n
        return rsl

    def backward(self, dJ_drsl):
        """Adjoint sweep of the last forward_record: returns (dJ_dlogeta
        (ne, nlat, nphi), dJ_dice (ncalls+1, nlat, nphi))."""
        g = np.ascontiguousarray(dJ_drsl, dtype=np.float64)
        n = getattr(self, "_ncalls", 0)
        if g.size != n * self.nphi * self.nlat:
            raise VilmaError("backward: dJ_drsl must be (ncalls, nlat, nphi) of the last forward_record")
        dp = np.empty((self.dims["ne"], self.nlat, self.nphi))
        di = np.empty((n + 1, self.nlat, self.nphi))
        self._check(self._lib.vh_backward(self._h, _cp(g), _cp(dp), _cp(di)), "backward")
        return dp, di

    def store_limit(self, gib):
        """HBM (GiB) the record pass may keep for the backward sweep."""
        self._check(self._lib.vh_adjoint_store_limit(self._h, float(gib)), "adjoint_store_limit")

    def rewind(self):
        """Back to the state at the start of the last forward_record."""
        self._check(self._lib.vh_forward_rewind(self._h), "forward_rewind")


def solid_earth_init(params, z_bed_eq, h_ice_eq, h_ice_init=None):
    return SolidEarth(params, z_bed_eq, h_ice_eq, h_ice_init)


def solid_earth_update(se, h_ice, dt_yr):
    se.update(h_ice, dt_yr)


def solid_earth_finalize(se):
    se.finalize()


class Ensemble:
    """An uneven ensemble of SolidEarth members (vh_ensemble_*): a pool of
    n_workers host threads steps the members, longest first, each on its own
    stream; results equal stepping them one after another. The members stay
    owned by the caller (finalize them after close())."""

    def __init__(self, members, n_workers=0):
        self._lib = load()
        self.members = list(members)
        n = len(self.members)
        hs = (C.c_void_p * n)(*[m._h for m in self.members])
        out = C.c_void_p()
        if self._lib.vh_ensemble_create(hs, n, int(n_workers), C.byref(out)) != 0:
            raise VilmaError("vh_ensemble_create: invalid member list")
        self._e = out.value
        self._dt = C.c_double * n

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.vh_ensemble_last_error(self._e)
            raise VilmaError("%s: %s" % (what, msg.decode() if msg else "failed"))

    def failed_member(self):
        return self._lib.vh_ensemble_failed_member(self._e)

    def update(self, h_ice, dt_yr):
        """One solid_earth_update of every member: h_ice[i] the (nlat, nphi)
        grid of member i, dt_yr a number or one per member."""
        n = len(self.members)
        arrs = [m._grid_in(a) for m, a in zip(self.members, h_ice)]
        if len(arrs) != n:
            raise VilmaError("ensemble update: %d ice fields for %d members" % (len(arrs), n))
        dts = self._dt(*([float(dt_yr)] * n if np.isscalar(dt_yr) else [float(x) for x in dt_yr]))
        ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
        self._check(self._lib.vh_ensemble_update(self._e, ptrs, dts), "ensemble update")

    def update_device(self, ptrs, dt_yr):
        """The same with device pointers (one per member)."""
        n = len(self.members)
        dts = self._dt(*([float(dt_yr)] * n if np.isscalar(dt_yr) else [float(x) for x in dt_yr]))
        self._check(self._lib.vh_ensemble_update_device(self._e, (C.c_void_p * n)(*ptrs), dts),
                    "ensemble update")

    def run_device(self, ncalls, hist_ptrs, dt_yr):
        """ncalls updates of every member; hist_ptrs[i] points to member i's
        [ncalls][nlat][nphi] ice history in device memory."""
        n = len(self.members)
        dts = self._dt(*([float(dt_yr)] * n if np.isscalar(dt_yr) else [float(x) for x in dt_yr]))
        self._check(self._lib.vh_ensemble_run_device(self._e, int(ncalls), (C.c_void_p * n)(*hist_ptrs), dts),
                    "ensemble run")

    def tune(self, ncalls, hist_ptrs, dt_yr):
        """The autotuner (vh_ensemble_tune): 3 x ncalls real updates of every
        member from hist_ptrs (as run_device); returns the chosen
        VH_T_SLE_BLOCKS, now set on every member."""
        n = len(self.members)
        dts = self._dt(*([float(dt_yr)] * n if np.isscalar(dt_yr) else [float(x) for x in dt_yr]))
        out = C.c_int(0)
        self._check(self._lib.vh_ensemble_tune(self._e, int(ncalls), (C.c_void_p * n)(*hist_ptrs), dts,
                                               C.byref(out)), "ensemble tune")
        return out.value

    def close(self):
        if self._e:
            self._lib.vh_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Remap:
    """Host lon-lat grid <-> Gauss grid (src/vilma_remap.f90): conservative
    to_gauss, bilinear to_ll. Host fields are (nlat_ll, nlon) numpy arrays
    (longitude fastest, latitude ascending); Gauss fields (nlat, nphi) in the
    model layout (north-first)."""

    def __init__(self, model, lon_deg, lat_deg):
        self._lib = load()
        self.nphi, self.nlat = model.nphi, model.nlat
        lon = np.ascontiguousarray(lon_deg, dtype=np.float64)
        lat = np.ascontiguousarray(lat_deg, dtype=np.float64)
        self.nlon, self.nlat_ll = lon.size, lat.size
        h = C.c_void_p()
        self._check(self._lib.vh_remap_create(self.nphi, self.nlat, self.nlon, self.nlat_ll, _cp(lon), _cp(lat),
                                              C.byref(h)), "remap_init")
        self._r = h

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.vh_last_error(None)
            raise VilmaError("%s: %s" % (what, msg.decode() if msg else "failed"))

    def to_gauss(self, f_ll, conserve_mass=False):
        """remap_to_gauss (src/vilma_remap.f90:101-138)."""
        f = np.ascontiguousarray(f_ll, dtype=np.float64)
        if f.size != self.nlon * self.nlat_ll:
            raise VilmaError("remap_to_gauss: source field shape /= map source grid")
        g = np.empty((self.nlat, self.nphi))
        self._check(self._lib.vh_remap_to_gauss(self._r, _cp(f), _cp(g), int(bool(conserve_mass))), "remap_to_gauss")
        return g

    def to_gauss_noalloc(self, f_ll, conserve_mass=False):
        """remap_to_gauss through the allocation-free entry point
        (vh_remap_to_gauss_noalloc): the same result, staged in the library's
        fixed per-GPU array."""
        f = np.ascontiguousarray(f_ll, dtype=np.float64)
        if f.size != self.nlon * self.nlat_ll:
            raise VilmaError("remap_to_gauss_noalloc: source field shape /= map source grid")
        g = np.empty((self.nlat, self.nphi))
        self._check(self._lib.vh_remap_to_gauss_noalloc(self._r, _cp(f), _cp(g), int(bool(conserve_mass))),
                    "remap_to_gauss_noalloc")
        return g

    def to_ll(self, f_gauss):
        """remap_to_ll (src/vilma_remap.f90:140-164)."""
        g = np.ascontiguousarray(f_gauss, dtype=np.float64)
        if g.size != self.nlat * self.nphi:
            raise VilmaError("remap_to_ll: source field shape /= Gauss grid")
        f = np.empty((self.nlat_ll, self.nlon))
        self._check(self._lib.vh_remap_to_ll(self._r, _cp(g), _cp(f)), "remap_to_ll")
        return f

    def close(self):
        if self._r:
            self._lib.vh_remap_destroy(self._r)
            self._r = None
