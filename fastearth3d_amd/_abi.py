"""ctypes mirror of include/vilma_hip.h: the vh_params record and the
entry-point signatures shared by every Python caller of the C-ABI."""
import ctypes as C

VH_MAX_LAYER = 16
SEC_PER_YEAR = 31556926.0  # src/vilma_constants.f90:34


class VhParams(C.Structure):
    """The `&vilma` physics/numerics record (vilma_param_class,
    src/vilma_params.f90:26-205) as laid out by struct vh_params."""
    _fields_ = [
        ("lmax", C.c_int), ("nlat", C.c_int), ("nphi", C.c_int), ("mmax", C.c_int), ("mres", C.c_int),
        ("eps_polar", C.c_double),
        ("earth", C.c_char * 64),
        ("n_layer", C.c_int),
        ("r_earth", C.c_double), ("r_core", C.c_double),
        ("r_bot", C.c_double * VH_MAX_LAYER), ("r_top", C.c_double * VH_MAX_LAYER),
        ("rho", C.c_double * VH_MAX_LAYER), ("mu", C.c_double * VH_MAX_LAYER),
        ("eta", C.c_double * VH_MAX_LAYER),
        ("rheology", C.c_int * VH_MAX_LAYER),
        ("n_elem_layer", C.c_int * VH_MAX_LAYER),
        ("scheme", C.c_char * 8),
        ("max_couple_iter", C.c_int),
        ("earth_response", C.c_char * 8),
        ("sle_n_outer", C.c_int), ("sle_n_inner", C.c_int),
        ("sle_tol", C.c_double),
        ("sle_max_mem_iter", C.c_int), ("sle_fixed_ocean", C.c_int), ("sle_subgrid", C.c_int),
        ("dt_init", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("rtol", C.c_double), ("atol", C.c_double), ("safety", C.c_double),
        ("grow_max", C.c_double), ("shrink_min", C.c_double), ("cfl", C.c_double),
        ("rotation", C.c_int),
        ("deg1_frame", C.c_char * 8),
        ("l_visc_3d", C.c_int),
        ("visc3d_tol", C.c_double),
        ("l_toroidal", C.c_int),
        ("visc_log10_min", C.c_double), ("visc_log10_max", C.c_double),
        ("visc3d_lid_depth", C.c_double), ("visc3d_lid_log10max", C.c_double),
        ("skip_tol", C.c_double),
    ]

    def grid(self):
        """Gauss grid (nphi, nlat) the model builds from this record
        (build_sht, src/vilma_coupling.f90:644-662; NLAT_MIN, src/vilma_sht.f90:72)."""
        nlat = self.nlat if self.nlat > 0 else 2 * self.lmax + 2
        nphi = self.nphi if self.nphi > 0 else 4 * self.lmax
        return nphi, max(nlat, 32)


# selector constants (enums of include/vilma_hip.h)
F_RSL, F_ZBED, F_C, F_C0, F_SROT, F_HICE = range(6)
(S_TIME, S_BSL, S_WORST_MASS_RESID, S_N_ACCEPT, S_N_REJECT, S_N_SOLVE,
 S_N_INNER_TOT, S_N_OUTER_TOT, S_POLAR_M_RE, S_POLAR_M_IM, S_N_SLE_STEAL, S_N_FLOOR, S_N_COUPLE_TOT,
 S_N_GRAPH_LAUNCH, S_N_GRAPH_CAPTURE) = range(15)
SP_SIGMA, SP_U, SP_N, SP_TOR = range(4)
D_LMAX, D_NLAT, D_NPHI, D_NLM, D_NR, D_NE, D_NK, D_NE3D, D_NLAM, D_NEM, D_NMN = range(11)
T_SLE_BLOCKS, T_GRAPH = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def declare_api(lib, prefix):
    """Attach argtypes/restype to the model entry points of a library whose
    functions are named `<prefix>create`, `<prefix>update`, ... (the HIP
    library uses prefix "vh_")."""
    def f(name, res, *args):
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = list(args)
        return fn
    vp = C.c_void_p
    f("update", C.c_int, vp, _dp, C.c_double)
    f("get_field", C.c_int, vp, C.c_int, _dp)
    f("get_scalar", C.c_int, vp, C.c_int, _dp)
    f("get_spectrum", C.c_int, vp, C.c_int, _dp)
    f("get_dims", C.c_int, vp, C.c_int, _ip)
    f("get_memory", C.c_int, vp, _dp)
    f("set_memory", C.c_int, vp, _dp)
    f("enable_visc_3d", C.c_int, vp, _dp)
    f("enable_visc_3d_nodes", C.c_int, vp, _dp, C.c_double, C.c_double, C.c_double)
    f("resp_begin_step", C.c_int, vp)
    f("get_drift", C.c_int, vp, _dp, _dp)
    f("get_gains", C.c_int, vp, _dp, _dp)
    return lib


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)

# kernel classes of vh_profile / vh_profile_read (VH_K_* of include/vilma_hip.h)
KERNEL_CLASSES = {0: "k_drift", 1: "k_advance", 2: "k_tsh_synth", 3: "k_t3_rows", 4: "k_tsh_analysis",
                  5: "k_t3_gather", 6: "k_sle_solve", 7: "k_t3_rows_adj", 8: "k_adv_adj", 9: "sle_adjoint",
                  10: "k_drift_adj", 11: "other"}
