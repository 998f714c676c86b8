"""`&vilma` namelist reader: the Python host's vilma_par_load.

Mirrors vilma_par_load (src/vilma_params.f90:209-331): the record starts from
the built-in defaults, is overlaid with the defaults file's group (when one is
given) and then with the run file's group; `dt_init`, `dt_min`, `dt_max` are
read in years and stored in seconds (:269-277); `deg1_frame` must be "cf" or
"cm" (:297-298). Keys the viscoelastic step does not use (file names, the
VILMA-v1 block, spin-up knobs) are accepted and ignored, as a host's shared
namelist file carries them.
"""
import re

from ._abi import VhParams, VH_MAX_LAYER, SEC_PER_YEAR

_INT = {"lmax", "nlat", "nphi", "mmax", "mres", "n_layer", "max_couple_iter", "sle_n_outer",
        "sle_n_inner", "sle_max_mem_iter"}
_REAL = {"eps_polar", "r_earth", "r_core", "sle_tol", "rtol", "atol", "safety", "grow_max",
         "shrink_min", "cfl", "visc3d_tol", "visc_log10_min", "visc_log10_max", "visc3d_lid_depth",
         "visc3d_lid_log10max", "skip_tol"}
_YEARS = {"dt_init", "dt_min", "dt_max"}
_LOGICAL = {"sle_fixed_ocean", "sle_subgrid", "rotation", "l_visc_3d", "l_toroidal"}
_STRING = {"earth": 64, "scheme": 8, "earth_response": 8, "deg1_frame": 8}
_REAL_ARR = {"r_bot", "r_top", "rho", "mu", "eta"}
_INT_ARR = {"rheology", "n_elem_layer"}


class NamelistError(ValueError):
    pass


def params_defaults():
    """vilma_param_class defaults overlaid with input/vilma_defaults.nml."""
    p = VhParams()
    p.lmax, p.nlat, p.nphi, p.mmax, p.mres, p.eps_polar = 64, 0, 0, -1, 1, -1.0
    p.earth = b"M3-L70-V01"
    p.n_layer, p.r_earth, p.r_core = 5, 6371.0e3, 3480.0e3
    rb = [6301.0e3, 5951.0e3, 5701.0e3, 3480.0e3, 0.0]
    rt = [6371.0e3, 6301.0e3, 5951.0e3, 5701.0e3, 3480.0e3]
    rho = [3037.0, 3438.0, 3871.0, 4978.0, 10750.0]
    mu = [5.0605e10, 7.0363e10, 1.0549e11, 2.2834e11, 0.0]
    eta = [1.0e40, 1.0e21, 1.0e21, 2.0e21, 0.0]
    rh = [0, 1, 1, 1, 2]
    for i in range(VH_MAX_LAYER):
        p.rheology[i] = 1
    for i in range(5):
        p.r_bot[i], p.r_top[i], p.rho[i], p.mu[i], p.eta[i], p.rheology[i] = rb[i], rt[i], rho[i], mu[i], eta[i], rh[i]
    p.scheme, p.max_couple_iter, p.earth_response = b"fe", 20, b"ve"
    p.sle_n_outer, p.sle_n_inner, p.sle_tol, p.sle_max_mem_iter = 3, 20, 1.0e-7, 20
    p.sle_fixed_ocean, p.sle_subgrid = 0, 1
    p.dt_init, p.dt_min, p.dt_max = 0.0, 0.0, 1.0e30 * SEC_PER_YEAR
    p.rtol, p.atol, p.safety, p.grow_max, p.shrink_min, p.cfl = 1.0e-4, 1.0e-3, 0.9, 5.0, 0.2, 1.0
    p.rotation = 1
    p.deg1_frame = b"cm"
    p.l_visc_3d, p.visc3d_tol, p.l_toroidal = 0, 1.0e-3, 0
    p.visc_log10_min, p.visc_log10_max = 19.5, 30.0
    p.visc3d_lid_depth, p.visc3d_lid_log10max = 0.0, 22.0
    p.skip_tol = 1.0e-4
    return p


def _strip_comment(line):
    out, q = [], None
    for ch in line:
        if q:
            out.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            out.append(ch)
        elif ch == "!":
            break
        else:
            out.append(ch)
    return "".join(out)


def _group_text(text, group):
    body = "\n".join(_strip_comment(l) for l in text.splitlines())
    m = re.search(r"&" + re.escape(group) + r"\b(.*?)(?:^\s*/|&end)", body, re.S | re.I | re.M)
    return None if m is None else m.group(1)


def _split_values(s):
    vals, cur, q = [], [], None
    for ch in s:
        if q:
            cur.append(ch)
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
            cur.append(ch)
        elif ch in ",\n":
            if "".join(cur).strip():
                vals.append("".join(cur).strip())
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        vals.append("".join(cur).strip())
    out = []
    for v in vals:
        out.extend(v.split()) if not v.startswith(("'", '"')) else out.append(v)
    return out


def _real(v):
    return float(v.lower().replace("d", "e"))


def _logical(v):
    v = v.strip().lower().strip(".")
    if v in ("true", "t"):
        return 1
    if v in ("false", "f"):
        return 0
    raise NamelistError("bad logical %r" % v)


def _assignments(body):
    # key[(index)] = values, up to the next key=
    pat = re.compile(r"([A-Za-z_][A-Za-z0-9_]*)\s*(\(\s*\d+\s*\))?\s*=", re.M)
    ms = list(pat.finditer(body))
    for i, m in enumerate(ms):
        end = ms[i + 1].start() if i + 1 < len(ms) else len(body)
        idx = int(m.group(2).strip("() ")) if m.group(2) else None
        yield m.group(1).lower(), idx, body[m.end():end]


def _apply(p, body):
    for key, idx, raw in _assignments(body):
        vals = _split_values(raw)
        if not vals:
            continue
        if key in _INT:
            setattr(p, key, int(vals[0]))
        elif key in _REAL:
            setattr(p, key, _real(vals[0]))
        elif key in _YEARS:
            setattr(p, key, _real(vals[0]) * SEC_PER_YEAR)
        elif key in _LOGICAL:
            setattr(p, key, _logical(vals[0]))
        elif key in _STRING:
            s = vals[0].strip().strip("'\"").encode()
            if len(s) >= _STRING[key]:
                raise NamelistError("%s too long" % key)
            setattr(p, key, s)
        elif key in _REAL_ARR or key in _INT_ARR:
            arr = getattr(p, key)
            conv = _real if key in _REAL_ARR else (lambda v: int(v))
            start = (idx - 1) if idx else 0
            for j, v in enumerate(vals):
                if start + j >= VH_MAX_LAYER:
                    raise NamelistError("%s has more than %d entries" % (key, VH_MAX_LAYER))
                arr[start + j] = conv(v)
        # other keys (solver, file names, VILMA-v1 block, spin-up) are ignored


def vilma_par_load(filename, defaults_file=None, group="vilma", params=None):
    """Load the `&vilma` group of `filename` over the defaults (and over the
    `defaults_file` group when given) into a VhParams record."""
    p = params if params is not None else params_defaults()
    for fn in ([defaults_file] if defaults_file else []) + [filename]:
        with open(fn) as fh:
            body = _group_text(fh.read(), group)
        if body is None:
            raise NamelistError("namelist group &%s not found in %s" % (group, fn))
        _apply(p, body)
    if p.deg1_frame not in (b"cf", b"cm"):
        raise NamelistError('vilma_params: deg1_frame must be "cf" or "cm"')
    return p
