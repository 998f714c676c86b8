"""Whole-run trajectories of the benchmark configurations 0, 1 and 2, written
by the oracle (the restated reference step) for the GPU full-run parity tests.

The run protocol (`run_record`) is shared by the generator below and the
`gpu` tests, so both sides drive a model identically: `solid_earth_init` from
the configuration's reference state, `enable_visc_3d` for configuration 2,
then every coupling call of the configuration (200 / 350 / 400 calls of
`solid_earth_update`, src/vilma_coupling.f90:336-437) with the seeded ice
history. Recorded per call: model time, bsl, polar motion, accepted and
rejected FE sub-steps, sea-level solves, inner iterations and coastline
passes (cumulative counters), max |rsl| and a CRC-32 of the ocean mask;
at four checkpoints (quarter, half, three quarters, end of the run): the
converged load spectrum sigma, the bit-packed mask and max |tau| of each of
the six memory arrays; at half and at the end: the rsl grid.

TEST INFRASTRUCTURE: the generator imports the oracle. Run once in the
build container (about 40 minutes, dominated by configuration 2):
    python tests/golden/fullrun.py [0 1 2]
Output: tests/golden/fullrun_c{0,1,2}.npz.
"""
import os
import sys
import time
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

# per-call scalar columns of the record
COLS = ("time", "bsl", "m_re", "m_im", "n_accept", "n_reject", "n_solve", "n_inner", "n_outer", "rsl_max", "mask_crc")
SCALARS = dict(time=0, bsl=1, n_accept=3, n_reject=4, n_solve=5, n_inner=6, n_outer=7, m_re=8, m_im=9)


def fixture_path(idx):
    return os.path.join(HERE, "fullrun_c%d.npz" % idx)


def checkpoints(steps):
    return [steps // 4, steps // 2, (3 * steps) // 4, steps]


def mask_bits(C):
    return np.packbits(np.asarray(C).ravel() != 0.0)


def run_record(model, cfg, calls=None, on_call=None):
    """Drive `model` (an oracle or HIP model handle, fresh from
    solid_earth_init) through the configuration's coupling calls and return
    the record described in the module docstring. `on_call(k, model)` is
    called after each call (the GPU tests compare as they go)."""
    steps = cfg["steps"] if calls is None else calls
    if cfg["pert"] is not None:
        model.enable_visc_3d(cfg["pert"])
    ck = set(checkpoints(cfg["steps"]))
    rows, sig, masks, taumax, rsl = [], {}, {}, {}, {}
    for k in range(steps):
        model.update(cfg["ice"](k), cfg["dt_yr"])
        r = model.rsl
        C = model.C
        row = [model.scalar(SCALARS[c]) if c in SCALARS else 0.0 for c in COLS]
        row[COLS.index("rsl_max")] = float(np.abs(r).max())
        row[COLS.index("mask_crc")] = float(zlib.crc32(mask_bits(C).tobytes()))
        rows.append(row)
        n = k + 1
        if n in ck:
            sig[n] = model.sigma_lm
            masks[n] = mask_bits(C)
            tau = model.memory()
            taumax[n] = np.abs(tau).reshape(6, -1).max(axis=1)
            if n in (cfg["steps"] // 2, cfg["steps"]):
                rsl[n] = r
        if on_call is not None:
            on_call(k, model)
    return dict(rows=np.array(rows), sigma=sig, mask=masks, taumax=taumax, rsl=rsl)


def save(idx, rec, seconds):
    out = {"rows": rec["rows"], "cols": np.array(COLS), "seconds": np.array(seconds)}
    for n, v in rec["sigma"].items():
        out["sigma_%d" % n] = v
    for n, v in rec["mask"].items():
        out["mask_%d" % n] = v
    for n, v in rec["taumax"].items():
        out["taumax_%d" % n] = v
    for n, v in rec["rsl"].items():
        out["rsl_%d" % n] = v
    np.savez_compressed(fixture_path(idx), **out)


def load(idx):
    f = np.load(fixture_path(idx), allow_pickle=False)
    rec = dict(rows=f["rows"], sigma={}, mask={}, taumax={}, rsl={})
    for key in f.files:
        for kind in ("sigma", "mask", "taumax", "rsl"):
            if key.startswith(kind + "_"):
                rec[kind][int(key.split("_")[1])] = f[key]
    return rec


def main(argv):
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fastearth3d_amd import configs
    which = [int(a) for a in argv] or [0, 1, 2]
    for idx in which:
        cfg = configs.config(idx)
        p = cfg["params"]
        m = O.OracleModel(p, cfg["z_bed_eq"], cfg["h_ice_eq"], cfg["h_ice_init"])
        t0 = time.time()

        def progress(k, model, t0=t0, idx=idx):
            if (k + 1) % 25 == 0:
                print("c%d call %d/%d  %.0f s  accepted %d inner %d" % (
                    idx, k + 1, cfg["steps"], time.time() - t0, model.scalar(3), model.scalar(6)), flush=True)

        rec = run_record(m, cfg, on_call=progress)
        save(idx, rec, time.time() - t0)
        print("wrote", fixture_path(idx), "%.0f s" % (time.time() - t0), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
