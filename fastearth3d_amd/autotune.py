"""Autotuner of the hot kernels' tile and launch shapes (build-time knobs).

    python -m fastearth3d_amd.autotune [--config 2] [--lmax 128] [--calls 8]
                                       [--write] [--report FILE]

Runs on a machine with the GPU. A coordinate search: for each knob in turn,
every candidate value (with the best values found so far for the other
knobs) is compiled into a library of its own (hipcc for gfx950, only the
translation unit the knob lives in is rebuilt) and timed in a fresh process
on the chosen workload: per-kernel HIP-event times over `calls` coupling
calls, after warm-up. A knob keeps the candidate with the smallest summed
device time of the kernels it shapes. --write rewrites csrc/vh_tuned.h with
the chosen values (rebuild afterwards); --report writes the measured table
as markdown.

The knobs (defaults in csrc/vh_tuned.h):
  VH_TSH_TB_SYNTH  batches per block of the tensor synthesis
  VH_TSH_SU        degrees per parity whose table values the synthesis loads together
  VH_DRIFT_RW      right-hand sides per block of the drift solve
  VH_TCHUNK        tensor batches per synthesis / row / analysis launch
  VH_TA_PF         k-steps of table fragments the tensor analysis keeps in flight
  VH_TA_JC         latitudes per LDS chunk of the tensor analysis
  VH_DRIFT_MINW    wavefronts per SIMD the drift's register budget must allow
  VH_DRIFT_AK      k-steps of the G_l fragment the drift's wavefront loads at a time
  VH_DRIFT_ONEPASS 1: the drift builds its forcing in one pass with LDS atomics
  VH_ROWS_RJ       latitudes per block of the grouped row pass
The sea-level solve's launch shape is a run-time knob (VH_T_SLE_BLOCKS),
tuned per ensemble by vh_ensemble_tune.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TUNED = os.path.join(HERE, "csrc", "vh_tuned.h")

# knob -> (translation unit, candidates, kernels it shapes)
KNOBS = {
    "VH_TSH_TB_SYNTH": ("vh_tensor.hip", [2, 3, 4], ["k_tsh_synth"]),
    "VH_TSH_SU": ("vh_tensor.hip", [1, 2, 4, 8], ["k_tsh_synth"]),
    "VH_DRIFT_RW": ("vh_drift.hip", [16, 32, 64], ["k_drift"]),
    "VH_TCHUNK": ("vh_api.hip", [36, 54, 72, 108], ["k_tsh_synth", "k_t3_rows", "k_tsh_analysis"]),
    "VH_TA_PF": ("vh_tensor.hip", [1, 2, 3, 4], ["k_tsh_analysis"]),
    "VH_TA_JC": ("vh_tensor.hip", [32, 48, 64, 96], ["k_tsh_analysis"]),
    "VH_DRIFT_MINW": ("vh_drift.hip", [1, 4], ["k_drift"]),
    "VH_DRIFT_AK": ("vh_drift.hip", [8, 16, 32], ["k_drift"]),
    "VH_DRIFT_ONEPASS": ("vh_drift.hip", [0, 1], ["k_drift"]),
    "VH_ROWS_RJ": ("vh_tensor.hip", [2, 4, 8], ["k_t3_rows"]),
}


def current_values():
    vals = {}
    for line in open(TUNED):
        m = re.match(r"#define (VH_\w+) (\d+)", line)
        if m:
            vals[m.group(1)] = int(m.group(2))
    return vals


def build_variant(vals, unit, workdir):
    """A library with every knob set by -D; objects of the other units are
    the in-tree ones (python -m fastearth3d_amd.build first)."""
    from . import build as B
    flags = ["-D%s=%d" % kv for kv in sorted(vals.items())]
    objs = []
    for s in B.SOURCES:
        if s == unit:
            o = os.path.join(workdir, s.replace(".hip", ".o"))
            cmd = [B.HIPCC] + B.FLAGS + flags + ["-c", os.path.join(B.CSRC, s), "-o", o]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode:
                raise RuntimeError("hipcc failed for %s %s:\n%s" % (s, flags, r.stderr))
        else:
            o = os.path.join(B.OBJ, s.replace(".hip", ".o"))
            if not os.path.exists(o):
                raise RuntimeError("%s is not built (python -m fastearth3d_amd.build)" % o)
        objs.append(o)
    lib = os.path.join(workdir, "libvilma_hip.so")
    r = subprocess.run([B.HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib] + objs,
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stderr)
    return lib


def time_library(lib, config, lmax, calls):
    """Per-kernel average device time (us) and wall time per call (ms) of
    the workload with library `lib`, in a child process."""
    r = subprocess.run([sys.executable, "-m", "fastearth3d_amd.autotune", "--time", lib, "--config", str(config),
                        "--lmax", str(lmax), "--calls", str(calls)], capture_output=True, text=True, cwd=ROOT,
                       timeout=600)
    if r.returncode:
        raise RuntimeError("timing run failed:\n" + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def _time_child(lib, config, lmax, calls):
    import ctypes as C
    import time

    import numpy as np

    from . import configs
    from . import lib as L_
    from ._abi import KERNEL_CLASSES
    L_.LIB = lib
    L = L_.load()
    c = configs.config(config, lmax=lmax)
    se = L_.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        se.enable_visc_3d(c["pert"])
    first = c["steps"] // 2   # mid-history ice
    for k in range(2):
        se.update(c["ice"](first + k), c["dt_yr"])
    ice = [c["ice"](first + 2 + k) for k in range(calls)]
    L.vh_profile(se._h, 1)
    t0 = time.perf_counter()
    for h in ice:
        se.update(h, c["dt_yr"])
    wall = (time.perf_counter() - t0) / calls
    L.vh_profile(se._h, 0)
    n = len(KERNEL_CLASSES)
    ms, cnt = (C.c_double * n)(), (C.c_double * n)()
    L.vh_profile_read(se._h, ms, cnt)
    out = {"ms_per_call": 1e3 * wall, "kernel_us_per_call": {}}
    for k, name in KERNEL_CLASSES.items():
        if cnt[k] > 0:
            out["kernel_us_per_call"][name] = 1e3 * ms[k] / calls
    se.finalize()
    print(json.dumps(out))


def write_tuned(vals):
    src = open(TUNED).read()
    for k, v in vals.items():
        src = re.sub(r"(#define %s )(\d+)" % k, r"\g<1>%d" % v, src)
    open(TUNED, "w").write(src)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--lmax", type=int, default=128)
    ap.add_argument("--calls", type=int, default=8)
    ap.add_argument("--write", action="store_true")
    ap.add_argument("--report")
    ap.add_argument("--knobs", help="comma-separated subset of the knobs to search (default: all)")
    ap.add_argument("--time", help=argparse.SUPPRESS)
    a = ap.parse_args()
    if a.time:
        _time_child(a.time, a.config, a.lmax, a.calls)
        return
    best = current_values()
    rows = []
    work = tempfile.mkdtemp(prefix="vh_autotune_")
    try:
        only = set(a.knobs.split(",")) if a.knobs else None
        for knob, (unit, cands, kernels) in KNOBS.items():
            if only is not None and knob not in only:
                continue
            scores = {}
            for v in cands:
                vals = dict(best, **{knob: v})
                d = os.path.join(work, "%s_%d" % (knob, v))
                os.makedirs(d, exist_ok=True)
                try:
                    t = time_library(build_variant(vals, unit, d), a.config, a.lmax, a.calls)
                except RuntimeError as e:   # e.g. a tile that does not fit: not a candidate
                    print("%s = %d: %s" % (knob, v, str(e).splitlines()[0]), flush=True)
                    continue
                us = sum(t["kernel_us_per_call"].get(k, 0.0) for k in kernels)
                scores[v] = us
                rows.append((knob, v, us, t["ms_per_call"]))
                print("%s = %d: %.1f us/call in %s, %.3f ms/call" % (knob, v, us, "+".join(kernels),
                                                                       t["ms_per_call"]), flush=True)
                shutil.rmtree(d, ignore_errors=True)
            if scores:
                best[knob] = min(scores, key=scores.get)
            print("chosen %s = %d" % (knob, best[knob]), flush=True)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    if a.write:
        write_tuned(best)
    if a.report:
        with open(a.report, "w") as f:
            f.write("# Autotuner: tile and launch shapes\n\n")
            f.write("Command: `python -m fastearth3d_amd.autotune --config %d --lmax %d --calls %d`"
                    " (device time per coupling call of the kernels each knob shapes, HIP events; wall time"
                    " per call)\n\n" % (a.config, a.lmax, a.calls))
            f.write("| knob | value | kernel us / call | ms / call |\n|---|---:|---:|---:|\n")
            for knob, v, us, ms in rows:
                f.write("| %s | %d%s | %.1f | %.3f |\n" % (knob, v, " (chosen)" if best[knob] == v else "", us, ms))
    print(json.dumps(best))


if __name__ == "__main__":
    main()
