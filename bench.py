#!/usr/bin/env python3
"""Throughput of the MI355X viscoelastic time step: simulated model-kyr per
wall-second of solid_earth_update on one GPU.

A step is one coupling call (solid_earth_update over dt_yr) of the workload:
by default the lmax 128 lateral-viscosity configuration (64 radial elements,
192 x 384 Gauss grid, sea level with migrating coastlines, rotation, dt 50 yr)
the metric is quoted on. The ice history is uploaded to HBM before the timed
region and handed over as device pointers (vh_update_device). W untimed warm-up
calls, then K timed calls bracketed by device synchronisations. The JSON line
adds the roofline of the dominant kernel (algorithmic bytes per launch over its
HIP-event-timed average duration) and the oracle's CPU throughput on a bounded
sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
NAMES = {0: "k_drift", 1: "k_advance", 2: "k_tsh_synth", 3: "k_t3_rows", 4: "k_tsh_analysis", 5: "k_t3_gather",
         6: "k_sle_solve"}


def alg_bytes(dims, cls, chunk_batches=None, sle_iters=None):
    """Algorithmic HBM bytes of one launch of kernel class cls.

    drift / advance: the Maxwell memory of the Maxwell elements (24 doubles
    per element and slot), the nodal drift at the memory nodes (4 doubles
    per node and slot), the surface drift (6 doubles per slot), the per-slot
    norm and, for the drift, the per-degree restricted solution operators
    ((2 nmn + 3) x 2 nmn doubles per degree, read once); the advance also
    reads the 3-D increments (12 complex per 3-D element and slot).
    Tensor kernels (per launch over `chunk_batches` element-shape batches):
    synthesis reads 4 complex coefficients per (l, m) and writes 6 complex
    dyad row spectra per (m, latitude), plus the P and dP tables once;
    the row pass reads and writes those spectra and reads one viscosity-rate
    row per latitude; the analysis is the synthesis reversed; the gather
    reads the 3-D elements' memory and writes 12 complex per element and slot.
    The fused sea-level loop (per launch over `sle_iters` iterations on
    average) reads and writes the grid fields of each iteration once.
    """
    nk, nem, nmn, nlm, L = dims["nk"], dims["nem"], dims["nmn"], dims["nlm"], dims["lmax"]
    ne3d, nlat, nphi = dims.get("ne3d", 0), dims["nlat"], dims["nphi"]
    tau = 24 * nem * nk * 8
    tables = 2 * nlm * ((nlat + 1) // 2) * 8
    spec = 6 * (L + 1) * nlat * 16
    nb = chunk_batches
    if cls == 0:   # drift: read tau + norm + operators, write nodal + surface drift
        return tau + 8 * nk + L * (2 * nmn + 3) * 2 * nmn * 8 + 4 * nmn * nk * 8 + 6 * nk * 8
    if cls == 1:   # advance: read + write tau, read nodal drift, sigma and 3-D increments, write norm
        return 2 * tau + 4 * nmn * nk * 8 + 16 * nlm + 8 * nk + 12 * ne3d * nk * 16
    if cls == 2 and nb:
        return nb * (4 * nlm * 16 + spec) + tables
    if cls == 3 and nb:
        return nb * 2 * spec + (nb // 3 + 1) * nlat * nphi * 8
    if cls == 4 and nb:
        return nb * (spec + 4 * nlm * 16) + tables
    if cls == 5:
        return ne3d * (24 * nk * 8 + 12 * nlm * 16) + 4 * nmn * nk * 8 + 16 * nlm
    if cls == 6 and sle_iters is not None:
        # sea-level loop: per iteration 8 grid fields and 3 spectra (load, C,
        # rsl, ice increment, subgrid term, Sraw, s_rot, the updated rsl;
        # sigma, u, N), plus the loop's start (mask and load transforms)
        per_it = 8 * nlat * nphi * 8 + 3 * nlm * 16
        return sle_iters * per_it + 2 * nlat * nphi * 8
    return None


def cpu_baseline(cfg, lmax, calls):
    """The oracle (OpenMP C restatement of the same step) timed on the host
    cores over `calls` coupling calls of the same workload, from the same
    initial state (model set-up and the lateral-viscosity enabling are not
    timed, as on the GPU side)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fastearth3d_amd import configs
    c = configs.config(cfg, lmax=lmax)
    m = O.OracleModel(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        m.enable_visc_3d(c["pert"])
    t0 = time.perf_counter()
    for k in range(calls):
        m.update(c["ice"](k), c["dt_yr"])
    el = time.perf_counter() - t0
    nsub = m.scalar(3)
    m.finalize()
    return calls * c["dt_yr"] / 1000.0 / el, el, nsub


def ensemble_bench(a, L):
    """Configuration 4: an uneven ensemble (lmax 64/96/128/192, 32-64
    elements, 16 of 64 members 3-D, sea level + rotation, dt 100 yr). Each
    member is its own model, stepped one after another on one host thread
    (the fused sea-level kernel of a member takes the whole device), every
    member's ice history resident in HBM. value = timed member calls per
    second / 200 (completed 200-call members per second)."""
    import torch
    from fastearth3d_amd import configs, lib
    n = a.members
    members = []
    for i in range(n):
        c = configs.ensemble_member(i, n)
        se = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
        if c["pert"] is not None:
            se.enable_visc_3d(c["pert"])
        ice = torch.from_numpy(np.stack([c["ice"](k) for k in range(a.warmup + a.steps)])).to("cuda:0")
        members.append((se, ice, c))
    torch.cuda.synchronize()

    def run(i, k0, k1):
        se, ice, c = members[i]
        for k in range(k0, k1):
            if L.vh_update_device(se._h, C.c_void_p(ice[k].data_ptr()), c["dt_yr"]):
                raise RuntimeError(se._err())
        return se.scalar(3)

    for i in range(n):
        run(i, 0, a.warmup)
    torch.cuda.synchronize()
    sub0 = sum(m[0].scalar(3) for m in members)
    t0 = time.perf_counter()
    for i in range(n):
        run(i, a.warmup, a.warmup + a.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nsub = sum(m[0].scalar(3) for m in members) - sub0
    calls = n * a.steps
    kyr = sum(m[2]["dt_yr"] for m in members) * a.steps / 1000.0
    mem = configs.ensemble_members(n)
    out = {
        "metric": "ens/s (64-member uneven ensemble, 1 GPU)", "value": round(calls / 200.0 / el, 4),
        "unit": "ens/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(el / a.steps * 1000.0, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c4 ensemble: %d members, lmax 64/96/128/192, 32-64 elements, %d 3-D, "
                               "SLE + rotation, dt 100 yr" % (n, sum(x[2] for x in mem)),
                   "members": n, "calls_per_member_timed": a.steps,
                   "sub_steps_timed": int(nsub)},
        "member_kyr_per_s": round(kyr / el, 4), "member_calls_per_s": round(calls / el, 2),
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--lmax", type=int, default=128)
    ap.add_argument("--cpu-calls", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--members", type=int, default=64, help="ensemble size (--config 4)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 times): the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    if a.copy_calib_mib > 0:
        n = a.copy_calib_mib * (1 << 20) // 8
        x = torch.arange(n, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_read": n * 8, "copy_calib_bytes_written": n * 8, "launches": 5}))
        return
    from fastearth3d_amd import configs, lib
    L = lib.load()
    L.vh_update_device.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
    L.vh_profile.argtypes = [C.c_void_p, C.c_int]
    L.vh_profile_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    if a.config == 4:
        ensemble_bench(a, L)
        return

    c = configs.config(a.config, lmax=a.lmax)
    p = c["params"]
    se = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        se.enable_visc_3d(c["pert"])
    nsteps = a.warmup + a.steps
    ice = torch.from_numpy(np.stack([c["ice"](k) for k in range(nsteps)])).to("cuda:0")
    torch.cuda.synchronize()

    def step(k):
        rc = L.vh_update_device(se._h, C.c_void_p(ice[k].data_ptr()), c["dt_yr"])
        if rc:
            raise RuntimeError(se._err())

    for k in range(a.warmup):
        step(k)
    L.vh_profile(se._h, 1)
    n_acc0 = se.scalar(3)
    n_in0 = se.scalar(6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.warmup, nsteps):
        step(k)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = (C.c_double * 8)()
    cnt = (C.c_double * 8)()
    L.vh_profile_read(se._h, ms, cnt)
    nsub = se.scalar(3) - n_acc0
    n_inner = se.scalar(6) - n_in0
    kyr = a.steps * c["dt_yr"] / 1000.0
    value = kyr / el

    d = dict(se.dims)
    d["nlat"], d["nphi"] = se.nlat, se.nphi
    nbat = 3 * d.get("ne3d", 0)
    # the tensor kernels run in chunks of up to 48 batches; the average launch
    # covers nbat / (number of chunks) batches
    nch = (nbat + 47) // 48 if nbat else 1
    tf = os.path.join(ROOT, "profiles", "c%d_lmax%d_traffic.json" % (a.config, a.lmax))
    traffic = json.load(open(tf))["kernels"] if os.path.exists(tf) else {}
    per_kernel = {}
    for k in NAMES:
        if cnt[k] <= 0:
            continue
        ab = alg_bytes(d, k, nbat / nch if nbat else None, n_inner / cnt[k] if k == 6 else None)
        if not ab:
            continue
        avg_s = ms[k] / cnt[k] / 1000.0
        ach = ab / avg_s / 1e9
        t = traffic.get(NAMES[k])
        per_kernel[NAMES[k]] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(ach / HBM_PEAK_GBS, 4),
                                "traffic": t["hbm_bytes_per_launch"] if t else None,
                                "avg_launch_us": round(avg_s * 1e6, 2), "launches": int(cnt[k]),
                                "alg_bytes_per_launch": int(ab), "total_ms": round(ms[k], 3)}
    # the dominant kernel by summed device time
    best = max(per_kernel, key=lambda n: per_kernel[n]["total_ms"], default=None)
    roof = dict(per_kernel[best], kernel=best) if best else None
    kernel_ms = {NAMES[k]: round(ms[k], 3) for k in NAMES if cnt[k] > 0}

    cpu = None
    if not a.no_cpu:
        v, t, ns = cpu_baseline(a.config, a.lmax, a.cpu_calls)
        cpu = {"value": round(v, 6), "unit": "model-kyr/s",
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())), "kind": "port",
               "sample": "first %d coupling call(s) of %g yr (%d FE sub-steps) of the same workload, %.1f s" % (
                   a.cpu_calls, c["dt_yr"], int(ns), t)}

    names = {0: "c0 1-D load-only lmax %d" % a.lmax, 1: "c1 1-D + SLE + rotation lmax %d" % a.lmax,
             2: "c2 3-D viscosity + SLE + rotation lmax %d" % a.lmax}
    out = {
        "metric": "model-kyr/s at lmax=128 3D visc (1 GPU)",
        "value": round(value, 4), "unit": "model-kyr/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(el / a.steps * 1000.0, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": names[a.config], "lmax": a.lmax, "nlat": se.nlat, "nphi": se.nphi,
                   "radial_elements": d["ne"], "maxwell_elements": d["nem"], "dt_yr": c["dt_yr"],
                   "sub_steps_timed": int(nsub), "ne3d": int(se.dims.get("ne3d", 0))},
        "roofline": roof, "kernel_roofline": per_kernel, "kernel_ms": kernel_ms, "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out))
    se.finalize()


if __name__ == "__main__":
    main()
