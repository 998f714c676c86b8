#!/usr/bin/env python3
"""Throughput of the MI355X viscoelastic time step: simulated model-kyr per
wall-second of solid_earth_update on one GPU.

A step is one coupling call (solid_earth_update over dt_yr) of the workload:
by default the lmax 128 lateral-viscosity configuration (64 radial elements,
192 x 384 Gauss grid, sea level with migrating coastlines, rotation, dt 50 yr)
the metric is quoted on. The ice history is uploaded to HBM before the timed
region and handed over as device pointers (vh_update_device). W untimed warm-up
calls, then K timed calls bracketed by device synchronisations. The JSON line
adds the roofline of the dominant kernel (algorithmic bytes — or, for the
tensor Legendre kernels, FP64 flops — per launch over its HIP-event-timed
average duration), the oracle's CPU throughput on a bounded sample of the
same workload, and the metric's two other parts: ens/s of the 64-member
uneven ensemble (with its ratio to stepping the members one after another)
and the adjoint / forward time ratio (--no-extras leaves them out;
--config 4 and --config 3 measure them alone, in full).
"""
import argparse
import copy
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
FP64_PEAK_TFS = 78.6
NAMES = {0: "k_drift", 1: "k_advance", 2: "k_tsh_synth", 3: "k_t3_rows", 4: "k_tsh_analysis", 5: "k_t3_gather",
         6: "k_sle_solve"}
ADJ_NAMES = {2: "k_tsh_synth (forward + adjoint rows)", 4: "k_tsh_analysis", 5: "k_t3_gather + k_t3_pack_adj",
             7: "k_t3_rows_adj", 8: "k_adv_adj + k_adv_nodes", 9: "sea-level transpose (scalar SHT kernels)",
             10: "k_drift_adj"}
NCLASS = 12
PROF_EVERY = 10   # calls per event-timed call in the forward bench


def _tuned(name):
    """A build-time knob of the library (fastearth3d_amd/csrc/vh_tuned.h)."""
    import re
    for line in open(os.path.join(ROOT, "fastearth3d_amd", "csrc", "vh_tuned.h")):
        m = re.match(r"#define %s (\d+)" % name, line)
        if m:
            return int(m.group(1))
    raise KeyError(name)


TCHUNK = _tuned("VH_TCHUNK")   # tensor batches per launch


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


def alg_flops(dims, cls, chunk_batches=None):
    """FP64 flops of one launch of a tensor Legendre kernel (synthesis, class 2;
    analysis, class 4): per element-shape batch 12 nlm nlat — six table
    products over the parity-folded latitudes, each a complex x real
    multiply-add (4 flops) per (l, m) and latitude pair. None for the other
    classes, which are priced by their HBM bytes."""
    if cls in (2, 4) and chunk_batches:
        return chunk_batches * 12.0 * dims["nlm"] * dims["nlat"]
    return None


def host_cpu():
    """The host's CPU model, NUMA nodes and the cores this process may use."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        numa = len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and d[4:].isdigit()])
    except OSError:
        numa = None
    return {"cpu_model": model, "numa_nodes": numa, "cores_available": len(os.sched_getaffinity(0))}


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


def ensemble_models(n, calls, mem=None):
    """The uneven ensemble's members (configs.ensemble_members): per member
    its workload and its ice history for `calls` coupling calls, resident in
    HBM ([calls][nlat][nphi], the domes times the history factor of each
    call, formed on the device: the same products as the host's)."""
    import torch
    from fastearth3d_amd import configs
    mem = mem or configs.ensemble_members(n)
    out = []
    for i in range(n):
        c = configs.ensemble_member(i, n, members=mem)
        fac = [float(configs.ice_history_factor((k + 1) * c["dt_yr"] / 1000.0)) for k in range(calls)]
        d = torch.from_numpy(c["domes"]).to("cuda:0")
        hist = d.unsqueeze(0) * torch.tensor(fac, dtype=torch.float64, device="cuda:0").view(-1, 1, 1)
        out.append((c, hist.contiguous()))
    torch.cuda.synchronize()
    return mem, out


def _ens_models(lib, ws, sle_blocks=0, graph=0):
    from fastearth3d_amd import _abi
    ms = []
    for c, _ in ws:
        m = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
        if c["pert"] is not None:
            m.enable_visc_3d(c["pert"])
        m.set_tuning(_abi.T_SLE_BLOCKS, sle_blocks)
        m.set_tuning(_abi.T_GRAPH, graph)
        ms.append(m)
    return ms


def ensemble_measure(L, ws, warmup, calls, workers, tune_calls=2, graph=0):
    """The same calls from the same state two ways: every member stepped
    alone, one after another (vh_update_device, one block per work item in
    the sea-level solve), and all members through vh_ensemble_run_device
    after the ensemble's autotuner (vh_ensemble_tune, 3 x tune_calls calls)
    has picked the solve's launch shape. Warm-up and tuning calls are run
    untimed by both. Returns (sequential s, ensemble s, sub-steps timed,
    per-member timed seconds of the sequential pass, same results, chosen
    VH_T_SLE_BLOCKS). graph: both ways with the sub-steps replayed as hipGraphs."""
    import torch
    from fastearth3d_amd import lib
    pre = warmup + 3 * tune_calls
    seq = _ens_models(lib, ws, graph=graph)
    for m, (c, h) in zip(seq, ws):
        for k in range(pre):
            m.update_device(h[k].data_ptr(), c["dt_yr"])
    torch.cuda.synchronize()
    n0 = sum(m.scalar(3) for m in seq)
    per = []
    t0 = time.perf_counter()
    for m, (c, h) in zip(seq, ws):
        t1 = time.perf_counter()
        for k in range(pre, pre + calls):
            m.update_device(h[k].data_ptr(), c["dt_yr"])
        per.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    nsub = sum(m.scalar(3) for m in seq) - n0
    ref = [(m.rsl, m.scalar(3)) for m in seq]
    for m in seq:
        m.finalize()
    par = _ens_models(lib, ws, graph=graph)
    ens = lib.Ensemble(par, n_workers=workers)
    dts = [c["dt_yr"] for c, _ in ws]
    if warmup:
        ens.run_device(warmup, [h.data_ptr() for _, h in ws], dts)
    chosen = ens.tune(tune_calls, [h[warmup].data_ptr() for _, h in ws], dts) if tune_calls > 0 else 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run_device(calls, [h[pre].data_ptr() for _, h in ws], dts)
    torch.cuda.synchronize()
    t_ens = time.perf_counter() - t0
    same = all(np.array_equal(m.rsl, r) and m.scalar(3) == s for m, (r, s) in zip(par, ref))
    ens.close()
    for m in par:
        m.finalize()
    return t_seq, t_ens, int(nsub), per, same, chosen


def ensemble_bench(a, L, emit=True):
    """The uneven ensemble: 64 members of their own earths (lithosphere
    60-120 km, upper / lower mantle 1e20-1e22 Pa s, 32-64 elements, lmax
    64/96/128/192, 16 with lateral viscosity), sea level + rotation, dt
    100 yr, each with its own topography and ice history (resident in HBM).
    value = completed 200-call members per second through
    vh_ensemble_run_device over --steps calls of every member (200: the whole
    workload); the same calls of the members stepped one after another give
    the sequential rate beside it."""
    n = a.members
    mem, ws = ensemble_models(n, a.warmup + 3 * a.tune_calls + a.steps)
    t_seq, t_ens, nsub, per, same, chosen = ensemble_measure(L, ws, a.warmup, a.steps, a.ens_workers, a.tune_calls,
                                                             a.graph)
    calls = n * a.steps
    kyr = sum(c["dt_yr"] for c, _ in ws) * a.steps / 1000.0
    crit = max(per)
    out = {
        "metric": "ens/s (64-member uneven ensemble, 1 GPU)", "value": round(calls / 200.0 / t_ens, 4),
        "unit": "ens/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t_ens / a.steps * 1000.0, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c4 uneven ensemble: %d members, lmax 64/96/128/192, 32-64 elements, %d 3-D, "
                               "own earths, topography and ice, SLE + rotation, dt 100 yr" % (n, sum(x[2] for x in mem)),
                   "members": n, "calls_per_member_timed": a.steps, "sub_steps_timed": nsub,
                   "workers": a.ens_workers, "tuned_sle_blocks": chosen,
                   "tuning_calls_untimed": 3 * a.tune_calls, "graph_replay": bool(a.graph)},
        "sequential_ens_per_s": round(calls / 200.0 / t_seq, 4),
        "speedup_vs_sequential": round(t_seq / t_ens, 3),
        # no schedule of independent members can beat the slowest member run alone
        "bound_by_slowest_member": round(t_seq / crit, 3),
        # share of the sequential time spent in the laterally heterogeneous
        # members, whose tensor kernels fill the device on their own
        "sequential_share_3d": round(sum(t for t, (c, _) in zip(per, ws) if c["pert"] is not None) / sum(per), 4),
        "ensemble_equals_sequential": same,
        "member_kyr_per_s": round(kyr / t_ens, 4), "member_calls_per_s": round(calls / t_ens, 2),
    }
    if emit:
        print(json.dumps(out))
    return out


def site_matrix(L, colat, lon):
    """E[s, i] (i in SHTns LM order) with f(site s) = Re(E @ a) for a real
    field of spectrum a: c_m P_lm(cos colat_s) e^{i m lon_s}, orthonormal
    P_lm without the Condon-Shortley phase, c_0 = 1, c_m = 2
    (the spectral point evaluation of sht_grid_eval_point, src/vilma_sht.f90:310-325)."""
    x, sn = np.cos(colat), np.sin(colat)
    cols = []
    pmm = np.full_like(x, 1.0 / np.sqrt(4.0 * np.pi))
    for m in range(L + 1):
        if m > 0:
            pmm = pmm * np.sqrt((2.0 * m + 1.0) / (2.0 * m)) * sn
        cm = 1.0 if m == 0 else 2.0
        e = cm * np.exp(1j * m * lon)
        pm2, pm1 = None, None
        for l in range(m, L + 1):
            if l == m:
                p = pmm
            elif l == m + 1:
                p = np.sqrt(2.0 * m + 3.0) * x * pmm
            else:
                a_ = np.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
                b_ = np.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
                p = a_ * (x * pm1 - b_ * pm2)
            cols.append(p * e)
            pm2, pm1 = pm1, p
    return np.stack(cols, axis=1)


def adjoint_bench(a, L, emit=True):
    """Configuration 3: reverse mode over 100 coupling calls of the lmax 128
    3-D workload (sea level with migrating coastlines and rotation, as the
    forward benchmark).
    Misfit J = sum (rsl - target)^2 at 500 uniform sites x 10 evenly spaced
    output calls, target from a forward run with another seeded delta field;
    gradients w.r.t. the lateral log-viscosity field (and delta's SH
    coefficients l <= 32 per element, by the transposed synthesis) and the
    ice history. value = wall(forward_record + misfit gradient + backward)
    / wall(plain forward) over the same calls from the same state
    (vh_forward_rewind); the misfit and dJ/drsl are two FP64 GEMMs on the
    device against the (500 x grid) site-evaluation matrix;
    memory checkpointed every 10 calls. Then central differences (h = 1e-4)
    along --fd-dirs seeded delta directions (l <= 32, elements away from the
    3-D threshold) and one ice direction, on fresh models."""
    import torch
    from fastearth3d_amd import configs, lib
    L.vh_forward_record_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_void_p]
    L.vh_backward_device.argtypes = [C.c_void_p] + [C.c_void_p] * 3
    L.vh_forward_rewind.argtypes = [C.c_void_p]
    ncalls, every, nsite, nout = a.adj_calls, 10, 500, 10
    c = configs.config(2, lmax=a.lmax)
    p = c["params"]
    nphi, nlat = p.grid()
    ng = nphi * nlat
    colat, lon = configs.gauss_grid(nlat, nphi)
    ice_np = np.stack([c["ice"](k) for k in range(ncalls)])
    ice = torch.from_numpy(ice_np).to("cuda:0")
    outs = [ncalls * (i + 1) // nout - 1 for i in range(nout)]
    rng = np.random.default_rng(71)
    zs = rng.uniform(-1.0, 1.0, nsite)
    site_colat, site_lon = np.arccos(zs), rng.uniform(0.0, 2.0 * np.pi, nsite)
    E = site_matrix(a.lmax, site_colat, site_lon)
    x_, w_ = np.polynomial.legendre.leggauss(nlat)
    W = np.repeat((w_[::-1] * 2.0 * np.pi / nphi)[:, None], nphi, axis=1)
    cm = np.array([1.0 if m == 0 else 2.0 for m in range(a.lmax + 1) for _ in range(m, a.lmax + 1)])

    def check(se, rc):
        if rc:
            raise RuntimeError(se._err())

    def site_values(se, rsl_grid):
        return (E @ se.sht_analysis(rsl_grid)).real

    # the site evaluation as one real (nsite, ng) matrix S on the device:
    # v = Re(E analysis(rsl)) = S rsl, row s = the transpose of the evaluation
    # at site s applied to e_s, W synthesis(conj(E_s) / c_m)
    ev = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    S_np = np.stack([W * ev.sht_synthesis(np.conj(E[k]) / cm) for k in range(nsite)]).reshape(nsite, ng)
    probe = configs.random_sh_field(colat, lon, 12, -2.0, 5)
    assert np.abs(S_np @ probe.ravel() - site_values(ev, probe)).max() <= 1e-10 * np.abs(probe).max()
    ev.finalize()
    S_d = torch.from_numpy(S_np).to("cuda:0")
    del S_np

    # target: another seeded delta field, same ice
    tgt_pert = configs.delta_visc(p, colat, lon, 1037)
    st = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    st.enable_visc_3d(tgt_pert)
    target = {}
    for k in range(ncalls):
        check(st, L.vh_update_device(st._h, C.c_void_p(ice[k].data_ptr()), c["dt_yr"]))
        if k in outs:
            target[k] = site_values(st, st.rsl)
    st.finalize()
    T_d = torch.from_numpy(np.stack([target[k] for k in outs])).to("cuda:0")

    se = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    se.enable_visc_3d(c["pert"])
    rsl_d = torch.empty((ncalls, nlat, nphi), dtype=torch.float64, device="cuda:0")
    dJ_d = torch.zeros_like(rsl_d)
    gp_d = torch.empty((se.dims["ne"], nlat, nphi), dtype=torch.float64, device="cuda:0")
    gi_d = torch.empty((ncalls + 1, nlat, nphi), dtype=torch.float64, device="cuda:0")

    def misfit_grad(rsl_out):
        """J and dJ/drsl (into dJ_d) on the device: two FP64 GEMMs with S."""
        dJ_d.zero_()
        R = rsl_out[outs].reshape(nout, ng) @ S_d.T - T_d   # site residuals [nout, nsite]
        dJ_d[outs] = ((2.0 * R) @ S_d).reshape(nout, nlat, nphi)
        return float((R * R).sum())

    def record():
        check(se, L.vh_forward_record_device(se._h, ncalls, C.c_void_p(ice.data_ptr()), c["dt_yr"], every,
                                             C.c_void_p(rsl_d.data_ptr())))

    def backward():
        check(se, L.vh_backward_device(se._h, C.c_void_p(dJ_d.data_ptr()), C.c_void_p(gp_d.data_ptr()),
                                       C.c_void_p(gi_d.data_ptr())))

    # warm-up cycle (allocates the checkpoints and the sub-step store), then rewind
    record()
    misfit_grad(rsl_d)
    backward()
    check(se, L.vh_forward_rewind(se._h))
    # plain forward over the window
    torch.cuda.synchronize()
    n0 = se.scalar(3)
    t0 = time.perf_counter()
    for k in range(ncalls):
        check(se, L.vh_update_device(se._h, C.c_void_p(ice[k].data_ptr()), c["dt_yr"]))
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    nsub = se.scalar(3) - n0
    check(se, L.vh_forward_rewind(se._h))
    # forward with checkpoints + backward with recompute (misfit gradient timed apart)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    record()
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    J = misfit_grad(rsl_d)
    torch.cuda.synchronize()
    t_mis = time.perf_counter() - t0
    t0 = time.perf_counter()
    backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    # per-kernel HIP-event timings from one more (untimed) record + backward of
    # the same window: events around every launch would slow the timed pass
    check(se, L.vh_forward_rewind(se._h))
    L.vh_profile(se._h, 1)
    record()
    backward()
    L.vh_profile(se._h, 0)
    ms = (C.c_double * NCLASS)()
    cnt = (C.c_double * NCLASS)()
    L.vh_profile_read(se._h, ms, cnt)
    gp = gp_d.cpu().numpy()
    n3 = se.dims["ne3d"]
    # gradient w.r.t. delta's SH coefficients (l <= 32) per element: Syn^T = c_m Anl(. / W)
    gdelta_norm = 0.0
    lm_l = np.array([l for m in range(a.lmax + 1) for l in range(m, a.lmax + 1)])
    for e in range(gp.shape[0]):
        q = cm * se.sht_analysis(gp[e] / W)
        gdelta_norm += float(np.sum(np.abs(q[lm_l <= 32]) ** 2))
    se.finalize()

    # Gradient check: the derivative is the one of the recorded branch, and a
    # perturbed run of 500 sub-steps flips coastline cells and skip-gate slots
    # (the misfit is not differentiable there), so the check runs the same
    # window with the coastline held (sle_fixed_ocean), the gate off
    # (skip_tol = 0) and the sea-level fixed point converged tightly
    # (sle_tol 1e-11: the adjoint differentiates the converged solve, not the
    # warm start of a truncated one), adjoint and central differences
    # (h = 1e-4) on fresh models from the initial state.
    p.sle_fixed_ocean = 1
    p.skip_tol = 0.0
    p.sle_tol = 1e-11
    if a.fd_dirs > 0:
        se = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
        se.enable_visc_3d(c["pert"])
        n0 = se.scalar(3)
        record()
        misfit_grad(rsl_d)
        branch = (se.scalar(3) - n0, se.dims["ne3d"])
        backward()
        gp = gp_d.cpu().numpy()
        gi = gi_d.cpu().numpy()
        se.finalize()

    def J_of(pert, icearr):
        m = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
        m.enable_visc_3d(pert)
        iced = torch.from_numpy(icearr).to("cuda:0")
        n0_ = m.scalar(3)
        Jv = 0.0
        for k in range(ncalls):
            check(m, L.vh_update_device(m._h, C.c_void_p(iced[k].data_ptr()), c["dt_yr"]))
            if k in outs:
                r = site_values(m, m.rsl) - target[k]
                Jv += float(np.sum(r * r))
        br = (m.scalar(3) - n0_, m.dims["ne3d"])
        m.finalize()
        return Jv, br

    h = 1e-4
    pert = c["pert"]
    spread = pert.max(axis=(1, 2)) - pert.min(axis=(1, 2))
    keep = np.abs(spread - p.visc3d_tol) > 0.5 * p.visc3d_tol
    helper = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    frng = np.random.default_rng(a.fd_seed)
    checks = []
    for dcount in range(a.fd_dirs):
        v = np.zeros_like(pert)
        for e in range(pert.shape[0]):
            if not keep[e]:
                continue
            q = frng.standard_normal(len(lm_l)) + 1j * frng.standard_normal(len(lm_l))
            q[lm_l > 32] = 0.0
            q[cm == 1.0] = q[cm == 1.0].real
            v[e] = helper.sht_synthesis(q)
        v /= np.abs(v).max()
        jp, bp = J_of(pert + h * v, ice_np)
        jm, bm = J_of(pert - h * v, ice_np)
        fd = (jp - jm) / (2.0 * h)
        ad = float(np.sum(gp * v))
        checks.append({"kind": "delta_lm (l<=32)", "adjoint": ad, "fd": fd,
                       "rel_err": abs(ad - fd) / max(abs(fd), 1e-300), "branch_kept": bp == branch and bm == branch})
    if a.fd_dirs > 0:
        land = (c["z_bed_eq"] > 0.0) & (ice_np.max(axis=0) > 10.0)
        vi = frng.standard_normal(ice_np.shape) * land
        hi = 0.1
        jp, bp = J_of(pert, ice_np + hi * vi)
        jm, bm = J_of(pert, ice_np - hi * vi)
        fd = (jp - jm) / (2.0 * hi)
        ad = float(np.sum(gi[1:] * vi))
        checks.append({"kind": "ice history", "adjoint": ad, "fd": fd,
                       "rel_err": abs(ad - fd) / max(abs(fd), 1e-300), "branch_kept": bp == branch and bm == branch})
    helper.finalize()

    kernel_ms = {ADJ_NAMES.get(k, NAMES.get(k)): round(ms[k], 3) for k in range(NCLASS) if cnt[k] > 0}
    ratio = (t_rec + t_mis + t_bwd) / t_fwd
    # roofline of the reverse sweep's dominant kernel, k_t3_rows_adj: per
    # launch it reads the adjoint and forward row spectra of its batches,
    # writes the adjoint rows back, reads the rate rows and updates the rate
    # gradient of its elements (all FP64)
    roof = None
    if cnt[7] > 0:
        nbat = 3 * n3
        nch = (nbat + TCHUNK - 1) // TCHUNK
        b = nbat / nch                                    # batches per launch, on average
        rows = b * 6 * (a.lmax + 1) * nlat * 16.0
        ab = 3.0 * rows + 3.0 * (b / 3.0) * ng * 8.0
        avg_s = ms[7] / cnt[7] / 1000.0
        ach = ab / avg_s / 1e9
        tf = os.path.join(ROOT, "profiles", "c3_lmax%d_adjoint_traffic.json" % a.lmax)
        tr = json.load(open(tf))["kernels"].get("k_t3_rows_adj") if os.path.exists(tf) else None
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": tr["hbm_bytes_per_launch"] if tr else None,
                "kernel": "k_t3_rows_adj", "avg_launch_us": round(avg_s * 1e6, 2), "launches": int(cnt[7]),
                "alg_bytes_per_launch": int(ab)}
    out = {
        "metric": "adjoint/forward time ratio (lmax 128 3-D, %d steps, 1 GPU)" % ncalls, "value": round(ratio, 4),
        "unit": "x", "n_gpus": 1, "steps": ncalls, "warmup": 1,
        "ms_per_step": round((t_rec + t_bwd) / ncalls * 1000.0, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c3 adjoint of the lmax %d 3-D + SLE + rotation workload, %d calls of %g yr, "
                               "checkpoint every %d calls, %d sites x %d output calls"
                               % (a.lmax, ncalls, c["dt_yr"], every, nsite, nout),
                   "lmax": a.lmax, "nlat": nlat, "nphi": nphi, "ne3d": int(n3), "sub_steps": int(nsub)},
        "forward_s": round(t_fwd, 4), "record_s": round(t_rec, 4), "backward_s": round(t_bwd, 4),
        "misfit_gradient_s": round(t_mis, 4), "ratio_excl_misfit": round((t_rec + t_bwd) / t_fwd, 4),
        "J": J, "grad_delta_lm_norm": gdelta_norm ** 0.5,
        "grad_check": {"h": h, "branch": "coastline held (sle_fixed_ocean), skip gate off, sle_tol 1e-11", "directions": checks,
                       "max_rel_err": max((x["rel_err"] for x in checks), default=None)},
        "kernel_ms_reverse_run": kernel_ms, "roofline": roof, "cpu_baseline": None,
    }
    if emit:
        print(json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--offset", type=int, default=200,
                    help="untimed calls before the warm-up (the timed window samples mid-history)")
    ap.add_argument("--lmax", type=int, default=128)
    ap.add_argument("--cpu-calls", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", type=int, default=0, choices=(0, 1),
                    help="1: replay each forward-Euler sub-step as a captured hipGraph (VH_T_GRAPH)")
    ap.add_argument("--members", type=int, default=64, help="ensemble size (--config 4)")
    ap.add_argument("--ens-workers", type=int, default=16, help="host threads of the ensemble (--config 4)")
    ap.add_argument("--tune-calls", type=int, default=2, help="calls per autotuner candidate (--config 4)")
    ap.add_argument("--ens-calls", type=int, default=10,
                    help="timed calls per member of the ensemble part of the default line")
    ap.add_argument("--no-extras", action="store_true",
                    help="default line without the ensemble and adjoint parts")
    ap.add_argument("--adj-calls", type=int, default=100, help="coupling calls of the adjoint window (--config 3)")
    ap.add_argument("--fd-dirs", type=int, default=20, help="finite-difference check directions (--config 3)")
    ap.add_argument("--fd-seed", type=int, default=5)
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (5 times): the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: the solid-earth step runs on one GPU (there is no multi-GPU path); use --gpus 1", file=sys.stderr)
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
    from fastearth3d_amd import _abi, configs, lib
    L = lib.load()
    L.vh_update_device.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
    L.vh_profile.argtypes = [C.c_void_p, C.c_int]
    L.vh_profile_read.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    if a.config == 4:
        ensemble_bench(a, L)
        return
    if a.config == 3:
        adjoint_bench(a, L)
        return

    c = configs.config(a.config, lmax=a.lmax)
    p = c["params"]
    se = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        se.enable_visc_3d(c["pert"])
    se.set_tuning(_abi.T_GRAPH, a.graph)
    # the timed calls start --offset calls into the history (run untimed
    # first), where the ice load and the sea-level iteration counts are those
    # of the bulk of the run rather than of its first, nearly ice-free calls
    off = max(0, min(a.offset, c["steps"] - a.warmup - a.steps))
    nsteps = off + a.warmup + a.steps
    ice = torch.from_numpy(np.stack([c["ice"](k) for k in range(nsteps)])).to("cuda:0")
    torch.cuda.synchronize()

    def step(k):
        rc = L.vh_update_device(se._h, C.c_void_p(ice[k].data_ptr()), c["dt_yr"])
        if rc:
            raise RuntimeError(se._err())

    for k in range(off + a.warmup):
        step(k)
    L.vh_profile(se._h, 1)   # clear the per-kernel totals
    L.vh_profile(se._h, 0)
    n_acc0 = se.scalar(3)
    n_in0 = se.scalar(6)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(off + a.warmup, nsteps):
        # per-kernel HIP events on every PROF_EVERY-th call of the timed region
        # (an event pair around each launch costs ~8 % of the call when on everywhere)
        sampled = (k - off - a.warmup) % PROF_EVERY == PROF_EVERY - 1 or (a.steps < PROF_EVERY and k == nsteps - 1)
        if sampled:
            L.vh_profile(se._h, 2)
        step(k)
        if sampled:
            L.vh_profile(se._h, 0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = (C.c_double * NCLASS)()
    cnt = (C.c_double * NCLASS)()
    L.vh_profile_read(se._h, ms, cnt)
    nsub = se.scalar(3) - n_acc0
    n_inner = se.scalar(6) - n_in0
    kyr = a.steps * c["dt_yr"] / 1000.0
    value = kyr / el

    d = dict(se.dims)
    d["nlat"], d["nphi"] = se.nlat, se.nphi
    nbat = 3 * d.get("ne3d", 0)
    # the tensor kernels run in chunks of up to TCHUNK batches; the average launch
    # covers nbat / (number of chunks) batches
    nch = (nbat + TCHUNK - 1) // TCHUNK if nbat else 1
    tf = os.path.join(ROOT, "profiles", "c%d_lmax%d_traffic.json" % (a.config, a.lmax))
    traffic = json.load(open(tf))["kernels"] if os.path.exists(tf) else {}
    per_kernel = {}
    for k in NAMES:
        if cnt[k] <= 0:
            continue
        # sea-level loop: iterations per launch over the whole timed region (one launch per sub-step)
        ab = alg_bytes(d, k, nbat / nch if nbat else None, n_inner / max(nsub, 1) if k == 6 else None)
        if not ab:
            continue
        avg_s = ms[k] / cnt[k] / 1000.0
        t = traffic.get(NAMES[k])
        fl = alg_flops(d, k, nbat / nch if nbat else None)
        if fl:   # the tensor Legendre kernels: FP64 MFMA-bound
            ach = fl / avg_s / 1e12
            per_kernel[NAMES[k]] = {"bound": "mfma", "achieved": round(ach, 2), "peak": FP64_PEAK_TFLOPS,
                                    "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 4),
                                    "traffic": t["hbm_bytes_per_launch"] if t else None,
                                    "avg_launch_us": round(avg_s * 1e6, 2), "launches": int(cnt[k]),
                                    "alg_flops_per_launch": int(fl), "alg_bytes_per_launch": int(ab),
                                    "total_ms": round(ms[k], 3)}
            continue
        ach = ab / avg_s / 1e9
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
               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())), "kind": "port", **host_cpu(),
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
                   "sub_steps_timed": int(nsub), "ne3d": int(se.dims.get("ne3d", 0)),
                   "timed_calls": [off + a.warmup, nsteps - 1], "sle_inner_per_sub_step": round(n_inner / max(nsub, 1), 2),
                   "graph_replay": bool(a.graph)},
        "roofline": roof, "kernel_roofline": per_kernel, "kernel_ms": kernel_ms, "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 2)
    se.finalize()
    if not a.no_extras:
        # the metric's other two parts, measured in the same run: the uneven
        # ensemble (--ens-calls calls of each of the 64 members, after one
        # warm-up call and the autotuner's calls) and the adjoint / forward
        # time ratio over 100 calls (gradient check left to --config 3)
        ea = copy.copy(a)
        ea.steps, ea.warmup = a.ens_calls, 1
        e = ensemble_bench(ea, L, emit=False)
        out["ens_per_s"] = e["value"]
        out["ens_sequential_per_s"] = e["sequential_ens_per_s"]
        out["ens_speedup_vs_sequential"] = e["speedup_vs_sequential"]
        out["ens_detail"] = {"members": e["config"]["members"], "calls_per_member_timed": a.ens_calls,
                             "sub_steps_timed": e["config"]["sub_steps_timed"],
                             "tuned_sle_blocks": e["config"]["tuned_sle_blocks"],
                             "bound_by_slowest_member": e["bound_by_slowest_member"],
                             "sequential_share_3d": e["sequential_share_3d"],
                             "ensemble_equals_sequential": e["ensemble_equals_sequential"],
                             "basis": "ens/s = members x timed calls / 200 / wall s (completed 200-call members per "
                                      "second at the measured rate)"}
        aa = copy.copy(a)
        aa.fd_dirs = 0
        r = adjoint_bench(aa, L, emit=False)
        out["adjoint_forward_ratio"] = r["value"]
        out["adjoint_detail"] = {"calls": r["steps"], "forward_s": r["forward_s"], "record_s": r["record_s"],
                                 "misfit_gradient_s": r["misfit_gradient_s"], "backward_s": r["backward_s"],
                                 "ratio_excl_misfit": r["ratio_excl_misfit"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
