"""Several models on one GPU at once give bit-identical results to the same
models stepped one after another: the models share no device state, and the
fused sea-level solve needs none of its blocks to be resident together
(blocks take over the work items of blocks that have not started), so it
completes on a device shared with other work.

* four small models from four host threads, each on its own stream;
* two models of the benchmarked size (lmax 128, 3-D viscosity, sea level,
  rotation) from two threads;
* a model whose solves have blocks that start late (the diagnostic
  VH_SLE_LATE_BLOCKS), so that the other blocks run their work items
  (VH_S_N_SLE_STEAL > 0);
* two processes, each with its own model, on the one GPU.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from fastearth3d_amd import configs

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _members():
    out = []
    for idx, L, seed in ((1, 24, 0), (1, 20, 5), (2, 16, 0), (2, 16, 9)):
        out.append(configs.config(idx, lmax=L, seed=seed))
    return out


def _run(c, calls=3, first=40):
    from fastearth3d_amd import lib
    se = lib.SolidEarth(c["params"], c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    if c["pert"] is not None:
        se.enable_visc_3d(c["pert"])
    for k in range(calls):
        se.update(c["ice"](k + first), c["dt_yr"])
    return se.rsl.copy(), se.memory().copy(), se.C.copy(), se.scalar(10)


def _same(a, b):
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)


def test_concurrent_models_match_sequential():
    ms = _members()
    seq = [_run(c) for c in ms]
    with ThreadPoolExecutor(len(ms)) as ex:
        par = list(ex.map(_run, ms))
    for s, p in zip(seq, par):
        _same(s, p)


def test_full_size_models_from_two_threads():
    ms = [configs.config(2, seed=0), configs.config(2, seed=3)]
    seq = [_run(c, calls=2, first=60) for c in ms]
    with ThreadPoolExecutor(2) as ex:
        par = list(ex.map(lambda c: _run(c, calls=2, first=60), ms))
    for s, p in zip(seq, par):
        _same(s, p)
    print("items run for non-resident blocks:", [p[3] for p in par])


@pytest.mark.parametrize("late", ["100:300", "100:200000"])
def test_solve_with_blocks_that_start_late(monkeypatch, late):
    """The diagnostic VH_SLE_LATE_BLOCKS=count:microseconds makes the last
    blocks of every sea-level solve (here 100 of 130: orders and latitude
    pairs) start late: 300 us (they rejoin during the solve, reading the
    published record) or 200 ms (the solve ends without them). The other
    blocks run their work items; the result is the same bit for bit."""
    c = configs.config(1, lmax=64)
    ref = _run(c, calls=3)
    monkeypatch.setenv("VH_SLE_LATE_BLOCKS", late)
    got = _run(c, calls=3)
    _same(ref, got)
    assert got[3] > 0


def test_two_processes_on_one_gpu(tmp_path):
    """Two child processes, each stepping its own model, at the same time;
    each equals this process's sequential run bit for bit."""
    c = configs.config(1, lmax=48)
    ref = _run(c, calls=4)
    outs = [str(tmp_path / ("p%d.npz" % i)) for i in range(2)]
    script = os.path.join(HERE, "child_model_run.py")
    procs = [subprocess.Popen([sys.executable, script, "1", "48", "4", o]) for o in outs]
    rcs = [p.wait(timeout=300) for p in procs]
    assert rcs == [0, 0]
    for o in outs:
        f = np.load(o)
        _same(ref, (f["rsl"], f["memory"], f["C"]))
