"""hipGraph replay of the forward-Euler sub-step (VH_T_GRAPH): a model that
replays each sub-step (ice interpolation, drift, fused sea-level solve, 3-D
tensor transforms and the memory advance) as one captured graph gives the
same results bit for bit as one launching the kernels one by one, after
every coupling call of the lateral-viscosity workload (lmax 16: sea level
with migrating coastlines, rotation, 3-D elements), with eager work (the
rotation channels, read-backs, getters) between the replays; the same with
the toroidal channels carried, across a re-enabling of the lateral field
(the graphs are dropped and captured again) and after switching the replay
off and on.
"""
import numpy as np
import pytest

from fastearth3d_amd import _abi, configs

pytestmark = pytest.mark.gpu


def _same(a, b, k):
    assert np.array_equal(a.rsl, b.rsl), k
    assert np.array_equal(a.C, b.C), k
    assert np.array_equal(a.sigma_lm, b.sigma_lm), k
    assert np.array_equal(a.memory(), b.memory()), k
    assert a.scalar(_abi.S_N_ACCEPT) == b.scalar(_abi.S_N_ACCEPT), k


@pytest.mark.parametrize("tor", [0, 1])
def test_graph_replay_bit_identical(tor):
    from fastearth3d_amd import lib
    c = configs.config(2, lmax=16)
    p = c["params"]
    p.l_toroidal = tor
    eager = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    graph = lib.SolidEarth(p, c["z_bed_eq"], c["h_ice_eq"], c["h_ice_init"])
    graph.set_tuning(_abi.T_GRAPH, 1)
    for m in (eager, graph):
        m.enable_visc_3d(c["pert"])
    k = 0
    for _ in range(4):
        for m in (eager, graph):
            m.update(c["ice"](k + 60), c["dt_yr"])
        _same(eager, graph, k)
        k += 1
    n_sub = graph.scalar(_abi.S_N_ACCEPT)
    assert graph.scalar(_abi.S_N_GRAPH_LAUNCH) == n_sub > 0
    assert 0 < graph.scalar(_abi.S_N_GRAPH_CAPTURE) < n_sub   # replays after the first call
    assert eager.scalar(_abi.S_N_GRAPH_LAUNCH) == 0
    # a new lateral field reallocates the tensor buffers: graphs captured again
    pert2 = 0.8 * c["pert"]
    for m in (eager, graph):
        m.enable_visc_3d(pert2)
    caps = graph.scalar(_abi.S_N_GRAPH_CAPTURE)
    for _ in range(3):
        for m in (eager, graph):
            m.update(c["ice"](k + 60), c["dt_yr"])
        _same(eager, graph, k)
        k += 1
    assert graph.scalar(_abi.S_N_GRAPH_CAPTURE) > caps
    # off and on again
    graph.set_tuning(_abi.T_GRAPH, 0)
    for m in (eager, graph):
        m.update(c["ice"](k + 60), c["dt_yr"])
    _same(eager, graph, k)
    k += 1
    graph.set_tuning(_abi.T_GRAPH, 1)
    for _ in range(2):
        for m in (eager, graph):
            m.update(c["ice"](k + 60), c["dt_yr"])
        _same(eager, graph, k)
        k += 1
    for m in (eager, graph):
        m.finalize()
