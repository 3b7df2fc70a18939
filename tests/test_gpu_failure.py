"""Failure paths of the C ABI, forced from the host (no kernel faults, traps or
hangs): simulated device allocation failure (RTNN_FAIL_ALLOC_BYTES), graphs and
contexts destroyed in either order, autotuner candidates a workload cannot use,
and the workloads the hierarchical Morton grid cannot hold (a box wider than
1024 top-level cells, k above the one-lane kernels' 64), which fall back to the
raster grid and the wider group kernels.

The reference has no counterpart: it throws on a CUDA error and exits
(sutil/Exception.h:44-55, main.cpp:154-158) and avoids running out of memory by
modelling it (util.cpp:551-676).
"""
import gc
import os

import numpy as np
import pytest

import oracle
import rtnn_amd

pytestmark = pytest.mark.gpu


@pytest.fixture
def fresh():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = rtnn_amd.Context(0)
    yield c
    c.close()


def test_forced_allocation_failure_reports_oom_and_recovers():
    # Every allocation of at least 64 MiB fails while the knob is set: the
    # search returns RTNN_ERR_OOM, the context stays usable once it is
    # cleared, and closing the context returns every tracked byte.
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 30, (200000, 3)).astype(np.float32)
    gc.collect()
    base = rtnn_amd.mem_stats()
    c = rtnn_amd.Context(0)
    try:
        c.set_points(pts)
        os.environ["RTNN_FAIL_ALLOC_BYTES"] = str(64 << 20)
        try:
            with pytest.raises(rtnn_amd.RtnnError) as ei:
                # 200k rows x 128 columns x 4 B = 102 MB of output buffers
                c.search(None, mode="radius", radius=1.0, k=128)
            assert ei.value.code == rtnn_amd.ERR_OOM
        finally:
            del os.environ["RTNN_FAIL_ALLOC_BYTES"]
        idx, d2, _ = c.search(None, mode="radius", radius=1.0, k=128)
        sel = np.arange(0, len(pts), 997, dtype=np.uint64)
        wi, wd = oracle.grid_search(pts, pts, oracle.MODE_RADIUS, 1.0, 128, qsel=sel)
        assert np.array_equal(idx[sel.astype(np.int64)], wi)
        assert np.array_equal(d2[sel.astype(np.int64)], wd)
    finally:
        c.close()
    gc.collect()
    assert rtnn_amd.mem_stats()["device_bytes"] == base["device_bytes"]
    assert rtnn_amd.mem_stats()["pinned_bytes"] == base["pinned_bytes"]


def test_forced_allocation_failure_in_set_points(fresh):
    pts = np.random.default_rng(4).uniform(0, 10, (100000, 3)).astype(np.float32)
    os.environ["RTNN_FAIL_ALLOC_BYTES"] = "1"  # every allocation
    try:
        with pytest.raises(rtnn_amd.RtnnError) as ei:
            fresh.set_points(pts)
        assert ei.value.code == rtnn_amd.ERR_OOM
    finally:
        del os.environ["RTNN_FAIL_ALLOC_BYTES"]
    fresh.set_points(pts)
    idx, _, _ = fresh.search(None, mode="knn", radius=0.5, k=4)
    assert idx.shape == (100000, 4)


def _graph_inputs():
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(7, 6, 500, 3000, h=0.1, jit=0.02)
    return torch.from_numpy(xyz).cuda(), offs


def test_graph_outlives_its_context():
    # The natural teardown order (graph after context) once freed a graph the
    # context had already freed.  Now the handle is rejected, not touched.
    L = rtnn_amd.lib()
    dxyz, offs = _graph_inputs()
    c = rtnn_amd.Context(0)
    g = c.graph_batched(dxyz, offs, mode="radius", radius=0.2, k=32)
    g.launch()
    h = g._h
    L.rtnn_destroy(c._h)  # the context (and with it the graph) first
    c._h = None
    assert L.rtnn_graph_launch(h, 1, None) == rtnn_amd.ERR_INVALID
    assert L.rtnn_graph_destroy(h) == rtnn_amd.ERR_INVALID
    g._h = None
    # a destroyed graph's handle is rejected the same way
    c2 = rtnn_amd.Context(0)
    try:
        g2 = c2.graph_batched(dxyz, offs, mode="radius", radius=0.2, k=32)
        h2 = g2._h
        assert L.rtnn_graph_destroy(h2) == rtnn_amd.OK
        g2._h = None
        assert L.rtnn_graph_destroy(h2) == rtnn_amd.ERR_INVALID
        assert L.rtnn_graph_launch(h2, 1, None) == rtnn_amd.ERR_INVALID
    finally:
        c2.close()


def test_context_closed_with_a_live_graph_then_graph_closed():
    # Python teardown in either order ends normally.
    dxyz, offs = _graph_inputs()
    c = rtnn_amd.Context(0)
    g = c.graph_batched(dxyz, offs, mode="knn", radius=0.25, k=8)
    g.launch()
    ref_i = g.idx.clone()
    c.close()
    g.close()
    c = rtnn_amd.Context(0)
    g = c.graph_batched(dxyz, offs, mode="knn", radius=0.25, k=8)
    g.launch()
    import torch
    assert torch.equal(g.idx, ref_i)
    g.close()
    c.close()


def test_wide_box_falls_back_to_the_raster_grid(fresh):
    # A box 2000 r wide needs 11 Morton bits per axis at the top level: the
    # partitioned and learned KNN (hierarchical grid by default for the
    # partition) take the raster grid and stay exact.
    rng = np.random.default_rng(8)
    pts = np.concatenate([rng.uniform([0, 0, 0], [2000, 3, 3], (60000, 3)),
                          rng.normal([1000, 1.5, 1.5], 0.3, (5000, 3))]).astype(np.float32)
    qs = rng.uniform([0, 0, 0], [2000, 3, 3], (20000, 3)).astype(np.float32)
    fresh.set_points(pts)
    pi, pd, _ = fresh.search(qs, mode="knn", radius=1.0, k=8)
    for levels in (0, 5, 8):
        fresh.set_options(grid_levels=levels)
        li, ld, _ = fresh.search_partitioned(qs, radius=1.0, k=8)
        assert np.array_equal(pi, li) and np.array_equal(pd, ld), levels
    fresh.set_options()
    sel = np.arange(0, len(qs), 53, dtype=np.uint64)
    wi, wd = oracle.bruteforce(pts, qs, oracle.MODE_KNN, 1.0, 8, qsel=sel, threads=8)
    assert np.array_equal(pi[sel.astype(np.int64)], wi)
    assert np.array_equal(pd[sel.astype(np.int64)], wd)
    # the autotuner skips no candidate that fails on this box: it finishes
    opt, ms = fresh.autotune(qs[:5000], method="partitioned", mode="knn", radius=1.0, k=8)
    assert ms > 0


@pytest.mark.parametrize("k", [100, 300])
def test_proposal_methods_beyond_k64(fresh, k):
    # k above the register top-k (64): the stages run on 16- or 64-lane
    # groups, on both grids; rows equal the plain KNN.
    rng = np.random.default_rng(k)
    pts = rng.uniform(0, 25, (80000, 3)).astype(np.float32)
    qs = rng.uniform(-2, 27, (8000, 3)).astype(np.float32)
    fresh.set_points(pts)
    pi, pd, _ = fresh.search(qs, mode="knn", radius=3.0, k=k)
    w = np.random.default_rng(1).normal(0, 0.2, rtnn_amd.MLP_NPARAMS).astype(np.float32)
    fresh.set_mlp_weights(np.concatenate([w, [0.25, 1.0]]).astype(np.float32))
    for levels in (1, 5):
        fresh.set_options(grid_levels=levels)
        for fn in (fresh.search_partitioned, fresh.search_learned):
            li, ld, st = fn(qs, radius=3.0, k=k)
            assert np.array_equal(pi, li) and np.array_equal(pd, ld), (levels, fn.__name__)
    fresh.set_options()
    sel = np.arange(0, len(qs), 97, dtype=np.uint64)
    wi, wd = oracle.bruteforce(pts, qs, oracle.MODE_KNN, 3.0, k, qsel=sel, threads=8)
    assert np.array_equal(pi[sel.astype(np.int64)], wi)
