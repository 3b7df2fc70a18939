"""GPU parity of the search path through the C ABI against the CPU oracle.

Bar: indices bit-exact and d2 bit-exact (FP32, no FMA on either side), rows in
(d2, index) order, -1 / +inf in empty slots.  Covers the reference's own data
file, the workload shapes at oracle-checkable sizes, every kernel variant
(thread-per-query and 4/16/64-lane groups, cell factors), edge cases (empty
rows, duplicates, one hot cell, far and NaN queries, k above the hit count,
buffer overflow of the group kernel), errors, batched clouds and determinism.
"""
import numpy as np
import pytest

import oracle
import rtnn_amd
from conftest import golden_d2

pytestmark = pytest.mark.gpu

MODES = {"radius": oracle.MODE_RADIUS, "knn": oracle.MODE_KNN}
# lanes per query (64 / queries_per_wave) -> largest k of that kernel
MAX_K = {1: 64, 2: 16, 4: 64, 8: 64, 16: 256, 64: 1024}


def check(ctx, pts, qs, mode, r, k, qsel=None, **search_kw):
    idx, d2, st = ctx.search(qs, mode=mode, radius=r, k=k, **search_kw)
    if hasattr(idx, "cpu"):
        idx, d2 = idx.cpu().numpy(), d2.cpu().numpy()
    qarr = pts if qs is None else (qs.cpu().numpy() if hasattr(qs, "cpu") else qs)
    parr = pts.cpu().numpy() if hasattr(pts, "cpu") else pts
    wi, wd = oracle.grid_search(parr, qarr, MODES[mode], r, k, qsel=qsel, threads=8)
    gi, gd = (idx, d2) if qsel is None else (idx[qsel], d2[qsel])
    bad = np.nonzero((gi != wi).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ, first {bad[:5]}: got {gi[bad[0]]} want {wi[bad[0]]}"
    assert np.array_equal(gd, wd)
    return st


@pytest.fixture(autouse=True)
def _reset_opts(ctx):
    ctx.set_options()
    yield
    ctx.set_options()


def test_samplepc_golden(ctx, golden, samplepc):
    ctx.set_points(samplepc)
    for case in golden["cases"]:
        for qpw in (0, 1, 4, 8, 16, 32, 64):
            if qpw and case["k"] > MAX_K[64 // qpw]:
                continue
            ctx.set_options(queries_per_wave=qpw)
            idx, d2, _ = ctx.search(None, mode=case["mode"], radius=case["r"], k=case["k"])
            assert idx.tolist() == case["idx"], (case["mode"], case["r"], case["k"], qpw)
            assert np.array_equal(d2, golden_d2(case["d2"]))
            idx, d2, _ = ctx.search(samplepc, mode=case["mode"], radius=case["r"], k=case["k"])
            assert idx.tolist() == case["idx"]


def test_device_tensors_in_and_out(ctx, samplepc):
    import torch
    p = torch.from_numpy(samplepc).cuda()
    ctx.set_points(p)
    idx, d2, _ = ctx.search(p, mode="knn", radius=10.0, k=3)
    assert idx.is_cuda and idx.cpu().tolist()[0] == [2, 4, 6]


def test_config1_range_full(ctx):
    # 100,000 points U[0,50)^3, queries = points, r = 2, K = 50 (range).
    pts = rtnn_amd.gen_uniform(1, 100000, 0.0, 50.0)
    ctx.set_points(pts)
    st = check(ctx, pts, None, "radius", 2.0, 50)
    assert st.search_launches == 1
    check(ctx, pts, None, "knn", 2.0, 50)


@pytest.mark.parametrize("qpw", [1, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("cf", [0.5, 1.0, 2.0, 3.0])
def test_kernel_variants_and_cell_factors(ctx, qpw, cf):
    rng = np.random.default_rng(int(cf * 10) + qpw)
    pts = rng.uniform(0, 20, (20000, 3)).astype(np.float32)
    qs = np.concatenate([pts[:3000], rng.uniform(-3, 23, (2000, 3)).astype(np.float32)])
    ctx.set_points(pts)
    # group kernels: wide and compact LDS buffers, and the wave-tiled kernel
    for buf, tile in (((0, 0), (1, 0), (1, 1), (1, 256), (1, 1024)) if qpw in (1, 4, 8, 16) else
                      ((0, 0), (1, 0)) if qpw == 32 else ((0, 0),)):
        ctx.set_options(cell_factor=cf, queries_per_wave=qpw, group_buffer=buf, lds_tile=tile)
        for mode in ("radius", "knn"):
            for k in (1, 5, 16, 33, 64):
                if k <= MAX_K[64 // qpw]:
                    check(ctx, pts, qs, mode, 1.5, k)


def test_large_k_group_kernel(ctx):
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 10, (30000, 3)).astype(np.float32)
    ctx.set_points(pts)
    # queries_per_wave 4 -> 16 lanes per query, 1 -> 64 lanes per query
    for qpw, k in [(4, 100), (4, 256), (1, 300), (1, 1024), (0, 200)]:
        ctx.set_options(queries_per_wave=qpw)
        check(ctx, pts, pts[:2000], "radius", 1.2, k)
        check(ctx, pts, pts[:500], "knn", 1.2, k)


def test_group_buffer_overflow_path(ctx):
    # Thousands of hits per query force the keep-k-smallest select of the LDS
    # buffer many times before the final sort.
    rng = np.random.default_rng(8)
    pts = rng.normal(0, 1.0, (60000, 3)).astype(np.float32)
    ctx.set_points(pts)
    # wide (2k + G keys) and compact (k + 2G keys) buffers, and the tiled kernel
    for buf, tile in ((2, 0), (1, 0), (1, 1)):
        for qpw in (1, 4, 8, 16, 32):  # 64, 16, 8, 4 and 2 lanes per query
            if qpw == 32 and tile:
                continue
            ctx.set_options(queries_per_wave=qpw, group_buffer=buf, lds_tile=tile)
            check(ctx, pts, pts[:1500], "radius", 2.0, 16)
            if MAX_K[64 // qpw] >= 64:
                check(ctx, pts, pts[:1500], "knn", 2.0, 64)
            check(ctx, pts, pts[:300], "radius", 2.0, 5)


def test_lattice_ties_in_buffer_cuts(ctx):
    # An exact integer lattice: whole shells of points at one d2 (6 at 1, 12 at
    # 2, 8 at 3, 6 at 4, 24 at 5, ...).  A full buffer's d2-threshold step
    # then cannot stop between tied entries, so it keeps whole shells and the
    # exact select takes over; the rows must still be the (d2, index) order.
    g = np.arange(0, 24, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = np.ascontiguousarray(pts[np.random.default_rng(5).permutation(len(pts))])
    qs = np.concatenate([pts[:700], pts[:300] + np.float32(0.5)])
    ctx.set_points(pts)
    for buf in (2, 1):
        for qpw in (1, 4, 8, 16, 32):
            ctx.set_options(queries_per_wave=qpw, group_buffer=buf)
            for k in (5, 16):
                check(ctx, pts, qs, "knn", 2.6, k)
                check(ctx, pts, qs, "radius", 2.6, k)
            if MAX_K[64 // qpw] >= 32:
                check(ctx, pts, qs[:400], "knn", 3.1, 32)


def test_knn_separate_queries_gaussian_mixture(ctx):
    # Config-2 shape at 1/10 size: clusters + background, separate queries,
    # K = 8, r = 20; many background rows are short (sentinels).
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 100000)
    qs = rtnn_amd.gen_gaussian_mixture(2, 21, 100000)
    ctx.set_points(pts)
    check(ctx, pts, qs, "knn", 20.0, 8)
    check(ctx, pts, qs, "radius", 20.0, 8)


def test_spheres_knn16(ctx):
    # Config-3 shape at 1/50 size: sphere surfaces, K = 16, r = 10.
    pts = rtnn_amd.gen_spheres(3, 30, 200000, nspheres=40)
    qs = rtnn_amd.gen_spheres(3, 31, 100000, nspheres=40)
    ctx.set_points(pts)
    check(ctx, pts, qs, "knn", 10.0, 16)


def test_caller_owned_outputs(ctx):
    import torch
    rng = np.random.default_rng(31)
    pts = rng.uniform(0, 10, (5000, 3)).astype(np.float32)
    ctx.set_points(torch.from_numpy(pts).cuda())
    ref_i, ref_d, _ = ctx.search(None, mode="knn", radius=1.0, k=8)
    oi = torch.full((5000, 8), 7, dtype=torch.int32, device="cuda")
    od = torch.zeros((5000, 8), dtype=torch.float32, device="cuda")
    for _ in range(2):  # the same buffers, written again
        i, d, _ = ctx.search(None, mode="knn", radius=1.0, k=8, out=(oi, od))
        assert i.data_ptr() == oi.data_ptr() and d.data_ptr() == od.data_ptr()
        assert torch.equal(i, ref_i) and torch.equal(d, ref_d)
    hi = np.empty((5000, 8), np.int32)
    i, d, _ = ctx.search(None, mode="knn", radius=1.0, k=8, out=(hi, None))
    assert d is None and np.array_equal(hi, ref_i.cpu().numpy())
    with pytest.raises(ValueError):
        ctx.search(None, mode="knn", radius=1.0, k=8, out=(oi[:, :4], od))
    with pytest.raises(ValueError):
        ctx.search(None, mode="knn", radius=1.0, k=8, out=(oi, od.cpu()))
    with pytest.raises(ValueError):  # a transposed view: right shape, not contiguous
        ctx.search(None, mode="knn", radius=1.0, k=8,
                   out=(oi, torch.zeros((8, 5000), dtype=torch.float32, device="cuda").t()))


@pytest.mark.parametrize("opts", [{}, {"queries_per_wave": 4, "lds_tile": 1}])
def test_edge_cases(ctx, opts):
    ctx.set_options(**opts)
    rng = np.random.default_rng(12)
    # duplicates + one hot cell
    base = rng.uniform(0, 1e-3, (500, 3)).astype(np.float32)
    pts = np.concatenate([base, base[:100], rng.uniform(0, 10, (500, 3)).astype(np.float32)])
    ctx.set_points(pts)
    for k in (3, 50, 128):
        check(ctx, pts, pts, "radius", 0.5, k)
        check(ctx, pts, pts, "knn", 0.5, k)
    # far queries, a NaN query, a query on a grid corner
    qs = np.array([[1e6, 1e6, 1e6], [-1e6, 0, 0], [np.nan, 0, 0], [0, 0, 0], [10, 10, 10],
                   [5, 5, 5]], np.float32)
    idx, d2, _ = ctx.search(qs, mode="radius", radius=2.0, k=8)
    assert (idx[:3] == -1).all() and np.isinf(d2[:3]).all()
    check(ctx, pts, qs[3:], "knn", 2.0, 8)
    # radius larger than the whole box
    check(ctx, pts, pts[:200], "knn", 1e4, 20)
    # a single point
    one = np.array([[1, 2, 3]], np.float32)
    ctx.set_points(one)
    idx, _, _ = ctx.search(None, mode="radius", radius=1.0, k=4)
    assert idx.tolist() == [[0, -1, -1, -1]]
    idx, _, _ = ctx.search(None, mode="knn", radius=1.0, k=4)
    assert idx.tolist() == [[-1, -1, -1, -1]]


def test_boundary_pairs_listed_and_exact(ctx):
    # Lattice spacing 0.1 with r = 0.2: many pairs sit on the sphere boundary.
    xyz, offs = rtnn_amd.gen_lattice_clouds(4, 1, 3000, 3000, jit=0.0)
    ctx.set_points(xyz)
    check(ctx, xyz, None, "radius", 0.2, 64)
    cnt, want = oracle.boundary_pairs(xyz, xyz, 0.2, max_out=1 << 20)
    assert cnt > 0  # the boundary is exercised (and the check above was exact)
    got = ctx.boundary_pairs(None, radius=0.2)
    assert np.array_equal(got.astype(np.uint64), want)
    # separate queries, a wider tolerance
    rng = np.random.default_rng(8)
    qs = xyz[rng.integers(0, len(xyz), 500)] + rng.uniform(-1e-3, 1e-3, (500, 3)).astype(np.float32)
    cnt, want = oracle.boundary_pairs(xyz, qs, 0.2, tol=1e-2, max_out=1 << 20)
    got = ctx.boundary_pairs(qs, radius=0.2, tol=1e-2)
    assert cnt > 0 and np.array_equal(got.astype(np.uint64), want)


def test_boundary_pairs_batched(ctx):
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(5, 3, 500, 3000)
    offs = np.insert(offs, 1, offs[1])  # an empty cloud
    got = ctx.boundary_pairs_batched(torch.from_numpy(xyz).cuda(), offs, radius=0.2, tol=1e-4)
    want = []
    for c in range(len(offs) - 1):
        a, b = int(offs[c]), int(offs[c + 1])
        if a == b:
            continue
        n, pr = oracle.boundary_pairs(xyz[a:b], xyz[a:b], 0.2, tol=1e-4, max_out=1 << 20)
        pr = pr.copy()
        pr[:, 0] += a
        want.append(pr)
    want = np.concatenate(want)
    assert len(want) > 0 and np.array_equal(got.astype(np.uint64), want)


def test_errors(ctx, samplepc):
    ctx.set_points(samplepc)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.search(None, mode="radius", radius=0.0, k=4)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.search(None, mode="radius", radius=1.0, k=0)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.search(None, mode="radius", radius=float("nan"), k=4)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.search(None, mode="radius", radius=1.0, k=5000)
    bad = samplepc.copy()
    bad[3, 1] = np.inf
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.set_points(bad)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.set_options(queries_per_wave=3)
    with pytest.raises(rtnn_amd.RtnnError, match="lds_tile"):
        ctx.set_options(lds_tile=300)


def test_deterministic(ctx):
    pts = rtnn_amd.gen_uniform(7, 50000, 0.0, 30.0)
    ctx.set_points(pts)
    a, ad, _ = ctx.search(None, mode="radius", radius=1.5, k=32)
    b, bd, _ = ctx.search(None, mode="radius", radius=1.5, k=32)
    assert np.array_equal(a, b) and np.array_equal(ad.view(np.int32), bd.view(np.int32))


@pytest.mark.parametrize("qpw,tile", [(0, 0), (1, 0), (16, 0), (16, 1), (4, 1), (16, 256),
                                      (4, 1024)])
def test_batched_clouds(ctx, qpw, tile):
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(9, 40, 1000, 8000)
    ctx.set_options(queries_per_wave=qpw, lds_tile=tile)
    t = torch.from_numpy(xyz).cuda()
    bi, bd, st = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)
    pi, pd, _ = ctx.search_percloud(t, offs, mode="radius", radius=0.2, k=64)
    bi, bd = bi.cpu().numpy(), bd.cpu().numpy()
    assert np.array_equal(bi, pi.cpu().numpy())
    assert np.array_equal(bd, pd.cpu().numpy())
    for c in (0, 17, 39):
        a, b = int(offs[c]), int(offs[c + 1])
        cloud = xyz[a:b]
        wi, wd = oracle.grid_search(cloud, cloud, oracle.MODE_RADIUS, 0.2, 64, threads=8)
        assert np.array_equal(bi[a:b], wi) and np.array_equal(bd[a:b], wd)
    ki, kd, _ = ctx.search_batched(t, offs, mode="knn", radius=0.25, k=8)
    a, b = int(offs[5]), int(offs[6])
    wi, wd = oracle.grid_search(xyz[a:b], xyz[a:b], oracle.MODE_KNN, 0.25, 8, threads=8)
    assert np.array_equal(ki.cpu().numpy()[a:b], wi)


@pytest.mark.parametrize("mode,k", [("radius", 64), ("knn", 8)])
def test_graph_batched_replays(ctx, mode, k):
    # The captured pipeline (device-side grid plan, sort, cell table, gather,
    # persistent search) equals the eager batched search on every replay, with
    # unrelated eager work and synchronisations between replays, and follows
    # new point values written into the bound tensor.
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(11, 30, 500, 6000)
    offs = np.insert(offs, 3, offs[3])  # an empty cloud
    t = torch.from_numpy(xyz).cuda()
    r = 0.2 if mode == "radius" else 0.25
    want_i, want_d, _ = ctx.search_batched(t, offs, mode=mode, radius=r, k=k)
    g = ctx.graph_batched(t, offs, mode=mode, radius=r, k=k)
    try:
        other = rtnn_amd.gen_uniform(3, 20000, 0.0, 10.0)
        for rep in range(4):
            g.idx.fill_(-7)
            g.d2.fill_(-7.0)
            st = g.launch()
            assert st.ms_total > 0
            assert torch.equal(g.idx, want_i), rep
            assert torch.equal(g.d2, want_d), rep
            ctx.set_points(other)  # eager work on the same context and stream
            ctx.search(None, mode="radius", radius=0.5, k=16)
            torch.cuda.synchronize()
        # new values in the bound tensor: one cloud stretched (needs a coarser
        # table), one translated far away
        a, b = int(offs[1]), int(offs[2])
        t[a:b] *= 3.0
        a, b = int(offs[5]), int(offs[6])
        t[a:b] += 1000.0
        g.launch()
        want_i, want_d, _ = ctx.search_batched(t, offs, mode=mode, radius=r, k=k)
        assert torch.equal(g.idx, want_i) and torch.equal(g.d2, want_d)
        h = t.cpu().numpy()
        wi, wd = oracle.grid_search(h[a:b], h[a:b], MODES[mode], r, k, threads=8)
        assert np.array_equal(g.idx[a:b].cpu().numpy(), wi)
        assert np.array_equal(g.d2[a:b].cpu().numpy(), wd)
        # a non-finite coordinate is reported by a synchronous replay
        t[0, 0] = float("nan")
        with pytest.raises(rtnn_amd.RtnnError, match="non-finite"):
            g.launch()
    finally:
        g.close()


@pytest.mark.slow
def test_config3_full_size_sampled(ctx):
    # Full config-3 shape: 10M sphere-surface points, 10M queries, K=16, r=10;
    # 4000 sampled rows checked bit-exactly, all rows checked for invariants.
    pts = rtnn_amd.gen_spheres(3, 30, 10_000_000)
    qs = rtnn_amd.gen_spheres(3, 31, 10_000_000)
    ctx.set_points(pts)
    idx, d2, st = ctx.search(qs, mode="knn", radius=10.0, k=16)
    sel = np.random.default_rng(0).choice(len(qs), 4000, replace=False).astype(np.uint64)
    wi, wd = oracle.grid_search(pts, qs, oracle.MODE_KNN, 10.0, 16, qsel=sel, threads=8)
    s = sel.astype(np.int64)
    assert np.array_equal(idx[s], wi) and np.array_equal(d2[s], wd)
    # size-independent invariants over every row
    assert (np.diff(d2, axis=1) >= 0).all()
    valid = idx >= 0
    assert (valid[:, 1:] <= valid[:, :-1]).all()  # sentinels only at the tail
    assert (d2[valid] > 0).all() and (d2[valid] < np.float32(100)).all()
    # the learned-radius search (the configuration the autotuner picks for
    # this workload) equals the plain one on every row
    ctx.mlp_train(10.0, 16, n_samples=1 << 21, steps=2000, batch=4096, lr=1e-3, tau=0.9, seed=1)
    ctx.set_options(cell_factor=2.0, queries_per_wave=32, group_buffer=2, grid_levels=1)
    li, ld, lst = ctx.search_learned(qs, radius=10.0, k=16)
    assert np.array_equal(li, idx) and np.array_equal(ld, d2)
    assert 0 < lst.fallback_queries < len(qs) // 5


def test_config2_full_size(ctx):
    # Full config-2 shape: 1M Gaussian-mixture points (64 x sigma 5 + 20 %
    # background), 1M queries, K = 8, r = 20.  2000 sampled rows equal the
    # oracle bit for bit; the partitioned and learned searches equal the plain
    # one on every row; invariants hold on every row.
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 1_000_000)
    qs = rtnn_amd.gen_gaussian_mixture(2, 21, 1_000_000)
    ctx.set_points(pts)
    idx, d2, _ = ctx.search(qs, mode="knn", radius=20.0, k=8)
    sel = np.random.default_rng(2).choice(len(qs), 2000, replace=False).astype(np.uint64)
    wi, wd = oracle.grid_search(pts, qs, oracle.MODE_KNN, 20.0, 8, qsel=sel, threads=8)
    s = sel.astype(np.int64)
    assert np.array_equal(idx[s], wi) and np.array_equal(d2[s], wd)
    pi, pd, _ = ctx.search_partitioned(qs, radius=20.0, k=8)
    assert np.array_equal(pi, idx) and np.array_equal(pd, d2)
    ctx.mlp_train(20.0, 8, n_samples=1 << 18, steps=300, batch=4096, seed=3)
    li, ld, st = ctx.search_learned(qs, radius=20.0, k=8)
    assert np.array_equal(li, idx) and np.array_equal(ld, d2)
    assert (d2[:, 1:] >= d2[:, :-1]).all()  # sorted; +inf sentinels at the tail
    valid = idx >= 0
    assert (valid[:, 1:] <= valid[:, :-1]).all()
    assert (valid.sum(axis=1) < 8).any()  # background rows with fewer than K in r


def test_config4_full_size(ctx):
    # Full config-4 shape: 4096 jittered lattices, sizes log-uniform in
    # [1e3, 2e5] (150.7M points), r = 0.2, K = 64, queries = points.  The one
    # persistent batched launch equals the per-cloud launches on every row,
    # and 40 sampled clouds x 50 rows equal the oracle.
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(4, 4096, 1000, 200000, h=0.1, jit=0.02)
    t = torch.from_numpy(xyz).cuda()
    bi, _, _ = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64, with_d2=False)
    pi, _, _ = ctx.search_percloud(t, offs, mode="radius", radius=0.2, k=64, with_d2=False)
    assert torch.equal(bi, pi)
    del pi
    rng = np.random.default_rng(4)
    for c in rng.choice(len(offs) - 1, 40, replace=False):
        a, b = int(offs[c]), int(offs[c + 1])
        sel = rng.choice(b - a, 50, replace=False).astype(np.uint64)
        wi, _ = oracle.bruteforce(xyz[a:b], xyz[a:b], oracle.MODE_RADIUS, 0.2, 64, qsel=sel,
                                  threads=8)
        rows = torch.from_numpy(sel.astype(np.int64) + a).cuda()
        assert np.array_equal(bi[rows].cpu().numpy(), wi)


def test_autotune_batched_keeps_result(ctx):
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(13, 12, 500, 4000)
    t = torch.from_numpy(xyz).cuda()
    want_i, want_d, _ = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)
    opt, ms = ctx.autotune_batched(t, offs, mode="radius", radius=0.2, k=64)
    assert ms > 0 and opt["cell_factor"] in (0.5, 1.0, 2.0)
    assert opt["queries_per_wave"] in (1, 4, 8, 16, 32, 64)
    got_i, got_d, _ = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)
    assert torch.equal(got_i, want_i) and torch.equal(got_d, want_d)


@pytest.fixture
def fresh_ctx():
    # Size classes stay on a context: these tests use one of their own.
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = rtnn_amd.Context(0)
    yield c
    c.close()


CLASS_OPTS = [{"cell_factor": 0.5, "queries_per_wave": 1},
              {"cell_factor": 1.0, "queries_per_wave": 16, "group_buffer": 1, "lds_tile": 256},
              {"cell_factor": 2.0, "queries_per_wave": 4, "group_buffer": 2}]


def test_batched_after_a_hierarchical_grid(fresh_ctx):
    # A batched search on a context whose last build was a hierarchical Morton
    # grid (partitioned KNN) must use the clouds' raster grids.
    import torch
    ctx = fresh_ctx
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 50_000)
    ctx.set_options(grid_levels=5, level_alpha=0.5)
    ctx.set_points(pts)
    ctx.search_partitioned(None, radius=20.0, k=8)
    ctx.set_options()
    xyz, offs = rtnn_amd.gen_lattice_clouds(9, 12, 500, 4000)
    t = torch.from_numpy(xyz).cuda()
    bi, bd, _ = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)
    pi, pd, _ = ctx.search_percloud(t, offs, mode="radius", radius=0.2, k=64)
    assert torch.equal(bi, pi) and torch.equal(bd, pd)
    # and the hierarchical search after it rebuilds its own grid
    wi, wd = oracle.grid_search(pts, pts[:500], oracle.MODE_KNN, 20.0, 8, threads=8)
    ctx.set_options(grid_levels=5, level_alpha=0.5)
    ctx.set_points(pts)
    gi, gd, _ = ctx.search_partitioned(pts[:500], radius=20.0, k=8)
    assert np.array_equal(np.asarray(gi), wi) and np.array_equal(np.asarray(gd), wd)


@pytest.mark.parametrize("mode,r,k", [("radius", 0.2, 64), ("knn", 0.25, 8)])
def test_batched_size_classes(fresh_ctx, mode, r, k):
    # Three classes with different cell edges, lanes and LDS use in one grid
    # build: the rows equal the per-cloud pipeline's and the oracle's.
    import torch
    ctx = fresh_ctx
    xyz, offs = rtnn_amd.gen_lattice_clouds(21, 40, 300, 9000)
    offs = np.insert(offs, 7, offs[7])  # an empty cloud
    t = torch.from_numpy(xyz).cuda()
    want_i, want_d, _ = ctx.search_percloud(t, offs, mode=mode, radius=r, k=k)
    ctx.set_size_classes([1500, 5000], CLASS_OPTS)
    up, opts = ctx.size_classes()
    assert up == [1500, 5000] and [o["queries_per_wave"] for o in opts] == [1, 16, 4]
    sizes = np.diff(offs)
    assert (sizes <= 1500).any() and ((sizes > 1500) & (sizes <= 5000)).any()
    assert (sizes > 5000).any()
    got_i, got_d, st = ctx.search_batched(t, offs, mode=mode, radius=r, k=k)
    assert st.search_launches == 3
    assert torch.equal(got_i, want_i) and torch.equal(got_d, want_d)
    gi = got_i.cpu().numpy()
    for c in (int(np.argmin(sizes + (sizes == 0) * 10**9)), int(np.argmax(sizes))):
        a, b = int(offs[c]), int(offs[c + 1])
        wi, wd = oracle.grid_search(xyz[a:b], xyz[a:b], MODES[mode], r, k, threads=8)
        assert np.array_equal(gi[a:b], wi)
    # the captured pipeline uses the same classes
    g = ctx.graph_batched(t, offs, mode=mode, radius=r, k=k)
    try:
        for rep in range(3):
            g.idx.fill_(-7)
            st = g.launch()
            assert st.search_launches == 3
            assert torch.equal(g.idx, want_i), rep
            assert torch.equal(g.d2, want_d), rep
            ctx.search_batched(t, offs, mode=mode, radius=r, k=k)  # eager work between
            torch.cuda.synchronize()
    finally:
        g.close()
    ctx.set_size_classes([], [])
    assert ctx.size_classes() == ([], [])
    _, _, st = ctx.search_batched(t, offs, mode=mode, radius=r, k=k)
    assert st.search_launches == 1


def test_size_class_errors(fresh_ctx):
    ctx = fresh_ctx
    with pytest.raises(rtnn_amd.RtnnError, match="increase"):
        ctx.set_size_classes([5000, 1500], CLASS_OPTS)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.set_size_classes([1500, 5000], [CLASS_OPTS[0], {"lds_tile": 300}, CLASS_OPTS[2]])
    with pytest.raises(ValueError):
        ctx.set_size_classes([1500], CLASS_OPTS)
    assert ctx.size_classes() == ([], [])
    # a class whose lanes cannot hold k fails the search, not silently
    import torch
    xyz, offs = rtnn_amd.gen_lattice_clouds(22, 6, 300, 3000)
    t = torch.from_numpy(xyz).cuda()
    ctx.set_size_classes([1000], [{"queries_per_wave": 32}, {}])
    with pytest.raises(rtnn_amd.RtnnError, match="k=64"):
        ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)


def test_autotune_batched_classes(fresh_ctx):
    import torch
    ctx = fresh_ctx
    xyz, offs = rtnn_amd.gen_lattice_clouds(23, 30, 300, 12000)
    t = torch.from_numpy(xyz).cuda()
    want_i, want_d, _ = ctx.search_percloud(t, offs, mode="radius", radius=0.2, k=64)
    opts, ms = ctx.autotune_batched_classes(t, offs, [2000, 6000], mode="radius", radius=0.2,
                                            k=64)
    assert len(opts) == 3 and all(m > 0 for m in ms)
    assert all(o["queries_per_wave"] in (1, 4, 8, 16, 32, 64) for o in opts)
    assert ctx.size_classes() == ([2000, 6000], opts)
    got_i, got_d, st = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)
    assert st.search_launches == 3
    assert torch.equal(got_i, want_i) and torch.equal(got_d, want_d)
    # the single-configuration tuner drops the classes
    ctx.autotune_batched(t, offs, mode="radius", radius=0.2, k=64)
    assert ctx.size_classes() == ([], [])


def test_search_async_on_a_torch_stream(ctx):
    # Enqueued on a caller's stream; the context's next (synchronous) call is
    # ordered after it without any host synchronisation in between.
    import torch
    pts = rtnn_amd.gen_uniform(17, 60000, 0.0, 30.0)
    qs = rtnn_amd.gen_uniform(18, 20000, -2.0, 32.0)
    tp, tq = torch.from_numpy(pts).cuda(), torch.from_numpy(qs).cuda()
    ctx.set_points(tp)
    want_i, want_d, _ = ctx.search(tq, mode="knn", radius=1.5, k=12)
    s = torch.cuda.Stream()
    idx = torch.full((len(qs), 12), -7, dtype=torch.int32, device="cuda")
    d2 = torch.zeros((len(qs), 12), dtype=torch.float32, device="cuda")
    ctx.search_async(tq, idx, d2, mode="knn", radius=1.5, stream=s)
    # the context's next call runs after the async search on the other stream
    ri, rd, _ = ctx.search(None, mode="radius", radius=1.0, k=8)
    s.synchronize()
    assert torch.equal(idx, want_i) and torch.equal(d2, want_d)
    wi, wd = oracle.grid_search(pts, pts, oracle.MODE_RADIUS, 1.0, 8, threads=8)
    assert np.array_equal(ri.cpu().numpy(), wi) and np.array_equal(rd.cpu().numpy(), wd)


def test_search_stream_chunks_equal_one_search(ctx, tmp_path):
    # Host queries searched in chunks (the last one short) with the copies of
    # one chunk overlapping the next chunk's search: rows equal one search,
    # into host arrays and into a result file.
    rng = np.random.default_rng(31)
    pts = rng.uniform(0, 30, (200000, 3)).astype(np.float32)
    qs = rng.uniform(-1, 31, (150001, 3)).astype(np.float32)
    ctx.set_points(pts)
    for mode, r, k in (("knn", 1.2, 16), ("radius", 1.0, 24)):
        wi, wd, _ = ctx.search(qs, mode=mode, radius=r, k=k)
        si, sd, st = ctx.search_stream(qs, mode=mode, radius=r, k=k, chunk=40000)
        assert np.array_equal(si, wi) and np.array_equal(sd, wd)
        si, sd, _ = ctx.search_stream(qs, mode=mode, radius=r, k=k, chunk=40000, with_d2=False)
        assert np.array_equal(si, wi) and sd is None
        path = tmp_path / f"rows_{mode}.bin"
        _, _, _ = ctx.search_stream(qs, mode=mode, radius=r, k=k, chunk=65536, path=str(path))
        fi, fd = rtnn_amd.read_results(str(path))
        assert np.array_equal(fi, wi) and np.array_equal(fd, wd)
    # proposal methods, one chunk and several
    wi, wd, _ = ctx.search(qs, mode="knn", radius=1.2, k=8)
    from oracle import mlp_ref as M
    w = np.random.default_rng(5).normal(0, 0.25, M.NP).astype(np.float32)
    w[M.OB3] = 0.3
    ctx.set_mlp_weights(np.concatenate([w, [0.25, 1.0]]).astype(np.float32))
    for method in ("partitioned", "learned"):
        for chunk in (0, 33333):
            si, sd, _ = ctx.search_stream(qs, method=method, radius=1.2, k=8, chunk=chunk)
            assert np.array_equal(si, wi) and np.array_equal(sd, wd)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.search_stream(qs, mode="radius", method="learned", radius=1.2, k=8)


def test_batched_stream_chunks_equal_one_batched_search(fresh_ctx, tmp_path):
    # Host clouds searched in chunks of whole clouds (a cloud larger than the
    # chunk on its own, empty clouds, a last short chunk): rows equal one
    # batched search, into host arrays and into a result file; size classes
    # apply per cloud inside every chunk.
    import torch
    ctx = fresh_ctx
    xyz, offs = rtnn_amd.gen_lattice_clouds(24, 60, 200, 12000, h=0.1, jit=0.02)
    offs = np.insert(offs, 5, offs[5])  # an empty cloud
    offs = np.insert(offs, len(offs), offs[-1])  # and one at the end
    t = torch.from_numpy(xyz).cuda()
    for mode, r, k in (("radius", 0.2, 64), ("knn", 0.25, 8)):
        wi, wd, _ = ctx.search_batched(t, offs, mode=mode, radius=r, k=k)
        wi, wd = wi.cpu().numpy(), wd.cpu().numpy()
        for chunk in (0, 10000, 50000):
            si, sd, st = ctx.search_batched_stream(xyz, offs, mode=mode, radius=r, k=k,
                                                   chunk=chunk)
            assert np.array_equal(si, wi) and np.array_equal(sd, wd), (mode, chunk)
        si, sd, _ = ctx.search_batched_stream(xyz, offs, mode=mode, radius=r, k=k, chunk=7000,
                                              with_d2=False)
        assert np.array_equal(si, wi) and sd is None
        path = tmp_path / f"batched_{mode}.bin"
        ctx.search_batched_stream(xyz, offs, mode=mode, radius=r, k=k, chunk=20000,
                                  path=str(path))
        fi, fd = rtnn_amd.read_results(str(path))
        assert np.array_equal(fi, wi) and np.array_equal(fd, wd)
    ctx.set_size_classes([1500, 5000], CLASS_OPTS)
    try:
        wi, wd, _ = ctx.search_batched(t, offs, mode="radius", radius=0.2, k=64)
        si, sd, st = ctx.search_batched_stream(xyz, offs, mode="radius", radius=0.2, k=64,
                                               chunk=30000)
        assert np.array_equal(si, wi.cpu().numpy()) and np.array_equal(sd, wd.cpu().numpy())
    finally:
        ctx.set_size_classes([], [])
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.search_batched_stream(xyz, offs[::-1].copy(), mode="radius", radius=0.2, k=64)


def test_mem_stats_track_the_library_allocations():
    # Device bytes rise with a context's buffers and fall back when it closes;
    # the peak covers the largest buffer set (memstatlib's running total).
    import gc
    base = rtnn_amd.mem_stats()
    rng = np.random.default_rng(41)
    pts = rng.uniform(0, 20, (100000, 3)).astype(np.float32)
    with rtnn_amd.Context(0) as c:
        c.set_points(pts)
        c.search(pts[:20000], mode="knn", radius=1.0, k=16)
        during = rtnn_amd.mem_stats()
        assert during["device_bytes"] >= base["device_bytes"] + 12 * len(pts)
        assert during["device_peak"] >= during["device_bytes"]
        c.search_stream(pts[:5000], mode="knn", radius=1.0, k=16, chunk=2000)
        assert rtnn_amd.mem_stats()["pinned_bytes"] > base["pinned_bytes"]
    gc.collect()
    after = rtnn_amd.mem_stats()
    assert after["device_bytes"] == base["device_bytes"]
    assert after["pinned_bytes"] == base["pinned_bytes"]
    assert after["device_peak"] >= during["device_bytes"]


# ---- pooled kernel (group_buffer 3): the candidates of 16/32/64 queries per
# wave tested as one pool; rows whose hit buffers overflow (more than 2k hits
# within the radius) are searched again by the group kernel.

@pytest.mark.parametrize("qpw", [16, 32, 64])
@pytest.mark.parametrize("mode,r,k", [("radius", 2.0, 50), ("radius", 1.2, 8), ("knn", 2.5, 16),
                                      ("knn", 1.5, 32), ("knn", 3.0, 4), ("radius", 2.0, 64)])
def test_pooled_kernel_matches_oracle(ctx, qpw, mode, r, k):
    rng = np.random.default_rng(100 + qpw + k)
    # uniform background + dense clusters (rows far beyond 2k hits: overflow
    # path) + far and duplicate queries
    pts = np.concatenate([rng.uniform(0, 40, (60000, 3)), rng.normal(20, 0.4, (8000, 3)),
                          np.repeat(rng.uniform(0, 40, (50, 3)), 20, axis=0)]).astype(np.float32)
    qs = np.concatenate([rng.uniform(-3, 43, (12000, 3)), pts[60000:62000],
                         [[500, 500, 500], [-100, 0, 0]]]).astype(np.float32)
    ctx.set_points(pts)
    for cf in (1.0, 2.0):
        ctx.set_options(cell_factor=cf, queries_per_wave=qpw, group_buffer=3)
        st = check(ctx, pts, qs, mode, r, k)
        assert st.search_launches == 2
        st = check(ctx, pts, None, mode, r, k, qsel=np.arange(0, len(pts), 7))


def test_pooled_kernel_many_rows_and_windows(ctx):
    # a radius spanning many cells (rows beyond one 256-row pass, pools beyond
    # one 1024-slot window) and queries with no candidate at all
    rng = np.random.default_rng(12)
    pts = rng.uniform(0, 30, (40000, 3)).astype(np.float32)
    qs = np.concatenate([rng.uniform(0, 30, (3000, 3)),
                         rng.uniform(100, 130, (300, 3))]).astype(np.float32)
    ctx.set_points(pts)
    for qpw in (16, 32):
        ctx.set_options(cell_factor=4.0, queries_per_wave=qpw, group_buffer=3)
        check(ctx, pts, qs, "knn", 6.0, 16)
        check(ctx, pts, qs, "radius", 3.0, 64)


def test_pooled_kernel_in_proposal_searches(ctx):
    pts = rtnn_amd.gen_spheres(3, 30, 400000, nspheres=80, box=800.0)
    qs = rtnn_amd.gen_spheres(3, 31, 60000, nspheres=80, box=800.0)
    ctx.set_points(pts)
    pi, pd, _ = ctx.search(qs, mode="knn", radius=10.0, k=16)
    ctx.mlp_train(10.0, 16, n_samples=1 << 16, steps=300, batch=4096, seed=3)
    for qpw in (16, 32, 64):
        ctx.set_options(cell_factor=2.0, queries_per_wave=qpw, group_buffer=3, grid_levels=1)
        for fn in (ctx.search_learned, ctx.search_partitioned):
            li, ld, st = fn(qs, radius=10.0, k=16)
            assert np.array_equal(pi, li) and np.array_equal(pd, ld), (qpw, fn.__name__)
