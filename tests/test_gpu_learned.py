"""GPU tests of the learned radius proposal and the density-partitioned KNN.

Numerics (tolerances written per test):
  * features: GPU log1p via the hardware log vs numpy log1p, atol 1e-5;
  * MFMA forward: BF16 weights/activations with FP32 accumulation vs the
    float64 reference with the same BF16 rounding; accumulation order can
    flip an activation's BF16 rounding, so atol 2e-2 (a wrong MFMA operand
    layout gives O(1) errors);
  * backward: FP32 GPU gradient vs the float64 straight-through gradient,
    relative L2 error < 2e-2 and elementwise atol 2e-3 * max|g|.
Exactness: learned and partitioned KNN must equal the plain KNN bit for bit,
including when every query falls back (r' forced to its floor) and when none
does (r' forced to r).
"""
import numpy as np
import pytest

import rtnn_amd
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rand_blob(seed, scale=0.25):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, M.NP).astype(np.float32)
    p[M.OB3] = 0.3
    return np.concatenate([p, [0.25, 1.0]]).astype(np.float32)


def test_features_match_reference(ctx):
    rng = np.random.default_rng(1)
    pts = rng.uniform(0, 60, (50000, 3)).astype(np.float32)
    qs = np.concatenate([pts[:2000], rng.uniform(-20, 80, (2000, 3)).astype(np.float32)])
    ctx.set_points(pts)
    got = ctx.mlp_features(_t(qs), 5.0).cpu().numpy()
    want = M.features(pts, qs, 5.0, 0.25)
    assert np.abs(got - want).max() < 1e-5


def test_mfma_forward_matches_reference(ctx):
    blob = rand_blob(2)
    ctx.set_mlp_weights(blob)
    rng = np.random.default_rng(3)
    feats = rng.uniform(0, 2, (5000, 32)).astype(np.float32)
    feats[:, 30:] = 0
    got = ctx.mlp_forward(_t(feats)).cpu().numpy()
    want = M.forward(blob[:M.NP], feats)
    err = np.abs(got - want)
    assert err.max() < 2e-2, err.max()
    assert np.median(err) < 2e-3
    # weights round trip
    w = ctx.mlp_weights()
    assert np.array_equal(w, blob)


def test_backward_matches_reference(ctx):
    blob = rand_blob(4)
    ctx.set_mlp_weights(blob)
    rng = np.random.default_rng(5)
    n = 1000
    feats = rng.uniform(0, 2, (n, 32)).astype(np.float32)
    feats[:, 30:] = 0
    o = M.forward(blob[:M.NP], feats)
    t = (o + np.where(rng.random(n) < 0.5, -0.3, 0.3)).astype(np.float32)
    g, loss = ctx.mlp_backward(_t(feats), _t(t), tau=0.9)
    g = g.cpu().numpy().astype(np.float64)
    want = M.grad(blob[:M.NP], feats, t, 0.9)
    assert np.linalg.norm(g - want) < 2e-2 * np.linalg.norm(want)
    assert np.abs(g - want).max() < 2e-3 * np.abs(want).max()
    assert abs(loss - M.loss(blob[:M.NP], feats, t, 0.9)) < 1e-4


def _plain_vs(ctx, qs, r, k, fn):
    import torch
    pi, pd, _ = ctx.search(qs, mode="knn", radius=r, k=k)
    li, ld, st = fn(qs, radius=r, k=k)
    assert np.array_equal(pi, li)
    assert np.array_equal(pd, ld)
    return st


def test_learned_training_and_exactness(ctx):
    pts = rtnn_amd.gen_spheres(3, 30, 300000, nspheres=60, box=600.0)
    qs = rtnn_amd.gen_spheres(3, 31, 100000, nspheres=60, box=600.0)
    ctx.set_points(pts)
    with rtnn_amd.Context(0) as fresh:
        fresh.set_points(pts)
        with pytest.raises(rtnn_amd.RtnnError, match="no weights"):
            fresh.search_learned(qs, radius=10.0, k=16)
    loss0 = ctx.mlp_train(10.0, 16, n_samples=1 << 15, steps=1, batch=4096, seed=7)
    loss = ctx.mlp_train(10.0, 16, n_samples=1 << 15, steps=400, batch=4096, seed=7)
    assert loss < loss0, (loss0, loss)
    st = _plain_vs(ctx, qs, 10.0, 16, ctx.search_learned)
    frac = st.fallback_queries / len(qs)
    assert frac < 0.5, frac
    # training is deterministic
    w1 = ctx.mlp_weights()
    ctx.mlp_train(10.0, 16, n_samples=1 << 15, steps=400, batch=4096, seed=7)
    assert np.array_equal(w1, ctx.mlp_weights())


def test_learned_forced_fallback_and_no_fallback(ctx):
    rng = np.random.default_rng(9)
    pts = rng.uniform(0, 40, (60000, 3)).astype(np.float32)
    qs = rng.uniform(-5, 45, (20000, 3)).astype(np.float32)
    ctx.set_points(pts)
    blob = np.zeros(M.NP + 2, np.float32)
    blob[M.NP] = 0.25
    blob[M.NP + 1] = 1.0
    ctx.set_mlp_weights(blob)  # o = 0 -> r' at its floor -> (almost) all fall back
    st = _plain_vs(ctx, qs, 3.0, 8, ctx.search_learned)
    assert st.fallback_queries > 0.9 * len(qs)
    blob[M.OB3] = 10.0  # o = 10 -> r' = r: every row is final after stage 1
    ctx.set_mlp_weights(blob)
    st = _plain_vs(ctx, qs, 3.0, 8, ctx.search_learned)
    pi, _, _ = ctx.search(qs, mode="knn", radius=3.0, k=8)
    assert int((pi[:, -1] < 0).sum()) > 0  # short rows exist, searched once at r
    assert st.fallback_queries == 0


@pytest.mark.parametrize("k", [65, 100, 300])
def test_proposal_methods_beyond_k64(ctx, k):
    # k above the register top-k: the default partitioned configuration (a
    # 5-level hierarchical grid) and the learned one search with 16- or
    # 64-lane groups, stage 1 and fallback; rows equal the plain KNN.
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 60000)
    qs = rtnn_amd.gen_gaussian_mixture(2, 21, 5000)
    ctx.set_points(pts)
    w = np.random.default_rng(6).normal(0, 0.2, M.NP).astype(np.float32)
    ctx.set_mlp_weights(np.concatenate([w, [0.25, 1.0]]).astype(np.float32))
    try:
        for opts in ({}, {"grid_levels": 1, "cell_factor": 4.0}):
            ctx.set_options(**opts)
            _plain_vs(ctx, qs, 20.0, k, ctx.search_partitioned)
            _plain_vs(ctx, qs, 20.0, k, ctx.search_learned)
    finally:
        ctx.set_options()


@pytest.mark.parametrize("k", [1, 8, 16, 32])
def test_partitioned_equals_plain(ctx, k):
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 100000)
    qs = rtnn_amd.gen_gaussian_mixture(2, 21, 50000)
    ctx.set_points(pts)
    st = _plain_vs(ctx, qs, 20.0, k, ctx.search_partitioned)
    # hierarchical grid (the default): stage 1, then the octree walk at r
    assert st.search_launches == 2
    ctx.set_options(grid_levels=1)
    try:
        st = _plain_vs(ctx, qs, 20.0, k, ctx.search_partitioned)
        assert st.search_launches == 6  # stage 1, four radius-growth stages, the last at r
    finally:
        ctx.set_options()
    rng = np.random.default_rng(k)
    u = rng.uniform(0, 30, (40000, 3)).astype(np.float32)
    ctx.set_points(u)
    _plain_vs(ctx, u[:5000], 2.0, k, ctx.search_partitioned)


@pytest.mark.parametrize("qpw", [0, 16, 4])
@pytest.mark.parametrize("levels,alpha", [(1, 0.0), (2, 1.0), (3, 0.5), (5, 1.0), (5, 2.0)])
def test_hierarchical_grid_exact(ctx, levels, alpha, qpw):
    # Partitioned and learned KNN on the hierarchical Morton grid (one sorted
    # point array, per-level tables, per-query level) equal the plain KNN, for
    # the thread-per-query kernel and the 4- and 16-lane group kernels.
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 100000)
    qs = np.concatenate([rtnn_amd.gen_gaussian_mixture(2, 21, 30000),
                         np.array([[-50, 0, 0], [1e4, 1e4, 1e4]], np.float32)])
    ctx.set_points(pts)
    ctx.set_options(grid_levels=levels, level_alpha=alpha, queries_per_wave=qpw)
    try:
        for k in (1, 8, 32):
            _plain_vs(ctx, qs, 20.0, k, ctx.search_partitioned)
        ctx.set_mlp_weights(rand_blob(11))
        _plain_vs(ctx, qs, 20.0, 8, ctx.search_learned)
    finally:
        ctx.set_options()


@pytest.mark.parametrize("leaf", [1, 64, 1 << 20])
def test_octree_leaf_sizes_exact(ctx, leaf):
    # The octree walk with tiny leaves (descend to the finest level), default
    # and huge ones (a top-level cell scanned chunk by chunk) equals the plain
    # KNN, also on Levy-flight clusters.
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 100000)
    qs = rtnn_amd.gen_gaussian_mixture(2, 21, 20000)
    ctx.set_points(pts)
    ctx.set_options(grid_levels=5, level_alpha=1.0, queries_per_wave=0, tree_leaf=leaf)
    try:
        for k in (1, 8, 32):
            _plain_vs(ctx, qs, 20.0, k, ctx.search_partitioned)
        lv = rtnn_amd.gen_levy(5, 300000, alpha=1.5, min_step=0.01, box=100.0)
        lq = rtnn_amd.gen_uniform(51, 20000, 0.0, 100.0)
        ctx.set_points(lv)
        ctx.set_options(grid_levels=7, queries_per_wave=0, tree_leaf=leaf)
        _plain_vs(ctx, lq, 10.0, 32, ctx.search_partitioned)
    finally:
        ctx.set_options()


@pytest.mark.parametrize("lanes", [2, 4, 8])
def test_octree_walk_with_several_lanes_exact(ctx, lanes):
    # G lanes per query split every leaf chunk's points, share the smallest
    # lane limit and merge their lists: rows equal the plain KNN for k up to
    # 64, small and huge leaves, the partitioned and learned methods (the
    # fallback's walks at r use the same lanes), clustered and far queries.
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 100000)
    qs = np.concatenate([rtnn_amd.gen_gaussian_mixture(2, 21, 20000),
                         [[1e4, 0, 0], [-50, -50, -50]]]).astype(np.float32)
    ctx.set_points(pts)
    w = np.random.default_rng(3).normal(0, 0.2, M.NP).astype(np.float32)
    ctx.set_mlp_weights(np.concatenate([w, [0.25, 1.0]]).astype(np.float32))
    try:
        for leaf in (0, 65536):
            ctx.set_options(grid_levels=5, level_alpha=1.0, queries_per_wave=0, tree_leaf=leaf,
                            tree_lanes=lanes)
            for k in (1, 8, 32, 64):
                _plain_vs(ctx, qs, 20.0, k, ctx.search_partitioned)
            _plain_vs(ctx, qs, 20.0, 16, ctx.search_learned)
        lv = rtnn_amd.gen_levy(5, 300000, alpha=1.5, min_step=0.01, box=100.0)
        lq = rtnn_amd.gen_uniform(51, 20000, 0.0, 100.0)
        ctx.set_points(lv)
        ctx.set_options(grid_levels=7, queries_per_wave=0, tree_leaf=65536, tree_lanes=lanes)
        _plain_vs(ctx, lq, 10.0, 32, ctx.search_partitioned)
        with pytest.raises(rtnn_amd.RtnnError):
            ctx.set_options(tree_lanes=3)
    finally:
        ctx.set_options()


@pytest.mark.parametrize("lanes", [1, 2])
def test_octree_growth_stage_fallback_exact(ctx, lanes):
    # tree_fallback = 1: the octree walk's short rows go through the radius-
    # growth stages instead of straight to r; rows still equal the plain KNN,
    # with the forced-fallback network (almost every row short) and a random
    # one, partitioned too.
    pts = rtnn_amd.gen_gaussian_mixture(2, 20, 100000)
    qs = np.concatenate([rtnn_amd.gen_gaussian_mixture(2, 21, 20000),
                         [[1e4, 0, 0], [-50, -50, -50]]]).astype(np.float32)
    ctx.set_points(pts)
    try:
        for blob_w in (np.zeros(M.NP, np.float32),
                       np.random.default_rng(4).normal(0, 0.2, M.NP).astype(np.float32)):
            ctx.set_mlp_weights(np.concatenate([blob_w, [0.25, 1.0]]).astype(np.float32))
            ctx.set_options(grid_levels=5, level_alpha=1.0, queries_per_wave=0,
                            tree_lanes=lanes, tree_fallback=1)
            for k in (8, 32):
                st = _plain_vs(ctx, qs, 20.0, k, ctx.search_learned)
                assert st.search_launches == 6
                if not blob_w.any():  # o = 0: r' at its floor, (almost) every row short
                    assert st.fallback_queries > 0.5 * len(qs)
            _plain_vs(ctx, qs, 20.0, 8, ctx.search_partitioned)
        ctx.set_options(grid_levels=5, level_alpha=1.0, queries_per_wave=0, tree_lanes=lanes)
        st = _plain_vs(ctx, qs, 20.0, 8, ctx.search_learned)
        assert st.search_launches == 2
        with pytest.raises(rtnn_amd.RtnnError, match="tree_fallback"):
            ctx.set_options(tree_fallback=2)
    finally:
        ctx.set_options()


def test_autotune_picks_a_valid_config(ctx):
    rng = np.random.default_rng(21)
    pts = rng.uniform(0, 40, (50000, 3)).astype(np.float32)
    ctx.set_points(pts)
    try:
        opt, ms = ctx.autotune(None, method="plain", mode="radius", radius=2.0, k=32)
        assert ms > 0 and opt["cell_factor"] > 0
        idx, d2, _ = ctx.search(None, mode="radius", radius=2.0, k=32)
        import oracle
        wi, wd = oracle.grid_search(pts, pts, oracle.MODE_RADIUS, 2.0, 32, threads=8)
        assert np.array_equal(idx, wi) and np.array_equal(d2, wd)
        opt, ms = ctx.autotune(pts[:5000], method="partitioned", mode="knn", radius=3.0, k=8)
        assert ms > 0
    finally:
        ctx.set_options()


def test_trained_network_gain(ctx):
    # What the trained network buys over the method without it.  Without the
    # MLP every query is searched at r (relative search volume 1).  With it, a
    # query is first searched at r' and re-run at r only if r' holds fewer
    # than k neighbours, so its expected work is about (r'/r)^3 plus 1 for a
    # miss.  Trained, that is well under half of the plain search and better
    # than an untrained network; the exact result is checked elsewhere.
    pts = rtnn_amd.gen_spheres(3, 30, 300000, nspheres=60, box=600.0)
    qs = rtnn_amd.gen_spheres(3, 31, 50000, nspheres=60, box=600.0)
    r, k = 10.0, 16
    ctx.set_points(pts)
    _, d2, _ = ctx.search(qs, mode="knn", radius=r, k=k)
    dk = np.sqrt(d2[:, k - 1].astype(np.float64))  # inf where fewer than k within r

    def work(rp):
        covered = dk < rp
        return float(np.mean((rp / r) ** 3) + np.mean(~covered)), float(np.mean(covered))

    ctx.set_mlp_weights(rand_blob(12))
    w0, c0 = work(ctx.propose_radius(qs, "learned", r, k))
    ctx.mlp_train(r, k, n_samples=1 << 16, steps=600, batch=4096, seed=3)
    rp = ctx.propose_radius(qs, "learned", r, k)
    assert (rp > 0).all() and (rp <= r).all()
    w1, c1 = work(rp)
    wp, cp = work(ctx.propose_radius(qs, "partitioned", r, k))
    print(f"relative work: trained {w1:.3f} (covers {c1:.2%}), untrained {w0:.3f} "
          f"(covers {c0:.2%}), density partition {wp:.3f} (covers {cp:.2%}), plain 1.0")
    assert w1 < 0.5 and c1 > 0.7
    assert w1 < w0


def test_train_on_query_sample_exact(ctx):
    # Uniform queries over clustered points: the network is trained on a
    # sample of the queries themselves; the learned search stays exact.
    pts = rtnn_amd.gen_levy(5, 200000, alpha=1.5, min_step=0.01, box=100.0)
    qs = rtnn_amd.gen_uniform(51, 40000, 0.0, 100.0)
    ctx.set_points(pts)
    loss = ctx.mlp_train(15.0, 8, steps=300, batch=2048, seed=5, queries=qs[:20000])
    assert np.isfinite(loss)
    _plain_vs(ctx, qs, 15.0, 8, ctx.search_learned)
    _plain_vs(ctx, qs, 15.0, 8, ctx.search_partitioned)


def test_levy_uniform_queries_k32_large_radius(ctx):
    # Config-5 shape at 1/30 of the points: Levy-flight points (alpha 1.5,
    # min step 0.01, periodic [0,1000)^3), uniform queries, K = 32, r = 100
    # (near-unbounded).  Learned (trained on a query sample) and partitioned
    # KNN equal the plain KNN on every row; sampled rows equal the oracle.
    import oracle
    pts = rtnn_amd.gen_levy(5, 1_000_000, alpha=1.5, min_step=0.01, box=1000.0)
    qs = rtnn_amd.gen_uniform(51, 60_000, 0.0, 1000.0)
    ctx.set_points(pts)
    pi, pd, _ = ctx.search(qs, mode="knn", radius=100.0, k=32)
    ctx.mlp_train(100.0, 32, steps=300, batch=2048, seed=2, queries=qs[:20000])
    for fn in (ctx.search_learned, ctx.search_partitioned):
        li, ld, _ = fn(qs, radius=100.0, k=32)
        assert np.array_equal(pi, li) and np.array_equal(pd, ld)
    # the octree walk of the hierarchical grid (one lane per query)
    ctx.set_options(grid_levels=7, queries_per_wave=0)
    try:
        for fn in (ctx.search_learned, ctx.search_partitioned):
            li, ld, _ = fn(qs, radius=100.0, k=32)
            assert np.array_equal(pi, li) and np.array_equal(pd, ld)
    finally:
        ctx.set_options()
    sel = np.random.default_rng(3).choice(len(qs), 300, replace=False).astype(np.uint64)
    wi, wd = oracle.bruteforce(pts, qs, oracle.MODE_KNN, 100.0, 32, qsel=sel, threads=8)
    s = sel.astype(np.int64)
    assert np.array_equal(pi[s], wi) and np.array_equal(pd[s], wd)


@pytest.mark.parametrize("ratio", [1.0, 0.7, 0.3])
def test_density_buckets_exact(ctx, ratio):
    # Density buckets (bucket_ratio): proposals r' >= ratio * r are searched at
    # r at once.  Rows equal the plain KNN for both proposal methods, on the
    # raster grid, the hierarchical grid's cube walks and the octree walk; the
    # radii rtnn_propose_radius reports are the proposals with that rule
    # applied.  learned_cap: the network's radius capped by the partition's,
    # exactly their minimum.
    pts = rtnn_amd.gen_levy(5, 300_000, alpha=1.5, min_step=0.01, box=300.0)
    qs = rtnn_amd.gen_uniform(52, 40_000, 0.0, 300.0)
    ctx.set_points(pts)
    r, k = 20.0, 16
    ctx.mlp_train(r, k, steps=200, batch=2048, seed=4, queries=qs[:10000])
    pi, pd, _ = ctx.search(qs, mode="knn", radius=r, k=k)
    try:
        for opts in ({"grid_levels": 1}, {"grid_levels": 5, "queries_per_wave": 16},
                     {"grid_levels": 5, "queries_per_wave": 0}):
            for method, fn in (("learned", ctx.search_learned),
                               ("partitioned", ctx.search_partitioned)):
                ctx.set_options(**opts)
                want = ctx.propose_radius(qs, method, r, k)
                ctx.set_options(bucket_ratio=ratio, **opts)
                got = ctx.propose_radius(qs, method, r, k)
                snap = want >= np.float32(ratio) * np.float32(r)
                assert np.array_equal(got[snap], np.full(int(snap.sum()), r, np.float32))
                assert np.array_equal(got[~snap], want[~snap])
                li, ld, st = fn(qs, radius=r, k=k)
                assert np.array_equal(pi, li) and np.array_equal(pd, ld), (opts, method)
                if ratio <= 0.3:
                    # most proposals land in the sparse bucket: few rows left
                    # short after the first launch
                    assert st.fallback_queries <= int((~snap).sum()), (opts, method)
            ctx.set_options(**opts)
            part = ctx.propose_radius(qs, "partitioned", r, k)
            net = ctx.propose_radius(qs, "learned", r, k)
            ctx.set_options(learned_cap=1, **opts)
            capped = ctx.propose_radius(qs, "learned", r, k)
            assert np.array_equal(capped, np.minimum(net, part)), opts
            assert (capped < net).any()
            li, ld, _ = ctx.search_learned(qs, radius=r, k=k)
            assert np.array_equal(pi, li) and np.array_equal(pd, ld), opts
    finally:
        ctx.set_options()


def test_bucket_ratio_validated(ctx):
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.set_options(bucket_ratio=1.5)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.set_options(bucket_ratio=-0.1)
    with pytest.raises(rtnn_amd.RtnnError):
        ctx.set_options(learned_cap=2)
    ctx.set_options()
