"""rtnn_search_noalloc (the search that allocates nothing, for callers that
cannot allocate) against rtnn_set_points + rtnn_search on the same inputs:
rows bit-identical, on two calls in a row (the second smaller), in both modes
and with separate queries; sizes beyond its scratch and non-finite points are
refused with an error; the library's tracked allocations do not change."""
import numpy as np
import pytest

import rtnn_amd

pytestmark = pytest.mark.gpu


@pytest.fixture
def nctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = rtnn_amd.Context(0)
    yield c
    c.close()


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _ref(ctx, pts, qs, mode, r, k):
    ctx.set_points(pts)
    idx, d2, _ = ctx.search(qs, mode=mode, radius=r, k=k, out_device=True)
    return idx, d2


@pytest.mark.parametrize("mode,r,k,opts", [("radius", 2.0, 50, {}),
                                           ("knn", 2.5, 16, {}),
                                           ("radius", 2.0, 64, {"queries_per_wave": 16}),
                                           ("knn", 3.0, 8, {"queries_per_wave": 4})])
def test_noalloc_equals_search_two_calls(nctx, mode, r, k, opts):
    import torch
    rng = np.random.default_rng(7)
    # first call: 400k points, 150k separate queries (some outside the box)
    p1 = _t(rng.uniform(0, 60, (400_000, 3)).astype(np.float32))
    q1 = _t(rng.uniform(-5, 65, (150_000, 3)).astype(np.float32))
    # second call, smaller: 60k clustered points searched against themselves
    p2 = _t(np.concatenate([rng.normal(10, 2, (50_000, 3)),
                            rng.uniform(0, 20, (10_000, 3))]).astype(np.float32))
    nctx.set_options(**opts)
    calls = [(p1, q1), (p2, None)]
    got = []
    before = rtnn_amd.mem_stats()["device_bytes"]
    alloc0 = torch.cuda.memory_allocated()
    outs = [(torch.empty((len(q if q is not None else p), k), dtype=torch.int32, device="cuda"),
             torch.empty((len(q if q is not None else p), k), dtype=torch.float32, device="cuda"))
            for p, q in calls]
    alloc1 = torch.cuda.memory_allocated()
    for (p, q), (oi, od) in zip(calls, outs):
        nctx.search_noalloc(p, q, mode=mode, radius=r, k=k, idx=oi, d2=od)
        got.append((oi, od))
    torch.cuda.synchronize()
    assert rtnn_amd.mem_stats()["device_bytes"] == before  # nothing allocated by the library
    assert torch.cuda.memory_allocated() == alloc1 >= alloc0
    for (p, q), (gi, gd) in zip(calls, got):
        wi, wd = _ref(nctx, p, q, mode, r, k)
        assert torch.equal(gi, wi) and torch.equal(gd, wd)
        assert (wi[:, 0] >= 0).any()


def test_noalloc_without_d2_and_repeated(nctx):
    import torch
    rng = np.random.default_rng(9)
    p = _t(rng.uniform(0, 30, (100_000, 3)).astype(np.float32))
    wi, _ = _ref(nctx, p, None, "knn", 1.5, 12)
    for _ in range(3):
        gi, gd = nctx.search_noalloc(p, None, mode="knn", radius=1.5, k=12)
        assert gd is None and torch.equal(gi, wi)


def test_noalloc_refuses_what_it_cannot_hold(nctx):
    import torch
    big = torch.zeros((4_000_000, 3), dtype=torch.float32, device="cuda")
    idx = torch.empty((1, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(rtnn_amd.RtnnError) as ei:
        nctx.search_noalloc(big, big[:1], mode="knn", radius=1.0, k=4, idx=idx)
    assert ei.value.code == rtnn_amd.ERR_UNSUPPORTED and "scratch" in str(ei.value)
    pts = np.random.default_rng(1).uniform(0, 5, (1000, 3)).astype(np.float32)
    pts[17, 1] = np.nan
    with pytest.raises(rtnn_amd.RtnnError) as ei:
        nctx.search_noalloc(_t(pts), None, mode="radius", radius=1.0, k=4)
    assert ei.value.code == rtnn_amd.ERR_INVALID
    with pytest.raises(rtnn_amd.RtnnError):
        nctx.search_noalloc(_t(pts[:0]), None, mode="radius", radius=1.0, k=4)
    # the next call after the refusals is exact again
    ok = _t(np.random.default_rng(2).uniform(0, 5, (5000, 3)).astype(np.float32))
    wi, wd = _ref(nctx, ok, None, "radius", 0.7, 20)
    gi, gd = nctx.search_noalloc(ok, None, mode="radius", radius=0.7, k=20,
                                 d2=torch.empty((5000, 20), dtype=torch.float32, device="cuda"))
    assert torch.equal(gi, wi) and torch.equal(gd, wd)
