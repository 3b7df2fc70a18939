"""The density partition's launch radii on the GPU, equal bit for bit to the
rule they restate.

Raster grid (grid_levels = 1): per query, the reference's calcSearchSize
(grid.cu:84-211) on a point-count grid of edge r/8 -- the cube of (2i+1)^3
cells around the query's cell grows until it holds >= k+1 points or its width
(2i+2)*cell exceeds 2r/sqrt(3) -- gives i; the launch radius is then
min(r, sqrt(3)*(i+1)*cell) (radiusFromMegacell with approx mode 0,
util.cpp:415-419; sort.cpp:366-371), or r when the width cap stopped the
growth.  i comes from the oracle's orc_search_size (itself pinned by closed
forms in test_oracle.py) on a count grid this test builds from the points.

Hierarchical grid (grid_levels >= 2): the per-level form of the same rule,
recounted here by brute force from the points on tiny grids: at the finest
level whose cell holding the query has >= k+1 points the radius is sqrt(3)
times that level's edge, where its 3x3x3 block does, twice that.

The radii are read through rtnn_propose_radius(..., RTNN_METHOD_PARTITIONED),
which returns what rtnn_search_partitioned searches first.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

F = np.float32
SQRT3 = F(1.7320508)
GROW = F(1.0) + F(1e-6)  # the kernels' (1 + 1e-6) margin on sqrt(3) * width


@pytest.fixture(autouse=True)
def _reset(ctx):
    ctx.set_options()
    yield
    ctx.set_options()


def _cells(x, mn, inv):
    return np.floor((x.astype(F) - mn) * inv).astype(np.float64)


def raster_expected(pts, qs, r, k):
    """Launch radius per query on the edge-r/8 count grid (origin: the points'
    box minimum; dims floor(extent / cell) + 1, FP32 as the device plans)."""
    r = F(r)
    cell = r / F(8.0)
    inv = F(1.0) / cell
    mn = pts.min(axis=0).astype(F)
    mx = pts.max(axis=0).astype(F)
    dims = (np.floor((mx - mn) * inv) + 1).astype(np.int64)
    pc = np.clip(_cells(pts, mn, inv), 0, dims - 1).astype(np.int64)
    counts = np.zeros((dims[2], dims[1], dims[0]), np.uint32)
    np.add.at(counts, (pc[:, 2], pc[:, 1], pc[:, 0]), 1)
    max_width = F(2.0) * r / SQRT3
    qc = np.clip(_cells(qs, mn, inv), -1, dims).astype(np.int64)
    out = np.empty(len(qs), F)
    for i, (x, y, z) in enumerate(qc):
        it = oracle.search_size(counts, int(x), int(y), int(z), float(cell), float(max_width), k)
        width = F(2 * it + 2) * cell
        if width > max_width:
            out[i] = r
        else:
            out[i] = min(r, SQRT3 * F(it + 1) * cell * GROW)
    return out


def hier_expected(pts, qs, r, k, levels):
    """Launch radius per query on a hierarchical Morton grid of `levels`
    levels (top edge r), counts recounted from the points at every level."""
    r = F(r)
    c0 = F(np.ldexp(r, -(levels - 1)))
    inv0 = F(1.0) / c0
    mn = pts.min(axis=0).astype(F)
    mx = pts.max(axis=0).astype(F)
    dmax = int((np.floor((mx - mn) * inv0) + 1).max())
    b = max(int(dmax - 1).bit_length(), levels - 1)
    assert b <= 9, "tiny grids only (every level's table built)"
    dim = 1 << b
    pc = np.clip(_cells(pts, mn, inv0), 0, dim - 1).astype(np.int64)
    ksq = SQRT3 * GROW
    out = np.empty(len(qs), F)
    level_counts = []
    for l in range(levels):
        d = dim >> l
        c = np.zeros((d, d, d), np.int64)
        np.add.at(c, (pc[:, 2] >> l, pc[:, 1] >> l, pc[:, 0] >> l), 1)
        level_counts.append(c)
    qf = _cells(qs, mn, inv0)
    for i in range(len(qs)):
        R = r
        if ((qf[i] >= 0) & (qf[i] < dim)).all():
            q0 = qf[i].astype(np.int64)
            for l in range(levels):
                cl = F(np.ldexp(c0, l))
                if ksq * cl >= r:
                    break
                x, y, z = q0 >> l
                c = level_counts[l]
                if c[z, y, x] >= k + 1:
                    R = ksq * cl
                    break
                if F(2.0) * ksq * cl >= r:
                    continue
                d = dim >> l
                blk = c[max(z - 1, 0):min(z + 2, d), max(y - 1, 0):min(y + 2, d),
                        max(x - 1, 0):min(x + 2, d)].sum()
                if blk >= k + 1:
                    R = F(2.0) * ksq * cl
                    break
        out[i] = min(R, r)
    return out


def _cloud(seed):
    # a dense cluster, a sparse background and an empty corner region, so the
    # shells stop at many different sizes
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal([12, 12, 12], 1.2, (12000, 3)),
                           rng.uniform(0, 24, (2500, 3)),
                           rng.normal([20, 4, 18], 0.4, (800, 3))]).astype(F)


def _queries(pts, seed):
    rng = np.random.default_rng(seed)
    mn, mx = pts.min(axis=0), pts.max(axis=0)
    interior = pts[rng.choice(len(pts), 1500, replace=False)]
    uniform = rng.uniform(mn, mx, (1500, 3))
    corners = np.array([[mn[0], mn[1], mn[2]], [mx[0], mx[1], mx[2]], [mn[0], mx[1], mn[2]],
                        [mx[0], mn[1], mx[2]]])
    far = np.array([[mn[0] - 50, mn[1], mn[2]], [mx[0] + 3, mx[1] + 3, mx[2] + 3],
                    [mn[0] - 1e4, 0, 0]])
    return np.concatenate([interior, uniform, corners, far]).astype(F)


@pytest.mark.parametrize("r,k,capped", [(4.0, 8, True), (4.0, 32, True), (2.5, 16, True),
                                         (12.0, 4, False)])
def test_partition_radius_raster_matches_calc_search_size(ctx, r, k, capped):
    pts = _cloud(11)
    qs = _queries(pts, 12)
    ctx.set_points(pts)
    ctx.set_options(grid_levels=1)
    got = ctx.propose_radius(qs, "partitioned", r, k)
    want = raster_expected(pts, qs, r, k)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, (len(bad), bad[:5], got[bad[:5]], want[bad[:5]])
    # the cases are exercised: megacells of several sizes and, where the cloud
    # is sparse enough for it, the width cap (radius r)
    assert len(np.unique(want)) >= 3 and (want < F(r)).any()
    assert (want == F(r)).any() == capped


@pytest.mark.parametrize("levels,r,k", [(3, 4.0, 8), (5, 8.0, 16), (4, 6.0, 1)])
def test_partition_radius_hierarchical_per_level_rule(ctx, levels, r, k):
    pts = _cloud(21)
    qs = _queries(pts, 22)
    ctx.set_points(pts)
    ctx.set_options(grid_levels=levels)
    got = ctx.propose_radius(qs, "partitioned", r, k)
    want = hier_expected(pts, qs, r, k, levels)
    bad = np.nonzero(got.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, (len(bad), bad[:5], got[bad[:5]], want[bad[:5]])
    assert len(np.unique(want)) >= 2


def test_partitioned_search_equals_plain_on_the_pinned_cloud(ctx):
    pts = _cloud(31)
    qs = _queries(pts, 32)
    ctx.set_points(pts)
    pi, pd, _ = ctx.search(qs, mode="knn", radius=4.0, k=16)
    for levels in (1, 4):
        ctx.set_options(grid_levels=levels)
        li, ld, _ = ctx.search_partitioned(qs, radius=4.0, k=16)
        assert np.array_equal(pi, li) and np.array_equal(pd, ld)
