"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so a plausible mistake in the oracle fails it:
  * golden vectors for the reference's own data file (samplepc.txt), produced
    by an independent numpy FP32 brute force (tests/golden/make_golden.py);
  * closed forms: d2 of axis-aligned offsets, FMA-free rounding, strictness of
    d2 < r*r, self handling per mode;
  * an independent numpy brute force on random inputs (set and order);
  * the grid oracle against the brute force oracle (cell sizes, far queries);
  * Morton known answers and the Compact1By2 round trip
    (helper_mortonCode.h:19-56);
  * the partition mask (grid.cu:84-211) against closed forms on uniform grids.
"""
import numpy as np
import pytest

import oracle
from conftest import golden_d2


def np_brute(pts, qs, mode, r, k):
    f = np.float32
    r2 = f(r) * f(r)
    out_i = np.full((len(qs), k), -1, np.int32)
    out_d = np.full((len(qs), k), np.inf, np.float32)
    for t, q in enumerate(qs):
        d = q[None, :] - pts
        sq = d * d
        d2 = (sq[:, 0] + sq[:, 1]) + sq[:, 2]
        keep = (d2 < r2) if mode == oracle.MODE_RADIUS else ((d2 > 0) & (d2 < r2))
        idx = np.nonzero(keep)[0]
        order = np.lexsort((idx, d2[idx]))[:k]
        sel = idx[order]
        out_i[t, :len(sel)] = sel
        out_d[t, :len(sel)] = d2[sel]
    return out_i, out_d


def test_golden_samplepc(golden, samplepc):
    assert np.array_equal(np.array(golden["points"], np.float32), samplepc)
    for case in golden["cases"]:
        mode = oracle.MODE_RADIUS if case["mode"] == "radius" else oracle.MODE_KNN
        idx, d2 = oracle.bruteforce(samplepc, samplepc, mode, case["r"], case["k"])
        assert idx.tolist() == case["idx"], case
        assert np.array_equal(d2, golden_d2(case["d2"])), case
        gi, gd = oracle.grid_search(samplepc, samplepc, mode, case["r"], case["k"])
        assert np.array_equal(gi, idx) and np.array_equal(gd, d2)


def test_golden_known_answers_samplepc(samplepc):
    # Hand-derived known answers for src/samplepc.txt at r=2 and r=10 (an
    # FP32 brute force in file order, independent of the oracle).
    idx, d2 = oracle.bruteforce(samplepc, samplepc, oracle.MODE_RADIUS, 2.0, 4)
    assert idx[1].tolist() == [1, 5, 3, -1]
    assert idx[4].tolist() == [4, 2, -1, -1]
    idx, d2 = oracle.bruteforce(samplepc, samplepc, oracle.MODE_KNN, 10.0, 3)
    assert idx[0].tolist() == [2, 4, 6]
    assert d2[0, :2].tolist() == [25.0, 36.0]
    assert idx[7].tolist() == [3, 5, 1]


def test_d2_closed_forms():
    assert oracle.d2([0, 0, 0], [3, 4, 0]) == 25.0
    assert oracle.d2([1, 2, 3], [1, 2, 3]) == 0.0
    assert oracle.d2([0, 0, 0], [0, 0, -2]) == 4.0
    # Rounding of each product/sum separately (no FMA): compare with numpy FP32.
    rng = np.random.default_rng(7)
    for _ in range(2000):
        q = rng.uniform(-1e3, 1e3, 3).astype(np.float32)
        p = rng.uniform(-1e3, 1e3, 3).astype(np.float32)
        d = q - p
        sq = d * d
        want = np.float32(np.float32(sq[0] + sq[1]) + sq[2])
        assert oracle.d2(q, p) == want
    # Operation order matters: (a + b) + c differs from a + (b + c) here.
    q = np.array([1.0, 1.0, 1.0], np.float32)
    p = np.array([1.0 - 2.0 ** -12, 1.0 + 2.0 ** -12, 1.0 + 3.0], np.float32)
    d = q - p
    sq = d * d
    assert oracle.d2(q, p) == np.float32(np.float32(sq[0] + sq[1]) + sq[2])


def test_strict_radius_and_self_rules():
    pts = np.array([[0, 0, 0], [2, 0, 0], [1, 0, 0], [0, 0, 0]], np.float32)
    # d2 == r*r is excluded (strict <), self/duplicates included in radius mode.
    idx, d2 = oracle.bruteforce(pts, pts[:1], oracle.MODE_RADIUS, 2.0, 4)
    assert idx[0].tolist() == [0, 3, 2, -1]
    # KNN drops d2 == 0: both the query itself and the duplicate point 3.
    idx, d2 = oracle.bruteforce(pts, pts[:1], oracle.MODE_KNN, 2.0, 4)
    assert idx[0].tolist() == [2, -1, -1, -1]
    # Ties at equal d2 resolve by index.
    pts = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, 0, 0]], np.float32)
    idx, _ = oracle.bruteforce(pts, pts[3:], oracle.MODE_KNN, 5.0, 2)
    assert idx[0].tolist() == [0, 1]


@pytest.mark.parametrize("mode", [oracle.MODE_RADIUS, oracle.MODE_KNN])
@pytest.mark.parametrize("k", [1, 7, 40])
def test_bruteforce_vs_numpy(mode, k):
    rng = np.random.default_rng(100 + k + mode)
    pts = rng.uniform(0, 10, (600, 3)).astype(np.float32)
    pts[::50] = pts[1::50]  # exact duplicates
    qs = np.concatenate([pts[:100], rng.uniform(-2, 12, (100, 3)).astype(np.float32)])
    want_i, want_d = np_brute(pts, qs, mode, 1.7, k)
    got_i, got_d = oracle.bruteforce(pts, qs, mode, 1.7, k, threads=3)
    assert np.array_equal(got_i, want_i)
    assert np.array_equal(got_d, want_d)


@pytest.mark.parametrize("cell", [0.0, 0.37, 3.0, 25.0])
def test_grid_oracle_equals_bruteforce(cell):
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 20, (3000, 3)).astype(np.float32)
    qs = np.concatenate([pts[:300], rng.uniform(-5, 25, (300, 3)).astype(np.float32)])
    for mode in (oracle.MODE_RADIUS, oracle.MODE_KNN):
        bi, bd = oracle.bruteforce(pts, qs, mode, 2.0, 16, threads=4)
        gi, gd = oracle.grid_search(pts, qs, mode, 2.0, 16, cell=cell, threads=4)
        assert np.array_equal(bi, gi) and np.array_equal(bd, gd)


def test_qsel_rows():
    rng = np.random.default_rng(9)
    pts = rng.uniform(0, 5, (500, 3)).astype(np.float32)
    sel = np.array([7, 3, 499, 0], np.uint64)
    fi, fd = oracle.bruteforce(pts, pts, oracle.MODE_KNN, 1.0, 5)
    si, sd = oracle.bruteforce(pts, pts, oracle.MODE_KNN, 1.0, 5, qsel=sel)
    assert np.array_equal(si, fi[sel.astype(np.int64)])


def compact1by2(x):
    x &= 0x09249249
    x = (x ^ (x >> 2)) & 0x030C30C3
    x = (x ^ (x >> 4)) & 0x0300F00F
    x = (x ^ (x >> 8)) & 0xFF0000FF
    x = (x ^ (x >> 16)) & 0x000003FF
    return x


def test_morton_known_answers_and_roundtrip():
    assert oracle.morton3(1, 0, 0) == 1
    assert oracle.morton3(0, 1, 0) == 2
    assert oracle.morton3(0, 0, 1) == 4
    assert oracle.morton3(1, 1, 1) == 7
    assert oracle.morton3(1023, 1023, 1023) == 0x3FFFFFFF
    assert oracle.morton3(2, 0, 0) == 8
    rng = np.random.default_rng(3)
    for x, y, z in rng.integers(0, 1024, (500, 3)):
        m = oracle.morton3(int(x), int(y), int(z))
        assert compact1by2(m) == x and compact1by2(m >> 1) == y and compact1by2(m >> 2) == z


def test_search_size_closed_forms():
    # Uniform count c per cell: the (2i+1)^3 block holds c*(2i+1)^3 points in
    # the interior, so iter is the smallest i with c*(2i+1)^3 >= k+1 unless the
    # width (2i+2)*cell exceeds maxWidth first.
    counts = np.full((21, 21, 21), 2, np.uint32)
    for k in [1, 2, 53, 54, 249, 250, 700]:
        want = 0
        while 2 * (2 * want + 1) ** 3 < k + 1:
            want += 1
        got = oracle.search_size(counts, 10, 10, 10, cell=1.0, max_width=100.0, knn=k)
        assert got == want, (k, got, want)
    # Width cap: maxWidth 5 with cell 1 stops at iter 2 (width 6 > 5).
    assert oracle.search_size(counts, 10, 10, 10, 1.0, 5.0, 10 ** 6) == 2
    # Border cells count only in-grid neighbours: corner (0,0,0), c=1 ->
    # block holds (i+1)^3 cells.
    ones = np.ones((8, 8, 8), np.uint32)
    assert oracle.search_size(ones, 0, 0, 0, 1.0, 100.0, 7) == 1   # 8 >= 8
    assert oracle.search_size(ones, 0, 0, 0, 1.0, 100.0, 8) == 2   # 27 >= 9


def test_boundary_pairs():
    pts = np.array([[0, 0, 0], [2.0, 0, 0], [1.9999999, 0, 0], [1.5, 0, 0]], np.float32)
    cnt, pairs = oracle.boundary_pairs(pts, pts[:1], 2.0, tol=1e-6)
    assert cnt == 2
    assert sorted(int(p) for _, p in pairs) == [1, 2]
