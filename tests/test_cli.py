"""The rtnn command-line driver: flag surface of the reference's optixNSearch
(util.cpp:137-383) and its result file."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden_d2

CLI = os.path.join(ROOT, "rtnn_amd", "rtnn")
SAMPLE = os.path.join(ROOT, "tests", "golden", "samplepc.txt")


def run(*args):
    return subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True, timeout=300)


def read_result(path):
    with open(path, "rb") as f:
        assert f.read(8) == b"RTNNRES1"
        nq, k, flags = struct.unpack("<QII", f.read(16))
        idx = np.frombuffer(f.read(4 * nq * k), np.int32).reshape(nq, k)
        d2 = np.frombuffer(f.read(4 * nq * k), np.float32).reshape(nq, k) if flags & 1 else None
    return idx, d2


def test_help_and_bad_flags():
    r = run("-h")
    assert r.returncode == 0 and "--searchmode" in r.stderr
    assert run("-f", SAMPLE, "--bogus", "1").returncode != 0
    assert run("-f", SAMPLE, "-sm", "nearest").returncode != 0
    assert run("-f", SAMPLE, "-crs", "1.0").returncode != 0
    assert run("-sm", "knn").returncode != 0  # no point file


def test_missing_and_malformed_files(tmp_path):
    r = run("-f", str(tmp_path / "nope.txt"))
    assert r.returncode != 0
    bad = tmp_path / "bad.txt"
    bad.write_text("1,2,3\n4,5\n")
    assert run("-f", str(bad)).returncode != 0
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    r = run("-f", str(empty))
    assert r.returncode == 0 and "empty query and/or points" in r.stdout


@pytest.mark.gpu
def test_cli_matches_golden(golden, tmp_path):
    out = str(tmp_path / "res.bin")
    r = run("-f", SAMPLE, "-r", "2", "-k", "50", "-c", "1", "-o", out)
    assert r.returncode == 0, r.stderr
    assert "Sanity check done." in r.stderr and "time total search time" in r.stdout
    idx, d2 = read_result(out)
    case = golden["cases"][0]
    assert idx.tolist() == case["idx"]
    assert np.array_equal(d2, golden_d2(case["d2"]))


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [[], ["-p", "0"], ["--qpw", "16"], ["-ab", "0", "-nb", "3"],
                                   ["-a", "0", "-gmu", "100", "-mc", "2", "-egs", "10"]])
def test_cli_knn_separate_queries(tmp_path, extra):
    import oracle
    rng = np.random.default_rng(3)
    p = rng.uniform(0, 20, (5000, 3)).astype(np.float32)
    q = rng.uniform(0, 20, (700, 3)).astype(np.float32)
    pf, qf, out = tmp_path / "p.txt", tmp_path / "q.txt", str(tmp_path / "o.bin")
    np.savetxt(pf, p, fmt="%.9g", delimiter=",")
    np.savetxt(qf, q, fmt="%.9g", delimiter=",")
    r = run("-f", str(pf), "-q", str(qf), "-sm", "knn", "-r", "3", "-c", "1", "-o", out, *extra)
    assert r.returncode == 0, r.stderr
    if "-nb" in extra:  # the partition's radius classes became grid levels
        assert "grid levels: 3" in r.stdout
    idx, d2 = read_result(out)
    k = idx.shape[1]
    wi, wd = oracle.bruteforce(p, q, oracle.MODE_KNN, 3.0, k, threads=4)
    assert np.array_equal(idx, wi) and np.array_equal(d2, wd)


@pytest.mark.gpu
def test_cli_crratio_autotune_and_boundary(tmp_path):
    # -cr sets the cell factor, -ac 1 runs the autotuner, --boundary lists the
    # pairs within 1e-6 r of r; the result stays the exact range search.
    import oracle
    g = np.stack(np.meshgrid(*[np.arange(12, dtype=np.float32) * np.float32(0.1)] * 3),
                 -1).reshape(-1, 3)
    pf, out, bf = tmp_path / "p.txt", str(tmp_path / "o.bin"), str(tmp_path / "b.txt")
    np.savetxt(pf, g, fmt="%.9g", delimiter=",")
    p = np.loadtxt(pf, delimiter=",", dtype=np.float64).astype(np.float32)
    for extra in (["-cr", "2"], ["-ac", "1"]):
        r = run("-f", str(pf), "-r", "0.2", "-k", "40", "-o", out, "--boundary", bf, *extra)
        assert r.returncode == 0, r.stderr
        idx, d2 = read_result(out)
        wi, wd = oracle.bruteforce(p, p, oracle.MODE_RADIUS, 0.2, 40, threads=4)
        assert np.array_equal(idx, wi) and np.array_equal(d2, wd)
        cnt, want = oracle.boundary_pairs(p, p, 0.2, max_out=1 << 20)
        got = np.loadtxt(bf, dtype=np.float64, ndmin=2)
        assert cnt > 0 and len(got) == cnt
        assert np.array_equal(got[:, :2].astype(np.uint64), want)
    assert "autotune: cell_factor" in r.stdout


@pytest.mark.gpu
def test_cli_filter_queries_reports_far_queries(tmp_path):
    # -fq 1: queries outside the points' (integer-rounded) box grown by r are
    # counted as the reference's filterRemoteQueries does; their rows stay, empty.
    rng = np.random.default_rng(4)
    p = rng.uniform(0, 10, (3000, 3)).astype(np.float32)
    q = np.concatenate([rng.uniform(0, 10, (300, 3)), rng.uniform(20, 30, (100, 3)),
                        [[-3.2, 5, 5], [13.9, 5, 5]]]).astype(np.float32)
    pf, qf, out = tmp_path / "p.txt", tmp_path / "q.txt", str(tmp_path / "o.bin")
    np.savetxt(pf, p, fmt="%.9g", delimiter=",")
    np.savetxt(qf, q, fmt="%.9g", delimiter=",")
    r = run("-f", str(pf), "-q", str(qf), "-r", "3", "-k", "8", "-fq", "1", "-o", out)
    assert r.returncode == 0, r.stderr
    # box [0, 10] (floor(max) + 1 = 10), grown by 3: [-3, 13]; -3.2 and 13.9 fall outside
    assert "Filter queries: 102 (" in r.stdout
    idx, _ = read_result(out)
    assert (idx[300:] == -1).all() and (idx[:300, 0] >= 0).all()


@pytest.mark.gpu
def test_cli_knn_default_path_on_a_wide_box(tmp_path):
    # The default KNN path (-p 1: the density partition, hierarchical Morton
    # grid) on a box 2000 r long, wider than the Morton grid's 1024 top-level
    # cells per axis: it falls back to the raster grid and stays exact.  The
    # reference's own printed scopes appear.
    import oracle
    rng = np.random.default_rng(6)
    p = np.concatenate([rng.uniform([0, 0, 0], [2000, 2, 2], (6000, 3)),
                        rng.normal([500, 1, 1], 0.2, (800, 3))]).astype(np.float32)
    q = rng.uniform([0, 0, 0], [2000, 2, 2], (900, 3)).astype(np.float32)
    pf, qf, out = tmp_path / "p.txt", tmp_path / "q.txt", str(tmp_path / "o.bin")
    np.savetxt(pf, p, fmt="%.9g", delimiter=",")
    np.savetxt(qf, q, fmt="%.9g", delimiter=",")
    p = np.loadtxt(pf, delimiter=",", dtype=np.float64).astype(np.float32)
    q = np.loadtxt(qf, delimiter=",", dtype=np.float64).astype(np.float32)
    r = run("-f", str(pf), "-q", str(qf), "-sm", "knn", "-r", "1", "-c", "1", "-o", out)
    assert r.returncode == 0, r.stderr
    for scope in ("sort points", "sort and/or partition queries", "search compute",
                  "result copy D2H", "total search time"):
        assert f"time {scope}: " in r.stdout, scope
    idx, d2 = read_result(out)
    wi, wd = oracle.bruteforce(p, q, oracle.MODE_KNN, 1.0, idx.shape[1], threads=4)
    assert np.array_equal(idx, wi) and np.array_equal(d2, wd)
