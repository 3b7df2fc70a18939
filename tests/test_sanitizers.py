"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

Three programs are built with gcc/g++ -fsanitize=address,undefined
-fno-sanitize-recover=all and run on this machine:

  * tests/sanitize/oracle_main.c + oracle/oracle.c: every oracle function on
    random and degenerate inputs, grid search equal to the brute force;
  * tests/sanitize/data_main.cpp + rtnn_amd/csrc/data.cpp: the point-file
    readers (malformed, blank, over-long, empty, missing, and changed-between-
    calls files), the writer round trip, every generator;
  * rtnn_amd/csrc/cli.cpp + data.cpp + tests/sanitize/abi_double.cpp (a test
    double of the C ABI that scans all points, so no GPU is needed): the CLI's
    host paths -- flags, bad flags and numbers, file errors, -fq, -c 1, -o,
    --boundary, MLP weight files.

Any sanitizer report fails the test (halt_on_error, a nonzero exit code, and
the report text checked for).  The reference's driver has no counterpart; it
asserts on unreadable files and trusts its inputs (util.cpp:101-135).
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-g", "-O1", "-ffp-contract=off"]
ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
KNN = 8  # compile-time K of the sanitized CLI build (RTNN_KNN)

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="no host compiler")


def _build(out, cmd):
    r = subprocess.run(cmd + ["-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return str(out)


def _run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=120, **kw)
    report = r.stderr + r.stdout
    for bad in ("AddressSanitizer", "LeakSanitizer", "runtime error:", "UndefinedBehavior"):
        assert bad not in report, report[-4000:]
    return r


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    inc = ["-I", os.path.join(ROOT, "include")]
    oracle = _build(d / "oracle_san", ["gcc", "-std=c11", *SAN, "-pthread",
                                       os.path.join(ROOT, "tests/sanitize/oracle_main.c"),
                                       os.path.join(ROOT, "oracle/oracle.c"), "-lm"])
    data = _build(d / "data_san", ["g++", "-std=c++17", *SAN, *inc,
                                   os.path.join(ROOT, "tests/sanitize/data_main.cpp"),
                                   os.path.join(ROOT, "rtnn_amd/csrc/data.cpp")])
    cli = _build(d / "cli_san", ["g++", "-std=c++17", *SAN, *inc, f"-DRTNN_KNN={KNN}",
                                 os.path.join(ROOT, "rtnn_amd/csrc/cli.cpp"),
                                 os.path.join(ROOT, "rtnn_amd/csrc/data.cpp"),
                                 os.path.join(ROOT, "tests/sanitize/abi_double.cpp")])
    return {"oracle": oracle, "data": data, "cli": cli, "dir": d}


def test_oracle_under_sanitizers(bins):
    r = _run([bins["oracle"]])
    assert r.returncode == 0, r.stderr[-3000:]
    assert "oracle checks passed" in r.stdout


def test_data_utilities_under_sanitizers(bins, tmp_path):
    r = _run([bins["data"], str(tmp_path)])
    assert r.returncode == 0, r.stderr[-3000:]
    assert "data checks passed" in r.stdout


def _write(path, pts):
    with open(path, "w") as f:
        for p in pts:
            f.write("%.9g,%.9g,%.9g\n" % tuple(p))


def _brute(pts, qs, r, k, knn):
    P = pts.astype(np.float32)
    out_i = np.full((len(qs), k), -1, np.int32)
    out_d = np.full((len(qs), k), np.inf, np.float32)
    for i, q in enumerate(qs.astype(np.float32)):
        d = P - q
        s = d * d
        d2 = (s[:, 0] + s[:, 1]) + s[:, 2]
        ok = (d2 < np.float32(r) * np.float32(r)) & ((d2 > 0) if knn else True)
        sel = np.nonzero(ok)[0]
        order = np.lexsort((sel, d2[sel]))[:k]
        out_i[i, :len(order)] = sel[order]
        out_d[i, :len(order)] = d2[sel[order]]
    return out_i, out_d


def _read_result(path):
    raw = open(path, "rb").read()
    assert raw[:8] == b"RTNNRES1"
    nq, k, flags = struct.unpack("<QII", raw[8:24])
    idx = np.frombuffer(raw[24:24 + 4 * nq * k], np.int32).reshape(nq, k)
    d2 = np.frombuffer(raw[24 + 4 * nq * k:], np.float32).reshape(nq, k)
    assert flags == 1 and len(raw) == 24 + 8 * nq * k
    return idx, d2


def test_cli_host_paths_under_sanitizers(bins, tmp_path):
    cli = bins["cli"]
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 6, (700, 3)).astype(np.float32)
    qs = np.concatenate([rng.uniform(-1, 7, (120, 3)), [[50, 50, 50], [-40, 3, 3]]]).astype(
        np.float32)
    pf, qf = tmp_path / "p.txt", tmp_path / "q.txt"
    _write(pf, pts)
    _write(qf, qs)
    pts = np.loadtxt(pf, delimiter=",", dtype=np.float64).astype(np.float32)
    qs = np.loadtxt(qf, delimiter=",", dtype=np.float64).astype(np.float32)

    # usage and flag errors end with a message, not a crash
    assert _run([cli, "--help"]).returncode == 0
    for bad in ([], ["-f"], ["-f", str(pf), "-r", "abc"], ["-f", str(pf), "-k", "0"],
                ["-f", str(pf), "-k", "12x"], ["-f", str(pf), "--crStep", "0.5"],
                ["-f", str(pf), "-sm", "nearest"], ["-f", str(pf), "--nope", "1"],
                ["-f", str(pf), "-d", "99999999999"]):
        r = _run([cli, *bad])
        assert r.returncode == 1, (bad, r.returncode, r.stderr[-500:])
    assert _run([cli, "-f", str(tmp_path / "missing.txt")]).returncode == 1
    (tmp_path / "bad.txt").write_text("1,2,3\n4,5\n")
    assert _run([cli, "-f", str(tmp_path / "bad.txt")]).returncode == 1
    (tmp_path / "empty.txt").write_text("")
    r = _run([cli, "-f", str(tmp_path / "empty.txt")])
    assert r.returncode == 0 and "empty query" in r.stdout

    # range search of the points themselves: check, result file, boundary file
    out, bnd = tmp_path / "r.bin", tmp_path / "b.txt"
    r = _run([cli, "-f", str(pf), "-r", "0.7", "-k", "9", "-c", "1", "-o", str(out),
              "--boundary", str(bnd)])
    assert r.returncode == 0 and "Sanity check done." in r.stderr, r.stderr[-2000:]
    idx, d2 = _read_result(out)
    wi, wd = _brute(pts, pts, 0.7, 9, knn=False)
    assert np.array_equal(idx, wi) and np.array_equal(d2, wd)
    assert "time search compute" in r.stdout and "boundary pairs" in r.stdout

    # KNN over separate queries, two far ones (-fq reports them), -ac, -ab/-nb
    out = tmp_path / "k.bin"
    r = _run([cli, "-f", str(pf), "-q", str(qf), "-sm", "knn", "-r", "1.2", "-c", "1",
              "-fq", "1", "-ac", "1", "-ab", "0", "-nb", "3", "-o", str(out)])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Filter queries: 2 " in r.stdout
    idx, d2 = _read_result(out)
    wi, wd = _brute(pts, qs, 1.2, KNN, knn=True)
    assert np.array_equal(idx, wi) and np.array_equal(d2, wd)

    # learned radius: weights saved, loaded, and a truncated weight file refused
    w = tmp_path / "w.bin"
    r = _run([cli, "-f", str(pf), "-sm", "knn", "-r", "1.0", "--learned", "1",
              "--mlp-out", str(w)])
    assert r.returncode == 0 and os.path.getsize(w) == 4 * 6211
    assert _run([cli, "-f", str(pf), "-sm", "knn", "--learned", "1", "--mlp-in",
                 str(w)]).returncode == 0
    (tmp_path / "short.bin").write_bytes(b"\0" * 100)
    assert _run([cli, "-f", str(pf), "-sm", "knn", "--learned", "1", "--mlp-in",
                 str(tmp_path / "short.bin")]).returncode == 1
    # the advanced flags the CLI accepts (scheduling flags of the reference)
    r = _run([cli, "-f", str(pf), "-a", "1", "-s", "2", "-sg", "2", "-egs", "0", "-crs", "2",
              "-gmu", "1", "-mc", "2", "-i", "0", "-m", "0", "-df", "1", "-g", "1", "-ps", "1",
              "-qs", "1", "-cr", "2", "--cell-factor", "1.5", "--qpw", "16", "-p", "0"])
    assert r.returncode == 0, r.stderr[-2000:]
