"""CPU tests of the host side: the C ABI library loads and exports every symbol
declared in include/*.h, the point-file reader follows the reference's
read_pc_data (util.cpp:101-135), and the seeded workload generators are
deterministic and shaped as documented.  No GPU needed."""
import os
import re

import numpy as np
import pytest

import rtnn_amd
from conftest import ROOT


def declared_functions():
    names = []
    for h in ("rtnn_hip.h", "rtnn_data.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(rtnn_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = rtnn_amd.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
        assert n in rtnn_amd.SIGNATURES, f"{n} has no ctypes signature"
    assert set(rtnn_amd.SIGNATURES) == set(names)
    assert b"gfx950" in L.rtnn_version()


def test_read_points_matches_reference_parse(tmp_path):
    src = os.path.join(ROOT, "tests", "golden", "samplepc.txt")
    pts = rtnn_amd.read_points(src)
    # sscanf("%lf") then make_float3(double -> float)
    want = np.loadtxt(src, delimiter=",", dtype=np.float64).astype(np.float32)
    assert pts.shape == (8, 3) and np.array_equal(pts, want)
    p = tmp_path / "a.txt"
    p.write_text("1,2,3\n\n4.5, 5e1 ,-6\n  \n7,8,9")
    got = rtnn_amd.read_points(str(p))
    assert got.tolist() == [[1, 2, 3], [4.5, 50, -6], [7, 8, 9]]
    bad = tmp_path / "bad.txt"
    bad.write_text("1,2,3\n4,5\n")
    with pytest.raises(ValueError, match="line 2"):
        rtnn_amd.read_points(str(bad))
    with pytest.raises(FileNotFoundError):
        rtnn_amd.read_points(str(tmp_path / "missing.txt"))
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    assert rtnn_amd.read_points(str(empty)).shape == (0, 3)


def test_read_points_file_grown_between_calls(tmp_path):
    # The counting call sizes the caller's buffer; a file that grows before
    # the filling call must not write past that buffer (-4 instead).
    import ctypes
    L = rtnn_amd.lib()
    p = tmp_path / "grow.txt"
    p.write_text("1,2,3\n4,5,6\n")
    n = ctypes.c_uint64(0)
    assert L.rtnn_read_points(str(p).encode(), None, ctypes.byref(n)) == 0 and n.value == 2
    with open(p, "a") as f:
        f.write("7,8,9\n10,11,12\n")
    buf = np.full(3 * 2 + 16, -7.0, dtype=np.float32)  # 2 points + a guard
    rc = L.rtnn_read_points(str(p).encode(), ctypes.c_void_p(buf.ctypes.data), ctypes.byref(n))
    assert rc == -4
    assert buf[:6].tolist() == [1, 2, 3, 4, 5, 6]
    assert (buf[6:] == -7.0).all()
    # a file that shrank: the second call reports the points it filled
    p.write_text("1,2,3\n")
    n = ctypes.c_uint64(2)
    assert L.rtnn_read_points(str(p).encode(), ctypes.c_void_p(buf.ctypes.data),
                              ctypes.byref(n)) == 0 and n.value == 1


def test_read_point_sets(tmp_path):
    # Each line: 3 * m numbers = point i of m sets; every number parsed
    # straight to float (stof), unlike the single-set reader's double.
    p = tmp_path / "sets.txt"
    p.write_text("0.1,2,3,4,5,6\n7, 8 ,9,10,11,0.3\n\n1e-3,-2,3.5,4,5,6\n")
    got = rtnn_amd.read_point_sets(str(p))
    assert got.shape == (2, 3, 3)
    want0 = np.array([["0.1", "2", "3"], ["7", "8", "9"], ["1e-3", "-2", "3.5"]]).astype(np.float32)
    want1 = np.array([[4, 5, 6], [10, 11, np.float32("0.3")], [4, 5, 6]], np.float32)
    assert np.array_equal(got[0], want0) and np.array_equal(got[1], want1)
    # decimal -> float directly: 0.1 parsed once, not via double
    assert got[0, 0, 0] == np.float32("0.1")
    # a first line whose count is not a multiple of 3 is padded with zeros
    q = tmp_path / "pad.txt"
    q.write_text("1,2,3,4\n5,6,7,8\n")
    got = rtnn_amd.read_point_sets(str(q))
    assert got.shape == (2, 2, 3) and got[1].tolist() == [[4, 0, 0], [8, 0, 0]]
    # another count of numbers, or a non-number, is an error naming the line
    for text in ("1,2,3,4,5,6\n1,2,3\n", "1,2,3\n4,x,6\n"):
        bad = tmp_path / "bad.txt"
        bad.write_text(text)
        with pytest.raises(ValueError, match="line 2"):
            rtnn_amd.read_point_sets(str(bad))
    # one set = the single-set file format
    src = os.path.join(ROOT, "tests", "golden", "samplepc.txt")
    one = rtnn_amd.read_point_sets(src)
    assert one.shape == (1, 8, 3)
    assert np.allclose(one[0], rtnn_amd.read_points(src), rtol=1e-7, atol=0)


def test_write_read_roundtrip(tmp_path):
    x = rtnn_amd.gen_uniform(3, 1000, -7.0, 1e4)
    p = str(tmp_path / "x.txt")
    rtnn_amd.write_points(p, x)
    assert np.array_equal(rtnn_amd.read_points(p), x)


def test_generators_deterministic_and_shaped():
    a = rtnn_amd.gen_uniform(1, 100000, 0.0, 50.0)
    assert np.array_equal(a, rtnn_amd.gen_uniform(1, 100000, 0.0, 50.0))
    assert a.min() >= 0 and a.max() < 50 and abs(a.mean() - 25) < 0.2
    g = rtnn_amd.gen_gaussian_mixture(2, 20, 50000)
    q = rtnn_amd.gen_gaussian_mixture(2, 21, 50000)
    assert not np.array_equal(g, q)
    assert np.array_equal(g, rtnn_amd.gen_gaussian_mixture(2, 20, 50000))
    s = rtnn_amd.gen_spheres(3, 30, 20000, nspheres=5, box=100.0, rmin=5, rmax=10, noise=0.0)
    assert np.isfinite(s).all()
    lv = rtnn_amd.gen_levy(5, 100000)
    assert lv.min() >= 0 and lv.max() <= 1000
    xyz, offs = rtnn_amd.gen_lattice_clouds(4, 16, 1000, 20000)
    sizes = np.diff(offs.astype(np.int64))
    assert sizes.min() >= 1000 and sizes.max() <= 20000 and len(xyz) == offs[-1]
    # jittered lattice: every point within 0.02h of a lattice site (h = 0.1)
    c0 = xyz[: sizes[0]]
    frac = np.abs(c0 / 0.1 - np.round(c0 / 0.1))
    assert frac.max() <= 0.02 + 1e-4


def test_sphere_points_lie_on_spheres():
    s = rtnn_amd.gen_spheres(8, 9, 5000, nspheres=1, box=100.0, rmin=10, rmax=10.0001,
                             noise=0.0)
    c = s.mean(axis=0)
    r = np.linalg.norm(s - c, axis=1)
    assert abs(r.mean() - 10) < 0.2 and r.std() < 0.2


def test_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(rtnn_amd.RtnnError):
        rtnn_amd.Context(0)
