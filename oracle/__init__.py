"""CPU oracle for RTNN neighbor search -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker; the product path never routes through
it.  Each wrapper names the function of oracle/oracle.c it calls; that file
cites the reference lines each one restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None

MODE_RADIUS = 0
MODE_KNN = 1


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        from . import build as _b
        _b.build()
    lib = ctypes.CDLL(_LIB_PATH)
    f32p = ctypes.POINTER(ctypes.c_float)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    i32p = ctypes.POINTER(ctypes.c_int32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.orc_d2.restype = ctypes.c_float
    lib.orc_d2.argtypes = [f32p, f32p]
    lib.orc_bruteforce.restype = ctypes.c_int
    lib.orc_bruteforce.argtypes = [f32p, ctypes.c_uint64, f32p, u64p, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                                   i32p, f32p, ctypes.c_int]
    lib.orc_grid_search.restype = ctypes.c_int
    lib.orc_grid_search.argtypes = [f32p, ctypes.c_uint64, f32p, u64p, ctypes.c_uint64,
                                    ctypes.c_int, ctypes.c_float, ctypes.c_uint32,
                                    ctypes.c_float, i32p, f32p, ctypes.c_int]
    lib.orc_morton3.restype = ctypes.c_uint32
    lib.orc_morton3.argtypes = [ctypes.c_uint32] * 3
    lib.orc_search_size.restype = ctypes.c_int
    lib.orc_search_size.argtypes = [u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_float, ctypes.c_float, ctypes.c_uint32]
    lib.orc_boundary_pairs.restype = ctypes.c_uint64
    lib.orc_boundary_pairs.argtypes = [f32p, ctypes.c_uint64, f32p, u64p, ctypes.c_uint64,
                                       ctypes.c_float, ctypes.c_double, u64p,
                                       ctypes.c_uint64]
    _lib = lib
    return lib


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _sel(qsel):
    if qsel is None:
        return None, None
    s = np.ascontiguousarray(qsel, dtype=np.uint64)
    return s, s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def d2(q, p):
    """orc_d2: FP32 ((dx*dx + dy*dy) + dz*dz), dx = q - p, no FMA."""
    q, qp = _f32(q)
    p, pp = _f32(p)
    return float(_load().orc_d2(qp, pp))


def bruteforce(points, queries, mode, r, k, qsel=None, threads=1):
    """orc_bruteforce -> (idx int32 [m, k], d2 float32 [m, k])."""
    lib = _load()
    pts, pp = _f32(points)
    qs, qp = _f32(queries)
    sel, sp = _sel(qsel)
    m = len(sel) if sel is not None else len(qs)
    idx = np.empty((m, k), dtype=np.int32)
    dd = np.empty((m, k), dtype=np.float32)
    rc = lib.orc_bruteforce(pp, len(pts), qp, sp, m, mode, r, k,
                            idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                            dd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), threads)
    if rc != 0:
        raise ValueError("orc_bruteforce failed")
    return idx, dd


def grid_search(points, queries, mode, r, k, qsel=None, cell=0.0, threads=8):
    """orc_grid_search -> (idx, d2); equal to bruteforce() by construction."""
    lib = _load()
    pts, pp = _f32(points)
    qs, qp = _f32(queries)
    sel, sp = _sel(qsel)
    m = len(sel) if sel is not None else len(qs)
    idx = np.empty((m, k), dtype=np.int32)
    dd = np.empty((m, k), dtype=np.float32)
    rc = lib.orc_grid_search(pp, len(pts), qp, sp, m, mode, r, k, cell,
                             idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                             dd.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), threads)
    if rc != 0:
        raise ValueError("orc_grid_search failed")
    return idx, dd


def morton3(x, y, z):
    """orc_morton3 (helper_mortonCode.h MortonCode3 on 10-bit coordinates)."""
    return int(_load().orc_morton3(x, y, z))


def search_size(counts, x, y, z, cell, max_width, knn):
    """orc_search_size on counts[z, y, x] (raster grid)."""
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    dz, dy, dx = c.shape
    return int(_load().orc_search_size(c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                       dx, dy, dz, x, y, z, cell, max_width, knn))


def boundary_pairs(points, queries, r, tol=1e-6, qsel=None, max_out=1 << 16):
    """orc_boundary_pairs -> (count, pairs[min(count, max_out), 2])."""
    lib = _load()
    pts, pp = _f32(points)
    qs, qp = _f32(queries)
    sel, sp = _sel(qsel)
    m = len(sel) if sel is not None else len(qs)
    out = np.zeros((max_out, 2), dtype=np.uint64)
    cnt = lib.orc_boundary_pairs(pp, len(pts), qp, sp, m, r, tol,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), max_out)
    return int(cnt), out[:min(int(cnt), max_out)]
