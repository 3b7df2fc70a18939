"""rtnn_amd -- Python host mirror of the gfx950 neighbor-search library.

Thin ctypes layer over the C ABI in include/rtnn_hip.h and include/rtnn_data.h
(librtnn_hip.so, built in-tree by rtnn_amd/build.py).  There is no fallback:
if the library is missing or no GPU is present, calls raise.

Search semantics mirror the reference driver (src/optixNSearch):
  mode "radius": include iff d2 < r*r, the query itself included
                 (geometry.cu:38-106), at most k results;
  mode "knn"   : include iff 0 < d2 < r*r (geometry.cu:157-191), k nearest.
Rows are original query indices, values original point indices in (d2, index)
order, empty slots -1 (the reference's UINT_MAX sentinel, search.cpp:14-18).
"""
import ctypes
import struct
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtnn_hip.so")

MODE_RADIUS = 0
MODE_KNN = 1
MLP_NBLOB = 6211
MLP_NPARAMS = 6209

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_OOM = -3
ERR_STATE = -4
ERR_UNSUPPORTED = -5

_lib = None


class RtnnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rtnn error {code}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("cell_factor", ctypes.c_float),
                ("queries_per_wave", ctypes.c_int),
                ("query_order", ctypes.c_int),
                ("time_stages", ctypes.c_int),
                ("max_cells", ctypes.c_uint32),
                ("grid_levels", ctypes.c_int),
                ("level_alpha", ctypes.c_float),
                ("group_buffer", ctypes.c_int),
                ("lds_tile", ctypes.c_int),
                ("tree_leaf", ctypes.c_int),
                ("bucket_ratio", ctypes.c_float),
                ("learned_cap", ctypes.c_int),
                ("tree_lanes", ctypes.c_int),
                ("tree_fallback", ctypes.c_int)]


def _options_struct(cell_factor=0.0, queries_per_wave=0, query_order=0, time_stages=0,
                    max_cells=0, grid_levels=0, level_alpha=0.0, group_buffer=0, lds_tile=0,
                    tree_leaf=0, bucket_ratio=0.0, learned_cap=0, tree_lanes=0,
                    tree_fallback=0):
    return Options(cell_factor, queries_per_wave, query_order, time_stages, max_cells,
                   grid_levels, level_alpha, group_buffer, lds_tile, tree_leaf, bucket_ratio,
                   learned_cap, tree_lanes, tree_fallback)


def _options_dict(o):
    return {name: getattr(o, name) for name, _ in Options._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double),
                ("ms_grid", ctypes.c_double),
                ("ms_qsort", ctypes.c_double),
                ("ms_search", ctypes.c_double),
                ("ms_mlp", ctypes.c_double),
                ("ms_fallback", ctypes.c_double),
                ("fallback_queries", ctypes.c_uint64),
                ("grid_dims", ctypes.c_uint32 * 3),
                ("cell_size", ctypes.c_float),
                ("sort_passes", ctypes.c_uint32),
                ("search_launches", ctypes.c_uint32),
                ("lanes_per_query", ctypes.c_uint32),
                ("ms_copy", ctypes.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["grid_dims"] = list(self.grid_dims)
        return d


_VP = ctypes.c_void_p
_U64 = ctypes.c_uint64
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_F = ctypes.c_float

# name -> (restype, argtypes); every symbol of include/rtnn_hip.h and rtnn_data.h
SIGNATURES = {
    "rtnn_create": (_I, [_I, ctypes.POINTER(_VP)]),
    "rtnn_destroy": (_I, [_VP]),
    "rtnn_last_error": (ctypes.c_char_p, [_VP]),
    "rtnn_create_error": (ctypes.c_char_p, []),
    "rtnn_set_options": (_I, [_VP, ctypes.POINTER(Options)]),
    "rtnn_synchronize": (_I, [_VP]),
    "rtnn_set_points": (_I, [_VP, _VP, _U64, _I]),
    "rtnn_search": (_I, [_VP, _VP, _U64, _I, _I, _F, _U32, _VP, _VP, _I,
                         ctypes.POINTER(Stats)]),
    "rtnn_search_async": (_I, [_VP, _VP, _U64, _I, _F, _U32, _VP, _VP, _VP]),
    "rtnn_search_noalloc": (_I, [_VP, _VP, _U64, _VP, _U64, _I, _F, _U32, _VP, _VP]),
    "rtnn_search_partitioned": (_I, [_VP, _VP, _U64, _I, _F, _U32, _VP, _VP, _I,
                                     ctypes.POINTER(Stats)]),
    "rtnn_mlp_train": (_I, [_VP, _F, _U32, _U32, _U32, _U32, _F, _F, _U64,
                            ctypes.POINTER(_F)]),
    "rtnn_mlp_train_queries": (_I, [_VP, _VP, _U64, _I, _F, _U32, _U32, _U32, _F, _F, _U64,
                                    ctypes.POINTER(_F)]),
    "rtnn_mlp_get_weights": (_I, [_VP, _VP]),
    "rtnn_mlp_set_weights": (_I, [_VP, _VP]),
    "rtnn_search_learned": (_I, [_VP, _VP, _U64, _I, _F, _U32, _VP, _VP, _I,
                                 ctypes.POINTER(Stats)]),
    "rtnn_propose_radius": (_I, [_VP, _VP, _U64, _I, _I, _F, _U32, _VP, _I]),
    "rtnn_search_stream": (_I, [_VP, _VP, _U64, _I, _I, _F, _U32, _U64, _VP, _VP,
                                ctypes.c_char_p, ctypes.POINTER(Stats)]),
    "rtnn_search_batched_stream": (_I, [_VP, _VP, _VP, _U32, _I, _F, _U32, _U64, _VP, _VP,
                                        ctypes.c_char_p, ctypes.POINTER(Stats)]),
    "rtnn_mem_stats": (_I, [ctypes.POINTER(_U64)] * 4),
    "rtnn_search_batched": (_I, [_VP, _VP, _VP, _U32, _I, _F, _U32, _VP, _VP,
                                 ctypes.POINTER(Stats)]),
    "rtnn_search_percloud": (_I, [_VP, _VP, _VP, _U32, _I, _F, _U32, _VP, _VP,
                                  ctypes.POINTER(Stats)]),
    "rtnn_graph_batched": (_I, [_VP, _VP, _VP, _U32, _I, _F, _U32, _VP, _VP,
                                ctypes.POINTER(_VP)]),
    "rtnn_graph_launch": (_I, [_VP, _I, ctypes.POINTER(Stats)]),
    "rtnn_graph_destroy": (_I, [_VP]),
    "rtnn_boundary_pairs": (_I, [_VP, _VP, _U64, _I, _F, ctypes.c_double, _VP, _U64, _I,
                                 ctypes.POINTER(_U64)]),
    "rtnn_boundary_pairs_batched": (_I, [_VP, _VP, _VP, _U32, _F, ctypes.c_double, _VP, _U64,
                                         _I, ctypes.POINTER(_U64)]),
    "rtnn_autotune": (_I, [_VP, _VP, _U64, _I, _I, _I, _F, _U32, _VP, _VP,
                           ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_double)]),
    "rtnn_autotune_batched": (_I, [_VP, _VP, _VP, _U32, _I, _F, _U32, _VP, _VP,
                                   ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_double)]),
    "rtnn_autotune_batched_classes": (_I, [_VP, _VP, _VP, _U32, _I, _F, _U32, _VP, _VP, _U32,
                                           _VP, _VP, _VP]),
    "rtnn_set_size_classes": (_I, [_VP, _U32, _VP, _VP]),
    "rtnn_get_size_classes": (_I, [_VP, ctypes.POINTER(_U32), _VP, _VP]),
    "rtnn_radix_sort_pairs": (_I, [_VP, _VP, _VP, _U64, _U32]),
    "rtnn_exclusive_scan": (_I, [_VP, _VP, _VP, _U64]),
    "rtnn_morton3": (_I, [_VP, _VP, _VP, _U64]),
    "rtnn_mlp_forward": (_I, [_VP, _VP, _U64, _VP]),
    "rtnn_mlp_backward": (_I, [_VP, _VP, _VP, _U64, _F, _VP, _VP]),
    "rtnn_mlp_features": (_I, [_VP, _VP, _U64, _F, _VP]),
    "rtnn_version": (ctypes.c_char_p, []),
    "rtnn_read_points": (_I, [ctypes.c_char_p, _VP, ctypes.POINTER(_U64)]),
    "rtnn_write_points": (_I, [ctypes.c_char_p, _VP, _U64]),
    "rtnn_read_point_sets": (_I, [ctypes.c_char_p, _VP, ctypes.POINTER(_U64),
                                  ctypes.POINTER(_U32)]),
    "rtnn_gen_uniform": (_I, [_U64, _U64, _F, _F, _VP]),
    "rtnn_gen_gaussian_mixture": (_I, [_U64, _U64, _U64, _U32, _F, _F, _F, _VP]),
    "rtnn_gen_spheres": (_I, [_U64, _U64, _U64, _U32, _F, _F, _F, _F, _VP]),
    "rtnn_gen_cloud_sizes": (_I, [_U64, _U32, _U32, _U32, _VP]),
    "rtnn_gen_lattice_clouds": (_I, [_U64, _U32, _VP, _F, _F, _VP]),
    "rtnn_gen_levy": (_I, [_U64, _U64, _F, _F, _F, _VP]),
}


def lib():
    """Loads librtnn_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with rtnn_amd/build.py")
    # torch bundles a HIP runtime of its own under the same soname
    # (libamdhip64.so.7).  Loaded first, it is the one the library binds to, so
    # the process holds one runtime; loaded after the system one, torch gets a
    # second runtime and the two cannot both open the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


# ------------------------------------------------------------------ helpers

def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _ptr(x):
    """(pointer, on_device) of a numpy array or torch tensor (kept alive by caller)."""
    if x is None:
        return None, 0
    if _is_torch(x):
        return ctypes.c_void_p(x.data_ptr()), int(x.is_cuda)
    return ctypes.c_void_p(x.ctypes.data), 0


def _as_points(a):
    if _is_torch(a):
        import torch
        a = a.to(torch.float32).contiguous()
        if a.dim() != 2 or a.shape[1] != 3:
            raise ValueError("points must have shape [n, 3]")
        return a
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("points must have shape [n, 3]")
    return a


def _mode(mode):
    if mode in ("radius", MODE_RADIUS):
        return MODE_RADIUS
    if mode in ("knn", MODE_KNN):
        return MODE_KNN
    raise ValueError("mode must be 'radius' or 'knn'")


class Context:
    """One search context on one GPU (not thread-safe; one per thread)."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.rtnn_create(device, ctypes.byref(h))
        if rc != OK:
            raise RtnnError(rc, L.rtnn_create_error().decode())
        self._h = h
        self._L = L
        self._points = None
        self._graphs = weakref.WeakSet()

    def close(self):
        if getattr(self, "_h", None):
            for g in list(self._graphs):
                g.close()
            self._L.rtnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise RtnnError(rc, self._L.rtnn_last_error(self._h).decode())

    def set_options(self, cell_factor=0.0, queries_per_wave=0, query_order=0,
                    time_stages=0, max_cells=0, grid_levels=0, level_alpha=0.0,
                    group_buffer=0, lds_tile=0, tree_leaf=0, bucket_ratio=0.0, learned_cap=0,
                    tree_lanes=0, tree_fallback=0):
        o = Options(cell_factor, queries_per_wave, query_order, time_stages, max_cells,
                    grid_levels, level_alpha, group_buffer, lds_tile, tree_leaf, bucket_ratio,
                    learned_cap, tree_lanes, tree_fallback)
        self._check(self._L.rtnn_set_options(self._h, ctypes.byref(o)))

    def synchronize(self):
        self._check(self._L.rtnn_synchronize(self._h))

    def set_size_classes(self, upper, options):
        """Batched-search size classes: class j holds clouds of at most upper[j]
        points (the last class every larger one) and uses options[j] (a dict of
        set_options keywords).  Empty lists drop the classes."""
        n = len(options)
        if n and len(upper) != n - 1:
            raise ValueError("need len(options) - 1 class bounds")
        up = np.ascontiguousarray(upper, dtype=np.uint64)
        arr = (Options * max(n, 1))()
        for j, o in enumerate(options):
            arr[j] = _options_struct(**o)
        self._check(self._L.rtnn_set_size_classes(
            self._h, n, ctypes.c_void_p(up.ctypes.data) if len(up) else None, arr))

    def size_classes(self):
        """(upper bounds, list of option dicts) of the current size classes."""
        n = _U32(0)
        self._check(self._L.rtnn_get_size_classes(self._h, ctypes.byref(n), None, None))
        up = np.zeros(max(n.value - 1, 1), dtype=np.uint64)
        arr = (Options * max(n.value, 1))()
        self._check(self._L.rtnn_get_size_classes(self._h, ctypes.byref(n),
                                                  ctypes.c_void_p(up.ctypes.data), arr))
        return ([int(u) for u in up[:max(n.value - 1, 0)]],
                [_options_dict(arr[j]) for j in range(n.value)])

    def set_points(self, points):
        p = _as_points(points)
        ptr, dev = _ptr(p)
        self._check(self._L.rtnn_set_points(self._h, ptr, len(p), dev))
        self._points = p
        return self

    def _outputs(self, like_device, nq, k, with_d2):
        if like_device:
            import torch
            idx = torch.empty((nq, k), dtype=torch.int32, device="cuda")
            d2 = torch.empty((nq, k), dtype=torch.float32, device="cuda") if with_d2 else None
        else:
            idx = np.empty((nq, k), dtype=np.int32)
            d2 = np.empty((nq, k), dtype=np.float32) if with_d2 else None
        return idx, d2

    def _run(self, fn, queries, r, k, with_d2, out_device, extra=(), out=None):
        stats = Stats()
        if queries is None:
            if self._points is None:
                raise RtnnError(ERR_STATE, "set_points() has not been called")
            nq, qptr, qdev = len(self._points), None, 0
            keep = None
        else:
            keep = _as_points(queries)
            qptr, qdev = _ptr(keep)
            nq = len(keep)
        if out is not None:
            # caller-owned outputs (reused across calls): (idx, d2 or None),
            # [nq, k] int32 / float32, both on the device or both on the host
            idx, d2 = out
            dev = _is_torch(idx) and idx.is_cuda
            ok = tuple(idx.shape) == (nq, k) and str(idx.dtype).endswith("int32") and (
                d2 is None or (tuple(d2.shape) == (nq, k) and str(d2.dtype).endswith("float32")
                               and (_is_torch(d2) and d2.is_cuda) == dev))
            contig = lambda t: t is None or (t.is_contiguous() if _is_torch(t)
                                             else t.flags.c_contiguous)
            if not ok or not contig(idx) or not contig(d2):
                raise ValueError(f"out must be contiguous ([{nq}, {k}] int32, [{nq}, {k}] float32 "
                                 "or None) on one device")
            out_device = dev
        else:
            if out_device is None:
                out_device = bool(qdev) if queries is not None else _is_torch(self._points)
            idx, d2 = self._outputs(out_device, nq, k, with_d2)
        ip, _ = _ptr(idx)
        dp, _ = _ptr(d2)
        rc = fn(self._h, qptr, nq, qdev, *extra, r, k, ip, dp, int(out_device),
                ctypes.byref(stats))
        self._check(rc)
        return idx, d2, stats

    def search(self, queries=None, mode="radius", radius=2.0, k=50, with_d2=True,
               out_device=None, out=None):
        """Plain grid search -> (idx [nq, k] int32, d2 [nq, k] float32 or None, Stats).
        out: optional (idx, d2) arrays or tensors to write into (then returned)."""
        return self._run(self._L.rtnn_search, queries, radius, k, with_d2, out_device,
                         extra=(_mode(mode),), out=out)

    def search_async(self, queries, idx, d2=None, mode="radius", radius=2.0, stream=None):
        """Enqueues a plain search of device tensor `queries` (None: the points)
        into device tensors idx [nq, k] (and d2) on `stream` (a torch CUDA
        stream; default: torch's current stream) and returns at once."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        if queries is None:
            qptr, nq = None, len(self._points)
        else:
            q = _as_points(queries)
            if not (_is_torch(q) and q.is_cuda):
                raise ValueError("search_async takes device tensors")
            qptr, nq = ctypes.c_void_p(q.data_ptr()), len(q)
            self._async_keep = q
        k = idx.shape[1]
        self._check(self._L.rtnn_search_async(
            self._h, qptr, nq, _mode(mode), radius, k, ctypes.c_void_p(idx.data_ptr()),
            ctypes.c_void_p(d2.data_ptr()) if d2 is not None else None,
            ctypes.c_void_p(st.cuda_stream)))

    def search_noalloc(self, points, queries=None, mode="radius", radius=2.0, k=50,
                       idx=None, d2=None):
        """rtnn_search_noalloc: points, queries (None: the points) and the
        outputs idx [nq, k] int32 / d2 [nq, k] float32 (None: not written) are
        device tensors; the library allocates nothing."""
        p = _as_points(points)
        q = None if queries is None else _as_points(queries)
        on_dev = lambda t: _is_torch(t) and t.is_cuda
        if not on_dev(p) or (q is not None and not on_dev(q)):
            raise ValueError("search_noalloc takes device tensors")
        nq = len(p) if q is None else len(q)
        if idx is None:
            import torch
            idx = torch.empty((nq, k), dtype=torch.int32, device=p.device)
        self._check(self._L.rtnn_search_noalloc(
            self._h, ctypes.c_void_p(p.data_ptr()), len(p),
            None if q is None else ctypes.c_void_p(q.data_ptr()), 0 if q is None else len(q),
            _mode(mode), radius, k, ctypes.c_void_p(idx.data_ptr()),
            None if d2 is None else ctypes.c_void_p(d2.data_ptr())))
        return idx, d2

    def search_partitioned(self, queries=None, radius=2.0, k=8, with_d2=True,
                           out_device=None, out=None):
        return self._run(self._L.rtnn_search_partitioned, queries, radius, k, with_d2,
                         out_device, out=out)

    def search_learned(self, queries=None, radius=10.0, k=16, with_d2=True, out_device=None,
                       out=None):
        return self._run(self._L.rtnn_search_learned, queries, radius, k, with_d2, out_device,
                         out=out)

    def search_stream(self, queries, mode="knn", method="plain", radius=2.0, k=50, chunk=0,
                      with_d2=True, path=None):
        """Host queries in, host rows out (or a result file at `path`), searched
        chunk by chunk with the copies of one chunk overlapping the search of
        the next -> (idx, d2 or None, Stats); (None, None, Stats) with a path."""
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.float32).reshape(-1, 3))
        meth = {"plain": 0, "partitioned": 1, "learned": 2}[method]
        nq = len(q)
        stats = Stats()
        idx = d2 = None
        if path is None:
            idx = np.empty((nq, k), dtype=np.int32)
            d2 = np.empty((nq, k), dtype=np.float32) if with_d2 else None
        rc = self._L.rtnn_search_stream(
            self._h, q.ctypes.data_as(ctypes.c_void_p), nq, _mode(mode), meth, radius, k, chunk,
            None if idx is None else idx.ctypes.data_as(ctypes.c_void_p),
            None if d2 is None else d2.ctypes.data_as(ctypes.c_void_p),
            None if path is None else os.fsencode(path), ctypes.byref(stats))
        self._check(rc)
        return idx, d2, stats

    def propose_radius(self, queries=None, method="learned", radius=10.0, k=16):
        """First-stage launch radius per query (input order) -> float32 array."""
        meth = {"partitioned": 1, "learned": 2}[method]
        if queries is None:
            nq, qptr, qdev, keep = len(self._points), None, 0, None
        else:
            keep = _as_points(queries)
            qptr, qdev = _ptr(keep)
            nq = len(keep)
        out = np.empty(nq, dtype=np.float32)
        self._check(self._L.rtnn_propose_radius(self._h, qptr, nq, qdev, meth, radius, k,
                                                ctypes.c_void_p(out.ctypes.data), 0))
        return out

    def mlp_train(self, radius, k, n_samples=1 << 21, steps=2000, batch=4096, lr=1e-3,
                  tau=0.9, seed=1, queries=None):
        """Trains the radius MLP on n_samples queries drawn from the points, or
        on `queries` when given; returns the final pinball loss."""
        loss = ctypes.c_float(0)
        if queries is None:
            self._check(self._L.rtnn_mlp_train(self._h, radius, k, n_samples, steps, batch, lr,
                                               tau, seed, ctypes.byref(loss)))
        else:
            q = _as_points(queries)
            qptr, qdev = _ptr(q)
            self._check(self._L.rtnn_mlp_train_queries(self._h, qptr, len(q), qdev, radius, k,
                                                       steps, batch, lr, tau, seed,
                                                       ctypes.byref(loss)))
        return float(loss.value)

    def mlp_weights(self):
        w = np.empty(MLP_NBLOB, dtype=np.float32)
        self._check(self._L.rtnn_mlp_get_weights(self._h, ctypes.c_void_p(w.ctypes.data)))
        return w

    def set_mlp_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != MLP_NBLOB:
            raise ValueError(f"weight blob must have {MLP_NBLOB} floats")
        self._check(self._L.rtnn_mlp_set_weights(self._h, ctypes.c_void_p(w.ctypes.data)))

    def _batched(self, fn, xyz, offsets, mode, radius, k, with_d2):
        import torch
        if not (_is_torch(xyz) and xyz.is_cuda):
            raise ValueError("batched search takes a CUDA tensor of points")
        xyz = xyz.to(torch.float32).contiguous()
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = int(offs[-1])
        idx = torch.empty((n, k), dtype=torch.int32, device=xyz.device)
        d2 = torch.empty((n, k), dtype=torch.float32, device=xyz.device) if with_d2 else None
        stats = Stats()
        rc = fn(self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(offs.ctypes.data),
                len(offs) - 1, _mode(mode), radius, k, ctypes.c_void_p(idx.data_ptr()),
                ctypes.c_void_p(d2.data_ptr()) if d2 is not None else None,
                ctypes.byref(stats))
        self._check(rc)
        return idx, d2, stats

    def search_batched(self, xyz, offsets, mode="radius", radius=0.2, k=64, with_d2=True):
        return self._batched(self._L.rtnn_search_batched, xyz, offsets, mode, radius, k,
                             with_d2)

    def search_batched_stream(self, xyz, offsets, mode="radius", radius=0.2, k=64, chunk=0,
                              with_d2=True, path=None):
        """Host clouds in, host rows out (or a result file at `path`), searched in
        chunks of whole clouds with the copies of one chunk overlapping the
        search of the next -> (idx, d2 or None, Stats); (None, None, Stats)
        with a path."""
        x = np.ascontiguousarray(np.asarray(xyz, dtype=np.float32).reshape(-1, 3))
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = int(offs[-1])
        stats = Stats()
        idx = d2 = None
        if path is None:
            idx = np.empty((n, k), dtype=np.int32)
            d2 = np.empty((n, k), dtype=np.float32) if with_d2 else None
        rc = self._L.rtnn_search_batched_stream(
            self._h, x.ctypes.data_as(ctypes.c_void_p), offs.ctypes.data_as(ctypes.c_void_p),
            len(offs) - 1, _mode(mode), radius, k, chunk,
            None if idx is None else idx.ctypes.data_as(ctypes.c_void_p),
            None if d2 is None else d2.ctypes.data_as(ctypes.c_void_p),
            None if path is None else os.fsencode(path), ctypes.byref(stats))
        self._check(rc)
        return idx, d2, stats

    def search_percloud(self, xyz, offsets, mode="radius", radius=0.2, k=64, with_d2=True):
        return self._batched(self._L.rtnn_search_percloud, xyz, offsets, mode, radius, k,
                             with_d2)

    def graph_batched(self, xyz, offsets, mode="radius", radius=0.2, k=64, with_d2=True):
        """Captures the batched pipeline in a HIP graph -> BatchedGraph (its
        .idx / .d2 tensors are rewritten by every .launch())."""
        import torch
        if not (_is_torch(xyz) and xyz.is_cuda and xyz.dtype == torch.float32
                and xyz.is_contiguous()):
            raise ValueError("graph_batched takes a contiguous float32 CUDA tensor")
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = int(offs[-1])
        idx = torch.empty((n, k), dtype=torch.int32, device=xyz.device)
        d2 = torch.empty((n, k), dtype=torch.float32, device=xyz.device) if with_d2 else None
        h = ctypes.c_void_p()
        self._check(self._L.rtnn_graph_batched(
            self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(offs.ctypes.data),
            len(offs) - 1, _mode(mode), radius, k, ctypes.c_void_p(idx.data_ptr()),
            ctypes.c_void_p(d2.data_ptr()) if d2 is not None else None, ctypes.byref(h)))
        g = BatchedGraph(self, h, xyz, idx, d2)
        self._graphs.add(g)
        return g

    def boundary_pairs(self, queries=None, radius=2.0, tol=1e-6):
        """Pairs (query row, point index) with |d - r| <= tol * r, d in FP64,
        ordered by (row, point) -> uint32 array [count, 2] (host)."""
        if queries is None:
            nq, qptr, qdev, keep = len(self._points), None, 0, None
        else:
            keep = _as_points(queries)
            qptr, qdev = _ptr(keep)
            nq = len(keep)
        cnt = _U64(0)
        self._check(self._L.rtnn_boundary_pairs(self._h, qptr, nq, qdev, radius, tol, None, 0, 0,
                                                ctypes.byref(cnt)))
        out = np.zeros((cnt.value, 2), dtype=np.uint32)
        if cnt.value:
            self._check(self._L.rtnn_boundary_pairs(self._h, qptr, nq, qdev, radius, tol,
                                                    ctypes.c_void_p(out.ctypes.data),
                                                    cnt.value, 0, ctypes.byref(cnt)))
        return out

    def boundary_pairs_batched(self, xyz, offsets, radius=0.2, tol=1e-6):
        """Batched form: rows global, point indices cloud-local."""
        import torch
        if not (_is_torch(xyz) and xyz.is_cuda):
            raise ValueError("batched boundary report takes a CUDA tensor of points")
        xyz = xyz.to(torch.float32).contiguous()
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        cnt = _U64(0)
        args = (self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(offs.ctypes.data),
                len(offs) - 1, radius, tol)
        self._check(self._L.rtnn_boundary_pairs_batched(*args, None, 0, 0, ctypes.byref(cnt)))
        out = torch.zeros((cnt.value, 2), dtype=torch.int32, device=xyz.device)
        if cnt.value:
            self._check(self._L.rtnn_boundary_pairs_batched(
                *args, ctypes.c_void_p(out.data_ptr()), cnt.value, 1, ctypes.byref(cnt)))
        return out.cpu().numpy().view(np.uint32)

    def autotune(self, queries=None, method="plain", mode="knn", radius=10.0, k=16):
        """Times the candidate configurations on this workload and keeps the
        fastest; returns (options dict, ms per search)."""
        import torch
        meth = {"plain": 0, "partitioned": 1, "learned": 2}[method]
        if queries is None:
            nq, qptr, qdev, keep = len(self._points), None, 0, None
        else:
            keep = _as_points(queries)
            qptr, qdev = _ptr(keep)
            nq = len(keep)
        idx = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        d2 = torch.empty((nq, k), dtype=torch.float32, device="cuda")
        best = Options()
        ms = ctypes.c_double(0)
        self._check(self._L.rtnn_autotune(self._h, qptr, nq, qdev, meth, _mode(mode), radius, k,
                                          ctypes.c_void_p(idx.data_ptr()),
                                          ctypes.c_void_p(d2.data_ptr()), ctypes.byref(best),
                                          ctypes.byref(ms)))
        return ({"cell_factor": best.cell_factor, "queries_per_wave": best.queries_per_wave,
                 "grid_levels": best.grid_levels, "level_alpha": best.level_alpha,
                 "group_buffer": best.group_buffer, "lds_tile": best.lds_tile,
                 "tree_leaf": best.tree_leaf}, ms.value)

    def autotune_batched(self, xyz, offsets, mode="radius", radius=0.2, k=64):
        """Times the batched-search configurations on this batch and keeps the
        fastest; returns (options dict, ms per search)."""
        import torch
        if not (_is_torch(xyz) and xyz.is_cuda):
            raise ValueError("batched autotune takes a CUDA tensor of points")
        xyz = xyz.to(torch.float32).contiguous()
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        idx = torch.empty((int(offs[-1]), k), dtype=torch.int32, device=xyz.device)
        best = Options()
        ms = ctypes.c_double(0)
        self._check(self._L.rtnn_autotune_batched(
            self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(offs.ctypes.data),
            len(offs) - 1, _mode(mode), radius, k, ctypes.c_void_p(idx.data_ptr()), None,
            ctypes.byref(best), ctypes.byref(ms)))
        del idx
        return ({"cell_factor": best.cell_factor, "queries_per_wave": best.queries_per_wave,
                 "group_buffer": best.group_buffer, "lds_tile": best.lds_tile}, ms.value)

    def autotune_batched_classes(self, xyz, offsets, upper, mode="radius", radius=0.2, k=64):
        """Tunes every size class (bounds `upper`, as set_size_classes) on the
        sub-batch of its clouds and installs the classes; returns (list of
        option dicts, list of ms per class sub-batch)."""
        import torch
        if not (_is_torch(xyz) and xyz.is_cuda):
            raise ValueError("batched autotune takes a CUDA tensor of points")
        xyz = xyz.to(torch.float32).contiguous()
        offs = np.ascontiguousarray(offsets, dtype=np.uint64)
        up = np.ascontiguousarray(upper, dtype=np.uint64)
        n = len(up) + 1
        idx = torch.empty((int(offs[-1]), k), dtype=torch.int32, device=xyz.device)
        best = (Options * n)()
        ms = (ctypes.c_double * n)()
        self._check(self._L.rtnn_autotune_batched_classes(
            self._h, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_void_p(offs.ctypes.data),
            len(offs) - 1, _mode(mode), radius, k, ctypes.c_void_p(idx.data_ptr()), None, n,
            ctypes.c_void_p(up.ctypes.data) if len(up) else None, best, ms))
        del idx
        return ([_options_dict(best[j]) for j in range(n)], [ms[j] for j in range(n)])

    # MLP building blocks (device tensors)
    def mlp_features(self, queries, radius):
        import torch
        q = _as_points(queries)
        out = torch.empty((len(q), 32), dtype=torch.float32, device="cuda")
        self._check(self._L.rtnn_mlp_features(self._h, ctypes.c_void_p(q.data_ptr()), len(q),
                                              radius, ctypes.c_void_p(out.data_ptr())))
        return out

    def mlp_forward(self, feats):
        import torch
        feats = feats.to(torch.float32).contiguous()
        out = torch.empty(feats.shape[0], dtype=torch.float32, device="cuda")
        self._check(self._L.rtnn_mlp_forward(self._h, ctypes.c_void_p(feats.data_ptr()),
                                             feats.shape[0], ctypes.c_void_p(out.data_ptr())))
        return out

    def mlp_backward(self, feats, target, tau=0.9):
        import torch
        feats = feats.to(torch.float32).contiguous()
        target = target.to(torch.float32).contiguous()
        grads = torch.empty(MLP_NPARAMS, dtype=torch.float32, device="cuda")
        loss = ctypes.c_float(0)
        self._check(self._L.rtnn_mlp_backward(self._h, ctypes.c_void_p(feats.data_ptr()),
                                              ctypes.c_void_p(target.data_ptr()),
                                              feats.shape[0], tau,
                                              ctypes.c_void_p(grads.data_ptr()),
                                              ctypes.byref(loss)))
        return grads, float(loss.value)

    # building blocks (device tensors)
    def radix_sort_pairs(self, keys, vals, bits=32):
        self._check(self._L.rtnn_radix_sort_pairs(self._h, ctypes.c_void_p(keys.data_ptr()),
                                                  ctypes.c_void_p(vals.data_ptr()),
                                                  keys.numel(), bits))

    def exclusive_scan(self, inp, out):
        self._check(self._L.rtnn_exclusive_scan(self._h, ctypes.c_void_p(inp.data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()), inp.numel()))

    def morton3(self, xyz, codes):
        self._check(self._L.rtnn_morton3(self._h, ctypes.c_void_p(xyz.data_ptr()),
                                         ctypes.c_void_p(codes.data_ptr()), codes.numel()))


# --------------------------------------------------------------- data utils

def read_points(path):
    """Point file -> float32 [n, 3] (reference read_pc_data semantics)."""
    L = lib()
    n = ctypes.c_uint64(0)
    rc = L.rtnn_read_points(path.encode(), None, ctypes.byref(n))
    if rc == -2:
        raise FileNotFoundError(path)
    if rc != 0:
        raise ValueError(f"cannot read {path} (code {rc})")
    out = np.empty((n.value, 3), dtype=np.float32)
    rc = L.rtnn_read_points(path.encode(), ctypes.c_void_p(out.ctypes.data), ctypes.byref(n))
    if rc == -1:
        raise ValueError(f"{path}: line {n.value} does not hold three numbers")
    if rc == -4:
        raise ValueError(f"{path} grew while it was read")
    if rc != 0:
        raise ValueError(f"cannot read {path} (code {rc})")
    return out[:n.value]


def mem_stats():
    """Device and pinned bytes the library holds now and at its peak (all
    contexts) -> dict."""
    v = [ctypes.c_uint64(0) for _ in range(4)]
    lib().rtnn_mem_stats(*[ctypes.byref(x) for x in v])
    return {"device_bytes": v[0].value, "device_peak": v[1].value,
            "pinned_bytes": v[2].value, "pinned_peak": v[3].value}


def read_results(path):
    """Reads a result file ("RTNNRES1": u64 nq, u32 k, u32 flags, int32 idx[nq][k],
    then float d2[nq][k] when flags bit 0) -> (idx, d2 or None)."""
    with open(path, "rb") as f:
        if f.read(8) != b"RTNNRES1":
            raise ValueError(f"{path}: not a result file")
        nq, k, flags = struct.unpack("<QII", f.read(16))
        idx = np.fromfile(f, dtype="<i4", count=nq * k).reshape(nq, k)
        d2 = np.fromfile(f, dtype="<f4", count=nq * k).reshape(nq, k) if flags & 1 else None
    return idx, d2


def read_point_sets(path):
    """File of point sets (each line: d numbers = d/3 points, one per set) ->
    float32 [nsets, n, 3] (reference tokenize/read_pc_data(..., int* d))."""
    L = lib()
    n, ns = ctypes.c_uint64(0), ctypes.c_uint32(0)
    rc = L.rtnn_read_point_sets(path.encode(), None, ctypes.byref(n), ctypes.byref(ns))
    if rc == -2:
        raise FileNotFoundError(path)
    if rc == -1:
        raise ValueError(f"{path}: line {n.value} is malformed or has another count of numbers")
    if rc != 0:
        raise ValueError(f"cannot read {path} (code {rc})")
    out = np.empty((ns.value, n.value, 3), dtype=np.float32)
    rc = L.rtnn_read_point_sets(path.encode(), ctypes.c_void_p(out.ctypes.data),
                                ctypes.byref(n), ctypes.byref(ns))
    if rc != 0:
        raise ValueError(f"cannot read {path} (code {rc})")
    return out


def write_points(path, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    rc = lib().rtnn_write_points(path.encode(), ctypes.c_void_p(pts.ctypes.data), len(pts))
    if rc != 0:
        raise OSError(f"cannot write {path}")


def _gen(fn, n, *args):
    out = np.empty((n, 3), dtype=np.float32)
    rc = fn(*args, ctypes.c_void_p(out.ctypes.data))
    if rc != 0:
        raise ValueError("generator failed")
    return out


def gen_uniform(seed, n, lo, hi):
    return _gen(lib().rtnn_gen_uniform, n, seed, n, lo, hi)


def gen_gaussian_mixture(dist_seed, sample_seed, n, ncomp=64, box=1000.0, sigma=5.0,
                         bg_frac=0.2):
    return _gen(lib().rtnn_gen_gaussian_mixture, n, dist_seed, sample_seed, n, ncomp, box,
                sigma, bg_frac)


def gen_spheres(dist_seed, sample_seed, n, nspheres=2000, box=2000.0, rmin=5.0, rmax=50.0,
                noise=0.05):
    return _gen(lib().rtnn_gen_spheres, n, dist_seed, sample_seed, n, nspheres, box, rmin,
                rmax, noise)


def gen_levy(seed, n, alpha=1.5, min_step=0.01, box=1000.0):
    return _gen(lib().rtnn_gen_levy, n, seed, n, alpha, min_step, box)


class BatchedGraph:
    """A captured batched pipeline; holds the point and output tensors it is
    bound to.  launch() replays it (reading xyz's current contents)."""

    def __init__(self, ctx, handle, xyz, idx, d2):
        self._ctx, self._h, self.xyz, self.idx, self.d2 = ctx, handle, xyz, idx, d2

    def launch(self, sync=True):
        st = Stats()
        self._ctx._check(self._ctx._L.rtnn_graph_launch(self._h, 1 if sync else 0,
                                                        ctypes.byref(st) if sync else None))
        return st

    def close(self):
        if self._h and getattr(self._ctx, "_h", None):
            self._ctx._L.rtnn_graph_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gen_lattice_clouds(seed, nclouds, nmin=1000, nmax=200000, h=0.1, jit=0.02):
    L = lib()
    offs = np.empty(nclouds + 1, dtype=np.uint64)
    rc = L.rtnn_gen_cloud_sizes(seed, nclouds, nmin, nmax, ctypes.c_void_p(offs.ctypes.data))
    if rc != 0:
        raise ValueError("bad cloud size range")
    xyz = np.empty((int(offs[-1]), 3), dtype=np.float32)
    L.rtnn_gen_lattice_clouds(seed, nclouds, ctypes.c_void_p(offs.ctypes.data), h, jit,
                              ctypes.c_void_p(xyz.ctypes.data))
    return xyz, offs
