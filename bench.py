"""Benchmark of the gfx950 neighbor search; prints one JSON line on stdout.

Default workload (the headline metric): KNN, K = 16, r = 10, 10,000,000
FP32 points on the surfaces of 2,000 spheres (radii U[5,50], centres
U[0,2000]^3, normal noise 0.05) and 10,000,000 separate queries from the same
distribution.  A step is one full search -- grid build + query ordering +
(radius proposal) + search (+ exact fallback) -- over device-resident inputs
into device-resident outputs.

  value      queries/s of the learned-radius search (plain grid beside it)
  roofline   the dominant search kernel's algorithmic bytes per launch over
             its average launch time (HIP events on the library stream)
  cpu_baseline  the CPU oracle's brute force (the reference's check path,
             check.cpp:20-139) on a bounded sample of the same queries

  python bench.py [--steps K] [--warmup W] [--workload c3|c1|c2|c4|c5]

The default run (c3) also measures the range search of 100k points (c1:
plain search and its graph replay) and the batched range search of 4096
clouds (c4: per-cloud launches, one batched launch, graph replay) and
reports them under extra_workloads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "KNN queries/s, 10M pts K=16; learned-radius vs plain grid, exact match"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_rate_gbs(torch):
    """Achievable HBM rate of a plain 1 GiB device copy (bytes read + written
    per second), timed with events; reported beside the 8 TB/s spec peak."""
    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    del src, dst
    return 2 * 4 * n / (ms * 1e-3) / 1e9


def workload(name):
    import rtnn_amd
    if name == "c3":
        return dict(pts=rtnn_amd.gen_spheres(3, 30, 10_000_000),
                    qs=rtnn_amd.gen_spheres(3, 31, 10_000_000), mode="knn", r=10.0, k=16,
                    methods=["plain", "learned"],
                    cfg={"workload": "KNN K=16 r=10: 10M FP32 points on 2000 sphere surfaces, "
                                     "10M queries from the same distribution",
                         "points": 10_000_000, "queries": 10_000_000, "k": 16, "radius": 10.0})
    if name == "c1":
        return dict(pts=rtnn_amd.gen_uniform(1, 100_000, 0.0, 50.0), qs=None, mode="radius",
                    r=2.0, k=50, methods=["plain", "graph"],
                    # the graph method: the pipeline (bbox, device-side grid
                    # plan, keys, sort, cell table, gather, search) of this one
                    # cloud captured once and replayed
                    graph_offsets=np.array([0, 100_000], dtype=np.uint64), graph_d2=True,
                    extra_steps=50,
                    cfg={"workload": "range r=2 K=50: 100k uniform points in [0,50)^3, "
                                     "queries = points",
                         "points": 100_000, "queries": 100_000, "k": 50, "radius": 2.0})
    if name == "c2":
        return dict(pts=rtnn_amd.gen_gaussian_mixture(2, 20, 1_000_000),
                    qs=rtnn_amd.gen_gaussian_mixture(2, 21, 1_000_000), mode="knn", r=20.0,
                    k=8, methods=["plain", "partitioned", "learned"],
                    cfg={"workload": "KNN K=8 r=20: 1M Gaussian-mixture points (64 x sigma 5 "
                                     "+ 20% background), 1M queries",
                         "points": 1_000_000, "queries": 1_000_000, "k": 8, "radius": 20.0})
    if name == "c5":
        return dict(pts=rtnn_amd.gen_levy(5, 30_000_000, alpha=1.5, min_step=0.01, box=1000.0),
                    qs=rtnn_amd.gen_uniform(51, 30_000_000, 0.0, 1000.0), mode="knn", r=100.0,
                    k=32, methods=["partitioned", "learned"], train_on_queries=True,
                    autotune=False,  # each candidate search at r = 100 takes seconds
                    # instead: the learned method's radius margin and density
                    # bucket, timed on 3M of the queries (about 0.1 s each)
                    proposal_tune={"sample": 3_000_000, "margins": [1.0, 1.5, 2.0, 3.0],
                                   "buckets": [0.0, 0.8]},
                    # the octree walk of a 6-level hierarchical grid (fixed
                    # configuration, not tuned): leaves of up to 65536
                    # points, scanned chunk by chunk, two lanes per query
                    options={"grid_levels": 6, "queries_per_wave": 0, "tree_leaf": 65536,
                             "tree_lanes": 2},
                    cfg={"workload": "KNN K=32 r=100: 30M Levy-flight points (Pareto alpha 1.5, "
                                     "min step 0.01, periodic [0,1000]^3), 30M uniform queries",
                         "points": 30_000_000, "queries": 30_000_000, "k": 32, "radius": 100.0})
    if name == "c4":
        xyz, offs = rtnn_amd.gen_lattice_clouds(4, 4096, 1000, 200000, h=0.1, jit=0.02)
        return dict(pts=xyz, qs=None, offsets=offs, mode="radius", r=0.2, k=64,
                    methods=["percloud", "batched", "classes", "graph"],
                    # as an extra leg of the default run: the configuration the
                    # batched autotuner picks for this batch, stored, and no
                    # size classes; 3 timed steps after 1 warm-up
                    extra_methods=["percloud", "batched", "graph"],
                    extra_options={"cell_factor": 1.0, "queries_per_wave": 16,
                                   "group_buffer": 1},
                    extra_steps=3, extra_cpu_sample=200_000,
                    # cloud-size classes of the per-class tuning: <= 6000,
                    # <= 35000 and larger points (thirds of the log range)
                    class_bounds=[6000, 35000],
                    cfg={"workload": "batched range r=0.2 K=64: 4096 jittered cubic lattices "
                                     "(h=0.1, jitter 0.02h), sizes log-uniform in [1e3, 2e5], "
                                     "queries = points",
                         "clouds": 4096, "points": int(offs[-1]), "k": 64, "radius": 0.2})
    raise SystemExit(f"unknown workload {name}")


# CPU baseline sample per workload: about 10-30 s of 16-thread brute force.
CPU_SAMPLE = {"c1": 100_000, "c2": 200_000, "c3": 15_000, "c4": 1_000_000, "c5": 6_000}


def c27_bytes_clouds(xyz, offs, r, k, sample, seed=5):
    """Range roofline bytes for many clouds: the exact C27 (counted inside each
    query's own cloud) of `sample` random queries, its mean scaled to all."""
    rng = np.random.default_rng(seed)
    n = int(offs[-1])
    qi = np.sort(rng.choice(n, min(sample, n), replace=False))
    cloud = np.searchsorted(offs, qi, side="right") - 1
    tot = 0
    for c in np.unique(cloud):
        a, b = int(offs[c]), int(offs[c + 1])
        sel = qi[cloud == c]
        pts = xyz[a:b]
        tot += c27_bytes(pts, xyz[sel], r, k) - len(sel) * (12 + 4 * k)
    c27_mean = tot / 12 / len(qi)
    return float(n * (12 + 4 * k) + 12 * c27_mean * n), c27_mean


def c27_bytes(pts, qs, r, k):
    """Range roofline bytes: sum over queries of 12 + 12*C27(q) + 4k, C27 = points
    in the 3x3x3 block of edge-r cells around the query's cell."""
    mn = pts.min(axis=0)
    cell = np.floor((pts - mn) / r).astype(np.int64)
    dims = cell.max(axis=0) + 3
    lin = lambda c: ((c[:, 2] * dims[1]) + c[:, 1]) * dims[0] + c[:, 0]
    counts = np.bincount(lin(cell + 1), minlength=int(np.prod(dims)))
    qc = np.clip(np.floor((qs - mn) / r).astype(np.int64), -1, dims - 3) + 1
    tot = np.zeros(len(qs), np.int64)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                c = qc + np.array([dx, dy, dz])
                ok = ((c >= 0) & (c < dims)).all(axis=1)
                tot[ok] += counts[lin(c[ok])]
    return float(len(qs) * (12 + 4 * k) + 12 * tot.sum())


def timed(fn, steps, warmup, label):
    import torch
    for i in range(warmup):
        fn()
        log(f"[bench] {label}: warmup {i + 1}/{warmup}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    out = None
    for _ in range(steps):
        out = fn()
        stats.append(out[2])
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    agg = {f: float(np.mean([getattr(s, f) for s in stats]))
           for f in ("ms_total", "ms_grid", "ms_qsort", "ms_search", "ms_mlp", "ms_fallback")}
    agg["fallback_queries"] = int(stats[-1].fallback_queries)
    agg["lanes_per_query"] = int(stats[-1].lanes_per_query)
    agg["cell_size"] = float(stats[-1].cell_size)
    agg["sort_passes"] = int(stats[-1].sort_passes)
    log(f"[bench] {label}: {steps} steps, {ms:.3f} ms/step; " +
        ", ".join(f"{k}={v:.3f}" for k, v in agg.items() if k.startswith("ms_")))
    return ms, agg, out


PMC_PROFILE = os.path.join(ROOT, "profiles", "counters_roofline.md")
CLOCK_GHZ = 2.4     # MI355X peak engine clock
SIMDS = 256 * 4     # 256 CUs x 4 SIMDs
VALU_CYCLES = 4     # a wave64 VALU instruction occupies a 16-lane SIMD 4 cycles


def pmc_row(workload_name, kernel):
    """Per-launch counters of `kernel` on this workload from the committed
    rocprofv3 summary profiles/counters_roofline.md: | workload | kernel |
    dispatches | HBM bytes | SQ_WAVES | SQ_INSTS_VALU | ... (HBM bytes =
    FETCH_SIZE and WRITE_SIZE scaled as that file states).  None if absent."""
    if not os.path.exists(PMC_PROFILE):
        return None
    for line in open(PMC_PROFILE):
        cols = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cols) >= 6 and cols[0] == workload_name and cols[1] == kernel:
            try:
                return {"hbm_bytes": float(cols[3]), "waves": float(cols[4]),
                        "valu_insts": float(cols[5])}
            except ValueError:
                return None
    return None


def proposal_tune(ctx, dq, r, k, spec, base_opts):
    """Times the learned search on a sample of the queries for each radius
    margin (times the trained one) and density-bucket ratio, one log line per
    candidate, and keeps the fastest (options and weights in the context)."""
    import torch
    t0 = time.perf_counter()
    sub = dq[:spec["sample"]]
    w0 = ctx.mlp_weights().copy()
    best = None
    for f in spec["margins"]:
        w = w0.copy()
        w[-1] = w0[-1] * f
        ctx.set_mlp_weights(w)
        for b in spec["buckets"]:
            ctx.set_options(**base_opts, bucket_ratio=b)
            ctx.search_learned(sub, radius=r, k=k, out_device=True)
            ms = min(ctx.search_learned(sub, radius=r, k=k, out_device=True)[2].ms_total
                     for _ in range(2))
            log(f"[bench] proposal tune: margin x{f} bucket {b}: {ms:.2f} ms on "
                f"{len(sub)} queries")
            if best is None or ms < best[0]:
                best = (ms, f, b)
    torch.cuda.empty_cache()
    ms, f, b = best
    w = w0.copy()
    w[-1] = w0[-1] * f
    ctx.set_mlp_weights(w)
    ctx.set_options(**base_opts, bucket_ratio=b)
    return {**base_opts, "radius_margin": round(float(w[-1]), 4), "margin_factor": f,
            "bucket_ratio": b,
            "sample_queries": len(sub), "sample_ms": round(ms, 3),
            "seconds": round(time.perf_counter() - t0, 3)}


def kernel_name(res, mode, opts):
    """Name of the dominant search kernel of a method's configuration."""
    o = opts or {}
    if res.get("lanes_per_query") == 1 and not o.get("group_buffer") == 3:
        if mode == "knn" and o.get("grid_levels", 0) >= 2:
            return "k_topk_treeg" if o.get("tree_lanes", 0) in (2, 4, 8) else "k_topk_tree"
        return "k_topk_thread"
    if o.get("group_buffer") == 3:
        return "k_topk_pool"
    if o.get("lds_tile"):
        return "k_topk_tiled"
    return "k_topk_group"


def measure(args, wname, torch, rtnn_amd, extra=False):
    """Runs every method of workload `wname`; returns the JSON line's dict."""
    fixed_opts = json.loads(args.options) if args.options else None
    autotune = args.autotune and fixed_opts is None
    w = workload(wname)
    if extra and w.get("extra_options") is not None:
        fixed_opts, autotune = w["extra_options"], False
    if args.methods and not extra:
        want = args.methods.split(",")
        bad = [m for m in want if m not in w["methods"]]
        if bad:
            raise SystemExit(f"workload {wname} has no method {bad}; "
                             f"it has {w['methods']}")
        w["methods"] = want
    if extra and w.get("extra_methods"):
        w["methods"] = w["extra_methods"]
    steps, warmup = (w.get("extra_steps", args.steps), 1) if extra else (args.steps, args.warmup)
    cpu_sample = (w.get("extra_cpu_sample") if extra else args.cpu_sample) or \
        CPU_SAMPLE.get(wname, 4000)
    pts, qs = w["pts"], w["qs"]
    k, r, mode = w["k"], w["r"], w["mode"]
    nq = len(qs) if qs is not None else len(pts)
    offs_all = w.get("offsets")
    goffs = offs_all if offs_all is not None else w.get("graph_offsets")
    log(f"[bench] workload {wname}: {len(pts)} points, {nq} queries")
    dp = torch.from_numpy(pts).cuda()
    dq = torch.from_numpy(qs).cuda() if qs is not None else None
    ctx = rtnn_amd.Context(0)
    if offs_all is None:
        ctx.set_points(dp)

    results, outputs, opts_used = {}, {}, {}
    train = None
    batched_opts = None
    graph = None
    # Output rows of the single-cloud methods: device tensors allocated once
    # per method and passed to every step (the C ABI writes into caller
    # buffers; the Python binding would otherwise allocate a pair per call).
    rows = len(pts) if qs is None else nq

    def outs():
        return (torch.empty((rows, k), dtype=torch.int32, device="cuda"),
                torch.empty((rows, k), dtype=torch.float32, device="cuda"))

    for m in w["methods"]:
        if m in ("plain", "partitioned", "learned"):
            ob = outs()
        if m == "plain":
            fn = lambda: ctx.search(dq, mode=mode, radius=r, k=k, out=ob)
        elif m == "partitioned":
            fn = lambda: ctx.search_partitioned(dq, radius=r, k=k, out=ob)
        elif m == "percloud":
            fn = lambda: ctx.search_percloud(dp, offs_all, mode=mode, radius=r, k=k,
                                             with_d2=False)
        elif m in ("batched", "classes"):
            fn = lambda: ctx.search_batched(dp, offs_all, mode=mode, radius=r, k=k,
                                            with_d2=False)
        elif m == "graph":  # the pipeline captured once (below), replayed per step
            fn = lambda: (graph.idx, graph.d2, graph.launch())
        else:
            t0 = time.perf_counter()
            # uniform queries over clustered points: train on a sample of the queries
            tq = dq[:args.train_samples] if w.get("train_on_queries") else None
            loss = ctx.mlp_train(r, k, n_samples=args.train_samples, steps=args.train_steps,
                                 batch=4096, lr=1e-3, tau=0.9, seed=1, queries=tq)
            train = {"seconds": round(time.perf_counter() - t0, 3), "final_pinball_loss": loss,
                     "samples": args.train_samples, "steps": args.train_steps, "batch": 4096,
                     "lr": 1e-3, "tau": 0.9}
            log(f"[bench] MLP trained on the GPU in {train['seconds']} s, loss {loss:.5f}")
            fn = lambda: ctx.search_learned(dq, radius=r, k=k, out=ob)
        ctx.set_options(**(fixed_opts or w.get("options", {})))
        tuned = None
        if m == "classes":
            bounds = w["class_bounds"]
            if autotune:
                t0 = time.perf_counter()
                copts, cms = ctx.autotune_batched_classes(dp, offs_all, bounds, mode=mode,
                                                          radius=r, k=k)
                keep = ("cell_factor", "queries_per_wave", "group_buffer", "lds_tile")
                tuned = {"class_bounds": bounds,
                         "classes": [{kk: o[kk] for kk in keep} for o in copts],
                         "class_search_ms": [round(v, 4) for v in cms],
                         "seconds": round(time.perf_counter() - t0, 3)}
                log(f"[bench] classes: autotuned {tuned}")
            else:
                o = fixed_opts or w.get("options", {})
                ctx.set_size_classes(bounds, [o] * (len(bounds) + 1))
        if autotune and m in ("batched", "graph"):
            if batched_opts is None:
                t0 = time.perf_counter()
                opt, tms = ctx.autotune_batched(dp, goffs, mode=mode, radius=r, k=k)
                batched_opts = (opt, round(tms, 4), round(time.perf_counter() - t0, 3))
                log(f"[bench] batched: autotuned {batched_opts}")
            ctx.set_options(**batched_opts[0])
            tuned = dict(batched_opts[0], search_ms=batched_opts[1], seconds=batched_opts[2])
        graph_cfg = None
        if m == "graph":  # capture with the options in force (tuned or default)
            # with the size classes when they timed faster than one configuration
            use_cls = ("classes" in results and "batched" in results and
                       results["classes"]["ms_per_step"] < results["batched"]["ms_per_step"])
            if not use_cls:
                ctx.set_size_classes([], [])
            graph_cfg = "size classes" if use_cls else "one configuration"
            log(f"[bench] graph: captured with {graph_cfg}")
            graph = ctx.graph_batched(dp, goffs, mode=mode, radius=r, k=k,
                                      with_d2=bool(w.get("graph_d2")))
        if m == "learned" and w.get("proposal_tune") and fixed_opts is None:
            tuned = proposal_tune(ctx, dq, r, k, w["proposal_tune"], w.get("options", {}))
        if autotune and w.get("autotune", True) and m in ("plain", "partitioned", "learned"):
            t0 = time.perf_counter()
            opt, tms = ctx.autotune(dq, method=m, mode=mode, radius=r, k=k)
            tuned = dict(opt, search_ms=round(tms, 4), seconds=round(time.perf_counter() - t0, 3))
            if m == "learned":
                tuned["radius_margin"] = round(float(ctx.mlp_weights()[-1]), 4)
            log(f"[bench] {m}: autotuned {tuned}")
        ms, agg, out = timed(fn, steps, warmup, f"{wname} {m}")
        results[m] = dict(ms_per_step=round(ms, 4), queries_per_s=round(nq / (ms * 1e-3), 1),
                          **{kk: (round(v, 4) if isinstance(v, float) else v)
                             for kk, v in agg.items()})
        if tuned:
            results[m]["autotuned"] = tuned
            opts_used[m] = tuned
        elif fixed_opts or w.get("options"):
            results[m]["options"] = dict(fixed_opts or w.get("options"))
            opts_used[m] = results[m]["options"]
        if graph_cfg:
            results[m]["captured_with"] = graph_cfg
        outputs[m] = (out[0], out[1])

    # The same head method with host-resident queries and outputs (the CLI's
    # path): PCIe copies included.  Reported beside `value`, never as it.
    host_leg = None
    head_m = w["methods"][-1]
    if not extra and offs_all is None and head_m in ("plain", "partitioned", "learned"):
        qh = qs if qs is not None else None
        call = {"plain": lambda: ctx.search(qh, mode=mode, radius=r, k=k, out_device=False),
                "partitioned": lambda: ctx.search_partitioned(qh, radius=r, k=k, out_device=False),
                "learned": lambda: ctx.search_learned(qh, radius=r, k=k, out_device=False)}[head_m]
        if qh is None:
            ctx.set_points(pts)  # host points: queries = points from the host
        call()
        t0 = time.perf_counter()
        for _ in range(2):
            call()
        hms = (time.perf_counter() - t0) * 1e3 / 2
        host_leg = {"method": head_m, "ms_per_step": round(hms, 3),
                    "queries_per_s": round(nq / (hms * 1e-3), 1),
                    "includes": "query upload and idx + d2 download over PCIe"}
        log(f"[bench] {head_m} with host buffers: {hms:.3f} ms/step")
        if qh is not None:
            # the same through the streamed entry point: chunks of 2M queries,
            # one chunk's rows crossing PCIe while the next is searched
            ch = 1 << 21
            sargs = dict(mode=mode, method=head_m, radius=r, k=k, chunk=ch)
            ctx.search_stream(qh, **sargs)
            t0 = time.perf_counter()
            for _ in range(2):
                ctx.search_stream(qh, **sargs)
            sms = (time.perf_counter() - t0) * 1e3 / 2
            host_leg["streamed"] = {"ms_per_step": round(sms, 3),
                                    "queries_per_s": round(nq / (sms * 1e-3), 1),
                                    "chunk_queries": ch}
            log(f"[bench] {head_m} streamed from/to host buffers: {sms:.3f} ms/step")
        if qh is None:
            ctx.set_points(dp)

    # every row of every method equals the first (plain / per-cloud) result
    base_i, base_d = outputs[w["methods"][0]]
    for m in w["methods"][1:]:
        oi, od = outputs[m]
        same = bool(torch.equal(oi, base_i) and
                    (base_d is None or od is None or torch.equal(od, base_d)))
        results[m]["identical_to_" + w["methods"][0] + "_all_rows"] = same
        log(f"[bench] {m} identical to {w['methods'][0]} on all {nq} rows: {same}")

    head = w["methods"][-1]
    # roofline of the dominant search kernel of the headline method (a graph
    # replay reports only its total: its search kernel is the batched one, or
    # for one cloud the plain one)
    kroof = head
    if head == "graph":
        kroof = "batched" if "batched" in results else w["methods"][0]
    kms = results[kroof]["ms_search"]
    c27_note = None
    if mode == "knn":
        algo = nq * (12 + 12 * len(pts) / nq + 8 * k)
        algo_def = "KNN compulsory bytes per query 12 + 12*N/Q + 8*K"
    elif offs_all is not None:
        algo, c27m = c27_bytes_clouds(pts, offs_all, r, k, sample=2_000_000)
        algo_def = "range bytes per query 12 + 12*C27(q) + 4*K"
        c27_note = f"C27 mean {c27m:.2f} from 2M sampled queries, scaled to all queries"
    else:
        algo = c27_bytes(pts, pts if qs is None else qs, r, k)
        algo_def = "range bytes per query 12 + 12*C27(q) + 4*K"
    achieved = algo / (kms * 1e-3) / 1e9
    kname = kernel_name(results[kroof], mode, opts_used.get(kroof))
    pmc = pmc_row(wname, kname)
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": pmc["hbm_bytes"] if pmc else None, "kernel": kname,
            "kernel_ms": round(kms, 4), "algorithmic_bytes": algo, "bytes_definition": algo_def}
    if pmc:
        # What the counters say about the same launch at this run's time: the
        # HBM bytes it moved (FETCH_SIZE / WRITE_SIZE, calibrated) and the
        # share of the chip's VALU issue slots its instructions took.
        roof["measured_hbm_gbs"] = round(pmc["hbm_bytes"] / (kms * 1e-3) / 1e9, 2)
        roof["measured_hbm_frac"] = round(roof["measured_hbm_gbs"] / HBM_PEAK_GBS, 5)
        roof["valu_issue_frac"] = round(pmc["valu_insts"] * VALU_CYCLES /
                                        (kms * 1e-3 * SIMDS * CLOCK_GHZ * 1e9), 4)
        roof["counters"] = os.path.relpath(PMC_PROFILE, ROOT)
    if c27_note:
        roof["c27"] = c27_note
    if not extra:
        copy_gbs = copy_rate_gbs(torch)
        roof["copy_rate_gbs"] = round(copy_gbs, 1)
        roof["frac_of_copy_rate"] = round(achieved / copy_gbs, 5)
        log(f"[bench] device copy rate {copy_gbs:.0f} GB/s (1 GiB copy)")

    # exactness on a sample + CPU baseline (oracle brute force on the sample)
    import oracle
    om = oracle.MODE_KNN if mode == "knn" else oracle.MODE_RADIUS
    # Host cores: every core this process may run on, up to the 16 a one-GPU
    # box shares out to a GPU's process (nproc there counts the whole host).
    nproc = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    cores = max(1, min(16, avail))
    cpu, mism = None, None
    if not args.no_cpu:
        rng = np.random.default_rng(1234)
        mism = {m: 0 for m in outputs}
        if offs_all is not None:  # brute force inside sampled clouds
            offs = offs_all
            clouds = rng.choice(len(offs) - 1, 40, replace=False)
            per = max(1, cpu_sample // len(clouds))
            t = time.perf_counter()
            ns = 0
            cloud_sel = []
            for c in clouds:
                a, b = int(offs[c]), int(offs[c + 1])
                sel = rng.choice(b - a, min(per, b - a), replace=False).astype(np.uint64)
                cloud_sel.append((a, b, sel))
                wi, wd = oracle.bruteforce(pts[a:b], pts[a:b], om, r, k, qsel=sel, threads=cores)
                rows = torch.from_numpy(sel.astype(np.int64) + a).cuda()
                for m, (oi, od) in outputs.items():
                    mism[m] += int((oi[rows].cpu().numpy() != wi).any(axis=1).sum())
                ns += len(sel)
            dt = time.perf_counter() - t
            sample = (f"{ns} random queries from 40 random clouds, brute force over their "
                      f"cloud ({dt:.1f} s)")
        else:
            qarr = pts if qs is None else qs
            ns = min(cpu_sample, nq)
            sel = rng.choice(nq, ns, replace=False).astype(np.uint64)
            t = time.perf_counter()
            wi, wd = oracle.bruteforce(pts, qarr, om, r, k, qsel=sel, threads=cores)
            dt = time.perf_counter() - t
            rows = torch.from_numpy(sel.astype(np.int64)).cuda()
            for m, (oi, od) in outputs.items():
                bad = (oi[rows].cpu().numpy() != wi).any(axis=1)
                if od is not None:
                    bad |= (od[rows].cpu().numpy() != wd).any(axis=1)
                mism[m] = int(bad.sum())
            sample = (f"{ns} random queries of this workload, brute force over all "
                      f"{len(pts)} points ({dt:.1f} s)")
        cpu = {"value": round(ns / dt, 3), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": sample, "host_nproc": nproc, "host_cpus_available": avail}
        if not extra:
            # the reference's check path runs on one thread: the same brute
            # force on one core over a subsample of about the same CPU time
            t = time.perf_counter()
            if offs_all is None:
                ns1 = max(1, ns // cores)
                oracle.bruteforce(pts, pts if qs is None else qs, om, r, k, qsel=sel[:ns1],
                                  threads=1)
            else:  # the same clouds, 1/cores of each cloud's sampled queries
                ns1 = 0
                for a, b, csel in cloud_sel:
                    sub = csel[:max(1, len(csel) // cores)]
                    oracle.bruteforce(pts[a:b], pts[a:b], om, r, k, qsel=sub, threads=1)
                    ns1 += len(sub)
            dt1 = time.perf_counter() - t
            cpu["single_thread"] = {"value": round(ns1 / dt1, 3), "unit": "queries/s",
                                    "cores": 1, "sample": f"{ns1} of the same queries "
                                                          f"({dt1:.1f} s)"}
            log(f"[bench] cpu baseline, one thread: {ns1 / dt1:.1f} q/s")
        log(f"[bench] cpu baseline {ns / dt:.1f} q/s on {cores} threads; mismatches {mism}")

    value = results[head]["queries_per_s"]
    steps_desc = {"learned": "grid build + query sort + radius proposal + search + exact fallback",
                  "partitioned": "grid build + query sort + density radius + search + exact "
                                 "fallback",
                  "batched": "per-cloud bbox + one grid build for all clouds + one persistent "
                             "search launch",
                  "classes": "per-cloud bbox + one grid build for all clouds + one "
                             "persistent search launch per cloud-size class",
                  "graph": "one HIP graph replay: bbox + device-side grid plan + cell keys + "
                           "radix sort + cell table + gather + one persistent search launch",
                  "plain": "grid build + query sort + search"}
    cfg = dict(w["cfg"], method=head, step=steps_desc.get(head, head))
    out = {
        "metric": METRIC if wname == "c3" else f"{mode} queries/s ({wname})",
        "value": value, "unit": "queries/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": results[head]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": cfg, "roofline": roof, "cpu_baseline": cpu,
        "methods": results, "sampled_mismatches": mism,
    }
    if train:
        out["mlp_training"] = train
    if host_leg:
        out["host_buffers"] = host_leg
    # grid build of the headline method: sorted keys/s (points sorted into the
    # cell grid; for the batched methods all clouds' points in one sort)
    gms = results[kroof]["ms_grid"]
    if gms > 0:
        out["grid_build"] = {"keys": len(pts), "ms": round(gms, 4),
                             "keys_per_s": round(len(pts) / (gms * 1e-3), 1),
                             "radix_passes": results[kroof]["sort_passes"],
                             "includes": "bbox, cell keys, radix sort, cell table, gather"
                                         + (", query ordering" if offs_all is None else "")}
    base = w["methods"][0]
    if head != base:
        out["speedup_vs_" + base] = round(results[base]["ms_per_step"] /
                                          results[head]["ms_per_step"], 3)
    if "batched" in results and "percloud" in results:
        out["speedup_batched_vs_percloud"] = round(results["percloud"]["ms_per_step"] /
                                                   results["batched"]["ms_per_step"], 3)
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="queries brute-forced on the CPU (default: about 10-30 s of work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-samples", type=int, default=1 << 21)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time the default configuration instead of the autotuned one")
    ap.add_argument("--options", default=None,
                    help="JSON of rtnn_options fields to time instead of autotuning, "
                         "e.g. '{\"queries_per_wave\": 16, \"group_buffer\": 1}'")
    ap.add_argument("--methods", default=None,
                    help="comma-separated subset of the workload's methods to run, in order")
    ap.add_argument("--no-extras", action="store_true",
                    help="the headline workload only (by default the range search of 100k "
                         "points and the 4096-cloud batched range search run after it, "
                         "reported under extra_workloads)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the counter calibration kernels (a 1 GiB device copy and "
                         "a 1 GiB gather of 128-byte blocks in random order)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (the search needs no collective: queries shard over "
              "GPUs with the points replicated on each); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    # one stderr line per autotuner candidate (a long sweep shows progress)
    os.environ.setdefault("RTNN_AUTOTUNE_LOG", "1")
    import torch
    import rtnn_amd

    if args.copy_calibration:
        n = 1 << 28  # 1 GiB of float32
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        # gather: 2^23 blocks of 128 bytes (one cache line each), in a random
        # order: 1 GiB read once line by line, 64 MiB of int64 block indices
        # read, 1 GiB written
        blocks = src.view(-1, 32)
        perm = torch.randperm(blocks.shape[0], device="cuda")
        gdst = torch.empty_like(blocks)
        for _ in range(3):
            torch.index_select(blocks, 0, perm, out=gdst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n,
                          "gather_bytes_read": 4 * n + 8 * blocks.shape[0],
                          "gather_bytes_written": 4 * n}))
        return

    if args.cpu_sample is None:
        args.cpu_sample = CPU_SAMPLE.get(args.workload, 4000)
    out = measure(args, args.workload, torch, rtnn_amd)
    if args.workload == "c3" and not args.no_extras and args.methods is None:
        # The other two targets measured in the same run: the range search
        # at >= 50 % of its roofline (100k points) and the batched range
        # search >= 3x per-cloud launches (4096 clouds), with stored
        # configurations so the run stays short.
        extras = {}
        for wn in ("c1", "c4"):
            e = measure(args, wn, torch, rtnn_amd, extra=True)
            keep = ("metric", "value", "unit", "ms_per_step", "config", "roofline",
                    "cpu_baseline", "methods", "sampled_mismatches", "grid_build",
                    "speedup_vs_plain", "speedup_vs_percloud", "speedup_batched_vs_percloud",
                    "steps", "warmup")
            extras[wn] = {kk: e[kk] for kk in keep if kk in e}
            torch.cuda.empty_cache()
        out["extra_workloads"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
