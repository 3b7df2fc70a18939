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

  python bench.py [--steps K] [--warmup W] [--workload c3|c1|c2|c4]
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
                    r=2.0, k=50, methods=["plain"],
                    cfg={"workload": "range r=2 K=50: 100k uniform points in [0,50)^3, "
                                     "queries = points",
                         "points": 100_000, "queries": 100_000, "k": 50, "radius": 2.0})
    if name == "c2":
        return dict(pts=rtnn_amd.gen_gaussian_mixture(2, 20, 1_000_000),
                    qs=rtnn_amd.gen_gaussian_mixture(2, 21, 1_000_000), mode="knn", r=20.0,
                    k=8, methods=["plain", "partitioned"],
                    cfg={"workload": "KNN K=8 r=20: 1M Gaussian-mixture points (64 x sigma 5 "
                                     "+ 20% background), 1M queries",
                         "points": 1_000_000, "queries": 1_000_000, "k": 8, "radius": 20.0})
    raise SystemExit(f"unknown workload {name}")


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
    log(f"[bench] {label}: {steps} steps, {ms:.3f} ms/step; " +
        ", ".join(f"{k}={v:.3f}" for k, v in agg.items() if k.startswith("ms_")))
    return ms, agg, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--cpu-sample", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-samples", type=int, default=1 << 21)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (scale factor for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X and the "
              "search needs no collective (queries shard with replicated points); run --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import rtnn_amd

    if args.copy_calibration:
        n = 1 << 28  # 1 GiB of float32
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n}))
        return

    w = workload(args.workload)
    pts, qs = w["pts"], w["qs"]
    k, r, mode = w["k"], w["r"], w["mode"]
    nq = len(qs) if qs is not None else len(pts)
    log(f"[bench] workload {args.workload}: {len(pts)} points, {nq} queries")
    dp = torch.from_numpy(pts).cuda()
    dq = torch.from_numpy(qs).cuda() if qs is not None else None
    ctx = rtnn_amd.Context(0)
    ctx.set_points(dp)

    results = {}
    outputs = {}
    train = None
    for m in w["methods"]:
        if m == "plain":
            fn = lambda: ctx.search(dq, mode=mode, radius=r, k=k, out_device=True)
        elif m == "partitioned":
            fn = lambda: ctx.search_partitioned(dq, radius=r, k=k, out_device=True)
        else:
            t0 = time.perf_counter()
            loss = ctx.mlp_train(r, k, n_samples=args.train_samples, steps=args.train_steps,
                                 batch=4096, lr=1e-3, tau=0.9, seed=1)
            train = {"seconds": round(time.perf_counter() - t0, 3), "final_pinball_loss": loss,
                     "samples": args.train_samples, "steps": args.train_steps, "batch": 4096,
                     "lr": 1e-3, "tau": 0.9}
            log(f"[bench] MLP trained on the GPU in {train['seconds']} s, loss {loss:.5f}")
            fn = lambda: ctx.search_learned(dq, radius=r, k=k, out_device=True)
        ms, agg, out = timed(fn, args.steps, args.warmup, m)
        results[m] = dict(ms_per_step=round(ms, 4), queries_per_s=round(nq / (ms * 1e-3), 1),
                          **{kk: (round(v, 4) if isinstance(v, float) else v)
                             for kk, v in agg.items()})
        outputs[m] = (out[0], out[1])

    # every row of every method equals the plain grid result
    base_i, base_d = outputs[w["methods"][0]]
    for m in w["methods"][1:]:
        same = bool(torch.equal(outputs[m][0], base_i) and torch.equal(outputs[m][1], base_d))
        results[m]["identical_to_plain_all_rows"] = same
        log(f"[bench] {m} identical to plain on all {nq} rows: {same}")

    head = w["methods"][-1]
    # roofline of the dominant search kernel of the headline method
    kms = results[head]["ms_search"]
    if mode == "knn":
        algo = nq * (12 + 12 * len(pts) / nq + 8 * k)
        algo_def = "KNN compulsory bytes per query 12 + 12*N/Q + 8*K"
    else:
        algo = c27_bytes(pts, pts if qs is None else qs, r, k)
        algo_def = "range bytes per query 12 + 12*C27(q) + 4*K"
    achieved = algo / (kms * 1e-3) / 1e9
    kname = "k_topk_thread" if results[head]["lanes_per_query"] == 1 else "k_topk_group"
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": kname,
            "kernel_ms": round(kms, 4), "algorithmic_bytes": algo, "bytes_definition": algo_def}

    # exactness on a sample + CPU baseline (oracle brute force on the sample)
    import oracle
    qarr = pts if qs is None else qs
    om = oracle.MODE_KNN if mode == "knn" else oracle.MODE_RADIUS
    ns = min(args.cpu_sample, nq)
    sel = np.random.default_rng(1234).choice(nq, ns, replace=False).astype(np.uint64)
    cores = min(16, os.cpu_count() or 1)
    cpu, mism = None, None
    if not args.no_cpu:
        t = time.perf_counter()
        wi, wd = oracle.bruteforce(pts, qarr, om, r, k, qsel=sel, threads=cores)
        dt = time.perf_counter() - t
        s = torch.from_numpy(sel.astype(np.int64)).cuda()
        mism = {}
        for m, (oi, od) in outputs.items():
            gi, gd = oi[s].cpu().numpy(), od[s].cpu().numpy()
            mism[m] = int(((gi != wi).any(axis=1) | (gd != wd).any(axis=1)).sum())
        cpu = {"value": round(ns / dt, 3), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"{ns} random queries of this workload, brute force over all "
                         f"{len(pts)} points ({dt:.1f} s)"}
        log(f"[bench] cpu baseline {ns / dt:.1f} q/s on {cores} threads; mismatches {mism}")

    value = results[head]["queries_per_s"]
    cfg = dict(w["cfg"], method=head, step="grid build + query sort + radius proposal + search "
                                          "+ exact fallback" if head == "learned" else
                                          "grid build + query sort + search")
    out = {
        "metric": METRIC if args.workload == "c3" else f"{mode} queries/s ({args.workload})",
        "value": value, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": results[head]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", "config": cfg, "roofline": roof, "cpu_baseline": cpu,
        "methods": results, "sampled_mismatches": mism, "mlp_training": train,
    }
    if "plain" in results and head != "plain":
        out["speedup_vs_plain"] = round(results["plain"]["ms_per_step"] /
                                        results[head]["ms_per_step"], 3)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
