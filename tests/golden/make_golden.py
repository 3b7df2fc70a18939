"""Writes samplepc_expected.json: known answers for the reference's only data
file (src/samplepc.txt, copied here as samplepc.txt).

The expected rows come from an independent FP32 brute force written here in
numpy (not the C oracle): d2 = ((dx*dx + dy*dy) + dz*dz) with every product
and sum rounded to float32, dx = q - p; radius mode keeps d2 < r*r (self
included, geometry.cu:38-106), KNN keeps 0 < d2 < r*r (geometry.cu:157-191,
check.cpp:38); rows are sorted by (d2, index) and padded with -1.

Run: python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def brute(pts, mode, r, k):
    f = np.float32
    r2 = f(r) * f(r)
    rows, dists = [], []
    for q in pts:
        d = q[None, :] - pts                      # float32 subtraction
        sq = d * d                                # float32 products
        d2 = (sq[:, 0] + sq[:, 1]) + sq[:, 2]     # float32 sums, in this order
        keep = d2 < r2 if mode == "radius" else (d2 > 0) & (d2 < r2)
        idx = np.nonzero(keep)[0]
        order = sorted(idx.tolist(), key=lambda i: (float(d2[i]), i))[:k]
        rows.append(order + [-1] * (k - len(order)))
        dists.append([float(d2[i]) for i in order] + [None] * (k - len(order)))
    return rows, dists


def main():
    pts = np.loadtxt(os.path.join(HERE, "samplepc.txt"), delimiter=",").astype(np.float32)
    cases = []
    for mode, r, k in [("radius", 2.0, 50), ("radius", 2.0, 2), ("knn", 2.0, 5),
                       ("knn", 10.0, 5), ("knn", 10.0, 2), ("radius", 10.0, 8),
                       ("knn", 100.0, 8)]:
        rows, dists = brute(pts, mode, r, k)
        cases.append({"mode": mode, "r": r, "k": k, "idx": rows, "d2": dists})
    out = {"source": "src/samplepc.txt of the reference (8 points)",
           "points": pts.tolist(), "cases": cases}
    with open(os.path.join(HERE, "samplepc_expected.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
