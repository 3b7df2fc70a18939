"""GPU parity of the building blocks behind the cell list: exclusive scan,
stable LSD radix sort and Morton codes, each against its CPU counterpart
(numpy cumsum, numpy stable argsort, the oracle's MortonCode3)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 4097, 100000, 3 * 4096 * 300 + 17])
def test_exclusive_scan(ctx, n):
    import torch
    rng = np.random.default_rng(n)
    x = rng.integers(0, 50, n, dtype=np.uint32)
    out = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    ctx.exclusive_scan(_t(x.view(np.int32)), out)
    got = out.cpu().numpy().view(np.uint32)
    want = np.concatenate([[0], np.cumsum(x, dtype=np.uint64)]).astype(np.uint32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n,bits", [(1, 8), (1000, 8), (4097, 13), (100000, 24),
                                    (1 << 20, 32), (777777, 30), (5000, 9), (70000, 10),
                                    (300001, 17), (300001, 19), (250000, 20), (400000, 25),
                                    (1 << 20, 27), (17_000_000, 20)])
def test_radix_sort_stable(ctx, n, bits):
    rng = np.random.default_rng(bits * 1000 + n)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    # few distinct keys (long equal runs: stability) or the full key range
    top = hi if (n <= 1000 or bits in (9, 10, 17, 19, 20, 25, 27)) else min(hi, 5000)
    keys = rng.integers(0, top, n, dtype=np.uint64).astype(np.uint32)
    if bits < 32:
        keys &= np.uint32((1 << bits) - 1)
    vals = np.arange(n, dtype=np.uint32)
    tk, tv = _t(keys.view(np.int32)), _t(vals.view(np.int32))
    ctx.radix_sort_pairs(tk, tv, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(tk.cpu().numpy().view(np.uint32), keys[order])
    assert np.array_equal(tv.cpu().numpy().view(np.uint32), vals[order])


@pytest.mark.parametrize("bits", [8, 20, 25, 27])
def test_radix_sort_low_bits_only(ctx, bits):
    # Sorting on the low bits must ignore the bits above them (also where the
    # passes' digits together span more bits) and keep ties stable.
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 1 << 32, 50000, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(len(keys), dtype=np.uint32)
    tk, tv = _t(keys.view(np.int32)), _t(vals.view(np.int32))
    ctx.radix_sort_pairs(tk, tv, bits)
    order = np.argsort(keys & np.uint32((1 << bits) - 1), kind="stable")
    assert np.array_equal(tv.cpu().numpy().view(np.uint32), vals[order])
    assert np.array_equal(tk.cpu().numpy().view(np.uint32), keys[order])


def test_morton3_matches_oracle(ctx):
    import torch
    rng = np.random.default_rng(11)
    xyz = rng.integers(0, 1024, (20000, 3), dtype=np.uint32)
    xyz[:4] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1023, 1023, 1023]]
    codes = torch.zeros(len(xyz), dtype=torch.int32, device="cuda")
    ctx.morton3(_t(xyz.view(np.int32)), codes)
    got = codes.cpu().numpy().view(np.uint32)
    want = np.array([oracle.morton3(int(a), int(b), int(c)) for a, b, c in xyz], np.uint32)
    assert np.array_equal(got, want)
