"""Pins the numpy MLP reference (oracle/mlp_ref.py) before the GPU kernels are
checked against it: BF16 rounding known answers, the analytic backward against
central finite differences, the straight-through BF16 gradient against the
unrounded one, and the features against a brute-force per-cell count."""
import numpy as np

from oracle import mlp_ref as M


def rand_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, M.NP).astype(np.float32)
    p[M.OB3] = 0.2
    return p


def test_bf16_known_answers():
    v = M.bf16(np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 1 + 2 ** -8 + 2 ** -20],
                        np.float32))
    assert v[0] == 1.0
    assert v[1] == 1.0                      # tie -> even (down)
    assert v[2] == 1 + 2 ** -6              # tie -> even (up)
    assert v[3] == -2.5
    assert v[4] == 1 + 2 ** -7              # above the tie -> up


def test_gelu_closed_forms():
    assert M.gelu(np.array(0.0)) == 0.0
    # x t(x) - (-x) t(-x) = x since t(-x) = 1 - t(x); exactly x and 0 beyond
    # |x| = 1 / (2A) = 3 / 1.702; t(1) = A + 1/2
    xs = np.linspace(-8, 8, 161)
    assert np.allclose(M.gelu(xs) - M.gelu(-xs), xs, atol=1e-12)
    edge = 0.5 / M.A
    assert abs(edge - 3.0 / 1.702) < 1e-12
    assert np.array_equal(M.gelu(np.array([edge, 3.0, 30.0])), [edge, 3.0, 30.0])
    assert np.array_equal(M.gelu(np.array([-edge, -3.0, -30.0])), [0.0, 0.0, 0.0])
    assert abs(M.gelu(np.array(1.0)) - (1.702 / 6 + 0.5)) < 1e-15
    # the derivative away from the two kinks at +-edge
    xs = xs[np.abs(np.abs(xs) - edge) > 1e-3]
    fd = (M.gelu(xs + 1e-7) - M.gelu(xs - 1e-7)) / 2e-7
    assert np.allclose(M.gelu_grad(xs), fd, atol=1e-6)
    assert M.gelu_grad(np.array(edge - 1e-9)) > 1.49  # 2 A x + 1/2 -> 3/2 at the kink
    assert M.gelu_grad(np.array(edge + 1e-9)) == 1.0


def test_gelu_hard_sigmoid_error_bound():
    # x clamp(1.702 x / 6 + 1/2, 0, 1) stays within 0.069 of the exact
    # x Phi(x) everywhere (the largest gap is at the kinks, |x| = 3 / 1.702);
    # the sigmoid form it replaces is within 0.021; another slope (1/4 or
    # 1/3) or a dropped factor x breaks the bound.
    xs = np.linspace(-10, 10, 20001)
    err = np.abs(M.gelu(xs) - M.gelu_erf(xs))
    assert 0.06 < err.max() < 0.069
    assert 1.7 < abs(xs[err.argmax()]) < 1.82
    sig = np.abs(xs * 0.5 * (1 + np.tanh(0.5 * 1.702 * xs)) - M.gelu_erf(xs)).max()
    assert 0.015 < sig < 0.021
    for a in (0.25, 1.0 / 3.0):
        assert np.abs(M.gelu(xs, a) - M.gelu_erf(xs)).max() > 0.069
    assert np.abs(M._t(xs) - M.gelu_erf(xs)).max() > 0.069


def test_analytic_gradient_matches_finite_differences():
    p = rand_params(1)
    rng = np.random.default_rng(2)
    feats = rng.uniform(0, 2, (64, M.FP)).astype(np.float32)
    feats[:, 30:] = 0
    o = M.forward(p, feats)
    t = o + np.where(rng.random(64) < 0.5, -0.5, 0.5)  # away from the pinball kink
    g = M.grad_exact(p, feats, t, 0.9)
    idx = list(rng.choice(M.NP, 60, replace=False)) + [M.OB3, M.OW3, M.OB1, M.OW1]
    fd = M.grad_fd(p, feats, t, 0.9, idx)
    assert np.allclose(g[idx], fd, rtol=1e-4, atol=1e-7)


def test_bf16_gradient_close_to_exact():
    p = M.bf16(rand_params(3)).astype(np.float32)  # weights exactly representable
    rng = np.random.default_rng(4)
    feats = rng.uniform(0, 2, (256, M.FP)).astype(np.float32)
    o = M.forward(p, feats)
    t = o + np.where(rng.random(256) < 0.5, -0.5, 0.5)
    a, b = M.grad(p, feats, t, 0.9), M.grad_exact(p, feats, t, 0.9)
    assert np.linalg.norm(a - b) < 3e-2 * np.linalg.norm(b)


def test_features_against_bruteforce_counts():
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 30, (3000, 3)).astype(np.float32)
    qs = np.concatenate([pts[:50], rng.uniform(-15, 45, (50, 3)).astype(np.float32)])
    r = 4.0
    f = M.features(pts, qs, r, 0.25)
    mn = pts.min(axis=0)
    pc = np.floor((pts - mn) / np.float32(r)).astype(int)
    for i, q in enumerate(qs):
        qc = np.floor((q - mn) / np.float32(r)).astype(int)
        for k in range(27):
            d = np.array([k % 3 - 1, (k // 3) % 3 - 1, k // 9 - 1])
            cnt = int((pc == qc + d).all(axis=1).sum())
            assert abs(f[i, k] - np.log1p(cnt) * 0.25) < 1e-9
        assert (0 <= f[i, 27:30]).all() and (f[i, 27:30] <= 1).all()
