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
    # x sigma(a x) - (-x) sigma(-a x) = x, and the limits x and 0
    xs = np.linspace(-8, 8, 161)
    assert np.allclose(M.gelu(xs) - M.gelu(-xs), xs, atol=1e-12)
    assert np.allclose(M.gelu(np.array([-30.0, 30.0])), [0.0, 30.0], atol=1e-9)
    assert abs(M.gelu(np.array(1.0)) - 1.0 / (1.0 + np.exp(-1.702))) < 1e-12
    fd = (M.gelu(xs + 1e-6) - M.gelu(xs - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(xs), fd, atol=1e-6)


def test_gelu_sigmoid_form_error_bound():
    # Hendrycks & Gimpel's x sigma(1.702 x) stays within 0.021 of the exact
    # x Phi(x) everywhere (the largest gap is near |x| = 2.1); a slope other
    # than 1.702 or a dropped factor x breaks the bound.
    xs = np.linspace(-10, 10, 20001)
    err = np.abs(M.gelu(xs) - M.gelu_erf(xs))
    assert err.max() < 0.021 and err.max() > 0.015
    assert 1.8 < abs(xs[err.argmax()]) < 2.4
    worse = np.abs(xs * M._sigma(1.6 * xs) - M.gelu_erf(xs)).max()
    assert worse > 0.021


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
