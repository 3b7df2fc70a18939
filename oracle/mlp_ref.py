"""Reference of the radius-proposal MLP -- TEST INFRASTRUCTURE ONLY.

The MLP is not part of the reference repository; it is specified by the task:
30 -> 64 -> 64 -> 1, GELU, BF16 weights with FP32 accumulation, pinball loss.
This module restates the exact network the GPU computes, in float64 numpy:
  * weights and the inputs of the two MFMA layers rounded to BF16 (round to
    nearest even); the last layer (64 -> 1, VALU on the GPU) takes the FP32
    activations;
  * GELU in the sigmoid form of Hendrycks & Gimpel (2016, section 2),
    x sigma(1.702 x), with the logistic replaced by the hard sigmoid of
    Howard et al.'s hard-swish (2019, section 5.2): x clamp(A x + 1/2, 0, 1),
    A = 1.702 / 6 (`gelu_erf` is the exact x Phi(x), for its error bound);
  * features: log1p(count) * count_scale of the 27 edge-r cells around the
    query's cell ((dx,dy,dz) = (f%3-1, f//3%3-1, f//9-1)), then the query's
    offset inside its cell;
  * loss: mean pinball loss at quantile tau;
  * gradients: analytic backward with straight-through BF16 rounding.
`grad_fd` differentiates `loss` numerically, which pins `grad` independently.
"""
import numpy as np

H, F, FP = 64, 30, 32
OW1, OB1, OW2, OB2, OW3, OB3, NP = 0, 1920, 1984, 6080, 6144, 6208, 6209
A = 1.702 / 6.0  # slope of the ramp


def bf16(x):
    """Round float32 values to BF16 (RNE), returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _t(x, a=A):
    return np.clip(a * np.asarray(x, np.float64) + 0.5, 0.0, 1.0)


def gelu(x, a=A):
    return np.asarray(x, np.float64) * _t(x, a)


def gelu_grad(x):
    x = np.asarray(x, np.float64)
    h = A * x + 0.5
    return np.where(h <= 0.0, 0.0, np.where(h >= 1.0, 1.0, 2.0 * A * x + 0.5))


def gelu_erf(x):
    """The exact GELU, x Phi(x) = 0.5 x (1 + erf(x / sqrt 2))."""
    from math import erf, sqrt
    x = np.asarray(x, np.float64)
    return 0.5 * x * (1.0 + np.vectorize(lambda v: erf(v / sqrt(2.0)))(x))


def unpack(p, dtype=np.float32):
    p = np.asarray(p, dtype=dtype)
    return (p[OW1:OB1].reshape(H, F), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def forward(p, feats, keep=False):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = bf16(np.asarray(feats, np.float32)[:, :F])
    z1 = x @ bf16(W1).T + b1
    h1 = bf16(gelu(z1))
    z2 = h1 @ bf16(W2).T + b2
    h2 = gelu(z2)
    o = h2 @ bf16(w3) + float(b3)
    if keep:
        return o, (x, z1, h1, z2, h2)
    return o


def loss(p, feats, t, tau):
    d = np.asarray(t, np.float64) - forward(p, feats)
    return float(np.mean(np.where(d >= 0, tau * d, (tau - 1) * d)))


def grad(p, feats, t, tau):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    o, (x, z1, h1, z2, h2) = forward(p, feats, keep=True)
    t = np.asarray(t, np.float64)
    n = len(t)
    g = np.where(o > t, 1 - tau, np.where(o < t, -tau, 0.0)) / n
    d2 = g[:, None] * bf16(w3)[None, :] * gelu_grad(z2)
    d1 = (d2 @ bf16(W2)) * gelu_grad(z1)
    out = np.zeros(NP)
    out[OW1:OB1] = (d1.T @ x).ravel()
    out[OB1:OW2] = d1.sum(0)
    out[OW2:OB2] = (d2.T @ h1).ravel()
    out[OB2:OW3] = d2.sum(0)
    out[OW3:OB3] = (g[:, None] * h2).sum(0)
    out[OB3] = g.sum()
    return out


def grad_fd(p, feats, t, tau, idx, eps=1e-5):
    """Central differences of `loss` w.r.t. parameters idx, with the BF16
    rounding of the weights bypassed (the analytic gradient treats it as the
    identity), i.e. on the network whose weights are exactly p.  The
    activation is piecewise quadratic in each weight, so a central difference
    is exact unless the step crosses a kink: the step is small (1e-5) and the
    perturbed weights stay in float64."""
    def loss_exact(q):
        W1, b1, W2, b2, w3, b3 = unpack(q, np.float64)  # the perturbation kept exact
        x = bf16(np.asarray(feats, np.float32)[:, :F])
        z1 = x @ W1.T + b1
        h1 = gelu(z1)
        z2 = h1 @ W2.T + b2
        o = gelu(z2) @ w3 + float(b3)
        d = np.asarray(t, np.float64) - o
        return float(np.mean(np.where(d >= 0, tau * d, (tau - 1) * d)))

    p64 = np.asarray(p, np.float64)
    out = []
    for i in idx:
        a = p64.copy(); a[i] += eps
        b = p64.copy(); b[i] -= eps
        out.append((loss_exact(a) - loss_exact(b)) / (2 * eps))
    return np.array(out)


def grad_exact(p, feats, t, tau):
    """Analytic gradient of the network grad_fd differentiates (no BF16
    rounding of weights or hidden activations)."""
    W1, b1, W2, b2, w3, b3 = [np.asarray(v, np.float64) for v in unpack(p)]
    x = bf16(np.asarray(feats, np.float32)[:, :F])
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    o = h2 @ w3 + float(b3)
    t = np.asarray(t, np.float64)
    g = np.where(o > t, 1 - tau, np.where(o < t, -tau, 0.0)) / len(t)
    d2 = g[:, None] * w3[None, :] * gelu_grad(z2)
    d1 = (d2 @ W2) * gelu_grad(z1)
    out = np.zeros(NP)
    out[OW1:OB1] = (d1.T @ x).ravel()
    out[OB1:OW2] = d1.sum(0)
    out[OW2:OB2] = (d2.T @ h1).ravel()
    out[OB2:OW3] = d2.sum(0)
    out[OW3:OB3] = (g[:, None] * h2).sum(0)
    out[OB3] = g.sum()
    return out


def features(points, queries, r, count_scale):
    """Feature rows (n x 32) on the edge-r count grid of `points` (origin =
    points' minimum, dims = floor((max - min) / r) + 1, FP32 cell arithmetic)."""
    pts = np.asarray(points, np.float32)
    q = np.asarray(queries, np.float32)
    mn = pts.min(axis=0)
    inv = np.float32(1.0) / np.float32(r)
    dims = (np.floor((pts.max(axis=0) - mn) * inv)).astype(np.int64) + 1
    pc = np.clip(np.floor((pts - mn) * inv).astype(np.int64), 0, dims - 1)
    lin = (pc[:, 2] * dims[1] + pc[:, 1]) * dims[0] + pc[:, 0]
    counts = np.bincount(lin, minlength=int(np.prod(dims)))
    tq = (q - mn) * inv
    fl = np.clip(np.floor(tq), -2, dims + 1)
    c = fl.astype(np.int64)
    off = np.clip(tq - fl, 0.0, 1.0)
    out = np.zeros((len(q), FP), np.float64)
    for f in range(27):
        d = np.array([f % 3 - 1, (f // 3) % 3 - 1, f // 9 - 1])
        cc = c + d
        ok = ((cc >= 0) & (cc < dims)).all(axis=1)
        v = np.zeros(len(q))
        v[ok] = counts[(cc[ok, 2] * dims[1] + cc[ok, 1]) * dims[0] + cc[ok, 0]]
        out[:, f] = np.log1p(v) * count_scale
    out[:, 27:30] = off
    return out
