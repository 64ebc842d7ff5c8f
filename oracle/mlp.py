"""FP64 CPU oracle of the learned warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Checks the device forward / backward kernels (yarnball_amd/csrc/mlp.hip).
The network has no counterpart in the reference; this restates the model the
device implements: y = W3^T relu(W2^T relu(W1^T xn + b1) + b2) + b3 with
xn = (x - mean) * istd, loss L = (1/n) sum_samples |y - t / oscale|^2.

`bf16=True` rounds exactly the quantities the device stores in bf16 (inputs,
weights, hidden activations, back-propagated gradients) so that the remaining
difference is only fp32-vs-fp64 accumulation.
"""
import numpy as np

F, H, O = 48, 128, 3


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def split(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H); o += F * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * O].reshape(H, O); o += H * O
    b3 = p[o:o + O]; o += O
    mean = p[o:o + F]; o += F
    istd = p[o:o + F]; o += F
    osc = p[o:o + O]
    return W1, b1, W2, b2, W3, b3, mean, istd, osc


def normalise(X, p):
    W1, b1, W2, b2, W3, b3, mean, istd, osc = split(p)
    return ((np.asarray(X, np.float32) - mean.astype(np.float32)) * istd.astype(np.float32))


def forward(X, p, rounding=True):
    W1, b1, W2, b2, W3, b3, mean, istd, osc = split(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    x = r(normalise(X, p))
    z1 = x @ r(W1) + b1
    h1 = r(np.maximum(z1, 0))
    z2 = h1 @ r(W2) + b2
    h2 = r(np.maximum(z2, 0))
    y = h2 @ r(W3) + b3
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def loss_and_grad(X, T, p, rounding=True):
    """Returns (sum of squared errors, gradient of L w.r.t. the trainable part)."""
    W1, b1, W2, b2, W3, b3, mean, istd, osc = split(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    n = len(X)
    a = forward(X, p, rounding)
    t = np.asarray(T, np.float64) / osc
    e = a["y"] - t
    sse = float((e ** 2).sum())
    dy = r(2.0 * e / n)
    g2 = r((dy @ r(W3).T) * (a["z2"] > 0))
    g1 = r((g2 @ r(W2).T) * (a["z1"] > 0))
    dW3 = a["h2"].T @ dy
    dW2 = a["h1"].T @ g2
    dW1 = a["x"].T @ g1
    grad = np.concatenate([dW1.ravel(), g1.sum(0), dW2.ravel(), g2.sum(0), dW3.ravel(),
                           dy.sum(0)])
    return sse, grad
