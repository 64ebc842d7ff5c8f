"""Device MLP kernels vs the FP64 oracle (oracle/mlp.py) with the same bf16
roundings: forward outputs and minibatch gradients must agree to fp32
accumulation accuracy; training must reduce the loss."""
import numpy as np
import pytest

from oracle import mlp
from yarnball_amd import learned


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 2.0, (n, 48)).astype(np.float32) + 0.5
    T = rng.normal(0, 1e-3, (n, 3)).astype(np.float32)
    p = learned.init_params(X, T, seed=seed)
    p[learned.B1:learned.W2] = rng.normal(0, 0.1, 128)
    p[learned.B2:learned.W3] = rng.normal(0, 0.1, 128)
    p[learned.B3:learned.MEAN] = rng.normal(0, 0.1, 3)
    return X, T, p


@pytest.mark.gpu
def test_mlp_forward_matches_oracle():
    X, T, p = random_problem(200, 1)
    Y = learned.forward(X, p)
    ref = mlp.forward(X, p)["y"]
    err = np.max(np.abs(Y - ref)) / np.max(np.abs(ref))
    print(f"forward rel err {err:.2e}")
    assert err < 1e-4


@pytest.mark.gpu
def test_mlp_gradient_matches_oracle():
    X, T, p = random_problem(256, 2)
    g, sse = learned.grad(X, T, p)
    sse_ref, g_ref = mlp.loss_and_grad(X, T, p)
    assert sse == pytest.approx(sse_ref, rel=1e-4)
    for name, a, b in [("W1", 0, 6144), ("b1", 6144, 6272), ("W2", 6272, 22656),
                       ("b2", 22656, 22784), ("W3", 22784, 23168), ("b3", 23168, 23171)]:
        d = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        print(f"grad {name}: rel err {d:.2e}")
        assert d < 2e-3, name


@pytest.mark.gpu
def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(3)
    n = 65536
    X = rng.normal(0, 1, (n, 48)).astype(np.float32)
    # smooth target that depends on a few inputs
    T = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), X[:, 2] * X[:, 3] * 0.3,
                  np.sin(X[:, 4])], axis=1).astype(np.float32) * 1e-3
    p0 = learned.init_params(X, T, seed=0)
    p, hist = learned.train(X, T, p0, epochs=6, batch=4096, lr=2e-3)
    print("loss per epoch", hist)
    assert hist[-1] < 0.2 * hist[0]
    # held-out error through the forward kernel
    Xv = rng.normal(0, 1, (4096, 48)).astype(np.float32)
    Tv = np.stack([np.tanh(Xv[:, 0] + 0.5 * Xv[:, 1]), Xv[:, 2] * Xv[:, 3] * 0.3,
                   np.sin(Xv[:, 4])], axis=1).astype(np.float32) * 1e-3
    Y = learned.forward(Xv, p) * p[learned.OSCALE:]
    rel = np.linalg.norm(Y - Tv) / np.linalg.norm(Tv)
    print(f"held-out relative error {rel:.3f}")
    assert rel < 0.5
