"""Pin the FP64 MLP oracle itself: its analytic gradient equals central
finite differences of its loss (no bf16 rounding)."""
import numpy as np

from oracle import mlp


def test_mlp_oracle_gradient_finite_difference():
    rng = np.random.default_rng(0)
    n = 12
    X = rng.normal(size=(n, mlp.F))
    T = rng.normal(size=(n, mlp.O))
    p = np.zeros(23270)
    p[:23171] = rng.normal(0, 0.2, 23171)
    p[23171:23219] = rng.normal(0, 0.1, 48)     # mean
    p[23219:23267] = rng.uniform(0.5, 2.0, 48)  # istd
    p[23267:] = rng.uniform(0.5, 2.0, 3)        # oscale

    def L(q):
        sse, _ = mlp.loss_and_grad(X, T, q, rounding=False)
        return sse / n

    _, g = mlp.loss_and_grad(X, T, p, rounding=False)
    idx = np.concatenate([rng.choice(6144, 20, replace=False), 6144 + rng.choice(128, 5),
                          6272 + rng.choice(16384, 20), 22656 + rng.choice(128, 5),
                          22784 + rng.choice(384, 10), [23168, 23169, 23170]])
    for i in idx:
        h = 1e-6
        qp = p.copy(); qp[i] += h
        qm = p.copy(); qm[i] -= h
        fd = (L(qp) - L(qm)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])
