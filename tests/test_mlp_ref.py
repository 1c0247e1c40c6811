"""The proposal network's CPU reference pinned by finite differences."""
import numpy as np

from oracle import mlp_ref as R


def test_backward_matches_finite_differences():
    sizes = (16, 32, 32, 8)
    n = sum(b * (a + 1) for a, b in zip(sizes[:-1], sizes[1:]))
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, n).astype(np.float64)
    x = rng.normal(size=(4, 16))
    d = rng.normal(size=(4, 8))

    def f(q):
        o, _ = R.forward_backward(q.astype(np.float32), sizes, x.astype(np.float32), d.astype(np.float32), False)
        return float((o.astype(np.float64) * d).sum())

    _, g = R.forward_backward(p.astype(np.float32), sizes, x.astype(np.float32), d.astype(np.float32), False)
    for i in rng.choice(n, 40, replace=False):
        h = 1e-2
        q = p.copy()
        q[i] += h
        fp = f(q)
        q[i] -= 2 * h
        fm = f(q)
        assert abs((fp - fm) / (2 * h) - g[i]) <= 2e-3 * max(1.0, abs(g[i]))


def test_wta_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    out = rng.normal(size=(5, 16 * 4))
    lab = rng.normal(size=(5, 3))
    _, g = R.wta_loss(out, lab)
    for idx in [(0, 0), (1, 5), (2, 10), (3, 63), (4, 30), (0, 3), (1, 7)]:
        h = 1e-6
        o = out.copy()
        o[idx] += h
        lp, _ = R.wta_loss(o, lab)
        o[idx] -= 2 * h
        lm, _ = R.wta_loss(o, lab)
        assert abs((lp - lm) / (2 * h) - g[idx]) < 1e-5


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    assert list(R.bf16(x)) == [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 0.0]
