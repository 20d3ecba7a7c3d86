"""Pins the router oracle: FP64 analytic gradient == central finite differences,
and the bf16 rounding helper matches the bfloat16 definition."""
import numpy as np

from oracle import router as R


def test_bf16_rounding_definition():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-3], np.float32)
    r = R.to_bf16(x)
    # representable with 8 significant bits and within half an ulp (2^-8 relative)
    for a, b in zip(x.astype(np.float64), r):
        m, e = np.frexp(b)
        assert float(m * 256).is_integer()
        assert abs(a - b) <= abs(a) * 2.0 ** -8
    assert r[1] == 1.0           # tie (1 + 2^-8) rounds to even
    assert r[2] == 1.0078125     # 1 + 1.5*2^-8 rounds up


def test_fp64_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, h, c, B = 6, 9, 5, 7
    p = rng.standard_normal(h * d + h + c * h + c) * 0.5
    X = rng.standard_normal((B, d)).astype(np.float32)
    labels = rng.integers(0, c + 1, size=(B, 3))
    labels[:, 0] = rng.integers(1, c + 1, size=B)       # at least one label per row
    _, g = R.loss_and_grad(p, X, labels, d, h, c)
    eps = 1e-6
    for i in rng.choice(p.size, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = R.loss_and_grad(pp, X, labels, d, h, c)
        lm, _ = R.loss_and_grad(pm, X, labels, d, h, c)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_targets_uniform_over_labels():
    T = R.targets(np.array([[2, 2, 0], [1, 3, 0]]), 4)
    assert T.tolist() == [[0, 1, 0, 0], [0.5, 0, 0.5, 0]]
