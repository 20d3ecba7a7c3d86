"""Synthetic workloads of the measured configurations (NumPy Philox streams).

Every generator is deterministic in its seed.  Arrays are returned point-major
(n, d) float32; pass ``X.T`` for the Julia d x n view (same memory).
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def gaussian(n: int, d: int, seed: int) -> np.ndarray:
    """i.i.d. N(0, 1) points (base 10,000 x 128 and its 1,000 queries)."""
    return _rng(seed).standard_normal((n, d), dtype=np.float32)


def clipped_mixture(n: int = 60_000, d: int = 784, clusters: int = 10, sigma: float = 0.15,
                    seed: int = 2) -> np.ndarray:
    """Mixture of isotropic Gaussians N(c, sigma^2 I), centres ~ U[0,1]^d, equal
    weights (labels drawn uniformly), clipped to [0, 1]."""
    g = _rng(seed)
    centres = g.random((clusters, d), dtype=np.float32)
    labels = g.integers(0, clusters, size=n)
    X = np.empty((n, d), np.float32)
    step = 8192
    for s in range(0, n, step):
        e = min(n, s + step)
        X[s:e] = centres[labels[s:e]] + sigma * g.standard_normal((e - s, d), dtype=np.float32)
    np.clip(X, 0.0, 1.0, out=X)
    return X


def zipf_anisotropic_mixture(n: int = 1_000_000, d: int = 256, components: int = 1000,
                             zipf_s: float = 1.1, seed: int = 3, return_labels: bool = False):
    """Mixture of `components` anisotropic Gaussians: centres ~ N(0, 4 I); per
    component per dimension scales ~ LogUniform[0.1, 1]; weights ~ i^-zipf_s
    (uneven inverted lists)."""
    g = _rng(seed)
    centres = (2.0 * g.standard_normal((components, d))).astype(np.float32)
    scales = np.exp(g.uniform(np.log(0.1), np.log(1.0), size=(components, d))).astype(np.float32)
    w = np.arange(1, components + 1, dtype=np.float64) ** (-zipf_s)
    w /= w.sum()
    labels = g.choice(components, size=n, p=w)
    X = np.empty((n, d), np.float32)
    step = 65536
    for s in range(0, n, step):
        e = min(n, s + step)
        lab = labels[s:e]
        X[s:e] = centres[lab] + scales[lab] * g.standard_normal((e - s, d), dtype=np.float32)
    return (X, labels) if return_labels else X


def mixture_queries(n: int, d: int = 256, components: int = 1000, zipf_s: float = 1.1,
                    seed: int = 3, query_seed: int = 33) -> np.ndarray:
    """Queries from the same mixture as zipf_anisotropic_mixture(seed) (same
    centres/scales/weights), drawn with an independent stream."""
    g = _rng(seed)
    centres = (2.0 * g.standard_normal((components, d))).astype(np.float32)
    scales = np.exp(g.uniform(np.log(0.1), np.log(1.0), size=(components, d))).astype(np.float32)
    w = np.arange(1, components + 1, dtype=np.float64) ** (-zipf_s)
    w /= w.sum()
    h = _rng(query_seed)
    lab = h.choice(components, size=n, p=w)
    return (centres[lab] + scales[lab] * h.standard_normal((n, d), dtype=np.float32)).astype(np.float32)


def unit_mixture(n: int, d: int = 256, components: int = 5000, noise: float = 0.1, seed: int = 5,
                 centres_seed: int = 55) -> np.ndarray:
    """normalize(c + noise * N(0, I)) with centres uniform on the sphere."""
    gc = _rng(centres_seed)
    C = gc.standard_normal((components, d)).astype(np.float32)
    C /= np.linalg.norm(C, axis=1, keepdims=True)
    g = _rng(seed)
    lab = g.integers(0, components, size=n)
    X = np.empty((n, d), np.float32)
    step = 65536
    for s in range(0, n, step):
        e = min(n, s + step)
        v = C[lab[s:e]] + noise * g.standard_normal((e - s, d), dtype=np.float32)
        X[s:e] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return X
