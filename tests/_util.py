"""Shared helpers of the test-suite."""

from __future__ import annotations

import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def min_image(d, L):
    if L is None:
        return d
    return d - L * np.round(d / L)


def random_cloud(m, seed, lo=-5.0, hi=5.0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(lo, hi, size=(m, 2))
    f = rng.uniform(0.01, 1.0, size=(m, 7))
    return pos, f
