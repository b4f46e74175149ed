"""Shared helpers of the test-suite: seeded cases and vector comparisons."""
from __future__ import annotations

import numpy as np

from hsrq3in_amd import workloads as w


def case(n: int, seed: int, m: int | None = None, kmax: int = 0, pert: float = 1e-12,
         frac: float = 1.0):
    """(H, shifts): Gaussian upper Hessenberg (graded if kmax) and its first m
    real eigenvalues perturbed by relative ``pert``."""
    h0 = w.gaussian_hessenberg(n, seed)
    ev = np.linalg.eigvals(h0)
    re = ev.real[ev.imag == 0.0]
    if m is None:
        m = re.size
    m = min(m, re.size)
    shifts = re[:m] * (1.0 + pert)
    h = w.graded(h0, kmax, seed, frac) if kmax else h0
    return h, np.ascontiguousarray(shifts)


def cosines(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-wise cos of the angle (overflow-safe for unit-norm columns)."""
    na = np.linalg.norm(a, axis=0)
    nb = np.linalg.norm(b, axis=0)
    return np.sum(a * b, axis=0) / (na * nb)


def one_minus_abs_cos(a, b):
    return 1.0 - np.abs(cosines(a, b))
