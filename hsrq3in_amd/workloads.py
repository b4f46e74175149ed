"""Synthetic problem generators for robust inverse iteration.

Every generator is seeded and deterministic (numpy PCG64), so a matrix built
here and one rebuilt on another host from the same seed are bit-identical;
that is what lets the (expensive, O(n^3)) shift lists be computed once and
stored as small fixture files.

Matrices are returned column-major (Fortran order), the layout of the C
entry point ``tiled_inverse_iteration`` (reference ``inverse_iteration.h:19-22``).

Generators
----------
* ``gaussian_hessenberg(n, seed)`` -- H[i, j] ~ N(0, 1) for i <= j + 1, zero
  below the subdiagonal (the workload of the plain and batched benchmarks).
* ``graded(H0, kmax, seed)`` -- H = D H0 D^-1, D = diag(2^k), k uniform in
  [-kmax, kmax].  Power-of-two scaling is exact in FP64, so the spectrum is
  that of H0 bit for bit and the eigenvector is D x0.
* ``constructed_spectrum(lam, seed)`` -- an unreduced upper Hessenberg matrix
  with a prescribed real spectrum, built in O(n^2) as H = U B U^-1 with B
  lower bidiagonal (diagonal = lam) and U upper bidiagonal.  An upper
  triangular similarity keeps Hessenberg form, and B is triangular, so the
  eigenvalues are lam up to the rounding of the construction.  Used where the
  O(n^3) eigenvalue computation of a Gaussian matrix is too costly (eigenvalue
  clusters, n >= 16000).
* ``real_shifts(H, m)`` -- the first m real eigenvalues (LAPACK dgeev via
  numpy: balancing + dhseqr) in returned order, perturbed by relative 1e-12.
"""
from __future__ import annotations

import os
import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "tests", "golden")

SHIFT_PERTURBATION = 1e-12


def gaussian_hessenberg(n: int, seed: int) -> np.ndarray:
    """Upper Hessenberg, N(0,1) on and above the subdiagonal, column-major."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    a = np.triu(a, -1)
    return np.asfortranarray(a)


def grading_exponents(n: int, kmax: int, seed: int, frac: float = 1.0) -> np.ndarray:
    """Integer exponents k_i in [-kmax, kmax]; a fraction ``1-frac`` of them 0."""
    rng = np.random.default_rng(seed ^ 0x5EED)
    k = rng.integers(-kmax, kmax + 1, size=n)
    if frac < 1.0:
        keep = rng.random(n) < frac
        k = np.where(keep, k, 0)
    return k.astype(np.int64)


def graded(h0: np.ndarray, kmax: int, seed: int, frac: float = 1.0) -> np.ndarray:
    """H = D H0 D^-1 with D = diag(2^k): exact, same eigenvalues as H0."""
    n = h0.shape[0]
    k = grading_exponents(n, kmax, seed, frac)
    h = np.ldexp(np.asarray(h0), (k[:, None] - k[None, :]).astype(np.int32))
    return np.asfortranarray(h)


def constructed_spectrum(lam: np.ndarray, seed: int, sub: float = 0.15,
                         ratio: float = 0.9, weak=None, weak_nu: float = 0.0) -> np.ndarray:
    """Unreduced upper Hessenberg H with eigenvalues ``lam`` (O(n^2) build).

    B = diag(lam) + subdiag(nu) (lower bidiagonal, so eig(B) = lam exactly);
    U = diag(d) + superdiag(u), |u_j / d_{j+1}| <= ratio < 1 (cond(U) modest).
    H = U B U^-1 is computed column by column from  X U = U B.
    ``weak``: indices whose couplings nu_{k-1}, nu_k are reduced to
    ``weak_nu`` (times the same random factor), so that eigenvalues closer
    together than nu (a cluster) keep well-conditioned eigenvectors.
    """
    nu, d, u = constructed_parameters(lam, seed, sub, ratio, weak, weak_nu)
    return constructed_from_parameters(lam, nu, d, u)


def constructed_parameters(lam: np.ndarray, seed: int, sub: float = 0.15, ratio: float = 0.9,
                           weak=None, weak_nu: float = 0.0):
    """(nu, d, u) of ``constructed_spectrum`` (seeded)."""
    lam = np.asarray(lam, dtype=np.float64)
    n = lam.size
    rng = np.random.default_rng(seed)
    scale = max(1.0, float(np.max(np.abs(lam)))) if n else 1.0
    nu = sub * scale * rng.uniform(0.5, 1.0, size=max(n - 1, 0))
    if weak is not None and n > 1:
        f = nu / (sub * scale)
        for k in np.asarray(weak, dtype=np.int64):
            for q in (k - 1, k):
                if 0 <= q < n - 1:
                    nu[q] = weak_nu * f[q]
    d = rng.uniform(1.0, 2.0, size=n)
    u = rng.uniform(-ratio, ratio, size=max(n - 1, 0)) * d[1:]
    return nu, d, u


def constructed_from_parameters(lam, nu, d, u) -> np.ndarray:
    """H = U B U^-1 for B = diag(lam) + subdiag(nu), U = diag(d) + superdiag(u).

    The device build (Context.constructed_hessenberg) performs the same
    operations in the same order and agrees bit for bit."""
    lam = np.asarray(lam, dtype=np.float64)
    n = lam.size
    # UB is tridiagonal: (UB)[i,i] = d_i lam_i + u_i nu_i, (UB)[i+1,i] = d_{i+1} nu_i,
    # (UB)[i,i+1] = u_i lam_{i+1}.
    x = np.zeros((n, n), dtype=np.float64, order="F")
    for j in range(n):
        col = np.zeros(min(j + 2, n))
        col[j] = d[j] * lam[j] + (u[j] * nu[j] if j < n - 1 else 0.0)
        if j + 1 < n:
            col[j + 1] = d[j + 1] * nu[j]
        if j >= 1:
            col[j - 1] = u[j - 1] * lam[j]
            col[:j + 1] -= u[j - 1] * x[:j + 1, j - 1]
        x[:col.size, j] = col / d[j]
    return x


def cluster_spectrum(n: int, m: int, n_cluster: int, seed: int,
                     spacing: float = 1e-8) -> tuple[np.ndarray, np.ndarray]:
    """Spectrum with ``n_cluster`` eigenvalues at relative spacing ``spacing``.

    Cluster members are interleaved with well separated eigenvalues so that
    the bidiagonal B stays far from a Jordan block.  Returns (lam, shifts)
    where shifts are the m prescribed eigenvalues (cluster first) perturbed by
    relative 1e-12.  Build H with ``cluster_hessenberg``.
    """
    rng = np.random.default_rng(seed)
    lam = rng.uniform(-np.sqrt(n), np.sqrt(n), size=n)
    lam_c = 0.5 * np.sqrt(n) + 1.0
    pos = cluster_positions(n, n_cluster)
    lam[pos] = lam_c * (1.0 + spacing * np.arange(n_cluster))
    rest = np.setdiff1d(np.arange(n), pos)
    chosen = np.concatenate([pos, rng.choice(rest, size=m - n_cluster, replace=False)])
    shifts = lam[chosen] * (1.0 + SHIFT_PERTURBATION)
    return lam, shifts


def cluster_positions(n: int, n_cluster: int) -> np.ndarray:
    """Diagonal positions of the cluster members in ``cluster_spectrum``."""
    stride = max(1, n // max(n_cluster, 1))
    return np.arange(n_cluster) * stride


def cluster_hessenberg(lam: np.ndarray, n_cluster: int, seed: int,
                       spacing: float = 1e-8) -> np.ndarray:
    """H with the spectrum of ``cluster_spectrum``.  B's couplings are
    ~max|lambda|/n (far below the typical distance of eigenvalues adjacent on
    B's diagonal, so no product of coupling/distance ratios grows) and those
    of cluster members 0.1 x their distance (spacing |lambda_c|), so every
    eigenvalue stays where it was put (condition numbers O(1)) and the
    shifts are within 1e-12 of eigenvalues of the stored H."""
    lam = np.asarray(lam)
    n = lam.size
    pos = cluster_positions(n, n_cluster)
    lam_c = float(np.max(np.abs(lam[pos]))) if n_cluster else 1.0
    return constructed_spectrum(lam, seed, sub=1.0 / max(n, 1), weak=pos,
                                weak_nu=0.1 * spacing * lam_c)


def real_shifts(h: np.ndarray, m: int) -> np.ndarray:
    """First m real eigenvalues (dgeev order), perturbed by relative 1e-12."""
    w = np.linalg.eigvals(np.asarray(h))
    real = w.real[w.imag == 0.0]
    if real.size < m:
        raise ValueError(f"matrix has only {real.size} real eigenvalues, {m} requested")
    return np.ascontiguousarray(real[:m] * (1.0 + SHIFT_PERTURBATION))


class ShiftCache:
    """The step before the solve for requests that bring no shifts: the real
    eigenvalues of H (LAPACK dgeev / dhseqr through numpy, O(n^3) on the host),
    computed once per matrix and kept in ``directory`` as ``<digest>.npz``
    (the digest is SHA-256 over the order and the bytes of H).  ``get(H, m)``
    returns the first m of them perturbed by relative 1e-12, as
    ``real_shifts``."""

    def __init__(self, directory: str):
        self.dir = directory
        os.makedirs(directory, exist_ok=True)
        self.hits = 0
        self.misses = 0

    @staticmethod
    def digest(h) -> str:
        import hashlib
        h = np.asfortranarray(h, dtype=np.float64)
        d = hashlib.sha256()
        d.update(np.array(h.shape, dtype=np.int64).tobytes())
        d.update(memoryview(h.ravel(order="F")))
        return d.hexdigest()

    def real_eigenvalues(self, h) -> np.ndarray:
        path = os.path.join(self.dir, self.digest(h) + ".npz")
        if os.path.exists(path):
            self.hits += 1
            with np.load(path, allow_pickle=False) as z:
                return np.array(z["real"])
        self.misses += 1
        w = np.linalg.eigvals(np.asarray(h))
        real = np.ascontiguousarray(w.real[w.imag == 0.0])
        tmp = path + ".part.npz"
        np.savez(tmp, real=real)
        os.replace(tmp, path)
        return real

    def get(self, h, m: int) -> np.ndarray:
        real = self.real_eigenvalues(h)
        if real.size < m:
            raise ValueError(f"matrix has only {real.size} real eigenvalues, {m} requested")
        return np.ascontiguousarray(real[:m] * (1.0 + SHIFT_PERTURBATION))


# ---------------------------------------------------------------------------
# Named workloads (sizes of the benchmark configurations)
# ---------------------------------------------------------------------------

def batch_sizes(count: int = 256, seed: int = 3, nmin: int = 200, nmax: int = 4000):
    """(n_p, m_p, graded_p, seed_p) for the batched workload."""
    rng = np.random.default_rng(seed)
    n = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count))).astype(int)
    f = rng.uniform(0.05, 0.2, size=count)
    m = np.maximum(1, np.round(f * n)).astype(int)
    g = rng.random(count) < 0.10
    seeds = 1000 + np.arange(count)
    return [(int(a), int(b), bool(c), int(d)) for a, b, c, d in zip(n, m, g, seeds)]


# Graded range of the serving workload.  D = diag(2^k) with |k| <= 450 keeps
# every entry 2^(k_i - k_j) H0_ij and every row sum ||H(I,:)||_oo below the
# overflow threshold Omega = 1e300 that the robust update protection assumes
# (reference robust.c:14-15, :244-262) up to n = 16000.  The nominal
# k in [-600, 600] would put entries of 2^1200 (Inf) into H.
SERVE_KMAX = 450


def serving_problem(n: int, seed: int, graded: bool):
    """(H, lam): an unreduced Hessenberg matrix of order n with the known real
    spectrum lam (uniform in [-sqrt(n), sqrt(n)], random order along the
    diagonal of B) built in O(n^2) by ``constructed_spectrum``; the
    subdiagonal of B is ~1/sqrt(n), far below the typical distance of
    eigenvalues adjacent on B's diagonal (~sqrt(n)), so the eigenvectors are
    well conditioned.  Graded variants are D H D^-1 with k in
    [-SERVE_KMAX, SERVE_KMAX] (same spectrum).  lam are used as shifts
    unperturbed: the rounding of the construction (~1e-13 absolute) already
    separates them from the eigenvalues of the stored H."""
    rng = np.random.default_rng(seed)
    r = np.sqrt(n)
    lam = rng.uniform(-r, r, size=n)
    h = constructed_spectrum(lam, seed + 1, sub=1.0 / n)
    if graded:
        h = graded_fn(h, SERVE_KMAX, seed + 2)
    return h, lam


graded_fn = graded


def serving_problem_device(ctx, n: int, seed: int, graded: bool, device: str = "cuda"):
    """``serving_problem`` built in HBM by the device generators (bit-identical
    to the host build: same parameters, same operations): (Ht, lam) with
    Ht[j, i] = H(i, j)."""
    rng = np.random.default_rng(seed)
    r = np.sqrt(n)
    lam = rng.uniform(-r, r, size=n)
    nu, d, u = constructed_parameters(lam, seed + 1, sub=1.0 / n)
    Ht = ctx.constructed_hessenberg(lam, nu, d, u, device=device)
    if graded:
        ctx.grade(Ht, grading_exponents(n, SERVE_KMAX, seed + 2))
    return Ht, lam


def serving_sizes(count: int, seed: int = 5, nmin: int = 500, nmax: int = 16000,
                  graded_frac: float = 0.15):
    """(n, graded, seed) of a pool of serving problems: n log-uniform in
    [nmin, nmax], a fraction graded."""
    rng = np.random.default_rng(seed)
    n = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count))).astype(int)
    g = rng.random(count) < graded_frac
    return [(int(a), bool(b), 7000 + i) for i, (a, b) in enumerate(zip(n, g))]


def load_shifts(name: str) -> np.ndarray:
    path = os.path.join(GOLDEN_DIR, name)
    with np.load(path, allow_pickle=False) as z:
        return np.array(z["shifts"])


class BatchWorkload:
    """The 256-problem batch: n log-uniform in [200, 4000], m = 5-20 % of n
    real-eigenvalue shifts (stored fixture), 10 % graded with k in
    [-400, 400].  Problem p is rebuilt from its seed on demand."""

    KMAX = 400

    def __init__(self, name: str = "shifts_batch256.npz"):
        path = os.path.join(GOLDEN_DIR, name)
        with np.load(path, allow_pickle=False) as z:
            self.shifts = np.array(z["shifts"])
            self.offsets = np.array(z["offsets"])
            self.n = np.array(z["n"])
            self.m = np.array(z["m"])
            self.graded = np.array(z["graded"])
            self.seeds = np.array(z["seeds"])

    def __len__(self):
        return len(self.n)

    def problem(self, p: int):
        n, seed = int(self.n[p]), int(self.seeds[p])
        h = gaussian_hessenberg(n, seed)
        if self.graded[p]:
            h = graded(h, self.KMAX, seed)
        s = self.shifts[self.offsets[p]:self.offsets[p + 1]]
        return h, np.ascontiguousarray(s)
