"""Pins the oracle's complex-shift solve (oracle_complex_inverse_iteration).

Checks that do not retype the algorithm: a dense complex LU solve of
(H - sigma I) x = eps3 (1, ..., 1) (direction and phase), LAPACK's
eigenvectors (scipy.linalg.eig), the closed form n = 1, conjugate symmetry,
the reduction to the real-shift oracle for wi = 0, exact graded similarity,
and finiteness with the residual bound where protections trigger.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from hsrq3in_amd import workloads as w

EPS = 2.0 ** -52


def phase_normalise(x):
    x = x / np.linalg.norm(x)
    p = int(np.argmax(np.abs(x)))
    return x * (np.conj(x[p]) / abs(x[p]))


def complex_eigs(h, count):
    ev = np.linalg.eigvals(h)
    cz = ev[ev.imag > 0]
    return cz[:count]


def residuals(h, sig, x):
    r = h @ x - x * sig
    return np.linalg.norm(r, axis=0) / (np.linalg.norm(h, np.inf) * np.linalg.norm(x, axis=0) * EPS)


@pytest.mark.parametrize("n,seed", [(2, 0), (17, 2), (60, 3), (200, 4)])
def test_matches_dense_solve(n, seed):
    h = w.gaussian_hessenberg(n, seed)
    cz = complex_eigs(h, 5)
    if cz.size == 0:
        pytest.skip("no complex pair")
    sig = cz * (1 + 1e-12)
    x, st = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    # dlaein's growth test needs |sigma - lambda| below ~ n eps ||H||
    assert (st & 1).all() or n < 50
    eps3 = np.linalg.norm(h, np.inf) * 2.0 ** -53
    for k, s in enumerate(sig):
        xd = np.linalg.solve(h - s * np.eye(n), np.full(n, eps3, dtype=complex))
        assert np.abs(phase_normalise(xd) - x[:, k]).max() <= 1e-12
    conv = (st & 1) == 1
    assert (residuals(h, sig, x)[conv] <= 10 * n).all()


def test_far_shift_matches_dense_solve():
    n = 80
    h = w.gaussian_hessenberg(n, 9)
    sig = np.array([0.3 + 25.0j, -40.0 + 0.5j, 2.0 - 3.0j])
    x, st = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    eps3 = np.linalg.norm(h, np.inf) * 2.0 ** -53
    for k, s in enumerate(sig):
        xd = np.linalg.solve(h - s * np.eye(n), np.full(n, eps3, dtype=complex))
        assert np.abs(phase_normalise(xd) - x[:, k]).max() <= 1e-12
    assert not (st & 1).all()          # no growth away from the spectrum


def test_matches_lapack_eigenvectors():
    n = 150
    h = w.gaussian_hessenberg(n, 21)
    lam, v = sl.eig(h)
    sel = np.where(lam.imag > 0)[0][:6]
    sig = lam[sel] * (1 + 1e-12)
    x, st = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    assert (st & 1).all()
    for j, k in enumerate(sel):
        c = abs(np.vdot(v[:, k], x[:, j])) / np.linalg.norm(v[:, k])
        assert 1.0 - c <= 1e-10


def test_n1_closed_form():
    x, st = oracle.complex_inverse_iteration(np.array([[2.5]]), [2.5], [1e-3])
    assert x.shape == (1, 1) and x[0, 0] == 1.0


def test_conjugate_shift_gives_conjugate_vector():
    h = w.gaussian_hessenberg(120, 5)
    cz = complex_eigs(h, 4)
    x, _ = oracle.complex_inverse_iteration(h, cz.real, cz.imag)
    xc, _ = oracle.complex_inverse_iteration(h, cz.real, -cz.imag)
    assert np.array_equal(xc, np.conj(x))


def test_real_shift_reduces_to_the_real_oracle():
    n = 300
    h = w.gaussian_hessenberg(n, 8)
    ev = np.linalg.eigvals(h)
    wr = ev.real[ev.imag == 0][:6] * (1 + 1e-12)
    x, st = oracle.complex_inverse_iteration(h, wr, np.zeros_like(wr))
    assert np.all(x.imag == 0.0) and (st & 1).all()
    xo, wo, sto = oracle.tiled_inverse_iteration(h, wr, 64, 8)
    c = np.abs(np.sum(x.real * xo, axis=0)) / np.linalg.norm(xo, axis=0)
    assert (1.0 - c).max() <= 1e-10


def test_graded_similarity():
    n, kmax = 120, 6
    h0 = w.gaussian_hessenberg(n, 13)
    cz = complex_eigs(h0, 4)
    hg = w.graded(h0, kmax, 13)
    k = w.grading_exponents(n, kmax, 13)
    x0, _ = oracle.complex_inverse_iteration(h0, cz.real, cz.imag)
    xg, st = oracle.complex_inverse_iteration(hg, cz.real, cz.imag)
    assert (st & 1).all()
    for j in range(cz.size):
        assert np.abs(phase_normalise(np.ldexp(1.0, k) * x0[:, j]) - xg[:, j]).max() <= 1e-8


def test_protections_keep_results_finite():
    # strongly graded, shifts at eigenvalues: protections and perturbed pivots
    n = 200
    h0 = w.gaussian_hessenberg(n, 31)
    cz = complex_eigs(h0, 6)
    hg = w.graded(h0, 400, 31)
    x, st = oracle.complex_inverse_iteration(hg, cz.real, cz.imag)
    assert np.isfinite(x).all()
    assert residuals(hg, cz, x).max() <= 10 * n


def test_exact_eigenvalue_of_triangular_matrix():
    # H upper triangular with a complex-free spectrum and sigma = H(0,0) + 0i:
    # the pivot R(0,0) vanishes and is raised to smin
    n = 40
    h = np.triu(w.gaussian_hessenberg(n, 2))
    h[np.arange(1, n), np.arange(n - 1)] = 0.0
    x, st = oracle.complex_inverse_iteration(h, [h[0, 0]], [0.0])
    assert st[0] & 2 and st[0] & 1
    e0 = np.zeros(n)
    e0[0] = 1.0
    assert np.abs(x[:, 0] - e0).max() <= 1e-12


def test_restart_vector_matches_dense_solve():
    # a shift off the spectrum fails the growth test every time: with r
    # restarts the result is the solve from dlaein's (r-1)-th restart vector
    n = 90
    h = w.gaussian_hessenberg(n, 4)
    sig = np.array([40.0 + 7.0j])
    eps3 = np.linalg.norm(h, np.inf) * 2.0 ** -53
    sq = np.sqrt(n)
    for r in (1, 2, 5):
        x, st = oracle.complex_inverse_iteration(h, sig.real, sig.imag, max_restarts=r)
        assert st[0] & 1 == 0
        y0 = np.full(n, eps3 / (sq + 1.0))
        y0[0] = eps3
        y0[n - 1 - (r - 1)] -= eps3 * sq
        xd = np.linalg.solve(h - sig[0] * np.eye(n), y0.astype(complex))
        assert np.abs(phase_normalise(xd) - x[:, 0]).max() <= 1e-12
