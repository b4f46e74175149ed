"""Device problem generators (hsrq_gen.hip through the C-ABI) against their
CPU restatements.

Bars: the constructed-spectrum build and the grading are bit-exact with the
numpy builds (same IEEE operations, no contraction, exact power-of-two
scaling); the Gaussian generator's integer stream is exact by construction
and its normals agree with numpy's log / sqrt / sin / cos to 8 ulp (the two
math libraries round differently in the last bits).
"""
import numpy as np
import pytest

import oracle
from hsrq3in_amd import HSRQ_CONVERGED, HSRQ_ERANGE, HsrqError, colmajor_to_host, workloads
from oracle import generators as gen

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ULP_TOL = 8


@pytest.fixture(scope="module")
def ctx():
    from hsrq3in_amd import Context
    c = Context(0)
    yield c
    c.close()


def within_ulps(a, b, k):
    return np.abs(a - b) <= k * np.spacing(np.abs(b))


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (255, 7), (777, (1 << 63) + 12345),
                                    (1024, 2**32 + 1)])
def test_gaussian_matches_restatement(ctx, n, seed):
    Ht = ctx.gaussian_hessenberg(n, seed)
    h = colmajor_to_host(Ht)
    ref = gen.gaussian_hessenberg(n, seed)
    assert np.array_equal(h == 0.0, ref == 0.0)
    ok = within_ulps(h, ref, ULP_TOL)
    assert ok.all(), (np.argwhere(~ok)[:5], h[~ok][:5], ref[~ok][:5])


def test_gaussian_deterministic_and_seeded(ctx):
    a = colmajor_to_host(ctx.gaussian_hessenberg(300, 5))
    b = colmajor_to_host(ctx.gaussian_hessenberg(300, 5))
    c = colmajor_to_host(ctx.gaussian_hessenberg(300, 6))
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_gaussian_statistics_at_size(ctx):
    # n = 4000: 8M entries; the Hessenberg part is N(0, 1)
    n = 4000
    Ht = ctx.gaussian_hessenberg(n, 99)
    ii = torch.arange(n, device="cuda")
    mask = ii[None, :] <= ii[:, None] + 1          # Ht[j, i]: i <= j + 1
    x = Ht[mask]
    assert torch.isfinite(x).all()
    assert float(x.mean().abs()) < 2e-3 and abs(float(x.var()) - 1.0) < 2e-3
    assert float(Ht[~mask].abs().max()) == 0.0


@pytest.mark.parametrize("n", [1, 2, 3, 64, 65, 513])
def test_constructed_bit_exact(ctx, n):
    lam = np.linspace(-np.sqrt(n), np.sqrt(n), n) + 0.25
    nu, d, u = workloads.constructed_parameters(lam, seed=n)
    h = colmajor_to_host(ctx.constructed_hessenberg(lam, nu, d, u))
    ref = workloads.constructed_from_parameters(lam, nu, d, u)
    assert np.array_equal(h.view(np.uint64), ref.view(np.uint64))


def test_constructed_cluster_bit_exact(ctx):
    n, m, nc = 1500, 50, 20
    lam, _ = workloads.cluster_spectrum(n, m, nc, seed=3)
    pos = workloads.cluster_positions(n, nc)
    lam_c = float(np.max(np.abs(lam[pos])))
    nu, d, u = workloads.constructed_parameters(lam, 3, sub=1.0 / n, weak=pos,
                                                weak_nu=0.1 * 1e-8 * lam_c)
    h = colmajor_to_host(ctx.constructed_hessenberg(lam, nu, d, u))
    assert np.array_equal(h, workloads.cluster_hessenberg(lam, nc, 3))


def test_grade_bit_exact_and_range(ctx):
    n = 600
    Ht = ctx.gaussian_hessenberg(n, 11)
    h0 = colmajor_to_host(Ht)
    k = workloads.grading_exponents(n, 200, seed=11).astype(np.int32)
    emax, emin = ctx.grade(Ht, k, exact=True)
    ref = np.ldexp(h0, (k[:, None] - k[None, :]))
    ref = np.where(np.tril(np.ones((n, n), bool), -2), h0, ref)
    assert np.array_equal(colmajor_to_host(Ht), ref)
    e = np.frexp(ref[ref != 0.0])[1]
    assert (emax, emin) == (e.max(), e.min())


def test_grade_subnormal_results_round_like_numpy(ctx):
    n = 200
    Ht = ctx.gaussian_hessenberg(n, 12)
    h0 = colmajor_to_host(Ht)
    # nondecreasing k: entries on and above the diagonal are scaled down (rows
    # of the top half by 2^-1060 in the last columns, into the subnormal range),
    # the subdiagonal up by at most 2^20
    k = np.minimum(0, -1060 + 20 * np.maximum(0, np.arange(n) - n // 2)).astype(np.int32)
    with pytest.raises(HsrqError) as ei:
        ctx.grade(Ht, k, exact=True)
    assert ei.value.code == HSRQ_ERANGE
    assert np.array_equal(colmajor_to_host(Ht), h0)          # untouched
    emax, emin = ctx.grade(Ht, k, exact=False)
    assert emin < -1021
    ref = np.ldexp(h0, (k[:, None] - k[None, :]))
    ref = np.where(np.tril(np.ones((n, n), bool), -2), h0, ref)
    assert np.array_equal(colmajor_to_host(Ht).view(np.uint64), ref.view(np.uint64))


def test_grade_subnormal_inputs_scaled_up_are_exact(ctx):
    n = 64
    Ht = ctx.gaussian_hessenberg(n, 14)
    Ht[40, 3] = 5e-324                  # H(3, 40): the smallest subnormal
    Ht[50, 7] = -3.0e-310
    h0 = colmajor_to_host(Ht)
    k = np.zeros(n, dtype=np.int32)
    k[3] = 60                           # row 3 up by 2^60, column 3 down
    k[7] = 60
    emax, emin = ctx.grade(Ht, k, exact=True)
    ref = np.ldexp(h0, (k[:, None] - k[None, :]))
    assert np.array_equal(colmajor_to_host(Ht), ref)
    assert emin == np.frexp(5e-324)[1] + 60


def test_grade_overflow_rejected(ctx):
    n = 100
    Ht = ctx.gaussian_hessenberg(n, 13)
    h0 = colmajor_to_host(Ht)
    k = np.zeros(n, dtype=np.int32)
    k[10] = 1030
    with pytest.raises(HsrqError) as ei:
        ctx.grade(Ht, k)
    assert ei.value.code == HSRQ_ERANGE
    assert np.array_equal(colmajor_to_host(Ht), h0)


def _solve(ctx, Ht, wr):
    m, n = wr.size, Ht.shape[0]
    wrt = torch.from_numpy(wr.copy()).cuda()
    vrt = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    st = ctx.solve(Ht, wrt, vrt, 128)
    res = ctx.residuals(Ht, wrt, vrt)
    assert (st & HSRQ_CONVERGED).all()
    assert np.isfinite(res).all() and res.max() <= 10 * n
    return np.asfortranarray(vrt.cpu().numpy().T), wrt.cpu().numpy()


def test_generated_problem_solves(ctx):
    # generate -> solve -> validate, then grade -> solve -> validate, in HBM
    n, m = 900, 40
    lam = np.arange(1.0, n + 1.0)
    nu, d, u = workloads.constructed_parameters(lam, seed=21)
    Ht = ctx.constructed_hessenberg(lam, nu, d, u)
    wr = lam[:m] * (1.0 + 1e-12)
    x, wo = _solve(ctx, Ht, wr)
    # the oracle on the host copy of the generated matrix
    xo, woo, sto = oracle.tiled_inverse_iteration(colmajor_to_host(Ht), wr, 128, 8)
    assert np.array_equal(wo, woo) and sto.nonconverged == 0
    c = np.abs(np.sum(x * xo, axis=0)) / (np.linalg.norm(x, axis=0) * np.linalg.norm(xo, axis=0))
    assert (1.0 - c).max() <= 1e-10
    k = workloads.grading_exponents(n, 30, seed=21).astype(np.int32)
    # far above the diagonal this H holds subnormal entries (U^-1 decays
    # geometrically), which the downward scalings round: not exact
    with pytest.raises(HsrqError):
        ctx.grade(Ht, k, exact=True)
    ctx.grade(Ht, k)
    # graded by up to 2^60, ||H|| eps exceeds the eigenvalue gaps: the
    # eigenvectors are not determined to working accuracy by a normwise
    # backward-stable solve, so the bar is the residual (and no Inf / NaN)
    _solve(ctx, Ht, wr)
