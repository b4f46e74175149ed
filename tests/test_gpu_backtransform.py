"""Back-transformation V = Z X (hsrq_backtransform_device through the C-ABI).

Bars: the FP64 GEMM error bound |V - Z X| <= gamma |Z| |X| elementwise with
gamma = 2 n eps against a host product; eigenvectors of A = Z H Z^T
(scipy's Hessenberg reduction) with residual <= 10 n eps and
1 - |cos| <= 1e-10 against the oracle's vectors transformed on the host; at
full size, the column norms (Z orthogonal) and the checksum e^T V = (e^T Z) X.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from hsrq3in_amd import HSRQ_CONVERGED, HsrqError, colmajor_to_device, colmajor_to_host

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

EPS = 2.0 ** -52


@pytest.fixture(scope="module")
def ctx():
    from hsrq3in_amd import Context
    c = Context(0)
    yield c
    c.close()


def rows(x: np.ndarray):
    """n x m column-major host matrix -> (m, n) device tensor, rows = columns."""
    return torch.from_numpy(np.ascontiguousarray(x.T)).cuda()


@pytest.mark.parametrize("n,m,ldz", [(1, 1, 1), (37, 5, 40), (500, 64, 500), (513, 130, 520)])
def test_gemm_bound(ctx, n, m, ldz):
    rng = np.random.default_rng(n)
    z = rng.standard_normal((n, n))
    x = rng.standard_normal((n, m))
    Zt = torch.zeros((n, ldz), dtype=torch.float64, device="cuda")
    Zt[:, :n] = torch.from_numpy(z.T.copy()).cuda()
    Vt = torch.full((m, n + 3), np.nan, dtype=torch.float64, device="cuda")
    ctx.backtransform(Zt, rows(x), Vt)
    vt = Vt.cpu().numpy()
    assert np.isnan(vt[:, n:]).all()                 # padding untouched
    v = vt[:, :n].T
    ref = z @ x
    bound = 2 * n * EPS * (np.abs(z) @ np.abs(x))
    assert (np.abs(v - ref) <= bound).all()


def test_rejects_overlap_and_bad_sizes(ctx):
    from hsrq3in_amd import lib
    n, m = 64, 8
    Z = torch.eye(n, dtype=torch.float64, device="cuda")
    X = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    L = lib()
    h = ctx.handle
    assert L.hsrq_backtransform_device(h, n, m, Z.data_ptr(), n, X.data_ptr(), n, X.data_ptr(),
                                       n) == -1           # V aliases X
    assert L.hsrq_backtransform_device(h, n, m, Z.data_ptr(), n, X.data_ptr(), n,
                                       Z.data_ptr() + 8 * n, n) == -1   # V inside Z
    assert L.hsrq_backtransform_device(h, n, m, Z.data_ptr(), n - 1, X.data_ptr(), n,
                                       X.data_ptr(), n) == -1
    assert L.hsrq_backtransform_device(h, n, 0, Z.data_ptr(), n, None, n, None, n) == 0
    with pytest.raises(HsrqError):
        ctx.backtransform(Z, X, X)


def test_eigenvectors_of_the_unreduced_matrix(ctx):
    n, m = 600, 30
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n))
    h, q = sl.hessenberg(a, calc_q=True)            # a = q h q^T
    h = np.asfortranarray(np.triu(h, -1))
    w = np.linalg.eigvals(h)
    wr = np.ascontiguousarray(w.real[w.imag == 0.0][:m] * (1.0 + 1e-12))
    m = wr.size
    Ht = colmajor_to_device(h)
    wrt = torch.from_numpy(wr.copy()).cuda()
    Xt = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    st = ctx.solve(Ht, wrt, Xt, 128)
    assert (st & HSRQ_CONVERGED).all()
    Vt = ctx.backtransform(colmajor_to_device(q), Xt)
    v = Vt.cpu().numpy().T
    lam = wrt.cpu().numpy()
    r = np.linalg.norm(a @ v - v * lam, axis=0)
    res = r / (np.linalg.norm(a, np.inf) * np.linalg.norm(v, axis=0) * EPS)
    assert np.isfinite(v).all() and res.max() <= 10 * n
    xo, wo, sto = oracle.tiled_inverse_iteration(h, wr, 128, 8)
    assert np.array_equal(wo, lam)
    vo = q @ xo
    c = np.abs(np.sum(v * vo, axis=0)) / (np.linalg.norm(v, axis=0) * np.linalg.norm(vo, axis=0))
    assert (1.0 - c).max() <= 1e-10


def test_full_size_invariants(ctx):
    # n = 8000, m = 800: Z = I - 2 w w^T (orthogonal), X random
    n, m = 8000, 800
    g = torch.Generator(device="cuda").manual_seed(3)
    wv = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    wv /= wv.norm()
    Zt = torch.eye(n, dtype=torch.float64, device="cuda") - 2.0 * torch.outer(wv, wv)  # symmetric
    Xt = torch.randn((m, n), dtype=torch.float64, device="cuda", generator=g)
    Vt = ctx.backtransform(Zt, Xt)
    assert torch.isfinite(Vt).all()
    rel = (Vt.norm(dim=1) / Xt.norm(dim=1) - 1.0).abs().max().item()
    assert rel <= 1e-12
    # checksum: e^T V = (e^T Z) X, e^T Z = e^T - 2 (e^T w) w^T
    ez = 1.0 - 2.0 * wv.sum() * wv
    lhs = Vt.sum(dim=1)
    rhs = Xt @ ez
    scale = (Xt.abs() @ ez.abs()) + Vt.abs().sum(dim=1)
    assert ((lhs - rhs).abs() <= 4 * n * EPS * scale).all()
    # and the explicit reflection, v = x - 2 (w^T x) w
    ref = Xt - 2.0 * torch.outer(Xt @ wv, wv)
    assert ((Vt - ref).abs().max() / Xt.abs().max()).item() <= 100 * EPS * np.sqrt(n)
