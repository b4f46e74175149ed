"""Complex shifts on the device (hsrq_solve_complex_device through the C-ABI)
against the oracle's complex-shift solve.

Bar: both vectors unit 2-norm with the first entry of largest modulus real
positive; max |x_gpu - x_oracle| <= 1e-10 (1e-8 on graded input), the same
status, residual ||(H - sigma I) x|| / (||H||_oo ||x|| eps) <= 10 n for the
converged shifts, no Inf/NaN.
"""
import functools

import numpy as np
import pytest

import oracle
from hsrq3in_amd import HSRQ_CONVERGED, colmajor_to_device
from hsrq3in_amd import workloads as w

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

EPS = 2.0 ** -52
TOL = 1e-10


@pytest.fixture(scope="module")
def ctx():
    from hsrq3in_amd import Context
    c = Context(0)
    yield c
    c.close()


def gpu_complex(ctx, h, sig, max_restarts=0):
    Ht = colmajor_to_device(h)
    wrt = torch.from_numpy(np.ascontiguousarray(sig.real)).cuda()
    wit = torch.from_numpy(np.ascontiguousarray(sig.imag)).cuda()
    vrt, st = ctx.solve_complex(Ht, wrt, wit, max_restarts=max_restarts)
    v = vrt.cpu().numpy()
    return (v[0::2] + 1j * v[1::2]).T, st


def residuals(h, sig, x):
    r = h @ x - x * sig
    return np.linalg.norm(r, axis=0) / (np.linalg.norm(h, np.inf) * np.linalg.norm(x, axis=0) * EPS)


@functools.lru_cache(maxsize=None)
def eig_case(n, seed, count):
    h = w.gaussian_hessenberg(n, seed)
    ev = np.linalg.eigvals(h)
    cz = ev[ev.imag > 0][:count]
    return h, cz * (1 + 1e-12)


@pytest.mark.parametrize("n,seed", [(2, 0), (17, 2), (200, 4), (511, 5), (513, 6), (1500, 7)])
def test_parity_complex_eigenvalues(ctx, n, seed):
    h, sig = eig_case(n, seed, 24)
    x, st = gpu_complex(ctx, h, sig)
    xo, sto = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    assert np.isfinite(x).all()
    assert np.array_equal(st, sto)
    assert np.abs(x - xo).max() <= TOL
    conv = (st & HSRQ_CONVERGED) == HSRQ_CONVERGED
    assert (residuals(h, sig, x)[conv] <= 10 * n).all()
    if n >= 200:
        assert conv.all()


@pytest.mark.parametrize("n", [1, 1025, 2049, 3000, 4096, 4097, 5000, 8192, 8193, 10000])
def test_parity_far_shifts(ctx, n):
    # sizes for every layout (512 threads with one row each; 1024 threads
    # with 1, 2, 4, 8 rows each) and the global-memory kernel above 8192;
    # shifts off the spectrum, so no eigenvalue computation is needed
    h = w.gaussian_hessenberg(n, n)
    rng = np.random.default_rng(n)
    sig = rng.uniform(-3, 3, 6) * np.sqrt(n) + 1j * rng.uniform(0.1, 2.0, 6) * np.sqrt(n)
    x, st = gpu_complex(ctx, h, sig)
    xo, sto = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    assert np.isfinite(x).all()
    assert np.array_equal(st, sto)
    assert np.abs(x - xo).max() <= TOL


@pytest.mark.parametrize("n", [700, 1500])
def test_parity_more_shifts_than_cus(ctx, n):
    # more shifts than compute units: up to n = 2048 the 512-thread register
    # layout (two workgroups per CU) runs instead of the 1024-thread one
    m = torch.cuda.get_device_properties(0).multi_processor_count + 40
    h = w.gaussian_hessenberg(n, n + 1)
    rng = np.random.default_rng(n)
    sig = rng.uniform(-3, 3, m) * np.sqrt(n) + 1j * rng.uniform(0.1, 2.0, m) * np.sqrt(n)
    x, st = gpu_complex(ctx, h, sig)
    xo, sto = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    assert np.isfinite(x).all()
    assert np.array_equal(st, sto)
    assert np.abs(x - xo).max() <= TOL
    # the same shifts a few at a time (the 1024-thread layout): the sweeps are
    # the same arithmetic, only the norm's sum is grouped over 16 wavefronts
    # instead of 8
    x8, st8 = gpu_complex(ctx, h, sig[:8])
    assert np.array_equal(st8, st[:8])
    assert np.abs(x8 - x[:, :8]).max() <= 1e-14


def test_conjugate_shifts_conjugate_vectors(ctx):
    h, sig = eig_case(200, 4, 24)
    x, _ = gpu_complex(ctx, h, sig)
    xc, _ = gpu_complex(ctx, h, np.conj(sig))
    assert np.array_equal(xc, np.conj(x))


def test_real_shifts_match_real_path(ctx):
    n = 700
    h = w.gaussian_hessenberg(n, 12)
    ev = np.linalg.eigvals(h)
    wr = ev.real[ev.imag == 0][:8] * (1 + 1e-12)
    x, st = gpu_complex(ctx, h, wr + 0j)
    assert np.all(x.imag == 0.0) and (st & HSRQ_CONVERGED).all()
    Ht = colmajor_to_device(h)
    wrt = torch.from_numpy(wr.copy()).cuda()
    vrt = torch.zeros((wr.size, n), dtype=torch.float64, device="cuda")
    ctx.solve(Ht, wrt, vrt, 128)
    xr = vrt.cpu().numpy().T
    c = np.abs(np.sum(x.real * xr, axis=0)) / np.linalg.norm(xr, axis=0)
    assert (1.0 - c).max() <= TOL


@pytest.mark.parametrize("n", [400, 1500])
def test_graded(ctx, n):
    # both kinds of layout: 512 threads (n <= 512), 1024 threads with y in LDS
    h0 = w.gaussian_hessenberg(n, 31)
    ev = np.linalg.eigvals(h0)
    sig = ev[ev.imag > 0][:12]
    h = w.graded(h0, 400, 31)
    x, st = gpu_complex(ctx, h, sig)
    xo, sto = oracle.complex_inverse_iteration(h, sig.real, sig.imag)
    assert np.isfinite(x).all()
    assert np.array_equal(st, sto)
    assert np.abs(x - xo).max() <= 1e-8
    assert residuals(h, sig, x).max() <= 10 * n


def test_rejects_and_empty(ctx):
    from hsrq3in_amd import lib
    L = lib()
    H = torch.zeros((1, 1), dtype=torch.float64, device="cuda")
    assert L.hsrq_solve_complex_device(ctx.handle, 5, H.data_ptr(), 5, 0, None, None, None, 5,
                                       0, None) == 0
    # leading dimension below n; negative restart count
    assert L.hsrq_solve_complex_device(ctx.handle, 5, H.data_ptr(), 4, 1, H.data_ptr(),
                                       H.data_ptr(), H.data_ptr(), 5, 0, None) == -1
    assert L.hsrq_solve_complex_device(ctx.handle, 1, H.data_ptr(), 1, 1, H.data_ptr(),
                                       H.data_ptr(), H.data_ptr(), 1, -1, None) == -1


@pytest.mark.parametrize("n", [300, 1300])
def test_leading_dimensions(ctx, n):
    # ldH > n and ldvr > n: the padding is neither read as data nor written
    ld, ldv = n + 7, n + 11
    h, sig = eig_case(n, 4 + 300, 8)
    x, st = gpu_complex(ctx, h, sig)
    Ht = torch.full((n, ld), np.nan, dtype=torch.float64, device="cuda")
    Ht[:, :n] = torch.from_numpy(np.ascontiguousarray(h.T)).cuda()
    vrt = torch.full((2 * sig.size, ldv), 7.0, dtype=torch.float64, device="cuda")
    wrt = torch.from_numpy(np.ascontiguousarray(sig.real)).cuda()
    wit = torch.from_numpy(np.ascontiguousarray(sig.imag)).cuda()
    vrt, st2 = ctx.solve_complex(Ht, wrt, wit, vrt)
    v = vrt.cpu().numpy()
    assert (v[:, n:] == 7.0).all()
    assert np.array_equal((v[0::2, :n] + 1j * v[1::2, :n]).T, x)
    assert np.array_equal(st, st2)


@pytest.mark.parametrize("n", [8000, 9000])
def test_known_spectrum_both_kernels(ctx, n):
    # a prescribed real spectrum (no eigenvalue computation): real shifts as
    # complex ones (wi = 0) on either side of the 8192 boundary equal the
    # real path's vectors, and converge
    from hsrq3in_amd import workloads as wl
    lam = np.arange(1.0, n + 1.0)
    h = wl.constructed_spectrum(lam, seed=n)
    wr = lam[[0, 1, n // 2, n - 1]] * (1 + 1e-12)
    x, st = gpu_complex(ctx, h, wr + 0j)
    assert (st & HSRQ_CONVERGED).all() and np.all(x.imag == 0.0)
    Ht = colmajor_to_device(h)
    wrt = torch.from_numpy(wr.copy()).cuda()
    vrt = torch.zeros((wr.size, n), dtype=torch.float64, device="cuda")
    ctx.solve(Ht, wrt, vrt, 128)
    xr = vrt.cpu().numpy().T
    c = np.abs(np.sum(x.real * xr, axis=0)) / np.linalg.norm(xr, axis=0)
    assert (1.0 - c).max() <= TOL


@pytest.mark.parametrize("n", [40, 700, 9000])
def test_restart_vectors(ctx, n):
    # shifts off the spectrum never pass the growth test: after r restarts the
    # vector is the solve from the (r-1)-th restart vector, in both codes;
    # converged shifts are not solved again
    h = w.gaussian_hessenberg(n, 77)
    rng = np.random.default_rng(3)
    far = rng.uniform(1, 2, 3) * np.sqrt(n) * 3 + 1j * rng.uniform(0.1, 1, 3) * np.sqrt(n)
    if n <= 700:
        ev = np.linalg.eigvals(h)
        near = ev[ev.imag > 0][:3] * (1 + 1e-13)
        sig = np.concatenate([far, near])
    else:
        sig = far
    x, st = gpu_complex(ctx, h, sig, max_restarts=3)
    xo, sto = oracle.complex_inverse_iteration(h, sig.real, sig.imag, max_restarts=3)
    assert np.array_equal(st, sto)
    assert np.abs(x - xo).max() <= TOL
    x0, _ = gpu_complex(ctx, h, sig)
    conv = (st & HSRQ_CONVERGED) == HSRQ_CONVERGED
    assert np.array_equal(x[:, conv], x0[:, conv])
    assert not np.array_equal(x[:, ~conv], x0[:, ~conv])
