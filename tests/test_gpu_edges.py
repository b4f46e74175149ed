"""Edge cases of the real-shift path on the GPU, each against the oracle and
against what the mathematics fixes.

* Triangular H with exact eigenvalue shifts: every rotation is the identity
  and the pivot of the shift's own row is exactly 0.  The reference raises it
  to smin and reports info = 1 (robust.c:462-465); here HSRQ_PERTURBED per
  shift, equal to the oracle's flag, and the vector equals the textbook
  back substitution of (H - mu I) with that pivot set to smin.
* Restarts that converge: the final vector is the solve from dlaein's
  restart vector (inverse_iteration.c:36-57) of the round that converged,
  entrywise against a dense LU solve with that right-hand side, and the shift
  moves behind the ones that converged at once (stable order).
* Padded leading dimensions (ldH > n, ldvr > n, as the reference driver pads
  them, main.c:76-77, utils.c:15-24): NaN-filled padding of H is never read,
  the padding of vr is never written, results equal the unpadded call bit for
  bit; also through the drop-in host entry point.
"""
import numpy as np
import pytest
import scipy.linalg as sla

import oracle
from hsrq3in_amd import HSRQ_CONVERGED, HSRQ_PERTURBED, colmajor_to_device
from tests.helpers import case
from tests.test_oracle import (RESTART_CASES, check_restart_vector, restart_candidates,
                               restart_problem)

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[1, 4], ids=["pack1", "pack4"])
def ctx(request):
    from hsrq3in_amd import Context
    c = Context(0, packing=request.param)
    yield c
    c.close()


def gpu_solve(ctx, H, wr, tlsz, ldH=None, ldvr=None, fill=0.0):
    n = H.shape[0]
    ldH = ldH or n
    ldvr = ldvr or n
    Ht = torch.full((n, ldH), np.nan, dtype=torch.float64, device="cuda")
    Ht[:, :n] = colmajor_to_device(H)
    wrt = torch.from_numpy(np.array(wr)).cuda()
    vrt = torch.full((len(wr), ldvr), fill, dtype=torch.float64, device="cuda")
    st = ctx.solve(Ht, wrt, vrt, tlsz)
    return vrt.cpu().numpy(), wrt.cpu().numpy(), st


@pytest.mark.parametrize("n,tile", [(150, 64), (300, 128), (260, 256)])
def test_triangular_exact_shifts_perturbed(ctx, n, tile):
    rng = np.random.default_rng(n)
    H = np.asfortranarray(np.triu(rng.standard_normal((n, n))))
    idx = np.array([0, 5, n // 2, n - 3, n - 1])
    exact = np.diag(H)[idx]
    near = np.diag(H)[[7, 11]] * (1 + 1e-12)
    wr = np.concatenate([exact, near])
    v, wo, st = gpu_solve(ctx, H, wr, tile)
    x = v.T
    xo, woo, sto, so, lsc = oracle.tiled_inverse_iteration_detail(H, wr, tile, 8)
    assert np.isfinite(x).all()
    assert np.array_equal(wo, woo) and np.array_equal(wo, wr)
    perturbed = (st & HSRQ_PERTURBED) != 0
    assert np.array_equal(perturbed, (so & 2) != 0)
    assert perturbed[:len(exact)].all() and not perturbed[len(exact):].any()
    assert (st & HSRQ_CONVERGED).all()
    eps3 = np.abs(H).sum(axis=1).max() * 2.0 ** -53
    smlnum = max(2 * np.finfo(float).tiny, np.finfo(float).tiny * n / 2.0 ** -53)
    for k, mu in enumerate(wr):
        A = H - mu * np.eye(n)
        smin = max(2.0 ** -53 * abs(mu), smlnum)
        d = np.diag(A).copy()
        d[np.abs(d) < smin] = smin
        np.fill_diagonal(A, d)
        ref = sla.solve_triangular(A, np.full(n, eps3))
        ref /= np.linalg.norm(ref)
        c = ref @ x[:, k]
        assert c > 0 and 1 - c <= 1e-10, (k, 1 - c)
        assert 1 - abs(xo[:, k] @ x[:, k]) <= 1e-10


def test_restart_converges_to_dlaein_solve(ctx):
    H, shifts = restart_problem()
    n = H.shape[0]
    # one problem: a shift that converges at once first, then the three
    # restart cases; the converged one stays in front, the others follow in
    # their order (each converges in its own round)
    ev = np.linalg.eigvals(H)
    lam = ev.real[ev.imag == 0]
    good = lam[np.argmax(np.abs(lam))] * (1 + 1e-14)
    wr = np.concatenate([shifts[::-1], [good]])
    v, wo, st = gpu_solve(ctx, H, wr, 64)
    xo, woo, sto, so, lsc = oracle.tiled_inverse_iteration_detail(H, wr, 64, 4)
    assert np.array_equal(wo, woo)
    assert wo[0] == good
    assert list(wo[1:]) == [shifts[0], shifts[1], shifts[2]]
    assert (st & HSRQ_CONVERGED).all() and sto.nonconverged == 0
    assert sto.restarts == 3
    x = v.T
    for (l0, d, rounds), s in zip(RESTART_CASES, shifts):
        k = int(np.nonzero(wo == s)[0][0])
        check_restart_vector(x[:, k], restart_candidates(H, s), rounds)
        assert np.abs(x[:, k] * np.sign(x[:, k] @ xo[:, k]) - xo[:, k]).max() <= 1e-14


@pytest.mark.parametrize("n,tile", [(333, 64), (1000, 128), (700, 256)])
def test_padded_leading_dimensions(ctx, n, tile):
    H, wr = case(n, seed=n + 5, m=20)
    v0, w0, s0 = gpu_solve(ctx, H, wr, tile)
    v1, w1, s1 = gpu_solve(ctx, H, wr, tile, ldH=n + 3, ldvr=n + 5, fill=7.0)
    assert (v1[:, n:] == 7.0).all()
    assert np.array_equal(v1[:, :n], v0)
    assert np.array_equal(w1, w0) and np.array_equal(s1, s0)


def test_dropin_padded_host_buffers():
    """The drop-in entry point on host buffers with ldH > n and ldvr > n: the
    padding rows of vr keep their sentinel, H's padding is never read (NaN),
    and the vectors equal the device call's."""
    import ctypes

    from hsrq3in_amd import Context, lib
    n, ld, ldv = 500, 509, 517
    H, wr = case(n, seed=41, m=16)
    Hp = np.full((ld, n), np.nan, order="F")
    Hp[:n, :] = H
    vr = np.full((ldv, len(wr)), 7.0, order="F")
    wr_io = wr.copy()
    dp = ctypes.POINTER(ctypes.c_double)
    lib().tiled_inverse_iteration(n, Hp.ctypes.data_as(dp), ld, len(wr), wr_io.ctypes.data_as(dp),
                                  64, 16, vr.ctypes.data_as(dp), ldv)
    assert (vr[n:, :] == 7.0).all()
    c = Context(0)
    v0, w0, _ = gpu_solve(c, H, wr, 64)
    c.close()
    assert np.array_equal(wr_io, w0)
    assert np.array_equal(vr[:n, :], v0.T)


@pytest.mark.parametrize("n,m,nc", [(300, 40, 12), (500, 60, 20)])
def test_many_restarts_against_oracle_and_sequential(ctx, n, m, nc):
    """A clustered spectrum whose first solve converges only 4 of the shifts:
    the other shifts need hundreds of restart rounds (the last at round n-1).
    The restart engine evaluates many rounds per launch; its vectors, shift
    order and status equal the one-round-at-a-time loop of the allocation-free
    entry point bit for bit, and the oracle's restart loop
    (inverse_iteration.c:160-328) to the bar (1e-10; cluster members 1e-8,
    their eigenvectors move by eps ||H|| / spacing)."""
    import ctypes

    from hsrq3in_amd import lib
    from hsrq3in_amd import workloads as w
    lam, shifts = w.cluster_spectrum(n, m, nc, 4)
    H = w.cluster_hessenberg(lam, nc, 4)
    v, wo, st = gpu_solve(ctx, H, shifts, 128)
    # the sequential loop (hsrq_solve_device_noalloc restarts one round at a time)
    Ht = colmajor_to_device(H)
    wrt = torch.from_numpy(shifts.copy()).cuda()
    vrt = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    st2 = np.zeros(m, dtype=np.int32)
    L = lib()
    torch.cuda.synchronize()
    assert L.hsrq_solve_device_noalloc(ctx.handle, n, Ht.data_ptr(), n, m, wrt.data_ptr(),
                                       vrt.data_ptr(), n, 128,
                                       st2.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    assert np.array_equal(st, st2)
    assert np.array_equal(wo, wrt.cpu().numpy())
    assert np.array_equal(v, vrt.cpu().numpy())
    # the oracle
    xo, woo, sto, so, lsc = oracle.tiled_inverse_iteration_detail(H, shifts, 128, 16)
    assert sto.restarts >= n - 2
    assert np.array_equal(wo, woo)
    assert np.array_equal(st & (HSRQ_CONVERGED | HSRQ_PERTURBED), so & 3)
    x = v.T
    conv = (st & HSRQ_CONVERGED) != 0
    d = 1 - np.abs(np.sum(x[:, conv] * xo[:, conv], axis=0))
    member = np.isin(wo[conv], shifts[:nc])
    print(f"n={n}: {conv.sum()}/{m} converged after {sto.restarts} rounds; 1-|cos| cluster "
          f"{d[member].max() if member.any() else 0:.2e}, others "
          f"{d[~member].max() if (~member).any() else 0:.2e}")
    assert (d[~member] <= 1e-10).all() and (d[member] <= 1e-8).all()


@pytest.mark.parametrize("n,tile", [(300, 64), (700, 128), (600, 256)])
def test_below_subdiagonal_never_read(ctx, n, tile):
    """NaN in every entry below the subdiagonal: the solve reads only the
    Hessenberg part (results bit for bit those of the zero-filled matrix), so
    the drop-in's Hessenberg-only H2D copy is exact."""
    H, wr = case(n, seed=n + 77, m=16, kmax=30)
    v0, w0, s0 = gpu_solve(ctx, H, wr, tile)
    Hn = H.copy(order="F")
    i, j = np.tril_indices(n, -2)
    Hn[i, j] = np.nan
    v1, w1, s1 = gpu_solve(ctx, Hn, wr, tile)
    assert np.array_equal(v0, v1) and np.array_equal(w0, w1) and np.array_equal(s0, s1)
