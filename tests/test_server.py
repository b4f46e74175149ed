"""Continuous-batching server (hsrq_server_*, hsrq3in_amd.serve).

Bar: a request's results through the server are bit-identical to the plain
device solve of the same problem (same tile), alone or in any mix of
concurrently running requests (tolerance: exact equality, since every
kernel's arithmetic for a problem is independent of what shares a launch);
every request completes when the HBM arena admits only a few at a time."""
import threading
import time

import numpy as np
import pytest

from hsrq3in_amd import HSRQ_CONVERGED, HSRQ_VALIDATED, HsrqError
from tests.helpers import case

torch = pytest.importorskip("torch")


def test_server_rejects_invalid_without_device():
    import ctypes
    from hsrq3in_amd import serve
    L = serve._bind()
    h = ctypes.c_void_p()
    assert L.hsrq_server_create(None, 1 << 20, 4, 0, ctypes.byref(h)) == -1
    assert L.hsrq_server_poll(None, 1) == -1
    assert L.hsrq_server_wait(None, 1, None, None, None) == -1


def _tile(n):
    return 128


def _plain(ctx, H, wr):
    from hsrq3in_amd import colmajor_to_device
    n = H.shape[0]
    Ht = colmajor_to_device(H)
    wrt = torch.from_numpy(wr.copy()).cuda()
    vrt = torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")
    st = ctx.solve(Ht, wrt, vrt, _tile(n))
    return np.asfortranarray(vrt.cpu().numpy().T), wrt.cpu().numpy(), st


CASES = [(150, 11, 0), (400, 25, 0), (1100, 30, 0), (257, 9, 300), (90, 40, 0), (700, 12, 200)]


@pytest.fixture(scope="module")
def ctx():
    from hsrq3in_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def problems():
    return [case(n, seed=500 + i, m=m, kmax=k) for i, (n, m, k) in enumerate(CASES)]


@pytest.mark.gpu
def test_server_equals_plain_solve(ctx, problems):
    from hsrq3in_amd.serve import Server
    srv = Server(ctx, hbm_bytes=2 << 30)
    ids = [srv.submit(H, wr) for H, wr in problems]
    for (H, wr), rid in zip(problems, ids):
        vr, wo, st, ms = srv.wait(rid)
        xp, wop, stp = _plain(ctx, H, wr)
        assert np.array_equal(wo, wop)
        assert np.array_equal(st & ~HSRQ_VALIDATED, stp)
        assert np.array_equal(vr, xp)
        assert (st & HSRQ_CONVERGED).all()
        # validated online: the device residual of every vector <= 10 n
        assert (st & HSRQ_VALIDATED).all()
        assert ms > 0
    s = srv.stats()
    assert s["completed"] == len(problems) and s["active"] == 0 and s["queued"] == 0
    assert s["p99_ms"] >= s["p50_ms"] > 0
    assert s["validated"] == sum(len(wr) for _, wr in problems)
    assert 0.0 < s["worst_residual"] <= 1.0
    srv.close()


@pytest.mark.gpu
def test_server_alone_vs_mix(ctx, problems):
    """Each request alone, then all of them arriving while others run."""
    from hsrq3in_amd.serve import Server
    alone = []
    srv = Server(ctx, hbm_bytes=2 << 30)
    for H, wr in problems:
        alone.append(srv.wait(srv.submit(H, wr)))
    srv.close()
    srv = Server(ctx, hbm_bytes=2 << 30)
    ids = []
    for H, wr in problems:      # staggered arrivals overlap with running work
        ids.append(srv.submit(H, wr))
        time.sleep(0.002)
    mixed = [srv.wait(r) for r in ids]
    assert srv.stats()["ticks"] > 0
    srv.close()
    for a, b in zip(alone, mixed):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2])


@pytest.mark.gpu
def test_server_admission_by_footprint(ctx):
    """An arena that holds about two requests: all eight still complete, the
    peak never exceeds the arena, requests had to queue."""
    from hsrq3in_amd.serve import Server
    probs = [case(600, seed=900 + i, m=20) for i in range(8)]
    # footprint of one host request: H + wr + X + ~32 n m workspace
    one = 600 * 600 * 8 + 20 * 8 + 600 * 20 * 8 + 40 * 600 * 20 * 8
    srv = Server(ctx, hbm_bytes=int(2.5 * one))
    ids = [srv.submit(H, wr) for H, wr in probs]
    queued_seen = srv.stats()["queued"]
    outs = []
    done = []

    def client(rid, k):
        outs.append((k, srv.wait(rid)))
        done.append(k)

    th = [threading.Thread(target=client, args=(r, k)) for k, r in enumerate(ids)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert len(done) == 8
    s = srv.stats()
    assert s["hbm_peak"] <= s["hbm_bytes"]
    assert queued_seen > 0
    srv.close()
    for k, (vr, wo, st, ms) in outs:
        H, wr = probs[k]
        xp, wop, _ = _plain(ctx, H, wr)
        assert np.array_equal(vr, xp)


@pytest.mark.gpu
def test_server_restarts_like_plain(ctx):
    from hsrq3in_amd.serve import Server
    H, wr = case(12, seed=3, pert=4e-16)
    wr = np.sort(wr)
    mixed = np.array([wr[0], 0.5 * (wr[0] + wr[1]), wr[1], 0.5 * (wr[1] + wr[2])])
    srv = Server(ctx, hbm_bytes=64 << 20)
    vr, wo, st, _ = srv.wait(srv.submit(H, mixed))
    srv.close()
    xp, wop, stp = _plain(ctx, H, mixed)
    assert np.array_equal(wo, wop) and np.array_equal(st & ~HSRQ_VALIDATED, stp)
    assert np.array_equal(vr, xp)
    # the two shifts between eigenvalues never converge and fail validation
    conv = (st & HSRQ_CONVERGED) != 0
    assert np.array_equal((st & HSRQ_VALIDATED) != 0, conv)


@pytest.mark.gpu
def test_server_device_requests(ctx, problems):
    from hsrq3in_amd import colmajor_to_device
    from hsrq3in_amd.serve import Server
    srv = Server(ctx, hbm_bytes=1 << 30)
    items, ids = [], []
    for H, wr in problems[:3]:
        n = H.shape[0]
        it = (colmajor_to_device(H), torch.from_numpy(wr.copy()).cuda(),
              torch.zeros((len(wr), n), dtype=torch.float64, device="cuda"))
        items.append(it)
        ids.append(srv.submit_device(*it))
    for (H, wr), it, rid in zip(problems[:3], items, ids):
        _, wo, st, _ = srv.wait(rid)
        xp, wop, _ = _plain(ctx, H, wr)
        assert np.array_equal(np.asfortranarray(it[2].cpu().numpy().T), xp)
        assert np.array_equal(it[1].cpu().numpy(), wop) and np.array_equal(wo, wop)
    srv.close()


@pytest.mark.gpu
def test_server_rejects_oversized(ctx):
    from hsrq3in_amd.serve import Server
    srv = Server(ctx, hbm_bytes=1 << 20)
    H, wr = case(600, seed=1, m=5)       # H alone is 2.9 MB
    with pytest.raises(HsrqError):
        srv.submit(H, wr)
    srv.close()


@pytest.mark.gpu
def test_request_without_shifts_alone(ctx):
    """A request with m = 0 completes even when it is the only traffic (its
    completion happens outside any tick's retirement and must still wake the
    waiting client)."""
    from hsrq3in_amd.serve import Server
    srv = Server(ctx, hbm_bytes=64 << 20, max_active=4)
    try:
        H, _ = case(100, seed=77, m=1)
        rid = srv.submit(H, np.zeros(0))
        done = {}

        def waiter():
            done["r"] = srv.wait(rid)

        t = threading.Thread(target=waiter, daemon=True)
        t.start()
        t.join(timeout=30.0)
        assert not t.is_alive(), "wait() for an m = 0 request did not return"
        vr, wr, st, _ = done["r"]
        assert vr.shape == (100, 0) and wr.size == 0 and st.size == 0
    finally:
        srv.close()
