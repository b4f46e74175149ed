"""SKIPPED: the persistent schedule these tests cover was taken out of the
library.  Its first GPU run (the first case below, n = 300, m = 20, tile 64)
did not return within 120 s, the cause was not found by reading the code or
by a CPU emulation of its work queues, so the schedule is not in the product
and these tests stay skipped.

The persistent schedule (hsrq_context_set_persistent): one launch in which
workgroups take diagonal-sweep and panel-pass items from two queues and wait
for each item's inputs on counters, against the launch schedule.

Both run the same kernel bodies on the same inputs, so the vectors, shift
order and status are bit-identical, for every tile size, queue lag, with
restarts (hsrq_solve_device), as plans, as HIP-graph replays, with several
problems in one lane, and on graded inputs; the oracle check is the usual
bar (1 - |cos| <= 1e-10).
"""
import numpy as np
import pytest

import oracle
from hsrq3in_amd import HSRQ_CONVERGED, colmajor_to_device
from tests.helpers import case, one_minus_abs_cos

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="persistent schedule removed: its first GPU run hung "
                                      "(> 120 s) and the cause is not found")]


def solve(ctx, H, wr, tile):
    n = H.shape[0]
    Ht = colmajor_to_device(H)
    wrt = torch.from_numpy(np.array(wr)).cuda()
    vrt = torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")
    st = ctx.solve(Ht, wrt, vrt, tile)
    return vrt.cpu().numpy(), wrt.cpu().numpy(), st


def contexts(lag):
    from hsrq3in_amd import Context
    return Context(0, packing=1), Context(0, packing=1, persistent=1, lag=lag)


@pytest.mark.parametrize("n,m,tile,kmax,lag", [(300, 20, 64, 0, 0), (1000, 60, 128, 0, 0),
                                               (1000, 60, 128, 0, 1000), (700, 30, 256, 0, 500),
                                               (1500, 100, 128, 400, 1900), (129, 9, 64, 0, 0)])
def test_persistent_equals_launch_schedule(n, m, tile, kmax, lag):
    H, wr = case(n, seed=n + m + kmax, m=m, kmax=kmax)
    a, b = contexts(lag)
    try:
        x0, w0, s0 = solve(a, H, wr, tile)
        x1, w1, s1 = solve(b, H, wr, tile)
    finally:
        a.close()
        b.close()
    assert np.array_equal(w0, w1) and np.array_equal(s0, s1)
    assert np.array_equal(x0, x1)
    assert (s1 & HSRQ_CONVERGED).all()
    xo, wo, _ = oracle.tiled_inverse_iteration(H, wr, tile, 8)
    assert np.array_equal(w1, wo)
    assert one_minus_abs_cos(x1.T, xo).max() <= 1e-10


def test_persistent_plan_graph_replays():
    """Several problems of different orders in one lane (each at its own step
    offset), as a plan replayed as a HIP graph with eager work between the
    replays: every replay equals the launch schedule's plan."""
    from hsrq3in_amd import Context
    sizes = [(900, 40, 3), (400, 25, 4), (1300, 50, 5), (200, 8, 6)]
    items, hosts = [], []
    for n, m, seed in sizes:
        H, wr = case(n, seed=seed, m=m)
        hosts.append((H, wr))
        items.append((colmajor_to_device(H), torch.from_numpy(wr.copy()).cuda(),
                      torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")))
    ref = Context(0, packing=1)
    ref.set_lanes(1)
    per = Context(0, packing=1, persistent=1, lag=700)
    per.set_lanes(1)
    try:
        p0 = ref.plan(items, 128)
        p0.execute()
        p0.sync()
        want = [v.clone() for _, _, v in items]
        st0 = p0.status()
        p0.close()
        p1 = per.plan(items, 128)
        for rep in range(3):
            for _, _, v in items:
                v.zero_()
            torch.cuda.synchronize()
            p1.execute(graph=rep > 0)
            p1.sync()
            _ = (items[0][0].sum() + 1.0).item()   # eager work between replays
            for (_, _, v), w in zip(items, want):
                assert torch.equal(v, w), rep
            assert np.array_equal(p1.status(), st0)
        p1.close()
    finally:
        ref.close()
        per.close()


def test_persistent_setting_validation():
    from hsrq3in_amd import Context, HsrqError
    c = Context(0)
    try:
        with pytest.raises(HsrqError):
            c.set_persistent(2, 0)
        with pytest.raises(HsrqError):
            c.set_persistent(1, 2000)
        c.set_persistent(1, 1999)
        c.set_persistent(0, 0)
    finally:
        c.close()
