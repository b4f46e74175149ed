"""Parity of the HIP path (libhsrq through its C-ABI) with the CPU oracle.

Bar (floating point): per eigenvector
1 - |cos angle(x_gpu, x_oracle)| <= 1e-10, the same sign where the shift is
separated from the eigenvalue by relative 1e-12, residual
||(H - wr_k I) x_k|| / (||H||_oo ||x_k|| 2^-52) <= 10 n, no Inf/NaN, the same
convergence status and the same (possibly reordered) shift order.
"""
import functools

import numpy as np
import pytest

import oracle
from hsrq3in_amd import HSRQ_CONVERGED
from tests.helpers import case, cosines, one_minus_abs_cos

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module", params=[1, 4], ids=["pack1", "pack4"])
def ctx(request):
    """Contexts for both diagonal-sweep layouts: 1 and 4 shifts per wavefront."""
    from hsrq3in_amd import Context
    c = Context(0, packing=request.param)
    yield c
    c.close()


def gpu_solve(ctx, H, wr, tlsz=64):
    from hsrq3in_amd import colmajor_to_device
    n = H.shape[0]
    Ht = colmajor_to_device(H)
    wrt = torch.from_numpy(np.array(wr)).cuda()
    vrt = torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")
    st = ctx.solve(Ht, wrt, vrt, tlsz)
    res = ctx.residuals(Ht, wrt, vrt)
    x = np.asfortranarray(vrt.cpu().numpy().T)
    return x, wrt.cpu().numpy(), st, res


def test_selftest(ctx):
    ctx.selftest()


# (321, 322, 323 at tile 128; 641, 642 at tile 256: the ragged first tile
# starts its sweep on the boundaries where the step loop changes the row
# slots it updates, s = 64, 65, 66 and 128, 129)
@pytest.mark.parametrize("n,tile", [(64, 64), (65, 64), (130, 64), (200, 64), (333, 128),
                                    (700, 64), (700, 128), (700, 192), (700, 256),
                                    (1000, 128), (321, 128), (322, 128), (323, 128),
                                    (641, 256), (642, 256)])
def test_parity_gaussian(ctx, n, tile):
    H, wr = case(n, seed=n + tile, m=24)
    x, wo, st, res = gpu_solve(ctx, H, wr, tile)
    xo, woo, sto = oracle.tiled_inverse_iteration(H, wr, tile, 8)
    assert np.isfinite(x).all()
    assert np.array_equal(wo, woo)
    assert (st & HSRQ_CONVERGED).all() and sto.nonconverged == 0
    c = cosines(x, xo)
    assert (1.0 - c).max() <= TOL, (1.0 - c)
    assert res.max() <= 10 * n
    ns, _ = oracle.residuals(H, wo, x)
    assert np.allclose(ns, res, rtol=1e-6, atol=1e-3 * n)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 17, 40])
def test_parity_small(ctx, n):
    H, wr = case(n, seed=7 * n + 1, pert=4e-16)
    x, wo, st, res = gpu_solve(ctx, H, wr, 64)
    xo, woo, sto = oracle.tiled_inverse_iteration(H, wr, 64, 8)
    assert np.isfinite(x).all()
    conv = (st & HSRQ_CONVERGED) != 0
    assert np.array_equal(wo, woo)
    assert conv.sum() == len(wr) - sto.nonconverged
    if conv.any():
        assert one_minus_abs_cos(x[:, conv], xo[:, conv]).max() <= TOL


# Bar 1e-10 (the north star's) for the graded ranges of the workloads.  An
# intermediate grading (k in [-60, 60]) makes the eigenvalues themselves
# ill-conditioned in H's norm (left and right eigenvectors D^-T y0, D x0 far
# from parallel), so two backward-stable runs differ by more there: measured
# up to 2.3e-10 at n = 1000; that case keeps 1e-9.
@pytest.mark.parametrize("n,kmax,tile,tol", [(200, 60, 64, 1e-9), (200, 400, 64, 1e-10),
                                             (500, 400, 128, 1e-10), (1000, 400, 64, 1e-10)])
def test_graded(ctx, n, kmax, tile, tol):
    H, wr = case(n, seed=n + kmax, m=16, kmax=kmax)
    x, wo, st, res = gpu_solve(ctx, H, wr, tile)
    xo, woo, sto = oracle.tiled_inverse_iteration(H, wr, tile, 8)
    assert np.isfinite(x).all() and np.isfinite(xo).all()
    assert (st & HSRQ_CONVERGED).all() and sto.nonconverged == 0
    assert res.max() <= 10 * n
    d = one_minus_abs_cos(x, xo)
    print("graded max 1-|cos| vs oracle", d.max())
    assert d.max() <= tol


def test_cluster(ctx):
    """20 of 50 shifts from a cluster of relative spacing 1e-8 (prescribed
    spectrum): the cluster's eigenvectors are sensitive at ~eps ||H|| / gap,
    so the bar is the residual and 1 - |cos| <= 1e-8 against the oracle."""
    from hsrq3in_amd import workloads as w
    lam, shifts = w.cluster_spectrum(900, 50, 20, 6)
    H = w.cluster_hessenberg(lam, 20, 6)
    x, wo, st, res = gpu_solve(ctx, H, shifts, 128)
    xo, woo, sto = oracle.tiled_inverse_iteration(H, shifts, 128, 8)
    assert np.isfinite(x).all()
    assert np.array_equal(wo, woo)
    assert (st & HSRQ_CONVERGED).all() and sto.nonconverged == 0
    assert res.max() <= 10 * 900
    assert one_minus_abs_cos(x, xo).max() <= 1e-8


def test_exact_norms_identical(ctx):
    from hsrq3in_amd import Context
    H, wr = case(300, seed=11, m=12, kmax=300)
    x1, _, _, _ = gpu_solve(ctx, H, wr, 64)
    c2 = Context(0, exact_norms=True, packing=ctx.packing)
    x2, _, _, _ = gpu_solve(c2, H, wr, 64)
    c2.close()
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("n,tile,kmax", [(700, 128, 200), (700, 128, 0), (517, 64, 60),
                                          (901, 256, 400), (1290, 128, 400), (250, 256, 30)])
def test_packings_agree(ctx, n, tile, kmax):
    """1 and 4 shifts per wavefront run the same arithmetic (no implicit FMA
    contraction, explicit fma()), so the vectors are bit-identical.  The
    four-shift sweep defers the rows above each 16-row block and falls back to
    full steps where a shift needs a protected step: graded inputs (kmax > 0)
    exercise the fall-backs at many places of the blocks, ragged first tiles
    the partial blocks."""
    from hsrq3in_amd import Context
    H, wr = case(n, seed=21 + n, m=20, kmax=kmax)
    out = []
    for pk in (1, 4):
        c = Context(0, packing=pk)
        x, _, st, _ = gpu_solve(c, H, wr, tile)
        out.append((x, st))
        c.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("lanes", [2, 3])
def test_lanes_agree(ctx, lanes):
    """Cutting the shifts into lanes on separate streams changes nothing."""
    from hsrq3in_amd import Context
    H, wr = case(600, seed=77, m=70, kmax=100)
    x1 = gpu_solve(ctx, H, wr, 128)[0]
    c = Context(0, packing=ctx.packing)
    c.set_lanes(lanes)
    x2, _, st, _ = gpu_solve(c, H, wr, 128)
    c.close()
    assert np.array_equal(x1, x2) and (st & HSRQ_CONVERGED).all()


@functools.lru_cache(maxsize=1)
def _many_shift_case():
    """n = 3000 and 600 of its real eigenvalues as shifts (computed once)."""
    return case(3000, seed=91, m=600)


@pytest.mark.parametrize("blocks_per_cu", [0, 1, 3])
def test_lookahead_agrees(ctx, blocks_per_cu):
    """The panel pass of the rows above on a second stream, capped to k
    workgroups per CU (so that its blocks loop over the items: n = 3000,
    m = 600 gives ~870 items per step, more than k * 256), gives the same
    vectors bit for bit as the automatic setting (and as no lookahead)."""
    from hsrq3in_amd import Context
    H, wr = _many_shift_case()
    x1, w1, st1, _ = gpu_solve(ctx, H, wr, 64)
    c = Context(0, packing=ctx.packing)
    c.set_lookahead(blocks_per_cu)
    x2, w2, st2, res = gpu_solve(c, H, wr, 64)
    c.close()
    assert np.array_equal(x1, x2) and np.array_equal(w1, w2) and np.array_equal(st1, st2)
    assert (st2 & HSRQ_CONVERGED).all() and res.max() <= 10 * 3000


def test_lookahead_batch_in_one_lane():
    """Problems of different orders in one lane advance in lockstep (each at
    its own tile column); with the lookahead each sweep waits only for the
    panel pass two steps back.  Bitwise the same as the default plan."""
    from hsrq3in_amd import Context, colmajor_to_device
    cases = [case(n, seed=500 + n, m=m) for n, m in [(300, 20), (700, 40), (1200, 60), (450, 25)]]
    def items():
        return [(colmajor_to_device(H), torch.from_numpy(wr.copy()).cuda(),
                 torch.zeros((len(wr), H.shape[0]), dtype=torch.float64, device="cuda"))
                for H, wr in cases]
    outs = []
    for lanes, look in [(0, 0), (1, 2), (1, 1)]:
        c = Context(0)
        if lanes:
            c.set_lanes(lanes)
        c.set_lookahead(look)
        a = items()
        p = c.plan(a, 128)
        p.execute(graph=True)
        p.sync()
        st = p.status()
        p.close()
        c.close()
        assert (st & HSRQ_CONVERGED).all()
        outs.append([v.cpu().numpy() for _, _, v in a])
    for other in outs[1:]:
        for xa, xb in zip(outs[0], other):
            assert np.array_equal(xa, xb)


@pytest.mark.parametrize("tile", [64, 128])
def test_fused_steps_match_two_streams(tile):
    """With the lookahead a plan runs each step as one launch (the sweeps of
    the step beside the panel pass of the step before, k_step); the profiled
    execution runs the same two kernels apart, the pass on a second stream.
    Both orders give the same vectors and status bit for bit."""
    from hsrq3in_amd import Context, colmajor_to_device
    H, wr = _many_shift_case()
    c = Context(0, packing=1)
    c.set_lookahead(3)
    vr = torch.zeros((len(wr), H.shape[0]), dtype=torch.float64, device="cuda")
    p = c.plan([(colmajor_to_device(H), torch.from_numpy(wr.copy()).cuda(), vr)], tile)
    p.execute(graph=True)
    p.sync()
    x1, st1 = vr.cpu().numpy(), p.status()
    vr.zero_()
    prof = p.execute_profiled()
    x2, st2 = vr.cpu().numpy(), p.status()
    p.close()
    c.close()
    assert prof["update"][1] > 0 and prof["diag"][1] > 0
    assert np.array_equal(x1, x2) and np.array_equal(st1, st2)
    assert (st1 & HSRQ_CONVERGED).all()


def test_autotune_choice_is_consistent():
    """The autotuner picks a supported configuration and a plan built after
    it gives the same vectors as an explicit request of that configuration."""
    from hsrq3in_amd import Context, colmajor_to_device
    cases = [case(n, seed=300 + n, m=m) for n, m in [(200, 12), (450, 30), (700, 40)]]
    def items():
        return [(colmajor_to_device(H), torch.from_numpy(wr.copy()).cuda(),
                 torch.zeros((len(wr), H.shape[0]), dtype=torch.float64, device="cuda"))
                for H, wr in cases]
    c = Context(0)
    tuned = c.autotune(items())
    assert tuned["tile"] in (64, 128, 256) and tuned["shifts_per_wave"] in (1, 4)
    assert 1 <= tuned["lanes"] <= 3 and tuned["ms"] > 0
    a = items()
    p = c.plan(a)
    p.execute(graph=True)
    p.sync()
    p.close()
    c.close()
    ref = Context(0, packing=tuned["shifts_per_wave"])
    b = items()
    p = ref.plan(b, tuned["tile"])
    p.execute()
    p.sync()
    p.close()
    ref.close()
    for (_, _, va), (_, _, vb) in zip(a, b):
        assert torch.equal(va, vb)


def test_batch_equals_alone(ctx):
    from hsrq3in_amd import colmajor_to_device
    cases = [case(n, seed=100 + n, m=m) for n, m in [(90, 7), (257, 20), (64, 5), (500, 33)]]
    items, alone = [], []
    for H, wr in cases:
        n = H.shape[0]
        items.append((colmajor_to_device(H), torch.from_numpy(wr.copy()).cuda(),
                      torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")))
        alone.append(gpu_solve(ctx, H, wr, 64)[0])
    sts = ctx.solve_batch(items, 64)
    for (Ht, wrt, vrt), xa, st in zip(items, alone, sts):
        xb = np.asfortranarray(vrt.cpu().numpy().T)
        assert (st & HSRQ_CONVERGED).all()
        assert np.array_equal(xa, xb)


def test_dropin_host_entry(ctx):
    from hsrq3in_amd import tiled_inverse_iteration
    H, wr = case(257, seed=5, m=10)
    xd, wod, _, _ = gpu_solve(ctx, H, wr, 64)
    xh, woh = tiled_inverse_iteration(H, wr.copy(), 64, 32)
    assert np.array_equal(wod, woh)
    assert np.array_equal(xd, xh)


def test_restart_semantics(ctx):
    # shifts half way between eigenvalues do not converge: restarts with the
    # dlaein vectors, unconverged shifts moved behind the converged ones
    H, wr = case(12, seed=3, pert=4e-16)
    wr = np.sort(wr)
    mixed = np.array([wr[0], 0.5 * (wr[0] + wr[1]), wr[1], 0.5 * (wr[1] + wr[2])])
    x, wo, st, _ = gpu_solve(ctx, H, mixed, 64)
    xo, woo, sto = oracle.tiled_inverse_iteration(H, mixed, 64, 8)
    assert np.array_equal(wo, woo)
    conv = (st & HSRQ_CONVERGED) != 0
    assert conv.sum() == len(mixed) - sto.nonconverged
    assert np.isfinite(x).all()


def test_graph_replay(ctx):
    from hsrq3in_amd import colmajor_to_device
    H, wr = case(400, seed=9, m=16)
    n = H.shape[0]
    Ht = colmajor_to_device(H)
    wrt = torch.from_numpy(wr.copy()).cuda()
    vrt = torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")
    plan = ctx.plan([(Ht, wrt, vrt)], 64)
    plan.execute(graph=False)
    plan.sync()
    ref = vrt.clone()
    for _ in range(3):
        vrt.zero_()
        torch.cuda.synchronize()
        plan.execute(graph=True)
        plan.sync()
        _ = (Ht.sum() + 1.0).item()   # eager work between replays
        assert torch.equal(vrt, ref)
    assert (plan.status() & HSRQ_CONVERGED).all()
    plan.close()
