"""FP32 panel-pass path (HSRQ_PREC_FP32 / HSRQ_PREC_FP32_RAW).

Bar, stated per case: the vectors are finite; in HSRQ_PREC_FP32 mode every
residual ||(H - mu I) x|| / (||H||_oo ||x|| 2^-52) stays <= 10 n (vectors
whose FP32 result fails it are solved again on the FP64 path) and 1 - |cos|
against the FP64 path stays <= 1e-10 on Gaussian matrices and prescribed
spectra.  Graded matrices (row norms spread over more than 2^20) are solved
on the FP64 path in HSRQ_PREC_FP32 mode, bit-identical to HSRQ_PREC_FP64.
The raw FP32 pass is held to the same 1 - |cos| <= 1e-10 on every kind,
graded ones included: the rows of the next diagonal tile are updated in
FP64 at the start of its sweep in every mode, so the FP32 error reaches the
rotations only through the damped cross-over terms.
"""
import numpy as np
import pytest

from hsrq3in_amd import HSRQ_CONVERGED, PREC_FP32, PREC_FP32_RAW
from hsrq3in_amd import workloads as w
from tests.helpers import case, one_minus_abs_cos

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def solve(prec, H, wr, tile=128):
    from hsrq3in_amd import Context, colmajor_to_device
    c = Context(0, precision=prec)
    n = H.shape[0]
    Ht = colmajor_to_device(H)
    wrt = torch.from_numpy(np.array(wr)).cuda()
    vrt = torch.zeros((len(wr), n), dtype=torch.float64, device="cuda")
    st = c.solve(Ht, wrt, vrt, tile)
    res = c.residuals(Ht, wrt, vrt)
    c.close()
    return np.asfortranarray(vrt.cpu().numpy().T), wrt.cpu().numpy(), st, res


@pytest.mark.parametrize("kind,n", [("gauss", 500), ("gauss", 1100), ("prescribed", 700),
                                    ("graded", 1300), ("graded-wide", 900)])
def test_fp32_against_fp64(kind, n):
    if kind == "gauss":
        H, shifts = case(n, seed=60 + n, m=30)
    else:
        H, lam = w.serving_problem(n, seed=40 + n, graded=False)
        if kind == "graded":
            H = w.graded(H, 20, 3)
        elif kind == "graded-wide":
            H = w.graded(H, 100, 3)
        shifts = np.ascontiguousarray(lam[:40])
    x64, w64, s64, r64 = solve(0, H, shifts)
    x32, w32, s32, r32 = solve(PREC_FP32, H, shifts)
    xr, wr_, sr, rr = solve(PREC_FP32_RAW, H, shifts)
    assert np.isfinite(x32).all() and np.isfinite(xr).all()
    assert np.array_equal(w64, w32) and np.array_equal(w64, wr_)
    assert (s32 & HSRQ_CONVERGED).all()
    e_raw = one_minus_abs_cos(xr, x64).max()
    e_cor = one_minus_abs_cos(x32, x64).max()
    print(f"{kind} n={n}: raw 1-|cos| {e_raw:.2e} resid {rr.max():.2e}; "
          f"checked {e_cor:.2e} resid {r32.max():.2e}; fp64 resid {r64.max():.2e}")
    assert r32.max() <= 10 * n
    assert e_raw <= 1e-10
    if kind.startswith("graded"):
        assert np.array_equal(x32, x64)
    else:
        assert e_cor <= 1e-10
        assert np.all(r32 <= np.maximum(rr, 10 * n))
