"""The spool-directory server (serve_spool) against direct solves: the
responses hold exactly what Server / Context.solve_complex return."""
import numpy as np
import pytest

from hsrq3in_amd import colmajor_to_device, serve
from tests.helpers import case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from hsrq3in_amd import Context
    c = Context(0)
    yield c
    c.close()


def test_spool_matches_direct(ctx, tmp_path):
    src, dst = tmp_path / "in", tmp_path / "out"
    src.mkdir()
    probs = {f"r{n}": case(n, seed=n, m=12) for n in (150, 700, 1300)}
    for name, (H, wr) in probs.items():
        serve.write_request(str(src / f"{name}.hsrq"), H, wr)
    Hc, _ = case(300, seed=9, m=1)
    ev = np.linalg.eigvals(Hc)
    cz = ev[ev.imag > 0][:5] * (1 + 1e-12)
    serve.write_request(str(src / "c300.hsrq"), Hc, cz.real, cz.imag)
    (src / "bad.hsrq").write_bytes(b"garbage")
    # no shifts: the server takes the first 10 real eigenvalues of H
    He, _ = case(500, seed=41, m=1)
    serve.write_request(str(src / "e500.hsrq"), He, None, m=10)
    done = serve.serve_spool(ctx, str(src), str(dst), hbm_bytes=1 << 30)
    assert done == {"bad": -1, "c300": 0, "e500": 0, "r150": 0, "r1300": 0, "r700": 0}
    # real: the server's own results
    srv = serve.Server(ctx, 1 << 30)
    for name, (H, wr) in probs.items():
        r = serve.read_response(str(dst / f"{name}.hsrs"))
        vr, wro, st, _ = srv.wait(srv.submit(H, wr))
        assert r["rc"] == 0 and np.array_equal(r["vr"], vr)
        assert np.array_equal(r["wr"], wro) and np.array_equal(r["status"], st)
        assert r["latency_ms"] > 0.0
    srv.close()
    # complex: the device complex solve
    r = serve.read_response(str(dst / "c300.hsrs"))
    vrt, st = ctx.solve_complex(colmajor_to_device(Hc), torch.from_numpy(cz.real.copy()).cuda(),
                                torch.from_numpy(cz.imag.copy()).cuda())
    v = vrt.cpu().numpy()
    assert np.array_equal(r["vr"], (v[0::2] + 1j * v[1::2]).T)
    assert np.array_equal(r["status"], st)
    # the shift-less request equals one with its shifts spelled out
    from hsrq3in_amd import workloads
    re_ = serve.read_response(str(dst / "e500.hsrs"))
    srv = serve.Server(ctx, 1 << 30)
    vr, wro, st, _ = srv.wait(srv.submit(He, workloads.real_shifts(He, 10)))
    srv.close()
    assert np.array_equal(re_["vr"], vr) and np.array_equal(re_["wr"], wro)
    # a second pass serves nothing new
    assert serve.serve_spool(ctx, str(src), str(dst)) == {"bad": -1}
