"""On-disk request / response format of the serving loop (hsrq3in_amd.serve):
round trips, header checks, truncation (no device needed)."""
import numpy as np
import pytest

from hsrq3in_amd import serve
from hsrq3in_amd import workloads as w


def test_request_round_trip(tmp_path):
    H = w.gaussian_hessenberg(37, 1)
    wr = np.linspace(-1, 1, 5)
    p = str(tmp_path / "a.hsrq")
    serve.write_request(p, H, wr)
    H2, wr2, wi2, m2 = serve.read_request(p)
    assert np.array_equal(np.asarray(H2), H) and np.array_equal(wr2, wr) and wi2 is None
    assert m2 == 5
    assert (tmp_path / "a.hsrq").stat().st_size == 24 + 8 * (37 * 37 + 5)


def test_complex_request_round_trip(tmp_path):
    H = w.gaussian_hessenberg(9, 2)
    wr, wi = np.array([0.5, -1.0]), np.array([2.0, 0.25])
    p = str(tmp_path / "c.hsrq")
    serve.write_request(p, H, wr, wi)
    H2, wr2, wi2, _ = serve.read_request(p)
    assert np.array_equal(np.asarray(H2), H)
    assert np.array_equal(wr2, wr) and np.array_equal(wi2, wi)


def test_response_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    vr = rng.standard_normal((11, 3))
    p = str(tmp_path / "r.hsrs")
    serve.write_response(p, vr, [1.0, 2.0, 3.0], [1, 3, 0], 1.25, 0)
    r = serve.read_response(p)
    assert r["n"] == 11 and r["m"] == 3 and r["rc"] == 0 and r["latency_ms"] == 1.25
    assert np.array_equal(r["vr"], vr) and list(r["status"]) == [1, 3, 0]
    # complex: 2m columns on disk, complex vectors when read
    vc = rng.standard_normal((11, 4))
    serve.write_response(p, vc, [1.0, 2.0], [1, 1], 0.5, 0, complex_shifts=True)
    r = serve.read_response(p)
    assert np.array_equal(r["vr"], vc[:, 0::2] + 1j * vc[:, 1::2])


def test_bad_files_rejected(tmp_path):
    p = tmp_path / "x.hsrq"
    p.write_bytes(b"NOTMAGIC" + bytes(16))
    with pytest.raises(ValueError):
        serve.read_request(str(p))
    H = w.gaussian_hessenberg(8, 3)
    serve.write_request(str(p), H, [0.1])
    data = p.read_bytes()
    p.write_bytes(data[:-8])                      # truncated
    with pytest.raises(ValueError):
        serve.read_request(str(p))
    with pytest.raises(ValueError):
        serve.read_response(str(p))               # a request is not a response
    with pytest.raises(ValueError):
        serve.write_response(str(tmp_path / "y.hsrs"), np.zeros((8, 2)), [0.1], [0], 0.0)


def test_no_partial_files_left(tmp_path):
    serve.write_request(str(tmp_path / "a.hsrq"), np.eye(4), [1.0])
    assert sorted(f.name for f in tmp_path.iterdir()) == ["a.hsrq"]


def test_request_without_shifts(tmp_path):
    H = w.gaussian_hessenberg(30, 5)
    p = str(tmp_path / "e.hsrq")
    serve.write_request(p, H, None, m=4)
    H2, wr2, wi2, m2 = serve.read_request(p)
    assert np.array_equal(np.asarray(H2), H) and wr2 is None and wi2 is None and m2 == 4
    assert (tmp_path / "e.hsrq").stat().st_size == 24 + 8 * 30 * 30
    with pytest.raises(ValueError):
        serve.write_request(p, H, None)                       # m missing


def test_shift_cache(tmp_path):
    H = w.gaussian_hessenberg(60, 6)
    c = w.ShiftCache(str(tmp_path / "cache"))
    a = c.get(H, 5)
    assert np.array_equal(a, w.real_shifts(H, 5))
    b = c.get(np.asfortranarray(H.copy()), 3)                # same matrix: a hit
    assert np.array_equal(b, a[:3]) and (c.hits, c.misses) == (1, 1)
    H2 = H.copy()
    H2[0, 0] += 1e-9                                          # another matrix: a miss
    c.get(H2, 1)
    assert c.misses == 2 and len(list((tmp_path / "cache").iterdir())) == 2
    with pytest.raises(ValueError):
        c.get(H, 61)
