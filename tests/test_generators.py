"""CPU checks of the generator restatements that the device generators are
tested against (oracle/generators.py, hsrq3in_amd/workloads.py)."""
import numpy as np
import pytest
from scipy import stats

from hsrq3in_amd import workloads
from oracle import generators as gen

# Random123 known-answer vectors for philox4x32 with 10 rounds:
# (counter, key) -> output, 32-bit words.
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    y = gen.philox4x32_10(np.array(ctr, dtype=np.uint32), key)
    assert tuple(int(v) for v in y) == out


def test_philox_vectorised_equals_scalar():
    rng = np.random.default_rng(1)
    ctr = rng.integers(0, 2**32, size=(17, 4), dtype=np.uint64).astype(np.uint32)
    key = (0x12345678, 0x9ABCDEF0)
    y = gen.philox4x32_10(ctr, key)
    for r in range(ctr.shape[0]):
        assert np.array_equal(y[r], gen.philox4x32_10(ctr[r], key))


def test_unit_open_interval_ends():
    lo = gen.unit_open(np.array([0], dtype=np.uint32), np.array([0], dtype=np.uint32))
    hi = gen.unit_open(np.array([0xFFFFFFFF], dtype=np.uint32),
                       np.array([0xFFFFFFFF], dtype=np.uint32))
    assert lo[0] == 2.0 ** -53 and hi[0] == 1.0 - 2.0 ** -53
    assert 0.0 < lo[0] and hi[0] < 1.0


def test_box_muller_pair_invariant():
    # the two variates of one counter share the radius: x_even^2 + x_odd^2 = -2 ln u1
    e = np.arange(0, 4000, dtype=np.uint64)
    x = gen.normals(e, seed=77)
    u1, _ = gen.uniforms(e[::2], seed=77)
    np.testing.assert_allclose(x[::2] ** 2 + x[1::2] ** 2, -2.0 * np.log(u1), rtol=1e-13)


def test_gaussian_sample_is_standard_normal():
    x = gen.normals(np.arange(200_000, dtype=np.uint64), seed=2024)
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1.0) < 0.015
    assert stats.kstest(x, "norm").pvalue > 1e-3
    # cosine and sine halves are uncorrelated
    assert abs(np.corrcoef(x[::2], x[1::2])[0, 1]) < 0.02


def test_gaussian_hessenberg_layout():
    n, seed = 9, (1 << 40) + 5
    h = gen.gaussian_hessenberg(n, seed)
    assert h.flags.f_contiguous
    assert np.all(np.tril(h, -2) == 0.0) and np.all(np.diag(h, -1) != 0.0)
    i, j = 4, 6
    assert h[i, j] == gen.normals(np.array([j * n + i], dtype=np.uint64), seed)[0]
    # seeds differing only in the high key word give different matrices
    assert not np.array_equal(h, gen.gaussian_hessenberg(n, seed - (1 << 40)))


def test_constructed_split_matches_whole():
    lam = np.linspace(-3.0, 5.0, 40)
    whole = workloads.constructed_spectrum(lam, seed=4, weak=[3, 17], weak_nu=1e-6)
    nu, d, u = workloads.constructed_parameters(lam, 4, weak=[3, 17], weak_nu=1e-6)
    assert np.array_equal(whole, workloads.constructed_from_parameters(lam, nu, d, u))


def test_constructed_spectrum_eigenvalues():
    lam = np.arange(1.0, 61.0)
    nu, d, u = workloads.constructed_parameters(lam, 9)
    h = workloads.constructed_from_parameters(lam, nu, d, u)
    assert np.all(np.tril(h, -2) == 0.0) and np.all(np.diag(h, -1) != 0.0)
    w = np.sort(np.linalg.eigvals(h).real)
    np.testing.assert_allclose(w, lam, rtol=0, atol=1e-9 * lam.max())
