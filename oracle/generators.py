"""numpy restatement of the device problem generators (hsrq3in_amd/csrc/hsrq_gen.hip).

TEST INFRASTRUCTURE ONLY: the checker of the device generators in tests/,
never a product path.

* ``philox4x32_10(ctr, key)`` -- the counter-based generator of Salmon, Moraes,
  Dror and Shaw ("Parallel random numbers: as easy as 1, 2, 3", SC'11):
  per round, two 32x32->64 products of counter words 0 and 2 with the
  multipliers 0xD2511F53 / 0xCD9E8D57, the high halves xored with the other
  counter words and the key, the key bumped by the Weyl constants 0x9E3779B9
  / 0xBB67AE85 between rounds.  Pinned by the published known-answer vectors
  of the Random123 distribution (tests/test_generators.py).
* ``gaussian_hessenberg(n, seed)`` -- Box-Muller on Philox output in the
  device generator's counter layout: entry (i, j) has e = j n + i, counter
  (e // 2, 0, 0, 0), key (seed low, seed high); words 0:1 give u1, 2:3 give
  u2 (the top 52 bits u of each, mapped to (0, 1) by (u + 1/2) 2^-52); even e takes
  sqrt(-2 ln u1) cos(2 pi u2), odd e the sine.  Pinned by the pair invariant
  x_2q^2 + x_2q+1^2 = -2 ln u1 and by the normal distribution of a large
  sample (tests/test_generators.py).

The reference has no counterpart: it draws its test matrix from the C
library's rand() (reference problemgenerator.c:15-19) and reduces it with
LAPACK (problemgenerator.c:305-356).
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr: np.ndarray, key) -> np.ndarray:
    """ctr: (..., 4) uint32 counters; key: (k0, k1).  Returns (..., 4) uint32."""
    x = [np.asarray(ctr[..., w], dtype=np.uint64) for w in range(4)]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + W0) & 0xFFFFFFFF
            k1 = (k1 + W1) & 0xFFFFFFFF
        p0 = M0 * x[0]
        p1 = M1 * x[2]
        x = [(p1 >> np.uint64(32)) ^ x[1] ^ np.uint64(k0), p1 & _MASK32,
             (p0 >> np.uint64(32)) ^ x[3] ^ np.uint64(k1), p0 & _MASK32]
    return np.stack(x, axis=-1).astype(np.uint32)


def unit_open(hi: np.ndarray, lo: np.ndarray) -> np.ndarray:
    u = (hi.astype(np.uint64) << np.uint64(20)) | (lo.astype(np.uint64) >> np.uint64(12))
    return (u.astype(np.float64) + 0.5) * 2.0 ** -52


def uniforms(e: np.ndarray, seed: int):
    """(u1, u2) of the pairs holding linear indices e."""
    p = np.asarray(e, dtype=np.uint64) >> np.uint64(1)
    ctr = np.zeros(p.shape + (4,), dtype=np.uint32)
    ctr[..., 0] = (p & _MASK32).astype(np.uint32)
    ctr[..., 1] = (p >> np.uint64(32)).astype(np.uint32)
    y = philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
    return unit_open(y[..., 0], y[..., 1]), unit_open(y[..., 2], y[..., 3])


def normals(e: np.ndarray, seed: int) -> np.ndarray:
    e = np.asarray(e, dtype=np.uint64)
    u1, u2 = uniforms(e, seed)
    rad = np.sqrt(-2.0 * np.log(u1))
    ang = 6.283185307179586 * u2
    return np.where((e & np.uint64(1)) == 1, rad * np.sin(ang), rad * np.cos(ang))


def gaussian_hessenberg(n: int, seed: int) -> np.ndarray:
    """Column-major n x n: N(0, 1) for i <= j + 1, zero below."""
    i, j = np.meshgrid(np.arange(n, dtype=np.uint64), np.arange(n, dtype=np.uint64),
                       indexing="ij")
    h = normals(j * np.uint64(n) + i, seed)
    h[i.astype(np.int64) > j.astype(np.int64) + 1] = 0.0
    return np.asfortranarray(h)
