"""Writes the shift fixtures of the benchmark workloads.

The shifts are the first m real eigenvalues of seeded Gaussian upper
Hessenberg matrices (numpy dgeev = balancing + LAPACK dhseqr), perturbed by
relative 1e-12.  The eigenvalue computation is O(n^3) (minutes at n = 8000),
so it runs once here and the lists are committed; the matrices themselves are
rebuilt bit-identically from their seeds wherever they are needed.

    python tests/golden/make_shifts.py [cfg1] [cfg2] [batch] [cfg4]

cfg4 (n = 20000, m = 2000) takes about an hour of host time.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hsrq3in_amd import workloads as w  # noqa: E402


def single(name, n, m, seed):
    t = time.time()
    h = w.gaussian_hessenberg(n, seed)
    s = w.real_shifts(h, m)
    np.savez(os.path.join(HERE, name), shifts=s, n=n, m=m, seed=seed)
    print(f"{name}: n={n} m={m} seed={seed} {time.time() - t:.1f}s", flush=True)


def batch():
    t = time.time()
    probs = w.batch_sizes()
    shifts, offs = [], [0]
    for (n, m, g, seed) in probs:
        h = w.gaussian_hessenberg(n, seed)
        s = w.real_shifts(h, m)
        shifts.append(s)
        offs.append(offs[-1] + m)
    np.savez(os.path.join(HERE, "shifts_batch256.npz"),
             shifts=np.concatenate(shifts), offsets=np.array(offs),
             n=np.array([p[0] for p in probs]), m=np.array([p[1] for p in probs]),
             graded=np.array([p[2] for p in probs]), seeds=np.array([p[3] for p in probs]))
    print(f"batch: {len(probs)} problems, {offs[-1]} shifts {time.time() - t:.1f}s", flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["cfg1", "cfg2", "batch"]
    if "cfg1" in which:
        single("shifts_n1000_s1.npz", 1000, 100, 1)
    if "cfg2" in which:
        single("shifts_n8000_s2.npz", 8000, 800, 2)
    if "batch" in which:
        batch()
    if "cfg4" in which:
        single("shifts_n20000_s4.npz", 20000, 2000, 4)
