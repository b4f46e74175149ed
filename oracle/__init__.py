"""ctypes binding of the CPU oracle (liboracle_hsrq.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, and only as the checker / baseline, never as a
product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_hsrq.so")

_dp = ctypes.POINTER(ctypes.c_double)


class OracleStats(ctypes.Structure):
    _fields_ = [("scalings", ctypes.c_longlong), ("perturbed", ctypes.c_longlong),
                ("restarts", ctypes.c_int), ("nonconverged", ctypes.c_int)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle_hsrq.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_tiled_inverse_iteration.restype = ctypes.c_int
        L.oracle_tiled_inverse_iteration.argtypes = [
            ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
            _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OracleStats)]
        L.oracle_tiled_inverse_iteration_ex.restype = ctypes.c_int
        L.oracle_tiled_inverse_iteration_ex.argtypes = [
            ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
            _dp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(OracleStats),
            ctypes.POINTER(ctypes.c_int), _dp, ctypes.c_int]
        L.oracle_dlartg.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.oracle_givens.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.oracle_protect_update.restype = ctypes.c_double
        L.oracle_protect_update.argtypes = [ctypes.c_double] * 3
        L.oracle_protect_double_update.restype = ctypes.c_double
        L.oracle_protect_double_update.argtypes = [ctypes.c_double] * 5
        L.oracle_solve_1x1_real_system.restype = ctypes.c_int
        L.oracle_solve_1x1_real_system.argtypes = [ctypes.c_double] * 3 + [_dp, _dp]
        L.oracle_nrm2_blue.restype = ctypes.c_double
        L.oracle_nrm2_blue.argtypes = [ctypes.c_int, _dp]
        L.oracle_dlange_inf.restype = ctypes.c_double
        L.oracle_dlange_inf.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
        L.oracle_partition.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int)]
        L.oracle_residuals.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                       _dp, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.oracle_complex_inverse_iteration.restype = ctypes.c_int
        L.oracle_complex_inverse_iteration.argtypes = [
            ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
            ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def tiled_inverse_iteration(H: np.ndarray, wr: np.ndarray, tlsz: int = 64,
                            rhs_tlsz: int = 32, nthreads: int = 0):
    """Returns (vr, wr_out, stats). H column-major (n x n), wr length m."""
    H = np.asfortranarray(H, dtype=np.float64)
    n = H.shape[0]
    wr_out = np.array(wr, dtype=np.float64, copy=True)
    m = wr_out.size
    vr = np.zeros((n, max(m, 1)), dtype=np.float64, order="F")
    st = OracleStats()
    rc = lib().oracle_tiled_inverse_iteration(n, _ptr(H), n, m, _ptr(wr_out), tlsz, rhs_tlsz,
                                              _ptr(vr), n, nthreads, ctypes.byref(st))
    if rc != 0:
        raise ValueError(f"oracle_tiled_inverse_iteration failed: {rc}")
    return vr[:, :m], wr_out, st


def tiled_inverse_iteration_detail(H: np.ndarray, wr: np.ndarray, tlsz: int = 64,
                                   rhs_tlsz: int = 32, nthreads: int = 0,
                                   max_restarts: int = -1):
    """As tiled_inverse_iteration, plus per shift (final order): status bits
    (1 converged, 2 a pivot raised to smin in the last solve) and log2 of the
    smallest tile scaling factor of the last solve.  max_restarts bounds the
    restart rounds (-1: n, the reference's limit; 0: first solve only).
    Returns (vr, wr_out, stats, status, log2scale)."""
    H = np.asfortranarray(H, dtype=np.float64)
    n = H.shape[0]
    wr_out = np.array(wr, dtype=np.float64, copy=True)
    m = wr_out.size
    vr = np.zeros((n, max(m, 1)), dtype=np.float64, order="F")
    status = np.zeros(max(m, 1), dtype=np.int32)
    lsc = np.zeros(max(m, 1), dtype=np.float64)
    st = OracleStats()
    rc = lib().oracle_tiled_inverse_iteration_ex(
        n, _ptr(H), n, m, _ptr(wr_out), tlsz, rhs_tlsz, _ptr(vr), n, nthreads, ctypes.byref(st),
        status.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(lsc), max_restarts)
    if rc != 0:
        raise ValueError(f"oracle_tiled_inverse_iteration_ex failed: {rc}")
    return vr[:, :m], wr_out, st, status[:m], lsc[:m]


def dlartg(f: float, g: float):
    c, s, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().oracle_dlartg(f, g, ctypes.byref(c), ctypes.byref(s), ctypes.byref(r))
    return c.value, s.value, r.value


def givens(f: float, g: float):
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().oracle_givens(f, g, ctypes.byref(c), ctypes.byref(s))
    return c.value, s.value


def protect_update(t, x, y):
    return lib().oracle_protect_update(t, x, y)


def solve_1x1(smin, t, lam, x):
    xv, sc = ctypes.c_double(x), ctypes.c_double(1.0)
    info = lib().oracle_solve_1x1_real_system(smin, t, lam, ctypes.byref(xv), ctypes.byref(sc))
    return xv.value, sc.value, info


def nrm2_blue(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().oracle_nrm2_blue(x.size, _ptr(x))


def dlange_inf(A: np.ndarray) -> float:
    A = np.asfortranarray(A, dtype=np.float64)
    return lib().oracle_dlange_inf(A.shape[0], A.shape[1], _ptr(A), A.shape[0])


def partition(n: int, num_tiles: int, tlsz: int):
    p = (ctypes.c_int * (num_tiles + 1))()
    lib().oracle_partition(n, num_tiles, tlsz, p)
    return list(p)


def residuals(H: np.ndarray, wr: np.ndarray, X: np.ndarray, nthreads: int = 0):
    """(north-star residual / (||H||_F ||x|| eps), validation.c:161 form)."""
    H = np.asfortranarray(H, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    n, m = X.shape
    ns = np.zeros(m)
    ref = np.zeros(m)
    lib().oracle_residuals(n, _ptr(H), H.shape[0], m, _ptr(wr), _ptr(X), X.shape[0],
                           _ptr(ns), _ptr(ref), nthreads)
    return ns, ref


def complex_inverse_iteration(H: np.ndarray, wr, wi, nthreads: int = 0, max_restarts: int = 0):
    """Complex shifts wr + i wi: returns (X complex n x m, status int array);
    X[:, k] unit 2-norm with its first entry of largest modulus real positive;
    unconverged shifts retried with dlaein's restart vectors up to max_restarts times."""
    H = np.asfortranarray(H, dtype=np.float64)
    n = H.shape[0]
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    wi = np.ascontiguousarray(wi, dtype=np.float64)
    m = wr.size
    vr = np.zeros((n, 2 * max(m, 1)), dtype=np.float64, order="F")
    st = np.zeros(max(m, 1), dtype=np.int32)
    rc = lib().oracle_complex_inverse_iteration(n, _ptr(H), n, m, _ptr(wr), _ptr(wi), _ptr(vr),
                                                n, max_restarts,
                                                st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                nthreads)
    if rc != 0:
        raise ValueError("oracle_complex_inverse_iteration: invalid arguments")
    return vr[:, 0:2 * m:2] + 1j * vr[:, 1:2 * m:2], st[:m]
