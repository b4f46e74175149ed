"""hsrq3in_amd -- robust tiled inverse iteration for upper Hessenberg matrices
on one MI355X (HIP / gfx950), behind the reference's C entry point.

The compute path is libhsrq.so (HIP kernels + C-ABI, ``include/hsrq.h`` and
``include/inverse_iteration.h``).  This module is a thin ctypes binding:

* ``tiled_inverse_iteration(H, wr, tlsz, rhs_tlsz)`` -- the reference's
  entry point (reference inverse_iteration.h:19-22) on host arrays, through
  the drop-in symbol of the same name;
* ``Context`` -- device-resident solves, batches, reusable plans (HIP-graph
  capture), on-device residuals;
* ``Context.gaussian_hessenberg / constructed_hessenberg / grade`` -- test
  problems generated in HBM (Philox Gaussian, prescribed spectrum, exact
  power-of-two grading with a range check);
* ``Context.backtransform`` -- eigenvectors of A = Z H Z^T from those of H
  (V = Z X, one FP64 GEMM);
* ``Server`` (``hsrq3in_amd.serve``) -- continuous-batching request loop.

There is no fallback: if libhsrq.so is missing or fails to load, every call
raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhsrq.so")

HSRQ_CONVERGED = 1
HSRQ_PERTURBED = 2
HSRQ_VALIDATED = 4     # server results: residual <= 10 n eps (validated online)
HSRQ_SCALE_SHIFT = 8   # status bits 8-23: binary orders of downscaling of the vector


def status_scale(st):
    """Binary orders of downscaling carried into the back-transform, per shift."""
    return (np.asarray(st) >> HSRQ_SCALE_SHIFT) & 0xFFFF
HSRQ_ERANGE = -8    # hsrq_grade_device: a graded entry would leave the FP64 range
PREC_FP64 = 0
PREC_FP32 = 1       # FP32 MFMA panel pass, FP64-residual check, FP64 re-solve of failures
PREC_FP32_RAW = 2   # FP32 MFMA panel pass only

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class HsrqError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = lib().hsrq_strerror(code).decode() if code < 0 else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


class Problem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("H", _vp), ("ldH", ctypes.c_int), ("m", ctypes.c_int),
                ("wr", _vp), ("vr", _vp), ("ldvr", ctypes.c_int)]


# Every symbol the C-ABI exports (checked by the CPU test-suite).
EXPORTS = [
    "tiled_inverse_iteration",
    "hsrq_context_create", "hsrq_context_destroy", "hsrq_context_set_tile",
    "hsrq_context_set_exact_norms", "hsrq_context_set_precision", "hsrq_context_set_lookahead",
    "hsrq_context_set_packing", "hsrq_context_set_lanes",
    "hsrq_strerror",
    "hsrq_selftest", "hsrq_solve_device", "hsrq_solve_device_noalloc", "hsrq_solve_batch_device",
    "hsrq_plan_create",
    "hsrq_plan_execute", "hsrq_plan_execute_graph", "hsrq_plan_execute_profiled",
    "hsrq_plan_sync", "hsrq_plan_status", "hsrq_plan_lanes", "hsrq_autotune",
    "hsrq_plan_workspace_bytes", "hsrq_plan_destroy", "hsrq_context_stream",
    "hsrq_residuals_device", "hsrq_residuals_batch_device",
    "hsrq_gaussian_hessenberg_device", "hsrq_constructed_hessenberg_device", "hsrq_grade_device",
    "hsrq_backtransform_device", "hsrq_solve_complex_device",
    "hsrq_server_create", "hsrq_server_submit", "hsrq_server_poll", "hsrq_server_wait",
    "hsrq_server_get_stats", "hsrq_server_destroy",
]

_lib = None


def lib():
    """Loads libhsrq.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m hsrq3in_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.tiled_inverse_iteration.restype = None
        L.tiled_inverse_iteration.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                              ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
        L.hsrq_strerror.restype = ctypes.c_char_p
        L.hsrq_strerror.argtypes = [ctypes.c_int]
        L.hsrq_context_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int]
        L.hsrq_context_destroy.restype = None
        L.hsrq_context_destroy.argtypes = [_vp]
        L.hsrq_context_set_tile.argtypes = [_vp, ctypes.c_int]
        L.hsrq_context_set_exact_norms.argtypes = [_vp, ctypes.c_int]
        L.hsrq_context_set_precision.argtypes = [_vp, ctypes.c_int]
        L.hsrq_context_set_lookahead.argtypes = [_vp, ctypes.c_int]
        L.hsrq_context_set_packing.argtypes = [_vp, ctypes.c_int]
        L.hsrq_context_set_lanes.argtypes = [_vp, ctypes.c_int]
        L.hsrq_context_stream.restype = _vp
        L.hsrq_context_stream.argtypes = [_vp]
        L.hsrq_selftest.argtypes = [_vp]
        L.hsrq_solve_device.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp,
                                        _vp, ctypes.c_int, ctypes.c_int, _ip]
        L.hsrq_solve_device_noalloc.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int,
                                                ctypes.c_int, _vp, _vp, ctypes.c_int,
                                                ctypes.c_int, _ip]
        L.hsrq_solve_batch_device.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Problem),
                                              ctypes.c_int, _ip]
        L.hsrq_plan_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Problem), ctypes.c_int,
                                       ctypes.POINTER(_vp)]
        for f in ("hsrq_plan_execute", "hsrq_plan_execute_graph", "hsrq_plan_sync"):
            getattr(L, f).argtypes = [_vp]
        L.hsrq_plan_execute_profiled.argtypes = [_vp, _dp, _ip]
        L.hsrq_plan_status.argtypes = [_vp, _ip]
        L.hsrq_plan_lanes.argtypes = [_vp, ctypes.c_int, _ip, _ip, _ip]
        L.hsrq_autotune.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Problem), _ip, _ip, _ip, _dp]
        L.hsrq_plan_workspace_bytes.restype = ctypes.c_size_t
        L.hsrq_plan_workspace_bytes.argtypes = [_vp]
        L.hsrq_plan_destroy.restype = None
        L.hsrq_plan_destroy.argtypes = [_vp]
        L.hsrq_residuals_device.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                            _vp, _vp, ctypes.c_int, _dp]
        L.hsrq_residuals_batch_device.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Problem), _dp]
        L.hsrq_gaussian_hessenberg_device.argtypes = [_vp, ctypes.c_int, ctypes.c_uint64, _vp,
                                                      ctypes.c_int]
        L.hsrq_constructed_hessenberg_device.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, _dp,
                                                         _vp, ctypes.c_int]
        L.hsrq_grade_device.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _ip, ctypes.c_int,
                                        _ip, _ip]
        L.hsrq_solve_complex_device.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                                _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _ip]
        L.hsrq_backtransform_device.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                                _vp, ctypes.c_int, _vp, ctypes.c_int]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise HsrqError(rc, what)


# ---------------------------------------------------------------------------
# Reference interface (host arrays)
# ---------------------------------------------------------------------------

def tiled_inverse_iteration(H: np.ndarray, wr: np.ndarray, tlsz: int = 64, rhs_tlsz: int = 32,
                            vr: np.ndarray | None = None):
    """Eigenvectors of the upper Hessenberg H for the real shifts wr.

    Mirrors the reference's C entry point (inverse_iteration.h:19-22): H is
    n x n column-major (a Fortran-ordered array), wr is updated in place
    (reordered, converged first, if restarts were needed) and returned with
    the n x m unit-norm eigenvector matrix vr.
    """
    H = np.asfortranarray(H, dtype=np.float64)
    n = H.shape[0]
    if H.ndim != 2 or H.shape[1] != n:
        raise ValueError("H must be square")
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    m = wr.size
    if vr is None:
        vr = np.zeros((n, m), dtype=np.float64, order="F")
    if not (vr.flags.f_contiguous and vr.shape == (n, m)):
        raise ValueError("vr must be a Fortran-ordered n x m array")
    lib().tiled_inverse_iteration(n, H.ctypes.data_as(_dp), n, m, wr.ctypes.data_as(_dp), tlsz,
                                  rhs_tlsz, vr.ctypes.data_as(_dp), n)
    return vr, wr


# ---------------------------------------------------------------------------
# Device API (torch tensors as device memory; column-major = transposed view)
# ---------------------------------------------------------------------------

def colmajor_to_device(a: np.ndarray, device="cuda"):
    """Fortran-ordered numpy matrix -> torch tensor T with T[j, i] = a[i, j]
    (the column-major layout the C-ABI expects; leading dimension a.shape[0])."""
    import torch
    a = np.asfortranarray(a, dtype=np.float64)
    return torch.from_numpy(a.T).to(device)


def colmajor_to_host(t) -> np.ndarray:
    """Inverse of colmajor_to_device."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def _check_matrix(t, name: str, n: int | None = None):
    """A device matrix in the transposed-storage layout: contiguous float64
    CUDA tensor (rows, ld) with ld >= rows (rows == n when n is given)."""
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda \
            or t.dim() != 2 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous 2-D float64 CUDA tensor")
    if n is not None and t.shape[0] != n:
        raise ValueError(f"{name} has {t.shape[0]} rows, expected {n}")


def _check_problem(Ht, wrt, vrt):
    """Shapes and types of one device problem (Ht (n, ldH >= n), wrt (m,),
    vrt (m, ldvr >= n)); returns (n, m)."""
    import torch
    _check_matrix(Ht, "Ht")
    n = Ht.shape[0]
    if Ht.shape[1] < n:
        raise ValueError("Ht must be (n, ldH) with ldH >= n")
    if not isinstance(wrt, torch.Tensor) or wrt.dtype != torch.float64 or not wrt.is_cuda \
            or not wrt.is_contiguous():
        raise ValueError("wrt must be a contiguous float64 CUDA tensor")
    m = wrt.numel()
    _check_matrix(vrt, "vrt", m)
    if m and vrt.shape[1] < n:
        raise ValueError("vrt must be (m, ldvr) with ldvr >= n")
    if not (Ht.device == wrt.device == vrt.device):
        raise ValueError("Ht, wrt and vrt must be on one device")
    return n, m


def _after_torch(stream_ptr: int, device):
    """Orders the library stream after the work torch has queued on its
    current stream (the library's streams are non-blocking: they do not wait
    for the fills and copies that produced the inputs by themselves)."""
    import torch
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream(device))
    torch.cuda.ExternalStream(stream_ptr, device=device).wait_event(ev)


class Context:
    """Device, stream and settings of libhsrq (hsrq_context).

    Every method that hands device tensors to the library first orders the
    context stream after torch's current stream, so inputs produced by torch
    are complete.  The synchronous methods return with the results in place;
    Plan.execute is asynchronous (Plan.sync)."""

    def __init__(self, device: int = 0, tile: int = 0, exact_norms: bool = False,
                 packing: int = 0, precision: int = 0):
        self._h = _vp()
        self.packing = packing
        _check(lib().hsrq_context_create(ctypes.byref(self._h), device), "hsrq_context_create")
        if tile:
            _check(lib().hsrq_context_set_tile(self._h, tile), "set_tile")
        if exact_norms:
            _check(lib().hsrq_context_set_exact_norms(self._h, 1), "set_exact_norms")
        if packing:
            _check(lib().hsrq_context_set_packing(self._h, packing), "set_packing")
        if precision:
            _check(lib().hsrq_context_set_precision(self._h, precision), "set_precision")

    def close(self):
        if self._h:
            lib().hsrq_context_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_exact_norms(self, on: bool):
        _check(lib().hsrq_context_set_exact_norms(self._h, int(on)), "set_exact_norms")

    def set_packing(self, shifts_per_wave: int):
        _check(lib().hsrq_context_set_packing(self._h, shifts_per_wave), "set_packing")
        self.packing = shifts_per_wave

    def set_tile(self, tile: int):
        _check(lib().hsrq_context_set_tile(self._h, tile), "set_tile")

    def reset_tuning(self):
        """Back to the automatic tile, sweep layout and lane choice."""
        self.set_tile(0)
        self.set_packing(0)
        self.set_lanes(0)

    def set_lanes(self, lanes: int):
        _check(lib().hsrq_context_set_lanes(self._h, lanes), "set_lanes")

    def set_lookahead(self, blocks_per_cu: int):
        """0: off; k in 1..8: panel pass of the rows above on a second stream,
        at most k workgroups per CU (plans of one lane); -1: automatic."""
        _check(lib().hsrq_context_set_lookahead(self._h, int(blocks_per_cu)), "set_lookahead")

    def set_precision(self, prec: int):
        _check(lib().hsrq_context_set_precision(self._h, prec), "set_precision")

    def selftest(self):
        _check(lib().hsrq_selftest(self._h), "hsrq_selftest")

    def stream_ptr(self) -> int:
        return lib().hsrq_context_stream(self._h) or 0

    def solve(self, Ht, wrt, vrt, tlsz: int = 0):
        """Solve on device tensors (column-major via transposed storage):
        Ht (n, ldH) with Ht[j, i] = H(i, j); wrt (m,); vrt (m, ldvr).
        Returns the per-shift status array."""
        n, m = _check_problem(Ht, wrt, vrt)
        st = np.zeros(max(m, 1), dtype=np.int32)
        _after_torch(self.stream_ptr(), Ht.device)
        rc = lib().hsrq_solve_device(self._h, n, Ht.data_ptr(), Ht.shape[1], m, wrt.data_ptr(),
                                     vrt.data_ptr(), vrt.shape[1], tlsz,
                                     st.ctypes.data_as(_ip))
        _check(rc, "hsrq_solve_device")
        return st[:m]

    @staticmethod
    def _problems(items):
        arr = (Problem * len(items))()
        for i, (Ht, wrt, vrt) in enumerate(items):
            _check_problem(Ht, wrt, vrt)
            arr[i] = Problem(Ht.shape[0], Ht.data_ptr(), Ht.shape[1], wrt.numel(), wrt.data_ptr(),
                             vrt.data_ptr(), vrt.shape[1])
        return arr

    def solve_batch(self, items, tlsz: int = 0):
        """items: list of (Ht, wrt, vrt). Returns one status array per problem."""
        arr = self._problems(items)
        total = sum(w.numel() for _, w, _ in items)
        st = np.zeros(max(total, 1), dtype=np.int32)
        if items:
            _after_torch(self.stream_ptr(), items[0][0].device)
        _check(lib().hsrq_solve_batch_device(self._h, len(items), arr, tlsz,
                                             st.ctypes.data_as(_ip)), "hsrq_solve_batch_device")
        out, pos = [], 0
        for _, w, _ in items:
            out.append(st[pos:pos + w.numel()])
            pos += w.numel()
        return out

    def autotune(self, items) -> dict:
        """Times the problems for every tile / sweep layout / lane count and
        keeps the fastest in this context (hsrq_autotune)."""
        arr = self._problems(items)
        t, pk, ln = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        if items:
            _after_torch(self.stream_ptr(), items[0][0].device)
        _check(lib().hsrq_autotune(self._h, len(items), arr, ctypes.byref(t), ctypes.byref(pk),
                                   ctypes.byref(ln), ctypes.byref(ms)), "hsrq_autotune")
        self.packing = pk.value
        return {"tile": t.value, "shifts_per_wave": pk.value, "lanes": ln.value,
                "ms": round(ms.value, 3)}

    def plan(self, items, tlsz: int = 0) -> "Plan":
        return Plan(self, items, tlsz)

    def residuals(self, Ht, wrt, vrt) -> np.ndarray:
        """||(H - wr_k I) x_k|| / (||H||_oo ||x_k|| 2^-52) per eigenvector (device)."""
        _, m = _check_problem(Ht, wrt, vrt)
        res = np.zeros(max(m, 1), dtype=np.float64)
        _after_torch(self.stream_ptr(), Ht.device)
        _check(lib().hsrq_residuals_device(self._h, Ht.shape[0], Ht.data_ptr(), Ht.shape[1], m,
                                           wrt.data_ptr(), vrt.data_ptr(), vrt.shape[1],
                                           res.ctypes.data_as(_dp)), "hsrq_residuals_device")
        return res[:m]


    def residuals_batch(self, items) -> list:
        """Residuals of many problems in one device pass (hsrq_residuals_batch_device):
        items as for solve_batch; one array per problem."""
        arr = self._problems(items)
        total = sum(w.numel() for _, w, _ in items)
        res = np.zeros(max(total, 1), dtype=np.float64)
        if items:
            _after_torch(self.stream_ptr(), items[0][0].device)
        _check(lib().hsrq_residuals_batch_device(self._h, len(items), arr,
                                                 res.ctypes.data_as(_dp)),
               "hsrq_residuals_batch_device")
        out, pos = [], 0
        for _, w, _ in items:
            out.append(res[pos:pos + w.numel()])
            pos += w.numel()
        return out

    def solve_complex(self, Ht, wrt, wit, vrt=None, max_restarts: int = 0):
        """Complex shifts wr + i wi (device tensors, m entries) on Ht (as in
        solve).  Returns (vrt, status): vrt (2m, n) with rows 2k / 2k+1 the real
        / imaginary parts of x_k (hsrq_solve_complex_device); unconverged shifts
        are retried with dlaein's restart vectors up to max_restarts times."""
        import torch
        _check_matrix(Ht, "Ht")
        n, m = Ht.shape[0], wrt.numel()
        if vrt is None:
            vrt = torch.zeros((2 * m, n), dtype=torch.float64, device=Ht.device)
        for t, name in ((wrt, "wrt"), (wit, "wit")):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous() or t.numel() != m:
                raise ValueError(f"{name} must be a contiguous float64 CUDA tensor of m entries")
        _check_matrix(vrt, "vrt", 2 * m)
        if Ht.shape[1] < n or vrt.shape[1] < n:
            raise ValueError("Ht and vrt need leading dimensions >= n")
        st = np.zeros(max(m, 1), dtype=np.int32)
        torch.cuda.synchronize(Ht.device)
        _check(lib().hsrq_solve_complex_device(self._h, n, Ht.data_ptr(), Ht.shape[1], m,
                                               wrt.data_ptr(), wit.data_ptr(), vrt.data_ptr(),
                                               vrt.shape[1], int(max_restarts),
                                               st.ctypes.data_as(_ip)),
               "hsrq_solve_complex_device")
        return vrt, st[:m]

    # -- problem generation (column-major via transposed storage, like solve) --

    def _drain(self):
        import torch
        torch.cuda.ExternalStream(self.stream_ptr()).synchronize()

    def gaussian_hessenberg(self, n: int, seed: int, device: str = "cuda"):
        """Tensor Ht (n, n), Ht[j, i] = H(i, j): N(0, 1) on and above the
        subdiagonal (Philox4x32-10 + Box-Muller, hsrq_gaussian_hessenberg_device)."""
        import torch
        Ht = torch.empty((n, n), dtype=torch.float64, device=device)
        torch.cuda.synchronize(Ht.device)
        _check(lib().hsrq_gaussian_hessenberg_device(self._h, n, seed & (2**64 - 1),
                                                     Ht.data_ptr(), n),
               "hsrq_gaussian_hessenberg_device")
        self._drain()
        return Ht

    def constructed_hessenberg(self, lam, nu, d, u, device: str = "cuda"):
        """Tensor Ht of H = U B U^-1 (B = diag(lam) + subdiag(nu), U = diag(d)
        + superdiag(u)): eigenvalues lam (hsrq_constructed_hessenberg_device)."""
        import torch
        lam, d = (np.ascontiguousarray(a, dtype=np.float64) for a in (lam, d))
        n = lam.size
        nu, u = (np.ascontiguousarray(a, dtype=np.float64) for a in (nu, u))
        if d.size != n or nu.size != max(n - 1, 0) or u.size != max(n - 1, 0):
            raise ValueError("lam, d need n entries and nu, u n-1")
        Ht = torch.empty((n, n), dtype=torch.float64, device=device)
        torch.cuda.synchronize(Ht.device)
        # a one-entry array stands in for nu, u when n == 1
        nu_, u_ = (a if a.size else np.zeros(1) for a in (nu, u))
        _check(lib().hsrq_constructed_hessenberg_device(
            self._h, n, lam.ctypes.data_as(_dp), nu_.ctypes.data_as(_dp),
            d.ctypes.data_as(_dp), u_.ctypes.data_as(_dp), Ht.data_ptr(), n),
            "hsrq_constructed_hessenberg_device")
        return Ht

    def grade(self, Ht, k, exact: bool = False):
        """In place: H(i, j) *= 2^(k_i - k_j) on the Hessenberg part, after the
        range check (hsrq_grade_device).  Returns (emax, emin), the extreme
        binary exponents of the graded entries; raises HsrqError (HSRQ_ERANGE,
        Ht untouched) on overflow or, if exact, on a result that rounds."""
        import torch
        k = np.ascontiguousarray(k, dtype=np.int32)
        _check_matrix(Ht, "Ht")
        n = Ht.shape[0]
        if k.size != n:
            raise ValueError("k needs n exponents")
        emax, emin = ctypes.c_int(), ctypes.c_int()
        torch.cuda.synchronize(Ht.device)
        _check(lib().hsrq_grade_device(self._h, n, Ht.data_ptr(), Ht.shape[1],
                                       k.ctypes.data_as(_ip), int(exact), ctypes.byref(emax),
                                       ctypes.byref(emin)), "hsrq_grade_device")
        return emax.value, emin.value


    def backtransform(self, Zt, Xt, Vt=None):
        """V = Z X on the context stream (hsrq_backtransform_device).  Zt (n, ldZ)
        with Zt[j, i] = Z(i, j); Xt (m, ldX) holds the eigenvectors of H as rows
        (like the solve's vrt); returns Vt (m, n) laid out the same way."""
        import torch
        _check_matrix(Zt, "Zt")
        _check_matrix(Xt, "Xt")
        n, m = Zt.shape[0], Xt.shape[0]
        if Vt is None:
            Vt = torch.empty((m, n), dtype=torch.float64, device=Xt.device)
        _check_matrix(Vt, "Vt", m)
        if min(Zt.shape[1], Xt.shape[1], Vt.shape[1]) < n:
            raise ValueError("Zt, Xt and Vt need leading dimensions >= n")
        torch.cuda.synchronize(Xt.device)
        _check(lib().hsrq_backtransform_device(self._h, n, m, Zt.data_ptr(), Zt.shape[1],
                                               Xt.data_ptr(), Xt.shape[1], Vt.data_ptr(),
                                               Vt.shape[1]), "hsrq_backtransform_device")
        self._drain()
        return Vt


class Plan:
    """Reusable device schedule for a set of problems (hsrq_plan)."""

    def __init__(self, ctx: Context, items, tlsz: int = 0):
        self.ctx = ctx
        self.items = items
        self._arr = Context._problems(items)
        self._h = _vp()
        _check(lib().hsrq_plan_create(ctx.handle, len(items), self._arr, tlsz,
                                      ctypes.byref(self._h)), "hsrq_plan_create")
        self.total_m = sum(w.numel() for _, w, _ in items)

    def execute(self, graph: bool = False):
        """Asynchronous on the context stream, after torch's current stream's
        queued work; sync() waits for it."""
        if self.items:
            _after_torch(self.ctx.stream_ptr(), self.items[0][0].device)
        f = lib().hsrq_plan_execute_graph if graph else lib().hsrq_plan_execute
        _check(f(self._h), "hsrq_plan_execute")

    KERNELS = ("norms", "init", "diag", "update", "backtransform")

    def execute_profiled(self) -> dict:
        """Synchronous run with HIP events per launch: {kernel: (ms, launches)}."""
        ms = (ctypes.c_double * 6)()
        cnt = (ctypes.c_int * 6)()
        _check(lib().hsrq_plan_execute_profiled(self._h, ms, cnt), "hsrq_plan_execute_profiled")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(self.KERNELS)}

    def sync(self):
        _check(lib().hsrq_plan_sync(self._h), "hsrq_plan_sync")

    def status(self) -> np.ndarray:
        st = np.zeros(max(self.total_m, 1), dtype=np.int32)
        _check(lib().hsrq_plan_status(self._h, st.ctypes.data_as(_ip)), "hsrq_plan_status")
        return st[:self.total_m]

    def workspace_bytes(self) -> int:
        return int(lib().hsrq_plan_workspace_bytes(self._h))

    def lanes(self) -> list:
        """[(tile, shifts per wavefront, problems)] per lane of the plan."""
        nl = lib().hsrq_plan_lanes(self._h, 0, None, None, None)
        if nl < 0:
            raise HsrqError(nl, "hsrq_plan_lanes")
        t, pk, pr = (ctypes.c_int * nl)(), (ctypes.c_int * nl)(), (ctypes.c_int * nl)()
        lib().hsrq_plan_lanes(self._h, nl, t, pk, pr)
        return [(t[i], pk[i], pr[i]) for i in range(nl)]

    def close(self):
        if self._h:
            lib().hsrq_plan_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
