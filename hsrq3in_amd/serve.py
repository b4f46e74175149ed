"""Continuous-batching request server (binding of hsrq_server_* in hsrq.h).

    srv = Server(ctx, hbm_bytes=8 << 30)
    rid = srv.submit(H, wr)            # host arrays: copied at admission
    vr, wr_final, status, ms = srv.wait(rid)
    srv.stats()                        # p50/p99 latency, eigvecs, HBM use

Device requests pass torch tensors (column-major through transposed
storage, as ``Context.solve``); their results are written in place.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import HsrqError, _check, _dp, _ip, _vp, lib

ON_HOST = 0
ON_DEVICE = 1


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("H", _vp), ("ldH", ctypes.c_int), ("m", ctypes.c_int),
                ("wr", _vp), ("vr", _vp), ("ldvr", ctypes.c_int), ("location", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64),
                ("eigvecs", ctypes.c_int64), ("active", ctypes.c_int64),
                ("queued", ctypes.c_int64), ("ticks", ctypes.c_int64),
                ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("span_s", ctypes.c_double), ("hbm_bytes", ctypes.c_size_t),
                ("hbm_in_use", ctypes.c_size_t), ("hbm_peak", ctypes.c_size_t)]


def _bind():
    L = lib()
    if getattr(L, "_server_bound", False):
        return L
    L.hsrq_server_create.argtypes = [_vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(_vp)]
    L.hsrq_server_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]
    L.hsrq_server_poll.argtypes = [_vp, ctypes.c_int64]
    L.hsrq_server_wait.argtypes = [_vp, ctypes.c_int64, _dp, _ip, ctypes.POINTER(ctypes.c_double)]
    L.hsrq_server_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
    L.hsrq_server_destroy.restype = None
    L.hsrq_server_destroy.argtypes = [_vp]
    L._server_bound = True
    return L


class Server:
    """hsrq_server: requests of any size, admitted by HBM footprint."""

    def __init__(self, ctx, hbm_bytes: int, max_active: int = 1024, tile: int = 0):
        self._L = _bind()
        self.ctx = ctx
        self._h = _vp()
        _check(self._L.hsrq_server_create(ctx.handle, int(hbm_bytes), max_active, tile,
                                          ctypes.byref(self._h)), "hsrq_server_create")
        self._keep = {}   # id -> buffers that must outlive the request

    def submit(self, H: np.ndarray, wr: np.ndarray) -> int:
        """Host request: H (n x n) and wr are copied to HBM at admission."""
        H = np.asfortranarray(H, dtype=np.float64)
        wr = np.ascontiguousarray(wr, dtype=np.float64)
        n, m = H.shape[0], wr.size
        vr = np.zeros((n, m), dtype=np.float64, order="F")
        req = Request(n, H.ctypes.data, n, m, wr.ctypes.data if m else None,
                      vr.ctypes.data if m else None, n, ON_HOST)
        rid = ctypes.c_int64()
        _check(self._L.hsrq_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)),
               "hsrq_server_submit")
        self._keep[rid.value] = (H, wr, vr)
        return rid.value

    def submit_device(self, Ht, wrt, vrt) -> int:
        """Device request on torch tensors (Ht[j, i] = H(i, j); vrt (m, ldvr))."""
        n, m = Ht.shape[0], wrt.numel()
        req = Request(n, Ht.data_ptr(), Ht.shape[1], m, wrt.data_ptr(), vrt.data_ptr(),
                      vrt.shape[1], ON_DEVICE)
        rid = ctypes.c_int64()
        _check(self._L.hsrq_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid)),
               "hsrq_server_submit")
        self._keep[rid.value] = (Ht, wrt, vrt)
        return rid.value

    def poll(self, rid: int) -> bool:
        rc = self._L.hsrq_server_poll(self._h, rid)
        if rc < 0:
            raise HsrqError(rc, "hsrq_server_poll")
        return rc == 1

    def wait(self, rid: int):
        """(vr, wr in final order, status, latency ms); vr is None for device
        requests (written in place)."""
        bufs = self._keep.pop(rid)
        m = bufs[1].size if isinstance(bufs[1], np.ndarray) else bufs[1].numel()
        wr_out = np.zeros(max(m, 1), dtype=np.float64)
        st = np.zeros(max(m, 1), dtype=np.int32)
        ms = ctypes.c_double()
        _check(self._L.hsrq_server_wait(self._h, rid, wr_out.ctypes.data_as(_dp),
                                        st.ctypes.data_as(_ip), ctypes.byref(ms)),
               "hsrq_server_wait")
        vr = bufs[2] if isinstance(bufs[2], np.ndarray) else None
        return vr, wr_out[:m], st[:m], ms.value

    def stats(self) -> dict:
        s = Stats()
        _check(self._L.hsrq_server_get_stats(self._h, ctypes.byref(s)), "hsrq_server_get_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def close(self):
        if self._h:
            self._L.hsrq_server_destroy(self._h)
            self._h = _vp()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
