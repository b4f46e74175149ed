"""Continuous-batching request server (binding of hsrq_server_* in hsrq.h).

    srv = Server(ctx, hbm_bytes=8 << 30)
    rid = srv.submit(H, wr)            # host arrays: copied at admission
    vr, wr_final, status, ms = srv.wait(rid)
    srv.stats()                        # p50/p99 latency, eigvecs, HBM use

On-disk requests and responses (a spool directory; format below):

    write_request("in/a.hsrq", H, wr)            # or H, wr, wi: complex shifts
    serve_spool(ctx, "in", "out")                # python -m hsrq3in_amd.serve in out
    read_response("out/a.hsrs")                  # vr, wr, status, latency

Device requests pass torch tensors (column-major through transposed
storage, as ``Context.solve``); their results are written in place.
"""
from __future__ import annotations

import ctypes
import os

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
                ("hbm_in_use", ctypes.c_size_t), ("hbm_peak", ctypes.c_size_t),
                ("validated", ctypes.c_int64), ("worst_residual", ctypes.c_double)]


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
        import torch

        from . import _check_problem
        n, m = _check_problem(Ht, wrt, vrt)
        # the server's streams do not wait for torch's: the inputs must be
        # complete before the worker reads them at admission
        torch.cuda.current_stream(Ht.device).synchronize()
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


# ---------------------------------------------------------------------------
# On-disk request / response format (a spool directory served in batches)
#
# Request file  (*.hsrq, little-endian):
#   8 bytes  magic "HSRQREQ1"
#   4 x int32  n, m, flags (bit 0: complex shifts; bit 1: no shifts given,
#              use the first m real eigenvalues of H), reserved (0)
#   n*n float64  H, column-major (leading dimension n)
#   m float64    wr              (absent with flag bit 1)
#   m float64    wi              (only with the complex flag)
# Response file (*.hsrs):
#   8 bytes  magic "HSRQRES1"
#   4 x int32  n, m, flags (as the request), rc (0, or a negative HSRQ_E* code)
#   m float64    wr in final order (restarts move converged shifts first)
#   m int32      status bits (1 converged, 2 perturbed pivot)
#   1 float64    latency in ms (submission to completion)
#   n*m float64  vr column-major; with the complex flag n*2m (column 2k the
#                real, 2k+1 the imaginary part of x_k)
# A response is written to a temporary name and renamed into place, so a
# reader never sees a partial file.
# ---------------------------------------------------------------------------

REQ_MAGIC = b"HSRQREQ1"
RES_MAGIC = b"HSRQRES1"
FLAG_COMPLEX = 1
FLAG_EIGENVALUE_SHIFTS = 2
_HDR = np.dtype([("magic", "S8"), ("n", "<i4"), ("m", "<i4"), ("flags", "<i4"),
                 ("aux", "<i4")])


def write_request(path: str, H: np.ndarray, wr, wi=None, m: int | None = None) -> None:
    """wr (and wi): the shifts; or wr=None and m: the server takes the first m
    real eigenvalues of H as shifts."""
    H = np.asfortranarray(H, dtype="<f8")
    n = H.shape[0]
    if H.shape != (n, n):
        raise ValueError("H must be square")
    if wr is None:
        if m is None or m < 0 or wi is not None:
            raise ValueError("without shifts give m (and no wi)")
        hdr = np.array([(REQ_MAGIC, n, m, FLAG_EIGENVALUE_SHIFTS, 0)], dtype=_HDR)
        tmp = path + ".part"
        with open(tmp, "wb") as f:
            f.write(hdr.tobytes())
            f.write(H.tobytes(order="F"))
        os.replace(tmp, path)
        return
    wr = np.ascontiguousarray(wr, dtype="<f8")
    flags = 0
    if wi is not None:
        wi = np.ascontiguousarray(wi, dtype="<f8")
        if wi.shape != wr.shape:
            raise ValueError("wr and wi differ in length")
        flags |= FLAG_COMPLEX
    hdr = np.array([(REQ_MAGIC, n, wr.size, flags, 0)], dtype=_HDR)
    tmp = path + ".part"
    with open(tmp, "wb") as f:
        f.write(hdr.tobytes())
        f.write(H.tobytes(order="F"))
        f.write(wr.tobytes())
        if wi is not None:
            f.write(wi.tobytes())
    os.replace(tmp, path)


def _read_header(path: str, magic: bytes):
    hdr = np.fromfile(path, dtype=_HDR, count=1)
    if hdr.size != 1 or hdr["magic"][0] != magic:
        raise ValueError(f"{path}: not a {magic.decode()} file")
    n, m, flags, aux = (int(hdr[f][0]) for f in ("n", "m", "flags", "aux"))
    if (n <= 0 or m < 0 or flags & ~(FLAG_COMPLEX | FLAG_EIGENVALUE_SHIFTS)
            or flags == FLAG_COMPLEX | FLAG_EIGENVALUE_SHIFTS
            or (magic == RES_MAGIC and flags & FLAG_EIGENVALUE_SHIFTS)):
        raise ValueError(f"{path}: invalid header (n={n}, m={m}, flags={flags})")
    return n, m, flags, aux


def read_request(path: str):
    """(H column-major, wr, wi or None); H is memory-mapped.  wr is None for a
    request that asks for H's first m real eigenvalues as shifts (then the
    fourth value is m)."""
    n, m, flags, _ = _read_header(path, REQ_MAGIC)
    if flags & FLAG_EIGENVALUE_SHIFTS:
        need = _HDR.itemsize + 8 * n * n
        if os.path.getsize(path) != need:
            raise ValueError(f"{path}: size {os.path.getsize(path)} != {need}")
        H = np.memmap(path, dtype="<f8", mode="r", offset=_HDR.itemsize, shape=(n, n), order="F")
        return H, None, None, m
    cplx = bool(flags & FLAG_COMPLEX)
    need = _HDR.itemsize + 8 * (n * n + m * (2 if cplx else 1))
    if os.path.getsize(path) != need:
        raise ValueError(f"{path}: size {os.path.getsize(path)} != {need}")
    off = _HDR.itemsize
    H = np.memmap(path, dtype="<f8", mode="r", offset=off, shape=(n, n), order="F")
    wr = np.fromfile(path, dtype="<f8", count=m, offset=off + 8 * n * n)
    wi = (np.fromfile(path, dtype="<f8", count=m, offset=off + 8 * (n * n + m))
          if cplx else None)
    return H, wr, wi, m


def write_response(path: str, vr, wr, status, latency_ms: float, rc: int = 0,
                   complex_shifts: bool = False) -> None:
    vr = np.asfortranarray(vr, dtype="<f8")
    n = vr.shape[0]
    wr = np.ascontiguousarray(wr, dtype="<f8")
    m = wr.size
    if vr.shape[1] != (2 * m if complex_shifts else m):
        raise ValueError("vr has the wrong number of columns")
    hdr = np.array([(RES_MAGIC, n, m, FLAG_COMPLEX if complex_shifts else 0, rc)], dtype=_HDR)
    tmp = path + ".part"
    with open(tmp, "wb") as f:
        f.write(hdr.tobytes())
        f.write(wr.tobytes())
        f.write(np.ascontiguousarray(status, dtype="<i4").tobytes())
        f.write(np.array([latency_ms], dtype="<f8").tobytes())
        f.write(vr.tobytes(order="F"))
    os.replace(tmp, path)


def read_response(path: str):
    """dict with n, m, rc, wr, status, latency_ms, vr (complex for complex shifts)."""
    n, m, flags, rc = _read_header(path, RES_MAGIC)
    cplx = bool(flags & FLAG_COMPLEX)
    cols = 2 * m if cplx else m
    need = _HDR.itemsize + 8 * m + 4 * m + 8 + 8 * n * cols
    if os.path.getsize(path) != need:
        raise ValueError(f"{path}: size {os.path.getsize(path)} != {need}")
    off = _HDR.itemsize
    wr = np.fromfile(path, dtype="<f8", count=m, offset=off)
    st = np.fromfile(path, dtype="<i4", count=m, offset=off + 8 * m)
    lat = float(np.fromfile(path, dtype="<f8", count=1, offset=off + 12 * m)[0])
    vr = np.fromfile(path, dtype="<f8", count=n * cols, offset=off + 12 * m + 8)
    vr = vr.reshape((n, cols), order="F")
    if cplx:
        vr = vr[:, 0::2] + 1j * vr[:, 1::2]
    return {"n": n, "m": m, "rc": rc, "wr": wr, "status": st, "latency_ms": lat, "vr": vr}


def serve_spool(ctx, in_dir: str, out_dir: str, hbm_bytes: int = 8 << 30,
                max_active: int = 1024, shift_cache=None) -> dict:
    """Serves every *.hsrq request in in_dir that has no response yet: real
    shifts through the continuous-batching server (all admitted together),
    complex shifts through hsrq_solve_complex_device.  Requests without
    shifts get H's first m real eigenvalues from ``shift_cache`` (a
    workloads.ShiftCache; one in out_dir/.shifts by default).  Writes
    <name>.hsrs into out_dir.  Returns {name: rc}."""
    import time
    from .workloads import ShiftCache
    os.makedirs(out_dir, exist_ok=True)
    names = sorted(f[:-5] for f in os.listdir(in_dir) if f.endswith(".hsrq"))
    names = [b for b in names if not os.path.exists(os.path.join(out_dir, b + ".hsrs"))]
    done = {}
    srv = Server(ctx, hbm_bytes, max_active)
    try:
        pending = []
        for b in names:
            try:
                H, wr, wi, m = read_request(os.path.join(in_dir, b + ".hsrq"))
            except ValueError:
                done[b] = -1
                continue
            if wr is None:
                if shift_cache is None:
                    shift_cache = ShiftCache(os.path.join(out_dir, ".shifts"))
                try:
                    wr = shift_cache.get(H, m)
                except ValueError:   # fewer real eigenvalues than m
                    n = H.shape[0]
                    write_response(os.path.join(out_dir, b + ".hsrs"), np.zeros((n, m)),
                                   np.zeros(m), np.zeros(m, np.int32), 0.0, -1)
                    done[b] = -1
                    continue
            if wi is not None:
                t0 = time.perf_counter()
                vr, wro, st, rc = _complex_request(ctx, H, wr, wi)
                write_response(os.path.join(out_dir, b + ".hsrs"), vr, wro, st,
                               1e3 * (time.perf_counter() - t0), rc, complex_shifts=True)
                done[b] = rc
                continue
            try:
                pending.append((b, srv.submit(np.asarray(H), wr), H.shape[0], wr.size))
            except HsrqError as e:
                n = H.shape[0]
                write_response(os.path.join(out_dir, b + ".hsrs"), np.zeros((n, wr.size)), wr,
                               np.zeros(wr.size, np.int32), 0.0, e.code)
                done[b] = e.code
        for b, rid, n, m in pending:
            vr, wro, st, ms = srv.wait(rid)
            write_response(os.path.join(out_dir, b + ".hsrs"), vr, wro, st, ms, 0)
            done[b] = 0
    finally:
        srv.close()
    return done


def _complex_request(ctx, H, wr, wi):
    import torch
    from . import colmajor_to_device
    n, m = H.shape[0], wr.size
    try:
        Ht = colmajor_to_device(np.array(H, order="F"))   # the request may be a read-only map
        wrt = torch.from_numpy(np.array(wr)).cuda()
        wit = torch.from_numpy(np.array(wi)).cuda()
        vrt, st = ctx.solve_complex(Ht, wrt, wit)
        return np.asfortranarray(vrt.cpu().numpy().T), wr, st, 0
    except HsrqError as e:
        return np.zeros((n, 2 * m)), wr, np.zeros(m, np.int32), e.code


def main(argv=None) -> int:
    import argparse
    from . import Context
    ap = argparse.ArgumentParser(description="serve a spool directory of *.hsrq requests")
    ap.add_argument("in_dir")
    ap.add_argument("out_dir")
    ap.add_argument("--hbm-gib", type=float, default=8.0)
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = Context(a.device)
    try:
        done = serve_spool(ctx, a.in_dir, a.out_dir, int(a.hbm_gib * (1 << 30)))
    finally:
        ctx.close()
    bad = {k: v for k, v in done.items() if v != 0}
    print(f"served {len(done)} requests, {len(bad)} failed")
    return 1 if bad else 0


if __name__ == "__main__":
    raise SystemExit(main())
