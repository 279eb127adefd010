"""Request loop on one device (bmi_server_* in include/bmi.h).

Assembly, assemble-and-apply and apply requests are submitted from any
thread and run concurrently on the device, up to ``slots`` at a time, each
admitted when a slot is free and its N x N matrix fits beside the matrices
already resident.  The reference has no request loop; this replaces the
host-side batching its benchmarks did by hand (GPU_scheduler.jl:160-206 is the
only scheduling code it has).
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _bmi
from .api import GPUConfiguration, _c_operator
from .spaces import P1, RWG, Space


class Server:
    def __init__(self, device: int = 0, slots: int = 4, memory_budget: int = 0):
        self.h = ctypes.c_void_p()
        o = _bmi.ServerOptions(int(slots), int(memory_budget))
        _bmi.check(_bmi.lib().bmi_server_create(int(device), ctypes.byref(o), ctypes.byref(self.h)))
        self._keep: dict = {}  # id -> host arrays that must outlive the request
        self._lock = threading.Lock()

    def _request(self, kind: int, op, X: Space | None, config: GPUConfiguration | None, x, matrix: int = 0):
        r = _bmi.Request()
        r.kind = kind
        keep = []
        if X is not None:
            config = config or GPUConfiguration()
            xyz = np.ascontiguousarray(X.mesh.vertices, dtype=np.float64)
            tri = np.ascontiguousarray(X.mesh.faces, dtype=np.int32)
            gidx = np.ascontiguousarray(X.idx, dtype=np.int64)
            cf = np.ascontiguousarray(X.coeff, dtype=np.float64)
            keep += [xyz, tri, gidx, cf]
            r.xyz, r.nv = xyz.ctypes.data, xyz.shape[0]
            r.tri, r.nt = tri.ctypes.data, tri.shape[0]
            r.space_kind = {P1: _bmi.SPACE_LAGRANGE_C0D1, RWG: _bmi.SPACE_RWG}[X.kind]
            r.gidx, r.coeff, r.nfun = gidx.ctypes.data, cf.ctypes.data, int(X.nfun)
            r.op = _c_operator(op)
            r.opts = config.options()
            r.dtype = config.dtype()
        r.matrix = int(matrix)
        y = None
        if x is not None:
            xf = np.asfortranarray(np.asarray(x, dtype=np.complex128).reshape(x.shape[0], -1))
            keep.append(xf)
            r.nrhs = xf.shape[1]
            r.x = xf.ctypes.data
            y = np.empty(xf.shape, dtype=np.complex128, order="F")
        rid = ctypes.c_uint64()
        _bmi.check(_bmi.lib().bmi_server_submit(self.h, ctypes.byref(r), y.ctypes.data if y is not None else None,
                                                ctypes.byref(rid)))
        with self._lock:
            self._keep[rid.value] = (y, keep)
        return rid.value

    def submit_assemble_apply(self, op, X: Space, x: np.ndarray, config: GPUConfiguration | None = None) -> int:
        """Assemble Z(op; X, X), return y = Z x (x: nfun or nfun x nrhs, nrhs <= 8)."""
        return self._request(_bmi.REQ_ASSEMBLE_APPLY, op, X, config, x)

    def submit_assemble(self, op, X: Space, config: GPUConfiguration | None = None) -> int:
        """Assemble and keep the matrix resident; the id names it for applies."""
        return self._request(_bmi.REQ_ASSEMBLE, op, X, config, None)

    def submit_apply(self, matrix: int, x: np.ndarray) -> int:
        return self._request(_bmi.REQ_APPLY, None, None, None, x, matrix=matrix)

    def wait(self, rid: int, timeout_ms: int = -1):
        """(result dict, y or None); raises BMIError if the request failed or
        (status EBUSY) if it did not complete within timeout_ms."""
        res = _bmi.Result()
        _bmi.check(_bmi.lib().bmi_server_wait(self.h, ctypes.c_uint64(rid), int(timeout_ms), ctypes.byref(res)))
        with self._lock:
            y, _ = self._keep.pop(rid, (None, None))
        d = res.as_dict()
        if d["status"] != _bmi.OK:
            raise _bmi.BMIError(d["status"], d["message"])
        return d, y

    def release(self, matrix: int) -> None:
        _bmi.check(_bmi.lib().bmi_server_release(self.h, ctypes.c_uint64(matrix)))

    def memory(self) -> tuple[int, int]:
        cur, peak = ctypes.c_int64(), ctypes.c_int64()
        _bmi.check(_bmi.lib().bmi_server_memory(self.h, ctypes.byref(cur), ctypes.byref(peak)))
        return cur.value, peak.value

    def close(self) -> None:
        if self.h and self.h.value:
            _bmi.lib().bmi_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
