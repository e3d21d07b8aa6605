"""Serving loop binding (rf_server_* of include/rf_capi.h).

A Server owns a device memory budget, the device-resident lattices of
methods 1 and 2, and a worker thread that batches queued requests into
grouped launches.  submit() is thread-safe and returns at once; wait()
blocks until the request's field is ready; read() copies it to the host;
release() frees it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _check_device
from ._capi import (FOM_NONE, FORM_MFMA_F16, Request, RequestResult, ServerConfig, ServerStats, check,
                    lib)


class Server:
    def __init__(self, ctx, hbm_budget: int = 0, max_inflight: int = 3, max_batch: int = 64,
                 form: int = FORM_MFMA_F16, max_nmc: int = 1024, xrange_norm=(10.0, 10.0, 10.0),
                 batch_terms: float = 2e11):
        self.ctx = ctx
        cfg = ServerConfig()
        cfg.hbmBudget, cfg.maxInflight, cfg.maxBatch, cfg.form, cfg.maxNmc = (
            int(hbm_budget), int(max_inflight), int(max_batch), int(form), int(max_nmc))
        cfg.xRangeNorm = (C.c_double * 3)(*[float(v) for v in xrange_norm])
        cfg.batchTerms = float(batch_terms)
        self._h = C.c_void_p()
        check(lib().rf_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)), ctx.handle)
        ctx._adopt(self)
        self._keep = {}

    def submit(self, x, method: int, Nmc: int, corrL, seed: int, margi: int = FOM_NONE,
               avg: float = 0.0, var: float = 1.0) -> int:
        """x: device tensor (N, 3) float64, or host ndarray (N, 3) float64
        (kept alive here until the request is released)."""
        r = Request()
        r.method, r.Nmc, r.margiFirst = int(method), int(Nmc), int(margi)
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float64)
            r.xOnDevice, r.x = 0, x.ctypes.data
        else:
            _check_device(x, "x", ("float64",), device=getattr(self.ctx, "device", None))
            r.xOnDevice, r.x = 1, x.data_ptr()
        if x.ndim != 2 or int(x.shape[1]) != 3:
            raise ValueError(f"x has shape {tuple(x.shape)}; expected (N, 3)")
        r.N = int(x.shape[0])
        r.corrL = (C.c_double * 3)(*[float(v) for v in corrL])
        r.fieldAvg, r.fieldVar, r.seed = float(avg), float(var), int(seed)
        rid = C.c_int64(0)
        check(lib().rf_server_submit(self._h, C.byref(r), C.byref(rid)), self.ctx.handle)
        self._keep[rid.value] = x
        return int(rid.value)

    def wait(self, rid: int, timeout: float = -1.0) -> RequestResult:
        res = RequestResult()
        check(lib().rf_server_wait(self._h, int(rid), float(timeout), C.byref(res)), self.ctx.handle)
        return res

    def read(self, rid: int) -> np.ndarray:
        res = self.wait(rid)
        if res.status != 0:
            raise RuntimeError(f"request {rid} failed with status {res.status}")
        out = np.empty((res.Nmc, res.N), dtype=np.float32)
        check(lib().rf_server_read(self._h, int(rid), out.ctypes.data), self.ctx.handle)
        return out

    def release(self, rid: int) -> None:
        check(lib().rf_server_release(self._h, int(rid)), self.ctx.handle)
        self._keep.pop(int(rid), None)

    def stats(self) -> ServerStats:
        st = ServerStats()
        check(lib().rf_server_get_stats(self._h, C.byref(st)), self.ctx.handle)
        return st

    def close(self) -> None:
        if self._h:
            lib().rf_server_destroy(self._h)
            self._h = C.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
