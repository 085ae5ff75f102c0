"""The correction server from Python: a thin view of the library's server
(include/pmsz.h: pmsz_server_create, pmsz_submit, pmsz_poll), which does the
continuous batching in C++ -- lanes of correction contexts with one host thread
each, jobs admitted by an HBM budget as lanes and memory free up, small jobs of
one value type batched into one run, runs on different lanes overlapping on the
device.  The reference corrects one field per process run (main,
pMSz.cu:2596-3188).

Every job's result equals the job corrected alone (tests/test_gpu_serve.py,
tests/test_gpu_configs.py check this).  Times (t_submit, t_start, t_done) are
seconds of CLOCK_MONOTONIC, the clock of time.perf_counter().
"""
from __future__ import annotations

import ctypes
import threading

from . import engine


class ServerConfig(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("hbm_budget", ctypes.c_int64), ("alone_voxels", ctypes.c_int64)]


class JobInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("batch_size", ctypes.c_int), ("lane", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double)]


def _lib():
    L = engine.lib()
    if not getattr(L, "_pmsz_serve_bound", False):
        vp, P = ctypes.c_void_p, ctypes.POINTER(engine.Params)
        L.pmsz_default_server_config.argtypes = [ctypes.POINTER(ServerConfig)]
        L.pmsz_default_server_config.restype = None
        L.pmsz_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.pmsz_server_destroy.argtypes = [vp]
        L.pmsz_submit.argtypes = [vp, P, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
        L.pmsz_poll.argtypes = [vp, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(engine.Stats),
                                ctypes.POINTER(JobInfo)]
        L.pmsz_server_stats.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int64)] * 3
        L._pmsz_serve_bound = True
    return L


class Job:
    """A submitted correction: ``g`` (a CUDA tensor) is corrected in place."""

    def __init__(self, server, jid, f, g, dc, t_submit):
        self._server, self.id = server, jid
        self.f, self.g, self.dc = f, g, dc  # kept alive until the job is done
        self.t_submit = t_submit
        self.t_start = self.t_done = None
        self.stats = None
        self.error = None
        self.batch_size = 0
        self.lane = -1
        self._done = False
        self._lock = threading.Lock()

    @property
    def voxels(self):
        return self.f.numel()

    def _poll(self, timeout_ms):
        with self._lock:
            if self._done:
                return True
            st, info = engine.Stats(), JobInfo()
            rc = _lib().pmsz_poll(self._server._h, self.id, int(timeout_ms), ctypes.byref(st),
                                  ctypes.byref(info))
            if rc == engine.EBUSY:
                return False
            if rc != engine.OK:
                raise engine.PmszError(rc, "pmsz_poll")
            self.stats = st.as_dict()
            self.batch_size, self.lane = info.batch_size, info.lane
            self.t_start, self.t_done = info.t_start, info.t_done
            if info.status not in (engine.OK, engine.NOT_CONVERGED):
                self.error = engine.PmszError(info.status, "served correction")
            self._done = True
            return True

    def wait(self, timeout=None):
        """The job's statistics once it is done (TimeoutError after `timeout` s)."""
        if not self._poll(-1 if timeout is None else int(timeout * 1000)):
            raise TimeoutError("correction job not finished")
        if self.error is not None:
            raise self.error
        return self.stats

    def done(self):
        return self._poll(0)

    @property
    def latency_s(self):
        return None if self.t_done is None else self.t_done - self.t_submit


class Server:
    def __init__(self, lanes=4, hbm_budget=64 << 30, max_batch=32, alone_voxels=1 << 20):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("the correction server needs a HIP device")
        torch.cuda.init()
        cfg = ServerConfig(lanes, max_batch, int(hbm_budget), int(alone_voxels))
        h = ctypes.c_void_p()
        engine._check(_lib().pmsz_server_create(ctypes.byref(cfg), ctypes.byref(h)),
                      "pmsz_server_create")
        self._h = h
        self.lanes, self.hbm_budget, self.max_batch = lanes, int(hbm_budget), max_batch
        self._jobs = []  # polled to completion before the server is destroyed
        self._mu = threading.Lock()

    def submit(self, f, g, abs_eb, *, dc=None, **kw):
        """Queue the correction of CUDA tensor ``g`` against ``f``; returns a Job.
        The tensors are read once torch's current stream has reached this call.

        Raises ValueError for a job that could never be admitted (its workspace
        alone exceeds the budget)."""
        import torch
        if kw.get("sub", (1, 1, 1)) != (1, 1, 1):
            raise ValueError("the server corrects whole fields (sub = 1,1,1)")
        if not (f.is_cuda and g.is_cuda and f.is_contiguous() and g.is_contiguous()):
            raise ValueError("server inputs must be contiguous CUDA tensors")
        p = engine.make_params(f.shape, f.dtype, abs_eb, **kw)
        need = engine.workspace_bytes([(tuple(f.shape), f.dtype)])
        if need > self.hbm_budget:
            raise ValueError(f"job needs {need} B of workspace, budget is {self.hbm_budget} B")
        stream = torch.cuda.current_stream(f.device).cuda_stream
        jid = ctypes.c_uint64(0)
        dptr = ctypes.c_void_p(dc.data_ptr()) if dc is not None else None
        import time
        t = time.perf_counter()
        rc = _lib().pmsz_submit(self._h, ctypes.byref(p), ctypes.c_void_p(f.data_ptr()),
                                ctypes.c_void_p(g.data_ptr()), dptr, ctypes.c_void_p(stream),
                                ctypes.byref(jid))
        engine._check(rc, "pmsz_submit")
        job = Job(self, jid.value, f, g, dc, t)
        with self._mu:
            self._jobs = [j for j in self._jobs if not j._done]
            self._jobs.append(job)
        return job

    def _stats(self):
        r, d, q = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib().pmsz_server_stats(self._h, ctypes.byref(r), ctypes.byref(d), ctypes.byref(q))
        return r.value, d.value, q.value

    @property
    def runs(self):
        return self._stats()[0]

    @property
    def jobs_done(self):
        return self._stats()[1]

    def close(self):
        """Waits for every submitted job, then frees the server."""
        if self._h is not None and self._h.value:
            with self._mu:
                pending = list(self._jobs)
            for j in pending:
                j._poll(-1)
            _lib().pmsz_server_destroy(self._h)
        self._h = None
        self._jobs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
