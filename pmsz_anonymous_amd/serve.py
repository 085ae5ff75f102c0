"""Serving loop for correction jobs: continuous batching over device lanes.

The reference corrects one field per process run (main, pMSz.cu:2596-3188).  A
server keeps ``lanes`` correction contexts (pmsz_ctx: each with its own HIP
stream and scratch state) and one host thread per lane.  Jobs are admitted as
lanes and memory free up: a lane that finishes its run takes every queued job it
can admit at once -- same value type and loop mode, at most ``max_batch`` jobs,
and only while the library-owned device memory of all running work (the
workspace of each lane's run, pmsz_workspace_bytes, or what the lane's context
already holds if that is more, idle lanes included) stays within ``hbm_budget`` -- and corrects them
as one batched run (pmsz_ctx_correct_batch).  A job of ``alone_voxels`` voxels or
more runs by itself: its full-volume passes fill the GPU anyway, and a field
alone takes the one-domain schedule (streamed applies, claimed work lists).
Runs on different lanes overlap on the device: the dense passes of one run fill
the GPU while another run is in its latency-bound tail.

Every job's result equals the job corrected alone (the batched run is exact per
field, and lanes share no state); tests/test_gpu_serve.py checks this.
"""
from __future__ import annotations

import collections
import threading
import time

from . import engine


class Job:
    """A submitted correction: ``g`` (a CUDA tensor) is corrected in place."""

    def __init__(self, f, g, abs_eb, kw, dc=None):
        self.f, self.g, self.abs_eb, self.kw, self.dc = f, g, abs_eb, kw, dc
        self.t_submit = time.perf_counter()
        self.t_start = None
        self.t_done = None
        self.stats = None
        self.error = None
        self.batch_size = 0
        self.lane = -1
        self._ev = threading.Event()

    @property
    def key(self):
        return (str(self.f.dtype), self.kw.get("mode", engine.LAZY),
                self.kw.get("max_sweeps", 100000))

    @property
    def voxels(self):
        return self.f.numel()

    def wait(self, timeout=None):
        if not self._ev.wait(timeout):
            raise TimeoutError("correction job not finished")
        if self.error is not None:
            raise self.error
        return self.stats

    def done(self):
        return self._ev.is_set()

    @property
    def latency_s(self):
        return None if self.t_done is None else self.t_done - self.t_submit


class Server:
    def __init__(self, lanes=4, hbm_budget=64 << 30, max_batch=32, alone_voxels=1 << 20):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("the correction server needs a HIP device")
        self.lanes = lanes
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.alone_voxels = int(alone_voxels)
        self._ctxs = [engine.Context() for _ in range(lanes)]
        self._reserved = [0] * lanes   # workspace of each lane's running group
        self._held = [0] * lanes       # bytes each lane's context holds (grow-only)
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self.runs = 0
        self.jobs_done = 0
        self._threads = [threading.Thread(target=self._lane, args=(k,), daemon=True)
                         for k in range(lanes)]
        for t in self._threads:
            t.start()

    def submit(self, f, g, abs_eb, *, dc=None, **kw):
        """Queue the correction of CUDA tensor ``g`` against ``f``; returns a Job.

        Raises ValueError for a job that could never be admitted (its workspace
        alone exceeds the budget)."""
        if kw.get("sub", (1, 1, 1)) != (1, 1, 1):
            raise ValueError("the server corrects whole fields (sub = 1,1,1)")
        need = engine.workspace_bytes([(tuple(f.shape), f.dtype)])
        if need > self.hbm_budget:
            raise ValueError(f"job needs {need} B of workspace, budget is {self.hbm_budget} B")
        job = Job(f, g, abs_eb, kw, dc)
        # the tensors are ready once the submitter's stream reaches this point
        import torch
        job.ready = torch.cuda.Event()
        job.ready.record()
        with self._cv:
            if self._stop:
                raise RuntimeError("server is closed")
            self._queue.append(job)
            self._cv.notify()
        return job

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        for c in self._ctxs:
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- admission -----------------------------------------------------------
    def _take(self, lane):
        """Jobs this lane admits now (caller holds the lock), oldest first."""
        if not self._queue:
            return []
        others = sum(max(r, h) for k, (r, h) in enumerate(zip(self._reserved, self._held))
                     if k != lane)
        have = self._held[lane]
        head = self._queue[0]
        key = head.key
        alone = head.voxels >= self.alone_voxels
        group, shapes = [], []
        for job in list(self._queue):
            # the head, and with a small head the small jobs of its key
            if job is not head and (alone or job.key != key or job.voxels >= self.alone_voxels):
                continue
            trial = shapes + [(tuple(job.f.shape), job.f.dtype)]
            need = max(have, engine.workspace_bytes(trial))
            if others + need > self.hbm_budget:
                break
            group.append(job)
            shapes = trial
            if alone or len(group) >= self.max_batch:
                break
        if group:
            for job in group:
                self._queue.remove(job)
            self._reserved[lane] = max(have, engine.workspace_bytes(shapes))
        return group

    def _lane(self, lane):
        ctx = self._ctxs[lane]
        while True:
            with self._cv:
                group = self._take(lane)
                while not group:
                    if self._stop and not self._queue:
                        return
                    if self._queue and self._held[lane]:
                        break  # memory this idle lane holds may be what blocks admission
                    self._cv.wait(0.05)
                    group = self._take(lane)
            if not group:
                ctx.trim()
                with self._cv:
                    self._held[lane] = 0
                    self._cv.notify_all()
                continue
            t0 = time.perf_counter()
            for job in group:
                job.t_start = t0
                job.lane = lane
                job.batch_size = len(group)
            try:
                per = [{k: v for k, v in j.kw.items() if k not in ("mode", "max_sweeps")}
                       for j in group]
                common = {k: v for k, v in group[0].kw.items() if k in ("mode", "max_sweeps")}
                dcs = None
                if any(j.dc is not None for j in group):
                    dcs = [j.dc for j in group]
                # the jobs' tensors were made on the submitting threads' streams
                for job in group:
                    job.ready.synchronize()
                sts = ctx.correct_batch([j.f for j in group], [j.g for j in group],
                                        [j.abs_eb for j in group], dcs=dcs, per_field=per,
                                        raise_not_converged=False, sync_torch=False, **common)
                err = None
            except Exception as e:  # reported on every job of the run
                sts, err = [None] * len(group), e
            t1 = time.perf_counter()
            held = ctx.bytes()
            with self._cv:
                self._reserved[lane] = 0
                self._held[lane] = held
                self.runs += 1
                self.jobs_done += len(group)
                self._cv.notify_all()
            for job, st in zip(group, sts):
                job.stats, job.error, job.t_done = st, err, t1
                job._ev.set()
