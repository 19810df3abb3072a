"""Batching of concurrent `integrate` calls onto one GPU launch.

The reference runs trajectories in separate processes, each calling
`_bdsim.integrate` for its own chain (python/bdsim/parallel.py:62-93,
ensemble.py:106-138). On one GPU the work of thousands of chains must go out
together, but the per-trajectory functions handed to `run_ensemble` are opaque
Python. So `gpu_map` runs each item in a worker thread, and every `integrate`
those workers make is routed here: a call parks its request, and when every
live worker is parked (or finished) -- or when the parked requests have seen no
new arrival for `flush_after_s` -- a waiting thread runs all pending requests
as ONE `bdsim_integrate_batch` call, then wakes the others. Trajectories are
independent and seeded by index, so the results do not depend on how calls
are grouped.
"""
import ctypes
import threading
import time

import numpy as np

from . import _abi

_local = threading.local()


def current():
    """The batcher the calling thread is enlisted in, or None."""
    return getattr(_local, "batcher", None)


class _Request:
    __slots__ = ("R", "phys", "sim", "rng", "done", "out", "error", "nonconverged")

    def __init__(self, R, phys, sim, rng):
        self.R, self.phys, self.sim, self.rng = R, phys, sim, rng
        self.done = False
        self.out = None
        self.error = None
        self.nonconverged = 0


class IntegrateBatcher:
    """Groups integrate() calls of `n_workers` cooperating threads."""

    def __init__(self, n_workers, precision=_abi.PRECISION_FP64, runner=None, flush_after_s=None):
        self.cv = threading.Condition()
        self.pending = []
        self.active = int(n_workers)
        self.precision = precision
        self._runner = runner or self._run_on_gpu
        # pending requests that saw no new arrival for this long are launched
        # without waiting for the remaining workers (None: wait for all)
        self.flush_after_s = flush_after_s
        self.last_arrival = time.monotonic()
        self.batches = 0            # diagnostics: launches made
        self.batched_calls = 0

    # ---- worker side -------------------------------------------------------
    def enlist(self):
        _local.batcher = self

    def leave(self):
        _local.batcher = None
        with self.cv:
            self.active -= 1
            batch = self._take_if_complete()
        if batch:
            self._execute(batch)

    def integrate(self, R, phys, sim, rng):
        req = _Request(R, phys, sim, rng)
        with self.cv:
            self.pending.append(req)
            self.last_arrival = time.monotonic()
            batch = self._take_if_complete()
        if batch:
            self._execute(batch)
        while True:
            with self.cv:
                if req.done:
                    break
                batch = self._take_if_stale()
                if batch is None:
                    self.cv.wait(self.flush_after_s)
                    continue
            self._execute(batch)
        if req.nonconverged:
            from .core import _report_nonconvergence
            _report_nonconvergence(req.nonconverged, R.shape[0], sim.implicit_loop_tol)
        if req.error is not None:
            raise req.error
        return req.out

    # ---- batching ------------------------------------------------------------
    def _take_if_complete(self):
        # caller holds the lock: run once every live worker is waiting here
        if self.pending and len(self.pending) >= self.active:
            batch, self.pending = self.pending, []
            return batch
        return None

    def _take_if_stale(self):
        # caller holds the lock: requests that waited flush_after_s without a
        # new arrival go out even though some workers are still busy
        if self.flush_after_s is None or not self.pending:
            return None
        if time.monotonic() - self.last_arrival < self.flush_after_s:
            return None
        batch, self.pending = self.pending, []
        return batch

    def _execute(self, batch):
        try:
            self._runner(batch)
        except Exception as exc:     # an API failure fails every request of the batch
            for r in batch:
                r.error = exc
        with self.cv:
            self.batches += 1
            self.batched_calls += len(batch)
            for r in batch:
                r.done = True
            self.cv.notify_all()

    def _run_on_gpu(self, batch):
        from .core import status_error
        lib = _abi.load_gpu()
        n = len(batch)
        nb = np.array([r.R.shape[0] for r in batch], dtype=np.int32)
        R = np.ascontiguousarray(np.concatenate([r.R.ravel() for r in batch]))
        packs, cache = [], {}
        for r in batch:
            key = (id(r.phys), r.R.shape[0])
            if key not in cache:
                cache[key] = _abi.PackedPhys(r.phys, r.R.shape[0])
            packs.append(cache[key])
        phys = (_abi.CPhys * n)(*[p.c for p in packs])
        sims = (_abi.CSim * n)(*[_abi.pack_sim(r.sim) for r in batch])
        rngs = (_abi.CRng * n)(*[_abi.CRng(*r.rng._state()) for r in batch])
        status = np.zeros(n, dtype=np.int32)
        cnt = (_abi.CCounters * n)()
        _abi.check(lib, lib.bdsim_integrate_batch(n, _abi.iptr(nb), _abi.dptr(R), phys, sims,
                                                  rngs, int(self.precision), _abi.iptr(status),
                                                  cnt, None, 0))
        off = 0
        for i, r in enumerate(batch):
            N = int(nb[i])
            r.rng._set_state(rngs[i].ix, rngs[i].iy, rngs[i].seed)
            r.out = R[3 * off:3 * (off + N)].reshape(N, 3).copy()
            r.nonconverged = int(cnt[i].nonconverged)
            r.error = status_error(int(status[i]), cnt[i])
            off += N
