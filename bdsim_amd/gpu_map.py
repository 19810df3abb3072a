"""Run per-trajectory Python functions so that their `integrate` calls reach the
GPU together.

The reference spreads an ensemble over processes: `parallel_map(fn, items,
backend="processes")` (python/bdsim/parallel.py:62-93) hands each trajectory's
opaque function to a worker, and every `integrate` it makes runs one chain on
one core. On one GPU the chains of an ensemble must be stepped together.
`gpu_map(call, items)` is the backend that the reference's `parallel_map`
switch (parallel.py:79-90) dispatches to for backend="gpu": it runs the items in a bounded pool of threads, each enlisted in one
`IntegrateBatcher`, so that `bdsim_amd.core.integrate` calls made by the items
are gathered and issued as batched launches. A launch goes out when every
worker is waiting in `integrate` or finished, or when requests have waited
for `flush_after_s` without a new arrival (so a slow Python function holds
the others back by at most that long). Each chain's arithmetic on the device
is independent of the others in its launch, so results are the same, bit for
bit, whatever the grouping.
"""
import threading

from . import _abi
from .batcher import IntegrateBatcher

DEFAULT_MAX_WORKERS = 1024
WORKER_STACK_BYTES = 512 << 10


def gpu_map(call, items, n_workers=None, precision=_abi.PRECISION_FP64, *,
            flush_after_s=0.02, runner=None):
    """[call(x) for x in items] with the items' integrate calls batched on the
    GPU. `n_workers` threads run items (default min(len(items), 1024)): it
    bounds how many chains one launch can hold. An exception raised by
    `call` propagates after every worker has stopped (the one of the lowest
    item index). `runner` replaces the GPU launch (tests)."""
    items = list(items)
    n = len(items)
    if n == 0:
        return []
    W = min(n, int(n_workers) if n_workers else DEFAULT_MAX_WORKERS)
    if W < 1:
        raise ValueError(f"n_workers must be positive, got {n_workers!r}")
    out = [None] * n
    errors = {}
    batcher = IntegrateBatcher(W, precision=precision, runner=runner, flush_after_s=flush_after_s)
    cursor = [0]
    lock = threading.Lock()

    def worker():
        batcher.enlist()
        try:
            while True:
                with lock:
                    i = cursor[0]
                    cursor[0] += 1
                if i >= n:
                    return
                try:
                    out[i] = call(items[i])
                except BaseException as exc:     # re-raised in the caller
                    errors[i] = exc
        finally:
            batcher.leave()

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(W)]
    old = threading.stack_size()
    try:
        threading.stack_size(WORKER_STACK_BYTES)     # applies to threads started now
        for t in threads:
            t.start()
    finally:
        threading.stack_size(old)
    for t in threads:
        t.join()
    if errors:
        raise errors[min(errors)]
    return out
