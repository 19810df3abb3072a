"""Parallel execution of trajectory ensembles (python/bdsim/parallel.py:1-93).

Backends:
  * "serial"    - a plain loop;
  * "processes" - a ProcessPoolExecutor, as in the reference (each worker's
                  integrate calls still run on the GPU, one chain at a time);
  * "gpu"       - the MI355X batch plug point: items run in worker threads, and the
                  `integrate` calls they make are gathered and executed as one
                  batched launch per round (see bdsim_amd/batcher.py). Results are
                  identical to "serial" because trajectories are independent and
                  seeded by index; they are just computed together.
"""
import threading

from . import _abi
from .batcher import IntegrateBatcher


class Failure:
    """Placeholder for a work item whose function raised (parallel.py:19-41)."""

    __slots__ = ("index", "error")

    def __init__(self, index, error):
        self.index = index
        self.error = error

    def __repr__(self):
        return f"Failure(index={self.index}, error={self.error!r})"

    @property
    def message(self):
        return f"{type(self.error).__name__}: {self.error}"


class _Guarded:
    """fn wrapped so exceptions come back as Failure (picklable)."""

    __slots__ = ("fn",)

    def __init__(self, fn):
        self.fn = fn

    def __call__(self, item):
        try:
            return self.fn(item)
        except Exception as exc:
            return Failure(None, exc)


def _gpu_map(call, items, n_workers, precision):
    n = len(items)
    if n == 0:
        return []
    W = min(n, n_workers or n)
    out = [None] * n
    errors = [None] * n
    batcher = IntegrateBatcher(W, precision=precision)
    next_item = [0]
    lock = threading.Lock()

    def worker():
        batcher.enlist()
        try:
            while True:
                with lock:
                    i = next_item[0]
                    next_item[0] += 1
                if i >= n:
                    return
                try:
                    out[i] = call(items[i])
                except BaseException as exc:    # re-raised in the caller (on_error="raise")
                    errors[i] = exc
        finally:
            batcher.leave()

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(W)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for e in errors:
        if e is not None:
            raise e
    return out


def parallel_map(fn, items, *, backend="serial", n_workers=None, on_error="raise",
                 precision=_abi.PRECISION_FP64):
    """Apply `fn` to each of `items`, returning results in order.

    backend="gpu" batches the `integrate` calls of all items onto the GPU
    (`n_workers` bounds the batch width; default: all items at once);
    `precision` selects the FP64 or mixed form for that backend.
    on_error="skip" puts a `Failure` in the slot of an item that raised.
    """
    items = list(items)
    if on_error not in ("raise", "skip"):
        raise ValueError(f"on_error must be 'raise' or 'skip', got {on_error!r}")
    call = fn if on_error == "raise" else _Guarded(fn)

    if backend == "serial":
        out = [call(x) for x in items]
    elif backend == "processes":
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=n_workers) as ex:
            out = list(ex.map(call, items))
    elif backend == "gpu":
        out = _gpu_map(call, items, n_workers, precision)
    else:
        raise ValueError(f"unknown backend {backend!r} (use 'serial', 'processes' or 'gpu')")

    for i, r in enumerate(out):
        if isinstance(r, Failure):
            r.index = i
    return out
