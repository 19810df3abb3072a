"""Per-bucket launch shapes of the integrate kernels, picked by timing.

The library launches the chains of a batch in buckets by block size (64, 128,
256, 512 threads for chains of up to that many beads). Each bucket's kernel
exists in builds for different register budgets (waves per SIMD: 1, 2 or 4 for
64-thread chains; 1 or 4 for the others); more registers per wave means fewer
spills, fewer means more chains resident at once. Chains of more than 512
beads, when they alone leave compute units idle, run split over several
workgroups each (waves per chain: 16 per workgroup times G workgroups); more
parts means less work per part but more exchanges to wait for. In the mixed
form the chains of the 128..512 buckets may also run on wider workgroups
(block x m threads, m = 1, 2, 4, 8 up to 1024: more waves per chain, the tiled
pair walk spread over them), which shortens a latency-bound step but holds
more waves per chain. Which wins depends on the mix, so `tune` times the
candidates on a representative batch -- one bucket at a time, the others on
the library's own choice, all buckets running side by side as they do in use
-- and keeps the fastest, and for batches of long chains the largest G. The
table ({block: waves per SIMD}, under key 1024 the workgroups per long chain,
under "w<block>" the width multiplier) applies with `apply` (bdsim_tune_set,
bdsim_tune_split, bdsim_tune_width) and round-trips through JSON (`save` /
`load`). Register budgets change time only (the same operations, another
register allocation); G and the width also change the order in which partner
sums are added (the width: exactly, through fixed-point sums). Without a
budget of its own the 64-thread bucket's chains of up to 51 beads (mixed
form) run with the step's pair weights in an LDS table (k_integrate_ptab);
an explicit budget selects the kernel that forms them in each mat-vec.
"""
import json
import time

import numpy as np

from . import _abi

CANDIDATES = {64: (1, 2, 4), 128: (1, 4), 256: (1, 4), 512: (1, 4)}
SPLIT_KEY = 1024                      # table key of the workgroups per long chain
SPLIT_CANDIDATES = (1, 2, 4, 8, 16)
WIDTH_CANDIDATES = {128: (1, 2, 4, 8), 256: (1, 2, 4), 512: (1, 2)}


def width_key(block):
    """Table key of the workgroup width multiplier of a block size."""
    return f"w{int(block)}"


def block_of(N):
    """The library's block size for a chain of N beads: 64, 128, 256, 512 or 1024."""
    b = 64
    while b < int(N):
        b *= 2
    return b


def apply(table, lib=None):
    """Set the library's budgets from {block: waves}, the workgroups per long
    chain from table[1024] and the width multipliers from table["w<block>"]
    (0 = automatic)."""
    lib = lib or _abi.load_gpu()
    for key, v in table.items():
        if isinstance(key, str):
            _abi.check(lib, lib.bdsim_tune_width(int(key[1:]), int(v)))
        elif int(key) == SPLIT_KEY:
            _abi.check(lib, lib.bdsim_tune_split(int(v)))
        else:
            _abi.check(lib, lib.bdsim_tune_set(int(key), int(v)))


def reset(lib=None):
    apply({**{b: 0 for b in CANDIDATES}, SPLIT_KEY: 0, **{width_key(b): 0 for b in WIDTH_CANDIDATES}},
          lib)


def current(lib=None):
    lib = lib or _abi.load_gpu()
    out = {b: int(lib.bdsim_tune_get(b)) for b in CANDIDATES}
    out[SPLIT_KEY] = int(lib.bdsim_tune_split_get())
    for b in WIDTH_CANDIDATES:
        out[width_key(b)] = int(lib.bdsim_tune_width_get(b))
    return out


def save(table, path):
    with open(path, "w") as f:
        json.dump({str(k): int(v) for k, v in table.items()}, f, indent=1, sort_keys=True)


def load(path):
    with open(path) as f:
        return {(int(k) if k.isdigit() else k): int(v) for k, v in json.load(f).items()}


def _timed(lib, N, R, phys, dt, precision, steps, warmup):
    n = len(N)
    Rf = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.float64).ravel() for r in R]))
    ph = (_abi.CPhys * n)(*phys)
    rng = (_abi.CRng * n)(*[_abi.CRng(1, -1, i) for i in range(n)])
    st = np.zeros(n, dtype=np.int32)
    Nc = np.ascontiguousarray(np.asarray(N, dtype=np.int32))

    def segment(t0, k):
        sims = (_abi.CSim * n)(*[_abi.CSim(t0, t0 + (k - 1) * dt, dt, 1e-3, 1)] * n)
        t = time.perf_counter()
        _abi.check(lib, lib.bdsim_integrate_batch(n, _abi.iptr(Nc), _abi.dptr(Rf), ph, sims, rng,
                                                  precision, _abi.iptr(st), None, None, 0))
        return time.perf_counter() - t

    if warmup > 0:
        segment(0.0, warmup)
    return segment(warmup * dt, steps)


def tune(N, R, phys, dt, precision=_abi.PRECISION_MIXED, steps=20, warmup=3, repeats=2,
         lib=None, verbose=False):
    """Fastest budget per bucket present in the batch (N: chain lengths, R:
    their positions, phys: CPhys per chain, dt), for a batch of chains of
    more than 512 beads only the fastest number of workgroups per chain, and
    (mixed form) the fastest width per bucket of 128..512 threads. Returns
    {block: waves} (with 1024: workgroups per chain; "w<block>": width
    multiplier); the library is left on that table."""
    lib = lib or _abi.load_gpu()
    blocks = sorted({block_of(n) for n in N if block_of(n) in CANDIDATES})
    table = {}
    reset(lib)
    # one untimed pass: the first integrate call of a form in a process is slow
    _timed(lib, N, R, phys, dt, precision, steps, warmup)
    if len(N) > 0 and all(block_of(n) == SPLIT_KEY for n in N):
        best = None
        for G in SPLIT_CANDIDATES:
            apply({SPLIT_KEY: G}, lib)
            t = min(_timed(lib, N, R, phys, dt, precision, steps, warmup) for _ in range(repeats))
            if verbose:
                print(f"autotune: {G} workgroups per chain: {t * 1e3 / steps:.3f} ms/step")
            if best is None or t < best[1]:
                best = (G, t)
        table[SPLIT_KEY] = best[0]
        apply({SPLIT_KEY: best[0]}, lib)
    for b in blocks:
        best = None
        for w in CANDIDATES[b]:
            apply({b: w}, lib)
            t = min(_timed(lib, N, R, phys, dt, precision, steps, warmup) for _ in range(repeats))
            if verbose:
                print(f"autotune: block {b} waves/SIMD {w}: {t * 1e3 / steps:.3f} ms/step")
            if best is None or t < best[1]:
                best = (w, t)
        table[b] = best[0]
        apply({b: best[0]}, lib)
    if precision != _abi.PRECISION_FP64:
        for b in blocks:
            if b not in WIDTH_CANDIDATES:
                continue
            best = None
            for m in WIDTH_CANDIDATES[b]:
                apply({width_key(b): m}, lib)
                t = min(_timed(lib, N, R, phys, dt, precision, steps, warmup) for _ in range(repeats))
                if verbose:
                    print(f"autotune: block {b} width x{m}: {t * 1e3 / steps:.3f} ms/step")
                if best is None or t < best[1]:
                    best = (m, t)
            table[width_key(b)] = best[0]
            apply({width_key(b): best[0]}, lib)
    return table
