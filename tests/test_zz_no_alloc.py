"""bdsim_integrate_batch_static (no device allocation: every buffer carved
from the library's static device array) against bdsim_integrate_batch: the
same positions, generator states, statuses, counters and step logs, bit for
bit, on consecutive calls of shrinking size -- the array keeps the previous
call's contents, which must not leak into the next -- including a chain split
over several workgroups (its barrier counters must start at zero each call),
and a batch too large for the array refused without a launch."""
import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from golden import cases as C

pytestmark = pytest.mark.gpu


def _items(specs):
    items, keep = [], []
    for k, (N, d) in enumerate(specs):
        p = H.Packed(d, N)
        keep.append(p)
        items.append((N, C.random_walk(N, 900 + k, 0.9), p.c,
                      H.csim(C.sim_dict(0.0, 0.011, 1e-3, 1e-3)), H.rng_of(k)))
    return items, keep


def _run(lib, fn, items, precision):
    n = len(items)
    nb = np.array([it[0] for it in items], dtype=np.int32)
    R = np.ascontiguousarray(np.concatenate([np.asarray(it[1], dtype=np.float64).ravel()
                                             for it in items]))
    phys = (A.CPhys * n)(*[it[2] for it in items])
    sims = (A.CSim * n)(*[it[3] for it in items])
    rngs = (A.CRng * n)(*[A.CRng(it[4].ix, it[4].iy, it[4].seed) for it in items])
    status = np.full(n, -7, dtype=np.int32)
    cnt = (A.CCounters * n)()
    stride = 16
    logs = (A.CStepLog * (n * stride))()
    rc = getattr(lib, fn)(n, A.iptr(nb), A.dptr(R), phys, sims, rngs, precision, A.iptr(status),
                          cnt, logs, stride)
    return rc, R, [(r.ix, r.iy, r.seed) for r in rngs], status, \
        [tuple(getattr(c, f) for f in A.COUNTER_FIELDS) for c in cnt], \
        [tuple(getattr(l, f) for f in A.LOG_FIELDS) for l in logs]


def _same(lib, items, precision):
    a = _run(lib, "bdsim_integrate_batch", items, precision)
    b = _run(lib, "bdsim_integrate_batch_static", items, precision)
    assert a[0] == 0 and b[0] == 0, lib.bdsim_last_error()
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2] and np.array_equal(a[3], b[3]) and a[4] == b[4] and a[5] == b[5]
    assert np.all(a[3] == 0)


def _mixed_specs():
    specs = []
    for k, N in enumerate([10, 50, 64, 65, 130, 300, 12, 257]):
        specs.append((N, C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3 if k % 3 else 0.0,
                                     ev=[3, 1, 2][k % 3], zstar=0.5, dstar=1.0,
                                     kappa=C.wi_kappa(0.5 * k))))
    for N in (20, 30, 40):      # Cholesky: factor in LDS (two sizes) and in the workspace
        specs.append((N, C.phys_dict(spring=2, sqrtb=50 ** 0.5, hstar=0.25, method=1, ev=1,
                                     zstar=1.0, dstar=0.5)))
    return specs


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED], ids=["fp64", "mixed"])
def test_static_equals_allocating_entry_point(gpu, precision):
    assert gpu.bdsim_static_scratch_bytes() == 256 << 20
    specs = _mixed_specs()
    items, keep = _items(specs)
    _same(gpu, items, precision)                     # first call: the full batch
    small, keep2 = _items(specs[5:8])                # second call: smaller
    _same(gpu, small, precision)
    split, keep3 = _items([(520, C.phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=1, zstar=0.3,
                                             dstar=0.5))])
    _same(gpu, split, precision)                     # one chain over 8 workgroups
    _same(gpu, split, precision)                     # and again: barriers cleared


def test_static_refuses_what_does_not_fit(gpu):
    N = 300
    items, keep = _items([(N, C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3))] * 60)
    rc, R, _, status, _, _ = _run(gpu, "bdsim_integrate_batch_static", items, A.PRECISION_FP64)
    assert rc == -3                                  # BDSIM_ENOMEM
    assert b"static scratch" in gpu.bdsim_last_error()
    assert np.all(status == -7)                      # nothing ran, nothing written
