"""Per-bucket register budgets: timed choice, table round trip, and results
that do not depend on the budget (bit for bit)."""
import numpy as np
import pytest

from bdsim_amd import _abi as A
from bdsim_amd import autotune, workloads as W

pytestmark = pytest.mark.gpu


def _run(lib, N, R, phys, dt, steps=6):
    n = len(N)
    Rf = np.ascontiguousarray(np.concatenate([r.ravel() for r in R]))
    ph = (A.CPhys * n)(*phys)
    rng = (A.CRng * n)(*[A.CRng(1, -1, i) for i in range(n)])
    st = np.zeros(n, dtype=np.int32)
    sims = (A.CSim * n)(*[A.CSim(0.0, (steps - 1) * dt, dt, 1e-3, 1)] * n)
    A.check(lib, lib.bdsim_integrate_batch(n, A.iptr(np.ascontiguousarray(N)), A.dptr(Rf), ph, sims,
                                           rng, A.PRECISION_MIXED, A.iptr(st), None, None, 0))
    assert np.all(st == 0)
    return Rf, [(r.ix, r.iy, r.seed) for r in rng]


def test_tune_table_and_identical_results(gpu, tmp_path):
    N, R, phys, dt = W.uneven_workload(n_traj=96)
    try:
        base, base_rng = _run(gpu, N, R, phys, dt)
        table = autotune.tune(N, R, phys, dt, steps=4, warmup=1, repeats=1, lib=gpu)
        assert set(table) == {autotune.block_of(n) for n in N} == {64, 128, 256}
        for b, w in table.items():
            assert w in autotune.CANDIDATES[b] and autotune.current(gpu)[b] == w
        p = tmp_path / "tune.json"
        autotune.save(table, p)
        assert autotune.load(p) == table
        # every budget gives the same numbers
        for b in (64, 128, 256):
            for w in autotune.CANDIDATES[b]:
                autotune.reset(gpu)
                autotune.apply({b: w}, gpu)
                out, rng = _run(gpu, N, R, phys, dt)
                assert np.array_equal(out, base) and rng == base_rng, (b, w)
    finally:
        autotune.reset(gpu)
    assert all(v == 0 for v in autotune.current(gpu).values())


def test_tune_rejects_bad_settings(gpu):
    assert gpu.bdsim_tune_set(96, 4) != 0
    assert gpu.bdsim_tune_set(128, 2) != 0
    assert gpu.bdsim_tune_set(64, 3) != 0
    assert gpu.bdsim_tune_get(96) == -1
