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
        blocks = {k for k in table if isinstance(k, int)}
        assert blocks == {autotune.block_of(n) for n in N} == {64, 128, 256}
        assert {k for k in table if isinstance(k, str)} == {autotune.width_key(128), autotune.width_key(256)}
        for b, w in table.items():
            cand = autotune.CANDIDATES[b] if isinstance(b, int) else autotune.WIDTH_CANDIDATES[int(b[1:])]
            assert w in cand and autotune.current(gpu)[b] == w
        p = tmp_path / "tune.json"
        autotune.save(table, p)
        assert autotune.load(p) == table
        # every budget and every width gives the same numbers
        for b in (64, 128, 256):
            for w in autotune.CANDIDATES[b]:
                autotune.reset(gpu)
                autotune.apply({b: w}, gpu)
                out, rng = _run(gpu, N, R, phys, dt)
                assert np.array_equal(out, base) and rng == base_rng, (b, w)
        for b in (128, 256):
            for m in autotune.WIDTH_CANDIDATES[b]:
                autotune.reset(gpu)
                autotune.apply({autotune.width_key(b): m}, gpu)
                out, rng = _run(gpu, N, R, phys, dt)
                assert np.array_equal(out, base) and rng == base_rng, (b, m)
    finally:
        autotune.reset(gpu)
    assert all(v == 0 for v in autotune.current(gpu).values())


def test_tune_rejects_bad_settings(gpu):
    assert gpu.bdsim_tune_set(96, 4) != 0
    assert gpu.bdsim_tune_set(128, 2) != 0
    assert gpu.bdsim_tune_set(64, 3) != 0
    assert gpu.bdsim_tune_get(96) == -1
    assert gpu.bdsim_tune_width(64, 2) != 0
    assert gpu.bdsim_tune_width(256, 3) != 0
    assert gpu.bdsim_tune_width_get(64) == -1


def test_tune_workgroups_per_long_chain(gpu, orc):
    """Batches of chains of N > 512 only: the tuner times 1, 2, 4, 8 and 16
    workgroups per chain and keeps one; every choice integrates to the
    oracle's positions within 1e-10 (the parts only reorder the partner
    sums)."""
    import helpers as H
    from golden import cases as C
    N_, n = 600, 3
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=1, zstar=0.3, dstar=0.5)
    p = H.Packed(d, N_)
    R = [W.self_avoiding_walk(N_, np.random.default_rng(90 + i)) for i in range(n)]
    N = np.full(n, N_, dtype=np.int32)
    try:
        table = autotune.tune(N, R, [p.c] * n, 1e-3, precision=A.PRECISION_FP64, steps=2, warmup=1,
                              repeats=1, lib=gpu)
        assert table[autotune.SPLIT_KEY] in autotune.SPLIT_CANDIDATES
        assert autotune.current(gpu)[autotune.SPLIT_KEY] == table[autotune.SPLIT_KEY]
        sim = A.CSim(0.0, 0.004, 1e-3, 1e-3, 1)
        want = [H.oracle_integrate(orc, N_, R[i], p.c, sim, H.rng_of(i))[0] for i in range(n)]
        for G in autotune.SPLIT_CANDIDATES:
            autotune.apply({autotune.SPLIT_KEY: G}, gpu)
            out, _, status, _, _ = H.gpu_batch(gpu, [(N_, R[i], p.c, sim, H.rng_of(i)) for i in range(n)])
            assert np.all(status == 0)
            for i in range(n):
                err = np.max(np.abs(out[i] - want[i])) / np.max(np.abs(want[i]))
                assert err <= 1e-10, (G, i, err)
    finally:
        autotune.reset(gpu)
    assert gpu.bdsim_tune_split_get() == 0
    assert gpu.bdsim_tune_split(17) != 0 and gpu.bdsim_tune_split(-1) != 0


def test_pair_weight_table_matches_the_walk(gpu):
    """Mixed-form 64-thread chains run on k_integrate_ptab (the step's RPY
    weights in an LDS table, formed once per step): up to 51 beads always,
    up to 64 when the launch's chains are resident at once anyway (as these
    60 are); an explicit register budget for the 64-thread bucket selects
    k_integrate, whose walk forms the weights itself. The table holds the
    same FP32 operations' results, so the two agree bit for bit, positions
    and generator states, at every N from 2 to 64 (odd and even: the half
    step k = N/2), with and without HI."""
    rs = np.random.default_rng(11)
    Ns = np.array([2, 3, 4, 5, 10, 17, 30, 31, 32, 49, 50, 51, 52, 57, 64] * 4, dtype=np.int32)
    R, phys = [], []
    for i, n in enumerate(Ns):
        R.append(W._walk(rs, int(n)))
        c = W.headline_phys()
        c.number_of_beads = int(n)
        c.hstar = 0.0 if i % 7 == 3 else 0.3
        c.flow.kappa0[1] = W.shear_rate(W.WI[i % len(W.WI)], int(n))
        phys.append(c)
    try:
        autotune.reset(gpu)
        base, base_rng = _run(gpu, Ns, R, phys, 1e-3, steps=12)
        for w in autotune.CANDIDATES[64]:
            autotune.apply({64: w}, gpu)
            out, rng = _run(gpu, Ns, R, phys, 1e-3, steps=12)
            assert np.array_equal(out, base) and rng == base_rng, w
    finally:
        autotune.reset(gpu)
