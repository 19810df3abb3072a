"""Host-side driver logic on the CPU: the integrate batcher (with a stand-in
runner, no GPU), parallel_map backends, parameter pickling, initial samplers."""
import pickle
import threading

import numpy as np
import pytest

import bdsim_amd as bd
from bdsim_amd import batcher as B
from bdsim_amd import parallel as par


def fake_runner(log):
    """Stand-in for the GPU launch: R -> R + 1, ix -> ix + 1; 3-bead chains fail."""
    def run(batch):
        log.append(len(batch))
        for r in batch:
            if r.R.shape[0] == 3:
                r.error = RuntimeError("cholesky_lower: matrix is not positive definite (pivot 2)")
                continue
            ix, iy, seed = r.rng._state()
            r.rng._set_state(ix + 1, iy, seed)
            r.out = r.R + 1.0
    return run


def test_batcher_groups_concurrent_calls():
    log = []
    W, K = 8, 5
    bt = B.IntegrateBatcher(W, runner=fake_runner(log))
    results = [None] * W

    def work(i):
        bt.enlist()
        R = np.full((4, 3), float(i))
        rng = bd.Rng(i)
        for _ in range(K):
            R = bd.integrate(R, bd.PhysParams(), bd.SimParams(), rng)
        results[i] = (R, rng._state())
        bt.leave()

    ts = [threading.Thread(target=work, args=(i,)) for i in range(W)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    assert log == [W] * K                  # K launches, each with every chain
    for i, (R, st) in enumerate(results):
        assert np.all(R == i + K) and st == (1 + K, -1, i)


def test_batcher_early_finishers_and_errors():
    log = []
    bt = B.IntegrateBatcher(3, runner=fake_runner(log))
    out = {}

    def work(i, n_calls, nb):
        bt.enlist()
        try:
            R = np.zeros((nb, 3))
            for _ in range(n_calls):
                R = bd.integrate(R, bd.PhysParams(), bd.SimParams(), bd.Rng(0))
            out[i] = R
        except RuntimeError as e:
            out[i] = e
        finally:
            bt.leave()

    ts = [threading.Thread(target=work, args=a) for a in [(0, 1, 4), (1, 3, 4), (2, 2, 3)]]
    [t.start() for t in ts]
    [t.join() for t in ts]
    assert log == [3, 1, 1]
    assert np.all(out[0] == 1) and np.all(out[1] == 3)
    assert isinstance(out[2], RuntimeError) and "pivot 2" in str(out[2])


def _per_traj(R0, phys, sim, rng):
    R = R0
    for _ in range(3):
        R = bd.integrate(R, phys, sim, rng)
    return float(R.sum())


def test_parallel_map_gpu_backend_routes_through_batcher(monkeypatch):
    log = []
    monkeypatch.setattr(B.IntegrateBatcher, "_run_on_gpu", lambda self, b: fake_runner(log)(b))
    items = list(range(6))

    def fn(i):
        return _per_traj(np.full((5, 3), float(i)), bd.PhysParams(), bd.SimParams(), bd.Rng(i))

    out = par.parallel_map(fn, items, backend="gpu")
    assert out == [float((i + 3) * 15) for i in items]
    assert log == [6, 6, 6]
    with pytest.raises(ValueError):
        par.parallel_map(fn, items, backend="cluster")


def test_parallel_map_skip_collects_failures(monkeypatch):
    monkeypatch.setattr(B.IntegrateBatcher, "_run_on_gpu", lambda self, b: fake_runner([])(b))

    def fn(n):
        return bd.integrate(np.zeros((n, 3)), bd.PhysParams(), bd.SimParams(), bd.Rng(0))

    out = par.parallel_map(fn, [4, 3, 5], backend="gpu", on_error="skip")
    assert isinstance(out[1], par.Failure) and out[1].index == 1
    ok, bad = bd.ensemble._drop_failures(out, "trajectories", max_fraction=0.5)
    assert len(ok) == 2 and len(bad) == 1
    with pytest.raises(RuntimeError, match="above"):
        bd.ensemble._drop_failures(out, "trajectories")


def test_parameter_pickle_tuples_match_reference_layout():
    # bindings/module.cpp:189-306
    p = bd.PhysParams()
    p.spring.type = bd.Spring.FENEFraenkel
    p.spring.sqrtb, p.spring.natural_length = 3.0, 1.5
    p.ev.type = bd.EV.SDK
    p.ev.phi = [0.1, 0.2, 0.3, 0.4]
    p.bend = bd.BendingParams(bd.Bending.OneMinusCosTheta, 2.0)
    p.flow = bd.flows.shear(2.0)
    p.external.add_time_varying(-1, [0.0, 1.0], [(0, 0, 0), (1, 2, 3)])
    p.external.add_stretch(0, -1, (0.0, 0.0, 5.0))
    p.hstar, p.hi_method = 0.25, bd.DelSMethod.Cholesky
    st = p.__getstate__()
    assert st[4:10] == (10, 0.25, 1, 1.0, 0.0025, False)
    assert st[0].__getstate__() == (6, 3.0, 1.5)
    assert st[1].__getstate__() == (3, 0.0, 0.0, 0.7, 1.5, 1, [0.1, 0.2, 0.3, 0.4])
    fl = st[3].__getstate__()
    assert fl[0] is True and fl[1][1] == 2.0 and fl[2] == [] and fl[3] == []
    ext = st[10].__getstate__()
    assert ext[0] == [-1, -1, 0] and ext[1] == [0, 1, 1] and ext[4] == [2, 0, 0]
    assert ext[3] == [0.0, 1.0] and ext[5] == [0.0, 0.0, 0.0, 1.0, 2.0, 3.0]
    assert ext[2][3:6] == [0.0, 0.0, 5.0] and ext[2][8] == -5.0
    q = pickle.loads(pickle.dumps(p))
    assert q.__getstate__()[4:10] == st[4:10]
    assert q.external.__getstate__() == ext
    s = bd.SimParams()
    assert s.__getstate__() == (0.0, 1.0, 0.01, 1e-3, True)
    r = bd.Rng(7)
    assert r.__getstate__() == (1, -1, 7)
    assert pickle.loads(pickle.dumps(r)).__getstate__() == (1, -1, 7)


def test_fene_fraenkel_sampler_distribution():
    """Support [max(0, s-dQ), s+dQ] and a KS test against
    psi(Q) ~ Q^2 (1-(Q-s)^2/dQ^2)^(dQ^2/2) (python/bdsim/initial.py:29-41)."""
    from scipy import integrate, stats
    s, dQ = 1.0, 2.0
    R = np.array([bd.fene_fraenkel_chain(201, s, dQ, seed=k) for k in range(20)])
    Q = np.linalg.norm(np.diff(R, axis=1), axis=2).ravel()
    assert Q.min() >= max(0.0, s - dQ) and Q.max() <= s + dQ
    psi = lambda q: q * q * (1 - (q - s) ** 2 / dQ ** 2) ** (dQ ** 2 / 2)  # noqa: E731
    Z = integrate.quad(psi, 0.0, s + dQ)[0]
    grid = np.linspace(0.0, s + dQ, 2001)
    cdf_grid = np.array([integrate.quad(psi, 0.0, x)[0] / Z for x in grid])
    assert stats.kstest(Q, lambda x: np.interp(x, grid, cdf_grid)).pvalue > 1e-3
    g = bd.gaussian_chain(50, seed=3)
    assert np.allclose(g.mean(axis=0), 0.0, atol=1e-12)
