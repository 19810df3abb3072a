"""Host-side driver logic on the CPU: the integrate batcher and gpu_map (with
stand-in runners, no GPU), parameter pickling, initial samplers."""
import pickle
import threading

import numpy as np
import pytest

import bdsim_amd as bd
from bdsim_amd import batcher as B


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


def test_gpu_map_routes_through_batcher():
    log = []
    items = list(range(6))

    def fn(i):
        R = np.full((5, 3), float(i))
        rng = bd.Rng(i)
        for _ in range(3):
            R = bd.integrate(R, bd.PhysParams(), bd.SimParams(), rng)
        return float(R.sum())

    out = bd.gpu_map(fn, items, runner=fake_runner(log))
    assert out == [float((i + 3) * 15) for i in items]
    assert log == [6, 6, 6]


def test_gpu_map_bounded_workers_and_errors():
    """n_workers bounds the threads (and so the launch width); an exception
    propagates after all workers stop, the one of the lowest item index."""
    log = []

    def fn(n):
        if n == 3:
            raise RuntimeError(f"planted {n}")
        return bd.integrate(np.zeros((n, 3)), bd.PhysParams(), bd.SimParams(), bd.Rng(0))

    out = bd.gpu_map(fn, [4, 5, 6, 7, 8], n_workers=2, runner=fake_runner(log))
    assert [o.shape[0] for o in out] == [4, 5, 6, 7, 8] and max(log) <= 2
    with pytest.raises(RuntimeError, match="planted 3"):
        bd.gpu_map(fn, [4, 3, 5, 3], runner=fake_runner([]))
    assert bd.gpu_map(fn, [], runner=fake_runner([])) == []


def test_batcher_flushes_when_a_worker_is_slow():
    """A request that waited flush_after_s with no new arrival goes out
    without the worker that is still busy in Python."""
    import time
    log = []

    def fn(i):
        if i == 0:
            time.sleep(0.5)                   # a slow per-trajectory function
        return bd.integrate(np.zeros((4, 3)), bd.PhysParams(), bd.SimParams(), bd.Rng(i))

    t = time.monotonic()
    out = bd.gpu_map(fn, [0, 1, 2, 3], flush_after_s=0.02, runner=fake_runner(log))
    assert len(out) == 4 and log[0] == 3 and sum(log) == 4
    assert time.monotonic() - t < 5.0


def _oracle_runner():
    """The batcher's launch done by the CPU oracle, one request at a time (a
    stand-in for the GPU that computes the real step)."""
    import helpers as H
    from bdsim_amd import _abi as A
    from bdsim_amd.core import status_error
    orc = H.oracle()

    def run(batch):
        for r in batch:
            N = r.R.shape[0]
            pk = A.PackedPhys(r.phys, N)
            st = A.CRng(*r.rng._state())
            R, rc, cnt, _ = H.oracle_integrate(orc, N, r.R, pk.c, A.pack_sim(r.sim), st)
            r.rng._set_state(st.ix, st.iy, st.seed)
            r.out = R
            r.error = status_error(int(rc), cnt)
    return run


def test_gpu_map_opaque_functions_equal_serial_calls():
    """Per-trajectory functions that make several integrate calls of different
    segment lengths, with Python decisions between them, give through gpu_map
    exactly what the same functions give called one after another."""
    import helpers as H
    from golden import cases as C
    orc_run = _oracle_runner()
    phys = bd.PhysParams()
    phys.spring.type, phys.spring.sqrtb = bd.Spring.FENE, 10.0
    phys.hstar = 0.25
    phys.flow = bd.flows.shear(0.5)

    def traj(i):
        R = C.random_walk(10, 40 + i)
        rng = bd.Rng(i)
        out = []
        t = 0.0
        for k in range(3 + i % 3):
            s = bd.SimParams()
            s.time_start, s.dt = t, 1e-3
            s.time_end = t + (5 + 3 * k + i) * 1e-3
            R = bd.integrate(R, phys, s, rng)
            t = s.time_end + s.dt
            out.append(float(np.sum((R[-1] - R[0]) ** 2)))
            if out[-1] > 1e9:               # a data-dependent branch
                break
        return R, rng.__getstate__(), out

    items = list(range(7))
    from bdsim_amd.batcher import IntegrateBatcher
    serial = []
    for i in items:                           # one at a time: a batcher of one worker
        b = IntegrateBatcher(1, runner=orc_run)
        b.enlist()
        try:
            serial.append(traj(i))
        finally:
            b.leave()
    batched = bd.gpu_map(traj, items, n_workers=4, runner=orc_run)
    for a, b in zip(serial, batched):
        assert np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2]
    del H


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
