"""The drivers on the GPU: `backend="gpu"` batches every trajectory's integrate
calls into shared launches and gives results identical to one-chain-at-a-time
execution ("serial"), bit for bit; failures are dropped as Failures."""
import os

import numpy as np
import pytest

import bdsim_amd as bd

pytestmark = pytest.mark.gpu


def _phys(N=12):
    p = bd.PhysParams()
    p.number_of_beads = N
    p.spring.type = bd.Spring.FENE
    p.spring.sqrtb = 10.0
    p.hstar = 0.25
    p.ev.type = bd.EV.Gauss
    p.ev.zstar, p.ev.dstar = 0.5, 0.5
    p.flow = bd.flows.shear(0.2)
    return p


def _sim(t1=0.099, dt=1e-3):
    s = bd.SimParams()
    s.time_start, s.time_end, s.dt = 0.0, t1, dt
    return s


def test_run_ensemble_gpu_equals_serial(gpu):
    phys, sim = _phys(), _sim()
    a = bd.run_ensemble(phys, sim, 16, bd.final_state, seed=3, backend="serial")
    b = bd.run_ensemble(phys, sim, 16, bd.final_state, seed=3, backend="gpu")
    assert len(a) == len(b) == 16
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    times = [0.02, 0.05, 0.1]
    c = bd.run_ensemble(phys, sim, 6, bd.sampled_states, args=(times,), backend="serial")
    d = bd.run_ensemble(phys, sim, 6, bd.sampled_states, args=(times,), backend="gpu",
                        n_workers=4)
    for x, y in zip(c, d):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)


def _fails_for_index(R0, phys, sim, rng):
    if rng.__getstate__()[2] == 2:     # seed 2 fails
        raise RuntimeError("planted failure")
    return bd.integrate(R0, phys, sim, rng)


def test_run_ensemble_gpu_drops_failures(gpu):
    phys, sim = _phys(), _sim(0.019)
    out = bd.run_ensemble(phys, sim, 12, _fails_for_index, backend="gpu", max_failed_fraction=0.5)
    assert len(out) == 11


def test_simulate_gpu_equals_serial(gpu, tmp_path):
    phys, sim = _phys(10), _sim(0.05)          # n_steps = round(0.05 / 1e-3) = 50
    out = bd.Output(directory=str(tmp_path / "ser"), write_every=10, write_forces=True)
    r1 = bd.simulate(phys, sim, n_traj=5, seed=7, initial=bd.Initial("fene_fraenkel",
                                                                     {"sigma": 1.0, "dQ": 2.0}),
                     output=out, backend="serial")
    out2 = bd.Output(directory=str(tmp_path / "gpu"), write_every=10, write_forces=True)
    r2 = bd.simulate(phys, sim, n_traj=5, seed=7, initial=bd.Initial("fene_fraenkel",
                                                                     {"sigma": 1.0, "dQ": 2.0}),
                     output=out2, backend="gpu")
    assert len(r1) == len(r2) == 5
    for t1, t2 in zip(r1, r2):
        assert np.array_equal(t1.step, [0, 10, 20, 30, 40, 50])
        assert np.array_equal(t1.positions, t2.positions)
        assert np.array_equal(t1.forces, t2.forces)
    assert os.path.exists(os.path.join(str(tmp_path / "gpu"), "manifest.json"))
