"""A run directory filled from a device-resident ensemble (storage.record_run):
every snapshot's positions are the oracle's, integrating the same chains in
the same chunks (one integrate segment per chunk on the shared flow clock,
python/bdsim/ensemble.py:106-123), within 1e-10, and the forces are the
oracle's total force at those positions within 1e-12 of their scale."""
from ctypes import byref

import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from bdsim_amd import h5, storage
from bdsim_amd.core import EV, Flow, PhysParams, SimParams, Spring
from golden import cases as C

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not h5.available(), reason="the HDF5 library is not installed")]


def test_record_run_snapshots_match_oracle(gpu, orc, tmp_path):
    N, n, n_steps, every = 14, 3, 25, 10
    p = PhysParams()
    p.number_of_beads = N
    p.spring.type, p.spring.sqrtb = Spring.FENE, 10.0
    p.hstar = 0.25
    p.ev.type, p.ev.zstar, p.ev.dstar = EV.Gauss, 0.5, 0.5
    p.flow = Flow.constant([[0.0, 0.3, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    sim = SimParams()
    sim.time_start, sim.time_end, sim.dt, sim.implicit_loop_tol = 0.2, 1.0, 1e-3, 1e-3
    R0 = np.array([C.random_walk(N, 40 + i, 0.9) for i in range(n)])
    seeds = [5, 6, 7]
    run = storage.record_run(str(tmp_path / "run"), R0, p, sim, seeds, n_steps, write_every=every,
                             write_forces=True)
    assert len(run) == n and run.manifest["n_steps"] == n_steps and run.manifest["write_forces"]
    pk = A.PackedPhys(p, N)
    for i, traj in enumerate(run):
        at = traj.attrs
        assert at["index"] == i and at["seed"] == seeds[i] and at["has_forces"] is True
        assert at["phys_n_beads"] == N and at["write_every"] == every
        assert np.array_equal(traj.step, [0, 10, 20, 25])
        assert np.allclose(traj.time, 0.2 + traj.step * 1e-3, rtol=0, atol=1e-15)
        R = R0[i].copy()
        st = A.CRng(1, -1, seeds[i])
        prev = 0
        for k, (step, t, Rk, Fk) in enumerate(traj):
            if step > prev:
                t0 = 0.2 + prev * 1e-3
                cs = A.CSim(t0, t0 + (step - prev - 1) * 1e-3, 1e-3, 1e-3, 1)
                R, rc, _, _ = H.oracle_integrate(orc, N, R, pk.c, cs, st)
                assert rc == 0
                prev = step
            assert np.max(np.abs(Rk - R)) <= 1e-10 * np.max(np.abs(R)), (i, step)
            F = np.empty_like(R)
            orc.bo_total_force(N, A.dptr(np.ascontiguousarray(R)), byref(pk.c), A.dptr(F))
            assert np.max(np.abs(Fk - F)) <= 1e-9 * np.max(np.abs(F)), (i, step)
