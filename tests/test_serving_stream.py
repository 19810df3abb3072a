"""The serving workload's job stream (host side, no GPU)."""
import numpy as np

from bdsim_amd import serving as S


def test_stream_distributions():
    jobs = S.c5_stream(400, seed=3)
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0)
    # Poisson at 50/s: mean gap 20 ms
    assert abs(np.mean(np.diff(arr)) - 0.02) < 0.004
    N = np.array([j.N for j in jobs])
    assert N.min() >= 10 and N.max() <= 500
    # log-uniform: the median sits near the geometric mean sqrt(10 * 500)
    assert 45 < np.median(N) < 95
    nt = np.array([j.n_traj for j in jobs])
    assert nt.min() >= 1 and nt.max() <= 64
    steps = np.array([j.meta["n_steps"] for j in jobs])
    assert steps.min() >= 1000 and steps.max() <= 100000
    hs = np.array([j.phys.hstar for j in jobs])
    assert hs.min() >= 0.0 and hs.max() <= 0.3
    for j in jobs[:20]:
        assert j.R0.shape == (j.n_traj, j.N, 3)
        assert np.allclose(j.R0.mean(axis=1), 0.0, atol=1e-12)
        assert j.bead_steps() == j.n_traj * j.N * j.meta["n_steps"]
        assert len(j.rngs) == j.n_traj
        k = j.phys.flow.constant_kappa()
        assert k[0][1] > 0.0 and abs(k[0][1] - j.meta["wi"] * np.sin(np.pi / (2 * j.N)) ** 2) < 1e-15
    # the same seed gives the same stream
    again = S.c5_stream(5, seed=3)
    for a, b in zip(again, jobs[:5]):
        assert a.arrival == b.arrival and np.array_equal(a.R0, b.R0)


def test_step_scale_keeps_at_least_one_step():
    jobs = S.c5_stream(50, step_scale=1e-5)
    assert all(j.meta["n_steps"] == 1 for j in jobs)
    assert all(j.bead_steps() == j.n_traj * j.N for j in jobs)
