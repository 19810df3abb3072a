"""The synthetic benchmark workloads (host side, no GPU)."""
import numpy as np

from bdsim_amd import workloads as W


def test_headline_starts_and_shear():
    R = W.start_configs(8, W.N_BEADS)
    assert R.shape == (8, 50, 3)
    assert np.allclose(R.mean(axis=1), 0.0, atol=1e-12)
    K = W.kappas(8)
    for i in range(8):
        assert K[i, 1] == W.WI[i % 4] * np.sin(np.pi / 100.0) ** 2
        assert np.count_nonzero(K[i]) == 1


def test_uneven_workload():
    N, R, phys, dt = W.uneven_workload(n_traj=400)
    assert N.min() >= 10 and N.max() <= 256 and dt == 1e-3
    # log-uniform: the median sits near the geometric mean sqrt(10 * 256)
    assert 35 < np.median(N) < 70
    for i in (0, 1, 2, 3, 399):
        assert R[i].shape == (N[i], 3)
        assert phys[i].number_of_beads == N[i]
        assert phys[i].flow.kappa0[1] == W.shear_rate(W.WI[i % 4], int(N[i]))
    # the full-size workload's first chains are the same chains
    again = W.uneven_workload(n_traj=400)
    assert np.array_equal(again[0], N) and np.array_equal(again[1][7], R[7])


def test_long_workload_self_avoiding():
    N, R, phys, dt = W.long_workload(n_traj=1)
    assert N[0] == 1000 and R[0].shape == (1000, 3) and dt == 5e-4
    X = R[0]
    assert np.allclose(np.linalg.norm(np.diff(X, axis=0), axis=1), 1.0)
    d = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2)
    far = d[np.triu_indices(1000, 2)]
    assert far.min() >= 0.9 - 1e-12
    p = phys[0]
    assert (p.ev_type, p.zstar, p.dstar, p.max_cutoff) == (2, 0.5, 1.0, 2.5)
    assert (p.bend_type, p.bend_stiffness, p.hstar) == (1, 5.0, 0.2)


def test_c1_workload():
    R0, phys, sim, nseg, per = W.c1_workload()
    assert R0.shape == (16, 10, 3) and nseg * per == 1000 and sim.dt == 0.01
    # FENE-Fraenkel bonds in [sigma - dQ, sigma + dQ] = [0, 3]
    q = np.linalg.norm(np.diff(R0, axis=1), axis=2)
    assert q.max() <= 3.0 and q.min() >= 0.0
    assert (int(phys.spring.type), phys.spring.sqrtb, phys.spring.natural_length) == (6, 2.0, 1.0)
    assert phys.flow.constant_kappa()[0][1] == 1.0 and phys.hstar == 0.3
