"""Synthetic workloads of the step's benchmarks and large-N tests.

headline: 4096 chains of N=50 beads, FENE b=100 (sqrtb=10), RPY h*=0.3 with the
Chebyshev square root, SDK excluded volume (d*=1, cutoffs 0.7/1.5), simple shear
with Wi in {0.1, 1, 10, 100} round robin, dt=1e-3.  The shear rate is Wi / tau_1
with tau_1 = 1/sin^2(pi/2N), the longest Rouse relaxation time in code units
(python/bdsim/dynamics.py:583-606, validation/hi_nan_probe.py:103).

uneven: 4096 chains with N log-uniform in [10, 256], the headline physics.

cholesky: 8192 chains of N=30 with the dense Cholesky square root (a 90 x 90
factor per chain and step), Gaussian excluded volume, planar extension.

c1: the reference's CPU-sized case, 16 chains of N=10 in 10 segments of 100
steps with snapshot observables after each.

long: 64 chains of N=1000 with bending and Lennard-Jones excluded volume,
self-avoiding starts.
"""
import numpy as np

from bdsim_amd import _abi as A

N_BEADS = 50
N_TRAJ = 4096
WI = [0.1, 1.0, 10.0, 100.0]


def headline_phys():
    c = A.CPhys()
    c.spring_type = 2
    c.sqrtb = 10.0
    c.natural_length = 0.0
    c.ev_type = 3
    c.zstar = 0.0
    c.dstar = 1.0
    c.min_cutoff = 0.7
    c.max_cutoff = 1.5
    c.contour_dist_for_ev = 1
    c.bend_type = 0
    c.flow.constant = 1
    c.number_of_beads = N_BEADS
    c.hstar = 0.3
    c.hi_method = 0
    c.ncheb_multiplier = 1.0
    c.fd_err_max = 0.0025
    c.equilibration = 0
    return c


def start_configs(n_traj, N):
    R = np.empty((n_traj, N, 3))
    for i in range(n_traj):
        rs = np.random.default_rng(i)
        b = rs.normal(scale=1.0, size=(N - 1, 3))
        R[i, 0] = 0.0
        R[i, 1:] = np.cumsum(b, axis=0)
        R[i] -= R[i].mean(axis=0)
    return np.ascontiguousarray(R)


def shear_rate(wi, N=N_BEADS):
    return wi * np.sin(np.pi / (2.0 * N)) ** 2


def kappas(n_traj):
    K = np.zeros((n_traj, 9))
    for i in range(n_traj):
        K[i, 1] = shear_rate(WI[i % len(WI)])
    return np.ascontiguousarray(K)


def _walk(rs, N, bond=1.0):
    R = np.zeros((N, 3))
    R[1:] = np.cumsum(rs.normal(scale=bond, size=(N - 1, 3)), axis=0)
    return R - R.mean(axis=0)


def self_avoiding_walk(N, rs, min_sep=0.9, tries=200):
    """Unit bonds grown one bead at a time, each new bead at least min_sep from
    every earlier one (a fresh random direction per rejected try)."""
    R = np.zeros((N, 3))
    for i in range(1, N):
        for _ in range(tries):
            u = rs.normal(size=3)
            x = R[i - 1] + u / np.linalg.norm(u)
            if i < 2 or np.min(np.linalg.norm(R[:i - 1] - x, axis=1)) >= min_sep:
                break
        R[i] = x
    return R - R.mean(axis=0)


def uneven_workload(n_traj=4096, seed=1):
    """4096 chains, N log-uniform in [10, 256] (seed 1), FENE b=100, RPY h*=0.3
    Chebyshev, SDK EV, shear Wi round robin, dt=1e-3."""
    rs = np.random.default_rng(seed)
    N = np.clip(np.rint(np.exp(rs.uniform(np.log(10), np.log(256), n_traj))), 10, 256).astype(np.int32)
    R = [_walk(np.random.default_rng(1000 + i), int(N[i])) for i in range(n_traj)]
    phys = []
    for i in range(n_traj):
        c = headline_phys()
        c.number_of_beads = int(N[i])
        c.flow.kappa0[1] = shear_rate(WI[i % len(WI)], int(N[i]))
        phys.append(c)
    return N, R, phys, 1e-3


def long_workload(n_traj=64, N=1000, seed=2):
    """64 chains of N=1000, FENE b=100, bending 5, LJ sigma=1 epsilon=0.5 cut
    2.5 sigma, RPY h*=0.2 Chebyshev, no flow, dt=5e-4, self-avoiding unit-bond
    starts (seed 2)."""
    rs = np.random.default_rng(seed)
    R = [self_avoiding_walk(N, rs) for _ in range(n_traj)]
    c = headline_phys()
    c.number_of_beads = N
    c.ev_type, c.zstar, c.dstar, c.max_cutoff = 2, 0.5, 1.0, 2.5
    c.bend_type, c.bend_stiffness = 1, 5.0
    c.hstar = 0.2
    return np.full(n_traj, N, dtype=np.int32), R, [c] * n_traj, 5e-4


def cholesky_workload(n_traj=8192, N=30, seed=4):
    """8192 chains of N=30, FENE b=50, RPY h*=0.25 with the dense Cholesky
    square root, Gaussian EV z*=1 d*=0.5, planar extension rate 0.5, dt=1e-3.
    Starts: Gaussian bonds of per-component variance b/(b+5), the second moment
    of the equilibrium FENE connector, one generator per chain (seed*1e6 + i)."""
    b = 50.0
    sd = np.sqrt(b / (b + 5.0))
    R = [_walk(np.random.default_rng(seed * 1_000_000 + i), N, bond=sd) for i in range(n_traj)]
    c = headline_phys()
    c.number_of_beads = N
    c.sqrtb = np.sqrt(b)
    c.ev_type, c.zstar, c.dstar = 1, 1.0, 0.5
    c.hstar, c.hi_method = 0.25, 1
    c.flow.kappa0[0], c.flow.kappa0[4] = 0.5, -0.5
    return np.full(n_traj, N, dtype=np.int32), R, [c] * n_traj, 1e-3


def c1_workload(n_traj=16, N=10):
    """16 trajectories (seeds 0..15) of N=10, FENE-Fraenkel springs sqrtb=2
    natural length 1 started from the FENE-Fraenkel sampler (sigma=1, dQ=2),
    RPY h*=0.3 Chebyshev, no EV, simple shear rate 1, dt=0.01; 1000 steps as
    10 segments of 100 (write_every=100). Returns (R0 (n, N, 3), PhysParams,
    SimParams, segments, steps per segment)."""
    from .core import DelSMethod, Flow, PhysParams, SimParams, Spring
    from .initial import fene_fraenkel_chain
    p = PhysParams()
    p.number_of_beads = N
    p.spring.type, p.spring.sqrtb, p.spring.natural_length = Spring.FENEFraenkel, 2.0, 1.0
    p.hstar, p.hi_method = 0.3, DelSMethod.Chebyshev
    p.flow = Flow.constant([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    s = SimParams()
    s.time_start, s.time_end, s.dt, s.implicit_loop_tol = 0.0, 9.99, 0.01, 1e-3
    R0 = np.array([fene_fraenkel_chain(N, 1.0, 2.0, seed=i) for i in range(n_traj)])
    return R0, p, s, 10, 100
