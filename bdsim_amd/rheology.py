"""Shear stress of a whole ensemble kept on the device.

The reference measures viscosities by running a per-trajectory function over
an ensemble (python/bdsim/rheology.py:36-160): each trajectory integrates to
the sample times and evaluates tau_xy = kramers_stress(R, total_force(R))[0, 1]
on the host; with variance reduction each trajectory carries a zero-flow
control chain driven by the same random stream (rheology.py:53-112).

`resident_stress_samples` does the same measurement with the ensemble held as
one resident batch in HBM: every sample time is one integrate segment of every
chain (the segments [t_k, t_{k+1}] that the reference's sampling walks), and
tau_xy is evaluated on the device, so one number per chain and sample crosses
to the host. With `control=True` every sheared chain gets a sibling in the
same batch with the same start and the same generator state and no flow: the
step draws exactly 3N deviates whatever the configuration, so the pair sees the
same Brownian kicks. Chain i starts from `start(N, seed + i)` with
Rng(seed + i), the seeding of the reference's ensembles.
"""
import numpy as np

from . import _abi
from .core import Rng
from .initial import gaussian_chain


def _gaussian_start(N, seed):
    return gaussian_chain(N, seed=seed)


def resident_stress_samples(phys, sim, n_traj, sample_times, *, seed=0, start=None,
                            control=False, precision=_abi.PRECISION_FP64):
    """tau_xy of n_traj chains at each sample time as an (n_traj, n_samples)
    array; with control=True a second array for the zero-flow siblings.
    `start(N, seed)` returns a chain's (N, 3) start (default: a Gaussian
    walk). The velocity gradient must be constant. Raises RuntimeError if a
    chain fails."""
    from .resident import ResidentEnsemble
    if not phys.flow.is_constant():
        raise ValueError("the resident path takes a constant velocity gradient")
    start = start or _gaussian_start
    N = int(phys.number_of_beads)
    R0 = np.array([start(N, seed + i) for i in range(n_traj)], dtype=np.float64)
    K = np.asarray(phys.flow.constant_kappa(), dtype=np.float64).reshape(1, 3, 3)
    kap = np.repeat(K, n_traj, axis=0)
    rngs = [Rng(seed + i) for i in range(n_traj)]
    if control:
        R0 = np.concatenate([R0, R0])
        kap = np.concatenate([kap, np.zeros_like(kap)])
        rngs = rngs + [Rng(seed + i) for i in range(n_traj)]
    tau = np.empty((len(R0), len(sample_times)))
    with ResidentEnsemble(R0, phys, sim, rngs, precision, kappas=kap) as ens:
        t = sim.time_start
        for k, ts in enumerate(sample_times):
            ens.run_segment(t, ts)
            tau[:, k] = ens.observe(force="total")["stress"][:, 0, 1]
            t = ts
        _, _, status, _ = ens.fetch()
    bad = np.flatnonzero(status != _abi.STATUS_OK)
    if bad.size:
        raise RuntimeError(f"{bad.size} resident chains failed (first: chain {int(bad[0])}, "
                           f"status {int(status[bad[0]])})")
    if control:
        return tau[:n_traj], tau[n_traj:]
    return tau


def resident_viscosity_samples(phys, sim, rate, n_traj, sample_times, *,
                               variance_reduction=False, **kwargs):
    """Per-sample viscosity contributions -tau_xy / rate, (n_traj, n_samples);
    with variance_reduction the control sibling's tau_xy is subtracted first.
    `phys.flow` must already be the shear at `rate`. Keyword arguments go to
    `resident_stress_samples` (seed, start, precision)."""
    times = list(sample_times)
    if variance_reduction:
        tf, te = resident_stress_samples(phys, sim, n_traj, times, control=True, **kwargs)
        return -(tf - te) / rate
    return -resident_stress_samples(phys, sim, n_traj, times, **kwargs) / rate
