"""Rheological measurements: shear stress, viscosity, variance reduction
(python/bdsim/rheology.py).

Per-trajectory functions with the reference's signatures (`stress_series`,
`viscosity_series`, `viscosity_series_vr`) for `run_ensemble`, and the
reference's wrappers (`shear_viscosity_series`, `equilibrium_stress_series`).

The wrappers take one more backend, `backend="gpu-resident"`: the whole
ensemble stays in HBM as one resident batch, every sample time is one segment
of every chain (the segments of `trajectory_samples`), and the shear stress
tau_xy = kramers_stress(R, total_force(R, phys))[0, 1] is evaluated on the
device, so only one number per chain and sample crosses to the host. With
variance reduction each sheared chain has a zero-flow sibling in the same
batch, started from the same configuration with the same generator state: the
integrator draws exactly 3N deviates per step whatever the configuration, so
the two see the same Brownian kicks (rheology.py:53-112). Chain i of either
path is the same trajectory (same start, seed + i, same segments).
"""
import copy

import numpy as np

from . import _abi
from . import properties as props
from .core import Flow, Rng, integrate, total_force
from .ensemble import INITIALIZERS, Initial, _segment, run_ensemble, trajectory_samples


# ---- per-trajectory functions (pass these to run_ensemble) -----------------

def stress_series(R0, phys, sim, rng, sample_times):
    """tau_xy at each sample time, for one chain (rheology.py:36-44)."""
    return [float(props.kramers_stress(R, total_force(R, phys))[0, 1])
            for _t, R in trajectory_samples(R0, phys, sim, rng, sample_times)]


def viscosity_series(R0, phys, sim, rng, sample_times, rate):
    """Per-sample -tau_xy/rate, for one chain (rheology.py:47-51)."""
    return [-float(props.kramers_stress(R, total_force(R, phys))[0, 1]) / rate
            for _t, R in trajectory_samples(R0, phys, sim, rng, sample_times)]


def viscosity_series_vr(R0, phys, sim, rng, sample_times, rate):
    """`viscosity_series` minus a zero-flow control chain driven by the same
    random stream (rheology.py:54-112): -(tau_xy - tau_xy^eq)/rate."""
    phys_eq = copy.deepcopy(phys)
    phys_eq.flow = Flow()
    R_f = np.asarray(R0, dtype=np.float64)
    R_e = R_f.copy()
    rng_e = copy.deepcopy(rng)
    out, t = [], sim.time_start
    for ts in sample_times:
        seg = _segment(sim, t, ts)
        R_f = integrate(R_f, phys, seg, rng)
        R_e = integrate(R_e, phys_eq, seg, rng_e)
        tau_f = props.kramers_stress(R_f, total_force(R_f, phys))[0, 1]
        tau_e = props.kramers_stress(R_e, total_force(R_e, phys))[0, 1]
        out.append(-float(tau_f - tau_e) / rate)
        t = ts
    return out


# ---- the ensemble on the device --------------------------------------------

def resident_stress_samples(phys, sim, n_traj, sample_times, *, seed=0, initial=None,
                            control=False, precision=_abi.PRECISION_FP64):
    """tau_xy of n_traj chains at each sample time, (n_traj, n_samples), from
    one resident batch; with control=True also the zero-flow siblings' tau_xy
    (a second array). Chain i starts from the `initial` sampler at seed + i
    with Rng(seed + i), as run_ensemble does. The flow must be constant."""
    from .resident import ResidentEnsemble
    if not phys.flow.is_constant():
        raise ValueError("the resident path takes a constant velocity gradient")
    initial = initial or Initial()
    N = int(phys.number_of_beads)
    R0 = np.array([INITIALIZERS[initial.method](N, seed + i, dict(initial.kwargs))
                   for i in range(n_traj)], dtype=np.float64)
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
        _, _, status, cnt = ens.fetch()
    bad = np.flatnonzero(status != _abi.STATUS_OK)
    if bad.size:
        raise RuntimeError(f"{bad.size} resident chains failed (first: chain {int(bad[0])}, "
                           f"status {int(status[bad[0]])})")
    if control:
        return tau[:n_traj], tau[n_traj:]
    return tau


# ---- wrappers (rheology.py:119-148) ------------------------------------------

def shear_viscosity_series(phys, sim, rate, n_traj, sample_times, *,
                           variance_reduction=False, **kwargs):
    """(n_traj, n_samples) per-sample viscosity contributions. `phys.flow`
    must already be a shear flow at `rate`; remaining keyword arguments go to
    `run_ensemble` (seed, initial, backend, ...), or, with
    backend="gpu-resident", to the resident path (seed, initial, precision)."""
    if kwargs.get("backend") == "gpu-resident":
        kw = {k: v for k, v in kwargs.items() if k in ("seed", "initial", "precision")}
        if variance_reduction:
            tf, te = resident_stress_samples(phys, sim, n_traj, list(sample_times), control=True, **kw)
            return -(tf - te) / rate
        return -resident_stress_samples(phys, sim, n_traj, list(sample_times), **kw) / rate
    fn = viscosity_series_vr if variance_reduction else viscosity_series
    out = run_ensemble(phys, sim, n_traj, fn, args=(list(sample_times), rate), **kwargs)
    return np.asarray(out)


def equilibrium_stress_series(phys, sim, n_traj, sample_times, **kwargs):
    """(n_traj, n_samples) tau_xy(t), for Green-Kubo; `phys.flow` should be the
    zero tensor (rheology.py:133-146)."""
    if kwargs.get("backend") == "gpu-resident":
        kw = {k: v for k, v in kwargs.items() if k in ("seed", "initial", "precision")}
        return resident_stress_samples(phys, sim, n_traj, list(sample_times), **kw)
    out = run_ensemble(phys, sim, n_traj, stress_series, args=(list(sample_times),), **kwargs)
    return np.asarray(out)


def shear_viscosity(phys, sim, rate, n_traj, sample_times, **kwargs):
    """Steady-state polymer shear viscosity as (mean, stderr) over trajectory
    means (rheology.py:149-160)."""
    from .ensemble import mean_stderr
    series = shear_viscosity_series(phys, sim, rate, n_traj, sample_times, **kwargs)
    return mean_stderr([float(np.mean(s)) for s in series])
