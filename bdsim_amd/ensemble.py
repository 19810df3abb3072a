"""Running trajectories and ensembles (python/bdsim/ensemble.py:1-379).

`simulate` (snapshots to disk) and `run_ensemble` (a per-trajectory function
reduced to results) with the reference's signatures, seeding (seed + i), segment
semantics and failure dropping. `backend="gpu"` runs all trajectories together:
every `integrate` call made at the same point of the trajectories' programs
goes to the GPU as one batched launch (bdsim_amd/batcher.py).
"""
import os
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _abi, storage
from .core import PhysParams, Rng, SimParams, integrate, total_force
from .initial import (fene_fraenkel_bending_chain, fene_fraenkel_chain,
                      fene_fraenkel_chain_aligned_x, gaussian_chain)
from .parallel import Failure, parallel_map


def _init_gaussian(n, seed, kw):              return gaussian_chain(n, seed=seed, **kw)
def _init_fene_fraenkel(n, seed, kw):         return fene_fraenkel_chain(n, seed=seed, **kw)
def _init_fene_fraenkel_x(n, seed, kw):       return fene_fraenkel_chain_aligned_x(n, seed=seed, **kw)
def _init_fene_fraenkel_bending(n, seed, kw): return fene_fraenkel_bending_chain(n, seed=seed, **kw)


INITIALIZERS = {
    "gaussian": _init_gaussian,
    "fene_fraenkel": _init_fene_fraenkel,
    "fene_fraenkel_x": _init_fene_fraenkel_x,
    "fene_fraenkel_bending": _init_fene_fraenkel_bending,
}


@dataclass
class Initial:
    """How to build each trajectory's starting chain (ensemble.py:56-65)."""
    method: str = "gaussian"
    kwargs: dict = field(default_factory=dict)


@dataclass
class Output:
    """What `simulate` records (ensemble.py:68-77)."""
    directory: str = None
    write_every: int = 0
    write_forces: bool = False
    compression: str = "gzip"


def _segment(sim: SimParams, t0: float, t1: float) -> SimParams:
    """`sim` restricted to [t0, t1] (ensemble.py:84-92)."""
    s = SimParams()
    s.dt = sim.dt
    s.implicit_loop_tol = sim.implicit_loop_tol
    s.update_center_of_mass = sim.update_center_of_mass
    s.time_start = t0
    s.time_end = t1
    return s


def n_steps_of(sim: SimParams) -> int:
    return int(round((sim.time_end - sim.time_start) / sim.dt))


def _snapshot(step, R, phys, sim, write_forces):
    t = sim.time_start + step * sim.dt
    F = total_force(R, phys) if write_forces else None
    return step, t, np.array(R, dtype=np.float64, copy=True), F


def _iter_snapshots(R0, phys, sim, rng, n_steps, write_every, write_forces):
    """Integrate `n_steps` steps in chunks of `write_every` (ensemble.py:106-123):
    a chunk of k steps is the segment [t0, t0 + (k-1) dt]."""
    R = np.asarray(R0, dtype=np.float64)
    w = write_every if (write_every and write_every > 0) else n_steps
    step = 0
    yield _snapshot(step, R, phys, sim, write_forces)
    while step < n_steps:
        chunk = min(w, n_steps - step)
        t0 = sim.time_start + step * sim.dt
        R = integrate(R, phys, _segment(sim, t0, t0 + (chunk - 1) * sim.dt), rng)
        step += chunk
        yield _snapshot(step, R, phys, sim, write_forces)


def trajectory_samples(R0, phys: PhysParams, sim: SimParams, rng: Rng, sample_times):
    """Configurations at each of `sample_times` (ensemble.py:126-138)."""
    R = np.asarray(R0, dtype=np.float64)
    out, t = [], sim.time_start
    for ts in sample_times:
        R = integrate(R, phys, _segment(sim, t, ts), rng)
        out.append((ts, R.copy()))
        t = ts
    return out


def _phys_summary(phys: PhysParams) -> dict:
    return {"n_beads": int(phys.number_of_beads), "spring_type": str(phys.spring.type),
            "sqrtb": float(phys.spring.sqrtb), "natural_length": float(phys.spring.natural_length),
            "hstar": float(phys.hstar), "hi_method": str(phys.hi_method)}


def _simulate_worker(job):
    (index, phys, sim, n_steps, directory, write_every, write_forces,
     compression, seed, method, kwargs) = job
    R0 = INITIALIZERS[method](phys.number_of_beads, seed, kwargs)
    rng = Rng(seed)
    steps, times, positions = [], [], []
    forces = [] if write_forces else None
    for st, t, R, F in _iter_snapshots(R0, phys, sim, rng, n_steps, write_every, write_forces):
        steps.append(st)
        times.append(t)
        positions.append(R)
        if write_forces:
            forces.append(F)
    attrs = {"index": index, "seed": seed, "n_steps": n_steps,
             "write_every": write_every or n_steps, "dt": float(sim.dt),
             **{f"phys_{k}": v for k, v in _phys_summary(phys).items()}}
    path = storage.trajectory_path(directory, index)
    storage.write_trajectory(path, steps, times, positions, forces, attrs=attrs,
                             compression=compression)
    return os.path.basename(path)


MAX_FAILED_FRACTION = 0.1


def _drop_failures(results, what, max_fraction=None):
    """Split results into successes and failures (ensemble.py:244-284)."""
    ok = [r for r in results if not isinstance(r, Failure)]
    bad = [r for r in results if isinstance(r, Failure)]
    if not bad:
        return ok, bad
    n = len(results)
    frac = len(bad) / n
    seen, lines = set(), []
    for f in bad:
        key = f.message.split("(")[0]
        if key not in seen:
            seen.add(key)
            lines.append(f"    trajectory {f.index}: {f.message.splitlines()[0]}")
    print(f"bdsim: {len(bad)} of {n} {what} failed and were dropped "
          f"({100 * frac:.0f}%). Distinct causes:", file=sys.stderr)
    for line in lines[:3]:
        print(line, file=sys.stderr)
    limit = MAX_FAILED_FRACTION if max_fraction is None else max_fraction
    if frac > limit:
        raise RuntimeError(f"{len(bad)} of {n} {what} failed ({100 * frac:.0f}%), above the "
                           f"{100 * limit:.0f}% limit. Reduce dt or tighten implicit_loop_tol.")
    if not ok:
        raise RuntimeError(f"every one of the {n} {what} failed")
    return ok, bad


def simulate(phys: PhysParams, sim: SimParams, *, n_traj: int = 1, seed: int = 0,
             n_steps: int = None, initial: Initial = None, output: Output = None,
             backend: str = "serial", n_workers: int = None, on_error: str = "skip",
             max_failed_fraction: float = None, precision=_abi.PRECISION_FP64):
    """Run `n_traj` chains seeded by `seed + i` and write their snapshots
    (ensemble.py:179-218). Returns a `storage.Run` over the run directory."""
    initial = initial or Initial()
    output = output or Output()
    if output.directory is None:
        raise ValueError("output.directory must be set (the run folder to write into)")
    n_steps = int(n_steps) if n_steps is not None else n_steps_of(sim)
    os.makedirs(output.directory, exist_ok=True)
    jobs = [(i, phys, sim, n_steps, output.directory, output.write_every,
             output.write_forces, output.compression, seed + i,
             initial.method, dict(initial.kwargs)) for i in range(n_traj)]
    files = parallel_map(_simulate_worker, jobs, backend=backend, n_workers=n_workers,
                         on_error=on_error, precision=precision)
    files, failed = _drop_failures(files, "trajectories", max_failed_fraction)
    manifest = {"n_trajectories": len(files), "n_trajectories_requested": n_traj,
                "n_trajectories_failed": len(failed), "seed": seed, "n_steps": n_steps,
                "dt": float(sim.dt), "write_every": output.write_every or n_steps,
                "write_forces": bool(output.write_forces),
                "initial": {"method": initial.method, "kwargs": initial.kwargs},
                "phys": _phys_summary(phys), "files": list(files)}
    storage.write_manifest(output.directory, manifest)
    return storage.read_run(output.directory)


def mean_stderr(values):
    arr = np.asarray(values, dtype=float)
    n = arr.size
    return float(arr.mean()), (float(arr.std(ddof=1) / np.sqrt(n)) if n > 1 else float("nan"))


def _run_worker(job):
    fn, phys, sim, seed, method, kwargs, args = job
    R0 = INITIALIZERS[method](phys.number_of_beads, seed, kwargs)
    return fn(R0, phys, sim, Rng(seed), *args)


def final_state(R0, phys, sim, rng):
    """A per_trajectory function: the final (N, 3) configuration."""
    return np.asarray(integrate(R0, phys, sim, rng), dtype=np.float64)


def sampled_states(R0, phys, sim, rng, sample_times):
    """A per_trajectory function: configurations at `sample_times`."""
    return [R for _t, R in trajectory_samples(R0, phys, sim, rng, sample_times)]


def run_ensemble(phys: PhysParams, sim: SimParams, n_traj: int, per_trajectory, *, args=(),
                 seed: int = 0, initial: Initial = None, backend="serial", n_workers=None,
                 on_error="skip", max_failed_fraction=None, precision=_abi.PRECISION_FP64):
    """Run `n_traj` chains, reducing each with `per_trajectory(R0, phys, sim,
    rng, *args)` (ensemble.py:330-379)."""
    initial = initial or Initial()
    jobs = [(per_trajectory, phys, sim, seed + i, initial.method, dict(initial.kwargs),
             tuple(args)) for i in range(n_traj)]
    out = parallel_map(_run_worker, jobs, backend=backend, n_workers=n_workers,
                       on_error=on_error, precision=precision)
    out, _ = _drop_failures(out, "trajectories", max_failed_fraction)
    return out
