"""The run directory of `simulate` (python/bdsim/storage.py:1-78), as HDF5.

A run is a directory with `manifest.json` (ensemble parameters and the list of
files) and one `traj_NNNNN.h5` per trajectory holding its snapshots:

    attrs       index, seed, n_steps, write_every, dt, phys_*, has_forces
    step        (n_snap,)        int64    integrator step of each snapshot
    time        (n_snap,)        float64  time of each snapshot
    positions   (n_snap, N, 3)   float64  gzip-compressed
    forces      (n_snap, N, 3)   float64  total bead forces, when written

The files are written and read by the HDF5 library itself (bdsim_amd.h5), with
h5py's conventions for attributes, so the reference's reader opens them.

`record_run` fills such a directory from a device-resident ensemble: the
chains step on the GPU in segments of `write_every` steps (each snapshot
chunk one integrate segment on the shared flow clock, as the reference's
_iter_snapshots does, ensemble.py:106-123) and only the snapshots cross to
the host; forces come from the device force kernel.
"""
import ctypes
import json
import os

import numpy as np

from . import _abi, h5


def trajectory_path(directory, index):
    """The conventional per-trajectory file name inside a run directory."""
    return os.path.join(directory, f"traj_{index:05d}.h5")


def write_trajectory(path, steps, times, positions, forces=None, attrs=None, compression="gzip"):
    """One trajectory's snapshots (steps/times of length n_snap, positions and
    forces (n_snap, N, 3) or forces None) and scalar attributes to `path`."""
    with h5.File(path, "w") as f:
        for k, v in (attrs or {}).items():
            f.set_attr(k, v)
        f.set_attr("has_forces", forces is not None)
        f.create_dataset("step", np.asarray(steps, dtype=np.int64))
        f.create_dataset("time", np.asarray(times, dtype=np.float64))
        f.create_dataset("positions", np.asarray(positions, dtype=np.float64), compression=compression)
        if forces is not None:
            f.create_dataset("forces", np.asarray(forces, dtype=np.float64), compression=compression)


def write_manifest(directory, manifest):
    os.makedirs(directory, exist_ok=True)
    with open(os.path.join(directory, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=2, default=str)


class Trajectory:
    """One traj_NNNNN.h5: `.step`, `.time`, `.positions`, `.forces` (or None)
    and `.attrs`, each read from the file when asked for; iterating yields
    (step, time, R, F or None) per snapshot."""

    def __init__(self, path):
        self.path = path

    def _get(self, name):
        with h5.File(self.path, "r") as f:
            return f.read(name) if name in f else None

    @property
    def attrs(self):
        with h5.File(self.path, "r") as f:
            return f.attrs()

    @property
    def step(self):
        return self._get("step")

    @property
    def time(self):
        return self._get("time")

    @property
    def positions(self):
        return self._get("positions")

    @property
    def forces(self):
        return self._get("forces")

    def __len__(self):
        return len(self.step)

    def __iter__(self):
        with h5.File(self.path, "r") as f:
            step, time, pos = f.read("step"), f.read("time"), f.read("positions")
            frc = f.read("forces") if "forces" in f else None
        for i in range(len(step)):
            yield int(step[i]), float(time[i]), pos[i], (None if frc is None else frc[i])


class Run:
    """A run directory: `.manifest` and its trajectories (len, iteration, [i])."""

    def __init__(self, directory):
        self.directory = directory
        with open(os.path.join(directory, "manifest.json")) as fh:
            self.manifest = json.load(fh)
        self.files = [os.path.join(directory, f) for f in self.manifest["files"]]

    def __len__(self):
        return len(self.files)

    def __iter__(self):
        return (Trajectory(p) for p in self.files)

    def __getitem__(self, i):
        return Trajectory(self.files[i])


def read_trajectory(path):
    return Trajectory(path)


def read_run(directory):
    return Run(directory)


def ensemble_average(run, fn, use_forces=False):
    """(times, mean, standard error) over the run's trajectories of fn(R) (or
    fn(R, F)) at each snapshot; the trajectories share their snapshot times."""
    rows, times = [], None
    for traj in run:
        vals, ts = zip(*[((fn(R, F) if use_forces else fn(R)), t) for _s, t, R, F in traj])
        rows.append(vals)
        times = np.asarray(ts) if times is None else times
    v = np.asarray(rows, dtype=float)
    se = v.std(axis=0, ddof=1) / np.sqrt(len(v)) if len(v) > 1 else np.full(v.shape[1], np.nan)
    return times, v.mean(axis=0), se


def phys_summary(phys):
    """The small description of PhysParams kept in attributes and manifests
    (the fields of the reference's _phys_summary, ensemble.py:145-154)."""
    return {"n_beads": int(phys.number_of_beads), "spring_type": str(phys.spring.type),
            "sqrtb": float(phys.spring.sqrtb), "natural_length": float(phys.spring.natural_length),
            "hstar": float(phys.hstar), "hi_method": str(phys.hi_method)}


def record_run(directory, R0, phys, sim, seeds, n_steps, write_every=None, write_forces=False,
               precision=_abi.PRECISION_FP64, compression="gzip", manifest_extra=None):
    """Integrate the chains R0 (n, N, 3) with generators Rng(seeds[i]) for
    n_steps steps on the GPU as one resident batch and write the run
    directory: snapshots at step 0, every `write_every` steps and at the last
    step. Returns a Run over it. Raises RuntimeError if a chain fails."""
    from .resident import ResidentEnsemble
    R0 = np.ascontiguousarray(R0, dtype=np.float64)
    n, N = R0.shape[0], R0.shape[1]
    w = write_every if (write_every and write_every > 0) else n_steps
    lib = _abi.load_gpu()
    pk = _abi.PackedPhys(phys, N)

    def forces_of(R):
        F = np.empty_like(R)
        _abi.check(lib, lib.bdsim_total_force_batch(n, N, _abi.dptr(R), ctypes.byref(pk.c), _abi.dptr(F)))
        return F

    steps, times, pos, frc = [0], [sim.time_start], [R0.copy()], [forces_of(R0)] if write_forces else None
    with ResidentEnsemble(R0, phys, sim, list(seeds), precision) as ens:
        step = 0
        while step < n_steps:
            chunk = min(w, n_steps - step)
            ens.run(chunk)
            step += chunk
            R, _, status, _ = ens.fetch()
            if np.any(status != 0):
                bad = int(np.flatnonzero(status != 0)[0])
                raise RuntimeError(f"trajectory {bad} failed with status {int(status[bad])} "
                                   f"before step {step}")
            steps.append(step)
            times.append(sim.time_start + step * sim.dt)
            pos.append(R.copy())
            if write_forces:
                frc.append(forces_of(R))
    os.makedirs(directory, exist_ok=True)
    files = []
    P = np.stack(pos, axis=1)                      # (n, n_snap, N, 3)
    Fa = np.stack(frc, axis=1) if write_forces else None
    summary = dict(phys_summary(phys), n_beads=N)
    for i in range(n):
        attrs = {"index": i, "seed": int(seeds[i]), "n_steps": int(n_steps), "write_every": int(w),
                 "dt": float(sim.dt), **{f"phys_{k}": v for k, v in summary.items()}}
        path = trajectory_path(directory, i)
        write_trajectory(path, steps, times, P[i], None if Fa is None else Fa[i], attrs=attrs,
                         compression=compression)
        files.append(os.path.basename(path))
    manifest = {"n_trajectories": n, "n_trajectories_requested": n, "n_trajectories_failed": 0,
                "seed": int(seeds[0]) if n else 0, "n_steps": int(n_steps), "dt": float(sim.dt),
                "write_every": int(w), "write_forces": bool(write_forces), "phys": summary,
                "files": files, **(manifest_extra or {})}
    write_manifest(directory, manifest)
    return read_run(directory)
