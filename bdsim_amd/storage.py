"""Run-directory output of `simulate` (python/bdsim/storage.py:33-65).

One file per trajectory plus `manifest.json`. HDF5 (`traj_NNNNN.h5`, datasets
step/time/positions[/forces], attributes) when h5py is importable; otherwise the
same arrays and attributes in `traj_NNNNN.npz` (h5py is absent in this image).
"""
import json
import os

import numpy as np

try:
    import h5py  # noqa: F401
    HAVE_H5PY = True
except ImportError:
    HAVE_H5PY = False


def trajectory_path(directory, index):
    return os.path.join(directory, f"traj_{index:05d}." + ("h5" if HAVE_H5PY else "npz"))


def write_trajectory(path, steps, times, positions, forces=None, attrs=None, compression="gzip"):
    data = {"step": np.asarray(steps, dtype=np.int64), "time": np.asarray(times, dtype=np.float64),
            "positions": np.asarray(positions, dtype=np.float64)}
    if forces is not None:
        data["forces"] = np.asarray(forces, dtype=np.float64)
    if HAVE_H5PY:
        import h5py
        with h5py.File(path, "w") as f:
            for k, v in data.items():
                f.create_dataset(k, data=v, compression=compression if v.ndim > 1 else None)
            for k, v in (attrs or {}).items():
                f.attrs[k] = v
    else:
        np.savez_compressed(path, attrs=np.array(json.dumps(attrs or {})), **data)


def write_manifest(directory, manifest):
    with open(os.path.join(directory, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=2, default=str)


class Trajectory:
    def __init__(self, path):
        self.path = path
        if path.endswith(".h5"):
            import h5py
            with h5py.File(path, "r") as f:
                self.step = f["step"][...]
                self.time = f["time"][...]
                self.positions = f["positions"][...]
                self.forces = f["forces"][...] if "forces" in f else None
                self.attrs = dict(f.attrs)
        else:
            z = np.load(path, allow_pickle=False)
            self.step, self.time, self.positions = z["step"], z["time"], z["positions"]
            self.forces = z["forces"] if "forces" in z.files else None
            self.attrs = json.loads(str(z["attrs"]))


class Run:
    def __init__(self, directory):
        self.directory = directory
        with open(os.path.join(directory, "manifest.json")) as f:
            self.manifest = json.load(f)

    def __len__(self):
        return len(self.manifest["files"])

    def trajectory(self, k):
        return Trajectory(os.path.join(self.directory, self.manifest["files"][k]))

    def __iter__(self):
        for k in range(len(self)):
            yield self.trajectory(k)


def read_run(directory):
    return Run(directory)


def read_trajectory(path):
    return Trajectory(path)
