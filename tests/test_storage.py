"""The run-directory format (python/bdsim/storage.py:1-78) written and read
through the HDF5 library: layout, types and attributes as the reference's
h5py writer leaves them (no device needed)."""
import json
import os

import numpy as np
import pytest

from bdsim_amd import h5, storage

pytestmark = pytest.mark.skipif(not h5.available(), reason="the HDF5 library is not installed")


def test_trajectory_file_layout_and_round_trip(tmp_path):
    rs = np.random.default_rng(0)
    steps = np.array([0, 10, 20, 25])
    times = 0.5 + steps * 1e-3
    R = rs.normal(size=(4, 12, 3))
    F = rs.normal(size=(4, 12, 3)) * 1e3
    attrs = {"index": 3, "seed": 7, "n_steps": 25, "write_every": 10, "dt": 1e-3,
             "phys_spring_type": "Spring.FENE", "phys_hstar": 0.25, "label": "ünïcode"}
    p = str(tmp_path / "traj_00003.h5")
    storage.write_trajectory(p, steps, times, R, F, attrs=attrs)
    assert open(p, "rb").read(8) == b"\x89HDF\r\n\x1a\n"          # the HDF5 signature
    t = storage.read_trajectory(p)
    assert t.step.dtype == np.int64 and np.array_equal(t.step, steps)
    assert t.time.dtype == np.float64 and np.array_equal(t.time, times)
    assert np.array_equal(t.positions, R) and np.array_equal(t.forces, F)
    got = t.attrs
    assert got == dict(attrs, has_forces=True)
    assert isinstance(got["has_forces"], bool) and isinstance(got["index"], int)
    snaps = list(t)
    assert len(snaps) == 4 and snaps[3][0] == 25 and np.array_equal(snaps[2][2], R[2])
    with h5.File(p, "r") as f:
        assert "positions" in f and "forces" in f and "velocities" not in f


def test_no_forces_and_compression(tmp_path):
    p = str(tmp_path / "t.h5")
    Z = np.zeros((200, 50, 3))
    storage.write_trajectory(p, np.arange(200), np.arange(200) * 0.01, Z)
    t = storage.read_trajectory(p)
    assert t.forces is None and t.attrs["has_forces"] is False
    assert np.array_equal(t.positions, Z)
    assert os.path.getsize(p) < Z.nbytes / 10           # positions are gzip-compressed


def test_run_directory_and_ensemble_average(tmp_path):
    d = str(tmp_path / "run")
    rs = np.random.default_rng(1)
    files = []
    for i in range(3):
        R = rs.normal(size=(5, 8, 3))
        os.makedirs(d, exist_ok=True)
        p = storage.trajectory_path(d, i)
        storage.write_trajectory(p, np.arange(5) * 2, np.arange(5) * 0.002, R, attrs={"index": i})
        files.append(os.path.basename(p))
    storage.write_manifest(d, {"n_trajectories": 3, "files": files})
    run = storage.read_run(d)
    assert len(run) == 3 and run.manifest["files"] == ["traj_00000.h5", "traj_00001.h5", "traj_00002.h5"]
    assert json.load(open(os.path.join(d, "manifest.json")))["n_trajectories"] == 3
    t, mean, se = storage.ensemble_average(run, lambda R: float(np.sum(R[:, 0])))
    vals = np.array([[np.sum(tr.positions[k][:, 0]) for k in range(5)] for tr in run])
    assert np.allclose(t, np.arange(5) * 0.002)
    assert np.allclose(mean, vals.mean(axis=0)) and np.allclose(se, vals.std(axis=0, ddof=1) / np.sqrt(3))


def test_unsupported_values_are_refused(tmp_path):
    with h5.File(str(tmp_path / "x.h5"), "w") as f:
        with pytest.raises(TypeError):
            f.set_attr("a", [1, 2])
        with pytest.raises(TypeError):
            f.create_dataset("s", np.array(["x"]))
