"""Shared fixtures. Markers: `gpu` = needs an MI355X-class HIP device."""
import os
import subprocess
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import helpers  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


@pytest.fixture(scope="session")
def orc():
    if not os.path.exists(helpers.ORACLE_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "oracle"], check=True)
    return helpers.oracle()


@pytest.fixture(scope="session")
def refl():
    lib = helpers.ref()
    if lib is None:
        pytest.skip("reference core build (oracle/_ref) not present here")
    return lib


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(helpers.GOLDEN, "golden.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def gpu():
    from bdsim_amd import _abi
    return _abi.load_gpu()
