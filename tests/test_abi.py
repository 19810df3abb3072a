"""The C-ABI library loads and exports every entry point declared in
include/bdsim_gpu.h (no device needed: symbols only, nothing is called)."""
import os
import re

import pytest

from bdsim_amd import _abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "bdsim_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bdsim_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("bdsim_init", "bdsim_integrate_batch", "bdsim_batch_create", "bdsim_batch_run",
                 "bdsim_total_force_batch", "bdsim_raw_noise_batch", "bdsim_rpy_apply_batch"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_abi.GPU_LIB_PATH):
        pytest.fail(f"{_abi.GPU_LIB_PATH} not built (python __graft_entry__.py builds it)")
    lib = _abi.load_gpu_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_header():
    import ctypes
    # sizes implied by the header (natural alignment on x86-64 / amdgcn)
    assert ctypes.sizeof(_abi.CRng) == 12
    assert ctypes.sizeof(_abi.CSim) == 40
    assert ctypes.sizeof(_abi.CCounters) == 10 * 8 + 8 + 8
    assert ctypes.sizeof(_abi.CStepLog) == 4 * 8 + 6 * 4
    assert ctypes.sizeof(_abi.CObservables) == 26 * 8


@pytest.mark.parametrize("t0,t1,dt,ok", [
    (0.0, 1.0, 0.01, True), (5.0, 5.0, 1e-3, True), (0.0, -1.0, 0.1, True),
    (0.0, 1.0, 0.0, False), (0.0, 1.0, -0.01, False), (0.0, 1.0, float("nan"), False),
    (0.0, float("inf"), 0.01, False), (float("nan"), 1.0, 0.01, False),
    (1e12, 1e12 + 1.0, 1e-6, False),          # t += dt would not move the clock
    (0.0, 1.0, 1e-10, False),                  # more than 1e9 steps
])
def test_segment_clock_checks(t0, t1, dt, ok):
    """The host-side checks that keep the device step loop finite (the same
    checks the library makes before any launch and returns BDSIM_EINVAL for)."""
    class S:
        time_start, time_end, implicit_loop_tol, update_center_of_mass = t0, t1, 1e-3, True
    S.dt = dt
    if ok:
        _abi.pack_sim(S)
    else:
        with pytest.raises(ValueError):
            _abi.pack_sim(S)
