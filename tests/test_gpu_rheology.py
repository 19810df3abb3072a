"""Rheology on the device (python/bdsim/rheology.py): the resident path gives
the per-trajectory path's numbers, and the paired zero-flow siblings reduce the
variance at low Wi without bias.

Tolerance: positions of the two paths are the same bit for bit (one kernel);
the stress sums run in another order on the device, so tau_xy agrees to 1e-10
relative of its scale.
"""
import numpy as np
import pytest

from bdsim_amd import rheology
from bdsim_amd.core import DelSMethod, EV, Flow, PhysParams, SimParams, Spring
from bdsim_amd.ensemble import Initial

pytestmark = pytest.mark.gpu


def _phys(N, rate, hstar=0.0, spring=Spring.Hook, sqrtb=1000.0, ev=EV.None_):
    p = PhysParams()
    p.number_of_beads = N
    p.spring.type, p.spring.sqrtb = spring, sqrtb
    p.hstar, p.hi_method = hstar, DelSMethod.Chebyshev
    p.ev.type, p.ev.zstar, p.ev.dstar = ev, 0.5, 0.5
    p.flow = Flow.constant([[0.0, rate, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    return p


def _sim(dt):
    s = SimParams()
    s.time_start, s.time_end, s.dt, s.implicit_loop_tol = 0.0, 1.0, dt, 1e-3
    return s


@pytest.mark.parametrize("vr", [False, True])
def test_resident_equals_per_trajectory(vr):
    p = _phys(12, 0.5, hstar=0.25, spring=Spring.FENE, sqrtb=10.0, ev=EV.Gauss)
    sim = _sim(0.01)
    times = [0.05, 0.1, 0.17]
    a = rheology.shear_viscosity_series(p, sim, 0.5, 6, times, variance_reduction=vr, seed=3,
                                        backend="gpu")
    b = rheology.shear_viscosity_series(p, sim, 0.5, 6, times, variance_reduction=vr, seed=3,
                                        backend="gpu-resident")
    assert a.shape == b.shape == (6, 3)
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a)), np.max(np.abs(a - b))


def test_equilibrium_stress_series_resident():
    p = _phys(8, 0.0)
    sim = _sim(0.01)
    a = rheology.equilibrium_stress_series(p, sim, 4, [0.1, 0.2], seed=1, backend="gpu")
    b = rheology.equilibrium_stress_series(p, sim, 4, [0.1, 0.2], seed=1, backend="gpu-resident")
    assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a))


def test_variance_reduction_at_low_wi():
    """Hookean N=10 at Wi = 0.03 (rate Wi sin^2(pi/2N)): the paired estimate
    is unbiased against the direct one and its scatter over trajectories is
    far smaller (the reference measured 31x in variance at Wi = 0.03)."""
    N, n = 10, 1024
    rate = 0.03 * np.sin(np.pi / (2 * N)) ** 2
    p = _phys(N, rate)
    sim = _sim(0.02)
    times = list(np.arange(1, 41) * 2.0)
    direct = rheology.shear_viscosity_series(p, sim, rate, n, times, seed=11,
                                             backend="gpu-resident").mean(axis=1)
    paired = rheology.shear_viscosity_series(p, sim, rate, n, times, variance_reduction=True,
                                             seed=11, backend="gpu-resident").mean(axis=1)
    vd, vp = direct.var(ddof=1), paired.var(ddof=1)
    se = np.sqrt(vd / n + vp / n)
    print(f"Wi 0.03: direct {direct.mean():.2f} +/- {np.sqrt(vd / n):.2f}, paired "
          f"{paired.mean():.3f} +/- {np.sqrt(vp / n):.3f}, variance ratio {vd / vp:.0f}")
    assert vd / vp > 10
    assert abs(direct.mean() - paired.mean()) < 3 * se
