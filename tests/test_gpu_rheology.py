"""Shear stress of a resident ensemble (bdsim_amd.rheology): the same numbers as
integrating each trajectory to the sample times and evaluating
tau_xy = sum_mu (R_mu - R_c)_x F_mu,y with the oracle's total force
(python/bdsim/rheology.py:36-51, properties.py:44-52), and paired zero-flow
siblings that cut the variance at low Wi without bias.

Tolerance: positions are within 1e-10 of the oracle over the segments; the
stress sums run in another order on the device, so tau_xy agrees to 1e-9 of its
scale.
"""
from ctypes import byref

import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from bdsim_amd import rheology
from bdsim_amd.core import DelSMethod, EV, Flow, PhysParams, SimParams, Spring
from bdsim_amd.initial import gaussian_chain

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


def _oracle_stress(orc, phys, sim, N, seed, times, flow=True):
    """tau_xy of chain `seed` at the sample times, on the oracle: one integrate
    segment [t_k, t_{k+1}] per sample, on one generator stream."""
    pk = A.PackedPhys(phys, N)
    if not flow:
        for k in range(9):
            pk.c.flow.kappa0[k] = 0.0
    R = gaussian_chain(N, seed=seed)
    st = A.CRng(1, -1, seed)
    out, t = [], sim.time_start
    for ts in times:
        R, rc, _, _ = H.oracle_integrate(orc, N, R, pk.c, A.CSim(t, ts, sim.dt, sim.implicit_loop_tol, 1), st)
        assert rc == 0
        F = np.empty_like(R)
        orc.bo_total_force(N, A.dptr(np.ascontiguousarray(R)), byref(pk.c), A.dptr(F))
        Rc = R - R.mean(axis=0)
        out.append(float(np.sum(Rc[:, 0] * F[:, 1])))
        t = ts
    return np.array(out)


@pytest.mark.parametrize("control", [False, True])
def test_resident_stress_equals_oracle(gpu, orc, control):
    N, n = 12, 6
    p = _phys(N, 0.5, hstar=0.25, spring=Spring.FENE, sqrtb=10.0, ev=EV.Gauss)
    sim = _sim(0.01)
    times = [0.05, 0.1, 0.17]
    got = rheology.resident_stress_samples(p, sim, n, times, seed=3, control=control)
    sheared = got[0] if control else got
    assert sheared.shape == (n, 3)
    for i in range(n):
        want = _oracle_stress(orc, p, sim, N, 3 + i, times)
        assert np.max(np.abs(sheared[i] - want)) <= 1e-9 * np.max(np.abs(want)), i
        if control:
            want0 = _oracle_stress(orc, p, sim, N, 3 + i, times, flow=False)
            assert np.max(np.abs(got[1][i] - want0)) <= 1e-9 * np.max(np.abs(want0)), i


def test_variance_reduction_at_low_wi(gpu):
    """Hookean N=10 at Wi = 0.03 (rate Wi sin^2(pi/2N)): the paired estimate
    is unbiased against the direct one and its scatter over trajectories is
    far smaller (the reference measured 31x in variance at Wi = 0.03)."""
    N, n = 10, 1024
    rate = 0.03 * np.sin(np.pi / (2 * N)) ** 2
    p = _phys(N, rate)
    sim = _sim(0.02)
    times = list(np.arange(1, 41) * 2.0)
    direct = rheology.resident_viscosity_samples(p, sim, rate, n, times, seed=11).mean(axis=1)
    paired = rheology.resident_viscosity_samples(p, sim, rate, n, times, variance_reduction=True,
                                                 seed=11).mean(axis=1)
    vd, vp = direct.var(ddof=1), paired.var(ddof=1)
    se = np.sqrt(vd / n + vp / n)
    print(f"Wi 0.03: direct {direct.mean():.2f} +/- {np.sqrt(vd / n):.2f}, paired "
          f"{paired.mean():.3f} +/- {np.sqrt(vp / n):.3f}, variance ratio {vd / vp:.0f}")
    assert vd / vp > 10
    assert abs(direct.mean() - paired.mean()) < 3 * se
