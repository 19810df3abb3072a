"""Snapshot observables on the device, and ensemble statistics of long runs.

Observables (bdsim_observables_batch / bdsim_batch_observe) against
properties.py's formulas evaluated in numpy on the same positions with the
oracle's total force: within 1e-12 relative per field (floating-point sums in
another order).

Statistics, the reference's own physics checks (validation/validate_physics.py:
20-70, docs/theory.tex:2057-2075): an ideal Hookean chain keeps <R^2> = 3(N-1)
and <Rg^2> = (N^2-1)/(2N) with and without hydrodynamic interaction, within 3
standard errors; its zero-shear viscosity is lambda_eta = (N^2-1)/3; the mixed
form's ensemble averages agree with the FP64 form's within combined standard
errors.
"""
from ctypes import byref

import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from bdsim_amd import properties as props
from bdsim_amd.core import EV, Bending, DelSMethod, Flow, PhysParams, SimParams, Spring
from bdsim_amd.resident import ResidentEnsemble, observe
from golden import cases as C

pytestmark = pytest.mark.gpu


def _phys(spring=Spring.Hook, sqrtb=1000.0, hstar=0.0, method=DelSMethod.Chebyshev, ev=EV.None_,
          zstar=0.0, dstar=0.0, bend=0.0, kappa=None):
    p = PhysParams()
    p.spring.type, p.spring.sqrtb = spring, sqrtb
    p.hstar, p.hi_method = hstar, method
    p.ev.type, p.ev.zstar, p.ev.dstar = ev, zstar, dstar
    if bend:
        p.bend.type, p.bend.stiffness = Bending.OneMinusCosTheta, bend
    if kappa is not None:
        p.flow = Flow.constant(kappa)
    return p


def _sim(dt, tol=1e-3):
    s = SimParams()
    s.time_start, s.dt, s.implicit_loop_tol = 0.0, dt, tol
    return s


def _chains(n, N, seed, bond=1.0):
    return np.ascontiguousarray(np.array([C.random_walk(N, seed + i, bond) for i in range(n)]))


def _numpy_observables(R, F):
    return {"rg2": props.radius_of_gyration_sq(R), "q2": props.end_to_end_sq(R),
            "gyration": props.gyration_tensor(R), "stretch": props.stretch(R),
            "stress": props.kramers_stress(R, F), "com": R.mean(axis=0)}


def _close(got, want, tol=1e-12, scale=0.0):
    """Max abs difference within tol of the larger of max|want| and scale."""
    want = np.asarray(want, dtype=np.float64)
    scale = max(np.max(np.abs(want)), scale, 1e-300)
    return float(np.max(np.abs(np.asarray(got) - want)) / scale) <= tol


@pytest.mark.parametrize("N", [10, 50, 130, 300])
def test_observables_match_properties(orc, N):
    p = _phys(spring=Spring.FENE, sqrtb=10.0, ev=EV.SDK, dstar=1.0, bend=2.0)
    R = _chains(6, N, 40 + N, 0.8) + np.array([1.5, -2.0, 0.7])   # centre of mass away from 0
    got = observe(R, p, force="total")
    spring = observe(R, p, force="spring")
    c = A.PackedPhys(p, N)
    for i in range(len(R)):
        F = np.empty((N, 3))
        orc.bo_total_force(N, A.dptr(np.ascontiguousarray(R[i])), byref(c.c), A.dptr(F))
        Fs = np.empty((N, 3))
        orc.bo_spring_force(N, A.dptr(np.ascontiguousarray(R[i])), byref(c.c), A.dptr(Fs))
        want = _numpy_observables(R[i], F)
        for k, v in want.items():
            assert _close(got[k][i], v, scale=np.abs(R[i]).max() if k == "com" else 0.0), (N, i, k)
        assert _close(spring["stress"][i], props.kramers_stress(R[i], Fs)), (N, i)
    # a force part changes the stress only
    assert np.array_equal(spring["rg2"], got["rg2"])


def test_resident_observe_matches_fetched_positions(orc):
    N, n = 20, 16
    p = _phys(spring=Spring.FENE, sqrtb=10.0, hstar=0.3, ev=EV.Gauss, zstar=0.5, dstar=0.5)
    kap = np.array([np.array(C.wi_kappa(0.1 * i)).reshape(3, 3) for i in range(n)])
    with ResidentEnsemble(_chains(n, N, 900), p, _sim(1e-3), list(range(n)), kappas=kap) as ens:
        ens.run(30)
        o = ens.observe()
        R, _, status, _ = ens.fetch()
    assert np.all(status == 0)
    c = A.PackedPhys(p, N)
    for i in range(n):
        F = np.empty((N, 3))
        orc.bo_total_force(N, A.dptr(np.ascontiguousarray(R[i])), byref(c.c), A.dptr(F))
        for k, v in _numpy_observables(R[i], F).items():
            # re-centred chains: the centre of mass is rounding noise around 0
            assert _close(o[k][i], v, scale=np.abs(R[i]).max() if k == "com" else 0.0), (i, k)


def _mean_se(x):
    x = np.asarray(x, dtype=np.float64)
    return float(x.mean()), float(x.std(ddof=1) / np.sqrt(len(x)))


def _equilibrium(N, n_traj, precision=A.PRECISION_FP64, hstar=0.0, method=DelSMethod.Chebyshev,
                 seed=1):
    """validate_physics.py:20-37 on the device: Hookean chains started in
    equilibrium, integrated to t=8 with dt=0.05; one sample per chain."""
    p = _phys(hstar=hstar, method=method)
    sim = _sim(0.05, 1e-4)
    R0 = _chains(n_traj, N, 10_000 * seed)
    with ResidentEnsemble(R0, p, sim, [seed + i for i in range(n_traj)], precision) as ens:
        ens.run(161)
        o = ens.observe()
        R, _, status, _ = ens.fetch()
    assert np.all(status == 0)
    # the device observables are those of the final positions
    assert _close(o["q2"], [props.end_to_end_sq(r) for r in R])
    return _mean_se(o["q2"]), _mean_se(o["rg2"])


@pytest.mark.parametrize("N", [5, 10, 20])
def test_equilibrium_moments_free_draining(N):
    (r2, r2e), (rg, rge) = _equilibrium(N, 4096)
    print(f"N={N} free-draining: <R^2> {r2:.3f} +/- {r2e:.3f} (theory {3.0 * (N - 1):.2f}), "
          f"<Rg^2> {rg:.4f} +/- {rge:.4f} (theory {0.5 * (N * N - 1) / N:.4f})")
    assert abs(r2 - 3.0 * (N - 1)) < 3 * r2e, (r2, r2e)
    assert abs(rg - 0.5 * (N * N - 1) / N) < 3 * rge, (rg, rge)


@pytest.mark.parametrize("method,precision", [(DelSMethod.Cholesky, A.PRECISION_FP64),
                                              (DelSMethod.Chebyshev, A.PRECISION_FP64),
                                              (DelSMethod.Chebyshev, A.PRECISION_MIXED)])
def test_equilibrium_moments_with_hi(method, precision):
    """HI leaves the equilibrium unchanged (validate_physics.py:56-68)."""
    N = 10
    (r2, r2e), (rg, rge) = _equilibrium(N, 4096, precision, hstar=0.2, method=method)
    print(f"N=10 h*=0.2 {method.name} precision {precision}: <R^2> {r2:.3f} +/- {r2e:.3f}, "
          f"<Rg^2> {rg:.4f} +/- {rge:.4f}")
    assert abs(r2 - 3.0 * (N - 1)) < 3 * r2e, (r2, r2e)
    assert abs(rg - 0.5 * (N * N - 1) / N) < 3 * rge, (rg, rge)


def test_rouse_zero_shear_viscosity():
    """lambda_eta = -<tau_xy>/rate = (N^2-1)/3 for a free-draining Hookean
    chain (theory.tex:2057-2075; the stress of a Hookean chain is linear in the
    rate at any Wi). N=10, rate 0.1 (Wi 4), 4096 chains: 5 longest relaxation
    times of start-up, then 10 samples 20 time units apart averaged per chain."""
    N, n, rate, dt = 10, 4096, 0.1, 0.01
    p = _phys(kappa=[[0, rate, 0], [0, 0, 0], [0, 0, 0]])
    with ResidentEnsemble(_chains(n, N, 77), p, _sim(dt, 1e-4), list(range(n))) as ens:
        ens.run(20_000)
        acc = np.zeros(n)
        for _ in range(10):
            ens.run(2000)
            acc += ens.observe(force="spring")["stress"][:, 0, 1]
        _, _, status, _ = ens.fetch()
    assert np.all(status == 0)
    eta, se = _mean_se(-acc / 10 / rate)
    print(f"Rouse N=10 lambda_eta {eta:.3f} +/- {se:.3f} (theory {(N * N - 1) / 3.0:.3f})")
    assert abs(eta - (N * N - 1) / 3.0) < 3 * se + 0.01 * (N * N - 1) / 3.0, (eta, se)


def _observed_averages(p, N, n, sim, steps, every, precision, kappas=None, seed=500):
    with ResidentEnsemble(_chains(n, N, seed), p, sim, list(range(n)), precision, kappas) as ens:
        acc = {"rg2": 0.0, "q2": 0.0, "txy": 0.0, "n1": 0.0}
        for _ in range(steps // every):
            ens.run(every)
            o = ens.observe()
            acc["rg2"] = acc["rg2"] + o["rg2"]
            acc["q2"] = acc["q2"] + o["q2"]
            acc["txy"] = acc["txy"] + o["stress"][:, 0, 1]
            acc["n1"] = acc["n1"] + (o["stress"][:, 0, 0] - o["stress"][:, 1, 1])
        _, _, status, _ = ens.fetch()
    assert np.all(status == 0)
    return {k: _mean_se(v / (steps // every)) for k, v in acc.items()}


@pytest.mark.parametrize("case", ["shear_chebyshev", "planar_cholesky"])
def test_mixed_form_ensemble_averages(case):
    """Mixed and FP64 forms from the same starts and streams: <Rg^2>, <Q^2>,
    shear stress and first normal stress difference agree within combined
    standard errors. Shear: the headline physics (N=50, FENE b=100, SDK,
    h*=0.3 Chebyshev, Wi 10). Planar extension: N=30, FENE b=50, Gauss EV,
    h*=0.25 Cholesky, rate 0.5."""
    if case == "shear_chebyshev":
        N, n = 50, 1024
        rate = 10.0 * np.sin(np.pi / (2 * N)) ** 2
        p = _phys(spring=Spring.FENE, sqrtb=10.0, hstar=0.3, ev=EV.SDK, dstar=1.0,
                  kappa=[[0, rate, 0], [0, 0, 0], [0, 0, 0]])
        steps, every = 2000, 200
    else:
        N, n = 30, 1024
        p = _phys(spring=Spring.FENE, sqrtb=np.sqrt(50.0), hstar=0.25, method=DelSMethod.Cholesky,
                  ev=EV.Gauss, zstar=1.0, dstar=0.5,
                  kappa=[[0.5, 0, 0], [0, -0.5, 0], [0, 0, 0]])
        steps, every = 1000, 100
    sim = _sim(1e-3)
    a = _observed_averages(p, N, n, sim, steps, every, A.PRECISION_FP64)
    b = _observed_averages(p, N, n, sim, steps, every, A.PRECISION_MIXED)
    for k in a:
        (ma, sa), (mb, sb) = a[k], b[k]
        print(f"{case} {k}: fp64 {ma:.6g} +/- {sa:.2g}, mixed {mb:.6g} +/- {sb:.2g}")
        assert abs(ma - mb) < np.hypot(sa, sb), (k, a[k], b[k])
