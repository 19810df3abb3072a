"""Failure paths of the device step against the oracle.

The reference ends a trajectory by aborting the process (rtsafe without a
bracket or out of iterations, numerics.hpp:33-36,71-72; a WLC-bounded spring
with sigma > 1, spring.cpp:101-105) or by throwing (cholesky_lower on a
degenerate tensor, linalg.cpp:16-24, here reached through the Chebyshev ->
Cholesky fallback, mobility.cpp:71-80). The library reports a per-trajectory
status instead and the rest of the batch runs on. Each failing chain runs in
one launch beside a healthy chain of the same length; its status, the step
and corrector pass it failed at, its mat-vec count and its generator state
must be the oracle's, in both precision forms. Positions of a failed trajectory
are not compared: the reference returns none.
"""
import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from golden import cases as C

pytestmark = pytest.mark.gpu

FAIL = {c["name"]: c for c in C.failure_cases()}


@pytest.mark.parametrize("prec", [A.PRECISION_FP64, A.PRECISION_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", sorted(FAIL))
def test_failure_status_matches_oracle(gpu, orc, name, prec):
    c = FAIL[name]
    N = c["N"]
    p = H.Packed(c["phys"], N)
    R0 = C.make_R0(c["R0"])
    sim = H.csim(c["sim"])
    st = H.rng_of(c["seed"])
    _, rc, cnto, _ = H.oracle_integrate(orc, N, R0, p.c, sim, st)
    assert rc == c["status"]
    hi = c["phys"]["hstar"] > 0.0
    ok = H.Packed(C.phys_dict(spring=2, sqrtb=10.0, hstar=0.2 if hi else 0.0), N)
    items = [(N, R0, p.c, sim, H.rng_of(c["seed"])),
             (N, C.random_walk(N, 77), ok.c, H.csim(C.sim_dict(0.0, 0.004, 1e-3, 1e-3)),
              H.rng_of(77))]
    out, rngs, status, cnts, _ = H.gpu_batch(gpu, items, precision=prec)
    assert status[0] == rc, (name, status[0], rc)
    g = cnts[0]
    fields = ["steps", "corrector_iters", "fallbacks", "nonconverged", "fail_index"]
    if prec == A.PRECISION_FP64 or not hi:
        # the mixed form's Lanczos products carry FP32 pair terms, so on the
        # degenerate HI chain its Krylov dimension and Chebyshev degrees may
        # differ; the FP64 form follows the oracle's every count
        fields += ["lanczos_iters", "lanczos_cold", "cheb_terms", "cheb_attempts"]
        # the reference forms D x0 again in every Chebyshev attempt
        # (hydrodynamics.cpp:292-299, counted so by the oracle); the device
        # reuses the product it formed for the FD test, one fewer per attempt.
        # (These chains fail before any upsilon, whose two products the
        # device forms as one.)
        assert g.matvecs == cnto.matvecs - cnto.cheb_attempts, (name, g.matvecs, cnto.matvecs)
    for f in fields:
        assert getattr(g, f) == getattr(cnto, f), (name, f, getattr(g, f), getattr(cnto, f))
    if rc == A.STATUS_NOT_POSITIVE_DEFINITE:
        # the fallback factor is FP64 in both forms: the oracle's pivot value
        assert g.fail_index == 55
        assert abs(g.fail_value - cnto.fail_value) <= 1e-9 * abs(cnto.fail_value)
    assert (rngs[0].ix, rngs[0].iy, rngs[0].seed) == (st.ix, st.iy, st.seed)
    assert status[1] == 0 and np.all(np.isfinite(out[1]))    # the batch went on


def test_singular_chebyshev_fallback_matches_reference(gpu, golden):
    """The reference's own outcome (golden): the fallback factor throws with
    pivot 55, after the step's generator draws."""
    c = FAIL["singular_cheb_fallback_not_pd"]
    p = H.Packed(c["phys"], 21)
    for prec in (A.PRECISION_FP64, A.PRECISION_MIXED):
        _, rngs, status, cnts, _ = H.gpu_batch(
            gpu, [(21, C.R0_SINGULAR, p.c, H.csim(c["sim"]), H.rng_of(c["seed"]))], precision=prec)
        assert int(golden["singular_cheb/rc"]) == 1 and status[0] == A.STATUS_NOT_POSITIVE_DEFINITE
        assert "pivot 55" in str(golden["singular_cheb/message"]) and cnts[0].fail_index == 55
        assert cnts[0].fallbacks == 1
        assert (rngs[0].ix, rngs[0].iy, rngs[0].seed) == tuple(golden["singular_cheb/rng"])


def test_library_rejects_clocks_that_would_not_end(gpu):
    """Entry points refuse, before any launch, a segment whose device step loop
    would not finish (BDSIM_EINVAL): dt not positive or not finite, non-finite
    ends, dt below the clock's resolution, more than 1e9 steps."""
    import ctypes
    from ctypes import byref
    p = H.Packed(C.phys_dict(spring=2, sqrtb=10.0), 8)
    R = np.ascontiguousarray(C.random_walk(8, 1))
    bad = [A.CSim(0.0, 1.0, 0.0, 1e-3, 1), A.CSim(0.0, 1.0, -1e-3, 1e-3, 1),
           A.CSim(0.0, 1.0, float("nan"), 1e-3, 1), A.CSim(0.0, float("inf"), 1e-3, 1e-3, 1),
           A.CSim(1e12, 1e12 + 1.0, 1e-6, 1e-3, 1), A.CSim(0.0, 1.0, 1e-10, 1e-3, 1)]
    for sim in bad:
        Rc = R.copy()
        rng = (A.CRng * 1)(H.rng_of(1))
        st = np.zeros(1, dtype=np.int32)
        rc = gpu.bdsim_integrate_batch(1, A.iptr(np.array([8], dtype=np.int32)), A.dptr(Rc),
                                       (A.CPhys * 1)(p.c), (A.CSim * 1)(sim), rng,
                                       A.PRECISION_FP64, A.iptr(st), None, None, 0)
        err = gpu.bdsim_last_error()
        assert rc == -1 and any(w in err for w in (b"dt", b"time", b"segment")), (sim.dt, rc, err)
        assert np.array_equal(Rc, R)
    b = ctypes.c_void_p()
    A.check(gpu, gpu.bdsim_batch_create(byref(b), 1, 8, A.dptr(R), byref(p.c), None,
                                        (A.CRng * 1)(H.rng_of(1)), A.PRECISION_FP64))
    try:
        for sim in bad:
            assert gpu.bdsim_batch_run(b, byref(sim)) == -1
    finally:
        gpu.bdsim_batch_destroy(b)
