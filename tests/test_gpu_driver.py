"""gpu_map on the GPU: per-trajectory Python functions whose integrate calls are
batched into shared launches give, call by call, the FP64 oracle's positions
(1e-10) and generator states; a function that raises propagates."""
import numpy as np
import pytest

import bdsim_amd as bd
import helpers as H
from bdsim_amd import _abi as A
from golden import cases as C

pytestmark = pytest.mark.gpu


def _phys(N):
    p = bd.PhysParams()
    p.number_of_beads = N
    p.spring.type = bd.Spring.FENE
    p.spring.sqrtb = 10.0
    p.hstar = 0.25
    p.ev.type = bd.EV.Gauss
    p.ev.zstar, p.ev.dstar = 0.5, 0.5
    p.flow = bd.flows.shear(0.2)
    return p


def _segments(i):
    """Trajectory i's segments: 2..4 calls of different lengths on one clock."""
    out, t = [], 0.0
    for k in range(2 + i % 3):
        s = bd.SimParams()
        s.time_start, s.dt = t, 1e-3
        s.time_end = t + (4 + 5 * k + i % 4) * 1e-3
        out.append(s)
        t = s.time_end + s.dt
    return out


def _trajectory(i):
    """An opaque per-trajectory function: integrate segment by segment, and
    between calls decide in Python (stretch the next segment's start when the
    chain is compact)."""
    N = 12 + (i % 3) * 20
    phys = _phys(N)
    R = C.random_walk(N, 500 + i, 0.9)
    rng = bd.Rng(i)
    calls = []
    for s in _segments(i):
        if np.sum((R[-1] - R[0]) ** 2) < 1.0:
            R = R * 1.01
        R_in = R.copy()
        st_in = rng.__getstate__()
        R = bd.integrate(R, phys, s, rng)
        calls.append((N, R_in, st_in, s, R.copy(), rng.__getstate__()))
    return calls


@pytest.mark.parametrize("n_workers", [None, 5])
def test_gpu_map_calls_match_oracle(gpu, orc, n_workers):
    items = list(range(16))
    results = bd.gpu_map(_trajectory, items, n_workers=n_workers)
    n_calls = 0
    for calls in results:
        for N, R_in, st_in, s, R_out, st_out in calls:
            pk = A.PackedPhys(_phys(N), N)
            st = A.CRng(*st_in)
            want, rc, _, _ = H.oracle_integrate(orc, N, R_in, pk.c, A.pack_sim(s), st)
            assert rc == 0
            err = np.max(np.abs(R_out - want)) / np.max(np.abs(want))
            assert err <= 1e-10, err
            assert st_out == (st.ix, st.iy, st.seed)
            n_calls += 1
    assert n_calls == sum(len(_segments(i)) for i in items)


def test_gpu_map_mixed_form_and_errors(gpu):
    def fn(i):
        if i == 5:
            raise ValueError("planted")
        R = bd.integrate(C.random_walk(20, i), _phys(20), _segments(0)[0], bd.Rng(i))
        return R

    with pytest.raises(ValueError, match="planted"):
        bd.gpu_map(fn, range(8))
    a = bd.gpu_map(fn, [0, 1, 2], precision=A.PRECISION_MIXED)
    b = bd.gpu_map(fn, [0, 1, 2])
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) / np.max(np.abs(y)) < 1e-5
