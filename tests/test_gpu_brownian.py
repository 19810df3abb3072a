"""The Brownian operator B(R) x of each method (mobility.cpp:58-88) through
bdsim_brownian_batch, against the oracle: free draining (identity), the
Cholesky factor L x (mobility.cpp:39-51, bo_cholesky_lower), and the
Chebyshev series of sqrt(D) x on cold Lanczos bounds (hydrodynamics.cpp:
317-337, bo_chebyshev_sqrt_times with an empty cache). FP64 to 1e-12
relative; the mixed form to 1e-5 against the same FP64 oracle."""
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from golden import cases as C

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _oracle_bx(orc, N, R, x, d):
    n = 3 * N
    if d["hstar"] <= 0.0:
        return x.copy(), 0
    D = np.empty((n, n))
    orc.bo_rpy_matrix(N, A.dptr(R), d["hstar"], A.dptr(D))
    y = np.empty(n)
    if d["method"] == 1:
        L = np.empty((n, n))
        assert orc.bo_cholesky_lower(n, A.dptr(D), A.dptr(L), None, None) == 0
        orc.bo_lower_matvec(n, A.dptr(L), A.dptr(x), A.dptr(y))
        return y.reshape(N, 3), 0
    L, fd = c_int(0), c_double(0.0)
    orc.bo_chebyshev_sqrt_times(n, A.dptr(D), A.dptr(x), d["mult"], d["fd"], None, None,
                                A.dptr(y), byref(L), byref(fd))
    assert fd.value <= d["fd"]
    return y.reshape(N, 3), L.value


def _gpu_bx(gpu, N, R, x, d, precision):
    n_traj = R.shape[0]
    p = H.Packed(d, N)
    y = np.empty_like(x)
    st = np.full(n_traj, -1, dtype=np.int32)
    A.check(gpu, gpu.bdsim_brownian_batch(n_traj, N, A.dptr(R), byref(p.c), precision,
                                          A.dptr(x), A.dptr(y), A.iptr(st)))
    return y, st


CASES = [  # (N, hstar, method)
    (2, 0.3, 0), (10, 0.3, 0), (50, 0.3, 0), (64, 0.2, 0), (100, 0.25, 0), (300, 0.3, 0),
    (600, 0.2, 0), (5, 0.25, 1), (21, 0.25, 1), (40, 0.25, 1), (50, 0.0, 0),
]


@pytest.mark.parametrize("N,h,method", CASES)
def test_brownian_operator(gpu, orc, N, h, method):
    n_traj = 3
    d = C.phys_dict(hstar=h, method=method)
    R = np.ascontiguousarray(np.stack([C.random_walk(N, 70 + N + k, 0.9) for k in range(n_traj)]))
    x = np.ascontiguousarray(np.random.default_rng(N + 5 * method).normal(size=(n_traj, N, 3)))
    want = [_oracle_bx(orc, N, R[k], x[k].ravel(), d)[0] for k in range(n_traj)]
    y, st = _gpu_bx(gpu, N, R, x, d, A.PRECISION_FP64)
    assert np.all(st == A.STATUS_OK), st
    for k in range(n_traj):
        assert rel(y[k], want[k]) <= 1e-12, (k, rel(y[k], want[k]))
    ym, st = _gpu_bx(gpu, N, R, x, d, A.PRECISION_MIXED)
    assert np.all(st == A.STATUS_OK), st
    for k in range(n_traj):
        assert rel(ym[k], want[k]) < 1e-5, (k, rel(ym[k], want[k]))


def test_chebyshev_form_is_the_matrix_square_root(gpu, orc):
    """With a long series (multiplier 2, tolerance 1e-8) the Chebyshev form
    is sqrt(D) x of the eigendecomposition D = V diag(w) V^T to 1e-7 (the
    oracle's own series lies 5.6e-9 from it on this chain), and
    |B x|^2 = x.D.x to the tolerance."""
    N = 50
    d = C.phys_dict(hstar=0.3, method=0, mult=2.0, fd=1e-8)
    R = np.ascontiguousarray(C.random_walk(N, 9, 0.9)[None])
    x = np.ascontiguousarray(np.random.default_rng(3).normal(size=(1, N, 3)))
    D = np.empty((3 * N, 3 * N))
    orc.bo_rpy_matrix(N, A.dptr(R[0]), 0.3, A.dptr(D))
    w, V = np.linalg.eigh(D)
    exact = V @ (np.sqrt(w) * (V.T @ x.ravel()))
    bx, st = _gpu_bx(gpu, N, R, x, d, A.PRECISION_FP64)
    assert st[0] == A.STATUS_OK
    assert rel(bx, exact) < 1e-7, rel(bx, exact)
    xDx = float(x.ravel() @ D @ x.ravel())
    assert abs(float(bx.ravel() @ bx.ravel()) - xDx) / xDx <= 1e-8


def test_cholesky_form_is_a_factor(gpu, orc):
    """B applied to the 3N basis vectors (one chain each, same positions)
    gives the columns of a lower-triangular Y with positive diagonal and
    Y Y^T = D: the Cholesky factor, which is unique."""
    N = 10
    n = 3 * N
    d = C.phys_dict(hstar=0.25, method=1)
    R1 = C.random_walk(N, 4, 0.9)
    R = np.ascontiguousarray(np.repeat(R1[None], n, axis=0))
    E = np.ascontiguousarray(np.eye(n).reshape(n, N, 3))
    Y, st = _gpu_bx(gpu, N, R, E, d, A.PRECISION_FP64)
    assert np.all(st == A.STATUS_OK)
    Y = Y.reshape(n, n).T  # column j = B e_j
    D = np.empty((n, n))
    orc.bo_rpy_matrix(N, A.dptr(np.ascontiguousarray(R1)), 0.25, A.dptr(D))
    assert np.all(np.triu(Y, 1) == 0.0)
    assert np.all(np.diag(Y) > 0.0)
    assert np.max(np.abs(Y @ Y.T - D)) <= 1e-14 * np.max(np.abs(D)) * n


def test_mixed_step_errors_against_the_oracle(gpu, orc):
    """The mixed form's per-step errors on headline physics (N=50, FENE b=100,
    h*=0.3 Chebyshev, SDK EV, shear Wi 0.1..100), each against the FP64
    oracle: D.F at configurations reached after 100 steps (F the total
    force), B.dW there (standard normal dW, cold bounds), and the positions
    after those 100 steps. Bar 1e-5 each; the values are printed."""
    N, steps = 50, 100
    errs = {"D.F": [], "B.dW": [], "positions": []}
    for i, wi in enumerate([0.1, 1.0, 10.0, 100.0]):
        d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3, ev=3, dstar=1.0, kappa=C.wi_kappa(wi))
        p = H.Packed(d, N)
        R0 = C.random_walk(N, 2000 + i)
        sim = H.csim(C.sim_dict(0.0, (steps - 1) * 1e-3, 1e-3, 1e-3))
        want, rc, _, _ = H.oracle_integrate(orc, N, R0, p.c, sim, H.rng_of(i))
        assert rc == 0
        out, _, status, _, _ = H.gpu_batch(gpu, [(N, R0, p.c, sim, H.rng_of(i))],
                                           precision=A.PRECISION_MIXED)
        assert status[0] == 0
        errs["positions"].append(rel(out[0], want))
        R = np.ascontiguousarray(want)
        F = np.empty_like(R)
        orc.bo_total_force(N, A.dptr(R), byref(p.c), A.dptr(F))
        D = np.empty((3 * N, 3 * N))
        orc.bo_rpy_matrix(N, A.dptr(R), 0.3, A.dptr(D))
        DF = np.empty_like(R)
        A.check(gpu, gpu.bdsim_rpy_apply_batch(1, N, 0.3, A.PRECISION_MIXED, A.dptr(R), A.dptr(F),
                                               A.dptr(DF)))
        errs["D.F"].append(rel(DF, D @ F.ravel()))
        dW = np.ascontiguousarray(np.random.default_rng(i).normal(size=(1, N, 3)))
        want_bw = _oracle_bx(orc, N, R, dW.ravel(), d)[0]
        bw, st = _gpu_bx(gpu, N, np.ascontiguousarray(R[None]), dW, d, A.PRECISION_MIXED)
        assert st[0] == A.STATUS_OK
        errs["B.dW"].append(rel(bw, want_bw))
    print({k: float(max(v)) for k, v in errs.items()})
    for k, v in errs.items():
        assert max(v) < 1e-5, (k, v)
