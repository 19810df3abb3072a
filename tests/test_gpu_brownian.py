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
