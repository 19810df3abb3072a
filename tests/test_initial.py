"""The initial-configuration samplers against the distributions they draw from.

They are pinned the way the reference pins its own samplers (python/bdsim/
initial.py is tested by its distributions, not by fixed numbers): support,
Kolmogorov-Smirnov tests against the densities, isotropy, and the exact bond
lengths and angles of assembled chains.
"""
import numpy as np
import pytest
from scipy import integrate, stats

import bdsim_amd as bd
from bdsim_amd import initial


def _bonds(R):
    return np.diff(R, axis=-2)


@pytest.mark.parametrize("s,dQ", [(1.0, 2.0), (3.0, 0.5), (0.0, 10.0)])
def test_ff_bond_lengths_follow_psi(s, dQ):
    """Support [max(0, s-dQ), s+dQ] and a KS test against
    psi(Q) ~ Q^2 (1-(Q-s)^2/dQ^2)^(dQ^2/2) (python/bdsim/initial.py:29-41)."""
    Q = initial.ff_bond_lengths(4000, s, dQ, np.random.default_rng(1))
    lo, hi = max(0.0, s - dQ), s + dQ
    assert Q.min() >= lo and Q.max() <= hi
    psi = lambda q: q * q * max(0.0, 1 - (q - s) ** 2 / dQ ** 2) ** (dQ ** 2 / 2)  # noqa: E731
    Z = integrate.quad(psi, lo, hi, limit=200)[0]
    grid = np.linspace(lo, hi, 1001)
    cdf = np.array([integrate.quad(psi, lo, x, limit=200)[0] / Z for x in grid])
    assert stats.kstest(Q, lambda x: np.interp(x, grid, cdf)).pvalue > 1e-3


def test_fene_fraenkel_chain_bonds_isotropic():
    R = np.array([bd.fene_fraenkel_chain(201, 1.0, 2.0, seed=k) for k in range(20)])
    assert np.all(R[:, 0] == 0.0)
    b = _bonds(R).reshape(-1, 3)
    u = b / np.linalg.norm(b, axis=1)[:, None]
    # each component of an isotropic unit vector is uniform on [-1, 1]
    for d in range(3):
        assert stats.kstest(u[:, d], "uniform", args=(-1.0, 2.0)).pvalue > 1e-3
    assert np.allclose(u.T @ u / len(u), np.eye(3) / 3.0, atol=0.02)
    assert np.array_equal(R[3], bd.fene_fraenkel_chain(201, 1.0, 2.0, seed=3))   # seeded


def test_aligned_chain_along_x():
    R = bd.fene_fraenkel_chain_aligned_x(50, 1.0, 0.5, seed=2)
    b = _bonds(R)
    assert np.all(b[:, 1:] == 0.0) and np.all((b[:, 0] >= 0.5) & (b[:, 0] <= 1.5))


def test_equilibrium_chain_lengths_and_angles_exact():
    rs = np.random.default_rng(4)
    L = rs.uniform(0.5, 2.0, 40)
    th = rs.uniform(0.0, np.pi, 39)
    R = bd.equilibrium_chain(L, th, seed=5)
    b = _bonds(R)
    assert np.allclose(np.linalg.norm(b, axis=1), L, rtol=1e-13)
    u = b / np.linalg.norm(b, axis=1)[:, None]
    cos = np.sum(u[1:] * u[:-1], axis=1)
    assert np.allclose(cos, np.cos(th), atol=1e-12)
    # no angles: every bond perpendicular to the previous one
    R0 = bd.equilibrium_chain(np.ones(10), None, seed=1)
    u0 = _bonds(R0)
    assert np.allclose(np.sum(u0[1:] * u0[:-1], axis=1), 0.0, atol=1e-12)


@pytest.mark.parametrize("C", [0.0, 2.0, 50.0])
def test_bending_chain_angles_follow_the_boltzmann_weight(C):
    """cos(theta) between successive bonds has density ~ exp(C u) on [-1, 1]:
    CDF (e^{C(u+1)} - 1) / (e^{2C} - 1)."""
    R = np.array([bd.fene_fraenkel_bending_chain(101, 1.0, 0.5, C, seed=k) for k in range(30)])
    b = _bonds(R)
    u = b / np.linalg.norm(b, axis=-1)[..., None]
    c = np.sum(u[:, 1:] * u[:, :-1], axis=-1).ravel()
    if C == 0.0:
        cdf = lambda x: (np.clip(x, -1, 1) + 1.0) / 2.0  # noqa: E731
    else:
        cdf = lambda x: np.expm1(C * (np.clip(x, -1, 1) + 1.0)) / np.expm1(2.0 * C)  # noqa: E731
    assert stats.kstest(c, cdf).pvalue > 1e-3
    L = np.linalg.norm(b, axis=-1)
    assert L.min() >= 0.5 and L.max() <= 1.5


def test_gaussian_chain_centred_with_gaussian_bonds():
    g = bd.gaussian_chain(50, seed=3)
    assert g.shape == (50, 3) and np.allclose(g.mean(axis=0), 0.0, atol=1e-12)
    b = _bonds(np.array([bd.gaussian_chain(200, 0.7, seed=k) for k in range(10)])).ravel()
    assert stats.kstest(b, "norm", args=(0.0, 0.7)).pvalue > 1e-3
