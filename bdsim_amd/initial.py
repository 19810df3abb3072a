"""Initial chain configurations: host-side inputs to the step.

The procedures of python/bdsim/initial.py:5-173 (Gaussian random walk;
FENE-Fraenkel equilibrium bond lengths by inverse-CDF sampling with isotropic
or x-aligned directions; bending-correlated equilibrium chains), written with
numpy's seeded generator. Their outputs are pinned distributionally by the
tests (support, KS against psi(Q)), not bit-for-bit: the reference pins them
only that way either.
"""
import numpy as np


def gaussian_chain(n: int, bond_std: float = 1.0, seed=None) -> np.ndarray:
    """Random walk of `n` beads with isotropic Gaussian bonds, centred."""
    rng = np.random.default_rng(seed)
    R = np.zeros((n, 3))
    R[1:] = np.cumsum(rng.normal(scale=bond_std, size=(n - 1, 3)), axis=0)
    return R - R.mean(axis=0)


def inverse_cdf_sample(x, pdf, u):
    """Values distributed as `pdf` on grid `x` from uniforms `u` (trapezoid CDF)."""
    cdf = np.concatenate(([0.0], np.cumsum(0.5 * (pdf[:-1] + pdf[1:]) * np.diff(x))))
    cdf /= cdf[-1]
    return np.interp(u, cdf, x)


def fene_fraenkel_bond_lengths(n_bonds, sigma, dQ, rng, nsteps=10000):
    """psi(Q) ~ Q^2 (1 - (Q-sigma)^2/dQ^2)^(dQ^2/2) on |Q - sigma| <= dQ."""
    alpha = dQ * dQ
    Q = np.linspace(max(0.0, sigma - dQ), sigma + dQ, nsteps)
    Q[0] += 1e-7
    Q[-1] -= 1e-7
    psi = Q * Q * (1.0 - (Q - sigma) ** 2 / alpha) ** (alpha / 2.0)
    return inverse_cdf_sample(Q, psi, rng.random(n_bonds))


def spherical_unit_vectors(n, rng):
    """`n` isotropic unit vectors (Marsaglia's method)."""
    out = np.empty((n, 3))
    filled = 0
    while filled < n:
        m = 2 * (n - filled) + 8
        x = rng.uniform(-1.0, 1.0, size=(m, 2))
        keep = x[(x[:, 0] ** 2 + x[:, 1] ** 2) < 1.0]
        take = min(len(keep), n - filled)
        s = keep[:take, 0] ** 2 + keep[:take, 1] ** 2
        root = np.sqrt(1.0 - s)
        out[filled:filled + take] = np.stack([2.0 * keep[:take, 0] * root,
                                              2.0 * keep[:take, 1] * root, 1.0 - 2.0 * s], axis=1)
        filled += take
    return out


def fene_fraenkel_chain(n: int, sigma: float, dQ: float, seed=None) -> np.ndarray:
    """FENE-Fraenkel equilibrium chain with isotropic bonds, bead 0 at the origin."""
    rng = np.random.default_rng(seed)
    Ql = fene_fraenkel_bond_lengths(n - 1, sigma, dQ, rng)
    R = np.zeros((n, 3))
    R[1:] = np.cumsum(spherical_unit_vectors(n - 1, rng) * Ql[:, None], axis=0)
    return R


def fene_fraenkel_chain_aligned_x(n: int, sigma: float, dQ: float, seed=None) -> np.ndarray:
    """FENE-Fraenkel equilibrium bond lengths, all bonds along x."""
    rng = np.random.default_rng(seed)
    R = np.zeros((n, 3))
    R[1:, 0] = np.cumsum(fene_fraenkel_bond_lengths(n - 1, sigma, dQ, rng))
    return R


def one_minus_cos_theta_angles(n, stiffness, rng):
    """Bond angles of U = C(1 - cos theta): cos theta has density ~ exp(C u)."""
    C = stiffness
    if C == 0.0:
        return np.arccos(rng.uniform(-1.0, 1.0, n))
    Y = rng.random(n)
    u = np.log(np.exp(-C) + Y * (np.exp(C) - np.exp(-C))) / C
    return np.arccos(np.clip(u, -1.0, 1.0))


def isotropic_directions(n, rng):
    cos_t = rng.uniform(-1.0, 1.0, n)
    sin_t = np.sqrt(1.0 - cos_t ** 2)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    return np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], axis=1)


def _rotation_z_to(u):
    c = u[2]
    if c > 1.0 - 1e-12:
        return np.eye(3)
    if c < -1.0 + 1e-12:
        return np.diag([1.0, -1.0, -1.0])
    v = np.array([-u[1], u[0], 0.0])
    vx = np.array([[0.0, -v[2], v[1]], [v[2], 0.0, -v[0]], [-v[1], v[0], 0.0]])
    return np.eye(3) + vx + vx @ vx * ((1.0 - c) / (v @ v))


def equilibrium_chain(lengths, angles=None, seed=None):
    """Chain from bond lengths and included angles (None: isotropic)."""
    lengths = np.asarray(lengths, dtype=float)
    nb = len(lengths)
    rng = np.random.default_rng(seed)
    u = np.empty((nb, 3))
    u[0] = isotropic_directions(1, rng)[0]
    for i in range(1, nb):
        theta = np.pi / 2 if angles is None else float(angles[i - 1])
        az = rng.uniform(0.0, 2.0 * np.pi)
        base = np.array([np.sin(theta) * np.cos(az), np.sin(theta) * np.sin(az), np.cos(theta)])
        u[i] = _rotation_z_to(u[i - 1]) @ base
    R = np.zeros((nb + 1, 3))
    R[1:] = np.cumsum(u * lengths[:, None], axis=0)
    return R


def fene_fraenkel_bending_chain(n, sigma, dQ, stiffness, seed=None):
    rng = np.random.default_rng(seed)
    lengths = fene_fraenkel_bond_lengths(n - 1, sigma, dQ, rng)
    angles = one_minus_cos_theta_angles(n - 2, stiffness, rng) if n > 2 else None
    return equilibrium_chain(lengths, angles, seed=seed)
