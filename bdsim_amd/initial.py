"""Initial chain configurations: host-side inputs to the step.

The distributions are those of python/bdsim/initial.py: a Gaussian random walk;
FENE-Fraenkel equilibrium bonds, lengths with density
psi(Q) ~ Q^2 (1 - (Q - sigma)^2 / dQ^2)^(dQ^2 / 2) on |Q - sigma| <= dQ and
isotropic (or x-aligned) directions; chains whose successive bonds meet at
angles theta with cos(theta) distributed as exp(C cos(theta)) on [-1, 1] (the
equilibrium of the bending potential C (1 - cos(theta))). They are drawn here by
methods of this module's own: bond lengths by rejection under a flat envelope
(no grid interpolation), directions as normalised Gaussian vectors, and each
bond placed in an orthonormal frame built around the previous one. Draws are
seeded and reproducible, but they are not the reference's numbers; the tests
pin the distributions (support, Kolmogorov-Smirnov against the densities,
isotropy, exact lengths and angles of assembled chains), which is also all that
the reference's own tests pin.
"""
import numpy as np


def gaussian_chain(n: int, bond_std: float = 1.0, seed=None) -> np.ndarray:
    """`n` beads joined by independent isotropic Gaussian bonds (each component
    of standard deviation `bond_std`), centre of mass at the origin."""
    rng = np.random.default_rng(seed)
    steps = rng.normal(0.0, bond_std, size=(max(n - 1, 0), 3))
    R = np.vstack([np.zeros((1, 3)), np.cumsum(steps, axis=0)])[:n]
    return R - R.mean(axis=0)


def _ff_log_density(Q, sigma, dQ):
    """log psi(Q) of the FENE-Fraenkel bond length, -inf outside the support."""
    x = (Q - sigma) / dQ
    with np.errstate(divide="ignore", invalid="ignore"):
        out = 2.0 * np.log(Q) + 0.5 * dQ * dQ * np.log1p(-x * x)
    return np.where((Q > 0.0) & (np.abs(x) < 1.0), out, -np.inf)


def ff_bond_lengths(n_bonds, sigma, dQ, rng):
    """FENE-Fraenkel bond lengths by rejection: candidates uniform on the
    support, kept with probability psi(Q) / M, where M bounds psi (its maximum
    located on a fine grid and refined around the best point, plus 2 %)."""
    lo, hi = max(0.0, sigma - dQ), sigma + dQ
    grid = np.linspace(lo, hi, 8193)
    lp = _ff_log_density(grid, sigma, dQ)
    k = int(np.argmax(lp))
    fine = np.linspace(grid[max(k - 1, 0)], grid[min(k + 1, len(grid) - 1)], 513)
    log_m = max(lp[k], float(np.max(_ff_log_density(fine, sigma, dQ)))) + np.log(1.02)
    out = np.empty(n_bonds)
    got = 0
    while got < n_bonds:
        want = n_bonds - got
        Q = rng.uniform(lo, hi, size=2 * want + 16)
        keep = Q[np.log(rng.random(Q.size)) < _ff_log_density(Q, sigma, dQ) - log_m]
        take = keep[:want]
        out[got:got + take.size] = take
        got += take.size
    return out


def unit_vectors(n, rng):
    """`n` isotropic unit vectors: standard normal 3-vectors, normalised (the
    normal density is rotation invariant)."""
    g = rng.standard_normal((n, 3))
    norm = np.linalg.norm(g, axis=1)
    while np.any(norm < 1e-12):            # probability zero, kept exact anyway
        bad = norm < 1e-12
        g[bad] = rng.standard_normal((int(bad.sum()), 3))
        norm = np.linalg.norm(g, axis=1)
    return g / norm[:, None]


def fene_fraenkel_chain(n: int, sigma: float, dQ: float, seed=None) -> np.ndarray:
    """FENE-Fraenkel equilibrium chain with isotropic bonds, bead 0 at the origin."""
    rng = np.random.default_rng(seed)
    bonds = unit_vectors(n - 1, rng) * ff_bond_lengths(n - 1, sigma, dQ, rng)[:, None]
    return np.vstack([np.zeros((1, 3)), np.cumsum(bonds, axis=0)])


def fene_fraenkel_chain_aligned_x(n: int, sigma: float, dQ: float, seed=None) -> np.ndarray:
    """FENE-Fraenkel equilibrium bond lengths, every bond along +x, bead 0 at
    the origin."""
    rng = np.random.default_rng(seed)
    R = np.zeros((n, 3))
    R[1:, 0] = np.cumsum(ff_bond_lengths(n - 1, sigma, dQ, rng))
    return R


def bending_cosines(n, stiffness, rng):
    """cos(theta) of `n` bond angles with density ~ exp(C u) on u in [-1, 1]:
    the inverse of F(u) = (e^{C(u+1)} - 1) / (e^{2C} - 1), written as
    u = 1 + log(Y + (1 - Y) e^{-2C}) / C so that large C does not overflow."""
    C = float(stiffness)
    Y = rng.random(n)
    if C == 0.0:
        return 2.0 * Y - 1.0
    u = 1.0 + np.log(Y + (1.0 - Y) * np.exp(-2.0 * C)) / C
    return np.clip(u, -1.0, 1.0)


def _frame(u):
    """Two unit vectors completing u to a right-handed orthonormal frame."""
    axis = np.zeros(3)
    axis[int(np.argmin(np.abs(u)))] = 1.0  # the coordinate axis least along u
    e1 = np.cross(u, axis)
    e1 /= np.linalg.norm(e1)
    return e1, np.cross(u, e1)


def _assemble(lengths, cos_theta, rng):
    """Bonds of the given lengths; bond i > 0 makes angle arccos(cos_theta[i-1])
    with bond i - 1 at a uniform azimuth; bond 0 isotropic. Bead 0 at the origin."""
    nb = len(lengths)
    u = np.empty((nb, 3))
    if nb:
        u[0] = unit_vectors(1, rng)[0]
    for i in range(1, nb):
        c = float(cos_theta[i - 1])
        s = np.sqrt(max(0.0, 1.0 - c * c))
        e1, e2 = _frame(u[i - 1])
        phi = rng.uniform(0.0, 2.0 * np.pi)
        v = c * u[i - 1] + s * (np.cos(phi) * e1 + np.sin(phi) * e2)
        u[i] = v / np.linalg.norm(v)
    return np.vstack([np.zeros((1, 3)), np.cumsum(u * lengths[:, None], axis=0)])


def equilibrium_chain(lengths, angles=None, seed=None):
    """A chain of len(lengths) + 1 beads from its bond lengths and the angles
    between successive bonds (len(lengths) - 1 of them, radians; None: every
    bond perpendicular to the one before, as in the reference). Azimuths and
    the first bond's direction are uniform; bead 0 at the origin."""
    lengths = np.asarray(lengths, dtype=float)
    nb = len(lengths)
    cos_theta = np.zeros(max(nb - 1, 0)) if angles is None else np.cos(np.asarray(angles, dtype=float))
    return _assemble(lengths, cos_theta, np.random.default_rng(seed))


def fene_fraenkel_bending_chain(n, sigma, dQ, stiffness, seed=None):
    """Equilibrium chain of n beads: FENE-Fraenkel bond lengths and bond angles
    of the bending potential C (1 - cos(theta)), C = stiffness."""
    rng = np.random.default_rng(seed)
    lengths = ff_bond_lengths(n - 1, sigma, dQ, rng)
    cos_theta = bending_cosines(max(n - 2, 0), stiffness, rng)
    return _assemble(lengths, cos_theta, rng)
