"""Constant velocity gradients of the standard flows, as Flow objects.

The integrator reads the flow as kappa with the solvent velocity at a bead
position r equal to kappa . r (the same names and tensors as
python/bdsim/flows.py, which callers of the reference use to build
PhysParams.flow).
"""
import numpy as np

from .core import Flow


def _gradient(entries) -> Flow:
    """A constant Flow whose kappa has the given {(row, col): value} entries."""
    K = np.zeros((3, 3))
    for (i, j), v in entries.items():
        K[i, j] = v
    return Flow.constant(K)


def equilibrium() -> Flow:
    """A solvent at rest: every entry of kappa is zero."""
    return Flow()


def shear(rate: float) -> Flow:
    """Steady simple shear of the given rate: the x velocity grows along y
    (kappa_xy = rate)."""
    return _gradient({(0, 1): rate})


def uniaxial_extension(rate: float) -> Flow:
    """Uniaxial elongation along x at the given rate, the two transverse axes
    contracting at half of it so that kappa is traceless."""
    return _gradient({(0, 0): rate, (1, 1): -0.5 * rate, (2, 2): -0.5 * rate})


def planar_extension(rate: float) -> Flow:
    """Planar elongation: stretching along x and compression along y at the
    same rate, no motion along z."""
    return _gradient({(0, 0): rate, (1, 1): -rate})
