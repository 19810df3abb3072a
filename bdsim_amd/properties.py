"""Per-configuration chain properties of (N, 3) positions (python/bdsim/properties.py)."""
import numpy as np


def end_to_end_sq(R: np.ndarray) -> float:
    d = R[-1] - R[0]
    return float(d @ d)


def gyration_tensor(R: np.ndarray) -> np.ndarray:
    Rc = R - R.mean(axis=0)
    return (Rc.T @ Rc) / len(R)


def radius_of_gyration_sq(R: np.ndarray) -> float:
    return float(np.trace(gyration_tensor(R)))


def stretch(R: np.ndarray) -> np.ndarray:
    Rc = R - R.mean(axis=0)
    return Rc.max(axis=0) - Rc.min(axis=0)


def kramers_stress(R: np.ndarray, F: np.ndarray) -> np.ndarray:
    """Kramers-Kirkwood stress tau_ij = sum_mu (R_mu - R_c)_i F_mu,j."""
    Rc = R - R.mean(axis=0)
    return Rc.T @ F


def shear_viscosity(R: np.ndarray, F: np.ndarray, rate: float) -> float:
    return -kramers_stress(R, F)[0, 1] / rate


def normal_stress_differences(R: np.ndarray, F: np.ndarray):
    s = kramers_stress(R, F)
    return s[0, 0] - s[1, 1], s[1, 1] - s[2, 2]
