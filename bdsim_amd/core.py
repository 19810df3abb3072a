"""The `_bdsim` surface of bdsim, backed by the MI355X library.

Same names, attributes, defaults and pickle tuples as the reference's nanobind
module (bindings/module.cpp:167-317), so code written against `bdsim` runs
unchanged. `integrate` and the force functions execute on the GPU through the
C ABI of include/bdsim_gpu.h; there is no CPU path in this module.
"""
import ctypes
import enum
import sys
import threading

import numpy as np

from . import _abi

# --------------------------------------------------------------------------
# enums (src/types.hpp:27-53, bindings/module.cpp:170-182)
# --------------------------------------------------------------------------


class Spring(enum.IntEnum):
    Hook = 1
    FENE = 2
    ILC = 3
    WLC = 4
    Fraenkel = 5
    FENEFraenkel = 6
    WLCbounded = 7


class EV(enum.IntEnum):
    None_ = 0
    Gauss = 1
    LJ = 2
    SDK = 3
    SDKstickers = 4


class Bending(enum.IntEnum):
    None_ = 0
    OneMinusCosTheta = 1


class DelSMethod(enum.IntEnum):
    Chebyshev = 0
    Cholesky = 1
    ExactSqrt = 2


# --------------------------------------------------------------------------
# parameter types (bindings/module.cpp:184-306)
# --------------------------------------------------------------------------


class SpringParams:
    """src/spring.hpp:12-19"""

    def __init__(self, type=Spring.Hook, sqrtb=10.0, natural_length=0.0):
        self.type = Spring(type)
        self.sqrtb = float(sqrtb)
        self.natural_length = float(natural_length)

    def __getstate__(self):
        return (int(self.type), self.sqrtb, self.natural_length)

    def __setstate__(self, s):
        self.__init__(Spring(s[0]), s[1], s[2])

    def __repr__(self):
        return f"SpringParams(type={self.type!s}, sqrtb={self.sqrtb}, natural_length={self.natural_length})"


class EVParams:
    """src/excluded_volume.hpp:17-25"""

    def __init__(self):
        self.type = EV.None_
        self.zstar = 0.0
        self.dstar = 0.0
        self.min_cutoff = 0.7
        self.max_cutoff = 1.5
        self.contour_dist_for_EV = 1
        self.phi = []

    def __getstate__(self):
        return (int(self.type), self.zstar, self.dstar, self.min_cutoff, self.max_cutoff,
                self.contour_dist_for_EV, list(self.phi))

    def __setstate__(self, s):
        self.__init__()
        (t, self.zstar, self.dstar, self.min_cutoff, self.max_cutoff,
         self.contour_dist_for_EV, phi) = s
        self.type = EV(t)
        self.phi = list(phi)


class BendingParams:
    """src/bending.hpp:9-12"""

    def __init__(self, type=Bending.None_, stiffness=0.0):
        self.type = Bending(type)
        self.stiffness = float(stiffness)

    def __getstate__(self):
        return (int(self.type), self.stiffness)

    def __setstate__(self, s):
        self.__init__(Bending(s[0]), s[1])


class Flow:
    """Velocity gradient kappa(t): constant, or an interpolated table
    (src/flow.hpp:17-43). Default: zero (equilibrium)."""

    def __init__(self):
        self._constant = True
        self._K0 = np.zeros((3, 3))
        self._times = []
        self._Ks = []

    @staticmethod
    def constant(kappa):
        f = Flow()
        f._K0 = np.array(kappa, dtype=np.float64).reshape(3, 3).copy()
        return f

    @staticmethod
    def time_varying(times, kappas):
        f = Flow()
        f._constant = False
        f._times = [float(t) for t in times]
        f._Ks = [np.array(k, dtype=np.float64).reshape(3, 3).copy() for k in kappas]
        return f

    def is_constant(self):
        return self._constant

    def constant_kappa(self):
        return self._K0

    def sample_times(self):
        return list(self._times)

    def sample_tensors(self):
        return [k for k in self._Ks]

    def __getstate__(self):   # bindings/module.cpp:143-151
        ks = []
        for K in self._Ks:
            ks.extend(float(x) for x in K.ravel())
        return (bool(self._constant), [float(x) for x in self._K0.ravel()],
                list(self._times), ks)

    def __setstate__(self, s):
        constant, k0, times, ks = s
        self.__init__()
        if constant:
            self._K0 = np.array(k0, dtype=np.float64).reshape(3, 3)
        else:
            self._constant = False
            self._times = list(times)
            self._Ks = [np.array(ks[9 * t:9 * t + 9], dtype=np.float64).reshape(3, 3)
                        for t in range(len(times))]


class _BeadForce:
    __slots__ = ("bead", "constant", "value", "times", "values")

    def __init__(self, bead, constant, value, times=(), values=()):
        self.bead = int(bead)
        self.constant = bool(constant)
        self.value = tuple(float(v) for v in value)
        self.times = [float(t) for t in times]
        self.values = [tuple(float(x) for x in v) for v in values]


class ExternalForce:
    """User-applied forces on individual beads (constant or time-varying).
    Negative bead indices count from the end, so -1 is the last bead.
    (src/external_force.hpp:31-57)"""

    def __init__(self):
        self._entries = []

    def add_constant(self, bead, force):
        """Apply a constant force vector to one bead."""
        self._entries.append(_BeadForce(bead, True, force))

    def add_time_varying(self, bead, times, forces):
        """Apply a force interpolated linearly between (time, vector) samples;
        held constant outside the tabulated range."""
        self._entries.append(_BeadForce(bead, False, (0.0, 0.0, 0.0), times, forces))

    def add_stretch(self, bead_minus, bead_plus, force):
        """Equal and opposite forces: +force on bead_plus, -force on bead_minus."""
        self.add_constant(bead_plus, force)
        self.add_constant(bead_minus, tuple(-float(f) for f in force))

    def empty(self):
        return not self._entries

    def entries(self):
        return list(self._entries)

    def __getstate__(self):   # bindings/module.cpp:84-99
        beads, is_const, consts, times, ntimes, values = [], [], [], [], [], []
        for e in self._entries:
            beads.append(e.bead)
            is_const.append(1 if e.constant else 0)
            consts.extend(e.value)
            ntimes.append(len(e.times))
            times.extend(e.times)
            for v in e.values:
                values.extend(v)
        return (beads, is_const, consts, times, ntimes, values)

    def __setstate__(self, s):
        beads, is_const, consts, times, ntimes, values = s
        self._entries = []
        ti = vi = 0
        for k, b in enumerate(beads):
            n = ntimes[k]
            e = _BeadForce(b, is_const[k] != 0, consts[3 * k:3 * k + 3],
                           times[ti:ti + n],
                           [tuple(values[vi + 3 * j:vi + 3 * j + 3]) for j in range(n)])
            ti += n
            vi += 3 * n
            self._entries.append(e)


class PhysParams:
    """src/model.hpp:20-35"""

    def __init__(self):
        self.spring = SpringParams()
        self.ev = EVParams()
        self.bend = BendingParams()
        self.flow = Flow()
        self.external = ExternalForce()
        self.number_of_beads = 10
        self.hstar = 0.0
        self.hi_method = DelSMethod.Chebyshev
        self.ncheb_multiplier = 1.0
        self.fd_err_max = 0.0025
        self.equilibration = False

    def __getstate__(self):   # bindings/module.cpp:270-274
        return (self.spring, self.ev, self.bend, self.flow, int(self.number_of_beads),
                float(self.hstar), int(self.hi_method), float(self.ncheb_multiplier),
                float(self.fd_err_max), bool(self.equilibration), self.external)

    def __setstate__(self, s):
        self.__init__()
        (self.spring, self.ev, self.bend, self.flow, self.number_of_beads, self.hstar,
         hm, self.ncheb_multiplier, self.fd_err_max, self.equilibration, self.external) = s
        self.hi_method = DelSMethod(hm)


class SimParams:
    """src/integrator.hpp:22-28"""

    def __init__(self):
        self.time_start = 0.0
        self.time_end = 1.0
        self.dt = 0.01
        self.implicit_loop_tol = 1e-3
        self.update_center_of_mass = True

    def __getstate__(self):
        return (float(self.time_start), float(self.time_end), float(self.dt),
                float(self.implicit_loop_tol), bool(self.update_center_of_mass))

    def __setstate__(self, s):
        self.__init__()
        (self.time_start, self.time_end, self.dt, self.implicit_loop_tol,
         self.update_center_of_mass) = s


class Rng:
    """Numerical Recipes ran_1 state (src/rng.hpp:25-63). The stream itself is
    drawn on the GPU; this object carries (ix, iy, seed) between calls exactly
    as the C++ Rng& does, so it pickles and deep-copies the same way."""

    def __init__(self, seed=123):
        self.reset(seed)

    def reset(self, seed):
        self._ix, self._iy, self._seed = 1, -1, int(seed)

    def __getstate__(self):
        return (self._ix, self._iy, self._seed)

    def __setstate__(self, s):
        self._ix, self._iy, self._seed = (int(x) for x in s)

    # state accessors used by the GPU backends
    def _state(self):
        return (self._ix, self._iy, self._seed)

    def _set_state(self, ix, iy, seed):
        self._ix, self._iy, self._seed = int(ix), int(iy), int(seed)

    def __repr__(self):
        return f"Rng(ix={self._ix}, iy={self._iy}, seed={self._seed})"


# --------------------------------------------------------------------------
# failure reporting, mirroring the reference's messages
# --------------------------------------------------------------------------

_nonconv_lock = threading.Lock()
_nonconvergence_count = 0


def _report_nonconvergence(count, N, tol):
    """integrator.cpp:27-43: throttled warnings, one per failed step."""
    global _nonconvergence_count
    with _nonconv_lock:
        for _ in range(int(count)):
            _nonconvergence_count += 1
            n = _nonconvergence_count
            show = n <= 3 or (n <= 1000 and n % 100 == 0) or n % 10000 == 0
            if not show:
                continue
            tail = (" -- the step was accepted anyway; this is where the trajectory "
                    "fails, anything later is a consequence") if n == 1 else ""
            print(f"bdsim warning: implicit corrector did not converge after {100 * N} "
                  f"iterations (tolerance {tol:.1e}){tail}", file=sys.stderr)
            if n == 3:
                print("bdsim warning: further non-convergence reports will be throttled",
                      file=sys.stderr)
            if n > 3:
                print(f"               ({n} occurrences so far)", file=sys.stderr)


def status_error(status, counters):
    """Exception for a per-trajectory failure status (None when OK)."""
    if status == _abi.STATUS_OK:
        return None
    if status == _abi.STATUS_NOT_POSITIVE_DEFINITE:
        return RuntimeError(
            f"cholesky_lower: matrix is not positive definite (pivot {counters.fail_index} = "
            f"{counters.fail_value:.6f}). For the RPY tensor this means the configuration is "
            "degenerate -- usually two beads at essentially the same position after a "
            "diverging step. Check for an earlier corrector non-convergence warning: that is "
            "where the trajectory actually failed.")
    if status == _abi.STATUS_RTSAFE_UNBRACKETED:
        return RuntimeError("root must be bracketed in rtsafe")
    if status == _abi.STATUS_RTSAFE_MAXIT:
        return RuntimeError("Max iterations exceeded in rtsafe")
    if status == _abi.STATUS_EXACT_SQRT:
        return RuntimeError("ExactSqrt square-root method not implemented")
    if status == _abi.STATUS_WLCB_SIGMA:
        return RuntimeError("sigma/L cannot be greater than 1 for WLC bounded")
    return RuntimeError(f"trajectory failed with status {status}")


# --------------------------------------------------------------------------
# entry points (bindings/module.cpp:45-76,128-132,308-316)
# --------------------------------------------------------------------------

def _positions(R):
    a = np.ascontiguousarray(np.asarray(R, dtype=np.float64))
    if a.ndim != 2 or a.shape[1] != 3:
        raise TypeError("R must be an (N, 3) float64 array")
    return a


def integrate(R, phys, sim, rng, precision=_abi.PRECISION_FP64):
    """Integrate a chain; returns the evolved (N,3) positions.

    bindings/module.cpp:45-50 on the GPU: `phys` is used with number_of_beads
    set to N (the caller's object is left unchanged), `rng` advances in place.
    """
    R = _positions(R)
    from . import batcher
    b = batcher.current()
    if b is not None:          # inside gpu_map: the calls of all items go out batched
        return b.integrate(R, phys, sim, rng)
    N = R.shape[0]
    lib = _abi.load_gpu()
    out = R.copy()
    pp = _abi.PackedPhys(phys, N)
    cs = _abi.pack_sim(sim)
    st = _abi.CRng(*rng._state())
    status = np.zeros(1, dtype=np.int32)
    cnt = _abi.CCounters()
    nb = np.array([N], dtype=np.int32)
    _abi.check(lib, lib.bdsim_integrate_batch(1, _abi.iptr(nb), _abi.dptr(out),
                                              ctypes.byref(pp.c), ctypes.byref(cs),
                                              ctypes.byref(st), int(precision),
                                              _abi.iptr(status), ctypes.byref(cnt), None, 0))
    rng._set_state(st.ix, st.iy, st.seed)
    if cnt.nonconverged:
        _report_nonconvergence(cnt.nonconverged, N, sim.implicit_loop_tol)
    err = status_error(int(status[0]), cnt)
    if err is not None:
        raise err
    return out


def _force(name, R, phys):
    R = _positions(R)
    N = R.shape[0]
    lib = _abi.load_gpu()
    F = np.empty_like(R)
    pp = _abi.PackedPhys(phys, N)
    _abi.check(lib, getattr(lib, name)(1, N, _abi.dptr(R), ctypes.byref(pp.c), _abi.dptr(F)))
    return F


def spring_force(R, spring):
    p = PhysParams()
    p.spring = spring
    return _force("bdsim_spring_force_batch", R, p)


def bending_force(R, bend):
    p = PhysParams()
    p.bend = bend
    return _force("bdsim_bending_force_batch", R, p)


def ev_force(R, ev, equilibration=False):
    p = PhysParams()
    p.ev = ev
    p.equilibration = bool(equilibration)
    return _force("bdsim_ev_force_batch", R, p)


def total_force(R, phys):
    """Intramolecular force (spring + bending + EV). Excludes external forces."""
    return _force("bdsim_total_force_batch", R, phys)


def external_force(phys, nbeads, t=0.0):
    """Applied external force on each bead at time t."""
    lib = _abi.load_gpu()
    N = int(nbeads)
    F = np.empty((N, 3))
    pp = _abi.PackedPhys(phys, N) if not phys.ev.phi else _abi.PackedPhys(_without_phi(phys), N)
    _abi.check(lib, lib.bdsim_external_force_batch(ctypes.byref(pp.c), N, float(t), _abi.dptr(F)))
    return F


def _without_phi(phys):
    import copy
    p = copy.copy(phys)
    p.ev = copy.copy(phys.ev)
    p.ev.phi = []
    return p
