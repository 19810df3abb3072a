"""An ensemble of equal-length chains kept in HBM between segments.

`ResidentEnsemble` holds positions, RNG states and workspace on the device
(bdsim_batch_*). `run(k)` integrates k steps as one segment [t, t + (k-1) dt],
the chunking of python/bdsim/ensemble.py:106-123: chunked stepping equals one
continuous run except for the cold spectral start each segment makes, exactly
as successive `integrate` calls do. `run_segment(t0, t1)` integrates the segment
[t0, t1] as `integrate` does with `_segment(sim, t0, t1)` (the sampling of
`trajectory_samples`, python/bdsim/ensemble.py:126-138). `observe()` evaluates the snapshot
observables of python/bdsim/properties.py:5-52 on the device, so ensemble
averages (Rg^2, Q^2, the Kramers stress behind the viscosities) need no copy
of the positions to the host.
"""
import ctypes
from ctypes import byref

import numpy as np

from . import _abi
from .core import Rng

_FORCES = {"total": _abi.FORCE_TOTAL, "spring": _abi.FORCE_SPRING,
           "bending": _abi.FORCE_BENDING, "ev": _abi.FORCE_EV}


def _rng_states(rngs, n):
    out = (_abi.CRng * n)()
    for i, r in enumerate(rngs):
        if isinstance(r, Rng):
            ix, iy, seed = r._state()
        elif isinstance(r, _abi.CRng):
            ix, iy, seed = r.ix, r.iy, r.seed
        else:                        # a seed: the state of Rng(seed) (rng.hpp:30-36)
            ix, iy, seed = 1, -1, int(r)
        out[i] = _abi.CRng(ix, iy, seed)
    return out


def observables_dict(recs, n):
    """Structured numpy view of n bdsim_observables records."""
    raw = np.ctypeslib.as_array(ctypes.cast(recs, ctypes.POINTER(ctypes.c_double)),
                                shape=(n, 26)).copy()
    return {"rg2": raw[:, 0], "q2": raw[:, 1], "gyration": raw[:, 2:11].reshape(n, 3, 3),
            "stretch": raw[:, 11:14], "stress": raw[:, 14:23].reshape(n, 3, 3),
            "com": raw[:, 23:26]}


class ResidentEnsemble:
    """n chains of N beads sharing one PhysParams, each with its own flow
    gradient (kappas, (n, 3, 3), optional) and RNG stream (Rng objects,
    states or seeds)."""

    def __init__(self, R0, phys, sim, rngs, precision=_abi.PRECISION_FP64, kappas=None):
        self.lib = _abi.load_gpu()
        R0 = np.ascontiguousarray(R0, dtype=np.float64)
        if R0.ndim != 3 or R0.shape[2] != 3:
            raise ValueError("R0 must have shape (n_traj, N, 3)")
        self.n, self.N = int(R0.shape[0]), int(R0.shape[1])
        self.sim = sim
        self.t = float(sim.time_start)
        self._phys = _abi.PackedPhys(phys, self.N)
        K = None
        if kappas is not None:
            K = np.ascontiguousarray(np.asarray(kappas, dtype=np.float64).reshape(self.n, 9))
        if len(rngs) != self.n:
            raise ValueError("one RNG per chain")
        st = _rng_states(rngs, self.n)
        self.h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.bdsim_batch_create(byref(self.h), self.n, self.N,
                                                         _abi.dptr(R0), byref(self._phys.c),
                                                         _abi.dptr(K), st, int(precision)))

    @property
    def time(self):
        return self.t

    def _queue(self, t0, t1):
        s = _abi.CSim(float(t0), float(t1), float(self.sim.dt), float(self.sim.implicit_loop_tol),
                      1 if self.sim.update_center_of_mass else 0)
        _abi.check(self.lib, self.lib.bdsim_batch_run(self.h, byref(s)))

    def run(self, n_steps):
        """Queue n_steps steps of every chain as one segment (returns at once)."""
        if n_steps <= 0:
            return
        self._queue(self.t, self.t + (n_steps - 1) * self.sim.dt)
        self.t += n_steps * self.sim.dt

    def run_segment(self, t0, t1):
        """Queue the segment [t0, t1] of every chain (the steps t = t0, t0 + dt,
        ... while t <= t1 + dt/2, as `integrate` takes a SimParams)."""
        self._queue(t0, t1)
        self.t = float(t1)

    def observe(self, force="total"):
        """Snapshot observables of every chain at the current step (dict of arrays)."""
        recs = (_abi.CObservables * self.n)()
        _abi.check(self.lib, self.lib.bdsim_batch_observe(self.h, _FORCES[force], recs))
        return observables_dict(recs, self.n)

    def fetch(self):
        """(R (n, N, 3), RNG states (n, 3), status (n,), counters)."""
        R = np.empty((self.n, self.N, 3))
        st = (_abi.CRng * self.n)()
        status = np.zeros(self.n, dtype=np.int32)
        cnt = (_abi.CCounters * self.n)()
        _abi.check(self.lib, self.lib.bdsim_batch_fetch(self.h, _abi.dptr(R), st,
                                                        _abi.iptr(status), cnt))
        return R, np.array([(s.ix, s.iy, s.seed) for s in st], dtype=np.int64), status, cnt

    def close(self):
        if self.h:
            self.lib.bdsim_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def observe(R, phys, force="total"):
    """Snapshot observables of positions R (n, N, 3) on the device."""
    lib = _abi.load_gpu()
    R = np.ascontiguousarray(R, dtype=np.float64)
    n, N = int(R.shape[0]), int(R.shape[1])
    p = _abi.PackedPhys(phys, N)
    recs = (_abi.CObservables * n)()
    _abi.check(lib, lib.bdsim_observables_batch(n, N, _abi.dptr(R), byref(p.c), _FORCES[force], recs))
    return observables_dict(recs, n)
