"""ctypes mirror of include/bdsim_gpu.h and the loaders of the native libraries.

The GPU library (bdsim_amd/csrc -> bdsim_amd/libbdsim_gpu.so) is the product.
`load_gpu()` raises if it is missing or if no GPU is visible: nothing in this
package falls back to a CPU implementation.
"""
import atexit
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GPU_LIB_PATH = os.path.join(HERE, "libbdsim_gpu.so")

_dp = POINTER(c_double)
_ip = POINTER(c_int32)


class CFlow(ctypes.Structure):
    _fields_ = [("constant", c_int32), ("n_times", c_int32), ("kappa0", c_double * 9),
                ("times", _dp), ("kappas", _dp)]


class CBeadForce(ctypes.Structure):
    _fields_ = [("bead", c_int32), ("constant", c_int32), ("value", c_double * 3),
                ("n_times", c_int32), ("times", _dp), ("values", _dp)]


class CPhys(ctypes.Structure):
    _fields_ = [("spring_type", c_int32), ("sqrtb", c_double), ("natural_length", c_double),
                ("ev_type", c_int32), ("zstar", c_double), ("dstar", c_double),
                ("min_cutoff", c_double), ("max_cutoff", c_double),
                ("contour_dist_for_ev", c_int32), ("n_phi", c_int32), ("phi", _dp),
                ("bend_type", c_int32), ("bend_stiffness", c_double),
                ("flow", CFlow),
                ("n_external", c_int32), ("external", POINTER(CBeadForce)),
                ("number_of_beads", c_int32), ("hstar", c_double), ("hi_method", c_int32),
                ("ncheb_multiplier", c_double), ("fd_err_max", c_double),
                ("equilibration", c_int32)]


class CSim(ctypes.Structure):
    _fields_ = [("time_start", c_double), ("time_end", c_double), ("dt", c_double),
                ("implicit_loop_tol", c_double), ("update_center_of_mass", c_int32)]


class CRng(ctypes.Structure):
    _fields_ = [("ix", c_int32), ("iy", c_int32), ("seed", c_int32)]


class CCounters(ctypes.Structure):
    _fields_ = [("steps", c_int64), ("matvecs", c_int64), ("lanczos_iters", c_int64),
                ("lanczos_cold", c_int64), ("lanczos_warm", c_int64),
                ("cheb_terms", c_int64), ("cheb_attempts", c_int64),
                ("fallbacks", c_int64), ("corrector_iters", c_int64),
                ("nonconverged", c_int64), ("fail_index", c_int32), ("pad_", c_int32),
                ("fail_value", c_double)]


class CStepLog(ctypes.Structure):
    _fields_ = [("t", c_double), ("lmin", c_double), ("lmax", c_double), ("fd_err", c_double),
                ("lanczos_cold", c_int32), ("lanczos_steps", c_int32), ("cheb_L", c_int32),
                ("cheb_attempts", c_int32), ("fallback", c_int32),
                ("corrector_iters", c_int32)]


COUNTER_FIELDS = [f for f, _ in CCounters._fields_ if f != "pad_"]
LOG_FIELDS = [f for f, _ in CStepLog._fields_]

STATUS_OK = 0
STATUS_NOT_POSITIVE_DEFINITE = 1
STATUS_RTSAFE_UNBRACKETED = 2
STATUS_RTSAFE_MAXIT = 3
STATUS_EXACT_SQRT = 4
STATUS_WLCB_SIGMA = 5
STATUS_NOT_RUN = 6
STATUS_SYNC_TIMEOUT = 7

PRECISION_FP64 = 0
PRECISION_MIXED = 1
PRECISION_MIXED_F32_FACTOR = 2


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or NULL for None)."""
    if a is None:
        return ctypes.cast(None, _dp)
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return ctypes.cast(None, _ip)
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


class PackedPhys:
    """A CPhys plus the numpy buffers its pointers refer to (kept alive here)."""

    def __init__(self, phys, nbeads):
        self.keep = []
        c = CPhys()
        sp = phys.spring
        c.spring_type = int(sp.type)
        c.sqrtb = float(sp.sqrtb)
        c.natural_length = float(sp.natural_length)
        ev = phys.ev
        c.ev_type = int(ev.type)
        c.zstar, c.dstar = float(ev.zstar), float(ev.dstar)
        c.min_cutoff, c.max_cutoff = float(ev.min_cutoff), float(ev.max_cutoff)
        c.contour_dist_for_ev = int(ev.contour_dist_for_EV)
        phi = np.ascontiguousarray(np.asarray(ev.phi, dtype=np.float64).ravel())
        if phi.size and int(ev.type) == 3 and phi.size != nbeads * nbeads:
            raise ValueError(f"EVParams.phi has {phi.size} entries; an SDK chain of "
                             f"{nbeads} beads needs {nbeads * nbeads} (row-major N*N)")
        c.n_phi = int(phi.size)
        c.phi = dptr(phi) if phi.size else dptr(None)
        self.keep.append(phi)
        c.bend_type = int(phys.bend.type)
        c.bend_stiffness = float(phys.bend.stiffness)
        fl = phys.flow
        c.flow.constant = 1 if fl.is_constant() else 0
        k0 = np.asarray(fl.constant_kappa(), dtype=np.float64).ravel()
        for i in range(9):
            c.flow.kappa0[i] = k0[i]
        ts = np.ascontiguousarray(np.asarray(fl.sample_times(), dtype=np.float64))
        ks = np.ascontiguousarray(np.asarray(fl.sample_tensors(), dtype=np.float64).ravel())
        c.flow.n_times = int(ts.size)
        c.flow.times = dptr(ts) if ts.size else dptr(None)
        c.flow.kappas = dptr(ks) if ks.size else dptr(None)
        self.keep += [ts, ks]
        entries = phys.external.entries()
        if entries:
            arr = (CBeadForce * len(entries))()
            for k, e in enumerate(entries):
                arr[k].bead = int(e.bead)
                arr[k].constant = 1 if e.constant else 0
                for d in range(3):
                    arr[k].value[d] = float(e.value[d])
                et = np.ascontiguousarray(np.asarray(e.times, dtype=np.float64))
                ev_ = np.ascontiguousarray(np.asarray(e.values, dtype=np.float64).ravel())
                arr[k].n_times = int(et.size)
                arr[k].times = dptr(et) if et.size else dptr(None)
                arr[k].values = dptr(ev_) if ev_.size else dptr(None)
                self.keep += [et, ev_]
            c.n_external = len(entries)
            c.external = ctypes.cast(arr, POINTER(CBeadForce))
            self.keep.append(arr)
        else:
            c.n_external = 0
            c.external = ctypes.cast(None, POINTER(CBeadForce))
        c.number_of_beads = int(nbeads)
        c.hstar = float(phys.hstar)
        c.hi_method = int(phys.hi_method)
        c.ncheb_multiplier = float(phys.ncheb_multiplier)
        c.fd_err_max = float(phys.fd_err_max)
        c.equilibration = 1 if phys.equilibration else 0
        self.c = c


MAX_SEGMENT_STEPS = 1.0e9


def check_clock(t0, t1, dt):
    """The checks the library makes before a launch (check_clock in
    bdsim_kernels.hip): the device loop `for (t = t0; t <= t1 + dt/2; t += dt)`
    must end, so dt is positive and finite, the ends are finite, t += dt moves
    the clock by at least dt/2 everywhere in the segment, and the segment has at
    most 1e9 steps. Raises ValueError otherwise."""
    t0, t1, dt = float(t0), float(t1), float(dt)
    if not (dt > 0.0) or not np.isfinite(dt):
        raise ValueError(f"dt must be positive and finite, got {dt!r}")
    if not (np.isfinite(t0) and np.isfinite(t1)):
        raise ValueError(f"time_start and time_end must be finite, got {t0!r}, {t1!r}")
    m = max(abs(t0), abs(t1)) + dt
    if not ((m + dt) - m >= 0.5 * dt):
        raise ValueError(f"dt={dt!r} is below the resolution of a clock near {m!r}")
    if (t1 - t0) / dt > MAX_SEGMENT_STEPS:
        raise ValueError("segment longer than 1e9 steps")


def pack_sim(sim):
    check_clock(sim.time_start, sim.time_end, sim.dt)
    return CSim(float(sim.time_start), float(sim.time_end), float(sim.dt),
                float(sim.implicit_loop_tol), 1 if sim.update_center_of_mass else 0)


# --------------------------------------------------------------------------
# the GPU library
# --------------------------------------------------------------------------

_gpu = None


class GPUUnavailable(RuntimeError):
    """The HIP library is not built, or no MI355X-class GPU is visible."""


class CServerStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int64), ("running", ctypes.c_int64),
                ("active_trajectories", ctypes.c_int64), ("bytes_reserved", ctypes.c_int64),
                ("bytes_in_use", ctypes.c_int64), ("bead_steps_done", ctypes.c_int64)]


class CObservables(ctypes.Structure):
    """bdsim_observables: one chain's snapshot observables."""
    _fields_ = [("rg2", c_double), ("q2", c_double), ("gyration", c_double * 9),
                ("stretch", c_double * 3), ("stress", c_double * 9), ("com", c_double * 3)]


FORCE_TOTAL = 0
FORCE_SPRING = 1
FORCE_BENDING = 2
FORCE_EV = 3


def _declare_gpu(lib):
    P = POINTER
    lib.bdsim_init.argtypes = [c_int]
    lib.bdsim_shutdown.argtypes = []
    lib.bdsim_last_error.restype = ctypes.c_char_p
    lib.bdsim_max_beads.restype = c_int
    lib.bdsim_tune_set.argtypes = [c_int, c_int]
    lib.bdsim_tune_get.argtypes = [c_int]
    lib.bdsim_tune_split.argtypes = [c_int]
    lib.bdsim_tune_split_get.argtypes = []
    lib.bdsim_tune_width.argtypes = [c_int, c_int]
    lib.bdsim_tune_width_get.argtypes = [c_int]
    lib.bdsim_integrate_batch.argtypes = [c_int, _ip, _dp, P(CPhys), P(CSim), P(CRng), c_int,
                                          _ip, P(CCounters), P(CStepLog), c_int]
    lib.bdsim_integrate_batch_static.argtypes = lib.bdsim_integrate_batch.argtypes
    lib.bdsim_static_scratch_bytes.restype = ctypes.c_size_t
    lib.bdsim_batch_create.argtypes = [P(c_void_p), c_int, c_int, _dp, P(CPhys), _dp,
                                       P(CRng), c_int]
    lib.bdsim_batch_run.argtypes = [c_void_p, P(CSim)]
    lib.bdsim_batch_run_steps.argtypes = [c_void_p, P(CSim), c_int, c_int]
    lib.bdsim_batch_sync.argtypes = [c_void_p]
    lib.bdsim_batch_fetch.argtypes = [c_void_p, _dp, P(CRng), _ip, P(CCounters)]
    lib.bdsim_batch_stream.argtypes = [c_void_p]
    lib.bdsim_batch_stream.restype = c_void_p
    lib.bdsim_batch_destroy.argtypes = [c_void_p]
    lib.bdsim_batch_destroy.restype = None
    for name in ("bdsim_total_force_batch", "bdsim_spring_force_batch",
                 "bdsim_bending_force_batch", "bdsim_ev_force_batch"):
        getattr(lib, name).argtypes = [c_int, c_int, _dp, P(CPhys), _dp]
    lib.bdsim_external_force_batch.argtypes = [P(CPhys), c_int, c_double, _dp]
    lib.bdsim_raw_noise_batch.argtypes = [c_int, c_int, c_int, P(CRng), _dp]
    lib.bdsim_rpy_apply_batch.argtypes = [c_int, c_int, c_double, c_int, _dp, _dp, _dp]
    lib.bdsim_brownian_batch.argtypes = [c_int, c_int, _dp, P(CPhys), c_int, _dp, _dp, _ip]
    lib.bdsim_server_create.argtypes = [P(c_void_p), c_int, ctypes.c_int64, c_int]
    lib.bdsim_server_submit.argtypes = [c_void_p, ctypes.c_int64, c_int, c_int, _dp, P(CPhys),
                                        P(CSim), P(CRng)]
    lib.bdsim_server_poll.argtypes = [c_void_p, c_int]
    lib.bdsim_server_completed.argtypes = [c_void_p, P(ctypes.c_int64), c_int]
    lib.bdsim_server_result.argtypes = [c_void_p, ctypes.c_int64, _dp, P(CRng), _ip,
                                        P(CCounters)]
    lib.bdsim_server_query.argtypes = [c_void_p, P(CServerStats)]
    lib.bdsim_server_destroy.argtypes = [c_void_p]
    lib.bdsim_server_destroy.restype = None
    lib.bdsim_observables_batch.argtypes = [c_int, c_int, _dp, P(CPhys), c_int, P(CObservables)]
    lib.bdsim_batch_observe.argtypes = [c_void_p, c_int, P(CObservables)]
    return lib


def load_gpu_library():
    """ctypes handle of libbdsim_gpu.so without touching the GPU (symbol checks)."""
    if not os.path.exists(GPU_LIB_PATH):
        raise GPUUnavailable(
            f"{GPU_LIB_PATH} is not built; run `python __graft_entry__.py` "
            f"(or `make -C bdsim_amd/csrc`) first")
    return _declare_gpu(ctypes.CDLL(GPU_LIB_PATH))


def load_gpu():
    """The initialised GPU library. Raises GPUUnavailable -- never falls back."""
    global _gpu
    if _gpu is None:
        # The batch path (one stream per block size) and the serving loop (one
        # per group, up to 12) launch side by side on streams of their own;
        # HIP runs the launches of streams that share one of its
        # GPU_MAX_HW_QUEUES hardware queues (4 by default) one after another.
        # Unless the process has chosen a value, ask for 16 (read by HIP at its
        # first call in the process, so it has no effect once HIP is running).
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
        lib = load_gpu_library()
        rc = lib.bdsim_init(int(os.environ.get("BDSIM_DEVICE", "0")))
        if rc != 0:
            raise GPUUnavailable(f"bdsim_init failed ({rc}): "
                                 f"{lib.bdsim_last_error().decode()}")
        _gpu = lib
        atexit.register(shutdown)
        # a launch-shape table saved by bdsim_amd.autotune applies to every
        # call of this process when BDSIM_TUNE_FILE names it
        path = os.environ.get("BDSIM_TUNE_FILE")
        if path:
            from . import autotune
            autotune.apply(autotune.load(path), lib)
    return _gpu


def shutdown():
    """Release the library's device state (bdsim_shutdown): registered to run
    at interpreter exit, before the HIP runtime unloads."""
    global _gpu
    if _gpu is not None:
        lib, _gpu = _gpu, None
        lib.bdsim_shutdown()


def check(lib, rc):
    if rc != 0:
        raise RuntimeError(f"bdsim GPU call failed ({rc}): {lib.bdsim_last_error().decode()}")
