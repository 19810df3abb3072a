"""Test-side ctypes bindings of the oracle (oracle/_build/liboracle.so), of the
reference core build (oracle/_ref/libbdref.so) and of the GPU library, plus
packing of the plain case descriptions in tests/golden/cases.py."""
import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_long

import numpy as np

from bdsim_amd import _abi as A

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(REPO, "oracle", "_build", "liboracle.so")
REF_PATH = os.path.join(REPO, "oracle", "_ref", "libbdref.so")
GOLDEN = os.path.join(REPO, "tests", "golden")

_dp = POINTER(c_double)
_P = POINTER


def oracle():
    lib = ctypes.CDLL(ORACLE_PATH)
    lib.bo_ran1.argtypes = [_P(A.CRng), c_int, _dp]
    lib.bo_raw_noise.argtypes = [_P(A.CRng), c_int, _dp]
    for n in ("bo_spring_force", "bo_bending_force", "bo_ev_force", "bo_total_force"):
        getattr(lib, n).argtypes = [c_int, _dp, _P(A.CPhys), _dp]
    lib.bo_external_force.argtypes = [_P(A.CPhys), c_int, c_double, _dp]
    lib.bo_kappa.argtypes = [_P(A.CFlow), c_double, _dp]
    lib.bo_rpy_matrix.argtypes = [c_int, _dp, c_double, _dp]
    lib.bo_matvec.argtypes = [c_int, _dp, _dp, _dp]
    lib.bo_cholesky_lower.argtypes = [c_int, _dp, _dp, _P(c_int), _dp]
    lib.bo_lower_matvec.argtypes = [c_int, _dp, _dp, _dp]
    lib.bo_spectral_bounds.argtypes = [c_int, _dp, _dp, _P(c_int), _dp, _dp, _P(c_int), _P(c_int)]
    lib.bo_chebyshev_sqrt_times.argtypes = [c_int, _dp, _dp, c_double, c_double, _dp,
                                            _P(c_int), _dp, _P(c_int), _dp]
    lib.bo_solve_implicit_r.argtypes = [c_int, c_double, c_double, c_double, _dp, _dp]
    lib.bo_solve_connector.argtypes = [_dp, c_double, c_int, c_double, c_double, _dp]
    lib.bo_integrate.argtypes = [c_int, _dp, _P(A.CPhys), _P(A.CSim), _P(A.CRng),
                                 _P(A.CCounters), _P(A.CStepLog), c_int]
    return lib


def ref():
    """The reference core build, or None where it was not built."""
    if not os.path.exists(REF_PATH):
        return None
    lib = ctypes.CDLL(REF_PATH)
    lib.ref_last_error.restype = ctypes.c_char_p
    lib.ref_ran1.argtypes = [_P(A.CRng), c_int, _dp]
    lib.ref_integrate.argtypes = [c_int, _dp, _P(A.CPhys), _P(A.CSim), _P(A.CRng)]
    for n in ("ref_spring_force", "ref_bending_force", "ref_ev_force", "ref_total_force"):
        getattr(lib, n).argtypes = [c_int, _dp, _P(A.CPhys), _dp]
    lib.ref_external_force.argtypes = [_P(A.CPhys), c_int, c_double, _dp]
    lib.ref_kappa.argtypes = [_P(A.CPhys), c_double, _dp]
    lib.ref_rpy_matrix.argtypes = [c_int, _dp, c_double, _dp]
    lib.ref_cholesky.argtypes = [c_int, _dp, _dp]
    lib.ref_chebyshev_sqrt_times.argtypes = [c_int, _dp, _dp, c_double, c_double, _dp,
                                             _P(c_int), _dp]
    lib.ref_spectral_bounds_seq.argtypes = [c_int, c_int, _dp, _dp, _P(c_long), _P(c_long)]
    lib.ref_solve_implicit_r.argtypes = [c_int, c_double, c_double, c_double, _dp, _dp]
    lib.ref_solve_connector.argtypes = [_dp, c_double, c_int, c_double, c_double, _dp]
    return lib


class Packed:
    """A CPhys built from a case dict, with its buffers kept alive."""

    def __init__(self, d, N):
        self.keep = []
        c = A.CPhys()
        c.spring_type = d["spring"]
        c.sqrtb = d["sqrtb"]
        c.natural_length = d["nat"]
        c.ev_type = d["ev"]
        c.zstar, c.dstar = d["zstar"], d["dstar"]
        c.min_cutoff, c.max_cutoff = d["min_cutoff"], d["max_cutoff"]
        c.contour_dist_for_ev = d["cd"]
        if d["phi"] is not None:
            phi = np.ascontiguousarray(np.asarray(d["phi"], dtype=np.float64).ravel())
            self.keep.append(phi)
            c.n_phi = phi.size
            c.phi = A.dptr(phi)
        c.bend_type = d["bend"]
        c.bend_stiffness = d["C"]
        for i in range(9):
            c.flow.kappa0[i] = d["kappa"][i]
        if d["flow_times"] is not None:
            ts = np.ascontiguousarray(np.asarray(d["flow_times"], dtype=np.float64))
            ks = np.ascontiguousarray(np.asarray(d["flow_kappas"], dtype=np.float64).ravel())
            self.keep += [ts, ks]
            c.flow.constant = 0
            c.flow.n_times = ts.size
            c.flow.times = A.dptr(ts)
            c.flow.kappas = A.dptr(ks)
        else:
            c.flow.constant = 1
        ext = d["external"]
        if ext:
            arr = (A.CBeadForce * len(ext))()
            for k, (bead, const, value, times, values) in enumerate(ext):
                arr[k].bead = bead
                arr[k].constant = 1 if const else 0
                for q in range(3):
                    arr[k].value[q] = value[q]
                if not const:
                    ts = np.ascontiguousarray(np.asarray(times, dtype=np.float64))
                    vs = np.ascontiguousarray(np.asarray(values, dtype=np.float64).ravel())
                    self.keep += [ts, vs]
                    arr[k].n_times = ts.size
                    arr[k].times = A.dptr(ts)
                    arr[k].values = A.dptr(vs)
            self.keep.append(arr)
            c.n_external = len(ext)
            c.external = ctypes.cast(arr, _P(A.CBeadForce))
        c.number_of_beads = N
        c.hstar = d["hstar"]
        c.hi_method = d["method"]
        c.ncheb_multiplier = d["mult"]
        c.fd_err_max = d["fd"]
        c.equilibration = d["equil"]
        self.c = c


def csim(s):
    return A.CSim(s["t0"], s["t1"], s["dt"], s["tol"], s["com"])


def rng_of(seed):
    """State of Rng(seed) (rng.hpp:30-36): ix=1, iy=-1, lazily seeded."""
    return A.CRng(1, -1, seed)


def oracle_integrate(lib, N, R, phys_c, sim_c, st, log_cap=0):
    R = np.ascontiguousarray(R, dtype=np.float64).copy()
    cnt = A.CCounters()
    logs = (A.CStepLog * max(log_cap, 1))()
    rc = lib.bo_integrate(N, A.dptr(R), byref(phys_c), byref(sim_c), byref(st), byref(cnt),
                          logs if log_cap else None, log_cap)
    return R, rc, cnt, (list(logs)[:min(log_cap, cnt.steps)] if log_cap else [])


def ref_integrate(lib, N, R, phys_c, sim_c, st):
    R = np.ascontiguousarray(R, dtype=np.float64).copy()
    rc = lib.ref_integrate(N, A.dptr(R), byref(phys_c), byref(sim_c), byref(st))
    return R, rc


def run_segments(integrate_fn, N, R0, phys_c, sim_d, seed, segments):
    """Integrate `segments` consecutive calls with the same SimParams on one
    continued RNG stream (the chained-call pattern of
    tests/test_integrator_rouse.cpp:105-112 and compare_fortran_oracle.py:189-205)."""
    st = rng_of(seed)
    R = R0.copy()
    outs = []
    for k in range(segments):
        res = integrate_fn(N, R, phys_c, csim(sim_d), st)
        R = res[0]
        outs.append(res)
        if res[1] != 0:
            break
    return outs, st


def gpu_batch(lib, items, precision=A.PRECISION_FP64, log_stride=0):
    """Run bdsim_integrate_batch on items = [(N, R0, CPhys, CSim, CRng)].
    Returns (list of R, rng states, status array, counters, logs)."""
    n = len(items)
    nb = np.array([it[0] for it in items], dtype=np.int32)
    R = np.ascontiguousarray(np.concatenate([np.asarray(it[1], dtype=np.float64).ravel() for it in items]))
    phys = (A.CPhys * n)(*[it[2] for it in items])
    sims = (A.CSim * n)(*[it[3] for it in items])
    rngs = (A.CRng * n)(*[A.CRng(it[4].ix, it[4].iy, it[4].seed) for it in items])
    status = np.zeros(n, dtype=np.int32)
    cnt = (A.CCounters * n)()
    logs = (A.CStepLog * max(1, n * log_stride))()
    rc = lib.bdsim_integrate_batch(n, A.iptr(nb), A.dptr(R), phys, sims, rngs, precision,
                                   A.iptr(status), cnt, logs if log_stride else None, log_stride)
    A.check(lib, rc)
    out, off = [], 0
    for N in nb:
        out.append(R[3 * off:3 * (off + N)].reshape(N, 3).copy())
        off += N
    return out, list(rngs), status, list(cnt), logs
