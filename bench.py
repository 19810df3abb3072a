"""Benchmark of the bdsim ChainIntegrator step on one MI355X.

Workload (the headline metric's configuration): 4096 chains of N=50 beads,
FENE springs b=100 (sqrtb=10), RPY hydrodynamics h*=0.3 with the Chebyshev
square root (multiplier 1, FD tolerance 0.0025), SDK excluded volume (d*=1,
cutoffs 0.7/1.5, phi=0), simple shear with Wi in {0.1, 1, 10, 100} assigned
round robin, dt=1e-3, implicit tolerance 1e-3, centre-of-mass update on.
The shear rate is Wi / tau_1 with tau_1 = 1/sin^2(pi/2N), the longest Rouse
relaxation time in code units -- the reference's own way of placing a run at a
Weissenberg number (python/bdsim/dynamics.py:583-606,
validation/hi_nan_probe.py:103).
Chain i uses Rng(seed=i) and a Gaussian random-walk start (bond std 1) drawn
from numpy's generator with seed i. Data is synthetic.

A "step" is one integrator step of every chain. Each chain integrates ONE
segment of W + K steps (one integrate call: a single cold Lanczos start, at its
first step, as the reference's run_ensemble integrates a trajectory), with
positions, RNG states and workspace resident in HBM; the first W steps run
untimed, then the K timed steps continue the segment in one launch
(bdsim_batch_run_steps). `value` is the mixed-precision form's
bead-steps/s = 4096 * 50 * K / seconds; the plain FP64 form is timed the same
way and reported beside it. `fresh_segment` times the same K steps as a segment
of their own (a separate integrate call, its cold start inside the timed
region), the cost of short segments.

roofline: algorithmic FLOPs of one launch = sum over chains of 18 N^2 per
mobility product (a dense 3N x 3N mat-vec) times the mat-vecs the run itself
counted; achieved = that / kernel time measured with HIP events on the batch's
stream. Peak: MI355X vector FP32 157.3 TFLOP/s (mixed) / FP64 78.6 TFLOP/s.

cpu_baseline: the reference core (oracle/_ref, built from the reference
sources) -- or, where that build is absent, the FP64 C restatement -- integrating
a bounded sample of the same workload in one process per host core (every CPU
the process may use, at most the GPU host's share of 16), with the host's nproc
and CPU model.

mixed_errors: the mixed form against the FP64 form on the first 256 chains --
relative errors of D.F, B.dW and positions after 100 steps, and the count of
chain-steps whose discrete decisions (Chebyshev degree, FD attempts, corrector
passes, fallbacks, Krylov dimension) differ.

serving: the first 40 jobs of the serving stream at their full step counts,
replayed in real time; p50 / p99 latency from arrival to result on the host.
"""
import argparse
import ctypes
import json
import os
import sys
import time
from ctypes import byref

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

# HIP places each stream on one of GPU_MAX_HW_QUEUES hardware queues (4 by
# default) and runs the launches of streams that share a queue one after
# another. The serving loop runs up to 12 groups (one per block size and
# method) side by side, each on a stream of its own, so this process asks for
# 16 queues; it must be set before the first HIP call of the process. Launches
# on a single stream (the throughput legs) do not depend on it.
HW_QUEUES = 16
os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

from bdsim_amd import _abi as A  # noqa: E402
from bdsim_amd.workloads import (N_BEADS, N_TRAJ, WI, headline_phys, kappas, shear_rate,  # noqa: E402
                                 cholesky_workload, long_workload, start_configs,
                                 uneven_workload)

PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6, "mixed_f32_factor": 157.3}
PRECISION = {"fp64": A.PRECISION_FP64, "mixed": A.PRECISION_MIXED,
             "mixed_f32_factor": A.PRECISION_MIXED_F32_FACTOR}


class HipEvents:
    """hipEvent timing on an arbitrary stream (the batch's own)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.h.hipEventElapsedTime(byref(ms), a, b) == 0
        return float(ms.value)


def copy_calibration(nbytes):
    """One device-to-device copy of `nbytes` (the runtime's blit kernel): the
    known-size reference that scales the FETCH_SIZE / WRITE_SIZE counters."""
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(byref(src), nbytes) == 0 and h.hipMalloc(byref(dst), nbytes) == 0
    assert h.hipMemcpy(dst, src, nbytes, 3) == 0          # hipMemcpyDeviceToDevice
    assert h.hipDeviceSynchronize() == 0
    h.hipFree(src)
    h.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)


def _counter_arrays(lib, b, n_traj):
    cnt = (A.CCounters * n_traj)()
    A.check(lib, lib.bdsim_batch_fetch(b, None, None, None, cnt))
    return {k: np.array([getattr(c, k) for c in cnt], dtype=np.float64)
            for k in ("matvecs", "steps", "lanczos_cold", "fallbacks", "corrector_iters")}


def run_gpu(lib, precision, steps, warmup, R0, K, ev, fresh=False):
    """Time `steps` steps of every chain on a resident batch.

    Default: the chains integrate ONE segment of warmup + steps steps (what a
    long run_ensemble integration is: one integrate call per trajectory, one
    cold Lanczos start); its first `warmup` steps run untimed, the timed
    launch continues the segment (bdsim_batch_run_steps resumes it from the
    stored time, warm seed and generator state). fresh=True: the timed steps
    are a segment of their own after a separate warm-up segment, so the timed
    region includes a segment's cold start."""
    phys = headline_phys()
    n_traj = R0.shape[0]
    rng = (A.CRng * n_traj)(*[A.CRng(1, -1, i) for i in range(n_traj)])
    b = ctypes.c_void_p()
    A.check(lib, lib.bdsim_batch_create(byref(b), n_traj, N_BEADS, A.dptr(R0), byref(phys),
                                        A.dptr(K), rng, precision))
    try:
        dt = 1e-3
        stream = lib.bdsim_batch_stream(b)
        e0, e1 = ev.create(), ev.create()
        if fresh:
            if warmup > 0:
                sim = A.CSim(0.0, (warmup - 1) * dt, dt, 1e-3, 1)
                A.check(lib, lib.bdsim_batch_run(b, byref(sim)))
            t_start = warmup * dt
            sim = A.CSim(t_start, t_start + (steps - 1) * dt, dt, 1e-3, 1)
            launch = lambda: lib.bdsim_batch_run(b, byref(sim))  # noqa: E731
            before = None
        else:
            sim = A.CSim(0.0, (warmup + steps - 1) * dt, dt, 1e-3, 1)
            if warmup > 0:
                A.check(lib, lib.bdsim_batch_run_steps(b, byref(sim), warmup, 0))
            launch = lambda: lib.bdsim_batch_run_steps(b, byref(sim), steps, 1 if warmup > 0 else 0)  # noqa: E731
        A.check(lib, lib.bdsim_batch_sync(b))
        if not fresh:
            before = _counter_arrays(lib, b, n_traj) if warmup > 0 else None
        w0 = time.perf_counter()
        ev.record(e0, stream)
        A.check(lib, launch())
        ev.record(e1, stream)
        A.check(lib, lib.bdsim_batch_sync(b))
        wall = time.perf_counter() - w0
        kernel_ms = ev.elapsed_ms(e0, e1)
        after = _counter_arrays(lib, b, n_traj)
        status = np.zeros(n_traj, dtype=np.int32)
        Rout = np.empty_like(R0)
        A.check(lib, lib.bdsim_batch_fetch(b, A.dptr(Rout), None, A.iptr(status), None))
    finally:
        lib.bdsim_batch_destroy(b)
    c = {k: after[k] - (before[k] if before is not None else 0.0) for k in after}
    flops = float(np.sum(18.0 * N_BEADS * N_BEADS * c["matvecs"]))
    n_steps = max(c["steps"].sum(), 1.0)
    return dict(wall_s=wall, kernel_ms=kernel_ms, flops=flops, status_ok=int(np.sum(status == 0)),
                steps_done=int(c["steps"].min()), matvecs_per_step=float(c["matvecs"].sum() / n_steps),
                finite=bool(np.all(np.isfinite(Rout))), lanczos_cold=int(c["lanczos_cold"].sum()),
                fallbacks=int(c["fallbacks"].sum()),
                corrector_per_step=float(c["corrector_iters"].sum() / n_steps))


def per_step_launches(lib, precision, n_steps, R0, K, ev):
    """The first n_steps steps of one segment of the headline batch as one
    launch per step (resumed segment): per step, the launch time (HIP events)
    and the mean and largest number of mat-vecs over the chains. Shows where
    a short segment's time goes: the cold first step, and the early steps
    from the random-walk start, whose launches last as long as their costliest
    chains."""
    phys = headline_phys()
    n_traj = R0.shape[0]
    rng = (A.CRng * n_traj)(*[A.CRng(1, -1, i) for i in range(n_traj)])
    b = ctypes.c_void_p()
    A.check(lib, lib.bdsim_batch_create(byref(b), n_traj, N_BEADS, A.dptr(R0), byref(phys),
                                        A.dptr(K), rng, precision))
    out = []
    try:
        sim = A.CSim(0.0, (n_steps - 1) * 1e-3, 1e-3, 1e-3, 1)
        stream = lib.bdsim_batch_stream(b)
        prev = None
        for k in range(n_steps):
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, stream)
            A.check(lib, lib.bdsim_batch_run_steps(b, byref(sim), 1, 1 if k else 0))
            ev.record(e1, stream)
            A.check(lib, lib.bdsim_batch_sync(b))
            c = _counter_arrays(lib, b, n_traj)["matvecs"]
            mv = c - (prev if prev is not None else 0.0)
            prev = c
            out.append({"step": k, "ms": round(ev.elapsed_ms(e0, e1), 4),
                        "matvecs_mean": round(float(mv.mean()), 3), "matvecs_max": int(mv.max())})
    finally:
        lib.bdsim_batch_destroy(b)
    return out


def run_workload(lib, name, precision, steps, warmup):
    """Integrate a mixed-size workload through bdsim_integrate_batch: one warm-up
    segment, then one timed segment; the time includes the host <-> device
    copies of positions and states (a few MB)."""
    N, R, phys, dt = {"uneven": uneven_workload, "long": long_workload,
                      "cholesky": cholesky_workload}[name]()
    n = len(N)
    Rf = np.ascontiguousarray(np.concatenate([r.ravel() for r in R]))
    ph = (A.CPhys * n)(*phys)
    rng = (A.CRng * n)(*[A.CRng(1, -1, i) for i in range(n)])
    st = np.zeros(n, dtype=np.int32)
    cnt = (A.CCounters * n)()

    def segment(t0, k):
        sims = (A.CSim * n)(*[A.CSim(t0, t0 + (k - 1) * dt, dt, 1e-3, 1)] * n)
        t = time.perf_counter()
        A.check(lib, lib.bdsim_integrate_batch(n, A.iptr(N), A.dptr(Rf), ph, sims, rng, precision,
                                               A.iptr(st), cnt, None, 0))
        return time.perf_counter() - t

    if warmup > 0:
        segment(0.0, warmup)
    wall = segment(warmup * dt, steps)
    # algorithmic FLOPs: 18 N^2 per D.x the run counted, plus n^3/3 + 2 n^2 per dense
    # factor and L x (n = 3N): one per step for the Cholesky method, one per fallback
    fl = 0.0
    for i, c in enumerate(cnt):
        nn = 3.0 * int(N[i])
        factors = int(c.fallbacks) + (int(c.steps) if phys[i].hi_method == 1 and phys[i].hstar > 0 else 0)
        fl += int(c.matvecs) * 18.0 * int(N[i]) ** 2 + factors * (nn ** 3 / 3.0 + 2.0 * nn * nn)
    return dict(bead_steps_per_s=float(N.sum()) * steps / wall, wall_s=wall, n_traj=n,
                beads=int(N.sum()), status_ok=int(np.sum(st == 0)), tflops=fl / wall / 1e12,
                finite=bool(np.all(np.isfinite(Rf))))


# ---- HBM traffic ------------------------------------------------------------

COUNTERS_PROFILE = os.path.join(REPO, "profiles", "k_integrate_headline_counters.md")


def traffic_per_chain_step(form):
    """(read, write) HBM bytes per chain-step of k_integrate on this workload,
    from the copy-calibrated FETCH_SIZE / WRITE_SIZE counter runs recorded in
    profiles/k_integrate_headline_counters.md, or None where that is absent."""
    import re
    try:
        text = open(COUNTERS_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic-per-chain-step mixed read=(\d+) write=(\d+) fp64 read=(\d+) write=(\d+)",
                  text)
    if not m:
        return None
    v = [int(x) for x in m.groups()]
    return (v[0], v[1]) if form == "mixed" else (v[2], v[3])


# ---- serving ---------------------------------------------------------------

def serving_leg(n_jobs, step_scale):
    """The serving stream replayed in real time on the mixed form: Poisson
    arrivals at 50 jobs/s (seed 3), N log-uniform in [10, 500], 1..64
    trajectories per job, n_steps log-uniform in [1e3, 1e5] scaled by
    step_scale, Wi log-uniform in [0.1, 100], h* uniform in [0, 0.3], FENE
    b=100, Gaussian EV; latency = arrival to result on the host."""
    from bdsim_amd import serving as S
    jobs = S.c5_stream(n_jobs, step_scale=step_scale)
    print(f"bench: serving {n_jobs} jobs", file=sys.stderr, flush=True)
    with S.Server(A.PRECISION_MIXED, hbm_bytes=32 << 30, steps_per_launch=256) as srv:
        rep = S.run_stream(srv, jobs, timeout_s=900.0,
                           progress=lambda k, n: print(f"bench: serving {k}/{n} jobs done",
                                                       file=sys.stderr, flush=True))
    bad = sum(int(np.sum(r.status != 0)) for r in rep["results"].values())
    return {"p50_s": rep["p50_s"], "p99_s": rep["p99_s"], "makespan_s": rep["makespan_s"],
            "bead_steps": rep["bead_steps"], "bead_steps_per_s": rep["bead_steps_per_s"],
            "failed_trajectories": bad, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
            "sample": (f"first {n_jobs} jobs of the serving stream (Poisson 50 jobs/s, seed 3; "
                       f"N log-uniform [10,500], 1..64 trajectories, Wi log-uniform [0.1,100], "
                       f"h* uniform [0,0.3], FENE b=100, Gauss EV, dt=1e-3) with n_steps "
                       f"log-uniform [1e3,1e5]"
                       + ("" if step_scale == 1.0 else f" scaled by {step_scale:g}")
                       + "; mixed form, arrivals replayed in real time, latency from arrival to "
                         "host result")}


# ---- CPU baseline ----------------------------------------------------------

def _cpu_worker(args):
    kind, idx, steps = args
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import helpers as H
    lib = H.ref() if kind == "reference" else H.oracle()
    phys = headline_phys()
    R0 = start_configs(max(idx) + 1, N_BEADS) if idx else None
    t0 = time.perf_counter()
    for i in idx:
        phys.flow.kappa0[1] = shear_rate(WI[i % len(WI)])
        R = np.ascontiguousarray(R0[i])
        st = A.CRng(1, -1, i)
        sim = A.CSim(0.0, (steps - 1) * 1e-3, 1e-3, 1e-3, 1)
        if kind == "reference":
            lib.ref_integrate(N_BEADS, A.dptr(R), byref(phys), byref(sim), byref(st))
        else:
            cnt = A.CCounters()
            lib.bo_integrate(N_BEADS, A.dptr(R), byref(phys), byref(sim), byref(st), byref(cnt),
                             None, 0)
    return time.perf_counter() - t0


def host_cpu():
    """(nproc, CPUs this process may run on, model name) of the host."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return os.cpu_count() or 1, usable, model


# Worker processes the GPU host lets one job run beside its GPU (its CPU share).
HOST_CPU_SHARE = 16


def cpu_baseline(n_traj_sample=768, steps=400):
    from concurrent.futures import ProcessPoolExecutor
    kind = "reference" if os.path.exists(os.path.join(REPO, "oracle", "_ref", "libbdref.so")) else "port"
    if kind == "port" and not os.path.exists(os.path.join(REPO, "oracle", "_build", "liboracle.so")):
        return None
    nproc, usable, model = host_cpu()
    cores = min(usable, HOST_CPU_SHARE, n_traj_sample)
    chunks = [list(range(k, n_traj_sample, cores)) for k in range(cores)]
    t0 = time.perf_counter()
    with ProcessPoolExecutor(max_workers=cores) as ex:
        busy = list(ex.map(_cpu_worker, [(kind, c, steps) for c in chunks]))
    wall = time.perf_counter() - t0
    bead_steps = n_traj_sample * N_BEADS * steps
    return {"value": bead_steps / wall, "unit": "bead-steps/s", "cores": cores, "kind": kind,
            "host": {"nproc": nproc, "usable_cpus": usable, "model": model,
                     "note": (f"one process per core on {cores} cores: every CPU this process "
                              f"may use, up to the GPU host's share of {HOST_CPU_SHARE} per GPU")},
            "sample": (f"{n_traj_sample} of the 4096 chains (same physics, seeds, starts), "
                       f"{steps} steps each in one integrate call, one process per core; "
                       f"{sum(busy):.1f} core-seconds of CPU work, wall {wall:.2f} s"
                       + ("" if kind == "reference" else
                          "; FP64 C restatement (reference build absent)"))}


def run_c1(precision):
    """The CPU-sized case on the GPU: 16 chains resident, 10 segments of 100
    steps, Rg^2 / Q^2 / tau_xy evaluated on the device after each segment; the
    time covers the whole run including the observables' copies. Beside it the
    same 16 x 10 segments integrated on one host core by the reference build
    (or the FP64 restatement)."""
    from bdsim_amd.resident import ResidentEnsemble
    from bdsim_amd.workloads import c1_workload
    R0, phys, sim, nseg, per = c1_workload()
    n, N = R0.shape[0], R0.shape[1]
    t = time.perf_counter()
    with ResidentEnsemble(R0, phys, sim, list(range(n)), precision) as ens:
        series = []
        for _ in range(nseg):
            ens.run(per)
            o = ens.observe(force="total")
            series.append((o["rg2"].mean(), o["q2"].mean(), o["stress"][:, 0, 1].mean()))
        _, _, status, _ = ens.fetch()
    wall = time.perf_counter() - t
    out = {"bead_steps_per_s": n * N * nseg * per / wall, "wall_s": wall, "n_traj": n,
           "status_ok": int(np.sum(status == 0)),
           "last_snapshot": {"rg2": float(series[-1][0]), "q2": float(series[-1][1]),
                             "tau_xy": float(series[-1][2])}}
    return out


def c1_cpu():
    """16 x 10 segments of 100 steps on one host core (reference build, else
    the FP64 restatement), from the same starts and seeds."""
    from bdsim_amd.workloads import c1_workload
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import helpers as H
    kind = "reference" if os.path.exists(H.REF_PATH) else "port"
    lib = H.ref() if kind == "reference" else H.oracle()
    R0, phys, sim, nseg, per = c1_workload()
    pk = A.PackedPhys(phys, R0.shape[1])
    n, N = R0.shape[0], R0.shape[1]
    t = time.perf_counter()
    for i in range(n):
        R = np.ascontiguousarray(R0[i]).copy()
        st = A.CRng(1, -1, i)
        for k in range(nseg):
            t0 = k * per * sim.dt
            s = A.CSim(t0, t0 + (per - 1) * sim.dt, sim.dt, sim.implicit_loop_tol, 1)
            if kind == "reference":
                lib.ref_integrate(N, A.dptr(R), byref(pk.c), byref(s), byref(st))
            else:
                cnt = A.CCounters()
                lib.bo_integrate(N, A.dptr(R), byref(pk.c), byref(s), byref(st), byref(cnt), None, 0)
    wall = time.perf_counter() - t
    return {"value": n * N * nseg * per / wall, "unit": "bead-steps/s", "cores": 1, "kind": kind,
            "sample": "the whole case: 16 chains x 10 segments of 100 steps, one process"}


def _batch_positions(lib, precision, R0, K, steps):
    """Positions of the headline chains after `steps` steps of one form."""
    phys = headline_phys()
    n_traj = R0.shape[0]
    rng = (A.CRng * n_traj)(*[A.CRng(1, -1, i) for i in range(n_traj)])
    b = ctypes.c_void_p()
    A.check(lib, lib.bdsim_batch_create(byref(b), n_traj, N_BEADS, A.dptr(R0), byref(phys),
                                        A.dptr(K), rng, precision))
    try:
        sim = A.CSim(0.0, (steps - 1) * 1e-3, 1e-3, 1e-3, 1)
        A.check(lib, lib.bdsim_batch_run(b, byref(sim)))
        A.check(lib, lib.bdsim_batch_sync(b))
        status = np.zeros(n_traj, dtype=np.int32)
        R = np.empty_like(R0)
        A.check(lib, lib.bdsim_batch_fetch(b, A.dptr(R), None, A.iptr(status), None))
    finally:
        lib.bdsim_batch_destroy(b)
    assert np.all(status == 0)
    return R


def mixed_errors(lib, n_traj=256, steps=100):
    """Relative errors of the mixed form against the FP64 form (which matches
    the FP64 CPU restatement to 1e-10 over 100 steps) on the first n_traj
    headline chains: positions after `steps` steps from the same start and
    generator streams, and, at the FP64 positions reached, one step's D.F
    (F the total force) and B.dW (dW standard normal, the Chebyshev series on
    cold bounds). Per chain |x_mixed - x_fp64| / |x_fp64|; max and median."""
    R0 = start_configs(n_traj, N_BEADS)
    K = kappas(n_traj)
    Rf = _batch_positions(lib, A.PRECISION_FP64, R0, K, steps)
    Rm = _batch_positions(lib, A.PRECISION_MIXED, R0, K, steps)
    phys = headline_phys()
    R = np.ascontiguousarray(Rf)
    F = np.empty_like(R)
    A.check(lib, lib.bdsim_total_force_batch(n_traj, N_BEADS, A.dptr(R), byref(phys), A.dptr(F)))
    dW = np.ascontiguousarray(np.random.default_rng(7).normal(size=R.shape))
    out = {}
    for form, prec in (("fp64", A.PRECISION_FP64), ("mixed", A.PRECISION_MIXED)):
        DF = np.empty_like(R)
        A.check(lib, lib.bdsim_rpy_apply_batch(n_traj, N_BEADS, phys.hstar, prec, A.dptr(R),
                                               A.dptr(F), A.dptr(DF)))
        BW = np.empty_like(R)
        st = np.zeros(n_traj, dtype=np.int32)
        A.check(lib, lib.bdsim_brownian_batch(n_traj, N_BEADS, A.dptr(R), byref(phys), prec,
                                              A.dptr(dW), A.dptr(BW), A.iptr(st)))
        assert np.all(st == 0)
        out[form] = (DF, BW)

    def summary(a, b):
        e = np.linalg.norm((a - b).reshape(n_traj, -1), axis=1) / \
            np.linalg.norm(b.reshape(n_traj, -1), axis=1)
        return {"max": float(e.max()), "median": float(np.median(e))}

    return {"sample": f"first {n_traj} headline chains; positions after {steps} steps",
            "against": "fp64 form",
            "D.F": summary(out["mixed"][0], out["fp64"][0]),
            "B.dW": summary(out["mixed"][1], out["fp64"][1]),
            "positions": summary(Rm, Rf),
            "decision_flips": decision_flips(lib, R0, K, steps)}


def decision_flips(lib, R0, K, steps):
    """Per chain-step discrete decisions of the two forms on the same chains
    (same starts, seeds and steps): Chebyshev degree L, FD attempts, corrector
    passes, Cholesky fallbacks and Krylov dimension -- counted where the mixed
    form's differs from the FP64 form's -- and the largest relative difference
    of the Chebyshev interval ends lmin, lmax (the Ritz bounds they come from
    are multisected in FP32 in the mixed form)."""
    n_traj = R0.shape[0]
    logs = {}
    for form, prec in (("fp64", A.PRECISION_FP64), ("mixed", A.PRECISION_MIXED)):
        ph = []
        for i in range(n_traj):
            c = headline_phys()
            c.flow.kappa0[1] = K[i, 1]
            ph.append(c)
        nb = np.full(n_traj, N_BEADS, dtype=np.int32)
        R = np.array(R0, dtype=np.float64).reshape(-1)       # a copy: the call writes R back
        sims = (A.CSim * n_traj)(*[A.CSim(0.0, (steps - 1) * 1e-3, 1e-3, 1e-3, 1)] * n_traj)
        rng = (A.CRng * n_traj)(*[A.CRng(1, -1, i) for i in range(n_traj)])
        st = np.zeros(n_traj, dtype=np.int32)
        lg = (A.CStepLog * (n_traj * steps))()
        A.check(lib, lib.bdsim_integrate_batch(n_traj, A.iptr(nb), A.dptr(R), (A.CPhys * n_traj)(*ph),
                                               sims, rng, prec, A.iptr(st), None, lg, steps))
        assert np.all(st == 0)
        logs[form] = np.array([(l.cheb_L, l.cheb_attempts, l.corrector_iters, l.fallback,
                                l.lanczos_steps, l.lmin, l.lmax) for l in lg]).reshape(n_traj, steps, 7)
    a, b = logs["mixed"], logs["fp64"]
    out = {"chain_steps": int(n_traj * steps)}
    for k, name in enumerate(("cheb_L", "fd_attempts", "corrector_passes", "fallbacks", "lanczos_dim")):
        out[name] = int(np.sum(a[:, :, k] != b[:, :, k]))
    out["chains_with_any_flip"] = int(np.sum(np.any(a[:, :, :5] != b[:, :, :5], axis=(1, 2))))
    out["bounds_rel_diff_max"] = float(np.max(np.abs(a[:, :, 5:] - b[:, :, 5:]) / np.abs(b[:, :, 5:])))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--traj", type=int, default=N_TRAJ)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="mixed,fp64")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-jobs", type=int, default=40)
    ap.add_argument("--serving-step-scale", type=float, default=1.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device copy of BYTES (counter calibration) and exit")
    ap.add_argument("--per-step-launches", type=int, default=0, metavar="K",
                    help="only time the first K steps of a headline segment, one launch per "
                         "step, and print one line per form")
    ap.add_argument("--workload", choices=["headline", "uneven", "long", "cholesky", "c1"],
                    default="headline",
                    help="uneven: 4096 chains with N log-uniform in [10,256]; "
                         "long: 64 chains of N=1000; cholesky: 8192 chains of N=30 with the "
                         "dense Cholesky square root; c1: 16 chains of N=10 in 10 segments "
                         "with snapshot observables (each prints its own line)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.workload != "headline":
        if args.gpus != 1:
            print("bench.py runs on one GPU only", file=sys.stderr)
            sys.exit(2)
        lib = A.load_gpu()
        out = {"workload": args.workload, "steps": args.steps, "warmup": args.warmup}
        forms = args.forms.split(",")
        # one untimed pass of every form first: the first integrate call of a
        # precision form in a process measured up to 15 % slow even after its
        # warm-up segment, which would favour whichever form ran first
        for form in forms:
            print(f"bench: {args.workload} {form} (untimed pass)", file=sys.stderr, flush=True)
            if args.workload == "c1":
                run_c1(PRECISION[form])
            else:
                run_workload(lib, args.workload, PRECISION[form], args.steps, args.warmup)
        for form in forms:
            prec = PRECISION[form]
            print(f"bench: {args.workload} {form}", file=sys.stderr, flush=True)
            if args.workload == "c1":
                out[form] = run_c1(prec)
            else:
                out[form] = run_workload(lib, args.workload, prec, args.steps, args.warmup)
        if args.workload == "c1":
            out["steps"], out["warmup"] = 1000, 0
            if not args.no_cpu:
                print("bench: c1 cpu", file=sys.stderr, flush=True)
                out["cpu_baseline"] = c1_cpu()
        print(json.dumps(out), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path "
              f"(trajectories would shard by index across GPUs). Got --gpus {args.gpus}.",
              file=sys.stderr)
        sys.exit(2)
    lib = A.load_gpu()
    ev = HipEvents()
    R0 = start_configs(args.traj, N_BEADS)
    K = kappas(args.traj)
    if args.per_step_launches:
        for form in args.forms.split(","):
            print(f"bench: {form} form, one launch per step", file=sys.stderr, flush=True)
            print(json.dumps({"form": form, "per_step": per_step_launches(
                lib, PRECISION[form], args.per_step_launches, R0, K, ev)}), flush=True)
        return
    res = {}
    for form in args.forms.split(","):
        prec = PRECISION[form]
        print(f"bench: {form} form, {args.traj} chains, warmup {args.warmup} + {args.steps} steps",
              file=sys.stderr, flush=True)
        res[form] = run_gpu(lib, prec, args.steps, args.warmup, R0, K, ev)
        print(f"bench: {form}: {json.dumps(res[form])}", file=sys.stderr, flush=True)
        assert res[form]["steps_done"] == args.steps, res[form]
    fresh = {}
    for form in args.forms.split(","):
        print(f"bench: {form} form, a fresh {args.steps}-step segment", file=sys.stderr, flush=True)
        f = run_gpu(lib, PRECISION[form], args.steps, args.warmup, R0, K, ev, fresh=True)
        fresh[form] = {"value": args.traj * N_BEADS * args.steps / (f["kernel_ms"] / 1e3),
                       "ms_per_step": f["kernel_ms"] / args.steps,
                       "matvecs_per_step": f["matvecs_per_step"],
                       "lanczos_cold": f["lanczos_cold"]}
    main_form = "mixed" if "mixed" in res else "fp64"
    r = res[main_form]
    bead_steps = args.traj * N_BEADS * args.steps
    value = bead_steps / (r["kernel_ms"] / 1e3)
    achieved = r["flops"] / (r["kernel_ms"] / 1e3) / 1e12
    cpu = None if args.no_cpu else cpu_baseline()
    traffic = traffic_per_chain_step(main_form)
    out = {
        "metric": "bead-steps/s 4096xN=50 RPY-Chebyshev, mixed vs FP64 GPU; p99 job latency",
        "value": value,
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["kernel_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_form == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.traj} chains x N=50 FENE b=100, RPY h*=0.3 Chebyshev, "
                               "SDK EV, shear Wi {0.1,1,10,100}/tau_1 round robin, dt=1e-3",
                   "segment": (f"each chain integrates one segment of {args.warmup}+{args.steps} "
                               f"steps (one integrate call, cold Lanczos start in the warm-up); "
                               f"the {args.steps} timed steps continue it"),
                   "form": main_form, "n_traj": args.traj, "n_beads": N_BEADS},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[main_form],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[main_form],
                     "traffic": (None if traffic is None else
                                 float(sum(traffic)) * args.traj * args.steps),
                     "traffic_note": ("HBM bytes of the timed launch: per chain-step read + "
                                      "write from copy-calibrated FETCH_SIZE (x2.0) / "
                                      "WRITE_SIZE (x1.0) runs, "
                                      "profiles/k_integrate_headline_counters.md")},
        "cpu_baseline": cpu,
    }
    if "fp64" in res and main_form != "fp64":
        f = res["fp64"]
        out["fp64"] = {"value": bead_steps / (f["kernel_ms"] / 1e3),
                       "ms_per_step": f["kernel_ms"] / args.steps,
                       "achieved_tflops": f["flops"] / (f["kernel_ms"] / 1e3) / 1e12,
                       "frac_of_78.6": f["flops"] / (f["kernel_ms"] / 1e3) / 1e12 / 78.6}
        out["mixed_over_fp64"] = out["value"] / out["fp64"]["value"]
    out["fresh_segment"] = dict(fresh, note=(f"the {args.steps} timed steps as a segment of their "
                                             "own (a separate integrate call after the warm-up): "
                                             "the timed region includes that segment's cold start"))
    if "mixed" in fresh and "fp64" in fresh:
        out["fresh_segment"]["mixed_over_fp64"] = fresh["mixed"]["value"] / fresh["fp64"]["value"]
    if "fp64" in res and "mixed" in res:
        print("bench: mixed-form errors", file=sys.stderr, flush=True)
        out["mixed_errors"] = mixed_errors(lib)
    out["counters"] = {k: {kk: v[kk] for kk in ("matvecs_per_step", "corrector_per_step",
                                                 "lanczos_cold", "fallbacks", "status_ok",
                                                 "finite", "wall_s")} for k, v in res.items()}
    srv = None if args.no_serving else serving_leg(args.serving_jobs, args.serving_step_scale)
    out["p50_job_latency_s"] = srv["p50_s"] if srv else None
    out["p99_job_latency_s"] = srv["p99_s"] if srv else None
    out["serving"] = srv
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
