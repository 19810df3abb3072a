"""Serving: continuous batching of integrate jobs on one GPU.

A job is `n_traj` trajectories of one chain length integrating ONE segment
(SimParams) each, from their own starting chains and Rng states: exactly what
`integrate` does for each of them, and what `ensemble.run_ensemble` asks of
its per-trajectory calls. The library's server (bdsim_server_* in
include/bdsim_gpu.h) keeps every admitted trajectory resident in HBM, admits
queued jobs as their memory footprint fits, and runs each chain-length bucket
as a sequence of persistent-kernel launches of a bounded number of steps on
its own stream, so jobs that arrive while others run join at the next launch
of their bucket. A trajectory's segment is resumed launch after launch from
its stored time, warm Lanczos seed and generator state, which makes the
result of a job identical, bit for bit, to integrating it alone.

`c5_stream` draws the Poisson job stream of the serving workload and
`run_stream` replays it in real time, reporting p50/p99 completion latency
(arrival to result on the host) and bead-steps/s.
"""
import ctypes
import time
from ctypes import byref
from dataclasses import dataclass, field

import numpy as np

from . import _abi, flows
from .core import EV, PhysParams, Rng, SimParams, Spring


@dataclass
class Job:
    """One serving request: n_traj chains of N beads, one segment each."""
    job_id: int
    R0: np.ndarray                    # (n_traj, N, 3)
    phys: PhysParams
    sim: SimParams
    rngs: list                        # n_traj Rng objects (states are read, not advanced)
    arrival: float = 0.0              # seconds after the stream starts
    meta: dict = field(default_factory=dict)

    @property
    def n_traj(self):
        return self.R0.shape[0]

    @property
    def N(self):
        return self.R0.shape[1]

    def bead_steps(self):
        steps = int(np.floor((self.sim.time_end - self.sim.time_start) / self.sim.dt + 0.5)) + 1
        return self.n_traj * self.N * steps


@dataclass
class JobResult:
    R: np.ndarray                     # (n_traj, N, 3)
    rng_states: list                  # (ix, iy, seed) per trajectory
    status: np.ndarray                # BDSIM_STATUS_* per trajectory
    counters: list


class Server:
    """The resident batch: submit jobs, poll, take results.

    hbm_bytes is reserved once for trajectory slots, positions, workspaces and
    tables; steps_per_launch bounds a launch of 64-thread chains (longer
    chains take proportionally fewer steps per launch)."""

    def __init__(self, precision=_abi.PRECISION_MIXED, hbm_bytes=8 << 30, steps_per_launch=256):
        self.lib = _abi.load_gpu()
        self.h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.bdsim_server_create(byref(self.h), int(precision),
                                                          int(hbm_bytes), int(steps_per_launch)))
        self._jobs = {}

    def submit(self, job: Job):
        R0 = np.ascontiguousarray(job.R0, dtype=np.float64)
        if R0.ndim != 3 or R0.shape[2] != 3:
            raise ValueError("R0 must be (n_traj, N, 3)")
        n, N = R0.shape[0], R0.shape[1]
        if len(job.rngs) != n:
            raise ValueError("one Rng per trajectory")
        packed = _abi.PackedPhys(job.phys, N)
        sim = _abi.pack_sim(job.sim)
        rng = (_abi.CRng * n)(*[_abi.CRng(*r._state()) for r in job.rngs])
        _abi.check(self.lib, self.lib.bdsim_server_submit(self.h, int(job.job_id), n, N,
                                                          _abi.dptr(R0), byref(packed.c),
                                                          byref(sim), rng))
        self._jobs[int(job.job_id)] = (n, N)

    def poll(self, wait_us=0):
        rc = self.lib.bdsim_server_poll(self.h, int(wait_us))
        if rc < 0:
            _abi.check(self.lib, rc)
        return rc

    def completed(self):
        out = []
        buf = (ctypes.c_int64 * 256)()
        while True:
            k = self.lib.bdsim_server_completed(self.h, buf, 256)
            if k < 0:
                _abi.check(self.lib, k)
            out += [int(buf[i]) for i in range(k)]
            if k < 256:
                return out

    def result(self, job_id) -> JobResult:
        n, N = self._jobs.pop(int(job_id))
        R = np.empty((n, N, 3))
        rng = (_abi.CRng * n)()
        st = np.empty(n, dtype=np.int32)
        cnt = (_abi.CCounters * n)()
        _abi.check(self.lib, self.lib.bdsim_server_result(self.h, int(job_id), _abi.dptr(R), rng,
                                                          _abi.iptr(st), cnt))
        return JobResult(R, [(r.ix, r.iy, r.seed) for r in rng], st, list(cnt))

    def stats(self):
        s = _abi.CServerStats()
        _abi.check(self.lib, self.lib.bdsim_server_query(self.h, byref(s)))
        return {k: int(getattr(s, k)) for k, _ in s._fields_}

    def close(self):
        if self.h:
            self.lib.bdsim_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serve(server: Server, jobs, timeout_s=600.0):
    """Submit every job at once and wait for all: {job_id: JobResult}."""
    for j in jobs:
        server.submit(j)
    out = {}
    t_end = time.monotonic() + timeout_s
    while len(out) < len(jobs):
        if time.monotonic() > t_end:
            raise TimeoutError(f"{len(jobs) - len(out)} jobs still running after {timeout_s} s")
        server.poll(wait_us=2000)
        for i in server.completed():
            out[i] = server.result(i)
    return out


# ---- the serving workload --------------------------------------------------

def c5_job_physics(N, wi, hstar):
    """FENE b=100, Gaussian EV z*=0.5 d*=0.5, RPY h* with the Chebyshev square
    root (free draining at h* = 0), simple shear at Weissenberg number wi
    (rate wi / tau_1, tau_1 = 1 / sin^2(pi / 2N))."""
    p = PhysParams()
    p.number_of_beads = N
    p.spring.type = Spring.FENE
    p.spring.sqrtb = 10.0
    p.ev.type = EV.Gauss
    p.ev.zstar, p.ev.dstar = 0.5, 0.5
    p.hstar = float(hstar)
    p.flow = flows.shear(wi * np.sin(np.pi / (2.0 * N)) ** 2)
    return p


def c5_stream(n_jobs, rate=50.0, seed=3, step_scale=1.0):
    """The first n_jobs of the serving stream: Poisson arrivals at `rate` per
    second; N log-uniform in [10, 500], n_traj uniform in 1..64, n_steps
    log-uniform in [1e3, 1e5] (times step_scale, at least 1), Wi log-uniform in
    [0.1, 100], h* uniform in [0, 0.3]; dt = 1e-3. Trajectory k of job j starts
    from a Gaussian random walk (bond std 1) and Rng(seed=1000 * j + k)."""
    rs = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for j in range(n_jobs):
        t += rs.exponential(1.0 / rate)
        N = int(np.clip(np.rint(np.exp(rs.uniform(np.log(10), np.log(500)))), 10, 500))
        n_traj = int(rs.integers(1, 65))
        n_steps = int(np.rint(np.exp(rs.uniform(np.log(1e3), np.log(1e5)))))
        n_steps = max(1, int(np.rint(n_steps * step_scale)))
        wi = float(np.exp(rs.uniform(np.log(0.1), np.log(100.0))))
        hstar = float(rs.uniform(0.0, 0.3))
        R0 = np.empty((n_traj, N, 3))
        for k in range(n_traj):
            b = rs.normal(size=(N - 1, 3))
            R0[k, 0] = 0.0
            R0[k, 1:] = np.cumsum(b, axis=0)
            R0[k] -= R0[k].mean(axis=0)
        sim = SimParams()
        sim.time_start, sim.dt = 0.0, 1e-3
        sim.time_end = (n_steps - 1) * sim.dt
        jobs.append(Job(j, R0, c5_job_physics(N, wi, hstar), sim,
                        [Rng(1000 * j + k) for k in range(n_traj)], arrival=t,
                        meta=dict(N=N, n_traj=n_traj, n_steps=n_steps, wi=wi, hstar=hstar)))
    return jobs


def run_stream(server: Server, jobs, timeout_s=600.0, progress=None):
    """Replay `jobs` in real time: each is submitted when its arrival time has
    passed; latency is the time from arrival to its result on the host.
    `progress(done, total)` is called as jobs complete."""
    jobs = sorted(jobs, key=lambda j: j.arrival)
    by_id = {j.job_id: j for j in jobs}
    lat = {}
    results = {}
    t0 = time.monotonic()
    nxt = 0
    while len(lat) < len(jobs):
        now = time.monotonic() - t0
        if now > timeout_s:
            raise TimeoutError(f"{len(jobs) - len(lat)} jobs unfinished after {timeout_s} s")
        while nxt < len(jobs) and jobs[nxt].arrival <= now:
            server.submit(jobs[nxt])
            nxt += 1
        wait = 1000 if nxt == len(jobs) else max(0, min(1000, int((jobs[nxt].arrival - now) * 1e6)))
        server.poll(wait_us=wait)
        for i in server.completed():
            results[i] = server.result(i)
            lat[i] = time.monotonic() - t0 - by_id[i].arrival
            if progress:
                progress(len(lat), len(jobs))
    wall = time.monotonic() - t0
    L = np.array([lat[j.job_id] for j in jobs])
    bead_steps = sum(int(c.steps) * by_id[i].N for i, r in results.items() for c in r.counters)
    return dict(latency_s=L.tolist(), p50_s=float(np.percentile(L, 50)),
                p99_s=float(np.percentile(L, 99)), makespan_s=wall,
                bead_steps=int(bead_steps), bead_steps_per_s=bead_steps / wall,
                results=results)
