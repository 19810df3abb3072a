"""The serving loop on the GPU: a job's result is the same, bit for bit, as
integrating its trajectories in one bdsim_integrate_batch call -- when it runs
alone, when it shares the resident batch with jobs of other chain lengths
(and is cut into many short launches), and when memory admission makes it wait
in the queue."""
import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from bdsim_amd import serving as S
from bdsim_amd.core import DelSMethod, Rng, SimParams

pytestmark = pytest.mark.gpu

# (N, n_traj, steps, h*, Wi, t0[, method]): buckets of 64..512 threads, free
# draining, a segment that does not start at 0, and Cholesky chains on both
# dense LDS kernels (3N <= 64, 3N <= 96) and beyond them (workspace factor)
SPECS = [(10, 3, 57, 0.3, 1.0, 0.0), (50, 4, 120, 0.3, 10.0, 0.25), (64, 2, 33, 0.0, 0.5, 0.0),
         (65, 2, 40, 0.25, 3.0, 0.0), (130, 2, 25, 0.2, 1.0, 1.0), (300, 1, 9, 0.3, 2.0, 0.0),
         (12, 3, 45, 0.25, 1.0, 0.0, DelSMethod.Cholesky), (30, 2, 38, 0.25, 2.0, 0.5, DelSMethod.Cholesky),
         (40, 1, 20, 0.2, 1.0, 0.0, DelSMethod.Cholesky)]


def make_job(j, N, n_traj, steps, hstar, wi, t0=0.0, method=None):
    rs = np.random.default_rng(77 + j)
    R0 = np.empty((n_traj, N, 3))
    for k in range(n_traj):
        R0[k] = np.concatenate([np.zeros((1, 3)), np.cumsum(rs.normal(size=(N - 1, 3)), axis=0)])
    sim = SimParams()
    sim.time_start, sim.dt = t0, 1e-3
    sim.time_end = t0 + (steps - 1) * 1e-3
    phys = S.c5_job_physics(N, wi, hstar)
    if method is not None:
        phys.hi_method = method
    return S.Job(j, R0, phys, sim, [Rng(100 * j + k) for k in range(n_traj)])


def reference(gpu, job, precision):
    packed = A.PackedPhys(job.phys, job.N)
    sim = A.pack_sim(job.sim)
    items = [(job.N, job.R0[k], packed.c, sim, A.CRng(*job.rngs[k]._state()))
             for k in range(job.n_traj)]
    out, rngs, status, _, _ = H.gpu_batch(gpu, items, precision=precision)
    return np.array(out), [(r.ix, r.iy, r.seed) for r in rngs], status


def same(res, ref):
    R, rng, st = ref
    assert np.array_equal(res.status, st)
    assert res.rng_states == rng
    assert np.array_equal(res.R, R)


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
def test_job_alone_and_in_the_mix_equals_integrate(gpu, precision):
    jobs = [make_job(j, *s) for j, s in enumerate(SPECS)]
    refs = {j.job_id: reference(gpu, j, precision) for j in jobs}
    with S.Server(precision, hbm_bytes=1 << 30, steps_per_launch=256) as srv:
        for job in jobs:
            out = S.serve(srv, [job])
            same(out[job.job_id], refs[job.job_id])
    with S.Server(precision, hbm_bytes=1 << 30, steps_per_launch=13) as srv:
        out = S.serve(srv, jobs)
        for job in jobs:
            same(out[job.job_id], refs[job.job_id])
        assert srv.stats()["running"] == 0


def test_memory_admission_queues_and_keeps_results(gpu):
    """A 64 MiB budget holds the workspaces of two 4-chain N=300 jobs: the
    third waits in the queue until one completes; every result is unchanged."""
    big = [make_job(10 + i, 300, 4, 12, 0.3, 1.0) for i in range(3)]
    small = [make_job(20 + i, 20, 8, 200, 0.2, 5.0) for i in range(3)]
    refs = {j.job_id: reference(gpu, j, A.PRECISION_MIXED) for j in big + small}
    with S.Server(A.PRECISION_MIXED, hbm_bytes=64 << 20, steps_per_launch=32) as srv:
        for j in big + small:
            srv.submit(j)
        srv.poll()
        st = srv.stats()
        assert st["queued"] >= 1 and st["running"] >= 1, st
        assert st["bytes_in_use"] <= st["bytes_reserved"]
        out = {}
        for _ in range(100000):
            srv.poll(wait_us=2000)
            for i in srv.completed():
                out[i] = srv.result(i)
            if len(out) == len(big) + len(small):
                break
        for j in big + small:
            same(out[j.job_id], refs[j.job_id])
        assert srv.stats()["bytes_in_use"] == 0


def test_submit_rejects_what_cannot_run(gpu):
    with S.Server(A.PRECISION_MIXED, hbm_bytes=64 << 20, steps_per_launch=32) as srv:
        with pytest.raises(RuntimeError, match="memory budget"):
            srv.submit(make_job(1, 500, 64, 10, 0.3, 1.0))
        bad = make_job(2, 10, 1, 10, 0.3, 1.0)
        bad.sim.dt = 0.0
        with pytest.raises(ValueError, match="dt"):     # the host-side clock check
            srv.submit(bad)
        ok = make_job(3, 10, 1, 10, 0.3, 1.0)
        srv.submit(ok)
        with pytest.raises(RuntimeError, match="already in use"):
            srv.submit(ok)
        out = S.serve(srv, [])
        assert out == {}


def test_stream_replay_reports_latency(gpu):
    jobs = S.c5_stream(6, step_scale=0.01)
    with S.Server(A.PRECISION_MIXED, hbm_bytes=4 << 30, steps_per_launch=128) as srv:
        rep = S.run_stream(srv, jobs)
    assert len(rep["latency_s"]) == 6 and min(rep["latency_s"]) > 0.0
    assert rep["p50_s"] <= rep["p99_s"]
    want = sum(j.bead_steps() for j in jobs)
    assert rep["bead_steps"] == want
    assert all(np.all(r.status == 0) for r in rep["results"].values())


def test_served_fp64_jobs_match_oracle(gpu, orc):
    """Jobs served together in the resident batch (FP64 form), cut into short
    launches, against the CPU oracle integrating each trajectory's segment:
    positions within 1e-10 relative, the same generator states."""
    jobs = [make_job(j, *s) for j, s in enumerate(SPECS) if s[0] <= 65]
    with S.Server(A.PRECISION_FP64, hbm_bytes=1 << 30, steps_per_launch=16) as srv:
        got = S.serve(srv, jobs)
    for job in jobs:
        res = got[job.job_id]
        pk = A.PackedPhys(job.phys, job.N)
        for k in range(job.n_traj):
            st = A.CRng(*job.rngs[k]._state())
            want, rc, _, _ = H.oracle_integrate(orc, job.N, job.R0[k], pk.c, A.pack_sim(job.sim), st)
            assert rc == 0 and res.status[k] == 0
            err = np.max(np.abs(res.R[k] - want)) / np.max(np.abs(want))
            assert err <= 1e-10, (job.job_id, k, err)
            assert res.rng_states[k] == (st.ix, st.iy, st.seed)


@pytest.mark.parametrize("precision", [A.PRECISION_FP64, A.PRECISION_MIXED])
def test_long_job_of_long_chains_served_split(gpu, precision):
    """A long job of few long chains (2 x N=300, 2000 steps) is held with a
    workspace copy per part and, on an otherwise idle device, launched split
    over 5 workgroups per chain (k_integrate_split, resumed launch after
    launch, the warm seed handed to every part's copy before each). The split
    sums partner terms in another order, so the result equals the unsplit
    integrate call within rounding: 1e-9 relative (FP64), 1e-6 (mixed), the
    same generator states."""
    job = make_job(60, 300, 2, 2000, 0.25, 2.0)
    R, rng, st = reference(gpu, job, precision)
    with S.Server(precision, hbm_bytes=1 << 30, steps_per_launch=256) as srv:
        out = S.serve(srv, [job])
    res = out[job.job_id]
    assert np.all(res.status == 0) and np.all(st == 0)
    assert res.rng_states == rng
    err = np.max(np.abs(res.R - R)) / np.max(np.abs(R))
    assert err < (1e-9 if precision == A.PRECISION_FP64 else 1e-6), err
