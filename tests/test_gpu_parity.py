"""GPU parity: the HIP library (through its C ABI) against the CPU oracle and
the golden fixtures written from the reference core.

Bars: the RNG stream is bit-exact; forces within 1e-12 relative (the
reference's own check is 1e-6); D.x within 1e-13 (FP64) / 1e-5 (mixed);
FP64 positions within 1e-10 relative of the reference over 100 steps with
identical per-step decisions (cold/warm Lanczos, Chebyshev terms and attempts,
fallbacks, corrector passes); the mixed form within 1e-5 over 100 steps.
"""
import ctypes
from ctypes import byref

import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from bdsim_amd import workloads as W
from golden import cases as C

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


# --------------------------------------------------------------------------
# RNG
# --------------------------------------------------------------------------

@pytest.mark.parametrize("N", [1, 2, 10, 50, 64, 65, 129, 300, 1000, 1024])
def test_noise_bit_exact(gpu, orc, N):
    seeds = [0, 5, -3, 4095, 123456789]
    n_steps = 3
    st = (A.CRng * len(seeds))(*[H.rng_of(s) for s in seeds])
    X = np.empty(len(seeds) * n_steps * 3 * N)
    A.check(gpu, gpu.bdsim_raw_noise_batch(len(seeds), N, n_steps, st, A.dptr(X)))
    X = X.reshape(len(seeds), n_steps, 3 * N)
    for k, s in enumerate(seeds):
        so = H.rng_of(s)
        for t in range(n_steps):
            xo = np.empty(3 * N)
            orc.bo_raw_noise(byref(so), N, A.dptr(xo))
            assert np.array_equal(X[k, t], xo), (N, s, t)
        assert (st[k].ix, st[k].iy, st[k].seed) == (so.ix, so.iy, so.seed)


def test_noise_matches_golden_uniforms(gpu, golden):
    """Invert the Ottinger transform's input: draws 3*mu..3*mu+2 of step 0 are
    the golden uniforms (for seeds whose fixtures cover 3*N draws)."""
    N = 600
    for k, s in enumerate(golden["rng_seeds"]):
        st = (A.CRng * 1)(H.rng_of(int(s)))
        X = np.empty(3 * N)
        A.check(gpu, gpu.bdsim_raw_noise_batch(1, N, 1, st, A.dptr(X)))
        y = golden["rng_draws"][k][:3 * N].astype(np.float64) - 0.5
        assert np.array_equal(X, (14.14855378 * y * y + 1.21569221) * y)


# --------------------------------------------------------------------------
# forces and mobility
# --------------------------------------------------------------------------

def test_forces_match_golden(gpu, golden):
    import json
    names = sorted({k.split("/")[1] for k in golden.files if k.startswith("force/")})
    fn = {"spring": "bdsim_spring_force_batch", "ev": "bdsim_ev_force_batch",
          "bend": "bdsim_bending_force_batch", "total": "bdsim_total_force_batch"}
    for nm in names:
        d = json.loads(str(golden[f"force/{nm}/phys"]))
        if f"force/{nm}/phi" in golden.files:
            d["phi"] = golden[f"force/{nm}/phi"]
        R = np.ascontiguousarray(golden[f"force/{nm}/R"])
        p = H.Packed(d, R.shape[0])
        F = np.empty_like(R)
        A.check(gpu, getattr(gpu, fn[str(golden[f"force/{nm}/kind"])])(
            1, R.shape[0], A.dptr(R), byref(p.c), A.dptr(F)))
        want = golden[f"force/{nm}/F"]
        assert rel(F, want) < 1e-12, nm


def test_external_force_and_flow_tables(gpu, golden):
    d = C.phys_dict(external=((0, True, (0.0, 0.0, -3.0), (), ()),
                              (-1, False, (0, 0, 0), (0.0, 0.5, 1.0),
                               ((0.0, 0.0, 0.0), (0.0, 0.0, 6.0), (1.0, 2.0, 3.0))),
                              (3, True, (1.0, -2.0, 0.5), (), ()),
                              (3, True, (0.5, 0.5, 0.5), (), ()),
                              (17, True, (9.0, 9.0, 9.0), (), ())))
    p = H.Packed(d, 6)
    for i, t in enumerate(golden["external/t"]):
        F = np.empty((6, 3))
        A.check(gpu, gpu.bdsim_external_force_batch(byref(p.c), 6, float(t), A.dptr(F)))
        assert np.array_equal(F, golden["external/F"][i])


@pytest.mark.parametrize("N,h", [(2, 0.3), (10, 0.3), (50, 0.3), (64, 0.2), (100, 0.25), (300, 0.3)])
def test_rpy_apply(gpu, orc, N, h):
    R = np.ascontiguousarray(C.random_walk(N, 40 + N, 0.8))
    x = np.ascontiguousarray(np.random.default_rng(N).normal(size=(N, 3)))
    D = np.empty((3 * N, 3 * N))
    orc.bo_rpy_matrix(N, A.dptr(R), h, A.dptr(D))
    want = D @ x.ravel()
    for prec, tol in ((A.PRECISION_FP64, 1e-13), (A.PRECISION_MIXED, 1e-5)):
        y = np.empty_like(x)
        A.check(gpu, gpu.bdsim_rpy_apply_batch(1, N, h, prec, A.dptr(R), A.dptr(x), A.dptr(y)))
        assert rel(y.ravel(), want) < tol, (prec, rel(y.ravel(), want))


@pytest.mark.parametrize("N", [65, 130, 257, 500])
def test_rpy_apply_operand_scales(gpu, orc, N):
    """Multi-wave chains in the mixed form add the partner rows of other waves
    as fixed-point integers; the unit follows the operand's largest entry, so
    operands from 1e-12 to 1e8 keep the 1e-5 bar (and FP64 its 1e-13)."""
    R = np.ascontiguousarray(C.random_walk(N, 80 + N, 0.8))
    x0 = np.random.default_rng(N + 1).normal(size=(N, 3))
    x0[N // 3] *= 50.0                  # one large entry, the rest well below it
    D = np.empty((3 * N, 3 * N))
    orc.bo_rpy_matrix(N, A.dptr(R), 0.3, A.dptr(D))
    for s in (1e-12, 1e-6, 1.0, 1e8):
        x = np.ascontiguousarray(x0 * s)
        want = D @ x.ravel()
        for prec, tol in ((A.PRECISION_FP64, 1e-13), (A.PRECISION_MIXED, 1e-5)):
            y = np.empty_like(x)
            A.check(gpu, gpu.bdsim_rpy_apply_batch(1, N, 0.3, prec, A.dptr(R), A.dptr(x), A.dptr(y)))
            assert rel(y.ravel(), want) < tol, (s, prec, rel(y.ravel(), want))


# --------------------------------------------------------------------------
# the integrator
# --------------------------------------------------------------------------

def _golden_items(golden):
    cases = {c["name"]: c for c in C.trajectory_cases()}
    return [(cases[str(n)], golden) for n in golden["traj_names"]]


def test_trajectories_fp64_parity(gpu, orc, golden):
    """Every golden trajectory, all chains in ONE batched launch per segment:
    positions <= 1e-10 relative of the reference, identical RNG state and the
    same discrete decisions as the oracle, step by step."""
    cases = [c for c, _ in _golden_items(golden)]
    keep = [H.Packed(c["phys"], c["N"]) for c in cases]
    Rs = [golden[f"traj/{c['name']}/R0"] for c in cases]
    sts = [H.rng_of(c["seed"]) for c in cases]
    oracle_sts = [H.rng_of(c["seed"]) for c in cases]
    n_seg = max(c["segments"] for c in cases)
    worst = {}
    for seg in range(n_seg):
        idx = [i for i, c in enumerate(cases) if seg < c["segments"]]
        items = [(cases[i]["N"], Rs[i], keep[i].c, H.csim(cases[i]["sim"]), sts[i]) for i in idx]
        out, rngs, status, cnts, logs = H.gpu_batch(gpu, items, log_stride=200)
        for j, i in enumerate(idx):
            c = cases[i]
            name = c["name"]
            assert status[j] == golden[f"traj/{name}/rc"][seg], name
            sts[i] = rngs[j]
            Ro, rco, cnto, lo = H.oracle_integrate(orc, c["N"], Rs[i], keep[i].c,
                                                   H.csim(c["sim"]), oracle_sts[i], log_cap=200)
            assert rco == status[j]
            if status[j] != 0:
                continue
            want = golden[f"traj/{name}/R"][seg]
            r = rel(out[j], want)
            worst[c["N"]] = max(worst.get(c["N"], 0.0), r)
            assert r <= 1e-10, (name, seg, r)
            lg = [logs[j * 200 + k] for k in range(cnts[j].steps)]
            assert len(lg) == len(lo) == cnto.steps
            for k, (a, b) in enumerate(zip(lg, lo)):
                assert (a.lanczos_cold, a.lanczos_steps, a.cheb_L, a.cheb_attempts, a.fallback,
                        a.corrector_iters) == (b.lanczos_cold, b.lanczos_steps, b.cheb_L,
                                               b.cheb_attempts, b.fallback, b.corrector_iters), (name, k)
                assert a.t == b.t
            Rs[i] = out[j]
    for c, st in zip(cases, sts):
        assert (st.ix, st.iy, st.seed) == tuple(golden[f"traj/{c['name']}/rng"]), c["name"]
    print("worst relative position error of the golden trajectories per N:",
          {N: f"{v:.1e}" for N, v in sorted(worst.items())})


def test_rouse_kat_on_gpu(gpu):
    # tests/test_integrator_rouse.cpp:49-94
    p = H.Packed(C.phys_dict(spring=1, sqrtb=1000.0), 10)
    st = H.rng_of(5)
    R = C.R0_ROUSE
    for ans in (C.ROUSE_ANS1, C.ROUSE_ANS2):
        out, rngs, status, _, _ = H.gpu_batch(gpu, [(10, R, p.c, H.csim(C.sim_dict(0.0, 1.0, 0.1, 1e-6)), st)])
        st = rngs[0]
        assert np.max(np.abs(out[0] - ans)) < 1e-6
        R = out[0]


def test_singular_configuration_fails_cleanly(gpu, golden):
    p = H.Packed(C.phys_dict(spring=1, sqrtb=1000.0, hstar=C.SINGULAR_HSTAR, method=1), 21)
    ok = H.Packed(C.phys_dict(spring=1, sqrtb=1000.0, hstar=0.2, method=1), 21)
    sim = H.csim(C.sim_dict(0.0, 0.0, 0.01, 1e-6))
    out, rngs, status, cnts, _ = H.gpu_batch(gpu, [(21, C.R0_SINGULAR, p.c, sim, H.rng_of(6)),
                                                   (21, C.random_walk(21, 3), ok.c, sim, H.rng_of(7))])
    assert status[0] == A.STATUS_NOT_POSITIVE_DEFINITE and cnts[0].fail_index == 55
    assert (rngs[0].ix, rngs[0].iy, rngs[0].seed) == tuple(golden["singular_integrate/rng"])
    assert status[1] == 0 and np.all(np.isfinite(out[1]))   # the batch went on
    fp64_pivot = cnts[0].fail_value
    # the mixed form factors in FP64 (the same pivot value); its FP32-factor
    # variant fails the FP32 pivot (-1.18e-6), factors again in FP64 (column by
    # column, in the workspace) and reports that pivot: the same to rounding
    # of the unit-sized entries it cancels from
    for prec in (A.PRECISION_MIXED, A.PRECISION_MIXED_F32_FACTOR):
        out, rngs, status, cnts, _ = H.gpu_batch(gpu, [(21, C.R0_SINGULAR, p.c, sim, H.rng_of(6)),
                                                       (21, C.random_walk(21, 3), ok.c, sim, H.rng_of(7))],
                                                 precision=prec)
        assert status[0] == A.STATUS_NOT_POSITIVE_DEFINITE and cnts[0].fail_index == 55
        assert abs(cnts[0].fail_value - fp64_pivot) <= 1e-13, (prec, cnts[0].fail_value)
        assert status[1] == 0 and np.all(np.isfinite(out[1]))


def test_uneven_batch_equals_single_runs(gpu):
    """Chains of different N, springs, EV, flows in one call give the same
    result, bit for bit, as each alone (no cross-talk between workgroups)."""
    specs = []
    rs = np.random.default_rng(5)
    for k, N in enumerate([10, 33, 64, 65, 130, 12, 257, 50]):
        d = C.phys_dict(spring=[2, 6, 1, 3][k % 4], sqrtb=10.0, nat=2.0 if k % 4 == 1 else 0.0,
                        hstar=[0.3, 0.0, 0.25, 0.2][k % 4], ev=[3, 1, 0, 2][k % 4], zstar=0.5,
                        dstar=1.0, kappa=C.wi_kappa(float(rs.uniform(0.1, 10))), method=k % 2)
        specs.append((N, C.random_walk(N, 600 + k), H.Packed(d, N),
                      H.csim(C.sim_dict(0.0, 0.019, 1e-3, 1e-3)), H.rng_of(k)))
    items = [(N, R, p.c, s, st) for N, R, p, s, st in specs]
    together, rt, stt, _, _ = H.gpu_batch(gpu, items)
    for j, it in enumerate(items):
        alone, ra, sa, _, _ = H.gpu_batch(gpu, [it])
        assert np.array_equal(alone[0], together[j]) and sa[0] == stt[j]
        assert (ra[0].ix, ra[0].iy) == (rt[j].ix, rt[j].iy)


def test_uneven_workload_sample_matches_oracle(gpu, orc):
    """The uneven workload (N log-uniform in [10, 256], every block size from
    64 to 256 in one call, buckets launched side by side): 32 of its chains,
    10 steps, each within 1e-10 of the FP64 oracle with the same RNG state."""
    N, R, phys, dt = W.uneven_workload(n_traj=32)
    assert len(set((int(n) + 63) // 64 for n in N)) == 4
    sim = A.CSim(0.0, 9 * dt, dt, 1e-3, 1)
    items = [(int(N[i]), R[i], phys[i], sim, A.CRng(1, -1, i)) for i in range(len(N))]
    out, rngs, status, cnts, _ = H.gpu_batch(gpu, items)
    assert np.all(status == 0)
    for i in range(len(N)):
        st = A.CRng(1, -1, i)
        want, rc, cnto, _ = H.oracle_integrate(orc, int(N[i]), R[i], phys[i], sim, st)
        assert rc == 0 and cnts[i].steps == cnto.steps == 10
        assert rel(out[i], want) <= 1e-10, (i, int(N[i]), rel(out[i], want))
        assert (rngs[i].ix, rngs[i].iy, rngs[i].seed) == (st.ix, st.iy, st.seed)


@pytest.mark.parametrize("N", [2, 5, 16, 21, 22, 27, 32, 33])
def test_cholesky_factor_sizes(gpu, orc, N):
    """The Cholesky method across the sizes whose factor is held in LDS (3N up
    to 64 and up to 96: padding rows, whole padding blocks, exact fit) and just
    beyond (N=33, factor in the workspace): 30 steps within 1e-10 of the FP64
    oracle with the same decisions; the mixed form (FP64 factor) and its
    FP32-factor variant within 1e-5."""
    d = C.phys_dict(spring=2, sqrtb=np.sqrt(50.0), hstar=0.25, method=1, ev=1, zstar=1.0,
                    dstar=0.5, kappa=[0.5, 0, 0, 0, -0.5, 0, 0, 0, 0])
    p = H.Packed(d, N)
    R0 = C.random_walk(N, 300 + N, 0.95)
    sim = H.csim(C.sim_dict(0.0, 0.029, 1e-3, 1e-3))
    want, rc, cnto, lo = H.oracle_integrate(orc, N, R0, p.c, sim, H.rng_of(N), log_cap=32)
    assert rc == 0 and cnto.steps == 30
    out, rngs, status, cnts, logs = H.gpu_batch(gpu, [(N, R0, p.c, sim, H.rng_of(N))], log_stride=32)
    assert status[0] == 0
    assert rel(out[0], want) <= 1e-10, rel(out[0], want)
    for k in range(30):
        assert logs[k].corrector_iters == lo[k].corrector_iters, k
    for prec in (A.PRECISION_MIXED, A.PRECISION_MIXED_F32_FACTOR):
        out, _, status, cnts, logs = H.gpu_batch(gpu, [(N, R0, p.c, sim, H.rng_of(N))],
                                                 precision=prec, log_stride=32)
        assert status[0] == 0 and rel(out[0], want) < 1e-5, (prec, rel(out[0], want))


def _long_items(golden, names):
    cases = {c["name"]: c for c in C.long_trajectory_cases()}
    items, keep = [], []
    for nm in names:
        c = cases[nm]
        p = H.Packed(c["phys"], c["N"])
        keep.append(p)
        items.append((c["N"], golden[f"long/{nm}/R0"], p.c, H.csim(c["sim"]), H.rng_of(c["seed"])))
    return items, keep


LOG_KEYS = ("lanczos_cold", "lanczos_steps", "cheb_L", "cheb_attempts", "fallback",
            "corrector_iters")


@pytest.mark.parametrize("which", ["together", "n1000_split"])
def test_long_trajectories_100_steps(gpu, golden, which):
    """100 steps of the large configurations against the reference: two chains
    of N=256 (four-wave workgroups, uneven-workload physics at Wi 1 and 100)
    and two of N=1000 (sixteen waves, bending + Lennard-Jones). 'together': all
    four in one launch, every chain on one workgroup; 'n1000_split': the two
    N=1000 chains alone, each split over 8 workgroups. FP64 within 1e-10 with
    the reference's per-step decisions and RNG state; mixed within 1e-5."""
    names = [str(n) for n in golden["long_names"]]
    if which == "n1000_split":
        names = [n for n in names if n.startswith("c4_")]
    items, keep = _long_items(golden, names)
    out, rngs, status, cnts, logs = H.gpu_batch(gpu, items, log_stride=128)
    assert np.all(status == 0)
    worst = {}
    for j, nm in enumerate(names):
        want = golden[f"long/{nm}/R"]
        r = rel(out[j], want)
        worst[nm] = r
        assert r <= 1e-10, (nm, r)
        assert (rngs[j].ix, rngs[j].iy, rngs[j].seed) == tuple(golden[f"long/{nm}/rng"]), nm
        lg_want = golden[f"long/{nm}/log"]
        assert cnts[j].steps == len(lg_want) == 100
        for k in range(100):
            a = logs[j * 128 + k]
            assert tuple(getattr(a, f) for f in LOG_KEYS) == tuple(lg_want[k]), (nm, k)
            assert a.t == golden[f"long/{nm}/t"][k]
    mixed, _, status_m, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
    assert np.all(status_m == 0)
    for j, nm in enumerate(names):
        r = rel(mixed[j], golden[f"long/{nm}/R"])
        worst[nm + "/mixed"] = r
        assert r < 1e-5, (nm, r)
    print("relative position error after 100 steps:", {k: f"{v:.1e}" for k, v in worst.items()})


@pytest.mark.parametrize("n,N", [(3, 700), (40, 520)])
def test_split_chains(gpu, orc, n, N):
    """Batches whose chains all have N > 512 and leave compute units free run
    each chain on several workgroups (8 for 3 chains, 6 for 40 on 256 CUs),
    every D.x split by partners: 20 steps of the first chains within 1e-10 of
    the FP64 oracle with the same decisions, mixed within 1e-5, RNG states as
    the oracle's."""
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=1, zstar=0.3, dstar=0.5,
                    kappa=C.wi_kappa(0.05))
    p = H.Packed(d, N)
    sim = H.csim(C.sim_dict(0.0, 0.019, 1e-3, 1e-3))
    R0 = [W.self_avoiding_walk(N, np.random.default_rng(50 + i)) for i in range(n)]
    items = [(N, R0[i], p.c, sim, H.rng_of(i)) for i in range(n)]
    out, rngs, status, cnts, logs = H.gpu_batch(gpu, items, log_stride=24)
    assert np.all(status == 0)
    mixed, _, status_m, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
    assert np.all(status_m == 0)
    for i in range(2):
        st = H.rng_of(i)
        want, rc, cnto, lo = H.oracle_integrate(orc, N, R0[i], p.c, sim, st, log_cap=24)
        assert rc == 0 and cnts[i].steps == cnto.steps == 20
        assert rel(out[i], want) <= 1e-10, (i, rel(out[i], want))
        assert (rngs[i].ix, rngs[i].iy, rngs[i].seed) == (st.ix, st.iy, st.seed)
        for k in range(20):
            a, b = logs[i * 24 + k], lo[k]
            assert (a.lanczos_steps, a.cheb_L, a.cheb_attempts, a.corrector_iters) == \
                (b.lanczos_steps, b.cheb_L, b.cheb_attempts, b.corrector_iters), (i, k)
        assert rel(mixed[i], want) < 1e-5, (i, rel(mixed[i], want))


def test_wide_workgroups_spread_the_pair_walk(gpu, orc):
    """Chains on workgroups with more waves than 64-bead tiles (N=329: 6 tiles
    on 8 waves; N=700: 11 tiles on 16, not split because the batch holds
    shorter chains too) spread the mixed form's tile walks over all waves, every
    row through the fixed-point sums: 20 steps within 1e-5 of the FP64 oracle
    (FP64 form within 1e-10 with the oracle's decisions), and a second run
    equal to the first bit for bit (the integer sums do not depend on the order
    in which the waves add)."""
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=1, zstar=0.3, dstar=0.5,
                    kappa=C.wi_kappa(0.05))
    sim = H.csim(C.sim_dict(0.0, 0.019, 1e-3, 1e-3))
    items, want = [], []
    for i, N in enumerate((329, 700, 329, 700)):
        p = H.Packed(d, N)
        R0 = W.self_avoiding_walk(N, np.random.default_rng(90 + i))
        items.append((N, R0, p.c, sim, H.rng_of(i)))
        if i < 2:
            st = H.rng_of(i)
            w, rc, cnto, lo = H.oracle_integrate(orc, N, R0, p.c, sim, st)
            assert rc == 0
            want.append(w)
    out, _, status, _, _ = H.gpu_batch(gpu, items)
    assert np.all(status == 0)
    m1, _, st1, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
    m2, _, st2, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
    assert np.all(st1 == 0) and np.all(st2 == 0)
    for i in range(2):
        assert rel(out[i], want[i]) <= 1e-10, (i, rel(out[i], want[i]))
        assert rel(m1[i], want[i]) < 1e-5, (i, rel(m1[i], want[i]))
    for a, b in zip(m1, m2):
        assert np.array_equal(a, b)


def test_resident_split_chains_match(gpu):
    """A resident batch of 3 chains of N=600 (split over 8 workgroups each)
    over two segments equals bdsim_integrate_batch on the same chains (split
    the same way) bit for bit."""
    n, N = 3, 600
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=2, zstar=0.5, dstar=1.0, max_cutoff=2.5)
    p = H.Packed(d, N)
    R0 = np.ascontiguousarray(np.array([W.self_avoiding_walk(N, np.random.default_rng(70 + i))
                                        for i in range(n)]))
    rng = (A.CRng * n)(*[H.rng_of(i) for i in range(n)])
    b = ctypes.c_void_p()
    A.check(gpu, gpu.bdsim_batch_create(byref(b), n, N, A.dptr(R0), byref(p.c), None, rng,
                                        A.PRECISION_MIXED))
    sim = A.CSim(0.0, 0.002, 1e-3, 1e-3, 1)
    try:
        A.check(gpu, gpu.bdsim_batch_run(b, byref(sim)))
        A.check(gpu, gpu.bdsim_batch_run(b, byref(sim)))
        Rb = np.empty_like(R0)
        st = (A.CRng * n)()
        status = np.zeros(n, dtype=np.int32)
        A.check(gpu, gpu.bdsim_batch_fetch(b, A.dptr(Rb), st, A.iptr(status), None))
    finally:
        gpu.bdsim_batch_destroy(b)
    assert np.all(status == 0)
    items = [(N, R0[i], p.c, sim, H.rng_of(i)) for i in range(n)]
    for _ in range(2):
        out, rngs, stt, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
        assert np.all(stt == 0)
        items = [(N, out[i], p.c, sim, rngs[i]) for i in range(n)]
    for i in range(n):
        assert np.array_equal(out[i], Rb[i]), i
        assert (rngs[i].ix, rngs[i].iy, rngs[i].seed) == (st[i].ix, st[i].iy, st[i].seed)


def _batch_segment(gpu, R0, p, prec, sim, cuts):
    """One segment on a resident batch: whole (cuts None) or cut into runs of
    the given step counts (the first starts it, the others resume)."""
    n, N = R0.shape[0], R0.shape[1]
    rng = (A.CRng * n)(*[H.rng_of(10 + i) for i in range(n)])
    b = ctypes.c_void_p()
    A.check(gpu, gpu.bdsim_batch_create(byref(b), n, N, A.dptr(R0), byref(p.c), None, rng, prec))
    try:
        if cuts is None:
            A.check(gpu, gpu.bdsim_batch_run(b, byref(sim)))
        else:
            for k, m in enumerate(cuts):
                A.check(gpu, gpu.bdsim_batch_run_steps(b, byref(sim), m, 1 if k else 0))
        R = np.empty_like(R0)
        st = (A.CRng * n)()
        status = np.zeros(n, dtype=np.int32)
        cnt = (A.CCounters * n)()
        A.check(gpu, gpu.bdsim_batch_fetch(b, A.dptr(R), st, A.iptr(status), cnt))
    finally:
        gpu.bdsim_batch_destroy(b)
    return R, [(s.ix, s.iy, s.seed) for s in st], status, [bytes(c) for c in cnt]


@pytest.mark.parametrize("kind", ["cheb64", "cheb130", "cholesky30", "split600"])
def test_resident_segment_cut_into_runs(gpu, kind):
    """bdsim_batch_run_steps: a 30-step segment cut into runs of 7, 13, 9 and
    5 steps (the last finds 1 step left) equals one bdsim_batch_run over it
    bit for bit: positions, generator states, status and counters, both forms
    (the resumed runs start from the stored time and warm Lanczos seed; split600:
    3 chains of N=600 each split over several workgroups, every part resuming
    from the same stored state)."""
    N = {"cheb64": 50, "cheb130": 130, "cholesky30": 30, "split600": 600}[kind]
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=3 if N == 50 else 1, zstar=0.5, dstar=1.0,
                    method=1 if kind == "cholesky30" else 0, kappa=C.wi_kappa(3.0))
    p = H.Packed(d, N)
    n = 3 if kind == "split600" else 5
    R0 = np.ascontiguousarray(np.array([C.random_walk(N, 900 + i, 0.9) for i in range(n)]))
    sim = A.CSim(0.0, 0.029, 1e-3, 1e-3, 1)
    precs = [A.PRECISION_FP64, A.PRECISION_MIXED]
    if kind == "cholesky30":
        precs.append(A.PRECISION_MIXED_F32_FACTOR)
    for prec in precs:
        whole = _batch_segment(gpu, R0, p, prec, sim, None)
        cut = _batch_segment(gpu, R0, p, prec, sim, [7, 13, 9, 5])
        assert np.all(whole[2] == 0)
        assert np.array_equal(whole[0], cut[0]), prec
        assert whole[1] == cut[1] and np.array_equal(whole[2], cut[2]) and whole[3] == cut[3], prec


def test_resident_resume_needs_its_segment(gpu):
    """Resuming without a started segment, or with another clock, is refused."""
    N, n = 20, 2
    p = H.Packed(C.phys_dict(spring=2, sqrtb=10.0, hstar=0.2), N)
    R0 = np.ascontiguousarray(np.array([C.random_walk(N, 5 + i) for i in range(n)]))
    rng = (A.CRng * n)(*[H.rng_of(i) for i in range(n)])
    b = ctypes.c_void_p()
    A.check(gpu, gpu.bdsim_batch_create(byref(b), n, N, A.dptr(R0), byref(p.c), None, rng,
                                        A.PRECISION_FP64))
    try:
        sim = A.CSim(0.0, 0.009, 1e-3, 1e-3, 1)
        assert gpu.bdsim_batch_run_steps(b, byref(sim), 3, 1) == -1
        assert gpu.bdsim_batch_run_steps(b, byref(sim), 0, 0) == -1
        A.check(gpu, gpu.bdsim_batch_run_steps(b, byref(sim), 3, 0))
        other = A.CSim(0.0, 0.019, 1e-3, 1e-3, 1)
        assert gpu.bdsim_batch_run_steps(b, byref(other), 3, 1) == -1
        A.check(gpu, gpu.bdsim_batch_run_steps(b, byref(sim), 3, 1))
        A.check(gpu, gpu.bdsim_batch_run(b, byref(sim)))          # a whole segment closes it
        assert gpu.bdsim_batch_run_steps(b, byref(sim), 3, 1) == -1
        A.check(gpu, gpu.bdsim_batch_sync(b))
    finally:
        gpu.bdsim_batch_destroy(b)


def test_small_launch_with_lds_workspace(gpu, orc):
    """A launch of few short chains keeps the hot workspace (state slots,
    Chebyshev coefficients, Lanczos basis) in LDS: the 16 chains of N=10 of the
    CPU-sized case, 100 steps, FP64 within 1e-10 of the oracle with its
    per-step decisions and generator states, mixed within 1e-5."""
    from bdsim_amd.workloads import c1_workload
    R0, phys, _, _, _ = c1_workload()
    n, N = R0.shape[0], R0.shape[1]
    pk = A.PackedPhys(phys, N)
    sim = A.CSim(0.0, 0.99, 0.01, 1e-3, 1)
    items = [(N, R0[i], pk.c, sim, H.rng_of(i)) for i in range(n)]
    out, rngs, status, cnts, logs = H.gpu_batch(gpu, items, log_stride=100)
    mixed, _, status_m, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
    assert np.all(status == 0) and np.all(status_m == 0)
    for i in range(n):
        st = H.rng_of(i)
        want, rc, cnto, lo = H.oracle_integrate(orc, N, R0[i], pk.c, sim, st, log_cap=100)
        assert rc == 0 and cnts[i].steps == cnto.steps == 100
        assert rel(out[i], want) <= 1e-10, (i, rel(out[i], want))
        assert (rngs[i].ix, rngs[i].iy, rngs[i].seed) == (st.ix, st.iy, st.seed)
        for k in range(100):
            a, b = logs[i * 100 + k], lo[k]
            assert (a.lanczos_cold, a.lanczos_steps, a.cheb_L, a.cheb_attempts, a.corrector_iters) == \
                (b.lanczos_cold, b.lanczos_steps, b.cheb_L, b.cheb_attempts, b.corrector_iters), (i, k)
        assert rel(mixed[i], want) < 1e-5, (i, rel(mixed[i], want))


def test_lds_workspace_at_its_largest(gpu, orc):
    """The largest chains whose hot workspace still goes to LDS (N=54: 64 KB
    of LDS per workgroup with the static part), three of them, headline
    physics, 30 steps: FP64 within 1e-10 of the oracle, mixed within 1e-5."""
    N, n = 54, 3
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3, ev=3, dstar=1.0, kappa=C.wi_kappa(10.0))
    p = H.Packed(d, N)
    sim = H.csim(C.sim_dict(0.0, 0.029, 1e-3, 1e-3))
    R0 = [C.random_walk(N, 60 + i) for i in range(n)]
    items = [(N, R0[i], p.c, sim, H.rng_of(i)) for i in range(n)]
    out, rngs, status, cnts, _ = H.gpu_batch(gpu, items)
    mixed, _, status_m, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
    assert np.all(status == 0) and np.all(status_m == 0)
    for i in range(n):
        st = H.rng_of(i)
        want, rc, cnto, _ = H.oracle_integrate(orc, N, R0[i], p.c, sim, st)
        assert rc == 0 and cnts[i].steps == cnto.steps == 30
        assert rel(out[i], want) <= 1e-10, (i, rel(out[i], want))
        assert (rngs[i].ix, rngs[i].iy, rngs[i].seed) == (st.ix, st.iy, st.seed)
        assert rel(mixed[i], want) < 1e-5, (i, rel(mixed[i], want))


def test_mixed_precision_within_1e5(gpu, orc):
    """Headline physics (N=50, FENE b=100, h*=0.3 Chebyshev, SDK, shear):
    the mixed form stays within 1e-5 relative of the FP64 oracle over 100 steps."""
    items, ref = [], []
    for i, wi in enumerate([0.1, 1.0, 10.0, 100.0]):
        d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3, ev=3, dstar=1.0, kappa=C.wi_kappa(wi))
        p = H.Packed(d, 50)
        R0 = C.random_walk(50, 1000 + i)
        sim = H.csim(C.sim_dict(0.0, 0.099, 1e-3, 1e-3))
        items.append((50, R0, p.c, sim, H.rng_of(i)))
        ref.append(H.oracle_integrate(orc, 50, R0, p.c, sim, H.rng_of(i))[0])
        items[-1] = items[-1] + (p,)
    out, _, status, _, _ = H.gpu_batch(gpu, [it[:5] for it in items], precision=A.PRECISION_MIXED)
    assert np.all(status == 0)
    for o, r in zip(out, ref):
        assert rel(o, r) < 1e-5, rel(o, r)


def test_mixed_form_every_golden_case(gpu, golden):
    """Every golden trajectory in the mixed form (FP32 pair terms for D x and
    excluded volume, FP64 elsewhere), one batched launch per segment, against
    the reference's FP64 positions. Over these 10-100 step segments the mixed
    form stays within 1e-5 relative (the largest error, N=50 SDK at Wi=10, is
    about 9e-7)."""
    cases = [c for c, _ in _golden_items(golden)]
    keep = [H.Packed(c["phys"], c["N"]) for c in cases]
    Rs = [golden[f"traj/{c['name']}/R0"] for c in cases]
    sts = [H.rng_of(c["seed"]) for c in cases]
    worst = {}
    for seg in range(max(c["segments"] for c in cases)):
        idx = [i for i, c in enumerate(cases) if seg < c["segments"]]
        items = [(cases[i]["N"], Rs[i], keep[i].c, H.csim(cases[i]["sim"]), sts[i]) for i in idx]
        out, rngs, status, _, _ = H.gpu_batch(gpu, items, precision=A.PRECISION_MIXED)
        for j, i in enumerate(idx):
            name = cases[i]["name"]
            assert status[j] == golden[f"traj/{name}/rc"][seg], name
            sts[i] = rngs[j]
            if status[j] != 0:
                continue
            key = (cases[i]["N"], name)
            worst[key] = max(worst.get(key, 0.0), rel(out[j], golden[f"traj/{name}/R"][seg]))
            Rs[i] = out[j]
    for c, st in zip(cases, sts):
        assert (st.ix, st.iy, st.seed) == tuple(golden[f"traj/{c['name']}/rng"]), c["name"]
    per_n = {}
    for (N, _), v in worst.items():
        per_n[N] = max(per_n.get(N, 0.0), v)
    print("mixed-form relative error, worst per N:", {N: f"{v:.1e}" for N, v in sorted(per_n.items())})
    print("mixed-form relative error per case:",
          {k[1]: f"{v:.1e}" for k, v in sorted(worst.items(), key=lambda kv: -kv[1])})
    for k, v in worst.items():
        assert v < 1e-5, (k, v)


def _helix(N, seed):
    """A tight helix (radius 0.6, 100 degrees and 0.25 rise per bead): bonds
    near 0.95, beads 3 and 4 apart at about 1.0, 2 apart at 1.3."""
    rs = np.random.default_rng(seed)
    k = np.arange(N)
    turn = np.deg2rad(100.0)
    R = np.stack([0.6 * np.cos(turn * k), 0.6 * np.sin(turn * k), 0.25 * k], axis=1)
    return np.ascontiguousarray(R + 0.02 * rs.normal(size=(N, 3)))


def _phi_table(N, seed):
    rs = np.random.default_rng(seed)
    a = rs.uniform(0.0, 1.0, size=(N, N))
    return (a + a.T) / 2


EV_VARIANTS = {
    "gauss": dict(ev=1, zstar=1.0, dstar=0.5),
    "lj_cd3": dict(ev=2, zstar=0.3, dstar=0.8, cd=3),
    "lj_soft_core_0": dict(ev=2, zstar=0.1, dstar=0.7, min_cutoff=0.0),
    "sdk_soft_core_0": dict(ev=3, dstar=0.95, min_cutoff=0.0),
    "sdk_phi_cd2": dict(ev=3, dstar=0.75, cd=2, phi="table"),
    "sdk_phi_equil": dict(ev=3, dstar=0.9, equil=1, phi="table"),
}


@pytest.mark.parametrize("name", sorted(EV_VARIANTS))
def test_ev_variants_both_forms(gpu, orc, name):
    """Excluded-volume variants beyond the golden set (zero soft-core radius,
    contour exclusion with an attraction table): the FP64 force matches the
    oracle to 1e-12; 100 HI steps match it to 1e-10 (FP64) and 1e-5 (mixed)."""
    N = 24
    kw = dict(EV_VARIANTS[name])
    if kw.get("phi") == "table":
        kw["phi"] = _phi_table(N, 5)
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3, kappa=C.wi_kappa(1.0), **kw)
    p = H.Packed(d, N)
    R0 = _helix(N, 9)
    F = np.empty_like(R0)
    A.check(gpu, gpu.bdsim_ev_force_batch(1, N, A.dptr(R0), byref(p.c), A.dptr(F)))
    Fo = np.empty_like(R0)
    orc.bo_ev_force(N, A.dptr(R0), byref(p.c), A.dptr(Fo))
    assert np.abs(Fo).max() > 0.0
    assert rel(F, Fo) < 1e-12
    sim = H.csim(C.sim_dict(0.0, 0.099, 1e-3, 1e-6))
    want, rc, _, _ = H.oracle_integrate(orc, N, R0, p.c, sim, H.rng_of(3))
    assert rc == 0
    for prec, tol in ((A.PRECISION_FP64, 1e-10), (A.PRECISION_MIXED, 1e-5)):
        out, _, status, _, _ = H.gpu_batch(gpu, [(N, R0, p.c, sim, H.rng_of(3))], precision=prec)
        assert status[0] == 0
        assert rel(out[0], want) < tol, (prec, rel(out[0], want))


def test_device_resident_batch_matches(gpu):
    """bdsim_batch_* (positions and RNG resident in HBM) reproduces
    bdsim_integrate_batch over two consecutive segments."""
    n, N = 6, 20
    d = C.phys_dict(spring=2, sqrtb=10.0, hstar=0.3, ev=3, dstar=1.0)
    p = H.Packed(d, N)
    R0 = np.ascontiguousarray(np.array([C.random_walk(N, 700 + i) for i in range(n)]))
    kap = np.ascontiguousarray(np.array([C.wi_kappa(0.5 * i) for i in range(n)]))
    rng = (A.CRng * n)(*[H.rng_of(i) for i in range(n)])
    b = ctypes.c_void_p()
    A.check(gpu, gpu.bdsim_batch_create(byref(b), n, N, A.dptr(R0), byref(p.c), A.dptr(kap), rng,
                                        A.PRECISION_FP64))
    sim = A.CSim(0.0, 0.009, 1e-3, 1e-3, 1)
    try:
        A.check(gpu, gpu.bdsim_batch_run(b, byref(sim)))
        A.check(gpu, gpu.bdsim_batch_run(b, byref(sim)))
        Rb = np.empty_like(R0)
        st = (A.CRng * n)()
        status = np.zeros(n, dtype=np.int32)
        A.check(gpu, gpu.bdsim_batch_fetch(b, A.dptr(Rb), st, A.iptr(status), None))
    finally:
        gpu.bdsim_batch_destroy(b)
    for i in range(n):
        di = dict(d, kappa=C.wi_kappa(0.5 * i))
        pi = H.Packed(di, N)
        s = H.rng_of(i)
        R = R0[i]
        for _ in range(2):
            out, rr, stt, _, _ = H.gpu_batch(gpu, [(N, R, pi.c, sim, s)])
            R, s = out[0], rr[0]
        assert np.array_equal(R, Rb[i]) and (s.ix, s.iy, s.seed) == (st[i].ix, st[i].iy, st[i].seed)


def test_core_api_integrate(gpu, orc):
    """The `_bdsim`-compatible Python surface drives the GPU and matches the oracle."""
    import pickle
    import bdsim_amd as bd
    phys = bd.PhysParams()
    phys.spring.type = bd.Spring.FENEFraenkel
    phys.spring.sqrtb = 2.0
    phys.spring.natural_length = 1.0
    phys.hstar = 0.3
    phys.flow = bd.Flow.constant(np.array([[0, 1.0, 0], [0, 0, 0], [0, 0, 0]]))
    sim = bd.SimParams()
    sim.time_start, sim.time_end, sim.dt = 0.0, 0.99, 0.01
    rng = bd.Rng(3)
    R0 = C.ff_chain(10, 103)
    phys2 = pickle.loads(pickle.dumps(phys))
    R = bd.integrate(R0, phys2, sim, rng)
    assert phys2.number_of_beads == 10   # the caller's object is untouched
    d = C.phys_dict(spring=6, sqrtb=2.0, nat=1.0, hstar=0.3, kappa=C.SHEAR)
    p = H.Packed(d, 10)
    st = H.rng_of(3)
    Ro, _, _, _ = H.oracle_integrate(orc, 10, R0, p.c, H.csim(C.sim_dict(0.0, 0.99, 0.01, 1e-3)), st)
    assert rel(R, Ro) < 1e-10
    assert rng.__getstate__() == (st.ix, st.iy, st.seed)
    with pytest.raises(RuntimeError, match="pivot 55"):
        ph = bd.PhysParams()
        ph.spring.sqrtb = 1000.0
        ph.hstar = C.SINGULAR_HSTAR
        ph.hi_method = bd.DelSMethod.Cholesky
        s1 = bd.SimParams()
        s1.time_end = 0.0
        bd.integrate(C.R0_SINGULAR, ph, s1, bd.Rng(6))
    Ft = bd.total_force(C.R0_FULL, phys)
    Fo = np.empty_like(C.R0_FULL)
    orc.bo_total_force(15, A.dptr(np.ascontiguousarray(C.R0_FULL)), byref(H.Packed(
        dict(d, phi=None), 15).c), A.dptr(Fo))
    assert rel(Ft, Fo) < 1e-12
