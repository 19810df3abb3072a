"""Write tests/golden/golden.npz from the reference bdsim core.

Runs only in a container where the reference tree is present: `make -C oracle
ref` builds oracle/_ref/libbdref.so from the reference sources, and this script
drives it through oracle/ref_shim.cpp. The fixture holds inputs and reference
outputs only (arrays, no code). Cases that would abort the reference process
(rtsafe) are screened with the oracle first and skipped.

    python tests/golden/make_golden.py
"""
import ctypes
import json
import os
import sys
from ctypes import byref, c_double, c_int, c_long

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from bdsim_amd import _abi as A  # noqa: E402
import helpers as H  # noqa: E402
from golden import cases as C  # noqa: E402

RNG_SEEDS = [0, 1, 5, 7, 10, 11, 4095, 8191, -3, 123456789]
RNG_DRAWS = 2000
RNG_LONG = 100000


def main():
    ref = H.ref()
    orc = H.oracle()
    if ref is None:
        raise SystemExit("oracle/_ref/libbdref.so not built (make -C oracle ref)")
    out = {}

    # ---- RNG streams (rng.cpp:41-70) ----
    draws = np.empty((len(RNG_SEEDS), RNG_DRAWS), dtype=np.float32)
    states = np.empty((len(RNG_SEEDS), 3), dtype=np.int32)
    for k, s in enumerate(RNG_SEEDS):
        st = H.rng_of(s)
        x = np.empty(RNG_DRAWS)
        ref.ref_ran1(byref(st), RNG_DRAWS, A.dptr(x))
        draws[k] = x.astype(np.float32)
        assert np.array_equal(draws[k].astype(np.float64), x)   # values are floats
        y = np.empty(RNG_LONG - RNG_DRAWS)
        ref.ref_ran1(byref(st), y.size, A.dptr(y))
        states[k] = (st.ix, st.iy, st.seed)
    out["rng_seeds"] = np.array(RNG_SEEDS, dtype=np.int32)
    out["rng_draws"] = draws
    out["rng_state_after_long"] = states

    # ---- forces (spring.cpp, excluded_volume.cpp, bending.cpp) ----
    force_cases = []
    Rs = {"full": C.R0_FULL, "walk10": C.random_walk(10, 7), "walk40": C.random_walk(40, 8, 0.7)}
    for rname, R in Rs.items():
        N = R.shape[0]
        for law, sqrtb, nat in [(1, 10.0, 0.0), (2, 10.0, 0.0), (3, 10.0, 0.0), (4, 10.0, 0.0),
                                (5, 3.0, 1.0), (6, 3.0, 1.5), (7, 10.0, 5.0)]:
            force_cases.append((f"spring{law}_{rname}", R, C.phys_dict(spring=law, sqrtb=sqrtb, nat=nat), "spring"))
        for ev, z, d in [(1, 1.0, 0.5), (2, 1.0, 1.0), (3, 0.0, 1.0)]:
            for eq in (0, 1):
                phi = None
                if ev == 3:
                    rs = np.random.default_rng(N)
                    phi = rs.uniform(0.0, 0.8, size=(N, N))
                force_cases.append((f"ev{ev}_eq{eq}_{rname}", R,
                                    C.phys_dict(ev=ev, zstar=z, dstar=d, equil=eq, phi=phi,
                                                max_cutoff=2.5 if ev == 2 else 1.5), "ev"))
        force_cases.append((f"ev2_cd3_{rname}", R, C.phys_dict(ev=2, zstar=0.5, dstar=1.0, cd=3), "ev"))
        force_cases.append((f"bend_{rname}", R, C.phys_dict(bend=1, C=2.5), "bend"))
        force_cases.append((f"total_{rname}", R,
                            C.phys_dict(spring=6, sqrtb=3.0, nat=1.5, bend=1, C=1.5, ev=2, zstar=1.0,
                                        dstar=1.0), "total"))
    fn = {"spring": ref.ref_spring_force, "ev": ref.ref_ev_force,
          "bend": ref.ref_bending_force, "total": ref.ref_total_force}
    for k, (name, R, d, kind) in enumerate(force_cases):
        N = R.shape[0]
        p = H.Packed(d, N)
        F = np.empty_like(R)
        fn[kind](N, A.dptr(np.ascontiguousarray(R)), byref(p.c), A.dptr(F))
        out[f"force/{name}/R"] = R
        out[f"force/{name}/F"] = F
        out[f"force/{name}/kind"] = np.array(kind)
        dd = dict(d)
        if dd["phi"] is not None:
            out[f"force/{name}/phi"] = np.asarray(dd["phi"])
            dd["phi"] = None
        out[f"force/{name}/phys"] = np.array(json.dumps(dd))

    # ---- external forces and kappa(t) tables ----
    ext_d = C.phys_dict(external=((0, True, (0.0, 0.0, -3.0), (), ()),
                                  (-1, False, (0, 0, 0), (0.0, 0.5, 1.0),
                                   ((0.0, 0.0, 0.0), (0.0, 0.0, 6.0), (1.0, 2.0, 3.0))),
                                  (3, True, (1.0, -2.0, 0.5), (), ()),
                                  (3, True, (0.5, 0.5, 0.5), (), ()),
                                  (17, True, (9.0, 9.0, 9.0), (), ())),
                        flow_times=[0.0, 0.3, 0.8],
                        flow_kappas=[[0.0] * 9, C.wi_kappa(1.0), [2.0, 0, 0, 0, -1.0, 0, 0, 0, -1.0]])
    ts = np.array([-1.0, 0.0, 0.1, 0.25, 0.3, 0.5, 0.55, 0.8, 1.0, 3.0])
    p = H.Packed(ext_d, 6)
    ext_F = np.empty((ts.size, 6, 3))
    kap = np.empty((ts.size, 9))
    for i, t in enumerate(ts):
        F = np.empty((6, 3))
        ref.ref_external_force(byref(p.c), 6, float(t), A.dptr(F))
        ext_F[i] = F
        K = np.empty(9)
        ref.ref_kappa(byref(p.c), float(t), A.dptr(K))
        kap[i] = K
    out["external/t"] = ts
    out["external/F"] = ext_F
    out["external/kappa"] = kap

    # ---- RPY tensor, D.x, Cholesky, Chebyshev, spectral bounds ----
    for name, R, hstar in [("walk10", C.random_walk(10, 11), 0.3), ("full", C.R0_FULL, 0.25),
                           ("walk50", C.random_walk(50, 12), 0.3), ("close", C.random_walk(12, 13, 0.3), 0.3)]:
        N = R.shape[0]
        n = 3 * N
        D = np.empty((n, n))
        ref.ref_rpy_matrix(N, A.dptr(np.ascontiguousarray(R)), hstar, A.dptr(D))
        x = np.random.default_rng(N).normal(size=(N, 3))
        L = np.empty((n, n))
        rc = ref.ref_cholesky(n, A.dptr(D), A.dptr(L))
        assert rc == 0
        cheb = np.empty((N, 3))
        nt = c_int()
        fd = c_double()
        ref.ref_chebyshev_sqrt_times(N, A.dptr(D), A.dptr(x), 1.0, 0.0025, A.dptr(cheb), byref(nt), byref(fd))
        out[f"hi/{name}/R"] = R
        out[f"hi/{name}/hstar"] = np.array(hstar)
        if N <= 12:
            out[f"hi/{name}/D"] = D
            out[f"hi/{name}/L"] = L
        out[f"hi/{name}/x"] = x
        out[f"hi/{name}/Dx"] = D @ x.ravel()
        out[f"hi/{name}/Lx"] = np.tril(L) @ x.ravel()
        out[f"hi/{name}/D_trace_fro"] = np.array([np.trace(D), np.linalg.norm(D)])
        out[f"hi/{name}/cheb"] = cheb
        out[f"hi/{name}/cheb_nterms"] = np.array(nt.value)
        out[f"hi/{name}/cheb_fd"] = np.array(fd.value)
    # warm-started bounds along a sequence of slightly moving configurations
    R = C.random_walk(24, 14)
    Ds, Rseq = [], []
    rs = np.random.default_rng(15)
    for k in range(12):
        Rk = R + 0.01 * k * rs.normal(size=R.shape)
        D = np.empty((72, 72))
        ref.ref_rpy_matrix(24, A.dptr(np.ascontiguousarray(Rk)), 0.15, A.dptr(D))
        Ds.append(D)
        Rseq.append(Rk)
    Ds = np.ascontiguousarray(np.array(Ds))
    bounds = np.empty(24)
    cold, warm = c_long(), c_long()
    ref.ref_spectral_bounds_seq(72, 12, A.dptr(Ds), A.dptr(bounds), byref(cold), byref(warm))
    out["bounds/R"] = np.array(Rseq)
    out["bounds/hstar"] = np.array(0.15)
    out["bounds/lminmax"] = bounds.reshape(12, 2)
    out["bounds/cold_warm"] = np.array([cold.value, warm.value])
    # the degenerate configuration must fail the Cholesky at pivot 55
    D = np.empty((63, 63))
    ref.ref_rpy_matrix(21, A.dptr(np.ascontiguousarray(C.R0_SINGULAR)), C.SINGULAR_HSTAR, A.dptr(D))
    L = np.empty((63, 63))
    rc = ref.ref_cholesky(63, A.dptr(D), A.dptr(L))
    msg = ref.ref_last_error().decode()
    out["singular/rc"] = np.array(rc)
    out["singular/message"] = np.array(msg)

    # ---- implicit spring solve r(gamma) per law (spring.cpp:123-172) ----
    gam = np.concatenate([np.linspace(0.0, 3.0, 61), [5.0, 10.0, 1e2, 1e3, 1e4, 1e6, 1e10, 1e20]])
    for law, natscl in [(1, 0.0), (2, 0.0), (3, 0.0), (4, 0.0), (5, 0.3), (6, 0.0), (6, 0.5), (7, 0.0), (7, 0.8)]:
        for dtby4 in (0.00025, 0.0025, 0.05):
            rr = np.empty(gam.size)
            ff = np.empty(gam.size)
            ok = np.ones(gam.size, dtype=bool)
            for i, g in enumerate(gam):
                a, b = c_double(), c_double()
                st = orc.bo_solve_implicit_r(law, dtby4, float(g), natscl, byref(a), byref(b))
                if st != 0 or not np.isfinite(a.value):
                    ok[i] = False
                    rr[i] = ff[i] = np.nan
                    continue
                ref.ref_solve_implicit_r(law, dtby4, float(g), natscl, byref(a), byref(b))
                rr[i], ff[i] = a.value, b.value
            key = f"spring_solve/{law}_{natscl:g}_{dtby4:g}"
            out[key + "/gamma"] = gam
            out[key + "/r"] = rr
            out[key + "/ff"] = ff
            out[key + "/ok"] = ok

    # ---- trajectories (integrator.cpp:61-177) ----
    names = []
    for case in C.trajectory_cases():
        N = case["N"]
        R0 = C.make_R0(case["R0"])
        p = H.Packed(case["phys"], N)
        # screen with the oracle: the reference aborts the process on rtsafe failures
        outs_o, _ = H.run_segments(lambda *a: H.oracle_integrate(orc, *a), N, R0, p.c,
                                   case["sim"], case["seed"], case["segments"])
        if any(o[1] not in (0, 1) for o in outs_o):
            print("skipping (would abort the reference):", case["name"])
            continue
        outs, st = H.run_segments(lambda *a: H.ref_integrate(ref, *a), N, R0, p.c,
                                  case["sim"], case["seed"], case["segments"])
        key = "traj/" + case["name"]
        out[key + "/R0"] = R0
        out[key + "/R"] = np.array([o[0] for o in outs])
        out[key + "/rc"] = np.array([o[1] for o in outs])
        out[key + "/rng"] = np.array([st.ix, st.iy, st.seed], dtype=np.int32)
        names.append(case["name"])
    out["traj_names"] = np.array(names)

    # 100-step trajectories of the large configurations; the oracle, bit for
    # bit the reference here, supplies the per-step decisions
    long_names = []
    for case in C.long_trajectory_cases():
        N = case["N"]
        R0 = C.make_R0(case["R0"])
        p = H.Packed(case["phys"], N)
        st = H.rng_of(case["seed"])
        R, rc = H.ref_integrate(ref, N, R0, p.c, H.csim(case["sim"]), st)
        so = H.rng_of(case["seed"])
        Ro, rco, cnt, lg = H.oracle_integrate(orc, N, R0, p.c, H.csim(case["sim"]), so, log_cap=200)
        assert rc == rco == 0 and np.array_equal(R, Ro), case["name"]
        assert (st.ix, st.iy, st.seed) == (so.ix, so.iy, so.seed)
        key = "long/" + case["name"]
        out[key + "/R0"] = R0
        out[key + "/R"] = R
        out[key + "/rng"] = np.array([st.ix, st.iy, st.seed], dtype=np.int32)
        out[key + "/log"] = np.array([[l.lanczos_cold, l.lanczos_steps, l.cheb_L, l.cheb_attempts,
                                       l.fallback, l.corrector_iters] for l in lg], dtype=np.int32)
        out[key + "/t"] = np.array([l.t for l in lg])
        long_names.append(case["name"])
        print("long case", case["name"], "steps", cnt.steps, flush=True)
    out["long_names"] = np.array(long_names)

    # the degenerate configuration under Chebyshev with an FD tolerance no
    # series meets: the fallback factor throws (linalg.cpp:16-24)
    fc = {c["name"]: c for c in C.failure_cases()}["singular_cheb_fallback_not_pd"]
    p = H.Packed(fc["phys"], 21)
    st = H.rng_of(fc["seed"])
    R, rc = H.ref_integrate(ref, 21, C.R0_SINGULAR, p.c, H.csim(fc["sim"]), st)
    out["singular_cheb/rc"] = np.array(rc)
    out["singular_cheb/message"] = np.array(ref.ref_last_error().decode())
    out["singular_cheb/rng"] = np.array([st.ix, st.iy, st.seed], dtype=np.int32)

    # the singular configuration through integrate (Cholesky, one step)
    sd = C.phys_dict(spring=1, sqrtb=1000.0, hstar=C.SINGULAR_HSTAR, method=1)
    p = H.Packed(sd, 21)
    st = H.rng_of(6)
    R, rc = H.ref_integrate(ref, 21, C.R0_SINGULAR, p.c, H.csim(C.sim_dict(0.0, 0.0, 0.01, 1e-6)), st)
    out["singular_integrate/rc"] = np.array(rc)
    out["singular_integrate/message"] = np.array(ref.ref_last_error().decode())
    out["singular_integrate/rng"] = np.array([st.ix, st.iy, st.seed], dtype=np.int32)

    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes;", len(names), "trajectory cases")


if __name__ == "__main__":
    main()
