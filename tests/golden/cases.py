"""Shared definitions of the golden test cases (inputs only).

Used by make_golden.py (which runs the reference bdsim core, oracle/_ref, to
write the expected outputs) and by the tests (which feed the same inputs to the
oracle and to the GPU library). Literal configurations are the reference's own
test fixtures: R0_FULL / R0_SINGULAR from validation/compare_fortran_oracle.py
:53-99, the Rouse chain and answers of tests/test_integrator_rouse.cpp:49-94.
"""
import numpy as np

R0_ROUSE = np.array([
    [0.0, 0.0, 0.0],
    [-6.586250662803650e-02, -4.911733418703079e-02, 2.791636288166050e-01],
    [-6.753685772418980e-01, 3.519757017493250e-01, -6.340299546718600e-01],
    [-2.844688922166820e00, -2.239686943590640e00, -7.809645235538480e-01],
    [-1.818384915590290e00, -2.201230965554710e00, 9.345296323299410e-01],
    [-3.210149317979810e00, -9.372534379363060e-01, -4.681020081043240e-01],
    [-1.490711003541950e00, -5.971195027232170e-01, -1.509506493806840e00],
    [-1.587433911859990e00, -1.017917685210700e00, -1.456502400338650e00],
    [-6.204553022980690e-01, -1.219155095517640e00, -7.125178799033161e-01],
    [-8.804024234414100e-01, -3.662307761609550e00, -7.348611745983360e-01]])

# tests/test_integrator_rouse.cpp:79-103 (two chained calls, 1e-6)
ROUSE_ANS1 = np.array([
    [2.014655995894740e00, 3.266955094952210e00, 6.195918791485830e-01],
    [6.898958025934120e-01, 8.753981552180370e-01, 4.785034222762240e-01],
    [-3.793564324297950e-01, 1.489672909282450e00, -6.390832004693100e-01],
    [-1.130337343526490e00, -6.146915212957660e-01, -5.621008935576799e-02],
    [-6.222945452675470e-01, -1.038001996615710e00, 5.028959245456580e-01],
    [-2.989398757279950e-01, -8.209160875818450e-01, 1.289920072746550e00],
    [2.453406088052380e-01, 2.086869672299050e-01, 4.316992884969550e-01],
    [-1.110128512089560e-01, -5.552336867493930e-01, -1.277926546344950e00],
    [-6.424329445947624e-02, -1.036741598612170e00, -1.699024281591250e00],
    [-1.915098093214250e-01, -2.620904430772130e00, 2.403765459336170e-01]])
ROUSE_ANS2 = np.array([
    [2.358466969676050e00, 3.680650287690870e00, 7.391286505833830e-01],
    [9.840637226840210e-02, 1.553897668605920e00, -4.535516789623883e-02],
    [3.703216834718498e-03, 1.100513005729840e00, -5.311631291023066e-02],
    [-1.762304275149680e00, -2.808383016377780e-02, 2.665681695031340e-01],
    [-1.122405351207410e00, -4.065182692363470e-01, 9.790414477264030e-01],
    [6.632472866300509e-02, -4.887170934587170e-01, 9.806758704379520e-01],
    [-5.063576758472540e-01, -9.911910816941520e-01, 1.377974053682720e-01],
    [2.248130548550660e-01, -9.103044958788740e-01, -1.223068999396410e00],
    [-2.639644443515610e-01, -1.153796153975820e00, -1.197611287878650e00],
    [1.905555228121980e00, -2.718416492074990e00, -5.090337805446200e-01]])

R0_FULL = np.array([
    [2.65804947113558121e-01, -1.51536415609088593e+00, 8.89717445791181061e-01],
    [-3.18394386622953773e-01, -2.64563101057035599e+00, 3.05587343736082095e-01],
    [-1.06546947390811186e+00, -2.81050526243381427e+00, -8.66886662385306539e-01],
    [-3.80618256309927572e-01, -3.82549992982231757e+00, -1.88090201782288413e-01],
    [-5.64124385340994294e-01, -2.59876961675185969e+00, 4.61107949479781398e-01],
    [-4.44828448905099816e-01, -1.75076197700389291e+00, 1.56865096105042268e+00],
    [1.09391750381430619e-01, -5.96707210933647536e-01, 1.00208448868080779e+00],
    [-4.45784116176970946e-01, 6.04428098566049954e-01, 5.44860555072567698e-01],
    [2.64959429939093938e-02, -1.65351468695490667e-01, -5.24901550975400810e-01],
    [-9.06054690540470387e-02, 1.14938236930606941e+00, -5.82515809656913852e-02],
    [-2.22990478928380753e-01, 2.45965545676120811e+00, 4.16783833978659535e-01],
    [1.21010683852806888e-01, 2.02117200844250133e+00, -8.67504087616699393e-01],
    [7.08211160006220863e-01, 3.22939227477250590e+00, -4.73296795910092771e-01],
    [1.69168368716962858e+00, 3.53380767141162089e+00, -1.42203967709754631e+00],
    [6.10216843728933034e-01, 2.91075275304230985e+00, -7.87822021056476252e-01]])

R0_SINGULAR = np.array([
    [3.37650519427394511e+03, 1.24952177883608638e+04, 3.20617475621448102e+04],
    [3.37624995705613946e+03, 1.24942727222040558e+04, 3.20593226363807626e+04],
    [3.37650519236580794e+03, 1.24952177878348302e+04, 3.20617475625506631e+04],
    [3.37624995686883312e+03, 1.24942727219768185e+04, 3.20593226364890361e+04],
    [3.37650519256962343e+03, 1.24952177881101998e+04, 3.20617475624218932e+04],
    [3.37624995602158970e+03, 1.24942727220022407e+04, 3.20593226365683076e+04],
    [3.37650519275310990e+03, 1.24952177881040006e+04, 3.20617475624049948e+04],
    [3.37624995676256776e+03, 1.24942727216464009e+04, 3.20593226366290000e+04],
    [3.37650519310864456e+03, 1.24952177884205630e+04, 3.20617475622441998e+04],
    [3.37624995673909280e+03, 1.24942727225040362e+04, 3.20593226362972237e+04],
    [3.37650519343873839e+03, 1.24952177891145147e+04, 3.20617475619389988e+04],
    [3.37624995507601216e+03, 1.24942727224786813e+04, 3.20593226364821530e+04],
    [3.37650519262076159e+03, 1.24952177891598494e+04, 3.20617475620074329e+04],
    [3.37624995656296505e+03, 1.24942727235068232e+04, 3.20593226359249456e+04],
    [3.37650519312531151e+03, 1.24952177882188316e+04, 3.20617475623210667e+04],
    [3.37624995772563670e+03, 1.24942727234146205e+04, 3.20593226358385036e+04],
    [3.37650519213303596e+03, 1.24952177875682719e+04, 3.20617475626790474e+04],
    [3.37624995747488310e+03, 1.24942727227043870e+04, 3.20593226361416964e+04],
    [3.37650519216934890e+03, 1.24952177873712444e+04, 3.20617475627520143e+04],
    [3.37624995681034625e+03, 1.24942727225795206e+04, 3.20593226362603054e+04],
    [3.37650519254607570e+03, 1.24952177876328660e+04, 3.20617475626104024e+04]])
SINGULAR_HSTAR = 1.18408796570133590e-01

SHEAR = [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def random_walk(n, seed, bond=1.0):
    """Gaussian random walk of n beads, centred (a synthetic input, seeded)."""
    rs = np.random.default_rng(seed)
    R = np.zeros((n, 3))
    R[1:] = np.cumsum(rs.normal(scale=bond, size=(n - 1, 3)), axis=0)
    return R - R.mean(axis=0)


def phys_dict(**kw):
    """Plain description of a PhysParams (turned into a CPhys by the tests)."""
    d = dict(spring=2, sqrtb=10.0, nat=0.0, ev=0, zstar=0.0, dstar=0.0, min_cutoff=0.7,
             max_cutoff=1.5, cd=1, phi=None, bend=0, C=0.0, kappa=[0.0] * 9,
             flow_times=None, flow_kappas=None, external=(), hstar=0.0, method=0, mult=1.0,
             fd=0.0025, equil=0)
    d.update(kw)
    return d


def sim_dict(t0=0.0, t1=1.0, dt=0.01, tol=1e-3, com=1):
    return dict(t0=t0, t1=t1, dt=dt, tol=tol, com=com)


def wi_kappa(wi):
    k = [0.0] * 9
    k[1] = wi
    return k


# ---- trajectory cases: (name, N, phys, sim, seeds, R0 maker, segments) ------
def trajectory_cases():
    cases = []
    # config 1 physics: FENE-Fraenkel sqrtb=2, Q0=1, HI h*=0.3 Chebyshev, shear 1
    for seed in range(4):
        cases.append(dict(name=f"c1_ff_hi_seed{seed}", N=10,
                          phys=phys_dict(spring=6, sqrtb=2.0, nat=1.0, hstar=0.3, kappa=SHEAR),
                          sim=sim_dict(0.0, 0.99, 0.01, 1e-3), seed=seed,
                          R0=("ff_chain", 10, seed + 100), segments=1))
    # headline physics: FENE b=100, h*=0.3 Chebyshev, SDK EV, Wi round robin
    for i, wi in enumerate([0.1, 1.0, 10.0, 100.0]):
        cases.append(dict(name=f"n50_fene_sdk_wi{wi:g}", N=50,
                          phys=phys_dict(spring=2, sqrtb=10.0, hstar=0.3, ev=3, dstar=1.0,
                                         kappa=wi_kappa(wi)),
                          sim=sim_dict(0.0, 0.099, 1e-3, 1e-3), seed=i,
                          R0=("walk", 50, 1000 + i), segments=1))
    # config 2 physics: N=30 FENE b=50, Cholesky, Gauss EV, planar extension 0.5
    for i in range(2):
        cases.append(dict(name=f"n30_chol_gauss_planar_{i}", N=30,
                          phys=phys_dict(spring=2, sqrtb=50 ** 0.5, hstar=0.25, method=1, ev=1,
                                         zstar=1.0, dstar=0.5,
                                         kappa=[0.5, 0, 0, 0, -0.5, 0, 0, 0, 0]),
                          sim=sim_dict(0.0, 0.099, 1e-3, 1e-3), seed=i,
                          R0=("walk", 30, 2000 + i), segments=1))
    # the validation 'full' case: FENE-Fraenkel + bending + LJ + HI Cholesky + shear
    cases.append(dict(name="full_case", N=15,
                      phys=phys_dict(spring=6, sqrtb=3.0, nat=1.5, bend=1, C=1.5, ev=2,
                                     zstar=1.0, dstar=1.0, hstar=0.25, method=1, kappa=SHEAR),
                      sim=sim_dict(0.0, 0.5, 0.005, 1e-12), seed=11, R0=("full",), segments=2))
    # the same physics with Chebyshev, segmented (cold Lanczos at each boundary)
    cases.append(dict(name="full_case_cheb_segments", N=15,
                      phys=phys_dict(spring=6, sqrtb=3.0, nat=1.5, bend=1, C=1.5, ev=2,
                                     zstar=1.0, dstar=1.0, hstar=0.25, method=0, kappa=SHEAR),
                      sim=sim_dict(0.0, 0.245, 0.005, 1e-10), seed=11, R0=("full",),
                      segments=2))
    # other spring laws with HI Chebyshev
    for law, sqrtb, nat in [(1, 10.0, 0.0), (3, 10.0, 0.0), (4, 10.0, 0.0), (5, 3.0, 1.0),
                            (7, 10.0, 5.0)]:
        cases.append(dict(name=f"law{law}_hi", N=12,
                          phys=phys_dict(spring=law, sqrtb=sqrtb, nat=nat, hstar=0.2,
                                         kappa=wi_kappa(2.0)),
                          sim=sim_dict(0.0, 0.099, 1e-3, 1e-6), seed=law,
                          R0=("walk", 12, 3000 + law), segments=1))
    # external forces (constant + time-varying) and a time-varying flow, free draining
    cases.append(dict(name="external_tv_flow", N=8,
                      phys=phys_dict(spring=2, sqrtb=10.0,
                                     external=((0, True, (0.0, 0.0, -3.0), (), ()),
                                               (-1, False, (0, 0, 0), (0.0, 0.5),
                                                ((0.0, 0.0, 0.0), (0.0, 0.0, 6.0))),
                                               (3, True, (1.0, -2.0, 0.5), (), ())),
                                     flow_times=[0.0, 0.3, 0.8],
                                     flow_kappas=[[0.0] * 9, wi_kappa(1.0),
                                                  [2.0, 0, 0, 0, -1.0, 0, 0, 0, -1.0]]),
                      sim=sim_dict(0.0, 0.99, 0.01, 1e-6, 0), seed=7,
                      R0=("walk", 8, 4000), segments=1))
    # bending + LJ with HI Chebyshev, larger chain (two waves per workgroup)
    cases.append(dict(name="n80_bend_lj_cheb", N=80,
                      phys=phys_dict(spring=2, sqrtb=10.0, bend=1, C=5.0, ev=2, zstar=0.5,
                                     dstar=1.0, max_cutoff=2.5, hstar=0.2),
                      sim=sim_dict(0.0, 0.0245, 5e-4, 1e-3), seed=2,
                      R0=("walk", 80, 5000), segments=1))
    # the Chebyshev -> Cholesky fallback of every step (mobility.cpp:71-80): an
    # FD tolerance no series can meet, so L grows to MAXCHEB and the step takes
    # the dense factor; a one-wave chain (config 1 physics) and a two-wave one
    cases.append(dict(name="fallback_n10_ff", N=10,
                      phys=phys_dict(spring=6, sqrtb=2.0, nat=1.0, hstar=0.3, kappa=SHEAR,
                                     fd=-1.0),
                      sim=sim_dict(0.0, 0.19, 0.01, 1e-3), seed=21,
                      R0=("ff_chain", 10, 121), segments=1))
    cases.append(dict(name="fallback_n70_gauss", N=70,
                      phys=phys_dict(spring=2, sqrtb=10.0, hstar=0.25, ev=1, zstar=0.5,
                                     dstar=0.8, kappa=wi_kappa(1.0), fd=-1.0),
                      sim=sim_dict(0.0, 0.009, 1e-3, 1e-3), seed=22,
                      R0=("walk", 70, 5100), segments=1))
    # a corrector that never meets its tolerance (tol 0): 100 N + 1 passes per
    # step, the step accepted regardless (integrator.cpp:140-163)
    cases.append(dict(name="nonconverged_n8_hi", N=8,
                      phys=phys_dict(spring=2, sqrtb=10.0, hstar=0.2, kappa=wi_kappa(1.0)),
                      sim=sim_dict(0.0, 0.004, 1e-3, 0.0), seed=23,
                      R0=("walk", 8, 5200), segments=1))
    return cases


def long_trajectory_cases():
    """100-step trajectories of the large configurations: two chains of
    N=256 with the uneven workload's physics (FENE b=100, h*=0.3 Chebyshev,
    SDK, shear Wi/tau_1 with Wi 1 and 100, dt=1e-3) and the first two chains of
    the N=1000 workload (FENE b=100, bending 5, LJ sigma 1 epsilon 0.5 cut 2.5,
    h*=0.2 Chebyshev, no flow, dt=5e-4, self-avoiding starts of seed 2). Their
    expected per-step decisions are stored with the positions, since the
    oracle takes minutes over them."""
    cases = []
    for i, wi in enumerate([1.0, 100.0]):
        k = [0.0] * 9
        k[1] = wi * np.sin(np.pi / (2.0 * 256)) ** 2
        cases.append(dict(name=f"c3_n256_wi{wi:g}", N=256,
                          phys=phys_dict(spring=2, sqrtb=10.0, hstar=0.3, ev=3, dstar=1.0,
                                         kappa=k),
                          sim=sim_dict(0.0, 0.099, 1e-3, 1e-3), seed=i,
                          R0=("walk", 256, 6000 + i), segments=1))
    for i in range(2):
        cases.append(dict(name=f"c4_n1000_{i}", N=1000,
                          phys=phys_dict(spring=2, sqrtb=10.0, bend=1, C=5.0, ev=2, zstar=0.5,
                                         dstar=1.0, max_cutoff=2.5, hstar=0.2),
                          sim=sim_dict(0.0, 0.0495, 5e-4, 1e-3), seed=i,
                          R0=("long", 1000, i), segments=1))
    return cases


def failure_cases():
    """Inputs that end a trajectory with a failure: the reference aborts the
    process (rtsafe, numerics.hpp:33-36,71-72; WLC bounded sigma > 1,
    spring.cpp:101-105) or throws (cholesky_lower, linalg.cpp:16-24); the
    library and the oracle report a status instead. WLC-bounded springs
    (law 7) free draining: sigma = 1.2 > 1; sigma = 1 (every residual of the
    bond equation evaluates to an infinity or NaN, so rtsafe's iterations never
    settle); sigma = -1 with a compact chain and dt = 1 (the bond equation is
    positive at both ends of the bracket), at the first corrector pass and,
    with sigma = -5 and dt = 0.01, after a completed step. The degenerate
    configuration under Chebyshev with an unattainable FD tolerance: the
    fallback factor fails at pivot 55."""
    return [
        dict(name="wlcb_sigma_above_1", N=8, phys=phys_dict(spring=7, sqrtb=10.0, nat=12.0),
             sim=sim_dict(0.0, 0.004, 1e-3, 1e-6), seed=31, R0=("walk", 8, 5300),
             status=5),
        dict(name="wlcb_sigma_1_rtsafe_maxit", N=8, phys=phys_dict(spring=7, sqrtb=10.0, nat=10.0),
             sim=sim_dict(0.0, 0.004, 1e-3, 1e-6), seed=32, R0=("walk", 8, 5301),
             status=3),
        dict(name="wlcb_negative_rtsafe_unbracketed", N=8,
             phys=phys_dict(spring=7, sqrtb=10.0, nat=-10.0),
             sim=sim_dict(0.0, 2.0, 1.0, 1e-6), seed=33, R0=("walk", 8, 5302, 0.05),
             status=2),
        dict(name="wlcb_negative_unbracketed_later", N=8,
             phys=phys_dict(spring=7, sqrtb=10.0, nat=-50.0),
             sim=sim_dict(0.0, 0.05, 1e-2, 1e-6), seed=34, R0=("walk", 8, 5302, 0.3),
             status=2),
        dict(name="singular_cheb_fallback_not_pd", N=21,
             phys=phys_dict(spring=1, sqrtb=1000.0, hstar=SINGULAR_HSTAR, method=0, fd=-1.0),
             sim=sim_dict(0.0, 0.0, 0.01, 1e-6), seed=6, R0=("singular",), status=1),
    ]


def ff_chain(n, seed, sigma=1.0, dQ=2.0):
    """FENE-Fraenkel equilibrium chain with random bond directions: inverse-CDF
    bond lengths (the procedure of python/bdsim/initial.py:29-74)."""
    rs = np.random.default_rng(seed)
    alpha = dQ * dQ
    lo, hi = max(0.0, sigma - dQ), sigma + dQ
    Q = np.linspace(lo, hi, 10000)
    Q[0] += 1e-7
    Q[-1] -= 1e-7
    psi = Q * Q * (1.0 - (Q - sigma) ** 2 / alpha) ** (alpha / 2.0)
    cdf = np.concatenate(([0.0], np.cumsum(0.5 * (psi[:-1] + psi[1:]) * np.diff(Q))))
    cdf /= cdf[-1]
    Ql = np.interp(rs.random(n - 1), cdf, Q)
    u = rs.normal(size=(n - 1, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    R = np.zeros((n, 3))
    R[1:] = np.cumsum(u * Ql[:, None], axis=0)
    return R


def make_R0(spec):
    kind = spec[0]
    if kind == "walk":
        return random_walk(spec[1], spec[2], *spec[3:])
    if kind == "ff_chain":
        return ff_chain(spec[1], spec[2])
    if kind == "full":
        return R0_FULL.copy()
    if kind == "singular":
        return R0_SINGULAR.copy()
    if kind == "long":
        # chain i of the N=1000 workload's self-avoiding starts (seed 2)
        from bdsim_amd.workloads import self_avoiding_walk
        rs = np.random.default_rng(2)
        for _ in range(spec[2]):
            self_avoiding_walk(spec[1], rs)
        return self_avoiding_walk(spec[1], rs)
    raise ValueError(kind)
