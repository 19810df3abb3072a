"""Pins the FP64 CPU oracle (oracle/bdsim_oracle.c).

Each part is checked against (a) the reference's own known-answer tests
(IsaacPincus/bdsim tests/*.cpp, values copied as data), (b) the golden fixtures
written from the reference core (tests/golden/golden.npz, bit-exact), and
(c) what the mathematics fixes independently (GF(2) matrix powers for the
xorshift, modular exponentiation for minstd, exact eigen-decompositions,
closed-form pair forces, analytic identities). CPU only.
"""
import ctypes
import json
from ctypes import byref, c_double, c_int

import numpy as np
import pytest

import helpers as H
from bdsim_amd import _abi as A
from golden import cases as C


def ran1(orc, seed, n, state=None):
    st = state if state is not None else H.rng_of(seed)
    x = np.empty(n)
    orc.bo_ran1(byref(st), n, A.dptr(x))
    return x, st


# --------------------------------------------------------------------------
# RNG (rng.cpp:24-70)
# --------------------------------------------------------------------------

def test_rng_kat_seed10(orc):
    # tests/test_rng.cpp:16-20 (Fortran tests.f90 oracle), tolerance 1e-6
    expected = [0.226473569869995, 0.357589989900589, 0.217218622565269, 0.688211202621460,
                8.224623650312424e-002, 0.238840281963348, 0.956106066703796,
                0.993510127067566, 0.167083948850632, 0.341053307056427]
    x, _ = ran1(orc, 10, 10)
    assert np.max(np.abs(x - expected)) < 1e-6


def test_rng_golden_streams_bit_exact(orc, golden):
    for k, s in enumerate(golden["rng_seeds"]):
        x, st = ran1(orc, int(s), golden["rng_draws"].shape[1])
        assert np.array_equal(x, golden["rng_draws"][k].astype(np.float64)), s
        y = np.empty(100000 - x.size)
        orc.bo_ran1(byref(st), y.size, A.dptr(y))
        assert (st.ix, st.iy, st.seed) == tuple(golden["rng_state_after_long"][k])


def _gf2_xorshift_matrix():
    I = np.eye(32, dtype=np.int64)

    def shl(k):
        M = np.zeros((32, 32), dtype=np.int64)
        for i in range(32 - k):
            M[i + k, i] = 1
        return M

    def shr(k):
        M = np.zeros((32, 32), dtype=np.int64)
        for i in range(k, 32):
            M[i - k, i] = 1
        return M
    return ((I + shl(5)) @ (I + shr(17)) @ (I + shl(13))) % 2


def _matpow2(M, k):
    R = np.eye(32, dtype=np.int64)
    while k:
        if k & 1:
            R = (R @ M) % 2
        M = (M @ M) % 2
        k >>= 1
    return R


def _bits(x):
    return np.array([(x >> i) & 1 for i in range(32)], dtype=np.int64)


def _from_bits(b):
    return int(sum(int(v) << i for i, v in enumerate(b)))


@pytest.mark.parametrize("seed", [0, 5, 10, 4095, -7])
def test_rng_closed_form(orc, seed):
    """Draw k restated as linear algebra: xorshift state T^k x0 over GF(2),
    minstd state iy0 * 16807^k mod (2^31-1), then am*result in binary32."""
    s = abs(seed)
    ix0 = (777755555 ^ s) & 0xFFFFFFFF
    iy0 = (888889999 ^ s) | 1
    T = _gf2_xorshift_matrix()
    am = np.float32(np.nextafter(np.float32(1.0), np.float32(-1.0))) / np.float32(2147483647)
    ks = [1, 2, 3, 17, 150, 1001]
    x, st = ran1(orc, seed, max(ks))
    for k in ks:
        xs = _from_bits((_matpow2(T, k) @ _bits(ix0)) % 2)
        iy = (iy0 * pow(16807, k, 2147483647)) % 2147483647
        result = ((xs ^ iy) & 0x7FFFFFFF) | 1
        u = np.float32(am) * np.float32(result)
        assert x[k - 1] == np.float64(u), k
    assert st.seed == s + 1


def test_rng_save_restore(orc):
    # tests/test_rng.cpp:51-97: a restored state replays the same stream
    _, st = ran1(orc, 10, 3)
    snap = A.CRng(st.ix, st.iy, st.seed)
    a, _ = ran1(orc, 0, 5, st)
    b, _ = ran1(orc, 0, 5, A.CRng(snap.ix, snap.iy, snap.seed))
    assert np.array_equal(a, b) and np.all(a != 0.0)


def test_raw_noise_is_ottinger_transform(orc):
    # mobility.cpp:14-29: X = (c1 y^2 + c2) y, y = u - 1/2, draws bead-major
    u, _ = ran1(orc, 3, 30)
    st = H.rng_of(3)
    X = np.empty(30)
    orc.bo_raw_noise(byref(st), 10, A.dptr(X))
    y = u - 0.5
    assert np.allclose(X, (14.14855378 * y * y + 1.21569221) * y, rtol=1e-15, atol=0)
    # second and fourth moments of the deviate (Ottinger's design): E[X^2]=1, E[X^4]=3
    xs = np.empty(3 * 200000)
    st = H.rng_of(99)
    orc.bo_raw_noise(byref(st), 200000, A.dptr(xs))
    assert abs(np.mean(xs ** 2) - 1.0) < 0.01
    assert abs(np.mean(xs ** 4) - 3.0) < 0.05


# --------------------------------------------------------------------------
# forces
# --------------------------------------------------------------------------

def _force(orc, name, R, d):
    R = np.ascontiguousarray(R, dtype=np.float64)
    p = H.Packed(d, R.shape[0])
    F = np.empty_like(R)
    getattr(orc, name)(R.shape[0], A.dptr(R), byref(p.c), A.dptr(F))
    return F


def golden_phys(golden, key):
    d = json.loads(str(golden[key + "/phys"]))
    if key + "/phi" in golden.files:
        d["phi"] = golden[key + "/phi"]
    return d


def test_forces_golden_bit_exact(orc, golden):
    names = sorted({k.split("/")[1] for k in golden.files if k.startswith("force/")})
    fn = {"spring": "bo_spring_force", "ev": "bo_ev_force", "bend": "bo_bending_force",
          "total": "bo_total_force"}
    assert len(names) > 40
    for nm in names:
        kind = str(golden[f"force/{nm}/kind"])
        d = golden_phys(golden, f"force/{nm}")
        F = _force(orc, fn[kind], golden[f"force/{nm}/R"], d)
        assert np.array_equal(F, golden[f"force/{nm}/F"]), nm


def test_spring_force_hookean_is_rouse(orc):
    # tests/test_spring_force.cpp: Hookean => F_i = q_i - q_{i-1}
    R = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [2, 1, 0]], dtype=float)
    F = _force(orc, "bo_spring_force", R, C.phys_dict(spring=1, sqrtb=1000.0))
    q = np.diff(R, axis=0)
    exp = np.zeros_like(R)
    exp[0] = q[0]
    exp[-1] = -q[-1]
    exp[1:-1] = q[1:] - q[:-1]
    assert np.allclose(F, exp, atol=1e-15)


BENDING_KATS = [  # tests/test_bending.cpp:31-73 (stiffness, R, expected F), 1e-6
    (1.0, [[0, 1, 0], [0, 0, 0], [1, 0, 0]], [[-1, 0, 0], [1, 1, 0], [0, -1, 0]]),
    (1.0, [[-3 ** 0.5 / 2, 0.5, 0], [0, 0, 0], [1, 0, 0]],
     [[-0.25, -3 ** 0.5 / 4, 0], [0.25, (3 ** 0.5 + 2) / 4, 0], [0, -0.5, 0]]),
    (1.0, [[0, 0, 0], [1, 0, 0], [1, -1, 0], [1 + 2 ** 0.5 / 2, -1 + 2 ** 0.5 / 2, 0]],
     [[0, 1, 0], [-1 - 2 ** 0.5 / 2, -1, 0], [(1 + 2 ** 0.5) / 2, 0.5, 0], [0.5, -0.5, 0]]),
    (1.0, [[0, 0, 0], [2 ** 0.5 / 2, 2 ** 0.5 / 2, 0], [2 ** 0.5 / 2 + 1, 2 ** 0.5 / 2, 0],
           [2 ** 0.5 / 2 + 1 + 3 ** 0.5 / 2, 2 ** 0.5 / 2 + 0.5, 0],
           [2 ** 0.5 / 2 + 1 + 3 ** 0.5, 2 ** 0.5 / 2, 0]],
     [[-0.5, 0.5, 0], [0.5, -2 ** 0.5 / 2 - 1, 0],
      [-0.25 - 3 ** 0.5 / 4, 1.25 + 3 ** 0.5 / 4 + 2 ** 0.5 / 2, 0],
      [0.25, -1.5 - 3 ** 0.5 / 4, 0], [3 ** 0.5 / 4, 0.75, 0]]),
    (2.0, [[-5.1645947419192162e-01, 7.9748452989683460e-01, -4.1121562443647974e-01],
           [-9.3988831202840872e-01, 1.1590716947448900e+00, -7.6310486638258535e-01],
           [-6.7034760405534555e-01, -1.4053776613761966e+00, -2.9724073830375936e-01],
           [1.1832133384385024e+00, -3.6608832152898463e-01, 6.1365584115473881e-01],
           [9.4348205183717349e-01, -1.8509024173654354e-01, 8.5790538796808580e-01]],
     [[1.0507847458911022e+00, 1.8075811550046665e+00, 5.9298653220518749e-01],
      [-1.2494310241030562e+00, -1.8333141087409899e+00, -6.1970590685649796e-01],
      [4.1982689539550588e-01, 3.4148812263349709e-01, -7.8361711242102261e-01],
      [-4.2420142390715965e+00, -2.7289683808745142e+00, -1.3478360326510446e+00],
      [4.0208336218880447e+00, 2.4132132119773404e+00, 2.1581725197233776e+00]]),
]


@pytest.mark.parametrize("k", range(len(BENDING_KATS)))
def test_bending_kats(orc, k):
    Cst, R, exp = BENDING_KATS[k]
    F = _force(orc, "bo_bending_force", np.array(R, dtype=float), C.phys_dict(bend=1, C=Cst))
    assert np.max(np.abs(F - np.array(exp))) < 1e-6


def test_bending_is_gradient_of_energy(orc):
    """F = -dU/dR for U = C sum (1 - cos theta): central differences."""
    R = C.random_walk(7, 21)
    Cst = 1.7

    def U(R):
        b = np.diff(R, axis=0)
        u = b / np.linalg.norm(b, axis=1)[:, None]
        return Cst * np.sum(1.0 - np.sum(u[1:] * u[:-1], axis=1))
    F = _force(orc, "bo_bending_force", R, C.phys_dict(bend=1, C=Cst))
    h = 1e-6
    G = np.zeros_like(R)
    for i in range(R.shape[0]):
        for d in range(3):
            Rp, Rm = R.copy(), R.copy()
            Rp[i, d] += h
            Rm[i, d] -= h
            G[i, d] = -(U(Rp) - U(Rm)) / (2 * h)
    assert np.max(np.abs(F - G)) < 1e-7


def test_ev_pair_forces_closed_form(orc):
    """Brute-force pair sums in another order against closed forms of the
    Gauss and LJ potentials' forces (excluded_volume.cpp:36-47)."""
    R = C.random_walk(9, 22, 0.9)
    for d in (C.phys_dict(ev=1, zstar=1.3, dstar=0.7),
              C.phys_dict(ev=2, zstar=0.8, dstar=1.0, max_cutoff=2.0)):
        F = _force(orc, "bo_ev_force", R, d)
        exp = np.zeros_like(R)
        for mu in range(R.shape[0]):
            for nu in range(R.shape[0]):
                if nu == mu:
                    continue
                q = R[nu] - R[mu]
                r = max(np.linalg.norm(q), 1e-6)
                if d["ev"] == 1:
                    fac = d["zstar"] / d["dstar"] ** 5 * np.exp(-r * r / (2 * d["dstar"] ** 2))
                else:  # -dU/dr / r of 4 z ((d/r)^12 - (d/r)^6), soft core below Rmin
                    rr = max(r, d["min_cutoff"])
                    fac = (48 * d["zstar"] * d["dstar"] ** 12 / rr ** 14
                           - 24 * d["zstar"] * d["dstar"] ** 6 / rr ** 8) if r <= d["max_cutoff"] else 0.0
                exp[mu] += -fac * q
        assert np.allclose(F, exp, rtol=1e-12, atol=1e-12)
        assert np.allclose(F.sum(axis=0), 0.0, atol=1e-12)   # Newton's third law


def test_ev_reference_sign_checks(orc):
    # tests/test_ev.cpp:15-48
    R = np.array([[0, 0, 0], [1, 0, 0], [2.1, 0, 0], [3, 0, 0], [0, 1.5, 0]], dtype=float)
    F = _force(orc, "bo_ev_force", R, C.phys_dict(ev=2, zstar=1.0, dstar=1.0, cd=4))
    assert abs(F[0, 1]) > np.finfo(float).eps and abs(F[-1, 1]) > np.finfo(float).eps
    assert np.all(F[1:-1, 0] == 0.0)


# --------------------------------------------------------------------------
# springs: implicit solve
# --------------------------------------------------------------------------

def _solve(orc, law, dtby4, g, nat):
    r, ff = c_double(), c_double()
    st = orc.bo_solve_implicit_r(law, dtby4, g, nat, byref(r), byref(ff))
    return st, r.value, ff.value


def test_spring_solve_kats(orc):
    # tests/test_spring.cpp:25-60
    st, r, ff = _solve(orc, 7, 0.001 / 4.0, 5.0, 0.8)
    assert abs(r - 0.999354532073543) < 1e-6 and abs(ff - 1.601291769648808e4) < 1e-2
    st, r, ff = _solve(orc, 7, 0.001 / 4.0, 0.0, 0.8)
    assert abs(r - 0.601489231399178) < 1e-6 and abs(ff + 4.000000000000779e3) < 1e-2
    st, r, ff = _solve(orc, 7, 0.001 / 4.0, 1.0e5, 0.8)
    assert abs(r - 0.999995917496677) < 1e-6 and abs(ff - 3.999976329984242e8) < 1e2
    st, r, ff = _solve(orc, 7, 0.2 / 4.0, 2.0 * np.finfo(float).eps, 0.0)
    assert abs(r) < 1e-6
    a, b = _solve(orc, 7, 0.05, 1.0, 0.0), _solve(orc, 4, 0.05, 1.0, 0.0)
    assert abs(a[1] - b[1]) < 1e-6 and abs(a[2] - b[2]) < 1e-6
    for i in range(1, 101, 7):
        for j in range(1, 101, 9):
            g = (j * 0.5) / (i * 0.1)
            assert abs(_solve(orc, 2, 0.05, g, 0.0)[1] - _solve(orc, 6, 0.05, g, 0.0)[1]) < 1e-6


@pytest.mark.parametrize("law,nat", [(2, 0.0), (3, 0.0), (4, 0.0), (6, 0.0), (6, 0.5), (1, 0.0),
                                     (5, 0.3), (7, 0.0)])
def test_spring_solve_satisfies_equation(orc, law, nat):
    """r (1 + dt/4 f(r)) = gamma at the returned root, away from the clamps."""
    fsh = {1: lambda q: 1.0, 2: lambda q: 1 / (1 - q * q), 3: lambda q: (3 - q * q) / (1 - q * q) / 3,
           4: lambda q: 1 / (6 * q) * (4 * q + 1 / (1 - q) ** 2 - 1),
           5: lambda q: 1 - nat / q, 6: lambda q: (1 - nat / q) / (1 - (q - nat) ** 2),
           7: lambda q: (2 / (3 * q)) * (((1 - q) ** -2 - 1) / 4 + q)}
    dtby4 = 0.0025
    for g in np.linspace(0.05, 2.0, 40):
        st, r, ff = _solve(orc, law, dtby4, g, nat)
        assert st == 0
        if law in (2, 3, 4, 6, 7) and (r >= 1.0 - 1e-6 - 1e-12 or r >= nat + 1.0 - 1e-6 - 1e-12):
            continue
        if law == 4 and g > 100:
            continue
        assert abs(r * (1 + dtby4 * fsh[law](r)) - g) < 1e-10 * max(1.0, g), (law, g, r)


def test_spring_solve_bounded(orc):
    # tests/test_hi_stability.cpp:71-90
    for g in [1.0, 1e2, 1e4, 1e6, 1e8, 1e10, 1e14, 1e20]:
        r = _solve(orc, 6, 0.01 / 4.0, g, 0.5)[1]
        assert 0.0 <= r <= 1.5 and np.isfinite(r)
    for g in [1.0, 1e4, 1e10, 1e20]:
        r = _solve(orc, 2, 0.01 / 4.0, g, 0.0)[1]
        assert 0.0 < r < 1.0


def test_spring_solve_golden(orc, golden):
    keys = sorted({k.rsplit("/", 1)[0] for k in golden.files if k.startswith("spring_solve/")})
    for key in keys:
        law, nat, dtby4 = key.split("/")[1].split("_")
        ok = golden[key + "/ok"]
        for g, r, ff, o in zip(golden[key + "/gamma"], golden[key + "/r"], golden[key + "/ff"], ok):
            if not o:
                continue
            st, r2, ff2 = _solve(orc, int(law), float(dtby4), float(g), float(nat))
            assert r2 == r and (ff2 == ff or (np.isnan(ff) and np.isnan(ff2))), (key, g)


def test_connector_direction(orc):
    # tests/test_direction.cpp: the new bond is parallel to Gamma
    for law, sqrtb, nat in [(1, 10.0, 0.0), (2, 10.0, 0.0), (3, 10.0, 0.0), (4, 10.0, 0.0),
                            (5, 10.0, 3.0), (6, 10.0, 3.0), (7, 10.0, 0.5)]:
        for g in [(1.0, 0, 0), (0, -2.0, 0), (1, 1, 1), (-3, 2, -1), (0.3, -0.7, 2.4), (5, -4, 0.1)]:
            for dt in (0.01, 0.1, 0.5):
                ga = np.array(g, dtype=float)
                out = np.empty(3)
                orc.bo_solve_connector(A.dptr(ga), dt, law, sqrtb, nat, A.dptr(out))
                s = np.linalg.norm(np.cross(out, ga)) / (np.linalg.norm(out) * np.linalg.norm(ga))
                assert s < 1e-12 and out @ ga > 0


# --------------------------------------------------------------------------
# hydrodynamics
# --------------------------------------------------------------------------

def rpy(orc, R, hstar):
    R = np.ascontiguousarray(R, dtype=np.float64)
    n = 3 * R.shape[0]
    D = np.empty((n, n))
    orc.bo_rpy_matrix(R.shape[0], A.dptr(R), hstar, A.dptr(D))
    return D


HI_R = np.array([[0.0, 0.0, 0.0],
                 [-8.7001417024242400e+00, 5.3611027493752204e+00, 2.0995067416916502e+00],
                 [-9.1605814460059198e-01, 1.1704273266182801e+01, -1.4985387607124800e+00],
                 [-5.4793423363635396e+00, 4.8982010793314998e+00, 4.0063480623078496e+00]])
HI_X0 = np.array([[-0.653795716820965, 2.25469853815588, 0.165603302292486],
                  [-0.587509005201644, -0.745084985707505, -0.157459965970764],
                  [-1.88644397151742, 3.16619, 1.23661201418787],
                  [-0.958635132046656, -0.253142128732732, 1.43910244145363]])


def test_hi_kats(orc):
    # tests/test_hi.cpp:19-77
    D = rpy(orc, HI_R, 0.2)
    blk01 = np.array([[0.043185241172496, -0.010895545687992, -0.004266896699380],
                      [-0.010895545687992, 0.032217586367374, 0.002629298741189],
                      [-0.004266896699380, 0.002629298741189, 0.026533338541433]])
    assert np.max(np.abs(D[0:3, 3:6] - blk01)) < 1e-6
    assert np.array_equal(D[3:6, 3:6], np.eye(3))
    L = np.empty_like(D)
    assert orc.bo_cholesky_lower(12, A.dptr(D), A.dptr(L), None, None) == 0
    y = np.empty(12)
    orc.bo_lower_matvec(12, A.dptr(L), A.dptr(HI_X0.ravel().copy()), A.dptr(y))
    chol_expected = np.array([[-0.206748358961671, 0.712998281762464, 0.052368362328964],
                              [-0.202521038852097, -0.210272818508389, -0.045641895488899],
                              [-0.611043311268904, 1.022458498065220, 0.391180874642083],
                              [-0.351959054078799, -0.035305408161349, 0.461910033760980]])
    assert np.max(np.abs(np.sqrt(0.1) * y - chol_expected.ravel())) < 1e-5
    out = np.empty(12)
    nt, fd = c_int(), c_double()
    x0 = HI_X0.ravel().copy()
    orc.bo_chebyshev_sqrt_times(12, A.dptr(D), A.dptr(x0), 1.0, 1e-6, None, None, A.dptr(out),
                                byref(nt), byref(fd))
    assert fd.value < 1e-4
    bb = np.empty(12)
    orc.bo_chebyshev_sqrt_times(12, A.dptr(D), A.dptr(out), 1.0, 1e-6, None, None, A.dptr(bb),
                                byref(nt), byref(fd))
    assert np.max(np.abs(bb - D @ x0)) < 1e-4


def test_rpy_closed_form_and_spectrum(orc):
    """RPY blocks against the textbook tensor; D symmetric positive definite;
    the Chebyshev square root against the exact one from an eigendecomposition."""
    R = C.random_walk(12, 31, 0.8)
    h = 0.3
    D = rpy(orc, R, h)
    a = np.sqrt(np.pi) * h
    for mu, nu in [(0, 5), (3, 4), (7, 11)]:
        r = R[nu] - R[mu]
        s = np.linalg.norm(r)
        rr = np.outer(r, r) / s ** 2
        if s >= 2 * a:
            exp = 0.75 * a / s * ((1 + 2 * a * a / (3 * s * s)) * np.eye(3) + (1 - 2 * a * a / (s * s)) * rr)
        else:
            exp = (1 - 9 * s / (32 * a)) * np.eye(3) + 3 * s / (32 * a) * rr
        assert np.allclose(D[3 * mu:3 * mu + 3, 3 * nu:3 * nu + 3], exp, rtol=1e-12, atol=1e-14)
    assert np.array_equal(D, D.T)
    w, V = np.linalg.eigh(D)
    assert w.min() > 0
    x0 = np.random.default_rng(1).normal(size=D.shape[0])
    out = np.empty_like(x0)
    nt, fd = c_int(), c_double()
    orc.bo_chebyshev_sqrt_times(D.shape[0], A.dptr(D), A.dptr(x0), 1.0, 1e-8, None, None, A.dptr(out),
                                byref(nt), byref(fd))
    exact = V @ (np.sqrt(w) * (V.T @ x0))
    assert np.linalg.norm(out - exact) / np.linalg.norm(exact) < 1e-4
    assert abs(out @ out - x0 @ D @ x0) / (x0 @ D @ x0) < 1e-8


def test_hi_golden(orc, golden):
    for name in ("walk10", "full", "walk50", "close"):
        R = golden[f"hi/{name}/R"]
        h = float(golden[f"hi/{name}/hstar"])
        D = rpy(orc, R, h)
        if f"hi/{name}/D" in golden.files:
            assert np.array_equal(D, golden[f"hi/{name}/D"])
        x = golden[f"hi/{name}/x"].ravel().copy()
        assert np.allclose(D @ x, golden[f"hi/{name}/Dx"], rtol=0, atol=1e-13)
        assert np.array_equal([np.trace(D), np.linalg.norm(D)], golden[f"hi/{name}/D_trace_fro"])
        L = np.empty_like(D)
        assert orc.bo_cholesky_lower(D.shape[0], A.dptr(D), A.dptr(L), None, None) == 0
        if f"hi/{name}/L" in golden.files:
            assert np.array_equal(L, golden[f"hi/{name}/L"])
        assert np.allclose(np.tril(L) @ x, golden[f"hi/{name}/Lx"], rtol=0, atol=1e-13)
        out = np.empty_like(x)
        nt, fd = c_int(), c_double()
        orc.bo_chebyshev_sqrt_times(D.shape[0], A.dptr(D), A.dptr(x), 1.0, 0.0025, None, None,
                                    A.dptr(out), byref(nt), byref(fd))
        assert np.array_equal(out, golden[f"hi/{name}/cheb"].ravel())
        assert nt.value == int(golden[f"hi/{name}/cheb_nterms"])
        assert fd.value == float(golden[f"hi/{name}/cheb_fd"])


def test_spectral_bounds_warm_golden_and_containment(orc, golden):
    seq = golden["bounds/R"]
    h = float(golden["bounds/hstar"])
    n = 3 * seq.shape[1]
    seed = np.zeros(n)
    valid = c_int(0)
    colds = 0
    for k, Rk in enumerate(seq):
        D = rpy(orc, Rk, h)
        lmin, lmax, steps, cold = c_double(), c_double(), c_int(), c_int()
        orc.bo_spectral_bounds(n, A.dptr(D), A.dptr(seed), byref(valid), byref(lmin), byref(lmax),
                               byref(steps), byref(cold))
        colds += cold.value
        assert (lmin.value, lmax.value) == tuple(golden["bounds/lminmax"][k])
        w = np.linalg.eigvalsh(D)    # tests/test_hi_stability.cpp:195-240
        assert lmin.value <= w.min() and lmax.value >= w.max()
    assert colds == golden["bounds/cold_warm"][0]


def test_singular_configuration_pivot(orc, golden):
    # fortran/README.md:100-104: the R0_SINGULAR tensor fails at (0-based) pivot 55
    D = rpy(orc, C.R0_SINGULAR, C.SINGULAR_HSTAR)
    L = np.empty_like(D)
    piv, val = c_int(), c_double()
    assert orc.bo_cholesky_lower(63, A.dptr(D), A.dptr(L), byref(piv), byref(val)) == -1
    assert piv.value == 55
    assert "pivot 55" in str(golden["singular/message"])


# --------------------------------------------------------------------------
# flow / external forces
# --------------------------------------------------------------------------

def test_flow_and_external_kats(orc, golden):
    # tests/test_flow.cpp: ramp 0 -> 2 over [0,1], clamped
    uni = lambda g: [g, 0, 0, 0, -g / 2, 0, 0, 0, -g / 2]  # noqa: E731
    ts = np.ascontiguousarray([0.0, 1.0])
    ks = np.ascontiguousarray(np.array([uni(0.0), uni(2.0)]).ravel())
    fl = A.CFlow(0, 2, (ctypes.c_double * 9)(), A.dptr(ts), A.dptr(ks))
    K = np.empty(9)
    for t, want in [(0.0, 0.0), (0.5, 1.0), (1.0, 2.0), (-3.0, 0.0), (5.0, 2.0)]:
        orc.bo_kappa(byref(fl), t, A.dptr(K))
        assert abs(K[0] - want) < 1e-12
    # external-force and kappa tables against the reference
    d = C.phys_dict(external=((0, True, (0.0, 0.0, -3.0), (), ()),
                              (-1, False, (0, 0, 0), (0.0, 0.5, 1.0),
                               ((0.0, 0.0, 0.0), (0.0, 0.0, 6.0), (1.0, 2.0, 3.0))),
                              (3, True, (1.0, -2.0, 0.5), (), ()),
                              (3, True, (0.5, 0.5, 0.5), (), ()),
                              (17, True, (9.0, 9.0, 9.0), (), ())),
                    flow_times=[0.0, 0.3, 0.8],
                    flow_kappas=[[0.0] * 9, C.wi_kappa(1.0), [2.0, 0, 0, 0, -1.0, 0, 0, 0, -1.0]])
    p = H.Packed(d, 6)
    for i, t in enumerate(golden["external/t"]):
        F = np.empty((6, 3))
        orc.bo_external_force(byref(p.c), 6, float(t), A.dptr(F))
        assert np.array_equal(F, golden["external/F"][i])
        orc.bo_kappa(byref(p.c.flow), float(t), A.dptr(K))
        assert np.array_equal(K, golden["external/kappa"][i])


# --------------------------------------------------------------------------
# the integrator
# --------------------------------------------------------------------------

def test_rouse_integration_kat(orc):
    # tests/test_integrator_rouse.cpp:49-94: two chained calls on one stream, 1e-6
    p = H.Packed(C.phys_dict(spring=1, sqrtb=1000.0), 10)
    outs, _ = H.run_segments(lambda *a: H.oracle_integrate(orc, *a), 10, C.R0_ROUSE, p.c,
                             C.sim_dict(0.0, 1.0, 0.1, 1e-6), 5, 2)
    assert np.max(np.abs(outs[0][0] - C.ROUSE_ANS1)) < 1e-6
    assert np.max(np.abs(outs[1][0] - C.ROUSE_ANS2)) < 1e-6


def test_trajectories_golden_bit_exact(orc, golden):
    cases = {c["name"]: c for c in C.trajectory_cases()}
    names = [str(n) for n in golden["traj_names"]]
    assert len(names) >= 15
    for name in names:
        c = cases[name]
        N = c["N"]
        R0 = golden[f"traj/{name}/R0"]
        assert np.array_equal(R0, C.make_R0(c["R0"]))
        p = H.Packed(c["phys"], N)
        outs, st = H.run_segments(lambda *a: H.oracle_integrate(orc, *a), N, R0, p.c, c["sim"],
                                  c["seed"], c["segments"])
        for k, o in enumerate(outs):
            assert o[1] == golden[f"traj/{name}/rc"][k]
            assert np.array_equal(o[0], golden[f"traj/{name}/R"][k]), name
        assert (st.ix, st.iy, st.seed) == tuple(golden[f"traj/{name}/rng"]), name


def test_singular_integrate(orc, golden):
    p = H.Packed(C.phys_dict(spring=1, sqrtb=1000.0, hstar=C.SINGULAR_HSTAR, method=1), 21)
    st = H.rng_of(6)
    R, rc, cnt, _ = H.oracle_integrate(orc, 21, C.R0_SINGULAR, p.c,
                                       H.csim(C.sim_dict(0.0, 0.0, 0.01, 1e-6)), st)
    assert rc == A.STATUS_NOT_POSITIVE_DEFINITE and cnt.fail_index == 55
    assert int(golden["singular_integrate/rc"]) == 1
    assert (st.ix, st.iy, st.seed) == tuple(golden["singular_integrate/rng"])


def test_free_draining_rouse_statistics(orc):
    """Hookean free-draining chain: <R_ee^2> = 3(N-1) at equilibrium
    (BENCHMARKS.md:24-33, validate_physics.py:45-70); 64 chains, 3 sigma."""
    N = 8
    p = H.Packed(C.phys_dict(spring=1, sqrtb=1000.0), N)
    vals = []
    for s in range(64):
        R, rc, _, _ = H.oracle_integrate(orc, N, C.random_walk(N, 900 + s), p.c,
                                         H.csim(C.sim_dict(0.0, 40.0, 0.02, 1e-6)), H.rng_of(s))
        vals.append(np.sum((R[-1] - R[0]) ** 2))
    m, se = np.mean(vals), np.std(vals, ddof=1) / np.sqrt(len(vals))
    assert abs(m - 3 * (N - 1)) < 3.5 * se


# --------------------------------------------------------------------------
# failure semantics (statuses instead of the reference's abort / throw)
# --------------------------------------------------------------------------

def _wlcb_force_sans_hookean(r, q0):
    """WLC-bounded f(r) of spring.cpp:30-37, in numpy (the form the golden
    spring forces pin)."""
    t = (r - q0) / (1.0 - q0)
    a = 1.0 / (1.0 - t) ** 2
    b = 1.0 / (1.0 + t) ** 2
    return (2.0 / (3.0 * r)) * ((a - 1.0) / 4.0 + t - q0 * (b - 1.0) / 4.0 + q0 * t)


def test_wlcb_unbracketed_iff_bond_equation_has_no_root(orc):
    """rtsafe's bracket test (numerics.hpp:33-36) on the transformed residual
    says 'no root' exactly when the bond equation r (1 + dt/4 f(r)) = gamma
    (spring.cpp:123-172) has none on the physical range: with sigma < 0 the
    left side stays above its value near r = 0, so a small gamma has no root."""
    dtby4, sigma = 0.25, -1.0
    r = np.linspace(1e-6, 1.0 - 1e-6, 200001)
    g_min = float(np.min(r * (1.0 + dtby4 * _wlcb_force_sans_hookean(r, sigma))))
    assert g_min > 0.1
    for gam in [0.0, 0.01, 0.5 * g_min, 0.9 * g_min]:
        st, _, _ = _solve(orc, 7, dtby4, gam, sigma)
        assert st == A.STATUS_RTSAFE_UNBRACKETED, gam
    for gam in [1.1 * g_min, 2.0 * g_min, 5.0]:
        st, rr, _ = _solve(orc, 7, dtby4, gam, sigma)
        assert st == 0
        res = rr * (1.0 + dtby4 * _wlcb_force_sans_hookean(rr, sigma)) - gam
        assert abs(res) < 1e-9 * max(1.0, gam), (gam, res)


def test_wlcb_sigma_limits(orc):
    """sigma > 1 is the reference's explicit abort (spring.cpp:101-105); at
    sigma = 1 the force law divides by 1 - sigma = 0, nothing is finite, and
    rtsafe runs out of iterations (numerics.hpp:71-72)."""
    assert _solve(orc, 7, 0.00025, 0.5, 1.2)[0] == A.STATUS_WLCB_SIGMA
    with np.errstate(all="ignore"):
        assert not np.any(np.isfinite(_wlcb_force_sans_hookean(np.linspace(0.01, 0.99, 50), 1.0)))
    assert _solve(orc, 7, 0.00025, 0.5, 1.0)[0] == A.STATUS_RTSAFE_MAXIT


@pytest.mark.parametrize("name", [c["name"] for c in C.failure_cases()])
def test_failure_cases_status(orc, golden, name):
    """Each failure case ends with its status on the oracle; the degenerate
    Chebyshev case also as the reference did (golden: it threw from the
    fallback factor at pivot 55, after the step's draws)."""
    c = {x["name"]: x for x in C.failure_cases()}[name]
    p = H.Packed(c["phys"], c["N"])
    st = H.rng_of(c["seed"])
    _, rc, cnt, _ = H.oracle_integrate(orc, c["N"], C.make_R0(c["R0"]), p.c, H.csim(c["sim"]), st)
    assert rc == c["status"]
    if name == "singular_cheb_fallback_not_pd":
        assert int(golden["singular_cheb/rc"]) == 1
        assert "pivot 55" in str(golden["singular_cheb/message"]) and cnt.fail_index == 55
        assert (st.ix, st.iy, st.seed) == tuple(golden["singular_cheb/rng"])
        assert cnt.fallbacks == 1


def test_forced_fallback_and_nonconvergence_counts(orc):
    """The golden fallback cases take the dense factor on every step after
    growing L to MAXCHEB; the tolerance-0 case runs 100 N + 1 corrector passes
    per step and counts each step as non-converged."""
    cases = {c["name"]: c for c in C.trajectory_cases()}
    for nm in ("fallback_n10_ff", "fallback_n70_gauss"):
        c = cases[nm]
        p = H.Packed(c["phys"], c["N"])
        _, rc, cnt, lg = H.oracle_integrate(orc, c["N"], C.make_R0(c["R0"]), p.c, H.csim(c["sim"]),
                                            H.rng_of(c["seed"]), log_cap=32)
        assert rc == 0 and cnt.fallbacks == cnt.steps > 0
        assert all(l.cheb_L == 500 and l.fallback == 1 for l in lg)
    c = cases["nonconverged_n8_hi"]
    p = H.Packed(c["phys"], c["N"])
    _, rc, cnt, _ = H.oracle_integrate(orc, c["N"], C.make_R0(c["R0"]), p.c, H.csim(c["sim"]),
                                       H.rng_of(c["seed"]))
    assert rc == 0 and cnt.nonconverged == cnt.steps == 5
    assert cnt.corrector_iters == 5 * (100 * c["N"] + 1)


def test_long_n256_golden_bit_exact(orc, golden):
    """The two N=256 100-step cases: oracle bit for bit the reference, with the
    stored per-step decisions (the N=1000 pair was checked the same way when
    the fixture was written)."""
    cases = {c["name"]: c for c in C.long_trajectory_cases()}
    for nm in [str(n) for n in golden["long_names"] if str(n).startswith("c3_")]:
        c = cases[nm]
        p = H.Packed(c["phys"], c["N"])
        assert np.array_equal(golden[f"long/{nm}/R0"], C.make_R0(c["R0"]))
        st = H.rng_of(c["seed"])
        R, rc, cnt, lg = H.oracle_integrate(orc, c["N"], golden[f"long/{nm}/R0"], p.c,
                                            H.csim(c["sim"]), st, log_cap=128)
        assert rc == 0 and np.array_equal(R, golden[f"long/{nm}/R"]), nm
        assert (st.ix, st.iy, st.seed) == tuple(golden[f"long/{nm}/rng"])
        got = np.array([[l.lanczos_cold, l.lanczos_steps, l.cheb_L, l.cheb_attempts, l.fallback,
                         l.corrector_iters] for l in lg])
        assert np.array_equal(got, golden[f"long/{nm}/log"])
