// bdsim_device.hpp -- device-side building blocks of the MI355X ChainIntegrator.
//
// Layout conventions (one workgroup integrates one chain):
//   * thread tid owns bead tid (tid < N) and, where bonds are concerned, bond tid
//     (tid < N-1); its 3-vectors live in registers;
//   * positions used by pair loops, the mat-vec operand and bond-level arrays are
//     staged in LDS as structure-of-arrays [3][BLOCK], so a pair loop reads the
//     partner bead with one broadcast LDS read per component;
//   * the Lanczos basis and the dense mobility (Cholesky path only) live in a
//     per-trajectory global workspace, each thread touching only its own rows.
//
// Citations are to IsaacPincus/bdsim.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bdgpu {

constexpr int MAXCHEB = 500;        // types.hpp:20
constexpr int LANCZOS_COLD = 30;    // hydrodynamics.cpp:79
constexpr int LANCZOS_WARM = 8;     // hydrodynamics.cpp:80
constexpr int MAX_BEADS = 1024;
constexpr double PI = 3.14159265358979323846;
constexpr double MYEPS = 1e-6;      // types.hpp:24

// ---- per-trajectory parameters, packed on the host ----------------------
struct ExtEntry {
    int32_t bead, constant, n_times, pad;
    double value[3];
    int64_t off;                    // times at blob[off], values at blob[off+n_times]
};

struct TrajParams {
    int32_t N, spring_type, ev_type, cd;
    int32_t equil, has_phi, bend_type, hi_method;
    int32_t flow_const, flow_n, n_ext, com;
    int32_t log_cap, pad0;
    double sqrtb, q0, natural_length;
    // excluded volume constants (excluded_volume.cpp:14-34), computed on the host
    double zsbyds5, pt5bydssq, LJa, LJb, SDKa, SDKb, Rmin, Rcutg, Rcutp, RcutpSDK;
    double Rmin2, Rcutg2, Rcutp2, RcutpSDK2;   // squared radii (pair loops compare r^2)
    double bend_C;
    double kappa0[9];
    double hstar, mult, fd_tol;
    double rpy_d[6];                // RPY pair-weight constants (RpyW), FP64 and FP32
    float rpy_f[6];
    double t0, t1, dt, tol;
    int64_t phi_off, flow_off, ext_off;
    int64_t R_off;                  // doubles into the positions buffer
    int64_t ws_off;                 // doubles into the workspace buffer
    int64_t log_off;                // entries into the step-log buffer (-1: none)
};

struct Counters {                   // mirrors bdsim_counters
    int64_t steps, matvecs, lanczos_iters, lanczos_cold, lanczos_warm, cheb_terms,
        cheb_attempts, fallbacks, corrector_iters, nonconverged;
    int32_t fail_index, pad_;
    double fail_value;
};

struct StepLog {                    // mirrors bdsim_step_log
    double t, lmin, lmax, fd_err;
    int32_t lanczos_cold, lanczos_steps, cheb_L, cheb_attempts, fallback, corrector_iters;
};

struct RngState { int32_t ix, iy, seed; };

// Where a resumable segment stands between launches: the accumulated step
// time (the reference's `t += dt`), whether the warm Lanczos seed in the
// workspace is valid, and whether the segment has ended (or failed).
struct SegState { double t; int32_t seed_valid, done; };

enum : int32_t { ST_OK = 0, ST_NOT_PD = 1, ST_RTSAFE_UNBRACKETED = 2, ST_RTSAFE_MAXIT = 3,
                 ST_EXACT_SQRT = 4, ST_WLCB_SIGMA = 5, ST_NOT_RUN = 6, ST_SYNC_TIMEOUT = 7 };

// Per-trajectory workspace in HBM (doubles). The state slots hold per-bead
// vectors that live across phases of a step (structure of arrays: slot k of
// bead mu at S + k*N + mu), so they need not occupy registers meanwhile.
enum : int { SLOT_FS = 0, SLOT_FE = 3, SLOT_DW = 6, SLOT_SEED = 9, SLOT_X0 = 12, N_SLOTS = 16 };
struct WsLayout {
    int64_t S, coef, aux, tri, V, dense, total;
};
__host__ __device__ inline WsLayout ws_layout(int N, bool hi) {
    WsLayout w;
    int64_t o = 0;
    w.S = o;      o += (int64_t)N_SLOTS * N;
    w.coef = o;   o += MAXCHEB + 8;                       // Chebyshev coefficients
    w.aux = o;    o += MAXCHEB + 8;                       // sqrt at the Chebyshev nodes
    w.tri = o;    o += 2 * 3 * LANCZOS_COLD + 4;           // inverse-iteration rows
    w.V = o;      o += hi ? (int64_t)LANCZOS_COLD * 3 * N : 0;   // Lanczos basis
    w.dense = o;  o += hi ? (int64_t)9 * N * N : 0;              // dense D / its factor
    w.total = (o + 31) & ~(int64_t)31;
    return w;
}

// ---- reductions ----------------------------------------------------------
// Butterfly sum: every lane ends with the same value (IEEE addition commutes),
// and the tree is fixed, so the result is deterministic run to run.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    // every pattern used here is a permutation within a row of 16 lanes, so
    // each lane has a valid source and no previous value is kept
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// All-lanes sum: quad swaps and row mirrors by DPP give every lane its row's
// sum (each stage adds two operands that are the same pair on both partner
// lanes), then the four row sums are read to scalars and added as
// (r0 + r1) + (r2 + r3): every lane holds bitwise the same total, and the
// order is fixed.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_mov<0xB1>(v);     // quad_perm [1,0,3,2]: lane ^ 1
    v += dpp_mov<0x4E>(v);     // quad_perm [2,3,0,1]: lane ^ 2
    v += dpp_mov<0x141>(v);    // row_half_mirror: sums of 8
    v += dpp_mov<0x140>(v);    // row_mirror: sums of 16
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Orders LDS accesses between the lanes of one wave: the LDS instructions of a
// wave complete in the order they are issued, so a lane reads what another
// lane of its wave wrote before, with no s_barrier and no wait for the write;
// the wavefront-scope fences keep the compiler from moving LDS accesses across
// this point. For workgroups of a single wave only.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block-wide sum over NW waves. sRed holds 2*NW doubles; `par` alternates the
// half used so consecutive calls need only one barrier each.
template <int NW>
__device__ __forceinline__ double block_sum(double v, double* sRed, int& par) {
    v = wave_sum(v);
    if constexpr (NW == 1) {
        return v;
    } else {
        double* buf = sRed + par * NW;
        par ^= 1;
        if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
        __syncthreads();
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < NW; ++i) t += buf[i];
        return t;
    }
}

// A pointer into device memory typed as such (global_load / global_store):
// pointers that reach the kernel inside a struct are otherwise generic, and
// generic (flat) accesses also wait on the LDS counter.
template <typename T>
using gbl_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gbl_ptr<T> as_gbl(T* p) { return (gbl_ptr<T>)p; }

// ---- RNG: Numerical Recipes ran_1 with exact jump-ahead -------------------
// Thread mu needs draws 3mu+1..3mu+3 of the step. The xorshift is linear over
// GF(2): state_k = T^k state_0, applied with the 32 columns of T^(3mu). The
// minstd part is iy_k = iy_0 * 16807^k mod (2^31-1). Both tables are built on
// the host. The draws then follow rng.cpp:41-70 exactly.
__device__ __forceinline__ uint32_t gf2_apply(gbl_ptr<const uint32_t> cols, uint32_t x) {
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 32; ++i) r ^= cols[i] & (0u - ((x >> i) & 1u));
    return r;
}

__device__ __forceinline__ int32_t minstd_mulmod(int32_t iy, uint32_t a) {
    const uint64_t p = (uint64_t)(uint32_t)iy * (uint64_t)a;
    uint64_t r = (p & 0x7fffffffull) + (p >> 31);
    r = (r & 0x7fffffffull) + (r >> 31);
    if (r >= 0x7fffffffull) r -= 0x7fffffffull;
    return (int32_t)r;
}

__device__ __forceinline__ void xorshift_step(uint32_t& ux) {
    ux ^= (ux << 13);
    ux ^= (ux >> 17);
    ux ^= (ux << 5);
}

__device__ __forceinline__ void schrage_step(int32_t& iy) {
    const int32_t k = iy / 127773;
    iy = 16807 * (iy - k * 127773) - 2836 * k;
    if (iy < 0) iy += 2147483647;
}

// rng.cpp:31-39
__device__ __forceinline__ void rng_maybe_init(int32_t& ix, int32_t& iy, int32_t& seed) {
    if (seed <= 0 || iy < 0) {
        const int32_t s = seed < 0 ? -seed : seed;
        iy = (888889999 ^ s) | 1;
        ix = 777755555 ^ s;
        seed = s + 1;
    }
}

// Three Ottinger deviates of bead mu (mobility.cpp:18-29) and the generator
// state after them. Contraction is off so X is bit-identical to the CPU's.
__device__ inline void bead_noise(gbl_ptr<const uint32_t> jx, gbl_ptr<const uint32_t> jy,
                                  int mu, int32_t ix0, int32_t iy0, double X[3],
                                  int32_t& ix_out, int32_t& iy_out) {
#pragma clang fp contract(off)
    uint32_t ux = gf2_apply(jx + 32 * mu, (uint32_t)ix0);
    int32_t iy = minstd_mulmod(iy0, jy[mu]);
    const float am = 0x1.fffffep-32f;   // nextafterf(1,-1) / 2^31, rng.cpp:24-28
    const double c1 = 14.14855378, c2 = 1.21569221;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        xorshift_step(ux);
        schrage_step(iy);
        const uint32_t mixed = ux ^ (uint32_t)iy;
        const int32_t result = (int32_t)((mixed & 0x7fffffffu) | 1u);
        const double u = (double)(am * (float)result);
        const double y = u - 0.5;
        X[d] = (c1 * y * y + c2) * y;
    }
    ix_out = (int32_t)ux;
    iy_out = iy;
}

// ---- springs (spring.cpp:14-207, numerics.cpp, numerics.hpp) ---------------
__device__ __forceinline__ double force_sans_hookean(int type, double q, double q0) {
    switch (type) {
        case 1: return 1.0;
        case 2: return 1.0 / (1.0 - q * q);
        case 3: return (3.0 - q * q) / (1.0 - q * q) / 3.0;
        case 4: return 1.0 / (6.0 * q) * (4.0 * q + 1.0 / (1.0 - q) / (1.0 - q) - 1.0);
        case 5: return 1.0 * (1.0 - (q0 / q));
        case 6: return (1.0 - q0 / q) / (1.0 - (q - q0) * (q - q0));
        case 7: {
            const double t = (q - q0) / (1.0 - q0);
            const double a = 1.0 / ((1.0 - t) * (1.0 - t));
            const double b = 1.0 / ((1.0 + t) * (1.0 + t));
            return (2.0 / (3.0 * q)) * ((a - 1.0) / 4.0 + t - q0 * (b - 1.0) / 4.0 + q0 * t);
        }
    }
    return 1.0;
}

__device__ __forceinline__ double cubic_at(const double* c, double x) {
    return ((c[3] * x + c[2]) * x + c[1]) * x + c[0];
}

__device__ inline double bisect_cubic(const double* c, double lo, double hi) {
    double flo = cubic_at(c, lo), fhi = cubic_at(c, hi);
    if (flo == 0.0) return lo;
    if (fhi == 0.0) return hi;
    if (flo * fhi > 0.0) return (fabs(flo) < fabs(fhi)) ? lo : hi;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double fm = cubic_at(c, mid);
        if (fm == 0.0 || (hi - lo) < 1e-15) return mid;
        if (flo * fm < 0.0) { hi = mid; }
        else { lo = mid; flo = fm; }
    }
    return 0.5 * (lo + hi);
}

__device__ inline double find_roots_cubic(const double* coeff, double lower, double upper) {
    const double a = coeff[2] / coeff[3];
    const double b = coeff[1] / coeff[3];
    const double c = coeff[0] / coeff[3];
    const double Q = (a * a - 3.0 * b) / 9.0;
    const double R = (2.0 * a * a * a - 9.0 * a * b + 27.0 * c) / 54.0;
    double x;
    if (R * R < Q * Q * Q) {
        const double theta = acos(R / sqrt(Q * Q * Q));
        for (int i = -1; i <= 1; ++i) {
            x = -2.0 * sqrt(Q) * cos((theta + (double)i * PI * 2.0) / 3.0) - a / 3.0;
            if (x >= lower && x <= upper) return x;
        }
        x = -2.0 * sqrt(Q) * cos((theta + PI * 2.0) / 3.0) - a / 3.0;
    } else {
        const double Au = -copysign(1.0, R) * cbrt(fabs(R) + sqrt(R * R - Q * Q * Q));
        const double Bu = (Au == 0.0) ? 0.0 : Q / Au;
        x = (Au + Bu) - a / 3.0;
    }
    if (x >= lower && x <= upper) return x;
    return bisect_cubic(coeff, lower, upper);
}

__device__ inline double polish_poly_root(const double* c, double x0, double atol) {
    double x = x0;
    for (int iter = 0; iter < 30; ++iter) {
        double p = c[3] * x + c[2];
        double p1 = c[3];
        p1 = p + p1 * x; p = c[1] + p * x;
        p1 = p + p1 * x; p = c[0] + p * x;
        if (fabs(p1) == 0.0) {
            p1 = 6.0 * c[3];
            p1 = 6.0 * p / p1;
            x -= copysign(1.0, p1) * cbrt(fabs(p1));
        } else {
            x -= p / p1;
        }
        if (fabs(p) <= atol) break;
    }
    return x;
}

__device__ inline void cubic_coefficients(int type, double dtby4, double gama, double natscl,
                                          double* c) {
    c[0] = c[1] = c[2] = 0.0;
    c[3] = 1.0;
    if (type == 2) {
        c[0] = gama; c[1] = -(1.0 + dtby4); c[2] = -gama;
    } else if (type == 3) {
        const double d = 3.0 + dtby4;
        c[0] = 3.0 * gama / d; c[1] = -(3.0 + 3.0 * dtby4) / d; c[2] = -3.0 * gama / d;
    } else if (type == 4) {
        const double d = 2.0 * (1.5 + dtby4);
        c[0] = -3.0 * gama / d;
        c[1] = 3.0 * (1.0 + dtby4 + 2.0 * gama) / d;
        c[2] = -1.5 * (4.0 + 3.0 * dtby4 + 2.0 * gama) / d;
    } else if (type == 6) {
        c[0] = dtby4 * natscl + gama * (1.0 - natscl * natscl);
        c[1] = -1.0 + natscl * natscl - dtby4 + 2.0 * gama * natscl;
        c[2] = -(2.0 * natscl + gama);
    }
}

__device__ __forceinline__ void wlcb_fd(double dt, double s, double gama, double x, double& f,
                                        double& df) {
    const double omx = 1.0 - x, opx = 1.0 + x - 2.0 * s, oms = 1.0 - s;
    f = 6.0 * x / dt + 0.25 * oms * oms / (omx * omx) - 0.25 + (x - s) / oms -
        s / 4.0 * oms * oms / (opx * opx) + s / 4.0 + s * (x - s) / oms - 6.0 * gama / dt;
    df = 6.0 / dt + 1.0 / oms + s / oms + oms * oms / (2.0 * (omx * omx * omx)) +
         (oms * oms * s) / (2.0 * ((1.0 - 2.0 * s + x) * (1.0 - 2.0 * s + x) * (1.0 - 2.0 * s + x)));
}

// spring.cpp:101-119 + numerics.hpp:24-73 (rtsafe); status instead of abort.
__device__ inline int solve_wlc_bounded(double dtby4, double gama, double sigma, double& root) {
    if (sigma > 1.0) return ST_WLCB_SIGMA;
    const double dt = dtby4 * 4.0;
    const double eps = 2.220446049250313e-16;
    const double lo = (0.0 < 2.0 * sigma - 1.0 + eps) ? 2.0 * sigma - 1.0 + eps : 0.0;
    const double x1 = lo, x2 = 1.0 - eps, xacc = 1e-15;
    double fl, fh, df;
    wlcb_fd(dt, sigma, gama, x1, fl, df);
    wlcb_fd(dt, sigma, gama, x2, fh, df);
    if ((fl > 0.0 && fh > 0.0) || (fl < 0.0 && fh < 0.0)) return ST_RTSAFE_UNBRACKETED;
    if (fl == 0.0) { root = x1; return ST_OK; }
    if (fh == 0.0) { root = x2; return ST_OK; }
    double xl, xh;
    if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
    double rts = 0.5 * (x1 + x2);
    double dxold = fabs(x2 - x1);
    double dx = dxold;
    double f;
    wlcb_fd(dt, sigma, gama, rts, f, df);
    for (int j = 0; j < 10000; ++j) {
        if (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0 ||
            fabs(2.0 * f) > fabs(dxold * df)) {
            dxold = dx;
            dx = 0.5 * (xh - xl);
            rts = xl + dx;
            if (xl == rts) { root = rts; return ST_OK; }
        } else {
            dxold = dx;
            dx = f / df;
            const double temp = rts;
            rts -= dx;
            if (temp == rts) { root = rts; return ST_OK; }
        }
        if (fabs(dx) < xacc) { root = rts; return ST_OK; }
        wlcb_fd(dt, sigma, gama, rts, f, df);
        if (f < 0.0) xl = rts; else xh = rts;
    }
    return ST_RTSAFE_MAXIT;
}

__device__ __forceinline__ double asymptotic_guess(int type, double dtby4, double gama) {
    if (type == 2) return 1.0 - dtby4 / 2.0 / gama;
    if (type == 3) return 1.0 - dtby4 / 3.0 / gama;
    if (type == 4) return 1.0 - sqrt(dtby4 / 6.0 / gama);
    return gama / (1.0 + dtby4);
}

// spring.cpp:123-172: reduced bond length r of r(1 + dtby4 f(r)) = gama.
__device__ inline int solve_implicit_r(int type, double dtby4, double gama, double natscl,
                                       double& r_out) {
    double r;
    if (type == 1) { r_out = gama / (1.0 + dtby4); return ST_OK; }
    if (type == 5) { r_out = (gama + dtby4 * natscl) / (1.0 + dtby4); return ST_OK; }
    if (type == 7) {
        const int st = solve_wlc_bounded(dtby4, gama, natscl, r);
        r_out = r;
        return st;
    }
    double c[4];
    cubic_coefficients(type, dtby4, gama, natscl, c);
    const double margin = 1e-6;
    if (type == 6) {
        const double lo = (natscl - 1.0 < 0.0) ? 0.0 : natscl - 1.0, hi = natscl + 1.0;
        r = find_roots_cubic(c, lo, hi);
        r = polish_poly_root(c, r, 1e-14);
        const double t = (r < lo + margin) ? lo + margin : r;
        r = (hi - margin < t) ? hi - margin : t;
    } else {
        r = (gama < 1.0) ? gama / (1.0 + dtby4) : asymptotic_guess(type, dtby4, gama);
        if (!(type == 4 && gama > 100.0)) r = polish_poly_root(c, r, 1e-14);
        const double t = (r < 0.0) ? 0.0 : r;
        r = (1.0 - margin < t) ? 1.0 - margin : t;
    }
    r_out = r;
    return ST_OK;
}

// spring.cpp:174-181: new bond vector parallel to gamma.
__device__ inline int solve_connector(const double g[3], double dt, int type, double sqrtb,
                                      double q0, double out[3]) {
    const double gmag = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    double r;
    const int st = solve_implicit_r(type, dt / 4.0, gmag / sqrtb, q0, r);
    if (fabs(r - 1.0) < 1e-6) r = 1.0 - 1e-6;
    const double s = r * sqrtb / gmag;
    out[0] = s * g[0]; out[1] = s * g[1]; out[2] = s * g[2];
    return st;
}

// spring.cpp:183-193 for one bond: F = f(|b|/sqrtb) b.
__device__ __forceinline__ void connector_force(const TrajParams& P, const double b[3],
                                                double F[3]) {
    double sq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    if (sq < 1.0e-12) sq = 1.0e-12;
    double r = sqrt(sq) / P.sqrtb;
    if (fabs(r - 1.0) < MYEPS) r = 1.0 - MYEPS;
    const double f = force_sans_hookean(P.spring_type, r, P.q0);
    F[0] = f * b[0]; F[1] = f * b[1]; F[2] = f * b[2];
}

// ---- LDS-qualified pointers ----------------------------------------------------
// A pointer into the workgroup's shared arrays, typed so that reads through it
// are LDS instructions (ds_read) wherever the access sits after inlining.
template <typename T>
using lds_cptr = const __attribute__((address_space(3))) T*;
template <typename T>
__device__ __forceinline__ lds_cptr<T> as_lds(const T* p) { return (lds_cptr<T>)p; }
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T*;
template <typename T>
__device__ __forceinline__ lds_ptr<T> as_lds_mut(T* p) { return (lds_ptr<T>)p; }

// ---- excluded volume (excluded_volume.cpp:36-58) ----------------------------
// Reciprocal of x > 0: v_rcp_f32 (1 ulp) in FP32; in FP64 v_rcp_f64 refined by
// two Newton steps (e = 1 - x r, r += r e), which brings it to within an ulp
// of 1/x.
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_fast(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// ---- flow and external forces (flow.cpp:5-26, external_force.cpp:8-52) ------
__device__ inline void kappa_at(const TrajParams& P, const double* __restrict__ blob, double t,
                                double K[9]) {
    if (P.flow_const || P.flow_n == 0) {
        for (int k = 0; k < 9; ++k) K[k] = P.kappa0[k];
        return;
    }
    const double* times = blob + P.flow_off;
    const double* Ks = times + P.flow_n;
    const int n = P.flow_n;
    if (n == 1 || t <= times[0]) { for (int k = 0; k < 9; ++k) K[k] = Ks[k]; return; }
    if (t >= times[n - 1]) { for (int k = 0; k < 9; ++k) K[k] = Ks[9 * (n - 1) + k]; return; }
    int i = 0;
    while (i < n - 1 && !(t >= times[i] && t < times[i + 1])) ++i;
    const double w = (t - times[i]) / (times[i + 1] - times[i]);
    for (int k = 0; k < 9; ++k) K[k] = Ks[9 * i + k] + w * (Ks[9 * (i + 1) + k] - Ks[9 * i + k]);
}

__device__ inline void external_at(const TrajParams& P, const ExtEntry* __restrict__ ext,
                                   const double* __restrict__ blob, int N, int mu, double t,
                                   double F[3]) {
    F[0] = F[1] = F[2] = 0.0;
    for (int k = 0; k < P.n_ext; ++k) {
        const ExtEntry e = ext[P.ext_off + k];
        const int i = e.bead < 0 ? N + e.bead : e.bead;
        if (i != mu) continue;
        double v[3];
        if (e.constant) {
            v[0] = e.value[0]; v[1] = e.value[1]; v[2] = e.value[2];
        } else {
            const int n = e.n_times;
            const double* ts = blob + e.off;
            const double* vs = ts + n;
            if (n == 0) { v[0] = v[1] = v[2] = 0.0; }
            else if (n == 1 || t <= ts[0]) { v[0] = vs[0]; v[1] = vs[1]; v[2] = vs[2]; }
            else if (t >= ts[n - 1]) {
                v[0] = vs[3 * (n - 1)]; v[1] = vs[3 * (n - 1) + 1]; v[2] = vs[3 * (n - 1) + 2];
            } else {
                int j = 0;
                while (j < n - 1 && !(t >= ts[j] && t < ts[j + 1])) ++j;
                const double w = (t - ts[j]) / (ts[j + 1] - ts[j]);
                for (int d = 0; d < 3; ++d) v[d] = vs[3 * j + d] + w * (vs[3 * (j + 1) + d] - vs[3 * j + d]);
            }
        }
        F[0] = F[0] + v[0]; F[1] = F[1] + v[1]; F[2] = F[2] + v[2];
    }
}

// ---- RPY mobility (hydrodynamics.cpp:12-54) --------------------------------
// Pair weights of the block D_{mu nu} = om1 I + om2 b b^T (b = R_nu - R_mu)
// from one reciprocal square root of sq = |b|^2, with the reference's
// constants folded into five products of h*:
//   far  (|b| >= 2 sqrt(pi) h*):  om1 = p (1 + C1 h*^2 / sq)
//                                 om2 = p / sq (1 - 2 pi h*^2 / sq),  p = (3 sqrt(pi)/4) h* / |b|
//   near (overlapping spheres):   om1 = 1 - (C3 / h*) |b|,  om2 = (C7 / h*) / |b|
// The far/near test compares sq with (2 sqrt(pi) h*)^2; the weights are
// continuous where the branches meet (om1 = 7/16, om2 |b|^2 = 3/16), so the
// choice at the boundary only moves results by rounding. sq is clamped to
// 1e-12 as in the reference. For the self pair mu == nu, b = 0, so the dyadic
// term vanishes and the block evaluates to om1(1e-12) I; adding
// `cdiag` = 1 - om1(1e-12) times the operand restores the identity block.
// K1..K5 and far2 for h* > 0, in this order.
__host__ __device__ inline void rpy_folded(double hstar, double k[6]) {
    const double RPI = 1.7724538509055160273;   // sqrt(pi)
    const double TPI = 2.0 * PI, C1 = TPI / 3.0, C3 = 0.75 / RPI * 0.375, C7 = 0.09375 / RPI;
    k[0] = RPI * 0.75 * hstar;
    k[1] = C1 * hstar * hstar;
    k[2] = -TPI * hstar * hstar;
    k[3] = -C3 / hstar;
    k[4] = C7 / hstar;
    k[5] = (2.0 * RPI * hstar) * (2.0 * RPI * hstar);
}

template <typename T>
struct RpyW {
    T K1, K12, K13, K4, K5, far2, clamp, cdiag;

    // k: the six rpy_folded constants in the arithmetic type
    __device__ __forceinline__ void init(const T* k) {
        K1 = k[0]; K12 = k[0] * k[1]; K13 = k[0] * k[2]; K4 = k[3]; K5 = k[4]; far2 = k[5];
        clamp = (T)1.0e-12;
        T om1, om2;
        weights((T)0, om1, om2);
        cdiag = (T)1 - om1;
    }

    // the identity-restoring diagonal coefficient of init()
    __device__ __forceinline__ static T cdiag_of(const T* k) {
        RpyW w;
        w.init(k);
        return w.cdiag;
    }

    __device__ __forceinline__ static T rsq(T x) {
        if constexpr (sizeof(T) == 4) return __frsqrt_rn(x);
        else return rsqrt(x);
    }

    __device__ __forceinline__ void weights(T sq, T& om1, T& om2) const {
        weights_floor(fmax(sq, clamp), om1, om2);
    }

    // sq already at least the clamp (a separation formed as 1e-12 + |b|^2).
    // Far field as polynomials in 1/|b|: om1 = K1 ir + K1 K2 ir^3,
    // om2 = ir^3 (K1 + K1 K3 ir^2).
    __device__ __forceinline__ void weights_floor(T sq, T& om1, T& om2) const {
        const T ir = rsq(sq);
        const T ir2 = ir * ir;
        const T ir3 = ir * ir2;
        const T om1f = fma(K12, ir3, K1 * ir);
        const T om2f = ir3 * fma(K13, ir2, K1);
        const T om1n = fma(K4, sq * ir, (T)1);
        const T om2n = K5 * ir;
        const bool far = sq >= far2;
        om1 = far ? om1f : om1n;
        om2 = far ? om2f : om2n;
    }
};

}  // namespace bdgpu
