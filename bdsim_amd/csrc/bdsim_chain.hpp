// bdsim_chain.hpp -- the device side of the MI355X ChainIntegrator step: the
// per-chain integrator (Chain), its kernels, and the piecewise kernels (forces,
// noise, mobility products, observables). Included by every translation unit
// of libbdsim_gpu.so; each instantiates the kernels it launches.
//
// One workgroup integrates one chain over a whole integrate() segment
// (integrator.cpp:66-85) without leaving the kernel: the step loop, the flow
// clock, the spectral warm-start cache and the RNG state all stay on chip.
// Thread tid owns bead tid. The RPY mobility is never stored: every D.x is a
// matrix-free pair loop over LDS-staged positions (hydrodynamics.cpp:12-54 +
// linalg.cpp:81-89 fused). The dense tensor is built only for the Cholesky
// method and for the rare Chebyshev -> Cholesky fallback (mobility.cpp:71-80).
//
// Occupancy: the LDS footprint is ~9 KB for chains of up to 64 beads and the
// per-bead vectors that live across phases of a step sit in a per-trajectory
// HBM slot, so a 64-thread workgroup fits in 128 VGPRs and 16 chains share a
// CU (4 waves per SIMD).
//
// Citations are to IsaacPincus/bdsim.
#pragma once
#include "bdsim_device.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

namespace bdgpu {

struct KArgs {
    const TrajParams* tp;
    const int32_t* list;        // trajectory index per workgroup
    double* R;                  // packed positions
    double* ws;                 // packed workspace
    const double* blob;         // phi / flow / external tables
    const ExtEntry* ext;
    RngState* rng;
    int32_t* status;
    Counters* cnt;
    StepLog* log;
    const uint32_t* jx;         // xorshift jump columns [MAX_BEADS][32]
    const uint32_t* jy;         // minstd jump multipliers [MAX_BEADS]
    int32_t sim_override;       // nonzero: use the sim fields below for all
    int32_t com;
    double t0, t1, dt, tol;
    // resumable segments (null seg: each launch runs whole segments)
    SegState* seg;              // per trajectory
    int32_t* done_out;          // per workgroup: 1 once its segment has ended
    int32_t max_steps;          // steps per launch at most
    // chains split over `split` workgroups (split > 1): per launch slot a
    // barrier counter and two exchange buffers of [split][3][1024] doubles
    int32_t split;
    uint32_t* bar;
    double* xbuf;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

template <int BLOCK, bool MIXED>
struct Smem {
    static constexpr int NW = BLOCK / 64;
    // single-wave chains walk bead pairs (mu, mu + k) on doubled FP32 copies
    static constexpr bool PAIRS = MIXED && BLOCK == 64;
    // mixed form, 2..16 waves: the pair walk by 64-bead tiles (apply_tiles)
    static constexpr bool TILES = MIXED && BLOCK >= 128 && BLOCK <= 1024;
    static_assert(!MIXED || PAIRS || TILES, "the mixed form walks bead pairs (64..1024 threads)");
    static constexpr int PB = (PAIRS || TILES) ? 2 * BLOCK : 1;
    static constexpr int AB = TILES ? BLOCK : 1;
    static constexpr int DB = MIXED ? 1 : BLOCK;   // FP64 operand (FP64 form)
    double R[3][BLOCK];         // configuration the mobility is evaluated at
    double P[3][BLOCK];         // force configuration; corrector positions
    double X[3][DB];            // mat-vec operand
    // mixed form, one wave (the pair walk): bead mu's FP32 position relative
    // to bead 0 and the operand's x in Pa[mu], the operand's y, z in Pb[mu],
    // both doubled (entry mu + N repeats mu), so that one 16-byte and one
    // 8-byte read at an immediate offset fetch a partner. Tiled walk: bead
    // 64 t + i at 128 t + i and 128 t + 64 + i (pads beyond N far away, with
    // a zero operand).
    f32x4 Pa[PB];
    union {
        double B[4][BLOCK];     // bond-level arrays (unit bonds + lengths, bonds)
        // tiled walk: partner-row sums of other waves in fixed point (the
        // scale set by the operand's largest entry), added as integers so
        // that the order of the waves does not matter
        long long Acc[3][AB];
    };
    union {
        double V[3][BLOCK];     // neighbour exchange; triangular-product operand
        float W[3][2 * BLOCK];  // mixed form: mat-vec operand, EV positions
        f32x2 Pb[PB];           // the pair walk's operand y, z
    };
    double Red[2 * NW];
    double Alpha[LANCZOS_COLD];
    double Beta[LANCZOS_COLD];
    double Scal[16];
    int32_t ScalI[8];
    int32_t Ex[TILES ? NW : 1]; // tiled walk: per wave, the largest biased exponent of the operand
    Counters cnt;               // work counters, updated by thread 0
    StepLog lg;                 // this step's decisions, written by thread 0
};

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// hydrodynamics.cpp:106-136; c, dd, y are this lane's rows (in LDS). The
// recurrences carry c[i-1], dd[i-1] and y[i+1] in registers (the same
// operations; the rows are kept for the back substitution and the result).
// T = float (the mixed form): the same recurrence in FP32 with one hardware
// reciprocal per row; the vectors only seed the next warm start. The tiny
// pivot guard becomes 1e-30.
template <typename T>
__device__ inline void tridiag_eigenvector(int m, lds_cptr<double> a, lds_cptr<double> b, double lambda,
                                           lds_ptr<T> c, lds_ptr<T> dd, lds_ptr<T> y) {
    constexpr bool F = sizeof(T) == 4;
    const T tiny = F ? (T)1.0e-30 : (T)1.0e-300;
    auto div = [](T x, T d) -> T {
        if constexpr (F) return x * __builtin_amdgcn_rcpf(d);
        else return x / d;
    };
    const double scale = (fabs(lambda) < 1.0) ? 1.0 : fabs(lambda);
    const T shift = (T)(lambda - 1.0e-7 * scale);
    const T y0 = (T)(1.0 / sqrt((double)m));
    for (int i = 0; i < m; ++i) y[i] = y0;
    for (int it = 0; it < 2; ++it) {
        T beta0 = (T)a[0] - shift;
        if (fabs(beta0) < tiny) beta0 = tiny;
        T cp = (m > 1 ? div((T)b[0], beta0) : (T)0);
        T dp = div(y[0], beta0);
        c[0] = cp;
        dd[0] = dp;
        for (int i = 1; i < m; ++i) {
            const T bi1 = (T)b[i - 1];
            T beta = ((T)a[i] - shift) - bi1 * cp;
            if (fabs(beta) < tiny) beta = tiny;
            T ci, di;
            if constexpr (F) {
                const T r = __builtin_amdgcn_rcpf(beta);
                ci = i < m - 1 ? (T)b[i] * r : (T)0;
                di = (y[i] - bi1 * dp) * r;
            } else {
                ci = (i < m - 1 ? b[i] / beta : 0.0);
                di = (y[i] - bi1 * dp) / beta;
            }
            c[i] = ci;
            dd[i] = di;
            cp = ci;
            dp = di;
        }
        T yn = dp;
        y[m - 1] = yn;
        for (int i = m - 2; i >= 0; --i) {
            yn = dd[i] - c[i] * yn;
            y[i] = yn;
        }
        T nrm = 0;
        for (int i = 0; i < m; ++i) nrm += y[i] * y[i];
        nrm = sqrt(nrm);
        if (!(nrm > (T)0) || !isfinite(nrm)) {
            for (int i = 0; i < m; ++i) y[i] = y0;
            return;
        }
        for (int i = 0; i < m; ++i) y[i] = div(y[i], nrm);
    }
}

// Extreme eigenvalues of the k x k Lanczos tridiagonal (hydrodynamics.cpp:240-262):
// the reference bisects the Gershgorin bracket 100 times on the Sturm count
// (>= 1 for the smallest, >= k for the largest). Here the 64 lanes of the first
// wave multisect instead -- lanes 0-31 the smallest, 32-63 the largest -- each
// round testing 32 interior points and keeping the sub-interval where the count
// first reaches the target, until the bracket stops shrinking (floating-point
// resolution); the value is its midpoint, as in the reference. Results:
// sm_out[0] = smallest, sm_out[1] = largest.
template <typename T>
__device__ __forceinline__ T sturm_div(T num, T d) { return num / d; }
template <>
__device__ __forceinline__ float sturm_div<float>(float num, float d) { return num * __builtin_amdgcn_rcpf(d); }

// T = double: the FP64 form's values. T = float (mixed form): the same
// multisection in FP32 with a reciprocal for the division -- the bounds are
// padded by 0.5 and 1.02 before use, so FP32 resolution is ample -- in about
// 5 rounds instead of 11.
template <typename T>
__device__ __noinline__ void ritz_multisect(int k, lds_cptr<double> al, lds_cptr<double> be,
                                            double* sm_out) {
    const int lane = threadIdx.x & 63, half = lane >> 5, j = lane & 31;
    double lo = al[0], hi = al[0];
    for (int i = 0; i < k; ++i) {
        const double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < k - 1 ? fabs(be[i]) : 0.0);
        if (al[i] - rad < lo) lo = al[i] - rad;
        if (hi < al[i] + rad) hi = al[i] + rad;
    }
    const int target = half == 0 ? 1 : k;
    T a = (T)lo, b = (T)hi;
    for (int round = 0; round < 64; ++round) {
        const T w = (b - a) / (T)33.0;
        const T x = a + (T)(j + 1) * w;
        int count = 0;
        {
            T d = (T)al[0] - x;
            if (d < (T)0) ++count;
            for (int i = 1; i < k; ++i) {
                if (d == (T)0) d = sizeof(T) == 8 ? (T)1.0e-300 : (T)1.0e-30f;
                const T bi = (T)be[i - 1];
                d = ((T)al[i] - x) - sturm_div<T>(bi * bi, d);
                if (d < (T)0) ++count;
            }
        }
        const bool ok = count >= target;
        const uint64_t m = __ballot(ok);
        const uint32_t hm = (uint32_t)(half ? (m >> 32) : m);
        const int f = hm ? __ffs(hm) - 1 : 32;
        const T nb = f < 32 ? a + (T)(f + 1) * w : b;
        const T na = f > 0 ? a + (T)f * w : a;
        const bool stuck = (nb == b && na == a) || !(na < nb);
        if (!stuck) { a = na; b = nb; }
        if (__all(stuck)) break;
    }
    if (j == 0) sm_out[half] = 0.5 * ((double)a + (double)b);
}

// spring.cpp:174-181 out of line: the cubic / rtsafe solvers are large and run
// once per bond per corrector pass.
__device__ __noinline__ int solve_connector_ool(double g0, double g1, double g2, double dt,
                                                int type, double sqrtb, double q0, double* out) {
    const double g[3] = {g0, g1, g2};
    double o[3];
    const int st = solve_connector(g, dt, type, sqrtb, q0, o);
    out[0] = o[0]; out[1] = o[1]; out[2] = o[2];
    return st;
}

// Rare-path helpers kept out of line. They take what they use explicitly
// (no Chain object), so the integrator's state stays in registers around the
// calls.
struct D3 { double v[3]; };

// Dense RPY tensor into the workspace (hydrodynamics.cpp:12-54, same entry
// formula and operand order), then its lower Cholesky factor in place
// (linalg.cpp:91-107, same per-entry sums, computed column by column).
// Returns false on a non-positive pivot, with its index/value counted.
template <int BLOCK, bool MIXED>
__device__ __noinline__ bool dense_factor_ool(Smem<BLOCK, MIXED>& sm, gbl_ptr<double> Dm, int N,
                                              double hstar) {
    const int n = 3 * N, tid = threadIdx.x;
    const bool own = tid < N;
    const double RPI = 1.7724538509055160273;
    const double TRPI = 2.0 * RPI, TPI = 2.0 * PI, C1 = TPI / 3.0, C3 = 0.75 / RPI * 0.375,
                 C7 = 0.09375 / RPI, RPI3by4 = RPI * 0.75;
    __syncthreads();
    if (own) {
        const int mu = tid;
        for (int nu = 0; nu < N; ++nu) {
            if (nu == mu) {
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) Dm[(size_t)(3 * mu + p) * n + 3 * nu + q] = p == q ? 1.0 : 0.0;
                continue;
            }
            const int lo = mu < nu ? mu : nu, hi_ = mu < nu ? nu : mu;
            double b[3];
            for (int d = 0; d < 3; ++d) b[d] = sm.R[d][hi_] - sm.R[d][lo];
            double sq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            if (sq < 1.0e-12) sq = 1.0e-12;
            const double rs = sqrt(sq);
            const double hsbyrs = hstar / rs, rsbyhs = rs / hstar;
            const double hsbyrs2 = hsbyrs * hsbyrs;
            double om1, om2;
            if (rsbyhs >= TRPI) {
                om1 = RPI3by4 * hsbyrs * (1.0 + C1 * hsbyrs2);
                om2 = RPI3by4 * (hsbyrs / (rs * rs)) * (1.0 - TPI * hsbyrs2);
            } else {
                om1 = 1.0 - C3 * rsbyhs;
                om2 = C7 * rsbyhs / (rs * rs);
            }
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    // entry (3mu+p, 3nu+q); reference index i is the lower bead's
                    const int i = (mu == lo) ? p : q, j = (mu == lo) ? q : p;
                    Dm[(size_t)(3 * mu + p) * n + 3 * nu + q] = om2 * b[i] * b[j] + (i == j ? om1 : 0.0);
                }
        }
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            double sum = Dm[(size_t)j * n + j];
            for (int k = 0; k < j; ++k) sum -= Dm[(size_t)j * n + k] * Dm[(size_t)j * n + k];
            const bool ok = sum > 0.0;
            sm.ScalI[4] = ok ? 1 : 0;
            sm.Scal[15] = sum;
            if (ok) Dm[(size_t)j * n + j] = sqrt(sum);
        }
        __syncthreads();
        if (!sm.ScalI[4]) {
            if (tid == 0) {
                sm.cnt.fail_index = j;
                sm.cnt.fail_value = sm.Scal[15];
            }
            return false;
        }
        const double Ljj = Dm[(size_t)j * n + j];
        for (int i = j + 1 + tid; i < n; i += BLOCK) {
            double sum = Dm[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) sum -= Dm[(size_t)i * n + k] * Dm[(size_t)j * n + k];
            Dm[(size_t)i * n + j] = sum / Ljj;
        }
        __syncthreads();
    }
    return true;
}

// y = L x with the factor in the workspace (linalg.cpp:109-117), own rows.
template <int BLOCK, bool MIXED>
__device__ __noinline__ D3 lower_matvec_ool(Smem<BLOCK, MIXED>& sm, gbl_ptr<double> Dm, int N,
                                            double x0, double x1, double x2) {
    const int tid = threadIdx.x;
    __syncthreads();
    sm.V[0][tid] = x0;
    sm.V[1][tid] = x1;
    sm.V[2][tid] = x2;
    __syncthreads();
    D3 y{{0.0, 0.0, 0.0}};
    if (tid >= N) return y;
    const int n = 3 * N;
    for (int p = 0; p < 3; ++p) {
        const int i = 3 * tid + p;
        const gbl_ptr<double> row = Dm + (size_t)i * n;
        double s = 0.0;
        for (int j = 0; j <= i; ++j) s += row[j] * sm.V[j % 3][j / 3];
        y.v[p] = s;
    }
    return y;
}

// ---- the dense factor held in LDS (Cholesky chains with 3N <= 16 NB) --------
// The RPY matrix, padded with unit rows to 16 nb (nb = ceil(3N/16) <= NB), is
// kept as its lower nb(nb+1)/2 blocks of 16 x 16: block (I,J) at
// 256 (I(I+1)/2 + J), element (r,c) of a block at 16 r + (c ^ r). The XOR puts
// the 16 lanes of a row-wise or a column-wise access on 16 distinct banks (bank
// pairs for doubles). The factor is FP64 (both forms); the mixed form's
// FP32-factor variant keeps it in FP32 (half the LDS: more chains per CU).
template <int NB, typename T>
struct DenseLds {
    T blk[NB * (NB + 1) / 2 * 256];
    double xs[16 * NB];         // operand and result of L x
};

__device__ __forceinline__ int dl_block(int I, int J) { return (I * (I + 1) / 2 + J) * 256; }
__device__ __forceinline__ int dl_elem(int r, int c) { return 16 * r + (c ^ r); }
__device__ __forceinline__ int dl_at(int i, int j) {
    return dl_block(i >> 4, j >> 4) + dl_elem(i & 15, j & 15);
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float readlane_any(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double readlane_any(double v, int lane) { return readlane_f64(v, lane); }

typedef double mfma_d4 __attribute__((ext_vector_type(4)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// One 16x16x4 product on the matrix cores, acc += A B: A lane l = (row l%16,
// k l/16), B lane l = (k l/16, column l%16). The accumulator layouts differ:
// FP64 lane l value q = (row l/16 + 4q, column l%16), FP32 (row 4(l/16) + q,
// column l%16); acc_row gives the row of value q.
__device__ __forceinline__ mfma_d4 mfma16(double a, double b, mfma_d4 acc) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ mfma_f4 mfma16(float a, float b, mfma_f4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
}
template <typename T>
__device__ __forceinline__ int acc_row(int g, int q) { return sizeof(T) == 8 ? g + 4 * q : 4 * g + q; }

// Dense RPY tensor (the entries of dense_factor_ool) and its lower Cholesky
// factor, in LDS, by 16 x 16 blocks, right-looking: per block column K the
// diagonal block is factored in registers (lane r holds row r; the pivots and
// the column broadcast by readlane; each entry subtracts its products in
// increasing k and divides by the pivot, linalg.cpp:91-107), the blocks below
// it are multiplied by the inverse of that factor, and every trailing block
// (I,J) takes -= L_IK L_JK^T, both on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64, four per block). The result equals the
// column-by-column factor up to rounding. Pivots are taken in column order,
// so a non-positive one is reported at the same index.
template <int BLOCK, bool MIXED, int NB, typename T>
__device__ __noinline__ bool dense_factor_lds(Smem<BLOCK, MIXED>& sm, lds_ptr<T> Lb, int N,
                                              double hstar) {
    static_assert(BLOCK == 64, "one wave per chain");
    constexpr bool F32 = sizeof(T) == 4;
    const int n = 3 * N, tid = threadIdx.x;
    const int nb = (n + 15) >> 4;
    const double RPI = 1.7724538509055160273;
    const double TRPI = 2.0 * RPI, TPI = 2.0 * PI, C1 = TPI / 3.0, C3 = 0.75 / RPI * 0.375,
                 C7 = 0.09375 / RPI, RPI3by4 = RPI * 0.75;
    const double ihstar = 1.0 / hstar;
    __syncthreads();
    // padding rows n .. 16 nb - 1: zero, unit diagonal
    {
        const int np = 16 * nb, rows = np - n;
        for (int e = tid; e < rows * np; e += BLOCK) {
            const int i = n + e / np, j = e % np;
            if (j <= i) Lb[dl_at(i, j)] = i == j ? (T)1 : (T)0;
        }
    }
    // lower triangle of D: bead pairs mu > nu spread over the wave (pair
    // index p = mu(mu-1)/2 + nu, lane l takes p = l, l + 64, ...)
    {
        int mu = 1, nu = tid;
        while (nu >= mu) { nu -= mu; ++mu; }
        // an FP32 factor's entries are built in FP32 (the separation formed
        // in FP64, then rounded)
        const T th = (T)hstar, tih = (T)ihstar;
        for (; mu < N;) {
            T b[3];
            for (int d = 0; d < 3; ++d) b[d] = (T)(sm.R[d][mu] - sm.R[d][nu]);
            T sq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            if (sq < (T)1.0e-12) sq = (T)1.0e-12;
            // one reciprocal square root for r, h*/r and 1/r^2 (within an ulp
            // of the square root and divisions of hydrodynamics.cpp:30-43)
            T ir;
            if constexpr (F32) ir = rsqrtf(sq);
            else ir = rsqrt(sq);
            const T ir2 = ir * ir;
            const T rs = sq * ir;
            const T hsbyrs = th * ir, rsbyhs = rs * tih;
            const T hsbyrs2 = hsbyrs * hsbyrs;
            T om1, om2;
            if (rsbyhs >= (T)TRPI) {
                om1 = (T)RPI3by4 * hsbyrs * ((T)1 + (T)C1 * hsbyrs2);
                om2 = (T)RPI3by4 * (hsbyrs * ir2) * ((T)1 - (T)TPI * hsbyrs2);
            } else {
                om1 = (T)1 - (T)C3 * rsbyhs;
                om2 = (T)C7 * rsbyhs * ir2;
            }
            // entry (3mu+p, 3nu+q); the reference's index i is the lower bead's (nu: q)
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q)
                    Lb[dl_at(3 * mu + p, 3 * nu + q)] = om2 * b[q] * b[p] + (q == p ? om1 : (T)0);
            nu += BLOCK;
            while (nu >= mu) { nu -= mu; ++mu; }
        }
        if (tid < N)
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q <= p; ++q) Lb[dl_at(3 * tid + p, 3 * tid + q)] = p == q ? (T)1 : (T)0;
    }
    __syncthreads();
    const int r = tid & 15, g = tid >> 4;
    for (int K = 0; K < nb; ++K) {
        // diagonal block: every lane loads row r (lanes 16.. duplicate 0..15)
        const int dk = dl_block(K, K);
        T row[16], rsv[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) row[c] = Lb[dk + dl_elem(r, c)];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            if (16 * K + c >= n) break;      // padding: already unit rows
            const T piv = readlane_any(row[c], c);
            if (!(piv > 0.0)) {
                if (tid == 0) {
                    sm.cnt.fail_index = 16 * K + c;
                    sm.cnt.fail_value = (double)piv;
                }
                __syncthreads();
                return false;
            }
            // one reciprocal square root for the pivot and the column
            // (L_jj = piv rs, L_ij = A_ij rs: within an ulp of the square
            // root and the division, off the dependent chain of columns)
            T rs;
            if constexpr (F32) rs = rsqrtf(piv);
            else rs = rsqrt(piv);
            rsv[c] = rs;
            row[c] = r == c ? piv * rs : row[c] * rs;
#pragma unroll
            for (int q = c + 1; q < 16; ++q) row[q] -= row[c] * readlane_any(row[c], q);
        }
        if (tid < 16) {
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c <= r) Lb[dk + dl_elem(r, c)] = row[c];
        }
        // blocks below: L_IK = A_IK L_KK^-T. Lane r forms row r of the
        // inverse of the diagonal factor (x L = e_r, solved from the last
        // column down), then every block below is one 16x16x16 product on the
        // matrix cores; the B operand of lane group g is column 4kk+g of that
        // row, picked from registers.
        {
            T dinv[16], xr[16];
            // 1 / L_jj: the pivot's reciprocal square root (unit padding rows: 1)
#pragma unroll
            for (int j = 0; j < 16; ++j) dinv[j] = 16 * K + j < n ? rsv[j] : (T)1;
#pragma unroll
            for (int j = 15; j >= 0; --j) {
                T sum = j == r ? (T)1 : (T)0;
#pragma unroll
                for (int k = j + 1; k < 16; ++k) sum -= xr[k] * readlane_any(row[j], k);
                xr[j] = sum * dinv[j];
            }
            T bop[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                bop[kk] = g == 0 ? xr[4 * kk] : g == 1 ? xr[4 * kk + 1] : g == 2 ? xr[4 * kk + 2] : xr[4 * kk + 3];
            for (int I = K + 1; I < nb; ++I) {
                const int bi = dl_block(I, K);
                std::conditional_t<F32, mfma_f4, mfma_d4> acc = {0, 0, 0, 0};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) acc = mfma16(Lb[bi + dl_elem(r, 4 * kk + g)], bop[kk], acc);
#pragma unroll
                for (int q = 0; q < 4; ++q) Lb[bi + dl_elem(acc_row<T>(g, q), r)] = acc[q];
            }
        }
        __syncthreads();
        // trailing blocks K < J <= I < nb: A_IJ -= L_IK L_JK^T. Operand
        // layouts of the 16x16x4 FP64 product: A lane l = (row l%16, k l/16),
        // B lane l = (k l/16, column l%16), C lane l value q = (row l/16 + 4q,
        // column l%16).
        for (int I = K + 1; I < nb; ++I) {
            const int bIK = dl_block(I, K);
            for (int J = K + 1; J <= I; ++J) {
                const int bIJ = dl_block(I, J), bJK = dl_block(J, K);
                std::conditional_t<F32, mfma_f4, mfma_d4> acc;
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] = Lb[bIJ + dl_elem(acc_row<T>(g, q), r)];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const T a = -Lb[bIK + dl_elem(r, 4 * kk + g)];
                    const T bv = Lb[bJK + dl_elem(r, 4 * kk + g)];
                    acc = mfma16(a, bv, acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) Lb[bIJ + dl_elem(acc_row<T>(g, q), r)] = acc[q];
            }
        }
        __syncthreads();
    }
    return true;
}

// y = L x with the factor in LDS (linalg.cpp:109-117: s += L[i][j] x[j] for
// j = 0..i), one row per lane, then regrouped by bead.
template <int BLOCK, bool MIXED, int NB, typename T>
__device__ __noinline__ D3 lower_matvec_lds(lds_ptr<T> Lb, lds_ptr<double> xs, int N, double x0,
                                            double x1, double x2) {
    static_assert(BLOCK == 64, "one wave per chain");
    const int tid = threadIdx.x, n = 3 * N;
    __syncthreads();
    if (tid < N) {
        xs[3 * tid] = x0;
        xs[3 * tid + 1] = x1;
        xs[3 * tid + 2] = x2;
    }
    __syncthreads();
    double yv[(16 * NB + 63) / 64];
#pragma unroll
    for (int pass = 0; pass < (16 * NB + 63) / 64; ++pass) {
        const int i = tid + 64 * pass;
        double s = 0.0;
        if (i < n) {
            const int I = i >> 4, rr = i & 15;
            for (int J = 0; J <= I; ++J) {
                const int base = dl_block(I, J) + 16 * rr;
                const int cend = J == I ? rr : 15;
                for (int c = 0; c <= cend; ++c) s += (double)Lb[base + (c ^ rr)] * xs[16 * J + c];
            }
        }
        yv[pass] = s;
    }
    __syncthreads();
#pragma unroll
    for (int pass = 0; pass < (16 * NB + 63) / 64; ++pass) {
        const int i = tid + 64 * pass;
        if (i < n) xs[i] = yv[pass];
    }
    __syncthreads();
    D3 y{{0.0, 0.0, 0.0}};
    if (tid < N)
        for (int p = 0; p < 3; ++p) y.v[p] = xs[3 * tid + p];
    return y;
}

// ---- the dense factor held in registers (Cholesky chains, FP64 factor) ------
// The padded tensor (16 NB rows; unit rows and columns beyond 3N) is held as
// its upper blocks J <= I of 16 x 16 in the accumulator layout of the FP64
// 16x16x4 product: lane l = 16 g + c holds entries (16 J + g + 4 q, 16 I + c),
// q = 0..3. Read in that layout, value kk of a block is chunk kk of the
// product's B operand, and of its A operand for the block's transpose, so the
// update A_KI -= U_JK^T U_JI is four products with no data movement. Block
// rows are formed left-looking, and each is factored unblocked in the upper
// form (U = L^T; row j of U: the pivot, the scaled row across all NB - K
// blocks at once, the rank-1 update of the rows below; linalg.cpp:91-107 by
// rows of U), one entry per lane and value, the row passed through a small
// LDS buffer. L x = U^T x accumulates block row by block row, so a block dies
// after its last use: at most 12 of the 21 blocks are live at NB = 6 (96
// VGPRs), which with the rest of the step fits the register budget of two
// waves per SIMD, and the chain needs 13 KB of LDS instead of 53 KB.
// TAB: with the pair-weight table of RegBuild<true>. slots: the step's state
// slots (F^S, F^E, dW; 9 N doubles, N <= 32) as the chain's workspace in LDS,
// so that the step's slot round trips stay on chip.
template <int NB, bool TAB>
struct RegLds {
    double row[NB][16];         // row j of each block of block row K, unscaled
    double xs[16 * NB];         // the operand, zero beyond 3N
    double yp[4][16 * NB];      // L x: partial sums of the four lane row groups
    double pw[TAB ? 2 * 6 * 32 : 1];   // RPY weights of the current block row's bead pairs
    double slots[9 * 32];
};

// The tensor's entries (as dense_factor_ool forms them, one reciprocal square
// root per entry) from the FP64 positions in LDS, and the lane's place.
// TAB: the weights come from a per-block-row table of bead pairs (one
// reciprocal square root per pair; the mixed form); otherwise each entry forms
// its pair's weights itself (the FP64 form, whose other live state leaves no
// registers for the table's schedule). Both give the same entries bit for bit.
template <bool TAB>
struct RegBuild {
    lds_cptr<double> Rb;        // sm.R: component d of bead mu at 64 d + mu
    lds_ptr<double> pw;         // RegLds::pw
    int n, N, c, g;
    double hstar, ihstar;

    // the RPY weights of one pair from its squared separation (clamped)
    __device__ __forceinline__ void weights(double sq, double& om1, double& om2) const {
        const double RPI = 1.7724538509055160273;
        const double TRPI = 2.0 * RPI, TPI = 2.0 * PI, C1 = TPI / 3.0, C3 = 0.75 / RPI * 0.375,
                     C7 = 0.09375 / RPI, RPI3by4 = RPI * 0.75;
        if (sq < 1.0e-12) sq = 1.0e-12;
        const double ir = rsqrt(sq);
        const double ir2 = ir * ir;
        const double rs = sq * ir;
        const double hsbyrs = hstar * ir, rsbyhs = rs * ihstar;
        const double hsbyrs2 = hsbyrs * hsbyrs;
        if (rsbyhs >= TRPI) {
            om1 = RPI3by4 * hsbyrs * (1.0 + C1 * hsbyrs2);
            om2 = RPI3by4 * (hsbyrs * ir2) * (1.0 - TPI * hsbyrs2);
        } else {
            om1 = 1.0 - C3 * rsbyhs;
            om2 = C7 * rsbyhs * ir2;
        }
    }

    // The RPY weights (om1, om2) of the bead pairs block row K needs: row
    // beads mu in [mlo, mlo + 6) (rows 16 K .. 16 K + 15) against beads nu in
    // [mlo, N), one reciprocal square root per pair (hydrodynamics.cpp:27-43
    // as dense_factor_ool forms them), pair (a, b) at 2 (32 a + b).
    __device__ __forceinline__ void table(int K) const {
        if constexpr (!TAB) return;
        const int mlo = 16 * K / 3;
        const int mhi = min(N - 1, (16 * K + 15) / 3);
        const int W = N - mlo;
        const int np = (mhi - mlo + 1) * W;
        wave_sync();                // the previous block row's entries are built
#pragma unroll 1
        for (int p = threadIdx.x; p < np; p += 64) {
            const int a = p / W, bb = p - a * W;
            const int mu = mlo + a, nu = mlo + bb;
            const double b0 = Rb[mu] - Rb[nu], b1 = Rb[64 + mu] - Rb[64 + nu],
                         b2 = Rb[128 + mu] - Rb[128 + nu];
            double om1, om2;
            weights(b0 * b0 + b1 * b1 + b2 * b2, om1, om2);
            pw[2 * (32 * a + bb)] = om1;
            pw[2 * (32 * a + bb) + 1] = om2;
        }
        wave_sync();
    }

    // block (K, I) of block row K (after table(K)) in the accumulator layout:
    // entry om2 b_q b_p + [p == q] om1 of the bead pair (mu, nu) = (i / 3,
    // j / 3), b = R_mu - R_nu; the identity on a bead's own 3 x 3 block and on
    // the padding
    __device__ __forceinline__ mfma_d4 operator()(int K, int I) const {
        // the lane's place as an opaque value: otherwise the index arithmetic
        // of all 21 blocks is hoisted out of the step loop and held (spilled)
        // across the whole step
        int cl = c, gl = g;
        asm volatile("" : "+v"(cl), "+v"(gl));
        const int mlo = 16 * K / 3;
        const int j = 16 * I + cl;
        const int jb = j / 3;
        const int qc = j - 3 * jb;
        const int nu = j < n ? jb : mlo;
        mfma_d4 v;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = 16 * K + gl + 4 * q;
            const int ib = i / 3;
            const int p = i - 3 * ib;
            const int mu = i < n ? ib : mlo;
            double om1, om2, bp, bq;
            if constexpr (TAB) {
                const int w = 2 * (32 * (mu - mlo) + (nu - mlo));
                om1 = pw[w];
                om2 = pw[w + 1];
                bp = Rb[64 * p + mu] - Rb[64 * p + nu];
                bq = Rb[64 * qc + mu] - Rb[64 * qc + nu];
            } else {
                const double b0 = Rb[mu] - Rb[nu], b1 = Rb[64 + mu] - Rb[64 + nu],
                             b2 = Rb[128 + mu] - Rb[128 + nu];
                weights(b0 * b0 + b1 * b1 + b2 * b2, om1, om2);
                bp = p == 0 ? b0 : (p == 1 ? b1 : b2);
                bq = qc == 0 ? b0 : (qc == 1 ? b1 : b2);
            }
            double e = om2 * bq * bp + (p == qc ? om1 : 0.0);
            if (i >= n || j >= n || ib == jb) e = i == j ? 1.0 : 0.0;
            // formed here: sunk to its first use, the entry would keep its
            // separation and weights live across the whole block row
            asm volatile("" : "+v"(e));
            v[q] = e;
        }
        return v;
    }
};

// Block row K of the register factor, left-looking: the row's blocks built
// and updated by the rows above (A_KI - sum_J U_JK^T U_JI, four products per
// term, in increasing J), its rows of U, and its part of L x. Recursion over
// K keeps every block index a compile-time constant, so the blocks stay in
// registers; U_JK dies once row K is formed, so (K + 1)(NB - K) blocks are
// live at most (12 at NB = 6). false on a non-positive pivot.
template <int K, int NB, bool TAB>
__device__ __forceinline__ bool regf_block_row(mfma_d4 (&M)[NB][NB], lds_ptr<double> yp, const RegBuild<TAB>& build,
                                               lds_ptr<double> rowb, lds_cptr<double> xs, Counters& cnt) {
    if constexpr (K == NB) {
        return true;
    } else {
        const int n = build.n, c = build.c, g = build.g;
        build.table(K);
#pragma unroll
        for (int I = K; I < NB; ++I) M[K][I] = build(K, I);
#pragma unroll
        for (int J = 0; J < K; ++J) {
            mfma_d4 na;
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) na[qq] = -M[J][K][qq];
#pragma unroll
            for (int I = K; I < NB; ++I) {
                mfma_d4 acc = M[K][I];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) acc = mfma16(na[kk], M[J][I][kk], acc);
                M[K][I] = acc;
            }
        }
        // rows jr = 4 q + gj of block row K, in order (rows of padding stay
        // unit rows); row jr is held by the lanes of row group gj, value q.
        // With q and gj unrolled, only the rows of value q depend on the
        // lane's row group: values qq < q are finished rows, qq > q rows below.
        const int rows = min(16, n - 16 * K);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int gj = 0; gj < 4; ++gj) {
                const int jr = 4 * q + gj;
                if (jr >= rows) break;
                if (g == gj) {
#pragma unroll
                    for (int I = K; I < NB; ++I) rowb[16 * I + c] = M[K][I][q];
                }
                wave_sync();
                double piv = rowb[16 * K + jr];
                piv = __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(piv)),
                                       __builtin_amdgcn_readfirstlane(__double2loint(piv)));
                if (!(piv > 0.0)) {
                    if (threadIdx.x == 0) {
                        cnt.fail_index = 16 * K + jr;
                        cnt.fail_value = piv;
                    }
                    wave_sync();
                    return false;
                }
                // one reciprocal square root for the pivot and the row
                const double rs = rsqrt(piv);
                double ur[NB];
#pragma unroll
                for (int I = K; I < NB; ++I) ur[I] = rowb[16 * I + c] * rs;
                // the holders scale row jr (its pivot becomes piv rs; its
                // entries left of the pivot are cleared after the block row)
                const double f = g == gj ? rs : 1.0;
#pragma unroll
                for (int I = K; I < NB; ++I) M[K][I][q] *= f;
                // rank-1 update of the rows below jr, right of the pivot
                const double urk = c > jr ? ur[K] : 0.0;
                {
                    const double v = rowb[16 * K + g + 4 * q] * rs;
                    const double u = g > gj ? v : 0.0;
                    M[K][K][q] -= u * urk;
#pragma unroll
                    for (int I = K + 1; I < NB; ++I) M[K][I][q] -= u * ur[I];
                }
#pragma unroll
                for (int qq = q + 1; qq < 4; ++qq) {
                    const double u = rowb[16 * K + g + 4 * qq] * rs;
                    M[K][K][qq] -= u * urk;
#pragma unroll
                    for (int I = K + 1; I < NB; ++I) M[K][I][qq] -= u * ur[I];
                }
            }
        }
        // U_KK is upper triangular
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) M[K][K][qq] = c < g + 4 * qq ? 0.0 : M[K][K][qq];
        // L x: block row K's part, U_KI^T x_K, added to this lane's partial
        // sums in LDS (lane 16 g + c owns entries (g, 16 I + c))
        {
            double xk[4];
#pragma unroll
            for (int qq = 0; qq < 4; ++qq) xk[qq] = xs[16 * K + g + 4 * qq];
#pragma unroll
            for (int I = K; I < NB; ++I) {
                double s = K == 0 ? 0.0 : yp[g * 16 * NB + 16 * I + c];
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) s += M[K][I][qq] * xk[qq];
                yp[g * 16 * NB + 16 * I + c] = s;
            }
        }
        // block rows of padding stay unit rows
        if (16 * (K + 1) >= n) return true;
        return regf_block_row<K + 1, NB, TAB>(M, yp, build, rowb, xs, cnt);
    }
}

// y = L x (own beads), L the lower Cholesky factor of the RPY tensor at sm.R.
// false on a non-positive pivot, taken in column order, its index and value
// in the counters.
template <int BLOCK, bool MIXED, int NB>
__device__ __forceinline__ bool dense_bdw_regs(Smem<BLOCK, MIXED>& sm, RegLds<NB, MIXED>* rlp, int N,
                                               double hstar, const double x[3], double y[3]) {
    static_assert(BLOCK == 64, "one wave per chain");
    const int n = 3 * N, tid = threadIdx.x;
    const int c = tid & 15, g = tid >> 4;
    lds_ptr<double> rowb = as_lds_mut(&rlp->row[0][0]);
    lds_ptr<double> xs = as_lds_mut(rlp->xs);
    lds_ptr<double> yp = as_lds_mut(&rlp->yp[0][0]);
    static_assert(BLOCK == 64, "sm.R rows of 64 beads");
    const RegBuild<MIXED> build{as_lds(sm.R[0]), as_lds_mut(rlp->pw), n, N, c, g, hstar, 1.0 / hstar};

    wave_sync();
    if (tid < N) {
        xs[3 * tid] = x[0];
        xs[3 * tid + 1] = x[1];
        xs[3 * tid + 2] = x[2];
    }
    for (int e = n + tid; e < 16 * NB; e += 64) xs[e] = 0.0;
    wave_sync();

    mfma_d4 M[NB][NB];
    if (!regf_block_row<0, NB, MIXED>(M, yp, build, rowb, xs, sm.cnt)) return false;
    wave_sync();
    y[0] = y[1] = y[2] = 0.0;
    if (tid < N)
        for (int d = 0; d < 3; ++d) {
            const int e = 3 * tid + d;
            y[d] = (yp[e] + yp[16 * NB + e]) + (yp[32 * NB + e] + yp[48 * NB + e]);
        }
    return true;
}

// Barrier of the parts of one split chain (thread 0 of each part): count in,
// then wait until the count reaches `target`. Bounded: false after about a
// second, so that a part that never arrives cannot hold the others forever.
// Everything the parts exchange goes through agent-scope atomic stores and
// loads, which are coherent across the XCDs by themselves; the caller has
// waited for its workgroup's stores to complete (s_waitcnt, then the
// workgroup barrier), so the arrival needs no release fence, and the reads
// that follow need no acquire (a fence at agent scope writes back and
// invalidates the whole L2 of the XCD, every workgroup's cached data).
__device__ __noinline__ bool chain_barrier(uint32_t* bar, uint32_t target) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int it = 0; it < (1 << 20); ++it) {
        if (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        __builtin_amdgcn_s_sleep(1);
    }
    return false;
}

// NB > 0: the kernel holds a DenseLds<NB> and the Cholesky factor of chains
// with 3N <= 16 NB lives there instead of in the workspace.
// SPLIT: one of A.split workgroups of the chain (part). Every part runs the
// whole step on its own copy of the state (the same values: the same inputs
// and the same operations), except that each D.x sums only its share of the
// partners; the parts' partial rows meet in HBM behind a per-chain barrier
// and every part adds them in part order. Part 0 writes the results.
// FT: the type of the LDS factor (NB > 0): double, or float in the mixed
// form's FP32-factor variant.
// LWS: the workspace up to its dense region (state slots, Chebyshev
// coefficients and nodes, inverse-iteration rows, Lanczos basis) lives in
// LDS (for launches of few chains, whose steps are bound by the latency of
// these round trips); the dense region stays in HBM.
// REGF (NB > 0, FT double): the Cholesky method's factor in registers
// (dense_bdw_regs) instead of LDS; the step draws the noise first and factors
// after, restoring the generator state when the factor fails.
// PTAB (mixed form, 64-thread chains): the RPY weights of the step's bead
// pairs in an LDS table (pair_tab), formed once per step in stage_R and read
// by every mat-vec of the step instead of recomputed.
template <int BLOCK, bool MIXED, int NB = 0, bool SPLIT = false, typename FT = double,
          bool LWS = false, bool REGF = false, bool PTAB = false>
struct Chain {
    static constexpr int NW = BLOCK / 64;
#ifndef BDSIM_MIXED_RITZ_F32
#define BDSIM_MIXED_RITZ_F32 1
#endif
    static constexpr bool RITZ32 = MIXED && BDSIM_MIXED_RITZ_F32;
    using S = Smem<BLOCK, MIXED>;
    Smem<BLOCK, MIXED>& sm;
    const TrajParams& P;        // global, read with scalar loads
    const KArgs& A;
    const int N, tid, n3;
    const bool own;
    const bool hi;
    const WsLayout wl;
    using ws_ptr = std::conditional_t<LWS, lds_ptr<double>, gbl_ptr<double>>;
    ws_ptr ws;                  // this trajectory's workspace (LWS: its part in LDS)
    gbl_ptr<double> wsg;        // the same in HBM (the dense region is used from here)
    lds_ptr<FT> dlb;            // DenseLds blocks (NB > 0)
    lds_ptr<double> dxs;        // DenseLds operand
    int par = 0;
    int part = 0, nparts = 1;   // SPLIT: this workgroup's share of the chain
    uint32_t epoch = 0;         // SPLIT: exchanges done
    bool sync_fail = false;     // SPLIT: a barrier timed out
    bool dense_ws = false;      // this step's factor went to the workspace
    using RL = RegLds<(NB > 0 ? NB : 1), MIXED>;
    RL* rl = nullptr;           // REGF: the register factor's LDS buffers
    static_assert(!PTAB || S::PAIRS, "the pair-weight table serves the single-wave pair walk");
    lds_ptr<f32x2> ptab;        // PTAB: (om1, om2) of pair (mu, mu + k) at mu * pt_stride() + k - 1

    __device__ __forceinline__ Chain(Smem<BLOCK, MIXED>& s, const TrajParams& p, const KArgs& a,
                     FT* dense_blk = nullptr, double* dense_xs = nullptr, double* lds_ws = nullptr,
                     void* reg_lds = nullptr, void* pair_tab = nullptr)
        : sm(s), P(p), A(a), N(p.N), tid(threadIdx.x), n3(3 * p.N), own((int)threadIdx.x < p.N),
          hi(p.hstar > 0.0), wl(ws_layout(p.N, p.hstar > 0.0)) {
        wsg = as_gbl(A.ws + P.ws_off);
        if constexpr (LWS) ws = as_lds_mut(lds_ws);
        else ws = wsg;
        dlb = as_lds_mut(dense_blk);
        rl = static_cast<RL*>(reg_lds);
        dxs = as_lds_mut(dense_xs);
        ptab = as_lds_mut(static_cast<f32x2*>(pair_tab));
    }

    // PTAB: a lane's row of the table holds floor(N / 2) entries, the walk's
    // steps k = 1..N/2; the row stride is odd, so that the 64 lanes' reads of
    // one step spread over the banks.
    __device__ __forceinline__ int pt_stride() const { return (N >> 1) | 1; }

    __device__ __forceinline__ double bsum(double v) { return block_sum<NW>(v, sm.Red, par); }
    // the workgroup's barrier; a single-wave chain needs only the ordering of
    // its own accesses (wave_sync), not a wait for its LDS writes
    __device__ __forceinline__ void sync() {
        if constexpr (NW == 1) wave_sync();
        else __syncthreads();
    }
    __device__ __forceinline__ bool sync_or(bool p) {
        if constexpr (NW == 1) {
            wave_sync();
            return __builtin_amdgcn_ballot_w64(p) != 0;
        } else {
            return __syncthreads_or(p);
        }
    }
    __device__ __forceinline__ void bump(int64_t& c, int64_t v = 1) { if (tid == 0) c += v; }
    __device__ __forceinline__ void bump32(int32_t& c) { if (tid == 0) c += 1; }
    template <typename T, typename U>
    __device__ __forceinline__ void put(T& dst, U v) { if (tid == 0) dst = (T)v; }

    // per-bead state slots (own bead only; no synchronisation needed)
    __device__ __forceinline__ void save(int slot, const double v[3]) {
        if (own)
            for (int d = 0; d < 3; ++d) ws[wl.S + (int64_t)(slot + d) * N + tid] = v[d];
    }
    __device__ __forceinline__ void load(int slot, double v[3]) {
        for (int d = 0; d < 3; ++d) v[d] = own ? ws[wl.S + (int64_t)(slot + d) * N + tid] : 0.0;
    }

    // LWS: the warm Lanczos seed is the only workspace content a step takes
    // from the one before; a resumed segment brings it in from HBM and every
    // launch leaves it there.
    __device__ __forceinline__ void ws_to_lds() {
        if constexpr (LWS) {
            if (own)
                for (int d = 0; d < 3; ++d) {
                    const int64_t o = wl.S + (int64_t)(SLOT_SEED + d) * N + tid;
                    ws[o] = wsg[o];
                }
            sync();
        }
    }
    __device__ __forceinline__ void ws_to_hbm() {
        if constexpr (LWS) {
            if (own)
                for (int d = 0; d < 3; ++d) {
                    const int64_t o = wl.S + (int64_t)(SLOT_SEED + d) * N + tid;
                    wsg[o] = ws[o];
                }
        }
    }

    // Stage the mobility configuration: sm.R from the caller's R, and for the
    // mixed form its FP32 copy relative to bead 0.
    __device__ __forceinline__ void stage_R(const double R[3]) {
        sync();
        for (int d = 0; d < 3; ++d) sm.R[d][tid] = R[d];
        sync();
        if constexpr (S::TILES) {
            const int e = 128 * (tid >> 6) + (tid & 63);
            f32x4 r;
            if (own) {
                r.x = (float)(sm.R[0][tid] - sm.R[0][0]);
                r.y = (float)(sm.R[1][tid] - sm.R[1][0]);
                r.z = (float)(sm.R[2][tid] - sm.R[2][0]);
            } else {
                // a pad: far from the chain and from the other pads
                r.x = 1.0e5f + 64.0f * (float)(tid & 63); r.y = 1.0e5f; r.z = 1.0e5f;
            }
            sm.Pa[e].x = r.x; sm.Pa[e].y = r.y; sm.Pa[e].z = r.z;
            sm.Pa[e + 64].x = r.x; sm.Pa[e + 64].y = r.y; sm.Pa[e + 64].z = r.z;
            sync();
        } else if constexpr (S::PAIRS) {
            if (own) {
                const float rx = (float)(sm.R[0][tid] - sm.R[0][0]);
                const float ry = (float)(sm.R[1][tid] - sm.R[1][0]);
                const float rz = (float)(sm.R[2][tid] - sm.R[2][0]);
                sm.Pa[tid].x = rx; sm.Pa[tid].y = ry; sm.Pa[tid].z = rz;
                sm.Pa[tid + N].x = rx; sm.Pa[tid + N].y = ry; sm.Pa[tid + N].z = rz;
            }
            sync();
            if constexpr (PTAB) {
                // the weights apply_pairs forms for the pair (mu, mu + k),
                // the same operations on the same FP32 separations
                if (hi && own) {
                    RpyW<float> W;
                    W.init(P.rpy_f);
                    const lds_cptr<f32x4> pa = as_lds(&sm.Pa[0]) + tid;
                    const f32x4 mine = pa[0];
                    const lds_ptr<f32x2> row = ptab + tid * pt_stride();
                    const int S2 = N >> 1;
                    for (int k = 1; k <= S2; ++k) {
                        const f32x4 q = pa[k];
                        const float bx = q.x - mine.x, by = q.y - mine.y, bz = q.z - mine.z;
                        float om1, om2;
                        W.weights_floor(fmaf(bz, bz, fmaf(by, by, fmaf(bx, bx, W.clamp))), om1, om2);
                        row[k - 1] = f32x2{om1, om2};
                    }
                }
                // apply() synchronises before the walk reads the table
            }
        }
    }

    // SPLIT: this workgroup is part p of G of its chain; its state copy lives
    // in its own workspace
    __device__ __forceinline__ void set_part(int p, int G) {
        part = p;
        nparts = G;
        static_assert(!LWS, "split chains keep their workspace in HBM");
        wsg = as_gbl(A.ws + P.ws_off + (int64_t)p * wl.total);
        ws = wsg;
    }

    // SPLIT: the parts' partial rows (own lanes) go to this launch slot's
    // exchange buffer (two, alternating), the parts meet at the chain's
    // barrier, and each part sums the partials in part order -- the same
    // value in every part. A barrier that times out marks the chain failed
    // (every later exchange is skipped, so every part ends).
    // The parts of a chain run on different XCDs, whose L2 caches are not
    // coherent with each other: the partials are written and read with
    // device-scope accesses, which bypass them.
    __device__ static void xput(double* p, double v) {
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ static double xget(double* p) {
        return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(p),
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }

    __device__ __forceinline__ void exchange(double& ax, double& ay, double& az) {
        const int slot = blockIdx.x / nparts;
        double* xb = A.xbuf + ((size_t)slot * 2 + (epoch & 1)) * (size_t)nparts * 3 * 1024;
        if (own) {
            xput(xb + (part * 3 + 0) * 1024 + tid, ax);
            xput(xb + (part * 3 + 1) * 1024 + tid, ay);
            xput(xb + (part * 3 + 2) * 1024 + tid, az);
        }
        __builtin_amdgcn_s_waitcnt(0);   // this wave's partials have completed
        sync();
        if (tid == 0) {
            bool ok = !sync_fail;
            if (ok) ok = chain_barrier(A.bar + slot, (epoch + 1) * (uint32_t)nparts);
            sm.ScalI[6] = ok ? 1 : 0;
        }
        sync();
        if (!sm.ScalI[6]) sync_fail = true;
        ++epoch;
        if (own) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int g = 0; g < nparts; ++g) {
                sx += xget(xb + (g * 3 + 0) * 1024 + tid);
                sy += xget(xb + (g * 3 + 1) * 1024 + tid);
                sz += xget(xb + (g * 3 + 2) * 1024 + tid);
            }
            ax = sx;
            ay = sy;
            az = sz;
        }
    }

    // y = D(sm.R) v   (Diffusion::apply, mobility.cpp:53-56). Each lane runs
    // over all N partners with its own bead included: the self pair evaluates
    // to om1(1e-12) I and cdiag * v restores the identity block (RpyW), so the
    // pair loop has no lane-dependent select.
    __device__ __forceinline__ void apply(const double v[3], double y[3]) {
        if (!hi) { y[0] = v[0]; y[1] = v[1]; y[2] = v[2]; return; }
        sync();
        if constexpr (S::TILES) {
            const int e = 128 * (tid >> 6) + (tid & 63);
            const float vx = own ? (float)v[0] : 0.0f;
            const f32x2 vyz = {own ? (float)v[1] : 0.0f, own ? (float)v[2] : 0.0f};
            sm.Pa[e].w = vx;
            sm.Pa[e + 64].w = vx;
            sm.Pb[e] = vyz;
            sm.Pb[e + 64] = vyz;
            for (int d = 0; d < 3; ++d) sm.Acc[d][tid] = 0;
            // the biased exponent of the largest |entry| (bit patterns of
            // non-negative floats order as integers)
            int eb = __float_as_int(fmaxf(fabsf(vx), fmaxf(fabsf(vyz.x), fabsf(vyz.y)))) >> 23;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) eb = max(eb, __shfl_xor(eb, o, 64));
            if ((tid & 63) == 0) sm.Ex[tid >> 6] = eb;
        } else if constexpr (S::PAIRS) {
            if (own) {
                const float vx = (float)v[0];
                const f32x2 vyz = {(float)v[1], (float)v[2]};
                sm.Pa[tid].w = vx;
                sm.Pa[tid + N].w = vx;
                sm.Pb[tid] = vyz;
                sm.Pb[tid + N] = vyz;
            }
        } else {
            for (int d = 0; d < 3; ++d) sm.X[d][tid] = v[d];
        }
        sync();
        bump(sm.cnt.matvecs);
        if constexpr (S::PAIRS) {
            apply_pairs(v, y);
            return;
        }
        if constexpr (!MIXED && BLOCK == 64 && !SPLIT) {
            apply_pairs64(v, y);
            return;
        }
        if constexpr (S::TILES) {
            apply_tiles(v, y);
            if constexpr (SPLIT) {
                // y holds this part's share of the pair terms; every thread
                // takes part in the exchange
                exchange(y[0], y[1], y[2]);
                if (own) { y[0] += v[0]; y[1] += v[1]; y[2] += v[2]; }
            }
            return;
        }
        // the FP64 form: each lane over its partners (SPLIT: this part's share)
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (own) {
            RpyW<double> W;
            W.init(P.rpy_d);
            const lds_cptr<double> r0 = as_lds(sm.R[0]), r1 = as_lds(sm.R[1]), r2 = as_lds(sm.R[2]);
            const lds_cptr<double> v0 = as_lds(sm.X[0]), v1 = as_lds(sm.X[1]), v2 = as_lds(sm.X[2]);
            const double rx = r0[tid], ry = r1[tid], rz = r2[tid];
            const int nlo = SPLIT ? part * N / nparts : 0;
            const int nhi = SPLIT ? (part + 1) * N / nparts : N;
#pragma unroll 2
            for (int nu = nlo; nu < nhi; ++nu) {
                const double bx = r0[nu] - rx, by = r1[nu] - ry, bz = r2[nu] - rz;
                double om1, om2;
                W.weights(bx * bx + by * by + bz * bz, om1, om2);
                const double xx = v0[nu], xy = v1[nu], xz = v2[nu];
                const double t = om2 * (bx * xx + by * xy + bz * xz);
                ax = fma(om1, xx, ax); ax = fma(t, bx, ax);
                ay = fma(om1, xy, ay); ay = fma(t, by, ay);
                az = fma(om1, xz, az); az = fma(t, bz, az);
            }
            if constexpr (!SPLIT) {
                ax = fma(W.cdiag, v[0], ax);
                ay = fma(W.cdiag, v[1], ay);
                az = fma(W.cdiag, v[2], az);
            }
        }
        if constexpr (SPLIT) {
            // every thread of the workgroup takes part in the exchange
            exchange(ax, ay, az);
            if (own) {
                const double cd = RpyW<double>::cdiag_of(P.rpy_d);
                ax = fma(cd, v[0], ax);
                ay = fma(cd, v[1], ay);
                az = fma(cd, v[2], az);
            }
        }
        y[0] = ax; y[1] = ay; y[2] = az;
    }

    // D x for a single-wave chain in the mixed form, each unordered bead pair
    // evaluated once. Lane mu visits partners nu = mu + k (mod N, read from
    // the doubled copies) for k = K..1, K = floor((N-1)/2): it adds the block
    // D_{mu nu} x_nu to its own row and forms the partner row's term
    // D_{nu mu} x_mu (the same block), which joins a register accumulator that
    // moves one lane up (mod N) per step, so that after step k lane j holds
    // terms for row j + k - 1. For even N the pairs at k = N/2 are met from
    // both ends, so that step adds own-row terms only. FP32 runs of 12 steps
    // are accumulated in FP64, the moving accumulator after a permute to the
    // lane of its row. y = x + the sum (identity diagonal blocks).
    __device__ __forceinline__ void apply_pairs(const double v[3], double y[3]) {
        RpyW<float> W;
        W.init(P.rpy_f);
        // partner k of this lane: one 16-byte read (position, operand x) and
        // one 8-byte read (operand y, z), both at immediate offsets from the
        // lane's own entries
        const int me = own ? tid : 0;
        const lds_cptr<f32x4> pa = as_lds(&sm.Pa[0]) + me;
        const lds_cptr<f32x2> pb = as_lds(&sm.Pb[0]) + me;
        // PTAB: this lane's row of the step's weights, entry k - 1 for step k
        const lds_cptr<f32x2> pt = PTAB ? (lds_cptr<f32x2>)ptab + me * pt_stride() - 1 : nullptr;
        const f32x4 mine = pa[0];
        const f32x2 mine2 = pb[0];
        const float rx = mine.x, ry = mine.y, rz = mine.z;
        const float xi = mine.w, yi = mine2.x, zi = mine2.y;
        const int K = (N - 1) / 2;
        // lane j takes the moving accumulator from lane j - 1 (mod N)
        const int rot = own ? 4 * (tid == 0 ? N - 1 : tid - 1) : 4 * tid;
        auto perm = [](int addr, float x) {
            return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
        };
        double ax = 0.0, ay = 0.0, az = 0.0;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        // own-row term of the pair (mu, mu + k); with `both` also the partner's
        auto pair = [&](int k, bool both, float& fx, float& fy, float& fz) {
            const f32x4 q = pa[k];
            const f32x2 q2 = pb[k];
            const float bx = q.x - rx, by = q.y - ry, bz = q.z - rz;
            // the walk never meets the self pair: 1e-12 + |b|^2 stands for the
            // reference's clamp max(|b|^2, 1e-12) (the same value to FP32
            // resolution wherever |b|^2 > 1e-5)
            float om1, om2;
            if constexpr (PTAB) {
                const f32x2 w = pt[k];
                om1 = w.x;
                om2 = w.y;
            } else {
                W.weights_floor(fmaf(bz, bz, fmaf(by, by, fmaf(bx, bx, W.clamp))), om1, om2);
            }
            const float xx = q.w, xy = q2.x, xz = q2.y;
            const float t = om2 * (bx * xx + by * xy + bz * xz);
            fx = fmaf(om1, xx, fx); fx = fmaf(t, bx, fx);
            fy = fmaf(om1, xy, fy); fy = fmaf(t, by, fy);
            fz = fmaf(om1, xz, fz); fz = fmaf(t, bz, fz);
            if (both) {
                const float u = om2 * (bx * xi + by * yi + bz * zi);
                sx = fmaf(om1, xi, sx); sx = fmaf(u, bx, sx);
                sy = fmaf(om1, yi, sy); sy = fmaf(u, by, sy);
                sz = fmaf(om1, zi, sz); sz = fmaf(u, bz, sz);
                sx = perm(rot, sx); sy = perm(rot, sy); sz = perm(rot, sz);
            }
        };
        // after step kend the moving terms are for row j + kend: bring them to
        // that lane and fold both FP32 runs into the FP64 sums
        auto flush = [&](int kend, float fx, float fy, float fz) {
            ax += (double)fx; ay += (double)fy; az += (double)fz;
            int from = tid;
            if (own && kend > 0) { from = tid - kend; if (from < 0) from += N; }
            ax += (double)perm(4 * from, sx);
            ay += (double)perm(4 * from, sy);
            az += (double)perm(4 * from, sz);
            sx = sy = sz = 0.0f;
        };
        if ((N & 1) == 0 && N >= 2) {
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
            pair(N / 2, false, fx, fy, fz);
            ax += (double)fx; ay += (double)fy; az += (double)fz;
        }
        int k = K;
        for (; k >= 12; k -= 12) {
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
            for (int u = 0; u < 12; ++u) pair(k - u, true, fx, fy, fz);
            flush(k - 12, fx, fy, fz);
        }
        if (k > 0) {
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
            for (; k > 0; --k) pair(k, true, fx, fy, fz);
            flush(0, fx, fy, fz);
        }
        if (own) { y[0] = v[0] + ax; y[1] = v[1] + ay; y[2] = v[2] + az; }
        else { y[0] = y[1] = y[2] = 0.0; }
    }

    // D x for a single-wave chain in the FP64 form, each unordered bead pair
    // evaluated once: the walk of apply_pairs (lane mu with partner
    // (mu + k) mod N for k = K..1, and for even N k = N/2 from both ends with
    // own-row terms only; the partner row's term on an accumulator that moves
    // one lane up per step), in FP64 throughout and with the partner read at
    // a wrapped index of the staged positions and operand. After step k = 1
    // lane j's moving accumulator holds row j's terms. y = x + the sum
    // (identity diagonal blocks, as Diffusion::apply has them).
    __device__ __forceinline__ void apply_pairs64(const double v[3], double y[3]) {
        RpyW<double> W;
        W.init(P.rpy_d);
        const lds_cptr<double> r0 = as_lds(sm.R[0]), r1 = as_lds(sm.R[1]), r2 = as_lds(sm.R[2]);
        const lds_cptr<double> v0 = as_lds(sm.X[0]), v1 = as_lds(sm.X[1]), v2 = as_lds(sm.X[2]);
        const int me = own ? tid : 0;
        const double rx = r0[me], ry = r1[me], rz = r2[me];
        const double xi = v0[me], yi = v1[me], zi = v2[me];
        // lane j takes the moving accumulator from lane j - 1 (mod N)
        const int rot = own ? 4 * (tid == 0 ? N - 1 : tid - 1) : 4 * tid;
        auto perm = [](int addr, double x) {
            const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(x));
            const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(x));
            return __hiloint2double(hi, lo);
        };
        double ax = 0.0, ay = 0.0, az = 0.0;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        auto pair = [&](int nu, bool both) {
            const double bx = r0[nu] - rx, by = r1[nu] - ry, bz = r2[nu] - rz;
            double om1, om2;
            W.weights(bx * bx + by * by + bz * bz, om1, om2);
            const double xx = v0[nu], xy = v1[nu], xz = v2[nu];
            const double t = om2 * (bx * xx + by * xy + bz * xz);
            ax = fma(om1, xx, ax); ax = fma(t, bx, ax);
            ay = fma(om1, xy, ay); ay = fma(t, by, ay);
            az = fma(om1, xz, az); az = fma(t, bz, az);
            if (both) {
                const double u = om2 * (bx * xi + by * yi + bz * zi);
                sx = fma(om1, xi, sx); sx = fma(u, bx, sx);
                sy = fma(om1, yi, sy); sy = fma(u, by, sy);
                sz = fma(om1, zi, sz); sz = fma(u, bz, sz);
                sx = perm(rot, sx); sy = perm(rot, sy); sz = perm(rot, sz);
            }
        };
        if ((N & 1) == 0 && N >= 2) {
            int nu = me + N / 2;
            if (nu >= N) nu -= N;
            pair(nu, false);
        }
        const int K = (N - 1) / 2;
        int nu = me + K;
        if (nu >= N) nu -= N;
#pragma unroll 4
        for (int k = K; k >= 1; --k) {
            pair(nu, true);
            nu = (nu == 0) ? N - 1 : nu - 1;
        }
        if (own) { y[0] = v[0] + (ax + sx); y[1] = v[1] + (ay + sy); y[2] = v[2] + (az + sz); }
        else { y[0] = y[1] = y[2] = 0.0; }
    }

    // D x for a chain of 2..16 waves in the mixed form, each unordered bead
    // pair evaluated once. The beads form tiles of 64 (one per wave, the last
    // padded). Wave w walks (1) its own tile as the one-wave pair walk does
    // (a ring of 64: offsets 31..1 with the partner-row terms on a register
    // accumulator rotated one lane per step, and offset 32 from both ends,
    // own-row terms only), and (2) every other tile t over 32 of the 64 lane
    // offsets k (partner (l + k) mod 64): k = 63..32 if w < t, else 64..33,
    // which between the two waves covers each cross pair once. The partner
    // rows of (2) belong to wave t: their rotated sums go, every 16 steps, to
    // Acc as fixed-point integers (exact and order-independent sums), read
    // back after a barrier. The fixed-point unit is 2^-ks with ks = 170 - e,
    // e the largest biased FP32 exponent of the operand's entries (|x| <
    // 2^(e-126)): a pair term's components are below 2 |x|max (RPY weights:
    // om1 <= 1, om2 |b|^2 < 0.21), so a row's sum over N <= 1024 partners
    // stays below 2^(1+10+e-126) units of 2^-ks, i.e. 2^55, and the unit is at
    // most 2^-43 of the largest entry, far below the FP32 rounding of the runs
    // it adds. FP32 runs of at most 16 steps are folded into FP64 sums, and
    // each walk's own-row sums go to Acc in fixed point as well, so every row
    // is one exact integer sum whichever wave walked each part.
    // y = x + the sum (identity diagonal blocks). A chain split over parts
    // (SPLIT): part p walks the tile pairs (w, t) with t = p (mod parts), own
    // tiles included, and y is its share of the sum alone (no x). A workgroup
    // with more waves than tiles spreads the walks over all its waves, with
    // the same result bit for bit.
    __device__ __forceinline__ void apply_tiles(const double v[3], double y[3]) {
        int e = 0;
#pragma unroll
        for (int i = 0; i < S::NW; ++i) e = max(e, sm.Ex[i]);
        const int ks = min(170 - e, 126);                // in [-85, 126]
        const float FIX = __int_as_float((127 + ks) << 23);         // 2^ks
        const double inv = __hiloint2double((1023 - ks) << 20, 0); // 2^-ks
        RpyW<float> W;
        W.init(P.rpy_f);
        const int w = tid >> 6, l = tid & 63;
        const int T = (N + 63) >> 6;
        // more waves than tiles (a chain on a wider workgroup than its beads
        // need): the T^2 walks of 32 pair steps (own tile w; the half of the
        // cross pair (w, t) walked from w) go round robin to all waves, and
        // every row's sums, own rows included, go through Acc -- where that
        // shortens the longest wave's share (ceil(T^2 / NW) < T; built for
        // workgroups of 8 waves or more, where that can happen by default)
        constexpr bool CAN_SPREAD = !SPLIT && S::NW >= 8;
        const bool spread = CAN_SPREAD && T * T <= S::NW * (T - 1);
        const int rot = 4 * ((l + 63) & 63);            // take from lane l - 1
        auto perm = [](int addr, float x) {
            return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
        };
        const lds_cptr<f32x4> pa0 = as_lds(&sm.Pa[0]);
        const lds_cptr<f32x2> pb0 = as_lds(&sm.Pb[0]);
        // LDS-qualified, so that these are ds_add_u64 (not flat atomics)
        const lds_ptr<unsigned long long> acc0 =
            as_lds_mut(reinterpret_cast<unsigned long long*>(&sm.Acc[0][0]));
        // one walk: this lane's bead of tile tw against tile tt (tt == tw:
        // the own tile), own-row sums into (ox, oy, oz); OWN: 1 own tile only,
        // 0 other tiles only, -1 either
        auto walk_t = [&](auto own_sel, int tw, int tt, double& ox, double& oy, double& oz) {
            constexpr int OWN = decltype(own_sel)::value;
            const f32x4 mine = pa0[128 * tw + l];
            const f32x2 mine2 = pb0[128 * tw + l];
            const float rx = mine.x, ry = mine.y, rz = mine.z;
            const float xi = mine.w, yi = mine2.x, zi = mine2.y;
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            const lds_cptr<f32x4> pa = pa0 + 128 * tt + l;
            const lds_cptr<f32x2> pb = pb0 + 128 * tt + l;
            auto pair = [&](int k, bool both, float& fx, float& fy, float& fz) {
                const f32x4 q = pa[k];
                const f32x2 q2 = pb[k];
                const float bx = q.x - rx, by = q.y - ry, bz = q.z - rz;
                float om1, om2;
                W.weights_floor(fmaf(bz, bz, fmaf(by, by, fmaf(bx, bx, W.clamp))), om1, om2);
                const float xx = q.w, xy = q2.x, xz = q2.y;
                const float t = om2 * (bx * xx + by * xy + bz * xz);
                fx = fmaf(om1, xx, fx); fx = fmaf(t, bx, fx);
                fy = fmaf(om1, xy, fy); fy = fmaf(t, by, fy);
                fz = fmaf(om1, xz, fz); fz = fmaf(t, bz, fz);
                if (both) {
                    const float u = om2 * (bx * xi + by * yi + bz * zi);
                    sx = fmaf(om1, xi, sx); sx = fmaf(u, bx, sx);
                    sy = fmaf(om1, yi, sy); sy = fmaf(u, by, sy);
                    sz = fmaf(om1, zi, sz); sz = fmaf(u, bz, sz);
                    sx = perm(rot, sx); sy = perm(rot, sy); sz = perm(rot, sz);
                }
            };
            if (OWN == 1 || (OWN == -1 && tt == tw)) {
                // (1) own tile: after step kend the rotated terms are for row
                // l + kend - 1 ... brought to their lane by one permute
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
                pair(32, false, fx, fy, fz);
                ox += (double)fx; oy += (double)fy; oz += (double)fz;
                auto flush = [&](int kend, float gx, float gy, float gz) {
                    ox += (double)gx; oy += (double)gy; oz += (double)gz;
                    const int from = 4 * ((l - kend) & 63);
                    ox += (double)perm(from, sx);
                    oy += (double)perm(from, sy);
                    oz += (double)perm(from, sz);
                    sx = sy = sz = 0.0f;
                };
                int k = 31;
                for (; k >= 16; k -= 16) {
                    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll 4
                    for (int u = 0; u < 16; ++u) pair(k - u, true, gx, gy, gz);
                    flush(k - 16, gx, gy, gz);
                }
                {
                    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                    for (; k > 0; --k) pair(k, true, gx, gy, gz);
                    flush(0, gx, gy, gz);
                }
            } else {
                // (2) another tile: 32 offsets, in two runs of 16; after a run
                // ending at offset kend, lane j's rotated sums are for row
                // (j - 1 + kend) mod 64 of tile tt
                const int khi = tw < tt ? 63 : 64;
#pragma unroll 1
                for (int r = 0; r < 2; ++r) {
                    const int k0 = khi - 16 * r;
                    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll 4
                    for (int u = 0; u < 16; ++u) pair(k0 - u, true, gx, gy, gz);
                    ox += (double)gx; oy += (double)gy; oz += (double)gz;
                    const lds_ptr<unsigned long long> acc = acc0 + 64 * tt + ((l + k0 - 16) & 63);
                    __hip_atomic_fetch_add(acc, (unsigned long long)(long long)__float2ll_rn(sx * FIX),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(acc + S::AB, (unsigned long long)(long long)__float2ll_rn(sy * FIX),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    __hip_atomic_fetch_add(acc + 2 * S::AB, (unsigned long long)(long long)__float2ll_rn(sz * FIX),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    sx = sy = sz = 0.0f;
                }
            }
        };
        auto walk = [&](int tw, int tt, double& ox, double& oy, double& oz) {
            walk_t(std::integral_constant<int, -1>{}, tw, tt, ox, oy, oz);
        };
        auto walk_own = [&](int tw, double& ox, double& oy, double& oz) {
            walk_t(std::integral_constant<int, 1>{}, tw, tw, ox, oy, oz);
        };
        auto walk_cross = [&](int tw, int tt, double& ox, double& oy, double& oz) {
            walk_t(std::integral_constant<int, 0>{}, tw, tt, ox, oy, oz);
        };
        // Every walk's own-row sums go to Acc in fixed point too, so a row's
        // sum is the same integer whichever wave walked each part: the result
        // does not depend on the workgroup's width (or on the order of waves).
        const double FIXd = (double)FIX;
        auto flush_own = [&](int tw, double ox, double oy, double oz) {
            const lds_ptr<unsigned long long> acc = acc0 + 64 * tw + l;
            __hip_atomic_fetch_add(acc, (unsigned long long)(long long)__double2ll_rn(ox * FIXd),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc + S::AB, (unsigned long long)(long long)__double2ll_rn(oy * FIXd),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(acc + 2 * S::AB, (unsigned long long)(long long)__double2ll_rn(oz * FIXd),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        };
        // this wave's walks: spread, units u = w, w + NW, ... as (u / T, u mod
        // T); otherwise (wave w < T) its own tile, then the other tiles
        // (SPLIT: the tiles t = part (mod parts), own tile included)
        if (spread) {
            for (int u = w; u < T * T; u += S::NW) {
                const int tw = u / T, tt = u - T * tw;
                double ox = 0.0, oy = 0.0, oz = 0.0;
                walk(tw, tt, ox, oy, oz);
                flush_own(tw, ox, oy, oz);
            }
        } else if (w < T) {
            if (!SPLIT || w % nparts == part) {
                double ox = 0.0, oy = 0.0, oz = 0.0;
                walk_own(w, ox, oy, oz);
                flush_own(w, ox, oy, oz);
            }
            for (int t = SPLIT ? part : 0; t < T; t += SPLIT ? nparts : 1) {
                if (t == w) continue;
                double ox = 0.0, oy = 0.0, oz = 0.0;
                walk_cross(w, t, ox, oy, oz);
                flush_own(w, ox, oy, oz);
            }
        }
        sync();
        if (own) {
            const double x0 = SPLIT ? 0.0 : v[0], x1 = SPLIT ? 0.0 : v[1], x2 = SPLIT ? 0.0 : v[2];
            y[0] = x0 + (double)sm.Acc[0][tid] * inv;
            y[1] = x1 + (double)sm.Acc[1][tid] * inv;
            y[2] = x2 + (double)sm.Acc[2][tid] * inv;
        } else {
            y[0] = y[1] = y[2] = 0.0;
        }
    }

    // Spring force on the beads from the bonds held by threads tid < N-1
    // (spring.cpp:183-207). Returns the bead force; Fc gets the own bond force.
    __device__ __forceinline__ void spring_bead_force(const double bond[3], double Fc[3], double F[3]) {
        Fc[0] = Fc[1] = Fc[2] = 0.0;
        if (tid < N - 1) connector_force(P, bond, Fc);
        sync();
        for (int d = 0; d < 3; ++d) sm.V[d][tid] = Fc[d];
        sync();
        for (int d = 0; d < 3; ++d) {
            double f = 0.0;
            if (tid == 0) f = sm.V[d][0];
            else if (tid < N - 1) f = sm.V[d][tid] - sm.V[d][tid - 1];
            else if (tid == N - 1) f = (-1.0) * sm.V[d][N - 2];
            F[d] = f;
        }
    }

    // Bending force (bending.cpp:5-43) at configuration C (LDS, [3][BLOCK]).
    __device__ __forceinline__ void bending_force(const double (*C)[BLOCK], double F[3]) {
        F[0] = F[1] = F[2] = 0.0;
        if (P.bend_type == 0 || N == 2) return;
        const double Cst = P.bend_C;
        sync();
        if (tid < N - 1) {
            double b[3];
            for (int d = 0; d < 3; ++d) b[d] = C[d][tid + 1] - C[d][tid];
            double sq = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
            if (sq < 1.0e-12) sq = 1.0e-12;
            const double len = sqrt(sq);
            for (int d = 0; d < 3; ++d) sm.B[d][tid] = (1.0 / len) * b[d];
            sm.B[3][tid] = len;
        }
        sync();
        if (!own) return;
        const int m = tid;
        auto u = [&](int k, int d) { return sm.B[d][k]; };
        auto len = [&](int k) { return sm.B[3][k]; };
        // bend cosines cs[c] = u_{c+1} . u_c (config.cpp:41-45) of the three
        // angles this bead takes part in
        double csl[3];
        for (int c = m - 2, i = 0; i < 3; ++c, ++i)
            csl[i] = (c >= 0 && c < N - 2)
                         ? u(c + 1, 0) * u(c, 0) + u(c + 1, 1) * u(c, 1) + u(c + 1, 2) * u(c, 2)
                         : 0.0;
        auto cs = [&](int c) { return csl[c - (m - 2)]; };
        double A_[3], B_[3], C_[3], D_[3];
        auto term = [&](double s, int a, double ca, int bb, double* out) {   // s*(ca*u_a - u_b)
            for (int d = 0; d < 3; ++d) out[d] = s * (ca * u(a, d) - u(bb, d));
        };
        auto term2 = [&](double s, double ca, int a, int bb, double* out) {  // s*((-ca)*u_a + u_b)
            for (int d = 0; d < 3; ++d) out[d] = s * ((-ca) * u(a, d) + u(bb, d));
        };
        if (m == 0) { term(Cst / len(0), 0, cs(0), 1, F); return; }
        if (m == N - 1) { term2(Cst / len(N - 2), cs(N - 3), N - 2, N - 3, F); return; }
        if (N == 3) {
            term(1.0 / len(1), 1, cs(0), 0, A_);
            term2(1.0 / len(0), cs(0), 0, 1, B_);
            for (int d = 0; d < 3; ++d) F[d] = Cst * (A_[d] + B_[d]);
            return;
        }
        if (m == 1) {
            term(1.0 / len(1), 1, cs(0), 0, A_);
            term2(1.0 / len(0), cs(0), 0, 1, B_);
            term(1.0 / len(1), 1, cs(1), 2, D_);
            for (int d = 0; d < 3; ++d) F[d] = Cst * (A_[d] + B_[d] + D_[d]);
            return;
        }
        if (m == N - 2) {
            term(1.0 / len(N - 2), N - 2, cs(N - 3), N - 3, A_);
            term2(1.0 / len(N - 3), cs(N - 3), N - 3, N - 2, B_);
            term2(1.0 / len(N - 3), cs(N - 4), N - 3, N - 4, C_);
            for (int d = 0; d < 3; ++d) F[d] = Cst * (A_[d] + B_[d] + C_[d]);
            return;
        }
        term(1.0 / len(m), m, cs(m - 1), m - 1, A_);
        term2(1.0 / len(m - 1), cs(m - 1), m - 1, m, B_);
        term2(1.0 / len(m - 1), cs(m - 2), m - 1, m - 2, C_);
        term(1.0 / len(m), m, cs(m), m + 1, D_);
        for (int d = 0; d < 3; ++d) F[d] = Cst * (A_[d] + B_[d] + C_[d] + D_[d]);
    }

    // Excluded-volume force on this lane's bead (excluded_volume.cpp:60-89) at
    // configuration C (LDS, [3][BLOCK]): the sum over partners nu with
    // |nu - mu| >= cd of -fac(|q|^2) q, q = C_nu - C_mu. The potential type,
    // its constants and the contour exclusion are resolved before the pair
    // loop. With cd <= 1 only the self pair is excluded, and it needs no test:
    // q = 0 there, so its term vanishes whatever the (finite) factor. The
    // mixed form works on FP32 positions relative to bead 0, with FP32 pair
    // terms summed over runs of 8 partners and accumulated in FP64.
    __device__ __forceinline__ void ev_force(const double (*C)[BLOCK], double F[3]) {
        F[0] = F[1] = F[2] = 0.0;
        const int et = P.ev_type;
        if (et != 1 && et != 2 && et != 3) return;
        if constexpr (MIXED) {
            sync();
            if (own)
                for (int d = 0; d < 3; ++d) sm.W[d][tid] = (float)(C[d][tid] - C[d][0]);
            sync();
        }
        if (own) {
            // The self pair takes the factor's value at Rmin, finite in FP32
            // for Rmin >= 0.1; a smaller soft-core radius keeps the explicit
            // exclusion test, which discards that value.
            const bool cdt = P.cd > 1 || (et != 1 && !(P.Rmin2 >= 1.0e-2));
            if (et == 1) {
                if (cdt) ev_loop<1, false, true>(C, F); else ev_loop<1, false, false>(C, F);
            } else if (et == 2) {
                if (cdt) ev_loop<2, false, true>(C, F); else ev_loop<2, false, false>(C, F);
            } else if (P.has_phi && !P.equil) {
                ev_loop<3, true, true>(C, F);
            } else {
                if (cdt) ev_loop<3, false, true>(C, F); else ev_loop<3, false, false>(C, F);
            }
        }
        // SPLIT: each part summed its share of the partners
        if constexpr (SPLIT) exchange(F[0], F[1], F[2]);
    }

    // The pair factor (excluded_volume.cpp:36-58) from the clamped r^2: Gauss
    // z*/d*^5 exp(-r^2 / 2d*^2); LJ and SDK a/r^14 - b/r^8 = (a/r^6 - b)/r^8
    // between Rmin and the cutoff (compared squared) and its value at Rmin
    // below, i.e. the same expression at max(r^2, Rmin^2), so there is one
    // select and no branch; with an attraction table the SDK tail
    // alpha phi sin(alpha r^2 + beta) between the cutoffs.
    template <int EVT, bool TAIL, bool CDT>
    __device__ __forceinline__ void ev_loop(const double (*C)[BLOCK], double F[3]) {
        using T = typename std::conditional<MIXED, float, double>::type;
        const T g_a = (T)P.zsbyds5, g_b = (T)(-P.pt5bydssq);
        const T ca = (T)(EVT == 2 ? P.LJa : P.SDKa), cb = (T)(EVT == 2 ? P.LJb : P.SDKb);
        const T rmin2 = (T)P.Rmin2;
        const T rcut2 = (T)(EVT == 2 ? (P.equil ? P.Rcutg2 : P.Rcutp2) : P.Rcutg2);
        // below Rmin the factor is constant even where Rmin exceeds the cutoff
        const T rcut2e = fmax(rcut2, rmin2);
        const T rtail2 = (T)P.RcutpSDK2;
        const double* phi = A.blob + P.phi_off;
        const int cd = P.cd;
        lds_cptr<T> x0, x1, x2;
        if constexpr (MIXED) {
            x0 = as_lds(sm.W[0]); x1 = as_lds(sm.W[1]); x2 = as_lds(sm.W[2]);
        } else {
            x0 = as_lds(C[0]); x1 = as_lds(C[1]); x2 = as_lds(C[2]);
        }
        const T px = x0[tid], py = x1[tid], pz = x2[tid];
        auto pair = [&](int nu, T& fx, T& fy, T& fz) {
            const T qx = x0[nu] - px, qy = x1[nu] - py, qz = x2[nu] - pz;
            const T sq = fmax(qx * qx + qy * qy + qz * qz, (T)1.0e-12);
            T fac;
            if constexpr (EVT == 1) {
                fac = g_a * exp(g_b * sq);
            } else {
                const T inv2 = rcp_fast(fmax(sq, rmin2));
                const T inv4 = inv2 * inv2, inv8 = inv4 * inv4;
                const T core = fma(ca * inv4, inv2, -cb) * inv8;
                fac = sq <= rcut2e ? core : (T)0;
                if constexpr (TAIL) {
                    if (sq <= rtail2 && sq > rcut2) {
                        const T SDK_ALPHA = (T)1.530633312, SDK_BETA = (T)1.213115524;
                        const T ph = (T)(nu > tid ? phi[tid * N + nu] : phi[nu * N + tid]);
                        fac += SDK_ALPHA * ph * sin(SDK_ALPHA * sq + SDK_BETA);
                    }
                }
            }
            if constexpr (CDT) {
                const int sep = nu > tid ? nu - tid : tid - nu;
                if (sep < cd) fac = (T)0;
            }
            fx = fma(-fac, qx, fx);
            fy = fma(-fac, qy, fy);
            fz = fma(-fac, qz, fz);
        };
        const int nlo = SPLIT ? part * N / nparts : 0;
        const int nhi = SPLIT ? (part + 1) * N / nparts : N;
        if constexpr (MIXED) {
            double ax = 0.0, ay = 0.0, az = 0.0;
            int nu = nlo;
            for (; nu + 8 <= nhi; nu += 8) {
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll
                for (int k = 0; k < 8; ++k) pair(nu + k, fx, fy, fz);
                ax += (double)fx; ay += (double)fy; az += (double)fz;
            }
            if (nu < nhi) {
                float fx = 0.0f, fy = 0.0f, fz = 0.0f;
                for (; nu < nhi; ++nu) pair(nu, fx, fy, fz);
                ax += (double)fx; ay += (double)fy; az += (double)fz;
            }
            F[0] = ax; F[1] = ay; F[2] = az;
        } else {
            double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll 2
            for (int nu = nlo; nu < nhi; ++nu) pair(nu, fx, fy, fz);
            F[0] = fx; F[1] = fy; F[2] = fz;
        }
    }

    // non-spring + external force at C and time t (model.hpp:56-64)
    __device__ __forceinline__ void explicit_force(const double (*C)[BLOCK], double t, double Fe[3]) {
        double Fev[3], Fb[3], Fx[3] = {0.0, 0.0, 0.0};
        ev_force(C, Fev);
        bending_force(C, Fb);
        if (P.n_ext > 0 && own) external_at(P, A.ext, A.blob, N, tid, t, Fx);
        for (int d = 0; d < 3; ++d) Fe[d] = (Fev[d] + Fb[d]) + Fx[d];
    }

    // Dense RPY tensor and its lower Cholesky factor: in LDS (3N <= 16 NB)
    // or in the workspace (dense_factor_ool); false on a non-positive pivot.
    // An FP32 LDS factor whose pivot fails is redone in FP64 in the
    // workspace, so the FP32 variant fails only where FP64 does.
    __device__ __forceinline__ bool dense_factor() {
        dense_ws = false;
        if constexpr (NB > 0 && !REGF) {
            if (n3 <= 16 * NB) {
                if (dense_factor_lds<BLOCK, MIXED, NB, FT>(sm, dlb, N, P.hstar)) return true;
                if constexpr (sizeof(FT) == 8) return false;
                dense_ws = true;
                const bool ok = dense_factor_ool<BLOCK, MIXED>(sm, wsg + wl.dense, N, P.hstar);
                if (ok) { put(sm.cnt.fail_index, 0); put(sm.cnt.fail_value, 0.0); }
                return ok;
            }
        }
        return dense_factor_ool<BLOCK, MIXED>(sm, wsg + wl.dense, N, P.hstar);
    }

    // y = L x with the factor where dense_factor left it, own rows.
    __device__ __forceinline__ void lower_matvec(const double x[3], double y[3]) {
        D3 r;
        if constexpr (NB > 0 && !REGF) {
            if (n3 <= 16 * NB && !dense_ws) r = lower_matvec_lds<BLOCK, MIXED, NB, FT>(dlb, dxs, N, x[0], x[1], x[2]);
            else r = lower_matvec_ool<BLOCK, MIXED>(sm, wsg + wl.dense, N, x[0], x[1], x[2]);
        } else {
            r = lower_matvec_ool<BLOCK, MIXED>(sm, wsg + wl.dense, N, x[0], x[1], x[2]);
        }
        y[0] = r.v[0]; y[1] = r.v[1]; y[2] = r.v[2];
    }

    // default Lanczos seed sin(0.7654321 (i+1)) + 0.5 (hydrodynamics.cpp:83-87)
    __device__ __forceinline__ void default_seed(double u[3]) {
        for (int d = 0; d < 3; ++d) u[d] = own ? sin(0.7654321 * (3 * tid + d + 1)) + 0.5 : 0.0;
    }

    // m-step Lanczos with full reorthogonalisation (hydrodynamics.cpp:146-220)
    // from the seed slot (warm) or the default seed. Returns the Krylov
    // dimension; extremal Ritz values out, Ritz vectors into the seed slot as
    // vmin + vmax (the next seed is built from their sum only).
    __device__ __forceinline__ int lanczos(int m, bool warm, double& rmin, double& rmax, double vsum[3]) {
        if (m > n3) m = n3;
        const ws_ptr Vb = ws + wl.V;
        double v[3], vprev[3] = {0.0, 0.0, 0.0}, w[3];
        if (warm) load(SLOT_SEED, v); else default_seed(v);
        double nv = sqrt(bsum(dot3(v, v)));
        if (!(nv > 0.0)) {
            default_seed(v);
            nv = sqrt(bsum(dot3(v, v)));
        }
        for (int d = 0; d < 3; ++d) v[d] /= nv;
        double b = 0.0;
        int k = 0;
        const int n = n3;
        for (int j = 0; j < m; ++j) {
            if (own)
                for (int d = 0; d < 3; ++d) Vb[(size_t)j * n + 3 * tid + d] = v[d];
            apply(v, w);
            bump(sm.cnt.lanczos_iters);
            const double a = bsum(dot3(v, w));
            if (tid == 0) sm.Alpha[j] = a;
            k = j + 1;
            for (int d = 0; d < 3; ++d) w[d] -= a * v[d] + b * vprev[d];
            // full reorthogonalisation, modified Gram-Schmidt in the
            // reference's order (q = 0..j, the last being v itself). The
            // basis vectors do not depend on w, so they are fetched four at a
            // time ahead of the dependent chain of dot products and updates.
            for (int q0 = 0; q0 <= j; q0 += 4) {
                double u[4][3];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int q = q0 + t;
                    for (int d = 0; d < 3; ++d)
                        u[t][d] = (q == j) ? v[d]
                                : (own && q < j) ? Vb[(size_t)q * n + 3 * tid + d] : 0.0;
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (q0 + t > j) break;
                    const double dq = bsum(dot3(w, u[t]));
                    for (int d = 0; d < 3; ++d) w[d] -= dq * u[t][d];
                }
            }
            b = sqrt(bsum(dot3(w, w)));
            if (b < 1.0e-10) break;
            for (int d = 0; d < 3; ++d) { vprev[d] = v[d]; v[d] = w[d] / b; }
            if (tid == 0) sm.Beta[j] = b;
        }
        sync();
        if (k == 1) {
            rmin = rmax = sm.Alpha[0];
            for (int d = 0; d < 3; ++d) {
                const double v0 = own ? Vb[3 * tid + d] : 0.0;
                vsum[d] = v0 + v0;
            }
            return k;
        }
        // rows c, dd, y of the two inverse iterations: the V area is free here
        static_assert(sizeof(sm.V) >= 6 * LANCZOS_COLD * sizeof(double), "tridiagonal rows in V");
        const lds_ptr<double> tri = as_lds_mut(&sm.V[0][0]);
        if (tid < 64) {
            if constexpr (RITZ32) ritz_multisect<float>(k, as_lds(sm.Alpha), as_lds(sm.Beta), sm.Scal);
            else ritz_multisect<double>(k, as_lds(sm.Alpha), as_lds(sm.Beta), sm.Scal);
        }
        sync();
        rmin = sm.Scal[0];
        rmax = sm.Scal[1];
        // Ritz vectors of the two extremes by inverse iteration (lanes 0, 1)
        if (tid < 2) {
            const lds_ptr<double> rows = tri + tid * 3 * LANCZOS_COLD;
            if constexpr (RITZ32) {
                const lds_ptr<float> fr = (lds_ptr<float>)rows;
                tridiag_eigenvector<float>(k, as_lds(sm.Alpha), as_lds(sm.Beta), tid == 0 ? rmin : rmax, fr,
                                           fr + LANCZOS_COLD, fr + 2 * LANCZOS_COLD);
            } else {
                tridiag_eigenvector<double>(k, as_lds(sm.Alpha), as_lds(sm.Beta), tid == 0 ? rmin : rmax, rows,
                                            rows + LANCZOS_COLD, rows + 2 * LANCZOS_COLD);
            }
        }
        sync();
        // lift the tridiagonal eigenvectors: v = V^T y (hydrodynamics.cpp:264-273)
        double vmin[3] = {0.0, 0.0, 0.0}, vmax[3] = {0.0, 0.0, 0.0};
        if (own) {
            // lane 0's y row holds the smallest, lane 1's the largest
            using Y = std::conditional_t<RITZ32, float, double>;
            const lds_cptr<Y> ymn = (lds_cptr<Y>)tri + 2 * LANCZOS_COLD;
            const lds_cptr<Y> ymx = (lds_cptr<Y>)(tri + 3 * LANCZOS_COLD) + 2 * LANCZOS_COLD;
            for (int j = 0; j < k; ++j) {
                const double a0 = ymn[j], a1 = ymx[j];
                for (int d = 0; d < 3; ++d) {
                    const double vj = Vb[(size_t)j * n + 3 * tid + d];
                    vmin[d] += a0 * vj;
                    vmax[d] += a1 * vj;
                }
            }
        }
        for (int d = 0; d < 3; ++d) vsum[d] = vmin[d] + vmax[d];
        return k;
    }

    // Chebyshev coefficients of sqrt on the interval (hydrodynamics.cpp:267-279)
    // into the workspace: one coefficient per thread.
    __device__ __forceinline__ void cheb_coefficients(int L, double da, double db) {
        const ws_ptr coef = ws + wl.coef;
        const ws_ptr aux = ws + wl.aux;
        sync();
        if constexpr (MIXED) {
            // FP32 nodes and sums: cos(pi m / (2(L+1))) = cos(2 pi t) with
            // t = (m mod 4(L+1)) / (4(L+1)) reduced exactly in integers, on
            // the hardware cosine (argument in turns)
            const int per = 4 * (L + 1);
            const float inv = 1.0f / (float)per;
            const float fda = (float)da, fdb = (float)db;
            for (int k = tid; k <= L; k += BLOCK) {
                const float c = __builtin_amdgcn_cosf((float)(2 * k + 1) * inv);
                aux[k] = (double)sqrtf(c / fda - fdb / fda);
            }
            sync();
            for (int j = tid; j <= L; j += BLOCK) {
                float s = 0.0f;
                for (int k = 0; k <= L; ++k)
                    s += (float)aux[k] * __builtin_amdgcn_cosf((float)((j * (2 * k + 1)) % per) * inv);
                s *= 2.0f / (float)(L + 1);
                if (j == 0) s *= 0.5f;
                coef[j] = (double)s;
            }
        } else {
            for (int k = tid; k <= L; k += BLOCK) aux[k] = sqrt(cos(PI * (k + 0.5) / (L + 1)) / da - db / da);
            sync();
            for (int j = tid; j <= L; j += BLOCK) {
                double s = 0.0;
                for (int k = 0; k <= L; ++k) s += aux[k] * cos(j * (k + 0.5) * PI / (L + 1));
                s *= 2.0 / (L + 1);
                if (j == 0) s /= 2.0;
                coef[j] = s;
            }
        }
        sync();
    }

    // B.x0 by the Chebyshev method with Lanczos bounds and FD control
    // (hydrodynamics.cpp:224-337, mobility.cpp:64-80), x0 in SLOT_X0. Returns
    // false when the fallback factorisation fails.
    __device__ __forceinline__ bool chebyshev_bx0(bool& seed_valid, double bx0[3]) {
        StepLog& lg = sm.lg;
        double rmin, rmax, vsum[3];
        bool cold = !seed_valid;
        int k = lanczos(seed_valid ? LANCZOS_WARM : LANCZOS_COLD, seed_valid, rmin, rmax, vsum);
        if (seed_valid && k < 4) {
            k = lanczos(LANCZOS_COLD, false, rmin, rmax, vsum);
            cold = true;
        }
        if (cold) bump(sm.cnt.lanczos_cold); else bump(sm.cnt.lanczos_warm);
        put(lg.lanczos_cold, cold ? 1 : 0);
        put(lg.lanczos_steps, k);
        {
            // next seed: normalize(vmin + vmax + 0.01 u/|u|)  (hydrodynamics.cpp:294-311)
            double u[3], s[3];
            default_seed(u);
            const double un = sqrt(bsum(dot3(u, u)));
            for (int d = 0; d < 3; ++d) s[d] = vsum[d] + 0.01 * u[d] / un;
            const double sn = sqrt(bsum(dot3(s, s)));
            if (sn > 0.0 && isfinite(sn)) {
                for (int d = 0; d < 3; ++d) s[d] = s[d] / sn;
                save(SLOT_SEED, s);
                seed_valid = true;
            } else {
                seed_valid = false;
            }
        }
        double lmin = 0.5 * rmin;
        double lmax = 1.02 * rmax;
        if (lmin < 1.0e-12) lmin = 1.0e-12;
        if (lmin >= lmax) lmin = lmax * 1.0e-6;
        put(lg.lmin, lmin);
        put(lg.lmax, lmax);

        double x0[3], Dx0[3];
        load(SLOT_X0, x0);
        apply(x0, Dx0);
        const double x0Dx0 = bsum(dot3(x0, Dx0));
        int L = (int)(sqrt(lmax / lmin) * P.mult + 0.5) + 1;
        if (L > MAXCHEB) L = MAXCHEB;
        if (L < 2) L = 2;
        const double da = 2.0 / (lmax - lmin);
        const double db = -(lmax + lmin) / (lmax - lmin);
        const ws_ptr coef = ws + wl.coef;
        double fd = 0.0;
        for (;;) {
            cheb_coefficients(L, da, db);
            double tprev[3], tcur[3], res[3];
            for (int d = 0; d < 3; ++d) {
                res[d] = coef[0] * x0[d];
                tprev[d] = x0[d];
                tcur[d] = da * Dx0[d] + db * x0[d];
                res[d] += coef[1] * tcur[d];
            }
            for (int l = 2; l <= L; ++l) {
                double Dv[3];
                apply(tcur, Dv);
                const double cl = coef[l];
                for (int d = 0; d < 3; ++d) {
                    double tn = da * Dv[d] + db * tcur[d];
                    tn = 2.0 * tn - tprev[d];
                    tprev[d] = tcur[d];
                    tcur[d] = tn;
                    res[d] += cl * tcur[d];
                }
            }
            bump(sm.cnt.cheb_attempts);
            bump(sm.cnt.cheb_terms, L);
            bump32(lg.cheb_attempts);
            fd = fabs(bsum(dot3(res, res)) - x0Dx0) / x0Dx0;
            for (int d = 0; d < 3; ++d) bx0[d] = res[d];
            if (fd <= P.fd_tol || L >= MAXCHEB) break;
            const int grow = (L / 4 > 1) ? L / 4 : 1;
            L = (L + grow < MAXCHEB) ? L + grow : MAXCHEB;
        }
        put(lg.cheb_L, L);
        put(lg.fd_err, fd);
        if (!(fd <= P.fd_tol)) {
            bump(sm.cnt.fallbacks);
            put(lg.fallback, 1);
            if (!dense_factor()) return false;
            lower_matvec(x0, bx0);
        }
        return true;
    }

    // B x0 of the chain's method at the staged configuration with an empty
    // spectral cache (mobility.cpp:58-88 without the sqrt(dt) factor): x0 for
    // free draining, L x0 for Cholesky, the Chebyshev series of sqrt(D) x0 on
    // cold Lanczos bounds. For parity checks of the forms.
    __device__ __forceinline__ int brownian(const double x0[3], double bx0[3]) {
        if (!hi) {
            for (int d = 0; d < 3; ++d) bx0[d] = x0[d];
            return ST_OK;
        }
        if (P.hi_method == 1) {
            if (!dense_factor()) return ST_NOT_PD;
            lower_matvec(x0, bx0);
            return ST_OK;
        }
        if (P.hi_method != 0) return ST_EXACT_SQRT;
        save(SLOT_X0, x0);
        bool seed_valid = false;
        return chebyshev_bx0(seed_valid, bx0) ? ST_OK : ST_NOT_PD;
    }

    // inlined into each kernel so that the kernel's waves-per-EU budget applies
    __device__ __forceinline__ void run(int ti) {
        double R[3] = {0.0, 0.0, 0.0};
        if (own)
            for (int d = 0; d < 3; ++d) R[d] = A.R[P.R_off + 3 * tid + d];
        int32_t ix = A.rng[ti].ix, iy = A.rng[ti].iy, rseed = A.rng[ti].seed;
        bool rng_ready = false;
        const bool resume = A.seg != nullptr;
        bool seed_valid = resume && A.seg[ti].seed_valid != 0;
        int status = ST_OK;

        const double t0 = A.sim_override ? A.t0 : P.t0;
        const double t1 = A.sim_override ? A.t1 : P.t1;
        const double dt = A.sim_override ? A.dt : P.dt;
        const double tol = A.sim_override ? A.tol : P.tol;
        const int com = A.sim_override ? A.com : P.com;
        const double s2i = 1.0 / sqrt(2.0);
        StepLog* logp = (A.log && P.log_off >= 0) ? A.log + P.log_off : nullptr;
        StepLog& lg = sm.lg;
        int step = 0;
        // a resumed segment continues from its stored time, bit for bit the
        // sequence of one uninterrupted call (the warm seed stays in the
        // workspace, the generator state in A.rng)
        double t = resume ? A.seg[ti].t : t0;
        int budget = resume ? A.max_steps : 0x7fffffff;

        for (; t <= t1 + dt / 2.0 && budget > 0; t += dt, --budget) {
            if (tid == 0) { lg = StepLog{}; lg.t = t; }
            // recenter (integrator.cpp:88-93): sequential sums, as the reference
            if (com) {
                sync();
                for (int d = 0; d < 3; ++d) sm.P[d][tid] = R[d];
                sync();
                if (tid < 3) {
                    double s = 0.0;
                    for (int mu = 0; mu < N; ++mu) s = s + sm.P[tid][mu];
                    sm.Scal[8 + tid] = (1.0 / N) * s;
                }
                sync();
                if (own)
                    for (int d = 0; d < 3; ++d) R[d] = R[d] - sm.Scal[8 + d];
            }
            stage_R(R);

            // forces at R (integrator.cpp:70-76), parked in the state slots
            {
                double bond[3] = {0.0, 0.0, 0.0}, Fc[3], F[3];
                if (tid < N - 1)
                    for (int d = 0; d < 3; ++d) bond[d] = sm.R[d][tid + 1] - R[d];
                spring_bead_force(bond, Fc, F);
                save(SLOT_FS, F);
                explicit_force(sm.R, t, F);
                save(SLOT_FE, F);
            }

            // mobility at R; the Cholesky method factors here (mobility.cpp:39-51)
            if (!REGF && hi && P.hi_method == 1) {
                if (!dense_factor()) { status = ST_NOT_PD; break; }
            }

            // B.dW (mobility.cpp:58-88). REGF: the factor follows the draws;
            // a failed factor leaves the generator where the reference's
            // (which factors before drawing) would be.
            const int32_t ix_b = ix, iy_b = iy, rseed_b = rseed;
            const bool ready_b = rng_ready;
            if (!rng_ready) { rng_maybe_init(ix, iy, rseed); rng_ready = true; }
            double bx0[3] = {0.0, 0.0, 0.0};
            {
                double x0[3] = {0.0, 0.0, 0.0};
                int32_t nix = ix, niy = iy;
                if (own) bead_noise(as_gbl(A.jx), as_gbl(A.jy), tid, ix, iy, x0, nix, niy);
                if (tid == N - 1) { sm.ScalI[0] = nix; sm.ScalI[1] = niy; }
                sync();
                ix = sm.ScalI[0];
                iy = sm.ScalI[1];
                if (!hi) {
                    for (int d = 0; d < 3; ++d) bx0[d] = x0[d];
                } else if (P.hi_method == 1) {
                    if constexpr (REGF) {
                        if (!dense_bdw_regs<BLOCK, MIXED, NB>(sm, rl, N, P.hstar, x0, bx0)) {
                            ix = ix_b;
                            iy = iy_b;
                            rseed = rseed_b;
                            rng_ready = ready_b;
                            status = ST_NOT_PD;
                            break;
                        }
                    } else {
                        lower_matvec(x0, bx0);
                    }
                } else if (REGF) {
                    // the register-factor kernel is built for the Cholesky
                    // method alone (its launcher sends it nothing else)
                    status = ST_NOT_RUN;
                    break;
                } else if (P.hi_method == 0) {
                    save(SLOT_X0, x0);
                } else {
                    status = ST_EXACT_SQRT;
                    break;
                }
            }
            if constexpr (!REGF) {
                if (hi && P.hi_method == 0) {
                    if (!chebyshev_bx0(seed_valid, bx0)) { status = ST_NOT_PD; break; }
                }
            }
            {
                const double sqrt_dt = sqrt(dt);
                double dW[3];
                for (int d = 0; d < 3; ++d) dW[d] = sqrt_dt * bx0[d];
                save(SLOT_DW, dW);
            }

            // predictor (integrator.cpp:96-101)
            double Rp[3];
            {
                double Fs[3], Fe[3], v[3], DF[3], dW[3];
                load(SLOT_FS, Fs);
                load(SLOT_FE, Fe);
                for (int d = 0; d < 3; ++d) v[d] = Fs[d] + Fe[d];
                apply(v, DF);
                load(SLOT_DW, dW);
                for (int d = 0; d < 3; ++d) {
                    const int r = 3 * d;
                    double K[3];
                    kappa_row(t, r, K);
                    const double fl0 = K[0] * R[0] + K[1] * R[1] + K[2] * R[2];
                    Rp[d] = own ? R[d] + dt * (fl0 + 0.25 * DF[d]) + s2i * dW[d] : 0.0;
                }
            }

            // upsilon (integrator.cpp:107-119). The reference forms D.F^S and
            // D.(F^E + F~^E) and adds 1/8 of each; D is linear, so one
            // product of the summed forces gives the same term (to rounding)
            // for one mat-vec fewer per step.
            double ups[3];
            {
                sync();
                for (int d = 0; d < 3; ++d) sm.P[d][tid] = Rp[d];
                sync();
                double Fep[3], Fe[3], Fs[3], DF[3], v[3], dW[3];
                explicit_force(sm.P, t + dt, Fep);
                load(SLOT_FE, Fe);
                load(SLOT_FS, Fs);
                for (int d = 0; d < 3; ++d) v[d] = Fs[d] + (Fe[d] + Fep[d]);
                apply(v, DF);
                load(SLOT_DW, dW);
                for (int d = 0; d < 3; ++d) {
                    double K0[3], K1[3];
                    kappa_row(t, 3 * d, K0);
                    kappa_row(t + dt, 3 * d, K1);
                    const double f0 = K0[0] * R[0] + K0[1] * R[1] + K0[2] * R[2];
                    const double f1 = K1[0] * Rp[0] + K1[1] * Rp[1] + K1[2] * Rp[2];
                    const double flow = 0.5 * (f0 + f1);
                    ups[d] = own ? R[d] + dt * (flow + 0.125 * DF[d]) + s2i * dW[d] : 0.0;
                }
            }

            // implicit connector solve (integrator.cpp:125-165)
            {
                double dups[3] = {0.0, 0.0, 0.0}, cb[3] = {0.0, 0.0, 0.0};
                if (tid < N - 1)
                    for (int d = 0; d < 3; ++d) cb[d] = sm.P[d][tid + 1] - Rp[d];
                const double anchor0 = sm.P[0][0], anchor1 = sm.P[1][0], anchor2 = sm.P[2][0];
                sync();
                for (int d = 0; d < 3; ++d) sm.V[d][tid] = ups[d];
                sync();
                if (tid < N - 1)
                    for (int d = 0; d < 3; ++d) dups[d] = sm.V[d][tid + 1] - ups[d];
                double Rprev[3] = {Rp[0], Rp[1], Rp[2]};
                const int max_iter = 100 * N;
                int iter = 0;
                bool fail = false;
                for (; iter <= max_iter; ++iter) {
                    double Fc[3], Fb[3], DFc[3];
                    spring_bead_force(cb, Fc, Fb);
                    apply(Fb, DFc);
                    bump(sm.cnt.corrector_iters);
                    bump32(lg.corrector_iters);
                    sync();
                    for (int d = 0; d < 3; ++d) sm.V[d][tid] = DFc[d];
                    sync();
                    int sst = ST_OK;
                    if (tid < N - 1) {
                        double g[3];
                        for (int d = 0; d < 3; ++d)
                            g[d] = dups[d] + (0.125 * dt) * (sm.V[d][tid + 1] - DFc[d]) + (0.25 * dt) * Fc[d];
                        sst = solve_connector_ool(g[0], g[1], g[2], dt, P.spring_type, P.sqrtb, P.q0, cb);
                    }
                    for (int d = 0; d < 3; ++d) sm.B[d][tid] = cb[d];
                    if (sst != ST_OK) sm.ScalI[2] = sst;
                    if (sync_or(sst != ST_OK)) { status = sm.ScalI[2]; fail = true; break; }
                    // positions from the anchor by the reference's sequential sum
                    if (tid < 3) {
                        double acc = tid == 0 ? anchor0 : (tid == 1 ? anchor1 : anchor2);
                        sm.P[tid][0] = acc;
                        for (int mu = 0; mu < N - 1; ++mu) {
                            acc = acc + sm.B[tid][mu];
                            sm.P[tid][mu + 1] = acc;
                        }
                    }
                    sync();
                    double dv[3] = {0.0, 0.0, 0.0};
                    if (own)
                        for (int d = 0; d < 3; ++d) { R[d] = sm.P[d][tid]; dv[d] = R[d] - Rprev[d]; }
                    const double increment = sqrt(bsum(dot3(dv, dv))) / N;
                    if (increment < tol) break;
                    for (int d = 0; d < 3; ++d) Rprev[d] = R[d];
                }
                if (fail) break;
                if (iter > max_iter) bump(sm.cnt.nonconverged);
            }
            bump(sm.cnt.steps);
            if (logp && step < P.log_cap && tid == 0 && part == 0) logp[step] = lg;
            ++step;
            if constexpr (SPLIT) {
                if (sync_fail) { status = ST_SYNC_TIMEOUT; break; }
            }
        }

        if (part != 0) return;      // SPLIT: part 0 writes the results
        if (own)
            for (int d = 0; d < 3; ++d) A.R[P.R_off + 3 * tid + d] = R[d];
        sync();
        if (tid == 0) {
            A.rng[ti].ix = ix;
            A.rng[ti].iy = iy;
            A.rng[ti].seed = rseed;
            A.status[ti] = status;
            A.cnt[ti] = sm.cnt;
            if (resume) {
                const int32_t done = (status != ST_OK || !(t <= t1 + dt / 2.0)) ? 1 : 0;
                A.seg[ti] = SegState{t, seed_valid ? 1 : 0, done};
                A.done_out[blockIdx.x / nparts] = done;
            }
        }
    }

    // row r..r+2 of kappa(t) (flow.cpp:5-26)
    __device__ __forceinline__ void kappa_row(double t, int r, double K[3]) {
        if (P.flow_const || P.flow_n == 0) {
            K[0] = P.kappa0[r]; K[1] = P.kappa0[r + 1]; K[2] = P.kappa0[r + 2];
            return;
        }
        double Kf[9];
        kappa_at(P, A.blob, t, Kf);
        K[0] = Kf[r]; K[1] = Kf[r + 1]; K[2] = Kf[r + 2];
    }
};

template <int BLOCK, bool MIXED>
__device__ __forceinline__ void init_counters(Smem<BLOCK, MIXED>& sm) {
    if (threadIdx.x == 0) { sm.cnt = Counters{}; sm.lg = StepLog{}; }
    __syncthreads();
}

// WPE: the waves per SIMD the register allocation is held to (1 = unconstrained).
template <int BLOCK, bool MIXED, int WPE>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) k_integrate(KArgs a) {
    __shared__ Smem<BLOCK, MIXED> sm;
    const int ti = a.list[blockIdx.x];
    const TrajParams* __restrict__ P = a.tp + ti;
    if (a.seg && a.seg[ti].done) {           // nothing left of this segment
        if (threadIdx.x == 0) a.done_out[blockIdx.x] = 1;
        return;
    }
    init_counters(sm);
    if (a.seg && threadIdx.x == 0) sm.cnt = a.cnt[ti];   // counts run on across launches
    Chain<BLOCK, MIXED> c(sm, *P, a);
    c.run(ti);
}

// 64-thread chains of the mixed form with the step's RPY pair weights in an
// LDS table (dynamic LDS: N * ((N / 2) | 1) float pairs, N the largest chain
// of the launch): the walk of each mat-vec reads them instead of forming them.
template <int WPE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) k_integrate_ptab(KArgs a) {
    __shared__ Smem<64, true> sm;
    extern __shared__ f32x2 ptab_lds[];
    const int ti = a.list[blockIdx.x];
    const TrajParams* __restrict__ P = a.tp + ti;
    if (a.seg && a.seg[ti].done) {
        if (threadIdx.x == 0) a.done_out[blockIdx.x] = 1;
        return;
    }
    init_counters(sm);
    if (a.seg && threadIdx.x == 0) sm.cnt = a.cnt[ti];
    Chain<64, true, 0, false, double, false, false, true> c(sm, *P, a, nullptr, nullptr, nullptr, nullptr,
                                                          ptab_lds);
    c.run(ti);
}

// 64-thread chains of a small launch with the hot part of the workspace in
// dynamic LDS (at least 8 ws_layout(N, true).dense bytes, N the largest
// chain of the launch; chains of the Cholesky method with the factor in LDS
// are not launched here). PT (mixed form): the step's pair-weight table
// follows the chain's workspace part (N * ((N / 2) | 1) float pairs more).
template <bool MIXED, bool PT = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1))) k_integrate_lws(KArgs a) {
    __shared__ Smem<64, MIXED> sm;
    extern __shared__ double lws[];
    const int ti = a.list[blockIdx.x];
    const TrajParams* __restrict__ P = a.tp + ti;
    if (a.seg && a.seg[ti].done) {
        if (threadIdx.x == 0) a.done_out[blockIdx.x] = 1;
        return;
    }
    init_counters(sm);
    if (a.seg && threadIdx.x == 0) sm.cnt = a.cnt[ti];
    Chain<64, MIXED, 0, false, double, true, false, PT> c(sm, *P, a, nullptr, nullptr, lws, nullptr,
                                                          PT ? lws + ws_layout(P->N, true).dense : nullptr);
    if (a.seg && a.seg[ti].seed_valid) c.ws_to_lds();
    c.run(ti);
    c.ws_to_hbm();
}

// Chains of the Cholesky method with 3N <= 16 NB: the factor in LDS, of
// type FT.
template <int BLOCK, bool MIXED, int NB, typename FT = double>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(1))) k_integrate_dense(KArgs a) {
    __shared__ Smem<BLOCK, MIXED> sm;
    __shared__ DenseLds<NB, FT> dl;
    const int ti = a.list[blockIdx.x];
    const TrajParams* __restrict__ P = a.tp + ti;
    if (a.seg && a.seg[ti].done) {
        if (threadIdx.x == 0) a.done_out[blockIdx.x] = 1;
        return;
    }
    init_counters(sm);
    if (a.seg && threadIdx.x == 0) sm.cnt = a.cnt[ti];
    Chain<BLOCK, MIXED, NB, false, FT> c(sm, *P, a, dl.blk, dl.xs);
    c.run(ti);
}

// Chains of the Cholesky method with 3N <= 16 NB, the FP64 factor in
// registers (dense_bdw_regs), 16-20 KB of LDS per chain (8 chains fit a CU).
// WPE: the waves per SIMD the register allocation is held to.
template <bool MIXED, int NB, int WPE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) k_integrate_dreg(KArgs a) {
    __shared__ Smem<64, MIXED> sm;
    __shared__ RegLds<NB, MIXED> rl;
    const int ti = a.list[blockIdx.x];
    const TrajParams* __restrict__ P = a.tp + ti;
    if (a.seg && a.seg[ti].done) {
        if (threadIdx.x == 0) a.done_out[blockIdx.x] = 1;
        return;
    }
    init_counters(sm);
    if (a.seg && threadIdx.x == 0) sm.cnt = a.cnt[ti];
    // the workspace in LDS is the state slots alone: the Cholesky step uses
    // no other part of it (no Lanczos basis, no Chebyshev coefficients)
    Chain<64, MIXED, NB, false, double, true, true> c(sm, *P, a, nullptr, nullptr, rl.slots, &rl);
    c.run(ti);
}

// Chains of N > 512 split over a.split workgroups each (few chains, many
// CUs). Launched cooperatively, so that all parts are resident together.
// Resumable like k_integrate (a.seg): every part reads the chain's segment
// state at its start (before the first exchange, which part 0 must pass before
// it writes the state back) and part 0 reports the chain's done flag.
template <bool MIXED>
__global__ void __launch_bounds__(1024) k_integrate_split(KArgs a) {
    __shared__ Smem<1024, MIXED> sm;
    const int G = a.split;
    const int ti = a.list[blockIdx.x / G];
    const TrajParams* __restrict__ P = a.tp + ti;
    if (a.seg && a.seg[ti].done) {
        if (threadIdx.x == 0 && blockIdx.x % G == 0) a.done_out[blockIdx.x / G] = 1;
        return;
    }
    init_counters(sm);
    if (a.seg && threadIdx.x == 0) sm.cnt = a.cnt[ti];
    Chain<1024, MIXED, 0, true> c(sm, *P, a);
    c.set_part(blockIdx.x % G, G);
    c.run(ti);
}

// ---- force / noise / mobility kernels for the piecewise entry points -------
enum : int { FORCE_TOTAL = 0, FORCE_SPRING = 1, FORCE_BENDING = 2, FORCE_EV = 3 };

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_forces(KArgs a, int which, double* out) {
    __shared__ Smem<BLOCK, false> sm;
    const int ti = blockIdx.x;
    const TrajParams* __restrict__ P = a.tp;
    init_counters(sm);
    Chain<BLOCK, false> c(sm, *P, a);
    const int N = P->N, tid = threadIdx.x;
    const bool own = tid < N;
    const int64_t off = (int64_t)ti * 3 * N;
    double R[3] = {0.0, 0.0, 0.0};
    if (own)
        for (int d = 0; d < 3; ++d) R[d] = a.R[off + 3 * tid + d];
    for (int d = 0; d < 3; ++d) sm.R[d][tid] = R[d];
    __syncthreads();
    double bond[3] = {0.0, 0.0, 0.0};
    if (tid < N - 1)
        for (int d = 0; d < 3; ++d) bond[d] = sm.R[d][tid + 1] - R[d];
    double Fc[3], Fs[3], Fev[3], Fb[3], F[3];
    c.spring_bead_force(bond, Fc, Fs);
    c.ev_force(sm.R, Fev);
    c.bending_force(sm.R, Fb);
    for (int d = 0; d < 3; ++d) {
        if (which == FORCE_SPRING) F[d] = Fs[d];
        else if (which == FORCE_BENDING) F[d] = Fb[d];
        else if (which == FORCE_EV) F[d] = Fev[d];
        else F[d] = Fs[d] + (Fev[d] + Fb[d]);   // bindings/module.cpp:67-76
    }
    if (own)
        for (int d = 0; d < 3; ++d) out[off + 3 * tid + d] = F[d];
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_noise(KArgs a, int N, int n_steps, double* out) {
    __shared__ int32_t s[2];
    const int ti = blockIdx.x, tid = threadIdx.x;
    int32_t ix = a.rng[ti].ix, iy = a.rng[ti].iy, seed = a.rng[ti].seed;
    if (n_steps > 0) rng_maybe_init(ix, iy, seed);
    for (int st = 0; st < n_steps; ++st) {
        double X[3];
        int32_t nix = ix, niy = iy;
        if (tid < N) {
            bead_noise(as_gbl(a.jx), as_gbl(a.jy), tid, ix, iy, X, nix, niy);
            for (int d = 0; d < 3; ++d)
                out[((int64_t)ti * n_steps + st) * 3 * N + 3 * tid + d] = X[d];
        }
        if (tid == N - 1) { s[0] = nix; s[1] = niy; }
        __syncthreads();
        ix = s[0];
        iy = s[1];
        __syncthreads();
    }
    if (tid == 0) { a.rng[ti].ix = ix; a.rng[ti].iy = iy; a.rng[ti].seed = seed; }
}

template <int BLOCK, bool MIXED>
__global__ void __launch_bounds__(BLOCK) k_rpy_apply(KArgs a, const double* x_in, double* out) {
    __shared__ Smem<BLOCK, MIXED> sm;
    const int ti = blockIdx.x;
    const TrajParams* __restrict__ P = a.tp;
    init_counters(sm);
    Chain<BLOCK, MIXED> c(sm, *P, a);
    const int N = P->N, tid = threadIdx.x;
    const int64_t off = (int64_t)ti * 3 * N;
    double x[3] = {0.0, 0.0, 0.0}, R[3] = {0.0, 0.0, 0.0};
    if (tid < N)
        for (int d = 0; d < 3; ++d) { R[d] = a.R[off + 3 * tid + d]; x[d] = x_in[off + 3 * tid + d]; }
    c.stage_R(R);
    double y[3];
    c.apply(x, y);
    if (tid < N)
        for (int d = 0; d < 3; ++d) out[off + 3 * tid + d] = y[d];
}

// B x for many chains of one physics (positions and x back to back), each
// with its own workspace.
template <int BLOCK, bool MIXED>
__global__ void __launch_bounds__(BLOCK) k_brownian(KArgs a, const double* x_in, double* out,
                                                    int32_t* status) {
    __shared__ Smem<BLOCK, MIXED> sm;
    const int ti = blockIdx.x;
    const TrajParams* __restrict__ P = a.tp;
    init_counters(sm);
    Chain<BLOCK, MIXED> c(sm, *P, a);
    c.ws = c.wsg = as_gbl(a.ws + (int64_t)ti * c.wl.total);
    const int N = P->N, tid = threadIdx.x;
    const int64_t off = (int64_t)ti * 3 * N;
    double R[3] = {0.0, 0.0, 0.0}, x[3] = {0.0, 0.0, 0.0};
    if (tid < N)
        for (int d = 0; d < 3; ++d) { R[d] = a.R[off + 3 * tid + d]; x[d] = x_in[off + 3 * tid + d]; }
    c.stage_R(R);
    double y[3];
    const int st = c.brownian(x, y);
    if (tid < N)
        for (int d = 0; d < 3; ++d) out[off + 3 * tid + d] = y[d];
    if (tid == 0) status[ti] = st;
}

// ---- snapshot observables (python/bdsim/properties.py:5-52) -----------------
// One record per chain, the layout of bdsim_observables.
struct Obs {
    double rg2, q2;
    double gyration[9];
    double stretch[3];
    double stress[9];
    double com[3];
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide maximum; shares block_sum's alternating halves of sRed.
template <int NW>
__device__ __forceinline__ double block_max(double v, double* sRed, int& par) {
    v = wave_max(v);
    if constexpr (NW == 1) {
        return v;
    } else {
        double* buf = sRed + par * NW;
        par ^= 1;
        if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
        __syncthreads();
        double t = buf[0];
#pragma unroll
        for (int i = 1; i < NW; ++i) t = fmax(t, buf[i]);
        return t;
    }
}

// Centre of mass, gyration tensor and its trace, squared end-to-end distance,
// extent of the centred chain, and the Kramers-Kirkwood stress
// tau_ij = sum_mu (R_mu - R_c)_i F_mu,j with F the total intramolecular force
// (spring + EV + bending, bindings/module.cpp:67-76) or one of its parts.
// per_chain: each chain has its own TrajParams (positions at its R_off);
// otherwise all chains share a.tp[0] and sit back to back.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_observe(KArgs a, int per_chain, int which, Obs* out) {
    constexpr int NW = BLOCK / 64;
    __shared__ Smem<BLOCK, false> sm;
    const int ti = blockIdx.x;
    const TrajParams* __restrict__ P = per_chain ? a.tp + ti : a.tp;
    init_counters(sm);
    Chain<BLOCK, false> c(sm, *P, a);
    const int N = P->N, tid = threadIdx.x;
    const bool own = tid < N;
    const int64_t off = per_chain ? P->R_off : (int64_t)ti * 3 * N;
    double R[3] = {0.0, 0.0, 0.0};
    if (own)
        for (int d = 0; d < 3; ++d) R[d] = a.R[off + 3 * tid + d];
    for (int d = 0; d < 3; ++d) sm.R[d][tid] = R[d];
    __syncthreads();
    double bond[3] = {0.0, 0.0, 0.0};
    if (tid < N - 1)
        for (int d = 0; d < 3; ++d) bond[d] = sm.R[d][tid + 1] - R[d];
    double Fc[3], Fs[3], Fev[3], Fb[3], F[3];
    c.spring_bead_force(bond, Fc, Fs);
    c.ev_force(sm.R, Fev);
    c.bending_force(sm.R, Fb);
    for (int d = 0; d < 3; ++d) {
        if (which == FORCE_SPRING) F[d] = Fs[d];
        else if (which == FORCE_BENDING) F[d] = Fb[d];
        else if (which == FORCE_EV) F[d] = Fev[d];
        else F[d] = Fs[d] + (Fev[d] + Fb[d]);
        if (!own) F[d] = 0.0;
    }
    int par = 0;
    Obs o;
    for (int d = 0; d < 3; ++d) o.com[d] = block_sum<NW>(R[d], sm.Red, par) / N;
    double rc[3];
    for (int d = 0; d < 3; ++d) rc[d] = own ? R[d] - o.com[d] : 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
            const double gij = block_sum<NW>(rc[i] * rc[j], sm.Red, par) / N;
            o.gyration[3 * i + j] = gij;
            o.gyration[3 * j + i] = gij;
        }
    o.rg2 = o.gyration[0] + o.gyration[4] + o.gyration[8];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.stress[3 * i + j] = block_sum<NW>(rc[i] * F[j], sm.Red, par);
    for (int d = 0; d < 3; ++d) {
        const double hi = block_max<NW>(own ? rc[d] : -INFINITY, sm.Red, par);
        const double lo = -block_max<NW>(own ? -rc[d] : -INFINITY, sm.Red, par);
        o.stretch[d] = hi - lo;
    }
    double e[3];
    for (int d = 0; d < 3; ++d) e[d] = sm.R[d][N - 1] - sm.R[d][0];
    o.q2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
    if (tid == 0) out[ti] = o;
}

}  // namespace bdgpu
