// bdsim_kernels.hip -- host side of the MI355X (gfx950) ChainIntegrator library
// and the C ABI of include/bdsim_gpu.h. The device code is in bdsim_chain.hpp;
// the integrate kernels are instantiated, per precision form and chain size,
// in the k_integrate_*.hip files (built in parallel).
//
// Citations are to IsaacPincus/bdsim.
#include "bdsim_chain.hpp"
#include "bdsim_launch.hpp"
#include "../../include/bdsim_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>


// =========================================================================
// host side
// =========================================================================
namespace {

using namespace bdgpu;

thread_local std::string g_err;
std::mutex g_mu;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(BDSIM_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        if (hipMalloc(&p, want) != hipSuccess) return fail(BDSIM_ENOMEM, "hipMalloc failed");
        cap = want;
        return 0;
    }
};

// kernel buckets of bdsim_integrate_batch: blocks 64..1024, then the two
// LDS-factor Cholesky kernels
constexpr int N_BUCKETS = 7;

// The device memory of bdsim_integrate_batch_static: one static array per
// device, 256 MiB, held from the library's load until the process ends. In
// host code `scratch` only names the symbol; its device address on each GPU
// comes from hipGetSymbolAddress at initialisation (g_scratch).
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)256 << 20;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];
void* g_scratch[MAX_DEVICES] = {};

struct Ctx {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipStream_t bstream[N_BUCKETS] = {};    // one per kernel bucket
    hipEvent_t bevent[N_BUCKETS] = {}, uploaded = nullptr;
    uint32_t* jx = nullptr;
    uint32_t* jy = nullptr;
    DevBuf tp, list, R, ws, blob, ext, rng, status, cnt, log, obs, bar, xbuf;
    int coop = 0, n_cu = 256;       // cooperative launch supported; compute units
} g;

// Columns of T^(3 mu) for the xorshift (13,17,5) map and 16807^(3 mu) mod M.
void build_jump_tables(std::vector<uint32_t>& jx, std::vector<uint32_t>& jy) {
    auto step = [](uint32_t x) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; return x; };
    jx.assign((size_t)MAX_BEADS * 32, 0);
    jy.assign(MAX_BEADS, 0);
    std::vector<uint32_t> cur(32);   // columns of T^k, k = 3 mu
    for (int i = 0; i < 32; ++i) cur[i] = 1u << i;
    uint64_t a = 1;
    const uint64_t M = 2147483647ull;
    for (int mu = 0; mu < MAX_BEADS; ++mu) {
        for (int i = 0; i < 32; ++i) jx[(size_t)mu * 32 + i] = cur[i];
        jy[mu] = (uint32_t)a;
        for (int s = 0; s < 3; ++s) {
            for (int i = 0; i < 32; ++i) cur[i] = step(cur[i]);   // T is linear: T(col) per column
            a = (a * 16807ull) % M;
        }
    }
}

int ensure_init() {
    if (g.ready) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(BDSIM_ENODEV, "no HIP device visible");
    if (g.device < 0) g.device = 0;
    HIP_TRY(hipSetDevice(g.device));
    HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    (void)hipDeviceGetAttribute(&g.coop, hipDeviceAttributeCooperativeLaunch, g.device);
    (void)hipDeviceGetAttribute(&g.n_cu, hipDeviceAttributeMultiprocessorCount, g.device);
    if (g.n_cu <= 0) g.n_cu = 256;
    for (int b = 0; b < N_BUCKETS; ++b) {
        HIP_TRY(hipStreamCreateWithFlags(&g.bstream[b], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&g.bevent[b], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&g.uploaded, hipEventDisableTiming));
    if (g.device < MAX_DEVICES && !g_scratch[g.device]) {
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch[g.device] = ptr;
    }
    std::vector<uint32_t> jx, jy;
    build_jump_tables(jx, jy);
    HIP_TRY(hipMalloc(&g.jx, jx.size() * sizeof(uint32_t)));
    HIP_TRY(hipMalloc(&g.jy, jy.size() * sizeof(uint32_t)));
    HIP_TRY(hipMemcpy(g.jx, jx.data(), jx.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(g.jy, jy.data(), jy.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    g.ready = true;
    return 0;
}

int block_for(int N) {
    if (N <= 64) return 64;
    if (N <= 128) return 128;
    if (N <= 256) return 256;
    if (N <= 512) return 512;
    return 1024;
}

// h* and the folded RPY pair-weight constants (zero without HI).
void set_rpy(TrajParams& t, double hstar) {
    t.hstar = hstar;
    double k[6] = {0, 0, 0, 0, 0, 0};
    if (hstar > 0.0) rpy_folded(hstar, k);
    for (int j = 0; j < 6; ++j) { t.rpy_d[j] = k[j]; t.rpy_f[j] = (float)k[j]; }
}

bool valid_precision(int p) {
    return p == BDSIM_PRECISION_FP64 || p == BDSIM_PRECISION_MIXED || p == BDSIM_PRECISION_MIXED_F32_FACTOR;
}
bool is_mixed(int p) { return p == BDSIM_PRECISION_MIXED || p == BDSIM_PRECISION_MIXED_F32_FACTOR; }

// Steps one segment may take in one kernel run at most (the step loop runs on
// the device, so an unbounded clock would never let the launch end).
constexpr double MAX_SEGMENT_STEPS = 1.0e9;

// A segment clock the device loop `for (t = t0; t <= t1 + dt/2; t += dt)`
// finishes: dt positive and finite, finite ends, every t += dt within the
// segment moving t by at least dt/2, and a bounded number of steps.
int check_clock(double t0, double t1, double dt) {
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(BDSIM_EINVAL, "dt must be positive and finite");
    if (!std::isfinite(t0) || !std::isfinite(t1)) return fail(BDSIM_EINVAL, "time_start and time_end must be finite");
    const double m = std::max(std::fabs(t0), std::fabs(t1)) + dt;
    if (!((m + dt) - m >= 0.5 * dt)) return fail(BDSIM_EINVAL, "dt is below the resolution of the segment's clock");
    if ((t1 - t0) / dt > MAX_SEGMENT_STEPS) return fail(BDSIM_EINVAL, "segment longer than 1e9 steps");
    return 0;
}

// Host packing of one trajectory (EV constants exactly as
// excluded_volume.cpp:14-34, evaluated with the host libm).
int pack_traj(const bdsim_phys& p, const bdsim_sim& s, int N, std::vector<double>& blob,
              std::vector<ExtEntry>& ext, TrajParams& t) {
    std::memset(&t, 0, sizeof t);
    if (int rc = check_clock(s.time_start, s.time_end, s.dt)) return rc;
    if (p.spring_type < 1 || p.spring_type > 7) return fail(BDSIM_EINVAL, "bad spring type");
    if (p.ev_type < 0 || p.ev_type > 4) return fail(BDSIM_EINVAL, "bad EV type");
    if (p.hi_method < 0 || p.hi_method > 2) return fail(BDSIM_EINVAL, "bad hi_method");
    t.N = N;
    t.spring_type = p.spring_type;
    t.sqrtb = p.sqrtb;
    t.natural_length = p.natural_length;
    t.q0 = p.natural_length / p.sqrtb;
    t.ev_type = p.ev_type;
    t.cd = p.contour_dist_for_ev;
    t.equil = p.equilibration ? 1 : 0;
    const double z = p.zstar, d = p.dstar;
    t.zsbyds5 = z / std::pow(d, 5.0);
    t.pt5bydssq = 0.5 / (d * d);
    t.LJa = 12.0 * 4.0 * z * std::pow(d, 12.0);
    t.LJb = 6.0 * 4.0 * z * std::pow(d, 6.0);
    t.SDKa = 12.0 * 4.0 * std::pow(d, 12.0);
    t.SDKb = 6.0 * 4.0 * std::pow(d, 6.0);
    t.Rmin = p.min_cutoff;
    t.Rcutp = p.max_cutoff;
    t.RcutpSDK = p.max_cutoff;
    t.Rcutg = d * std::pow(2.0, 1.0 / 6.0);
    set_rpy(t, p.hstar);
    t.Rmin2 = t.Rmin * t.Rmin;
    t.Rcutg2 = t.Rcutg * t.Rcutg;
    t.Rcutp2 = t.Rcutp * t.Rcutp;
    t.RcutpSDK2 = t.RcutpSDK * t.RcutpSDK;
    t.has_phi = 0;
    t.phi_off = 0;
    if (p.ev_type == BDSIM_EV_SDK && p.n_phi > 0) {
        if (p.n_phi != N * N || !p.phi) return fail(BDSIM_EINVAL, "phi must hold N*N entries");
        t.has_phi = 1;
        t.phi_off = (int64_t)blob.size();
        blob.insert(blob.end(), p.phi, p.phi + p.n_phi);
    }
    t.bend_type = p.bend_type;
    t.bend_C = p.bend_stiffness;
    t.flow_const = (p.flow.constant || p.flow.n_times == 0) ? 1 : 0;
    for (int k = 0; k < 9; ++k) t.kappa0[k] = p.flow.kappa0[k];
    t.flow_n = t.flow_const ? 0 : p.flow.n_times;
    t.flow_off = (int64_t)blob.size();
    if (!t.flow_const) {
        blob.insert(blob.end(), p.flow.times, p.flow.times + p.flow.n_times);
        blob.insert(blob.end(), p.flow.kappas, p.flow.kappas + 9 * p.flow.n_times);
    }
    t.n_ext = p.n_external;
    t.ext_off = (int64_t)ext.size();
    for (int k = 0; k < p.n_external; ++k) {
        const bdsim_bead_force& e = p.external[k];
        ExtEntry x{};
        x.bead = e.bead;
        x.constant = e.constant ? 1 : 0;
        x.n_times = e.constant ? 0 : e.n_times;
        for (int q = 0; q < 3; ++q) x.value[q] = e.value[q];
        x.off = (int64_t)blob.size();
        if (!e.constant && e.n_times > 0) {
            blob.insert(blob.end(), e.times, e.times + e.n_times);
            blob.insert(blob.end(), e.values, e.values + 3 * e.n_times);
        }
        ext.push_back(x);
    }
    t.hi_method = p.hi_method;
    t.mult = p.ncheb_multiplier;
    t.fd_tol = p.fd_err_max;
    t.t0 = s.time_start;
    t.t1 = s.time_end;
    t.dt = s.dt;
    t.tol = s.implicit_loop_tol;
    t.com = s.update_center_of_mass ? 1 : 0;
    t.log_off = -1;
    t.log_cap = 0;
    return 0;
}

// Register budget per block size set by the caller (bdsim_tune_set): waves
// per SIMD, 0 = the library's own choice. Index: 64, 128, 256, 512.
std::atomic<int> g_tune[4] = {{0}, {0}, {0}, {0}};
int tune_index(int block) {
    switch (block) {
        case 64: return 0;
        case 128: return 1;
        case 256: return 2;
        case 512: return 3;
        default: return -1;
    }
}

long long device_simds() {
    static const int simds = [] {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        return 4 * (cus > 0 ? cus : 256);
    }();
    return simds;
}

// More waves than 3 per SIMD of the device: the launch would not be resident
// at once with the unconstrained register allocation.
bool crowded(long long waves) { return waves > 3LL * device_simds(); }

// Register budget of the integrate kernel for 64-thread workgroups: waves per
// SIMD the allocation is held to -- the largest that keeps the whole launch
// resident: 1 (no spills) for up to one wave per SIMD of the device, 2 for up
// to two, else 4 (the fastest measured at 4096 x N=50). BDSIM_WPE=1|2|4 in the
// environment fixes it.
int wpe_choice(long long waves) {
    static const int fixed = [] {
        const char* e = std::getenv("BDSIM_WPE");
        const int w = e ? std::atoi(e) : 0;
        return (w == 1 || w == 2 || w == 4) ? w : 0;
    }();
    if (fixed) return fixed;
    const long long simds = device_simds();
    return waves <= simds ? 1 : waves <= 2 * simds ? 2 : 4;
}

// Cholesky chains whose factor fits the LDS of a one-wave workgroup: the
// DenseLds size (in 16 x 16 blocks per side) of their kernel, or 0.
// BDSIM_DENSE_LDS=0 in the environment keeps every factor in the workspace.
int dense_nb_for(int N, double hstar, int hi_method) {
    static const bool enabled = [] {
        const char* e = std::getenv("BDSIM_DENSE_LDS");
        return !(e && e[0] == '0');
    }();
    if (!enabled || !(hstar > 0.0) || hi_method != 1) return 0;
    if (3 * N <= 64) return 4;
    if (3 * N <= 96) return 6;
    return 0;
}

// Dynamic LDS for the hot workspace of k_integrate_lws when a launch of
// 64-thread chains (largest N = nmax) is small enough for it: every chain
// resident at once at one wave per SIMD with its workspace in LDS, and the
// launch the only one running (waves = its own). 0: keep the workspace in HBM.
// A workgroup's static and dynamic LDS together stay within 64 KB (the
// dynamic size a launch may ask for without a kernel attribute).
constexpr size_t LWS_MAX_TOTAL = 64 * 1024;
constexpr size_t LDS_PER_CU = 160 * 1024;
size_t ptab_bytes_for(int nmax, bool with_lws = false, unsigned n_wg = 0);
template <bool MIXED>
size_t lws_bytes_for(int nmax, unsigned n_wg, long long waves) {
    static const bool enabled = [] {
        const char* e = std::getenv("BDSIM_LDS_WORKSPACE");
        return !(e && e[0] == '0');
    }();
    if (!enabled || nmax <= 0 || waves != (long long)n_wg) return 0;
    // the mixed form's pair-weight table, when used, follows the workspace part
    const size_t dyn = sizeof(double) * (size_t)ws_layout(nmax, true).dense +
                       (MIXED ? ptab_bytes_for(nmax, true) : 0);
    if (sizeof(Smem<64, MIXED>) + dyn > LWS_MAX_TOTAL) return 0;
    const size_t per_cu = std::min<size_t>(4, LDS_PER_CU / (sizeof(Smem<64, MIXED>) + dyn));
    return (size_t)n_wg <= (size_t)g.n_cu * per_cu ? sizeof(double) * (size_t)ws_layout(nmax, true).dense : 0;
}

// Dynamic LDS for the pair-weight table of k_integrate_ptab (mixed form,
// 64-thread chains, largest N = nmax): N * ((N / 2) | 1) float pairs, used
// when the table and the chain's static LDS together leave room for 8 chains
// per compute unit (N <= 51), which the two-waves-per-SIMD register budget
// then holds, or at any N <= 64 when the launch's n_wg chains are resident at
// once even so (the table then costs no occupancy the launch would use); 0:
// the walk forms the weights itself. BDSIM_PAIR_TABLE=0 in the environment
// turns the table off, =2 uses it at every N <= 64.
// with_lws: the table of a small launch that keeps the workspace in LDS too
// (k_integrate_lws, at most 4 chains per compute unit): used at every N.
constexpr size_t PTAB_LDS_MAX = LDS_PER_CU / 8;
size_t ptab_bytes_for(int nmax, bool with_lws, unsigned n_wg) {
    static const int mode = [] {
        const char* e = std::getenv("BDSIM_PAIR_TABLE");
        return e ? std::atoi(e) : 1;
    }();
    if (mode == 0 || nmax < 2 || nmax > 64) return 0;
    const size_t bytes = sizeof(f32x2) * (size_t)nmax * (size_t)((nmax >> 1) | 1);
    const size_t per_wg = sizeof(Smem<64, true>) + bytes;
    if (with_lws || mode == 2 || per_wg <= PTAB_LDS_MAX) return bytes;
    // fewer chains per compute unit than 8: only when the launch fits anyway
    return (size_t)n_wg <= (size_t)g.n_cu * (LDS_PER_CU / per_wg) ? bytes : 0;
}

// Workgroup width per block size set by the caller (bdsim_tune_width): the
// chains of that block run on workgroups of block x m threads (at most 1024),
// 0 = the library's choice. Index as g_tune.
std::atomic<int> g_width[4] = {{0}, {0}, {0}, {0}};

// Workgroup width (threads) for the mixed form's multi-wave chains of a grid of
// `chains` chains of at most nmax beads (0: block's own), launched with
// `waves` waves in all: wider while the wider workgroup shortens the longest
// wave's share of the tiled pair walk (T^2 <= NW (T - 1), T tiles, NW >= 8
// waves: the walk spreads over all waves) and the whole launch still holds no
// more than one wave per SIMD of the device (a latency-bound launch: the extra
// waves land on idle SIMDs). BDSIM_WIDE=0 in the environment keeps block.
int width_choice(int block, int nmax, unsigned chains, long long waves) {
    static const bool enabled = [] {
        const char* e = std::getenv("BDSIM_WIDE");
        return !(e && e[0] == '0');
    }();
    const int ti = tune_index(block);
    const int m = ti >= 0 ? g_width[ti].load() : 0;
    if (m > 0) return std::min(1024, block * m);
    if (!enabled || block < 128) return block;
    const int T = ((nmax > 0 ? nmax : block) + 63) / 64;
    int wide = block;
    while (wide < 1024) {
        const int nw = 2 * wide / 64;
        const long long more = (long long)chains * (wide / 64);
        if (!(nw >= 8 && T * T <= nw * (T - 1)) || waves + more > device_simds()) break;
        waves += more;
        wide *= 2;
    }
    return wide;
}

// waves: the waves of everything launched together with this grid (buckets
// side by side), or -1 for this grid alone.
// f32_factor: the mixed form's variant with the LDS Cholesky factor in FP32.
// nmax: the largest chain of the grid (0: unknown).
template <bool MIXED>
void launch_integrate(int block, dim3 grid, hipStream_t stream, const KArgs& a, int nb = 0,
                      long long waves = -1, bool f32_factor = false, int nmax = 0) {
    if (waves < 0) waves = (long long)grid.x * (block / 64);
    if (nb == 4 || nb == 6) {
        launch_integrate_small<MIXED>(1, nb, f32_factor, grid, stream, a);
        return;
    }
    const int ti = tune_index(block);
    const int tuned = ti >= 0 ? g_tune[ti].load() : 0;
    if (block == 64) {
        const size_t lws = tuned ? 0 : lws_bytes_for<MIXED>(nmax, grid.x, waves);
        const size_t pt = (MIXED && !tuned) ? ptab_bytes_for(nmax, lws > 0, grid.x) : 0;
        const int wpe = tuned ? tuned : wpe_choice(waves);
        // with the table at most 8 chains share a compute unit: the register
        // budget of 2 waves per SIMD (1 for a launch of at most one per SIMD)
        launch_integrate_small<MIXED>(pt ? std::min(wpe, 2) : wpe, 0, false, grid, stream, a, lws, pt);
        return;
    }
    // Multi-wave chains: the register budget of 4 waves per SIMD when the
    // launch has more waves than 3 per SIMD could hold (all chains then stay
    // resident; measured 2x on 4096 chains of N in [10, 256]), otherwise the
    // unconstrained build; 1024-thread chains always the latter. The mixed
    // form may run them on wider workgroups (width_choice).
    const int wide = MIXED ? width_choice(block, nmax, grid.x, waves) : block;
    waves += (long long)grid.x * ((wide - block) / 64);
    const bool four = tuned ? tuned == 4 : crowded(waves);
    launch_integrate_large<MIXED>(wide, four && wide < 1024 ? 4 : 1, grid, stream, a);
}

// Largest number of workgroups per split chain set by the caller
// (bdsim_tune_split): 0 = the library's choice (8; every part holds its own
// workspace copy, 9 N^2 doubles of it the dense region).
std::atomic<int> g_split_max{0};

// Workgroups per chain for a batch whose only chains are `count` chains of
// N > 512 (1 = no split): as many as fill the compute units, at most 8 (or
// the tuned maximum, up to 16). BDSIM_SPLIT=0 in the environment turns the
// split off.
int split_factor(int count, int total) {
    static const bool enabled = [] {
        const char* e = std::getenv("BDSIM_SPLIT");
        return !(e && e[0] == '0');
    }();
    if (!enabled || !g.coop || count == 0 || count != total) return 1;
    const int cap = g_split_max.load();
    const int G = std::min(cap > 0 ? cap : 8, g.n_cu / count);
    return G >= 2 ? G : 1;
}

// The split launch: count chains x G parts, cooperative (all resident); if
// the runtime refuses a cooperative launch of that size, the chains run
// whole, one workgroup each (their first workspace copy).
// bar: count counters; xbuf: count * split_xbuf_doubles(G) doubles.
size_t split_xbuf_doubles(int G) { return (size_t)2 * G * 3 * 1024; }

// Zeroes n words (used for the split chains' barrier counters instead of a
// memset, so that a launch sequence holding it can be captured in a graph).
__global__ void k_zero_u32(uint32_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Opens a resumable segment at t0 on n trajectories: cold spectral cache,
// nothing done yet.
__global__ void k_seg_open(SegState* seg, int n, double t0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seg[i] = SegState{t0, 0, 0};
}

template <bool MIXED>
int launch_split(int count, int G, hipStream_t stream, KArgs a, uint32_t* bar, double* xbuf) {
    hipLaunchKernelGGL(k_zero_u32, dim3((count + 255) / 256), dim3(256), 0, stream, bar, count);
    HIP_TRY(hipGetLastError());
    a.split = G;
    a.bar = bar;
    a.xbuf = xbuf;
    const hipError_t e = launch_split_kernel<MIXED>(dim3(count * G), stream, a);
    if (e == hipSuccess) return 0;
    (void)hipGetLastError();
    a.split = 1;
    launch_integrate<MIXED>(1024, dim3(count), stream, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <bool MIXED>
int launch_bucket(int block, int count, hipStream_t stream, const KArgs& a, int nb, long long waves,
                  bool f32_factor, int nmax) {
    if (count == 0) return 0;
    launch_integrate<MIXED>(block, dim3(count), stream, a, nb, waves, f32_factor, nmax);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <typename T>
T* as(DevBuf& b) { return reinterpret_cast<T*>(b.p); }

// Buffer sizes (bytes) of one bdsim_integrate_batch call.
struct BatchSizes {
    size_t tp = 0, list = 0, R = 0, ws = 0, blob = 0, ext = 0, rng = 0, status = 0, cnt = 0,
           log = 0, bar = 0, xbuf = 0;
};

// Device buffers of one call (wherever they were carved from).
struct BatchBufs {
    TrajParams* tp = nullptr;
    int32_t* list = nullptr;
    double* R = nullptr;
    double* ws = nullptr;
    double* blob = nullptr;
    ExtEntry* ext = nullptr;
    RngState* rng = nullptr;
    int32_t* status = nullptr;
    Counters* cnt = nullptr;
    StepLog* log = nullptr;
    uint32_t* bar = nullptr;
    double* xbuf = nullptr;
};

// Host side of one call: packed parameters, the launch order by bucket, the
// split factor and the sizes of the device buffers.
struct HostBatch {
    int n_traj = 0;
    std::vector<TrajParams> tps;
    std::vector<double> blob;
    std::vector<ExtEntry> ext;
    std::vector<int32_t> order;
    int counts[N_BUCKETS] = {};
    int nmax[N_BUCKETS] = {};     // largest chain of each bucket
    int split = 1;
    int64_t roff = 0;
    size_t nlog = 1;
    BatchSizes sz;
};

constexpr int BUCKET_BLOCK[N_BUCKETS] = {64, 128, 256, 512, 1024, 64, 64};
constexpr int BUCKET_DNB[N_BUCKETS] = {0, 0, 0, 0, 0, 4, 6};

size_t static_bytes_needed(const BatchSizes& z) {
    size_t t = 0;
    for (size_t b : {z.tp, z.list, z.R, z.ws, z.blob, z.ext, z.rng, z.status, z.cnt, z.log, z.bar, z.xbuf})
        t += (b + 255) & ~(size_t)255;
    return t;
}

// Checks the arguments and packs the call on the host (no device work).
int plan_batch(int n_traj, const int32_t* nbeads, const double* R, const bdsim_phys* phys,
               const bdsim_sim* sim, const bdsim_rng_state* rng, int precision,
               const int32_t* status, const bdsim_step_log* log, int log_stride, HostBatch& hb) {
    if (n_traj < 0 || !nbeads || !R || !phys || !sim || !rng || !status)
        return fail(BDSIM_EINVAL, "null argument");
    if (!valid_precision(precision))
        return fail(BDSIM_EINVAL, "bad precision");
    hb.n_traj = n_traj;
    if (n_traj == 0) return 0;
    hb.tps.resize(n_traj);
    for (int i = 0; i < n_traj; ++i) {
        const int N = nbeads[i];
        if (N < 2 || N > MAX_BEADS) return fail(BDSIM_EINVAL, "nbeads out of range [2, max_beads]");
        if (int rc = pack_traj(phys[i], sim[i], N, hb.blob, hb.ext, hb.tps[i])) return rc;
        hb.tps[i].R_off = hb.roff;
        hb.roff += 3 * N;
        if (log && log_stride > 0) {
            hb.tps[i].log_off = (int64_t)i * log_stride;
            hb.tps[i].log_cap = log_stride;
        }
    }
    if (hb.blob.empty()) hb.blob.push_back(0.0);
    if (hb.ext.empty()) hb.ext.push_back(ExtEntry{});
    // Buckets by kernel (block size; Cholesky chains with the factor in LDS
    // apart), each launched on its own stream so that buckets overlap on the
    // device; within a bucket the longest chains go first (workgroups are
    // dispatched in order, so the short ones fill in behind).
    std::vector<int> bucket(n_traj);
    for (int i = 0; i < n_traj; ++i) {
        const int nb = dense_nb_for(nbeads[i], phys[i].hstar, phys[i].hi_method);
        int b = 0;
        while (!(BUCKET_BLOCK[b] == block_for(nbeads[i]) && BUCKET_DNB[b] == nb)) ++b;
        bucket[i] = b;
    }
    for (int b = 0; b < N_BUCKETS; ++b) {
        const size_t first = hb.order.size();
        for (int i = 0; i < n_traj; ++i)
            if (bucket[i] == b) {
                hb.order.push_back(i);
                ++hb.counts[b];
                hb.nmax[b] = std::max(hb.nmax[b], (int)nbeads[i]);
            }
        std::stable_sort(hb.order.begin() + first, hb.order.end(),
                         [&](int32_t x, int32_t y) { return nbeads[x] > nbeads[y]; });
    }
    // Chains of N > 512 alone in the batch and too few to fill the device:
    // each is split over G workgroups (one per CU), with G workspaces.
    hb.split = split_factor(hb.counts[4], (int)hb.order.size());
    int64_t wsoff = 0;
    for (int i = 0; i < n_traj; ++i) {
        hb.tps[i].ws_off = wsoff;
        const int copies = (bucket[i] == 4 && hb.split > 1) ? hb.split : 1;
        wsoff += copies * ws_layout(nbeads[i], phys[i].hstar > 0.0).total;
    }
    hb.nlog = (log && log_stride > 0) ? (size_t)n_traj * log_stride : 1;
    BatchSizes& z = hb.sz;
    z.tp = sizeof(TrajParams) * n_traj;
    z.list = sizeof(int32_t) * n_traj;
    z.R = sizeof(double) * hb.roff;
    z.ws = sizeof(double) * std::max<int64_t>(wsoff, 1);
    z.blob = sizeof(double) * hb.blob.size();
    z.ext = sizeof(ExtEntry) * hb.ext.size();
    z.rng = sizeof(RngState) * n_traj;
    z.status = sizeof(int32_t) * n_traj;
    z.cnt = sizeof(Counters) * n_traj;
    z.log = sizeof(StepLog) * hb.nlog;
    if (hb.split > 1) {
        z.bar = sizeof(uint32_t) * hb.counts[4];
        z.xbuf = sizeof(double) * hb.counts[4] * split_xbuf_doubles(hb.split);
    }
    return 0;
}

// Uploads, launches every bucket, copies the results back and waits.
int run_batch(const HostBatch& hb, const BatchBufs& d, double* R, bdsim_rng_state* rng, int precision,
              int32_t* status, bdsim_counters* counters, bdsim_step_log* log, int log_stride) {
    static_assert(sizeof(RngState) == sizeof(bdsim_rng_state), "rng layout");
    static_assert(sizeof(Counters) == sizeof(bdsim_counters), "counter layout");
    static_assert(sizeof(StepLog) == sizeof(bdsim_step_log), "log layout");
    const int n_traj = hb.n_traj;
    HIP_TRY(hipMemcpyAsync(d.tp, hb.tps.data(), hb.sz.tp, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(d.list, hb.order.data(), hb.sz.list, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(d.R, R, hb.sz.R, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(d.blob, hb.blob.data(), hb.sz.blob, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(d.ext, hb.ext.data(), hb.sz.ext, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(d.rng, rng, hb.sz.rng, hipMemcpyHostToDevice, g.stream));
    if (log && log_stride > 0) {
        // entries past a trajectory's last step read back as zeros, not as
        // what an earlier call left in the buffer
        const int words = (int)(hb.sz.log / sizeof(uint32_t));
        hipLaunchKernelGGL(k_zero_u32, dim3((words + 255) / 256), dim3(256), 0, g.stream,
                           reinterpret_cast<uint32_t*>(d.log), words);
        HIP_TRY(hipGetLastError());
    }

    KArgs a{};
    a.tp = d.tp;
    a.R = d.R;
    a.ws = d.ws;
    a.blob = d.blob;
    a.ext = d.ext;
    a.rng = d.rng;
    a.status = d.status;
    a.cnt = d.cnt;
    a.log = (log && log_stride > 0) ? d.log : nullptr;
    a.jx = g.jx;
    a.jy = g.jy;
    a.sim_override = 0;
    int starts[N_BUCKETS];
    long long waves = 0;
    for (int b = 0, st = 0; b < N_BUCKETS; ++b) {
        starts[b] = st;
        st += hb.counts[b];
        waves += (long long)hb.counts[b] * (BUCKET_BLOCK[b] / 64);
    }
    HIP_TRY(hipEventRecord(g.uploaded, g.stream));
    for (int b = N_BUCKETS - 1; b >= 0; --b) {    // dense and largest blocks first
        if (hb.counts[b] == 0) continue;
        HIP_TRY(hipStreamWaitEvent(g.bstream[b], g.uploaded, 0));
        a.list = d.list + starts[b];
        int rc;
        if (b == 4 && hb.split > 1) {
            rc = is_mixed(precision)
                     ? launch_split<true>(hb.counts[b], hb.split, g.bstream[b], a, d.bar, d.xbuf)
                     : launch_split<false>(hb.counts[b], hb.split, g.bstream[b], a, d.bar, d.xbuf);
        } else {
            rc = is_mixed(precision)
                     ? launch_bucket<true>(BUCKET_BLOCK[b], hb.counts[b], g.bstream[b], a, BUCKET_DNB[b], waves,
                                           precision == BDSIM_PRECISION_MIXED_F32_FACTOR, hb.nmax[b])
                     : launch_bucket<false>(BUCKET_BLOCK[b], hb.counts[b], g.bstream[b], a, BUCKET_DNB[b], waves,
                                            false, hb.nmax[b]);
        }
        if (rc) {
            (void)hipDeviceSynchronize();     // buckets already queued must end before reuse
            return rc;
        }
        HIP_TRY(hipEventRecord(g.bevent[b], g.bstream[b]));
        HIP_TRY(hipStreamWaitEvent(g.stream, g.bevent[b], 0));
    }
    HIP_TRY(hipMemcpyAsync(R, d.R, hb.sz.R, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(rng, d.rng, hb.sz.rng, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(status, d.status, hb.sz.status, hipMemcpyDeviceToHost, g.stream));
    if (counters) HIP_TRY(hipMemcpyAsync(counters, d.cnt, hb.sz.cnt, hipMemcpyDeviceToHost, g.stream));
    if (log && log_stride > 0) HIP_TRY(hipMemcpyAsync(log, d.log, hb.sz.log, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    (void)n_traj;
    return 0;
}

}  // namespace

struct bdsim_batch {
    int n_traj = 0, N = 0, precision = 0, block = 64, dense_nb = 0, split = 1;
    hipStream_t stream = nullptr;
    TrajParams* tp = nullptr;
    int32_t* list = nullptr;
    double* R = nullptr;
    double* ws = nullptr;
    double* blob = nullptr;
    ExtEntry* ext = nullptr;
    RngState* rng = nullptr;
    int32_t* status = nullptr;
    Counters* cnt = nullptr;
    void* obs = nullptr;            // observables records, allocated on first use
    uint32_t* bar = nullptr;        // split chains: barrier counters, exchange buffers
    double* xbuf = nullptr;
    SegState* seg = nullptr;        // resumable runs (bdsim_batch_run_steps), on first use
    int32_t* done = nullptr;
    bdsim_sim seg_sim{};            // the segment those runs are cutting
    bool seg_open = false;
};

extern "C" {

int bdsim_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g.ready && device != g.device) return fail(BDSIM_EINVAL, "already initialised on another device");
    g.device = device;
    return ensure_init();
}

int bdsim_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g.ready) return 0;
    int rc = 0;
    if (hipSetDevice(g.device) != hipSuccess) rc = fail(BDSIM_EHIP, "hipSetDevice failed at shutdown");
    // everything queued by this library (all its streams) ends first
    if (hipDeviceSynchronize() != hipSuccess) rc = fail(BDSIM_EHIP, "hipDeviceSynchronize failed at shutdown");
    DevBuf* bufs[] = {&g.tp, &g.list, &g.R, &g.ws, &g.blob, &g.ext, &g.rng, &g.status,
                      &g.cnt, &g.log, &g.obs, &g.bar, &g.xbuf};
    for (DevBuf* b : bufs) {
        if (b->p) (void)hipFree(b->p);
        b->p = nullptr;
        b->cap = 0;
    }
    if (g.jx) (void)hipFree(g.jx);
    if (g.jy) (void)hipFree(g.jy);
    g.jx = nullptr;
    g.jy = nullptr;
    for (int b = 0; b < N_BUCKETS; ++b) {
        if (g.bstream[b]) (void)hipStreamDestroy(g.bstream[b]);
        if (g.bevent[b]) (void)hipEventDestroy(g.bevent[b]);
        g.bstream[b] = nullptr;
        g.bevent[b] = nullptr;
    }
    if (g.uploaded) (void)hipEventDestroy(g.uploaded);
    if (g.stream) (void)hipStreamDestroy(g.stream);
    g.uploaded = nullptr;
    g.stream = nullptr;
    g.ready = false;
    return rc;
}

const char* bdsim_last_error(void) { return g_err.c_str(); }

int bdsim_max_beads(void) { return MAX_BEADS; }

int bdsim_tune_set(int block, int waves_per_simd) {
    const int i = tune_index(block);
    if (i < 0) return fail(BDSIM_EINVAL, "tune: block must be 64, 128, 256 or 512");
    const bool ok = waves_per_simd == 0 || waves_per_simd == 1 || waves_per_simd == 4 ||
                    (block == 64 && waves_per_simd == 2);
    if (!ok) return fail(BDSIM_EINVAL, "tune: waves per SIMD must be 0, 1 or 4 (or 2 for block 64)");
    g_tune[i].store(waves_per_simd);
    return 0;
}

int bdsim_tune_get(int block) {
    const int i = tune_index(block);
    return i < 0 ? -1 : g_tune[i].load();
}

int bdsim_tune_width(int block, int multiplier) {
    const int i = tune_index(block);
    if (i < 1) return fail(BDSIM_EINVAL, "tune: width is set for blocks 128, 256 or 512");
    if (!(multiplier == 0 || multiplier == 1 || multiplier == 2 || multiplier == 4 || multiplier == 8))
        return fail(BDSIM_EINVAL, "tune: width multiplier must be 0 (automatic), 1, 2, 4 or 8");
    g_width[i].store(multiplier);
    return 0;
}

int bdsim_tune_width_get(int block) {
    const int i = tune_index(block);
    return i < 1 ? -1 : g_width[i].load();
}

int bdsim_tune_split(int max_workgroups) {
    if (max_workgroups < 0 || max_workgroups > 16)
        return fail(BDSIM_EINVAL, "tune: workgroups per chain must be 0 (automatic) or 1..16");
    g_split_max.store(max_workgroups);
    return 0;
}

int bdsim_tune_split_get(void) { return g_split_max.load(); }

int bdsim_integrate_batch(int n_traj, const int32_t* nbeads, double* R, const bdsim_phys* phys,
                          const bdsim_sim* sim, bdsim_rng_state* rng, int precision,
                          int32_t* status, bdsim_counters* counters, bdsim_step_log* log,
                          int log_stride) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init()) return rc;
    HostBatch hb;
    if (int rc = plan_batch(n_traj, nbeads, R, phys, sim, rng, precision, status, log, log_stride, hb))
        return rc;
    if (n_traj == 0) return 0;
    BatchBufs d;
    if (int rc = g.tp.ensure(hb.sz.tp)) return rc;
    if (int rc = g.list.ensure(hb.sz.list)) return rc;
    if (int rc = g.R.ensure(hb.sz.R)) return rc;
    if (int rc = g.ws.ensure(hb.sz.ws)) return rc;
    if (int rc = g.blob.ensure(hb.sz.blob)) return rc;
    if (int rc = g.ext.ensure(hb.sz.ext)) return rc;
    if (int rc = g.rng.ensure(hb.sz.rng)) return rc;
    if (int rc = g.status.ensure(hb.sz.status)) return rc;
    if (int rc = g.cnt.ensure(hb.sz.cnt)) return rc;
    if (int rc = g.log.ensure(hb.sz.log)) return rc;
    if (hb.sz.bar) {
        if (int rc = g.bar.ensure(hb.sz.bar)) return rc;
        if (int rc = g.xbuf.ensure(hb.sz.xbuf)) return rc;
    }
    d.tp = as<TrajParams>(g.tp);
    d.list = as<int32_t>(g.list);
    d.R = as<double>(g.R);
    d.ws = as<double>(g.ws);
    d.blob = as<double>(g.blob);
    d.ext = as<ExtEntry>(g.ext);
    d.rng = as<RngState>(g.rng);
    d.status = as<int32_t>(g.status);
    d.cnt = as<Counters>(g.cnt);
    d.log = as<StepLog>(g.log);
    d.bar = hb.sz.bar ? as<uint32_t>(g.bar) : nullptr;
    d.xbuf = hb.sz.bar ? as<double>(g.xbuf) : nullptr;
    return run_batch(hb, d, R, rng, precision, status, counters, log, log_stride);
}

int bdsim_integrate_batch_static(int n_traj, const int32_t* nbeads, double* R,
                                 const bdsim_phys* phys, const bdsim_sim* sim,
                                 bdsim_rng_state* rng, int precision, int32_t* status,
                                 bdsim_counters* counters, bdsim_step_log* log, int log_stride) {
    std::lock_guard<std::mutex> lk(g_mu);
    // no initialisation here: it allocates (streams, jump tables)
    if (!g.ready) return fail(BDSIM_EINVAL, "bdsim_integrate_batch_static: call bdsim_init first");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != g.device) HIP_TRY(hipSetDevice(g.device));
    if (g.device < 0 || g.device >= MAX_DEVICES || !g_scratch[g.device])
        return fail(BDSIM_EINVAL, "bdsim_integrate_batch_static: no scratch address for this device");
    HostBatch hb;
    if (int rc = plan_batch(n_traj, nbeads, R, phys, sim, rng, precision, status, log, log_stride, hb))
        return rc;
    if (n_traj == 0) return 0;
    // every buffer of the call carved from the static array, 256-byte aligned
    unsigned char* base = static_cast<unsigned char*>(g_scratch[g.device]);
    size_t off = 0;
    bool fits = true;
    auto carve = [&](size_t bytes) -> void* {
        if (bytes == 0) return nullptr;
        off = (off + 255) & ~(size_t)255;
        if (bytes > STATIC_SCRATCH_BYTES || off > STATIC_SCRATCH_BYTES - bytes) { fits = false; return nullptr; }
        void* p = base + off;
        off += bytes;
        return p;
    };
    BatchBufs d;
    d.tp = static_cast<TrajParams*>(carve(hb.sz.tp));
    d.list = static_cast<int32_t*>(carve(hb.sz.list));
    d.R = static_cast<double*>(carve(hb.sz.R));
    d.ws = static_cast<double*>(carve(hb.sz.ws));
    d.blob = static_cast<double*>(carve(hb.sz.blob));
    d.ext = static_cast<ExtEntry*>(carve(hb.sz.ext));
    d.rng = static_cast<RngState*>(carve(hb.sz.rng));
    d.status = static_cast<int32_t*>(carve(hb.sz.status));
    d.cnt = static_cast<Counters*>(carve(hb.sz.cnt));
    d.log = static_cast<StepLog*>(carve(hb.sz.log));
    d.bar = static_cast<uint32_t*>(carve(hb.sz.bar));
    d.xbuf = static_cast<double*>(carve(hb.sz.xbuf));
    if (!fits)
        return fail(BDSIM_ENOMEM, "bdsim_integrate_batch_static: the batch needs " +
                                      std::to_string(static_bytes_needed(hb.sz)) + " bytes of device memory, more than the " +
                                      std::to_string(STATIC_SCRATCH_BYTES) + "-byte static scratch");
    return run_batch(hb, d, R, rng, precision, status, counters, log, log_stride);
}

size_t bdsim_static_scratch_bytes(void) { return STATIC_SCRATCH_BYTES; }

// ---- device-resident batches ---------------------------------------------

int bdsim_batch_create(bdsim_batch** out, int n_traj, int nbeads, const double* R0,
                       const bdsim_phys* phys, const double* kappa, const bdsim_rng_state* rng,
                       int precision) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init()) return rc;
    if (!out || n_traj <= 0 || !R0 || !phys || !rng) return fail(BDSIM_EINVAL, "bad argument");
    if (nbeads < 2 || nbeads > MAX_BEADS) return fail(BDSIM_EINVAL, "nbeads out of range");
    if (!valid_precision(precision))
        return fail(BDSIM_EINVAL, "bad precision");
    const int N = nbeads;
    bdsim_sim s0{0.0, 0.0, 0.01, 1e-3, 1};
    std::vector<double> blob;
    std::vector<ExtEntry> ext;
    TrajParams proto;
    if (int rc = pack_traj(*phys, s0, N, blob, ext, proto)) return rc;
    std::vector<TrajParams> tps(n_traj, proto);
    // chains of N > 512 too few to fill the device run split (G workspaces)
    const int split = block_for(N) == 1024 ? split_factor(n_traj, n_traj) : 1;
    const int64_t wsper = split * ws_layout(N, phys->hstar > 0.0).total;
    std::vector<int32_t> list(n_traj);
    for (int i = 0; i < n_traj; ++i) {
        tps[i].R_off = (int64_t)i * 3 * N;
        tps[i].ws_off = (int64_t)i * wsper;
        if (kappa) {
            tps[i].flow_const = 1;
            tps[i].flow_n = 0;
            for (int k = 0; k < 9; ++k) tps[i].kappa0[k] = kappa[9 * i + k];
        }
        list[i] = i;
    }
    if (blob.empty()) blob.push_back(0.0);
    if (ext.empty()) ext.push_back(ExtEntry{});
    bdsim_batch* b = new bdsim_batch();
    b->n_traj = n_traj;
    b->N = N;
    b->precision = precision;
    b->block = block_for(N);
    b->dense_nb = dense_nb_for(N, phys->hstar, phys->hi_method);
    b->split = split;
    auto cleanup = [&](int code, const char* what) {
        bdsim_batch_destroy(b);
        return fail(code, what);
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(BDSIM_EHIP, "stream create failed");
    if (hipMalloc(&b->tp, sizeof(TrajParams) * n_traj) != hipSuccess ||
        hipMalloc(&b->list, sizeof(int32_t) * n_traj) != hipSuccess ||
        hipMalloc(&b->R, sizeof(double) * 3 * N * (size_t)n_traj) != hipSuccess ||
        hipMalloc(&b->ws, sizeof(double) * std::max<size_t>(1, (size_t)wsper * n_traj)) != hipSuccess ||
        hipMalloc(&b->blob, sizeof(double) * blob.size()) != hipSuccess ||
        hipMalloc(&b->ext, sizeof(ExtEntry) * ext.size()) != hipSuccess ||
        hipMalloc(&b->rng, sizeof(RngState) * n_traj) != hipSuccess ||
        hipMalloc(&b->status, sizeof(int32_t) * n_traj) != hipSuccess ||
        hipMalloc(&b->cnt, sizeof(Counters) * n_traj) != hipSuccess)
        return cleanup(BDSIM_ENOMEM, "hipMalloc failed for batch");
    if (split > 1 &&
        (hipMalloc(&b->bar, sizeof(uint32_t) * n_traj) != hipSuccess ||
         hipMalloc(&b->xbuf, sizeof(double) * n_traj * split_xbuf_doubles(split)) != hipSuccess))
        return cleanup(BDSIM_ENOMEM, "hipMalloc failed for batch");
    if (hipMemcpy(b->tp, tps.data(), sizeof(TrajParams) * n_traj, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->list, list.data(), sizeof(int32_t) * n_traj, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->R, R0, sizeof(double) * 3 * N * (size_t)n_traj, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->blob, blob.data(), sizeof(double) * blob.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->ext, ext.data(), sizeof(ExtEntry) * ext.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->rng, rng, sizeof(RngState) * n_traj, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(BDSIM_EHIP, "upload failed for batch");
    *out = b;
    return 0;
}

int bdsim_batch_run(bdsim_batch* b, const bdsim_sim* sim) {
    if (!b || !sim) return fail(BDSIM_EINVAL, "null argument");
    if (int rc = check_clock(sim->time_start, sim->time_end, sim->dt)) return rc;
    b->seg_open = false;            // a whole segment ends any one being cut into runs
    KArgs a{};
    a.tp = b->tp;
    a.list = b->list;
    a.R = b->R;
    a.ws = b->ws;
    a.blob = b->blob;
    a.ext = b->ext;
    a.rng = b->rng;
    a.status = b->status;
    a.cnt = b->cnt;
    a.log = nullptr;
    a.jx = g.jx;
    a.jy = g.jy;
    a.sim_override = 1;
    a.com = sim->update_center_of_mass ? 1 : 0;
    a.t0 = sim->time_start;
    a.t1 = sim->time_end;
    a.dt = sim->dt;
    a.tol = sim->implicit_loop_tol;
    if (b->split > 1)
        return is_mixed(b->precision)
                   ? launch_split<true>(b->n_traj, b->split, b->stream, a, b->bar, b->xbuf)
                   : launch_split<false>(b->n_traj, b->split, b->stream, a, b->bar, b->xbuf);
    if (is_mixed(b->precision))
        launch_integrate<true>(b->block, dim3(b->n_traj), b->stream, a, b->dense_nb, -1,
                               b->precision == BDSIM_PRECISION_MIXED_F32_FACTOR, b->N);
    else launch_integrate<false>(b->block, dim3(b->n_traj), b->stream, a, b->dense_nb, -1, false, b->N);
    HIP_TRY(hipGetLastError());
    return 0;
}

int bdsim_batch_run_steps(bdsim_batch* b, const bdsim_sim* sim, int max_steps, int resume) {
    if (!b || !sim) return fail(BDSIM_EINVAL, "null argument");
    if (max_steps <= 0) return fail(BDSIM_EINVAL, "max_steps must be positive");
    if (int rc = check_clock(sim->time_start, sim->time_end, sim->dt)) return rc;
    if (resume) {
        if (!b->seg_open) return fail(BDSIM_EINVAL, "no segment to resume");
        const bdsim_sim& o = b->seg_sim;
        if (sim->time_start != o.time_start || sim->time_end != o.time_end || sim->dt != o.dt ||
            sim->implicit_loop_tol != o.implicit_loop_tol ||
            (sim->update_center_of_mass != 0) != (o.update_center_of_mass != 0))
            return fail(BDSIM_EINVAL, "resume with the segment's own bdsim_sim");
    }
    if (!b->seg || !b->done) {
        if (!b->seg && hipMalloc(&b->seg, sizeof(SegState) * b->n_traj) != hipSuccess) b->seg = nullptr;
        if (!b->done && hipMalloc(&b->done, sizeof(int32_t) * b->n_traj) != hipSuccess) b->done = nullptr;
        if (!b->seg || !b->done) return fail(BDSIM_ENOMEM, "hipMalloc failed for the segment state");
    }
    const int n = b->n_traj;
    if (!resume) {
        // counters run on across the launches of one segment: start them at zero
        static_assert(sizeof(Counters) % 4 == 0, "Counters is zeroed as 32-bit words");
        const int words = (int)(sizeof(Counters) / 4) * n;
        hipLaunchKernelGGL(k_zero_u32, dim3((words + 255) / 256), dim3(256), 0, b->stream,
                           reinterpret_cast<uint32_t*>(b->cnt), words);
        hipLaunchKernelGGL(k_seg_open, dim3((n + 255) / 256), dim3(256), 0, b->stream, b->seg, n,
                           sim->time_start);
        HIP_TRY(hipGetLastError());
        b->seg_sim = *sim;
        b->seg_open = true;
    }
    KArgs a{};
    a.tp = b->tp;
    a.list = b->list;
    a.R = b->R;
    a.ws = b->ws;
    a.blob = b->blob;
    a.ext = b->ext;
    a.rng = b->rng;
    a.status = b->status;
    a.cnt = b->cnt;
    a.log = nullptr;
    a.jx = g.jx;
    a.jy = g.jy;
    a.sim_override = 1;
    a.com = sim->update_center_of_mass ? 1 : 0;
    a.t0 = sim->time_start;
    a.t1 = sim->time_end;
    a.dt = sim->dt;
    a.tol = sim->implicit_loop_tol;
    a.seg = b->seg;
    a.done_out = b->done;
    a.max_steps = max_steps;
    if (b->split > 1)
        return is_mixed(b->precision)
                   ? launch_split<true>(n, b->split, b->stream, a, b->bar, b->xbuf)
                   : launch_split<false>(n, b->split, b->stream, a, b->bar, b->xbuf);
    if (is_mixed(b->precision))
        launch_integrate<true>(b->block, dim3(n), b->stream, a, b->dense_nb, -1,
                               b->precision == BDSIM_PRECISION_MIXED_F32_FACTOR, b->N);
    else launch_integrate<false>(b->block, dim3(n), b->stream, a, b->dense_nb, -1, false, b->N);
    HIP_TRY(hipGetLastError());
    return 0;
}

int bdsim_batch_sync(bdsim_batch* b) {
    if (!b) return fail(BDSIM_EINVAL, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int bdsim_batch_fetch(bdsim_batch* b, double* R, bdsim_rng_state* rng, int32_t* status,
                      bdsim_counters* counters) {
    if (!b) return fail(BDSIM_EINVAL, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (R) HIP_TRY(hipMemcpy(R, b->R, sizeof(double) * 3 * b->N * (size_t)b->n_traj, hipMemcpyDeviceToHost));
    if (rng) HIP_TRY(hipMemcpy(rng, b->rng, sizeof(RngState) * b->n_traj, hipMemcpyDeviceToHost));
    if (status) HIP_TRY(hipMemcpy(status, b->status, sizeof(int32_t) * b->n_traj, hipMemcpyDeviceToHost));
    if (counters) HIP_TRY(hipMemcpy(counters, b->cnt, sizeof(Counters) * b->n_traj, hipMemcpyDeviceToHost));
    return 0;
}

void* bdsim_batch_stream(bdsim_batch* b) { return b ? (void*)b->stream : nullptr; }

void bdsim_batch_destroy(bdsim_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    (void)hipFree(b->tp); (void)hipFree(b->list); (void)hipFree(b->R); (void)hipFree(b->ws);
    (void)hipFree(b->blob); (void)hipFree(b->ext); (void)hipFree(b->rng);
    (void)hipFree(b->status); (void)hipFree(b->cnt); (void)hipFree(b->obs);
    (void)hipFree(b->bar); (void)hipFree(b->xbuf);
    (void)hipFree(b->seg); (void)hipFree(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// ---- piecewise entry points ------------------------------------------------

// Uploads n_traj chains of one length sharing one PhysParams (positions back
// to back) and fills the kernel arguments for them. Called with g_mu held.
static int upload_shared(int n_traj, int nbeads, const double* R, const bdsim_phys* phys,
                         KArgs& a) {
    if (int rc = ensure_init()) return rc;
    if (n_traj <= 0 || nbeads < 2 || nbeads > MAX_BEADS || !R || !phys)
        return fail(BDSIM_EINVAL, "bad argument");
    std::vector<double> blob;
    std::vector<ExtEntry> ext;
    TrajParams tp;
    bdsim_sim s0{0.0, 0.0, 0.01, 1e-3, 1};
    if (int rc = pack_traj(*phys, s0, nbeads, blob, ext, tp)) return rc;
    if (blob.empty()) blob.push_back(0.0);
    if (ext.empty()) ext.push_back(ExtEntry{});
    const size_t nd = (size_t)n_traj * 3 * nbeads;
    if (int rc = g.tp.ensure(sizeof(TrajParams))) return rc;
    if (int rc = g.R.ensure(sizeof(double) * nd)) return rc;
    if (int rc = g.ws.ensure(sizeof(double) * nd)) return rc;
    if (int rc = g.blob.ensure(sizeof(double) * blob.size())) return rc;
    if (int rc = g.ext.ensure(sizeof(ExtEntry) * ext.size())) return rc;
    HIP_TRY(hipMemcpyAsync(g.tp.p, &tp, sizeof tp, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.R.p, R, sizeof(double) * nd, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.blob.p, blob.data(), sizeof(double) * blob.size(), hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.ext.p, ext.data(), sizeof(ExtEntry) * ext.size(), hipMemcpyHostToDevice, g.stream));
    a = KArgs{};
    a.tp = as<TrajParams>(g.tp);
    a.R = as<double>(g.R);
    a.ws = as<double>(g.ws);
    a.blob = as<double>(g.blob);
    a.ext = as<ExtEntry>(g.ext);
    return 0;
}

static int forces_common(int n_traj, int nbeads, const double* R, const bdsim_phys* phys,
                         double* F, int which) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!F) return fail(BDSIM_EINVAL, "bad argument");
    KArgs a;
    if (int rc = upload_shared(n_traj, nbeads, R, phys, a)) return rc;
    const size_t nd = (size_t)n_traj * 3 * nbeads;
    dim3 grid(n_traj);
    switch (block_for(nbeads)) {
        case 64: hipLaunchKernelGGL(k_forces<64>, grid, dim3(64), 0, g.stream, a, which, a.ws); break;
        case 128: hipLaunchKernelGGL(k_forces<128>, grid, dim3(128), 0, g.stream, a, which, a.ws); break;
        case 256: hipLaunchKernelGGL(k_forces<256>, grid, dim3(256), 0, g.stream, a, which, a.ws); break;
        case 512: hipLaunchKernelGGL(k_forces<512>, grid, dim3(512), 0, g.stream, a, which, a.ws); break;
        default: hipLaunchKernelGGL(k_forces<1024>, grid, dim3(1024), 0, g.stream, a, which, a.ws); break;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(F, g.ws.p, sizeof(double) * nd, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

static int launch_observe(int block, int n_traj, hipStream_t stream, const KArgs& a, int per_chain,
                          int which, Obs* out) {
    const dim3 grid(n_traj);
    switch (block) {
        case 64: hipLaunchKernelGGL(k_observe<64>, grid, dim3(64), 0, stream, a, per_chain, which, out); break;
        case 128: hipLaunchKernelGGL(k_observe<128>, grid, dim3(128), 0, stream, a, per_chain, which, out); break;
        case 256: hipLaunchKernelGGL(k_observe<256>, grid, dim3(256), 0, stream, a, per_chain, which, out); break;
        case 512: hipLaunchKernelGGL(k_observe<512>, grid, dim3(512), 0, stream, a, per_chain, which, out); break;
        default: hipLaunchKernelGGL(k_observe<1024>, grid, dim3(1024), 0, stream, a, per_chain, which, out); break;
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

static_assert(sizeof(Obs) == sizeof(bdsim_observables), "observables record layout");

}  // extern "C"

template <bool MIXED>
void launch_brownian(int block, int n, hipStream_t st, const KArgs& a, const double* x, double* y,
                     int32_t* status) {
    const dim3 grid(n);
    switch (block) {
        case 64: hipLaunchKernelGGL((k_brownian<64, MIXED>), grid, dim3(64), 0, st, a, x, y, status); break;
        case 128: hipLaunchKernelGGL((k_brownian<128, MIXED>), grid, dim3(128), 0, st, a, x, y, status); break;
        case 256: hipLaunchKernelGGL((k_brownian<256, MIXED>), grid, dim3(256), 0, st, a, x, y, status); break;
        case 512: hipLaunchKernelGGL((k_brownian<512, MIXED>), grid, dim3(512), 0, st, a, x, y, status); break;
        default: hipLaunchKernelGGL((k_brownian<1024, MIXED>), grid, dim3(1024), 0, st, a, x, y, status); break;
    }
}

extern "C" {

int bdsim_brownian_batch(int n_traj, int nbeads, const double* R, const bdsim_phys* phys,
                         int precision, const double* x, double* y, int32_t* status) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!x || !y || !status || (!valid_precision(precision)))
        return fail(BDSIM_EINVAL, "bad argument");
    KArgs a;
    if (int rc = upload_shared(n_traj, nbeads, R, phys, a)) return rc;
    const size_t nd = (size_t)n_traj * 3 * nbeads;
    const int64_t wsper = ws_layout(nbeads, phys->hstar > 0.0).total;
    if (int rc = g.ws.ensure(sizeof(double) * std::max<size_t>(1, (size_t)wsper * n_traj))) return rc;
    if (int rc = g.obs.ensure(sizeof(double) * 2 * nd + sizeof(int32_t) * n_traj)) return rc;
    a.ws = as<double>(g.ws);
    double* dx = as<double>(g.obs);
    double* dy = dx + nd;
    int32_t* dst = reinterpret_cast<int32_t*>(dy + nd);
    HIP_TRY(hipMemcpyAsync(dx, x, sizeof(double) * nd, hipMemcpyHostToDevice, g.stream));
    if (is_mixed(precision)) launch_brownian<true>(block_for(nbeads), n_traj, g.stream, a, dx, dy, dst);
    else launch_brownian<false>(block_for(nbeads), n_traj, g.stream, a, dx, dy, dst);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, dy, sizeof(double) * nd, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(status, dst, sizeof(int32_t) * n_traj, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

int bdsim_observables_batch(int n_traj, int nbeads, const double* R, const bdsim_phys* phys,
                            int force_kind, bdsim_observables* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!out || force_kind < BDSIM_FORCE_TOTAL || force_kind > BDSIM_FORCE_EV)
        return fail(BDSIM_EINVAL, "bad argument");
    KArgs a;
    if (int rc = upload_shared(n_traj, nbeads, R, phys, a)) return rc;
    if (int rc = g.obs.ensure(sizeof(Obs) * n_traj)) return rc;
    if (int rc = launch_observe(block_for(nbeads), n_traj, g.stream, a, 0, force_kind, as<Obs>(g.obs)))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, g.obs.p, sizeof(Obs) * n_traj, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

int bdsim_batch_observe(bdsim_batch* b, int force_kind, bdsim_observables* out) {
    if (!b || !out || force_kind < BDSIM_FORCE_TOTAL || force_kind > BDSIM_FORCE_EV)
        return fail(BDSIM_EINVAL, "bad argument");
    if (!b->obs && hipMalloc(&b->obs, sizeof(Obs) * b->n_traj) != hipSuccess) {
        b->obs = nullptr;
        return fail(BDSIM_ENOMEM, "hipMalloc failed for observables");
    }
    KArgs a{};
    a.tp = b->tp;
    a.R = b->R;
    a.ws = b->ws;
    a.blob = b->blob;
    a.ext = b->ext;
    if (int rc = launch_observe(b->block, b->n_traj, b->stream, a, 1, force_kind, static_cast<Obs*>(b->obs)))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, b->obs, sizeof(Obs) * b->n_traj, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int bdsim_total_force_batch(int n, int N, const double* R, const bdsim_phys* p, double* F) {
    return forces_common(n, N, R, p, F, FORCE_TOTAL);
}
int bdsim_spring_force_batch(int n, int N, const double* R, const bdsim_phys* p, double* F) {
    return forces_common(n, N, R, p, F, FORCE_SPRING);
}
int bdsim_bending_force_batch(int n, int N, const double* R, const bdsim_phys* p, double* F) {
    return forces_common(n, N, R, p, F, FORCE_BENDING);
}
int bdsim_ev_force_batch(int n, int N, const double* R, const bdsim_phys* p, double* F) {
    return forces_common(n, N, R, p, F, FORCE_EV);
}

// External forces are a handful of table lookups; they are evaluated by the
// same device function the step uses, on one workgroup.
namespace {
__global__ void k_external(KArgs a, double t) {
    const TrajParams P = a.tp[0];
    const int mu = threadIdx.x + blockIdx.x * blockDim.x;
    if (mu >= P.N) return;
    double F[3];
    external_at(P, a.ext, a.blob, P.N, mu, t, F);
    for (int d = 0; d < 3; ++d) a.ws[3 * mu + d] = F[d];
}
}  // namespace

int bdsim_external_force_batch(const bdsim_phys* phys, int nbeads, double t, double* F) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init()) return rc;
    if (!phys || !F || nbeads < 1 || nbeads > (1 << 24)) return fail(BDSIM_EINVAL, "bad argument");
    std::vector<double> blob;
    std::vector<ExtEntry> ext;
    TrajParams tp;
    bdsim_sim s0{0.0, 0.0, 0.01, 1e-3, 1};
    bdsim_phys p = *phys;
    p.n_phi = 0;   // phi is irrelevant to the external force
    if (int rc = pack_traj(p, s0, nbeads, blob, ext, tp)) return rc;
    if (blob.empty()) blob.push_back(0.0);
    if (ext.empty()) ext.push_back(ExtEntry{});
    if (int rc = g.tp.ensure(sizeof(TrajParams))) return rc;
    if (int rc = g.ws.ensure(sizeof(double) * 3 * nbeads)) return rc;
    if (int rc = g.blob.ensure(sizeof(double) * blob.size())) return rc;
    if (int rc = g.ext.ensure(sizeof(ExtEntry) * ext.size())) return rc;
    HIP_TRY(hipMemcpyAsync(g.tp.p, &tp, sizeof tp, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.blob.p, blob.data(), sizeof(double) * blob.size(), hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.ext.p, ext.data(), sizeof(ExtEntry) * ext.size(), hipMemcpyHostToDevice, g.stream));
    KArgs a{};
    a.tp = as<TrajParams>(g.tp);
    a.ws = as<double>(g.ws);
    a.blob = as<double>(g.blob);
    a.ext = as<ExtEntry>(g.ext);
    hipLaunchKernelGGL(k_external, dim3((nbeads + 255) / 256), dim3(256), 0, g.stream, a, t);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(F, g.ws.p, sizeof(double) * 3 * nbeads, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

int bdsim_raw_noise_batch(int n_traj, int nbeads, int n_steps, bdsim_rng_state* rng, double* X) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init()) return rc;
    if (n_traj <= 0 || nbeads < 1 || nbeads > MAX_BEADS || n_steps < 0 || !rng || !X)
        return fail(BDSIM_EINVAL, "bad argument");
    const size_t nd = (size_t)n_traj * n_steps * 3 * nbeads;
    if (int rc = g.ws.ensure(sizeof(double) * std::max<size_t>(nd, 1))) return rc;
    if (int rc = g.rng.ensure(sizeof(RngState) * n_traj)) return rc;
    HIP_TRY(hipMemcpyAsync(g.rng.p, rng, sizeof(RngState) * n_traj, hipMemcpyHostToDevice, g.stream));
    KArgs a{};
    a.ws = as<double>(g.ws);
    a.rng = as<RngState>(g.rng);
    a.jx = g.jx;
    a.jy = g.jy;
    dim3 grid(n_traj);
    switch (block_for(nbeads)) {
        case 64: hipLaunchKernelGGL(k_noise<64>, grid, dim3(64), 0, g.stream, a, nbeads, n_steps, a.ws); break;
        case 128: hipLaunchKernelGGL(k_noise<128>, grid, dim3(128), 0, g.stream, a, nbeads, n_steps, a.ws); break;
        case 256: hipLaunchKernelGGL(k_noise<256>, grid, dim3(256), 0, g.stream, a, nbeads, n_steps, a.ws); break;
        case 512: hipLaunchKernelGGL(k_noise<512>, grid, dim3(512), 0, g.stream, a, nbeads, n_steps, a.ws); break;
        default: hipLaunchKernelGGL(k_noise<1024>, grid, dim3(1024), 0, g.stream, a, nbeads, n_steps, a.ws); break;
    }
    HIP_TRY(hipGetLastError());
    if (nd) HIP_TRY(hipMemcpyAsync(X, g.ws.p, sizeof(double) * nd, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipMemcpyAsync(rng, g.rng.p, sizeof(RngState) * n_traj, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

int bdsim_rpy_apply_batch(int n_traj, int nbeads, double hstar, int precision, const double* R,
                          const double* x, double* y) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init()) return rc;
    if (n_traj <= 0 || nbeads < 1 || nbeads > MAX_BEADS || !R || !x || !y)
        return fail(BDSIM_EINVAL, "bad argument");
    if (!valid_precision(precision))
        return fail(BDSIM_EINVAL, "bad precision");
    TrajParams tp;
    std::memset(&tp, 0, sizeof tp);
    tp.N = nbeads;
    set_rpy(tp, hstar);
    const size_t nd = (size_t)n_traj * 3 * nbeads;
    if (int rc = g.tp.ensure(sizeof(TrajParams))) return rc;
    if (int rc = g.R.ensure(sizeof(double) * nd)) return rc;
    if (int rc = g.blob.ensure(sizeof(double) * nd)) return rc;
    if (int rc = g.ws.ensure(sizeof(double) * nd)) return rc;
    HIP_TRY(hipMemcpyAsync(g.tp.p, &tp, sizeof tp, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.R.p, R, sizeof(double) * nd, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipMemcpyAsync(g.blob.p, x, sizeof(double) * nd, hipMemcpyHostToDevice, g.stream));
    KArgs a{};
    a.tp = as<TrajParams>(g.tp);
    a.R = as<double>(g.R);
    a.blob = as<double>(g.blob);
    a.ws = as<double>(g.ws);
    dim3 grid(n_traj);
    const bool mixed = is_mixed(precision);
#define LAUNCH_A(BS)                                                                         \
    if (mixed) hipLaunchKernelGGL((k_rpy_apply<BS, true>), grid, dim3(BS), 0, g.stream, a, a.blob, a.ws);  \
    else hipLaunchKernelGGL((k_rpy_apply<BS, false>), grid, dim3(BS), 0, g.stream, a, a.blob, a.ws);
    switch (block_for(nbeads)) {
        case 64: LAUNCH_A(64); break;
        case 128: LAUNCH_A(128); break;
        case 256: LAUNCH_A(256); break;
        case 512: LAUNCH_A(512); break;
        default: LAUNCH_A(1024); break;
    }
#undef LAUNCH_A
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, g.ws.p, sizeof(double) * nd, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return 0;
}

}  // extern "C"

#include "bdsim_server.hpp"
