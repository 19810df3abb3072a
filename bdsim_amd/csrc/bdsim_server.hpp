// Serving loop: a resident batch that admits integrate jobs while others run.
//
// Host side of bdsim_server_* (include/bdsim_gpu.h). Included at the end of
// bdsim_kernels.hip: it launches the same k_integrate instantiations through
// launch_integrate and packs trajectories with pack_traj.
//
// Layout in HBM, fixed at creation from the memory budget:
//   slots   per-trajectory records (TrajParams, RngState, status, Counters,
//           SegState), a job taking a contiguous range of slots;
//   R       positions, 3N doubles per trajectory, a job's range contiguous;
//   ws      workspaces, ws_layout(N, hi).total doubles per trajectory;
//   blob    phi / flow / external-force tables of each job;
//   ext     external-force entries of each job.
// Every region is handed out first-fit and returned when the job completes,
// so admission is by HBM footprint: a queued job is admitted (in arrival
// order, later jobs may pass one that does not fit yet) as soon as its
// ranges fit.
//
// Stepping: trajectories are grouped by block size (64..1024 threads, from
// N); Cholesky chains with 3N <= 96 form two groups of their own on the
// register-factor kernel. Small jobs of long chains (N > 256, few chains)
// hold G workspace copies per chain and form split groups (one per G): a
// launch of such a group runs its chains split over G co-resident workgroups
// each (k_integrate_split) while the device has idle SIMDs, else whole. Each group runs its own sequence of launches on its own stream with a
// per-launch step budget; a trajectory's segment is resumed launch after
// launch (SegState), which is bit for bit one uninterrupted integrate call, so
// a job's result does not depend on what else is resident. poll() retires the
// groups whose launch has finished: it collects finished trajectories,
// completes jobs whose trajectories are all done (results copied to host),
// frees their ranges, admits queued jobs of that group and launches again.
// Groups run concurrently.

#include <chrono>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>

namespace {

// First-fit allocator over [0, cap) in abstract units.
class RangeAlloc {
    std::map<int64_t, int64_t> free_;   // offset -> length
    int64_t cap_ = 0, used_ = 0;

public:
    void reset(int64_t cap) {
        free_.clear();
        cap_ = cap;
        used_ = 0;
        if (cap > 0) free_[0] = cap;
    }
    int64_t alloc(int64_t n) {
        if (n <= 0) return 0;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const int64_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_ += n;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t n) {
        if (n <= 0) return;
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t cap() const { return cap_; }
    int64_t used() const { return used_; }
};

constexpr int64_t WS_ALIGN = 32;   // doubles (256 B)

int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct SrvJob {
    int64_t id = 0;
    int n_traj = 0, N = 0, block = 64;
    int nb = 0;                       // > 0: Cholesky chains on the register-factor kernel
    int split = 1;                    // > 1: G workspace copies per chain (a split group)
    // packed per-trajectory parameters (blob/ext offsets relative to the job)
    std::vector<TrajParams> tp;
    std::vector<double> blob;
    std::vector<ExtEntry> ext;
    std::vector<double> R0;
    std::vector<RngState> rng0;
    double t0 = 0.0;
    int64_t wsper = 0;
    // ranges once admitted
    int64_t slot = -1, R_off = -1, ws_off = -1, blob_off = -1, ext_off = -1;
    int64_t ws_len = 0, blob_len = 0, ext_len = 0;
    int remaining = 0;
    std::vector<char> done;           // per trajectory
    // results
    std::vector<double> R;
    std::vector<RngState> rng;
    std::vector<int32_t> status;
    std::vector<Counters> cnt;
    int state = 0;                    // 0 queued, 1 running, 2 complete
};

struct SrvGroup {
    int block = 64;
    int nb = 0;                       // Cholesky kernel of 16 nb rows (k_integrate_dreg<*, nb, 2>)
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool in_flight = false;
    std::vector<int32_t> active;      // slots of the running trajectories
    std::vector<int32_t> launched;    // slots of the launch in flight
    int32_t* d_list = nullptr;        // device: launched slots
    int32_t* d_done = nullptr;        // device: done flag per launched slot
    int64_t cap = 0;
    std::vector<int32_t> h_done;
    int steps = 1;
    // split groups: workgroups per chain, and the barrier counters and
    // exchange buffers of a split launch (scap chains)
    int split = 1;
    uint32_t* d_bar = nullptr;
    double* d_xbuf = nullptr;
    int64_t scap = 0;
};

// groups 0-4: k_integrate by block size; 5 and 6: the Cholesky chains with
// 3N <= 64, 3N <= 96; 7-11: split groups of chains of 257-320, 321-384,
// 385-448, 449-512 and 513-1024 beads (G = 5, 6, 7, 8, 8 workgroups per chain:
// the tiles t = p (mod G) for part p, so G <= T tiles keeps every part busy)
constexpr int SRV_GROUPS = 12;
constexpr int SRV_SPLIT0 = 7;

// Workgroups per chain of a job of n_traj chains of N beads over `steps`
// steps: a split group's G for small, long jobs of long chains (N > 256, at
// most 16 chains, at least 4 compute units per chain, 2000 steps or more: the
// jobs whose latency a split shortens), else 1. A split launch sums the
// partner terms in another order, so such a job's result can differ from an
// unsplit run in the last bits (within the forms' tolerances).
// BDSIM_SERVER_SPLIT=0 in the environment turns it off.
constexpr int SRV_SPLIT_MIN_STEPS = 2000;
int srv_split_of(int N, int n_traj, int64_t steps) {
    static const bool enabled = [] {
        const char* e = std::getenv("BDSIM_SERVER_SPLIT");
        return !(e && e[0] == '0');
    }();
    if (!enabled || !g.coop || N <= 256 || n_traj > 16 || (int64_t)n_traj * 4 > g.n_cu ||
        steps < SRV_SPLIT_MIN_STEPS)
        return 1;
    return std::min(8, (N + 63) / 64);
}

int split_group_of(int N) {
    return N <= 320 ? SRV_SPLIT0 : N <= 384 ? SRV_SPLIT0 + 1 : N <= 448 ? SRV_SPLIT0 + 2
                                   : N <= 512 ? SRV_SPLIT0 + 3 : SRV_SPLIT0 + 4;
}

int group_of(int block, int nb) {
    if (nb == 4) return 5;
    if (nb == 6) return 6;
    switch (block) {
        case 64: return 0;
        case 128: return 1;
        case 256: return 2;
        case 512: return 3;
        default: return 4;
    }
}

}  // namespace

struct bdsim_server {
    std::mutex mu;
    int precision = 0;
    int64_t slot_cap = 0;
    TrajParams* tp = nullptr;
    RngState* rng = nullptr;
    int32_t* status = nullptr;
    Counters* cnt = nullptr;
    SegState* seg = nullptr;
    double* R = nullptr;
    double* ws = nullptr;
    double* blob = nullptr;
    ExtEntry* ext = nullptr;
    RangeAlloc a_slot, a_R, a_ws, a_blob, a_ext;
    std::vector<int64_t> slot_job;    // job id per slot
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, SrvJob> jobs;
    std::deque<int64_t> completed;
    SrvGroup grp[SRV_GROUPS];
    int64_t n_running = 0, n_active_traj = 0;
    int64_t bead_steps_done = 0;

    ~bdsim_server() {
        for (auto& gr : grp) {
            if (gr.stream) (void)hipStreamSynchronize(gr.stream);
            if (gr.ev) (void)hipEventDestroy(gr.ev);
            if (gr.stream) (void)hipStreamDestroy(gr.stream);
            (void)hipFree(gr.d_list);
            (void)hipFree(gr.d_done);
            (void)hipFree(gr.d_bar);
            (void)hipFree(gr.d_xbuf);
        }
        for (void* p : {(void*)tp, (void*)rng, (void*)status, (void*)cnt, (void*)seg, (void*)R,
                        (void*)ws, (void*)blob, (void*)ext})
            (void)hipFree(p);
    }
};

namespace {

int srv_group_of(const SrvJob& j) {
    return j.split > 1 ? split_group_of(j.N) : group_of(j.block, j.nb);
}

// Before a split launch: every part's workspace copy takes the warm Lanczos
// seed from copy 0 (the only workspace content a step takes from the one
// before; whole launches keep it in copy 0 alone). One workgroup per chain.
__global__ void k_seed_fanout(const TrajParams* tp, const int32_t* list, double* ws, int G) {
    const TrajParams& P = tp[list[blockIdx.x]];
    const WsLayout wl = ws_layout(P.N, P.hstar > 0.0);
    const double* src = ws + P.ws_off + wl.S + (int64_t)SLOT_SEED * P.N;
    for (int p = 1; p < G; ++p) {
        double* dst = ws + P.ws_off + (int64_t)p * wl.total + wl.S + (int64_t)SLOT_SEED * P.N;
        for (int e = threadIdx.x; e < 3 * P.N; e += blockDim.x) dst[e] = src[e];
    }
}

// Try to place a queued job; false when some range does not fit yet.
bool srv_admit(bdsim_server* s, SrvJob& j) {
    const int64_t nR = (int64_t)j.n_traj * 3 * j.N;
    const int64_t slot = s->a_slot.alloc(j.n_traj);
    if (slot < 0) return false;
    const int64_t roff = s->a_R.alloc(nR);
    const int64_t wstride = align_up(j.wsper * j.split, WS_ALIGN);
    const int64_t wlen = wstride * j.n_traj;
    const int64_t woff = roff < 0 ? -1 : s->a_ws.alloc(wlen);
    const int64_t blen = (int64_t)j.blob.size();
    const int64_t boff = woff < 0 ? -1 : s->a_blob.alloc(blen);
    const int64_t elen = (int64_t)j.ext.size();
    const int64_t eoff = boff < 0 ? -1 : s->a_ext.alloc(elen);
    if (roff < 0 || woff < 0 || boff < 0 || eoff < 0) {
        if (eoff >= 0) s->a_ext.release(eoff, elen);
        if (boff >= 0) s->a_blob.release(boff, blen);
        if (woff >= 0) s->a_ws.release(woff, wlen);
        if (roff >= 0) s->a_R.release(roff, nR);
        s->a_slot.release(slot, j.n_traj);
        return false;
    }
    j.slot = slot; j.R_off = roff; j.ws_off = woff; j.blob_off = boff; j.ext_off = eoff;
    j.ws_len = wlen; j.blob_len = blen; j.ext_len = elen;
    // relocate the job's tables into the arenas
    std::vector<TrajParams> tps = j.tp;
    for (int i = 0; i < j.n_traj; ++i) {
        TrajParams& t = tps[i];
        t.R_off = roff + (int64_t)i * 3 * j.N;
        t.ws_off = woff + wstride * i;       // copy p of a split chain at + p * wsper
        t.phi_off += boff;
        t.flow_off += boff;
        t.ext_off += eoff;
        t.log_off = -1;
    }
    std::vector<ExtEntry> ext = j.ext;
    for (auto& e : ext) e.off += boff;
    std::vector<SegState> seg(j.n_traj, SegState{j.t0, 0, 0});
    std::vector<int32_t> st(j.n_traj, 0);
    std::vector<Counters> cz(j.n_traj, Counters{});
    SrvGroup& gr = s->grp[srv_group_of(j)];
    hipStream_t q = gr.stream;
    bool ok = hipMemcpyAsync(s->tp + slot, tps.data(), sizeof(TrajParams) * j.n_traj, hipMemcpyHostToDevice, q) == hipSuccess &&
              hipMemcpyAsync(s->R + roff, j.R0.data(), sizeof(double) * nR, hipMemcpyHostToDevice, q) == hipSuccess &&
              hipMemcpyAsync(s->rng + slot, j.rng0.data(), sizeof(RngState) * j.n_traj, hipMemcpyHostToDevice, q) == hipSuccess &&
              hipMemcpyAsync(s->status + slot, st.data(), sizeof(int32_t) * j.n_traj, hipMemcpyHostToDevice, q) == hipSuccess &&
              hipMemcpyAsync(s->cnt + slot, cz.data(), sizeof(Counters) * j.n_traj, hipMemcpyHostToDevice, q) == hipSuccess &&
              hipMemcpyAsync(s->seg + slot, seg.data(), sizeof(SegState) * j.n_traj, hipMemcpyHostToDevice, q) == hipSuccess &&
              (blen == 0 || hipMemcpyAsync(s->blob + boff, j.blob.data(), sizeof(double) * blen, hipMemcpyHostToDevice, q) == hipSuccess) &&
              (elen == 0 || hipMemcpyAsync(s->ext + eoff, ext.data(), sizeof(ExtEntry) * elen, hipMemcpyHostToDevice, q) == hipSuccess);
    // the staging vectors above are local: let the copies finish before they go
    ok = ok && hipStreamSynchronize(q) == hipSuccess;
    if (!ok) return false;
    for (int i = 0; i < j.n_traj; ++i) {
        s->slot_job[slot + i] = j.id;
        gr.active.push_back((int32_t)(slot + i));
    }
    j.remaining = j.n_traj;
    j.done.assign(j.n_traj, 0);
    j.state = 1;
    s->n_running += 1;
    s->n_active_traj += j.n_traj;
    return true;
}

// Copy a finished job's results out and give its ranges back.
int srv_complete(bdsim_server* s, SrvJob& j, hipStream_t q) {
    const int64_t nR = (int64_t)j.n_traj * 3 * j.N;
    j.R.resize(nR);
    j.rng.resize(j.n_traj);
    j.status.resize(j.n_traj);
    j.cnt.resize(j.n_traj);
    HIP_TRY(hipMemcpyAsync(j.R.data(), s->R + j.R_off, sizeof(double) * nR, hipMemcpyDeviceToHost, q));
    HIP_TRY(hipMemcpyAsync(j.rng.data(), s->rng + j.slot, sizeof(RngState) * j.n_traj, hipMemcpyDeviceToHost, q));
    HIP_TRY(hipMemcpyAsync(j.status.data(), s->status + j.slot, sizeof(int32_t) * j.n_traj, hipMemcpyDeviceToHost, q));
    HIP_TRY(hipMemcpyAsync(j.cnt.data(), s->cnt + j.slot, sizeof(Counters) * j.n_traj, hipMemcpyDeviceToHost, q));
    HIP_TRY(hipStreamSynchronize(q));
    for (int i = 0; i < j.n_traj; ++i) s->slot_job[j.slot + i] = -1;
    s->a_slot.release(j.slot, j.n_traj);
    s->a_R.release(j.R_off, nR);
    s->a_ws.release(j.ws_off, j.ws_len);
    s->a_blob.release(j.blob_off, j.blob_len);
    s->a_ext.release(j.ext_off, j.ext_len);
    for (const auto& c : j.cnt) s->bead_steps_done += c.steps * (int64_t)j.N;
    j.state = 2;
    j.tp.clear(); j.blob.clear(); j.ext.clear(); j.R0.clear(); j.rng0.clear();
    s->n_running -= 1;
    s->n_active_traj -= j.n_traj;
    s->completed.push_back(j.id);
    return 0;
}

// One group: finish its last launch, admit, launch again.
int srv_service(bdsim_server* s, int gi) {
    SrvGroup& gr = s->grp[gi];
    if (gr.in_flight) {
        const hipError_t e = hipEventQuery(gr.ev);
        if (e == hipErrorNotReady) return 0;
        if (e != hipSuccess) return fail(BDSIM_EHIP, std::string("serving launch: ") + hipGetErrorString(e));
        gr.in_flight = false;
        const size_t n = gr.launched.size();
        gr.h_done.resize(n);
        HIP_TRY(hipMemcpy(gr.h_done.data(), gr.d_done, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        std::vector<int32_t> still;
        still.reserve(gr.active.size());
        for (size_t k = 0; k < n; ++k) {
            const int32_t sl = gr.launched[k];
            if (!gr.h_done[k]) { still.push_back(sl); continue; }
            auto it = s->jobs.find(s->slot_job[sl]);
            if (it == s->jobs.end()) continue;
            SrvJob& j = it->second;
            const int i = (int)(sl - j.slot);
            if (!j.done[i]) { j.done[i] = 1; j.remaining -= 1; }
            if (j.remaining == 0)
                if (int rc = srv_complete(s, j, gr.stream)) return rc;
        }
        gr.active.swap(still);
    }
    // admit queued jobs of this group, in arrival order
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        SrvJob& j = s->jobs[*it];
        if (srv_group_of(j) == gi && srv_admit(s, j)) it = s->queue.erase(it);
        else ++it;
    }
    if (gr.active.empty()) return 0;
    const int64_t n = (int64_t)gr.active.size();
    if (n > gr.cap) {
        (void)hipFree(gr.d_list);
        (void)hipFree(gr.d_done);
        gr.d_list = nullptr;
        gr.d_done = nullptr;
        const int64_t c = std::max<int64_t>(n, 2 * gr.cap);
        if (hipMalloc(&gr.d_list, sizeof(int32_t) * c) != hipSuccess ||
            hipMalloc(&gr.d_done, sizeof(int32_t) * c) != hipSuccess)
            return fail(BDSIM_ENOMEM, "serving: launch list allocation failed");
        gr.cap = c;
    }
    gr.launched = gr.active;
    HIP_TRY(hipMemcpyAsync(gr.d_list, gr.launched.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, gr.stream));
    KArgs a{};
    a.tp = s->tp;
    a.list = gr.d_list;
    a.R = s->R;
    a.ws = s->ws;
    a.blob = s->blob;
    a.ext = s->ext;
    a.rng = s->rng;
    a.status = s->status;
    a.cnt = s->cnt;
    a.log = nullptr;
    a.jx = g.jx;
    a.jy = g.jy;
    a.sim_override = 0;
    a.seg = s->seg;
    a.done_out = gr.d_done;
    a.max_steps = gr.steps;
    // waves of all groups' active trajectories: the groups run side by side
    long long waves = 0;
    for (const SrvGroup& o : s->grp) waves += (long long)o.active.size() * (o.block / 64);
    // a split group runs split while the whole device holds no more than one
    // wave per SIMD (its parts, workgroups of 1024 threads and most of a
    // CU's LDS each, then find idle compute units)
    if (gr.split > 1 && waves <= device_simds()) {
        if (n > gr.scap) {
            (void)hipFree(gr.d_bar);
            (void)hipFree(gr.d_xbuf);
            gr.d_bar = nullptr;
            gr.d_xbuf = nullptr;
            const int64_t c = std::max<int64_t>(n, 2 * gr.scap);
            if (hipMalloc(&gr.d_bar, sizeof(uint32_t) * c) != hipSuccess ||
                hipMalloc(&gr.d_xbuf, sizeof(double) * c * split_xbuf_doubles(gr.split)) != hipSuccess)
                return fail(BDSIM_ENOMEM, "serving: split buffers allocation failed");
            gr.scap = c;
        }
        hipLaunchKernelGGL(k_seed_fanout, dim3((unsigned)n), dim3(256), 0, gr.stream, s->tp, gr.d_list, s->ws,
                           gr.split);
        HIP_TRY(hipGetLastError());
        const int rc = is_mixed(s->precision)
                           ? launch_split<true>((int)n, gr.split, gr.stream, a, gr.d_bar, gr.d_xbuf)
                           : launch_split<false>((int)n, gr.split, gr.stream, a, gr.d_bar, gr.d_xbuf);
        if (rc) return rc;
    } else {
        // the launch's longest chain sizes the LDS of the 64-thread kernels
        // (workspace part, pair-weight table)
        int nmax = 0;
        if (gr.block == 64)
            for (const int32_t sl : gr.launched) {
                auto it = s->jobs.find(s->slot_job[sl]);
                if (it != s->jobs.end()) nmax = std::max(nmax, it->second.N);
            }
        if (is_mixed(s->precision))
            launch_integrate<true>(gr.block, dim3((unsigned)n), gr.stream, a, gr.nb, waves,
                                   s->precision == BDSIM_PRECISION_MIXED_F32_FACTOR, nmax);
        else
            launch_integrate<false>(gr.block, dim3((unsigned)n), gr.stream, a, gr.nb, waves, false, nmax);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(gr.ev, gr.stream));
    gr.in_flight = true;
    return 0;
}

}  // namespace

extern "C" {

int bdsim_server_create(bdsim_server** out, int precision, int64_t hbm_bytes, int steps_per_launch) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (int rc = ensure_init()) return rc;
    if (!out) return fail(BDSIM_EINVAL, "null argument");
    if (!valid_precision(precision))
        return fail(BDSIM_EINVAL, "bad precision");
    if (hbm_bytes < ((int64_t)64 << 20)) return fail(BDSIM_EINVAL, "memory budget below 64 MiB");
    if (steps_per_launch < 1) return fail(BDSIM_EINVAL, "steps_per_launch must be >= 1");
    bdsim_server* s = new bdsim_server();
    s->precision = precision;
    // budget split: workspaces take what the rest leaves
    const int64_t slot_bytes = sizeof(TrajParams) + sizeof(RngState) + sizeof(int32_t) + sizeof(Counters) + sizeof(SegState);
    s->slot_cap = std::min<int64_t>(1 << 18, hbm_bytes / 64 / slot_bytes);
    const int64_t R_cap = hbm_bytes / 32 / 8;
    const int64_t blob_cap = hbm_bytes / 32 / 8;
    const int64_t ext_cap = std::max<int64_t>(1024, hbm_bytes / 256 / (int64_t)sizeof(ExtEntry));
    const int64_t used = s->slot_cap * slot_bytes + R_cap * 8 + blob_cap * 8 + ext_cap * (int64_t)sizeof(ExtEntry);
    const int64_t ws_cap = (hbm_bytes - used) / 8;
    auto bad = [&](const char* what) { delete s; return fail(BDSIM_ENOMEM, what); };
    if (hipMalloc(&s->tp, sizeof(TrajParams) * s->slot_cap) != hipSuccess ||
        hipMalloc(&s->rng, sizeof(RngState) * s->slot_cap) != hipSuccess ||
        hipMalloc(&s->status, sizeof(int32_t) * s->slot_cap) != hipSuccess ||
        hipMalloc(&s->cnt, sizeof(Counters) * s->slot_cap) != hipSuccess ||
        hipMalloc(&s->seg, sizeof(SegState) * s->slot_cap) != hipSuccess ||
        hipMalloc(&s->R, sizeof(double) * R_cap) != hipSuccess ||
        hipMalloc(&s->blob, sizeof(double) * blob_cap) != hipSuccess ||
        hipMalloc(&s->ext, sizeof(ExtEntry) * ext_cap) != hipSuccess ||
        hipMalloc(&s->ws, sizeof(double) * ws_cap) != hipSuccess)
        return bad("serving: HBM reservation failed");
    s->a_slot.reset(s->slot_cap);
    s->a_R.reset(R_cap);
    s->a_blob.reset(blob_cap);
    s->a_ext.reset(ext_cap);
    s->a_ws.reset(ws_cap / WS_ALIGN * WS_ALIGN);
    s->slot_job.assign(s->slot_cap, -1);
    const int blocks[SRV_GROUPS] = {64, 128, 256, 512, 1024, 64, 64, 512, 512, 512, 512, 1024};
    const int nbs[SRV_GROUPS] = {0, 0, 0, 0, 0, 4, 6, 0, 0, 0, 0, 0};
    const int splits[SRV_GROUPS] = {1, 1, 1, 1, 1, 1, 1, 5, 6, 7, 8, 8};
    for (int i = 0; i < SRV_GROUPS; ++i) {
        SrvGroup& gr = s->grp[i];
        gr.block = blocks[i];
        gr.nb = nbs[i];
        gr.split = splits[i];
        // about the same time per launch in every group: the pair work of a
        // chain step grows with BLOCK^2 spread over BLOCK / 64 waves
        gr.steps = std::max(1, steps_per_launch * 64 / blocks[i]);
        if (hipStreamCreateWithFlags(&gr.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&gr.ev, hipEventDisableTiming) != hipSuccess)
            return bad("serving: stream creation failed");
    }
    *out = s;
    return 0;
}

int bdsim_server_submit(bdsim_server* s, int64_t job_id, int n_traj, int nbeads, const double* R0,
                        const bdsim_phys* phys, const bdsim_sim* sim, const bdsim_rng_state* rng) {
    if (!s || !R0 || !phys || !sim || !rng) return fail(BDSIM_EINVAL, "null argument");
    if (n_traj < 1) return fail(BDSIM_EINVAL, "n_traj must be >= 1");
    if (nbeads < 2 || nbeads > MAX_BEADS) return fail(BDSIM_EINVAL, "nbeads out of range");
    if (int rc = check_clock(sim->time_start, sim->time_end, sim->dt)) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->jobs.count(job_id)) return fail(BDSIM_EINVAL, "job id already in use");
    SrvJob j;
    j.id = job_id;
    j.n_traj = n_traj;
    j.N = nbeads;
    j.block = block_for(nbeads);
    j.nb = dense_nb_for(nbeads, phys->hstar, phys->hi_method);
    j.split = srv_split_of(nbeads, n_traj,
                           (int64_t)std::floor((sim->time_end - sim->time_start) / sim->dt + 0.5) + 1);
    TrajParams proto;
    if (int rc = pack_traj(*phys, *sim, nbeads, j.blob, j.ext, proto)) return rc;
    j.tp.assign(n_traj, proto);
    j.wsper = ws_layout(nbeads, phys->hstar > 0.0).total;
    j.t0 = sim->time_start;
    j.R0.assign(R0, R0 + (size_t)n_traj * 3 * nbeads);
    j.rng0.resize(n_traj);
    for (int i = 0; i < n_traj; ++i) j.rng0[i] = RngState{rng[i].ix, rng[i].iy, rng[i].seed};
    if ((int64_t)n_traj > s->a_slot.cap() || (int64_t)n_traj * 3 * nbeads > s->a_R.cap() ||
        align_up(j.wsper * j.split, WS_ALIGN) * n_traj > s->a_ws.cap() ||
        (int64_t)j.blob.size() > s->a_blob.cap() || (int64_t)j.ext.size() > s->a_ext.cap())
        return fail(BDSIM_ENOMEM, "job larger than the whole memory budget");
    s->jobs.emplace(job_id, std::move(j));
    s->queue.push_back(job_id);
    return 0;
}

int bdsim_server_poll(bdsim_server* s, int wait_us) {
    if (!s) return fail(BDSIM_EINVAL, "null argument");
    const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(wait_us > 0 ? wait_us : 0);
    for (;;) {
        size_t before;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            before = s->completed.size();
            for (int gi = 0; gi < SRV_GROUPS; ++gi)
                if (int rc = srv_service(s, gi)) return rc;
            if (s->completed.size() > before || wait_us <= 0) return (int)s->completed.size();
            if (s->queue.empty() && s->n_running == 0) return (int)s->completed.size();
        }
        if (std::chrono::steady_clock::now() >= until) return (int)s->completed.size();
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
}

int bdsim_server_completed(bdsim_server* s, int64_t* ids, int max_ids) {
    if (!s || (!ids && max_ids > 0)) return fail(BDSIM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int n = 0;
    while (n < max_ids && !s->completed.empty()) {
        ids[n++] = s->completed.front();
        s->completed.pop_front();
    }
    return n;
}

int bdsim_server_result(bdsim_server* s, int64_t job_id, double* R_out, bdsim_rng_state* rng_out,
                        int32_t* status, bdsim_counters* counters) {
    if (!s) return fail(BDSIM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return fail(BDSIM_EINVAL, "unknown job id");
    SrvJob& j = it->second;
    if (j.state != 2) return fail(BDSIM_EINVAL, "job not complete");
    if (R_out) std::memcpy(R_out, j.R.data(), sizeof(double) * j.R.size());
    for (int i = 0; i < j.n_traj; ++i) {
        if (rng_out) rng_out[i] = bdsim_rng_state{j.rng[i].ix, j.rng[i].iy, j.rng[i].seed};
        if (status) status[i] = j.status[i];
        if (counters) std::memcpy(&counters[i], &j.cnt[i], sizeof(bdsim_counters));
    }
    s->jobs.erase(it);
    return 0;
}

int bdsim_server_query(bdsim_server* s, bdsim_server_stats* st) {
    if (!s || !st) return fail(BDSIM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->queued = (int64_t)s->queue.size();
    st->running = s->n_running;
    st->active_trajectories = s->n_active_traj;
    st->bytes_reserved = (s->a_ws.cap() + s->a_R.cap() + s->a_blob.cap()) * 8 +
                         s->a_ext.cap() * (int64_t)sizeof(ExtEntry);
    st->bytes_in_use = (s->a_ws.used() + s->a_R.used() + s->a_blob.used()) * 8 +
                       s->a_ext.used() * (int64_t)sizeof(ExtEntry);
    st->bead_steps_done = s->bead_steps_done;
    return 0;
}

void bdsim_server_destroy(bdsim_server* s) { delete s; }

}  // extern "C"
