/*
 * bdsim_gpu.h -- C ABI of the MI355X ChainIntegrator library (libbdsim_gpu.so).
 *
 * The library integrates batches of independent bead-spring chains with the
 * semi-implicit predictor-corrector step of bdsim on one AMD Instinct MI355X
 * (gfx950).  Every entry point takes plain pointers and sizes; parameter objects
 * are plain-old-data mirrors of bdsim's C++ parameter structs, so a caller that
 * holds a bdsim PhysParams / SimParams / Rng can fill them field by field.
 *
 * Reference interfaces replaced (IsaacPincus/bdsim):
 *   bdsim_integrate_batch     <- time_integrate_chain(Vec3Field&, const PhysParams&,
 *                                 const SimParams&, Rng&)      src/integrator.hpp:32-33
 *                                and its binding _bdsim.integrate  bindings/module.cpp:45-50
 *   bdsim_total_force_batch   <- _bdsim.total_force            bindings/module.cpp:67-76
 *   bdsim_spring_force_batch  <- _bdsim.spring_force           bindings/module.cpp:52-55
 *   bdsim_bending_force_batch <- _bdsim.bending_force          bindings/module.cpp:56-60
 *   bdsim_ev_force_batch      <- _bdsim.ev_force               bindings/module.cpp:61-66
 *   bdsim_external_force_batch<- _bdsim.external_force         bindings/module.cpp:128-132
 *   bdsim_raw_noise_batch     <- raw_noise + Rng::ran_1        src/mobility.cpp:18-29,
 *                                                              src/rng.cpp:41-70
 *   bdsim_rpy_apply_batch     <- Diffusion::apply (D.F)        src/mobility.cpp:53-56
 *
 * Error convention: every entry point returns 0 on success and a negative
 * BDSIM_E* code on an API error (bad argument, no device, out of memory); the
 * text of the last error is available from bdsim_last_error().  Per-trajectory
 * failures of the physics (a non positive-definite diffusion tensor, an
 * unbracketed spring root, ...) are NOT API errors: they are reported in the
 * status[] array, one BDSIM_STATUS_* code per trajectory, and never abort the
 * batch.
 */
#ifndef BDSIM_GPU_H
#define BDSIM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums: same integer values as bdsim's src/types.hpp:27-53 ---------- */
enum { BDSIM_SPRING_HOOK = 1, BDSIM_SPRING_FENE = 2, BDSIM_SPRING_ILC = 3,
       BDSIM_SPRING_WLC = 4, BDSIM_SPRING_FRAENKEL = 5,
       BDSIM_SPRING_FENE_FRAENKEL = 6, BDSIM_SPRING_WLC_BOUNDED = 7 };
enum { BDSIM_EV_NONE = 0, BDSIM_EV_GAUSS = 1, BDSIM_EV_LJ = 2, BDSIM_EV_SDK = 3,
       BDSIM_EV_SDK_STICKERS = 4 };
enum { BDSIM_BENDING_NONE = 0, BDSIM_BENDING_ONE_MINUS_COS = 1 };
enum { BDSIM_HI_CHEBYSHEV = 0, BDSIM_HI_CHOLESKY = 1, BDSIM_HI_EXACT_SQRT = 2 };

/* Precision forms of the device step. */
enum { BDSIM_PRECISION_FP64 = 0,   /* plain FP64: parity form                */
       BDSIM_PRECISION_MIXED = 1,  /* FP32 RPY/Chebyshev/EV pair sums with   */
                                   /* FP64 accumulation; FP64 Lanczos bounds, */
                                   /* Cholesky, connector solve, update      */
       BDSIM_PRECISION_MIXED_F32_FACTOR = 2 };
                                   /* the mixed form with the Cholesky factor */
                                   /* of chains with 3N <= 96 held in FP32    */
                                   /* (half the LDS: 5 chains per CU instead  */
                                   /* of 3); a pivot that fails in FP32 is    */
                                   /* factored again in FP64                  */

/* Per-trajectory status codes. */
enum {
    BDSIM_STATUS_OK = 0,
    BDSIM_STATUS_NOT_POSITIVE_DEFINITE = 1, /* cholesky_lower threw (linalg.cpp:16-24) */
    BDSIM_STATUS_RTSAFE_UNBRACKETED = 2,    /* rtsafe abort (numerics.hpp:33-36)     */
    BDSIM_STATUS_RTSAFE_MAXIT = 3,          /* rtsafe abort (numerics.hpp:71-72)     */
    BDSIM_STATUS_EXACT_SQRT = 4,            /* ExactSqrt abort (mobility.cpp:83-85)  */
    BDSIM_STATUS_WLCB_SIGMA = 5,            /* sigma/L > 1 abort (spring.cpp:101-105)*/
    BDSIM_STATUS_NOT_RUN = 6,               /* batch entry was skipped               */
    BDSIM_STATUS_SYNC_TIMEOUT = 7           /* a chain split over several workgroups
                                               lost a part (they were not resident
                                               together); no counterpart in the
                                               reference                            */
};

/* API error codes (negative return values). */
enum { BDSIM_OK = 0, BDSIM_EINVAL = -1, BDSIM_ENODEV = -2, BDSIM_ENOMEM = -3,
       BDSIM_EHIP = -4 };

/* ---- parameter PODs ---------------------------------------------------- */

/* Flow (src/flow.hpp:17-43): constant kappa, or a (time, kappa) table that is
 * linearly interpolated and clamped at its ends. kappa is row-major 3x3.     */
typedef struct bdsim_flow {
    int32_t constant;          /* nonzero: kappa0 at every t                   */
    int32_t n_times;           /* table length (ignored when constant)         */
    double kappa0[9];
    const double* times;       /* [n_times], strictly ascending                */
    const double* kappas;      /* [n_times * 9]                                */
} bdsim_flow;

/* One applied bead force (src/external_force.hpp:24-30). */
typedef struct bdsim_bead_force {
    int32_t bead;              /* negative counts from the end (-1 = last)     */
    int32_t constant;
    double value[3];
    int32_t n_times;
    const double* times;       /* [n_times]                                    */
    const double* values;      /* [n_times * 3]                                */
} bdsim_bead_force;

/* PhysParams (src/model.hpp:20-35) with SpringParams (spring.hpp:12-19),
 * EVParams (excluded_volume.hpp:17-25) and BendingParams (bending.hpp:9-12)
 * flattened in. number_of_beads is overwritten by the chain length, as the
 * binding does (bindings/module.cpp:47).                                     */
typedef struct bdsim_phys {
    int32_t spring_type;
    double sqrtb;
    double natural_length;

    int32_t ev_type;
    double zstar, dstar, min_cutoff, max_cutoff;
    int32_t contour_dist_for_ev;
    int32_t n_phi;             /* 0, or N*N (SDK attractive strengths)         */
    const double* phi;         /* row-major [N*N], read for pairs mu<nu         */

    int32_t bend_type;
    double bend_stiffness;

    bdsim_flow flow;
    int32_t n_external;
    const bdsim_bead_force* external;

    int32_t number_of_beads;
    double hstar;
    int32_t hi_method;
    double ncheb_multiplier;
    double fd_err_max;
    int32_t equilibration;
} bdsim_phys;

/* SimParams (src/integrator.hpp:22-28). */
typedef struct bdsim_sim {
    double time_start, time_end, dt, implicit_loop_tol;
    int32_t update_center_of_mass;
} bdsim_sim;

/* Rng state (ix, iy, seed) exactly as Rng::State (src/rng.hpp:46-48). */
typedef struct bdsim_rng_state { int32_t ix, iy, seed; } bdsim_rng_state;

/* Per-trajectory work counters of one integrate call. matvecs counts every
 * 3N x 3N mobility product D.x the device forms (Lanczos, D.x0, Chebyshev
 * terms, predictor, upsilon, corrector); it is the n_mv of the roofline
 * accounting. Per step it is two fewer than the reference forms when the
 * Chebyshev square root runs: D.x0 is formed once for the FD test and the
 * series, and the upsilon's two products are one product of the summed
 * forces (D is linear).                                                     */
typedef struct bdsim_counters {
    int64_t steps;
    int64_t matvecs;
    int64_t lanczos_iters;     /* Lanczos mat-vecs                              */
    int64_t lanczos_cold;      /* cold Lanczos runs (incl. warm redone cold)    */
    int64_t lanczos_warm;
    int64_t cheb_terms;        /* sum over all FD attempts of L                 */
    int64_t cheb_attempts;
    int64_t fallbacks;         /* Chebyshev -> Cholesky fallbacks               */
    int64_t corrector_iters;   /* corrector passes (each one D.x)               */
    int64_t nonconverged;      /* corrector ran out of iterations               */
    int32_t fail_index;        /* pivot index for NOT_POSITIVE_DEFINITE         */
    int32_t pad_;
    double fail_value;         /* pivot value for NOT_POSITIVE_DEFINITE         */
} bdsim_counters;

/* Optional per-step decision log (one entry per step of one trajectory). */
typedef struct bdsim_step_log {
    double t;
    double lmin, lmax;         /* Chebyshev interval (0 when not Chebyshev)    */
    double fd_err;             /* final FD error                                */
    int32_t lanczos_cold;      /* 1 if this step's Lanczos ran cold             */
    int32_t lanczos_steps;     /* Krylov dimension of the accepted run          */
    int32_t cheb_L;            /* final term count                              */
    int32_t cheb_attempts;
    int32_t fallback;
    int32_t corrector_iters;   /* corrector passes this step                    */
} bdsim_step_log;

/* ---- library lifetime -------------------------------------------------- */

/* Selects the device and uploads constant tables. Idempotent. */
int bdsim_init(int device);
/* Waits for all work the library queued, then releases its streams, events
 * and device buffers (resident batches and servers are the caller's to
 * destroy). A later call initialises again. No counterpart in the reference,
 * whose state dies with the process; here device memory and HIP objects are
 * given back before the HIP runtime and any profiler unload. */
int bdsim_shutdown(void);
/* Text of the last API error of the calling thread ("" if none). */
const char* bdsim_last_error(void);
/* Largest supported chain length (beads per trajectory). */
int bdsim_max_beads(void);

/* Register budget of the integrate kernels per block size (64, 128, 256 or
 * 512 threads, i.e. chains of up to that many beads): waves per SIMD, 1 or
 * 4 (and 2 for 64); 0 restores the library's own choice (4 for 64; for the
 * others 4 when a launch holds more waves than 3 per SIMD, else 1). Results do
 * not depend on it. Set by bdsim_amd.autotune from timed runs; no counterpart
 * in the reference.                                                          */
int bdsim_tune_set(int block, int waves_per_simd);
int bdsim_tune_get(int block);           /* current setting, -1 for a bad block */
/* Workgroups per chain for batches of chains of N > 512 that leave compute
 * units idle (the chain split over co-resident workgroups): at most
 * max_workgroups (1 = no split, 2..16), 0 restores the library's choice (as
 * many as fill the compute units, at most 8). Every part holds its own
 * workspace copy (about 8 (9 N^2 + 120 N) bytes). Results do not depend on it
 * beyond rounding of the partial-row sums (each part sums 1/G of the
 * partners). Set by bdsim_amd.autotune; no counterpart in the reference.     */
int bdsim_tune_split(int max_workgroups);
int bdsim_tune_split_get(void);
/* Workgroup width for the mixed form's chains of block size 128, 256 or 512
 * (chains of 65..512 beads): they run on workgroups of block x multiplier
 * threads (1, 2, 4 or 8; at most 1024), the tiled pair walk spread over the
 * extra waves; 0 restores the library's choice (wider while that shortens the
 * walk and the launch stays within one wave per SIMD). The FP64 form keeps
 * its block. Results do not depend on it beyond the fixed-point rounding of
 * the pair sums (2^-43 of the operand). Set by bdsim_amd.autotune; no
 * counterpart in the reference.                                             */
int bdsim_tune_width(int block, int multiplier);
int bdsim_tune_width_get(int block);    /* current setting, -1 for a bad block */

/* ---- the step: batched time_integrate_chain ---------------------------- */

/* Integrates n_traj independent chains, each over its own segment
 * [sim[i].time_start, sim[i].time_end] exactly as time_integrate_chain does
 * (fresh spectral cache per call, accumulated t += dt clock).
 *   nbeads[i]   chain length of trajectory i (2 .. bdsim_max_beads())
 *   R           host, packed: trajectory i occupies 3*nbeads[i] doubles
 *               starting at 3*offset(i), offset = prefix sum of nbeads;
 *               positions in on entry, integrated positions out
 *   phys, sim   host arrays [n_traj]
 *   rng         host [n_traj], advanced in place like Rng& is
 *   precision   BDSIM_PRECISION_FP64 or BDSIM_PRECISION_MIXED
 *   status      host [n_traj] out, BDSIM_STATUS_*
 *   counters    host [n_traj] out, or NULL
 *   log         host, or NULL; trajectory i writes its steps at
 *               log[i*log_stride ...], at most log_stride entries
 * Blocks until the results are back on the host.                            */
int bdsim_integrate_batch(int n_traj, const int32_t* nbeads, double* R,
                          const bdsim_phys* phys, const bdsim_sim* sim,
                          bdsim_rng_state* rng, int precision,
                          int32_t* status, bdsim_counters* counters,
                          bdsim_step_log* log, int log_stride);

/* The same call with no device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Every
 * device buffer the call needs -- staged positions and generator states,
 * per-trajectory parameters, workspaces, step logs, the split chains'
 * barriers and exchange buffers -- is carved from one static device array of
 * bdsim_static_scratch_bytes() (256 MiB) per device, which the library holds
 * from its load until the process ends. Arguments and results are those of
 * bdsim_integrate_batch (the same kernels, the same bits). bdsim_init must
 * have been called (initialisation allocates); a batch that needs more than
 * the array returns BDSIM_ENOMEM and launches nothing. Each call overwrites
 * the array: calls from several threads of one process are serialised by the
 * library's lock, and any work that used the array concurrently would race
 * on it.                                                                     */
int bdsim_integrate_batch_static(int n_traj, const int32_t* nbeads, double* R,
                                 const bdsim_phys* phys, const bdsim_sim* sim,
                                 bdsim_rng_state* rng, int precision,
                                 int32_t* status, bdsim_counters* counters,
                                 bdsim_step_log* log, int log_stride);
/* Size in bytes of the static array bdsim_integrate_batch_static uses. */
size_t bdsim_static_scratch_bytes(void);

/* ---- device-resident batches (for the serving loop and benchmarks) ----- */

/* An opaque batch whose positions, RNG states and workspace stay in HBM
 * between calls. All chains in one batch share (phys, sim) except the flow,
 * which may differ per trajectory (per_traj_kappa), and the RNG seed.        */
typedef struct bdsim_batch bdsim_batch;

/* Creates a device-resident batch. R0 host [n_traj*N*3], rng host [n_traj],
 * kappa host [n_traj*9] or NULL (then phys->flow applies to all).           */
int bdsim_batch_create(bdsim_batch** out, int n_traj, int nbeads,
                       const double* R0, const bdsim_phys* phys,
                       const double* kappa, const bdsim_rng_state* rng,
                       int precision);
/* Runs one integrate segment [sim->time_start, sim->time_end] on every
 * trajectory of the batch on the batch's stream; does not synchronise.     */
int bdsim_batch_run(bdsim_batch* b, const bdsim_sim* sim);
/* Runs the next max_steps steps at most of the segment [sim->time_start,
 * sim->time_end] on every trajectory, on the batch's stream (no
 * synchronisation). resume = 0 starts the segment (cold spectral cache,
 * counters from zero); resume = 1, with the same *sim, continues each
 * trajectory where the previous call stopped (its time, warm Lanczos seed and
 * generator state), so a segment cut into calls gives, bit for bit, the
 * result of one bdsim_batch_run over it (what one time_integrate_chain call,
 * src/integrator.hpp:32-33, does for each chain). Not for batches of split
 * chains (N > 512 spread over several workgroups): BDSIM_EINVAL.            */
int bdsim_batch_run_steps(bdsim_batch* b, const bdsim_sim* sim, int max_steps, int resume);
/* Waits for the batch's stream. */
int bdsim_batch_sync(bdsim_batch* b);
/* Copies positions / rng / status / counters back (any may be NULL). */
int bdsim_batch_fetch(bdsim_batch* b, double* R, bdsim_rng_state* rng,
                      int32_t* status, bdsim_counters* counters);
/* Device-side HIP stream handle (hipStream_t) of the batch, for timing. */
void* bdsim_batch_stream(bdsim_batch* b);
void bdsim_batch_destroy(bdsim_batch* b);

/* ---- serving: jobs admitted into a resident batch while others run ------
 *
 * No counterpart in the reference, whose job API is simulate / run_ensemble
 * (python/bdsim/ensemble.py:179-218,330-379) over one integrate call per
 * trajectory segment. A job is n_traj trajectories of one chain length, each
 * integrating ONE segment [sim->time_start, sim->time_end] from its own R0
 * and rng: the result is identical, bit for bit, to bdsim_integrate_batch
 * on the same job, whatever else is resident. Memory for slots, positions,
 * workspaces and tables is reserved once (hbm_bytes) and handed to queued
 * jobs in arrival order as their footprint fits (a later job may pass one
 * that does not fit yet). Trajectories run in launches of at most
 * steps_per_launch steps (scaled down by BLOCK/64 for longer chains); a job
 * submitted meanwhile joins at its group's next launch. Groups of different
 * chain-length buckets run on their own streams, concurrently.
 * submit may be called from another thread than poll.                       */
typedef struct bdsim_server bdsim_server;

typedef struct bdsim_server_stats {
    int64_t queued, running, active_trajectories;
    int64_t bytes_reserved, bytes_in_use;   /* positions, workspaces, tables */
    int64_t bead_steps_done;                /* over completed jobs           */
} bdsim_server_stats;

int bdsim_server_create(bdsim_server** out, int precision, int64_t hbm_bytes,
                        int steps_per_launch);
/* Queues a job (inputs are copied). BDSIM_ENOMEM if it could never fit. */
int bdsim_server_submit(bdsim_server* s, int64_t job_id, int n_traj, int nbeads,
                        const double* R0, const bdsim_phys* phys,
                        const bdsim_sim* sim, const bdsim_rng_state* rng);
/* Services every group whose launch has finished (collect, admit, launch
 * again), waiting up to wait_us for a job to complete. Returns the number of
 * completed jobs not yet taken by bdsim_server_completed, or < 0 on error. */
int bdsim_server_poll(bdsim_server* s, int wait_us);
/* Takes up to max_ids completed job ids, oldest first; returns the count. */
int bdsim_server_completed(bdsim_server* s, int64_t* ids, int max_ids);
/* Results of a completed job (any pointer may be NULL), then forgets it:
 * R_out [n_traj*N*3], rng_out / status / counters [n_traj].                 */
int bdsim_server_result(bdsim_server* s, int64_t job_id, double* R_out,
                        bdsim_rng_state* rng_out, int32_t* status,
                        bdsim_counters* counters);
int bdsim_server_query(bdsim_server* s, bdsim_server_stats* stats);
void bdsim_server_destroy(bdsim_server* s);

/* ---- forces and pieces, batched over chains of one length -------------- */

/* R host [n_traj*N*3] -> F host [n_traj*N*3]. */
int bdsim_total_force_batch(int n_traj, int nbeads, const double* R,
                            const bdsim_phys* phys, double* F);
int bdsim_spring_force_batch(int n_traj, int nbeads, const double* R,
                             const bdsim_phys* phys, double* F);
int bdsim_bending_force_batch(int n_traj, int nbeads, const double* R,
                              const bdsim_phys* phys, double* F);
int bdsim_ev_force_batch(int n_traj, int nbeads, const double* R,
                         const bdsim_phys* phys, double* F);
/* External force of phys on nbeads beads at time t -> F host [nbeads*3]. */
int bdsim_external_force_batch(const bdsim_phys* phys, int nbeads, double t,
                               double* F);
/* n_steps consecutive raw_noise fields per trajectory:
 * rng host [n_traj] in/out, X host [n_traj*n_steps*nbeads*3].               */
int bdsim_raw_noise_batch(int n_traj, int nbeads, int n_steps,
                          bdsim_rng_state* rng, double* X);
/* y = D(R).x for each chain (RPY with hstar; identity when hstar <= 0). */
int bdsim_rpy_apply_batch(int n_traj, int nbeads, double hstar, int precision,
                          const double* R, const double* x, double* y);
/* y = B(R).x for each chain: the Brownian operator of phys's method
 * (Diffusion::brownian_step, mobility.cpp:58-88, without the sqrt(dt)
 * factor): x when hstar <= 0, L x for Cholesky, the Chebyshev series of
 * sqrt(D) x on cold Lanczos bounds (hydrodynamics.cpp:317-337). One physics
 * for all chains; status [n_traj] per chain. For checks of the forms.        */
int bdsim_brownian_batch(int n_traj, int nbeads, const double* R,
                         const bdsim_phys* phys, int precision,
                         const double* x, double* y, int32_t* status);

/* ---- snapshot observables ----------------------------------------------
 * One record per chain, evaluated on the device where the positions are:
 * what python/bdsim/properties.py:5-52 computes from a snapshot
 * (end_to_end_sq, gyration_tensor, radius_of_gyration_sq, stretch,
 * kramers_stress) together with the force _bdsim.total_force returns
 * (bindings/module.cpp:67-76) or one of its parts, so that ensemble averages
 * need no copy of the positions to the host (the snapshots of
 * python/bdsim/ensemble.py:67-83 reduced to what rheology and the
 * equilibrium checks read).                                                 */
enum { BDSIM_FORCE_TOTAL = 0,   /* spring + EV + bending                     */
       BDSIM_FORCE_SPRING = 1,
       BDSIM_FORCE_BENDING = 2,
       BDSIM_FORCE_EV = 3 };

typedef struct bdsim_observables {
    double rg2;            /* trace of the gyration tensor                   */
    double q2;             /* |R_N - R_1|^2                                  */
    double gyration[9];    /* (1/N) sum (R-Rc)(R-Rc)^T, row-major            */
    double stretch[3];     /* max - min of the centred chain per axis        */
    double stress[9];      /* tau_ij = sum_mu (R_mu - R_c)_i F_mu,j          */
    double com[3];         /* centre of mass R_c                             */
} bdsim_observables;

/* R host [n_traj*N*3] (one PhysParams for all chains) -> out [n_traj]. */
int bdsim_observables_batch(int n_traj, int nbeads, const double* R,
                            const bdsim_phys* phys, int force_kind,
                            bdsim_observables* out);
/* The same for a resident batch's current positions (after the runs queued
 * on its stream); only the records cross to the host.                        */
int bdsim_batch_observe(bdsim_batch* b, int force_kind, bdsim_observables* out);

#ifdef __cplusplus
}
#endif

#endif /* BDSIM_GPU_H */
