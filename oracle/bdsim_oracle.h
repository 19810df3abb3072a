/*
 * bdsim_oracle.h -- FP64 CPU restatement of bdsim's ChainIntegrator step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity oracle for the GPU
 * library: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it, and only as the checker (or the timed CPU baseline), never on the
 * product path.  Each function names the reference file:line it restates.
 *
 * Parameter PODs are the ones of include/bdsim_gpu.h, so a test fills one
 * struct and hands it to both implementations.
 */
#ifndef BDSIM_ORACLE_H
#define BDSIM_ORACLE_H

#include "../include/bdsim_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* rng.cpp:31-70 -- n uniform deviates, advancing the state. */
void bo_ran1(bdsim_rng_state* r, int n, double* x);
/* mobility.cpp:18-29 -- one uncorrelated noise field X0 (N x 3). */
void bo_raw_noise(bdsim_rng_state* r, int N, double* X);

/* Forces on (N x 3) positions. */
void bo_spring_force(int N, const double* R, const bdsim_phys* p, double* F);
void bo_bending_force(int N, const double* R, const bdsim_phys* p, double* F);
void bo_ev_force(int N, const double* R, const bdsim_phys* p, double* F);
void bo_total_force(int N, const double* R, const bdsim_phys* p, double* F);
void bo_external_force(const bdsim_phys* p, int N, double t, double* F);
/* flow.cpp:5-26 -- kappa(t), row-major. */
void bo_kappa(const bdsim_flow* f, double t, double K[9]);

/* Dense mobility pieces (n = 3N). */
void bo_rpy_matrix(int N, const double* R, double hstar, double* D);
void bo_matvec(int n, const double* M, const double* x, double* y);
/* returns 0, or -1 with the failing pivot index/value */
int bo_cholesky_lower(int n, const double* M, double* L, int* pivot, double* value);
void bo_lower_matvec(int n, const double* L, const double* x, double* y);

/* hydrodynamics.cpp:224-264.  seed: in/out warm-start vector [n];
 * *seed_valid in/out (0 = empty cache).  Returns the padded interval.        */
void bo_spectral_bounds(int n, const double* D, double* seed, int* seed_valid,
                        double* lmin, double* lmax, int* steps, int* cold);
/* hydrodynamics.cpp:317-337 with a cold (no cache) or warm cache. */
void bo_chebyshev_sqrt_times(int n, const double* D, const double* x0, double mult,
                             double fd_tol, double* seed, int* seed_valid,
                             double* out, int* nterms, double* fd_err);

/* spring.cpp:123-172 -- returns a BDSIM_STATUS_* code. */
int bo_solve_implicit_r(int type, double dtby4, double gama, double natscl,
                        double* r, double* ff);
/* spring.cpp:174-181 */
int bo_solve_connector(const double gamma[3], double dt, int type, double sqrtb,
                       double natural_length, double out[3]);

/* integrator.cpp:61-177 -- one integrate call on one chain.  Returns a
 * BDSIM_STATUS_* code; counters/log may be NULL.                             */
int bo_integrate(int N, double* R, const bdsim_phys* phys, const bdsim_sim* sim,
                 bdsim_rng_state* rng, bdsim_counters* cnt,
                 bdsim_step_log* log, int log_cap);

/* Many chains in a row (same layout as bdsim_integrate_batch). */
int bo_integrate_batch(int n_traj, const int32_t* nbeads, double* R,
                       const bdsim_phys* phys, const bdsim_sim* sim,
                       bdsim_rng_state* rng, int32_t* status,
                       bdsim_counters* counters);

#ifdef __cplusplus
}
#endif
#endif
