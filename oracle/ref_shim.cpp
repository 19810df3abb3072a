// ref_shim.cpp -- C ABI over the reference bdsim C++ core, for the test suite.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile together with the
// reference sources where they lie (IsaacPincus/bdsim src/*.cpp) into
// oracle/_ref/libbdref.so.  It lets the tests check the C restatement
// (oracle/bdsim_oracle.c) and write golden fixtures against the reference
// itself.  Nothing on the product path loads it.
//
// Parameter PODs are those of include/bdsim_gpu.h; to_phys() converts them to
// the reference's PhysParams (src/model.hpp:20-35).
#include "../include/bdsim_gpu.h"

#include "hydrodynamics.hpp"
#include "integrator.hpp"
#include "linalg.hpp"
#include "model.hpp"
#include "spring.hpp"

#include <cstring>
#include <stdexcept>
#include <string>

using namespace bdsim;

namespace {

thread_local std::string g_err;

Vec3Field to_field(int N, const double* R) {
    Vec3Field f(N);
    for (int i = 0; i < N; ++i) f[i] = {R[3 * i], R[3 * i + 1], R[3 * i + 2]};
    return f;
}

void from_field(const Vec3Field& f, double* out) {
    for (size_t i = 0; i < f.size(); ++i)
        for (int d = 0; d < 3; ++d) out[3 * i + d] = f[i][d];
}

Mat3 to_mat3(const double* k) {
    Mat3 m{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = k[3 * i + j];
    return m;
}

PhysParams to_phys(const bdsim_phys* p, int N) {
    PhysParams q;
    q.spring.type = static_cast<Spring>(p->spring_type);
    q.spring.sqrtb = p->sqrtb;
    q.spring.natural_length = p->natural_length;
    q.ev.type = static_cast<EV>(p->ev_type);
    q.ev.zstar = p->zstar;
    q.ev.dstar = p->dstar;
    q.ev.min_cutoff = p->min_cutoff;
    q.ev.max_cutoff = p->max_cutoff;
    q.ev.contour_dist_for_EV = p->contour_dist_for_ev;
    if (p->n_phi > 0) q.ev.phi.assign(p->phi, p->phi + p->n_phi);
    q.bend.type = static_cast<Bending>(p->bend_type);
    q.bend.stiffness = p->bend_stiffness;
    if (p->flow.constant || p->flow.n_times == 0) {
        q.flow = Flow::constant(to_mat3(p->flow.kappa0));
    } else {
        std::vector<dp> times(p->flow.times, p->flow.times + p->flow.n_times);
        std::vector<Mat3> Ks(p->flow.n_times);
        for (int t = 0; t < p->flow.n_times; ++t) Ks[t] = to_mat3(p->flow.kappas + 9 * t);
        q.flow = Flow::time_varying(times, Ks);
    }
    for (int k = 0; k < p->n_external; ++k) {
        const bdsim_bead_force& e = p->external[k];
        if (e.constant) {
            q.external.add_constant(e.bead, Vec3{e.value[0], e.value[1], e.value[2]});
        } else {
            std::vector<dp> times(e.times, e.times + e.n_times);
            std::vector<Vec3> vals(e.n_times);
            for (int j = 0; j < e.n_times; ++j)
                vals[j] = {e.values[3 * j], e.values[3 * j + 1], e.values[3 * j + 2]};
            q.external.add_time_varying(e.bead, times, vals);
        }
    }
    q.number_of_beads = N;
    q.hstar = p->hstar;
    q.hi_method = static_cast<DelSMethod>(p->hi_method);
    q.ncheb_multiplier = p->ncheb_multiplier;
    q.fd_err_max = p->fd_err_max;
    q.equilibration = p->equilibration != 0;
    return q;
}

SimParams to_sim(const bdsim_sim* s) {
    SimParams q;
    q.time_start = s->time_start;
    q.time_end = s->time_end;
    q.dt = s->dt;
    q.implicit_loop_tol = s->implicit_loop_tol;
    q.update_center_of_mass = s->update_center_of_mass != 0;
    return q;
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

void ref_ran1(bdsim_rng_state* st, int n, double* x) {
    Rng r;
    r.restore({st->ix, st->iy, st->seed});
    r.ran_1(n, x);
    const Rng::State s = r.save();
    st->ix = s.ix; st->iy = s.iy; st->seed = s.seed;
}

// bindings/module.cpp:45-50 -- returns 0, or 1 when the core threw.
int ref_integrate(int N, double* R, const bdsim_phys* p, const bdsim_sim* s,
                  bdsim_rng_state* st) {
    Vec3Field f = to_field(N, R);
    PhysParams phys = to_phys(p, N);
    SimParams sim = to_sim(s);
    Rng r;
    r.restore({st->ix, st->iy, st->seed});
    int rc = 0;
    try {
        time_integrate_chain(f, phys, sim, r);
        from_field(f, R);
    } catch (const std::exception& e) {
        g_err = e.what();
        rc = 1;
    }
    const Rng::State o = r.save();
    st->ix = o.ix; st->iy = o.iy; st->seed = o.seed;
    return rc;
}

void ref_spring_force(int N, const double* R, const bdsim_phys* p, double* F) {
    const Vec3Field f = to_field(N, R);
    from_field(spring_force(chain_geometry(f), to_phys(p, N).spring), F);
}

void ref_bending_force(int N, const double* R, const bdsim_phys* p, double* F) {
    const Vec3Field f = to_field(N, R);
    const ChainGeometry g = chain_geometry(f);
    from_field(BendingForce(to_phys(p, N).bend).force({f, g}), F);
}

void ref_ev_force(int N, const double* R, const bdsim_phys* p, double* F) {
    const Vec3Field f = to_field(N, R);
    const ChainGeometry g = chain_geometry(f);
    ExcludedVolume ev(to_phys(p, N).ev, p->equilibration != 0, N);
    from_field(ev.force({f, g}), F);
}

// bindings/module.cpp:67-76
void ref_total_force(int N, const double* R, const bdsim_phys* p, double* F) {
    const Vec3Field f = to_field(N, R);
    PhysicalModel model(to_phys(p, N));
    const ChainGeometry g = chain_geometry(f);
    Vec3Field out = spring_force(g, model.spring());
    const Vec3Field Fe = model.non_spring_force({f, g});
    for (size_t i = 0; i < out.size(); ++i) out[i] = out[i] + Fe[i];
    from_field(out, F);
}

void ref_external_force(const bdsim_phys* p, int N, double t, double* F) {
    from_field(PhysicalModel(to_phys(p, N)).external_force(t), F);
}

void ref_kappa(const bdsim_phys* p, double t, double* K) {
    const Mat3 m = PhysicalModel(to_phys(p, 2)).kappa(t);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K[3 * i + j] = m[i][j];
}

void ref_rpy_matrix(int N, const double* R, double hstar, double* D) {
    const Matrix M = rpy_diffusion_matrix(to_field(N, R), hstar);
    std::memcpy(D, M.a.data(), sizeof(double) * M.a.size());
}

// returns 0, or 1 when cholesky_lower threw
int ref_cholesky(int n, const double* M, double* L) {
    Matrix A(n);
    std::memcpy(A.a.data(), M, sizeof(double) * n * n);
    try {
        const Matrix C = cholesky_lower(A);
        std::memcpy(L, C.a.data(), sizeof(double) * n * n);
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
    return 0;
}

// Cold (no cache) Chebyshev square root: hydrodynamics.cpp:317-337.
void ref_chebyshev_sqrt_times(int N, const double* D, const double* x0, double mult,
                              double tol, double* out, int* nterms, double* fd) {
    Matrix A(3 * N);
    std::memcpy(A.a.data(), D, sizeof(double) * 9 * N * N);
    const ChebyshevResult cr = chebyshev_sqrt_times(A, to_field(N, x0), mult, tol, nullptr);
    from_field(cr.value, out);
    *nterms = cr.nterms;
    *fd = cr.fd_error;
}

// A sequence of spectral_bounds calls sharing one SpectralCache (warm start):
// D holds n_mat matrices of size n x n; bounds out[2*k], out[2*k+1].
void ref_spectral_bounds_seq(int n, int n_mat, const double* D, double* out,
                             long* cold_runs, long* warm_runs) {
    SpectralCache cache;
    for (int k = 0; k < n_mat; ++k) {
        Matrix A(n);
        std::memcpy(A.a.data(), D + (size_t)k * n * n, sizeof(double) * n * n);
        const auto b = spectral_bounds(A, &cache);
        out[2 * k] = b.first;
        out[2 * k + 1] = b.second;
    }
    *cold_runs = cache.cold_runs;
    *warm_runs = cache.warm_runs;
}

void ref_solve_implicit_r(int type, double dtby4, double gama, double natscl, double* r,
                          double* ff) {
    const SpringSolution s = solve_implicit_r(static_cast<Spring>(type), dtby4, gama, natscl);
    *r = s.length;
    *ff = s.force_factor;
}

void ref_solve_connector(const double* gamma, double dt, int type, double sqrtb,
                         double natural_length, double* out) {
    SpringParams sp{static_cast<Spring>(type), sqrtb, natural_length};
    const Vec3 q = solve_connector(Vec3{gamma[0], gamma[1], gamma[2]}, dt, sp);
    out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
}

}  // extern "C"
