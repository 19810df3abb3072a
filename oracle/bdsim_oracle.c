/*
 * bdsim_oracle.c -- FP64 CPU restatement of bdsim's ChainIntegrator step.
 *
 * TEST INFRASTRUCTURE ONLY (see bdsim_oracle.h).  The arithmetic follows the
 * reference expression by expression -- same association, same sequential sums,
 * same libm calls -- so that, built with gcc at -O2 for baseline x86-64 (no FMA
 * contraction available), it reproduces the reference C++ core bit for bit.
 * tests/test_oracle.py checks that against the golden arrays written from
 * oracle/_ref (test_trajectories_golden_bit_exact, test_forces_golden_bit_exact).
 *
 * Citations are to IsaacPincus/bdsim.
 */
#include "bdsim_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BO_PI 3.14159265358979323846
#define BO_MYEPS 1e-6
#define BO_MAXCHEB 500
#define BO_LANCZOS_COLD 30
#define BO_LANCZOS_WARM 8

/* ======================================================================= */
/* RNG: Numerical Recipes ran_1 (rng.cpp:24-70)                              */
/* ======================================================================= */

/* rng.cpp:24-28: am = nearest(1.0,-1.0)/real(IM) in single precision. */
static float bo_am(void) { return nextafterf(1.0f, -1.0f) / (float)2147483647; }

/* rng.cpp:31-39 */
static void bo_rng_maybe_init(bdsim_rng_state* r) {
    if (r->seed <= 0 || r->iy < 0) {
        const int32_t s = abs(r->seed);
        r->iy = (888889999 ^ s) | 1;
        r->ix = 777755555 ^ s;
        r->seed = s + 1;
    }
}

/* rng.cpp:41-70 */
void bo_ran1(bdsim_rng_state* r, int n, double* x) {
    bo_rng_maybe_init(r);
    uint32_t ux = (uint32_t)r->ix;
    const float am = bo_am();
    for (int c = 0; c < n; ++c) {
        ux ^= (ux << 13);
        ux ^= (ux >> 17);
        ux ^= (ux << 5);
        const int32_t k = r->iy / 127773;
        r->iy = 16807 * (r->iy - k * 127773) - 2836 * k;
        if (r->iy < 0) r->iy += 2147483647;
        const uint32_t mixed = ux ^ (uint32_t)r->iy;
        const int32_t result = (int32_t)((mixed & 0x7fffffffu) | 1u);
        x[c] = (double)(am * (float)result);
    }
    r->ix = (int32_t)ux;
}

/* mobility.cpp:14-29: Ottinger deviate X = (c1 y^2 + c2) y, y = u - 1/2. */
void bo_raw_noise(bdsim_rng_state* r, int N, double* X) {
    const double c1 = 14.14855378, c2 = 1.21569221;
    double u[3];
    for (int mu = 0; mu < N; ++mu) {
        bo_ran1(r, 3, u);
        for (int d = 0; d < 3; ++d) {
            const double y = u[d] - 0.5;
            X[3 * mu + d] = (c1 * y * y + c2) * y;
        }
    }
}

/* ======================================================================= */
/* small vector helpers (vec.hpp)                                            */
/* ======================================================================= */

static double n2(const double* a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static double dotn(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* ======================================================================= */
/* geometry (config.cpp:8-45)                                                */
/* ======================================================================= */

static double clamped_length(const double* b) {
    double sq = n2(b);
    if (sq < 1.0e-12) sq = 1.0e-12;
    return sqrt(sq);
}

/* config.cpp:15-25: bonds b_i = R_{i+1} - R_i and clamped lengths. */
static void chain_geometry(int N, const double* R, double* bond, double* len) {
    for (int i = 0; i < N - 1; ++i) {
        for (int d = 0; d < 3; ++d) bond[3 * i + d] = R[3 * (i + 1) + d] - R[3 * i + d];
        len[i] = clamped_length(bond + 3 * i);
    }
}

/* ======================================================================= */
/* springs (spring.cpp:14-207, numerics.cpp, numerics.hpp:24-73)             */
/* ======================================================================= */

/* spring.cpp:14-37 */
static double force_sans_hookean(int type, double q, double q0) {
    switch (type) {
        case BDSIM_SPRING_HOOK: return 1.0;
        case BDSIM_SPRING_FENE: return 1.0 / (1.0 - q * q);
        case BDSIM_SPRING_ILC: return (3.0 - q * q) / (1.0 - q * q) / 3.0;
        case BDSIM_SPRING_WLC:
            return 1.0 / (6.0 * q) * (4.0 * q + 1.0 / (1.0 - q) / (1.0 - q) - 1.0);
        case BDSIM_SPRING_FRAENKEL: return 1.0 * (1.0 - (q0 / q));
        case BDSIM_SPRING_FENE_FRAENKEL:
            return (1.0 - q0 / q) / (1.0 - (q - q0) * (q - q0));
        case BDSIM_SPRING_WLC_BOUNDED: {
            const double t = (q - q0) / (1.0 - q0);
            return (2.0 / (3.0 * q)) *
                   ((pow(1.0 - t, -2.0) - 1.0) / 4.0 + t -
                    q0 * (pow(1.0 + t, -2.0) - 1.0) / 4.0 + q0 * t);
        }
    }
    return 1.0;
}

/* spring.cpp:183-193 */
static void connector_forces(int nb, const double* bond, const double* len, int type,
                             double sqrtb, double q0, double* Fb) {
    for (int i = 0; i < nb; ++i) {
        double r = len[i] / sqrtb;
        if (fabs(r - 1.0) < BO_MYEPS) r = 1.0 - BO_MYEPS;
        const double f = force_sans_hookean(type, r, q0);
        for (int d = 0; d < 3; ++d) Fb[3 * i + d] = f * bond[3 * i + d];
    }
}

/* spring.cpp:195-203 */
static void bead_forces_from_connectors(int N, const double* Fb, double* F) {
    memset(F, 0, sizeof(double) * 3 * N);
    if (N == 1) return;
    for (int d = 0; d < 3; ++d) F[d] = Fb[d];
    for (int nu = 1; nu < N - 1; ++nu)
        for (int d = 0; d < 3; ++d) F[3 * nu + d] = Fb[3 * nu + d] - Fb[3 * (nu - 1) + d];
    for (int d = 0; d < 3; ++d) F[3 * (N - 1) + d] = (-1.0) * Fb[3 * (N - 2) + d];
}

/* numerics.cpp:8-23 */
static double cubic_at(const double* c, double x) { return ((c[3] * x + c[2]) * x + c[1]) * x + c[0]; }
static double bisect_cubic(const double* c, double lo, double hi) {
    double flo = cubic_at(c, lo), fhi = cubic_at(c, hi);
    if (flo == 0.0) return lo;
    if (fhi == 0.0) return hi;
    if (flo * fhi > 0.0) return (fabs(flo) < fabs(fhi)) ? lo : hi;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double fm = cubic_at(c, mid);
        if (fm == 0.0 || (hi - lo) < 1e-15) return mid;
        if (flo * fm < 0.0) { hi = mid; fhi = fm; }
        else { lo = mid; flo = fm; }
    }
    (void)fhi;
    return 0.5 * (lo + hi);
}

/* numerics.cpp:25-59 */
static double find_roots_cubic(const double* coeff, double lower, double upper) {
    const double a = coeff[2] / coeff[3];
    const double b = coeff[1] / coeff[3];
    const double c = coeff[0] / coeff[3];
    const double Q = (a * a - 3.0 * b) / 9.0;
    const double R = (2.0 * a * a * a - 9.0 * a * b + 27.0 * c) / 54.0;
    double x;
    if (R * R < Q * Q * Q) {
        const double theta = acos(R / sqrt(Q * Q * Q));
        for (int i = -1; i <= 1; ++i) {
            x = -2.0 * sqrt(Q) * cos((theta + (double)i * BO_PI * 2.0) / 3.0) - a / 3.0;
            if (x >= lower && x <= upper) return x;
        }
        x = -2.0 * sqrt(Q) * cos((theta + BO_PI * 2.0) / 3.0) - a / 3.0;
    } else {
        const double Au = -copysign(1.0, R) * cbrt(fabs(R) + sqrt(R * R - Q * Q * Q));
        const double Bu = (Au == 0.0) ? 0.0 : Q / Au;
        x = (Au + Bu) - a / 3.0;
    }
    if (x >= lower && x <= upper) return x;
    return bisect_cubic(coeff, lower, upper);
}

/* numerics.cpp:61-85 */
static double polish_poly_root(const double* c, double x0, double atol) {
    double x = x0;
    for (int iter = 0; iter < 30; ++iter) {
        double p = c[3] * x + c[2];
        double p1 = c[3];
        for (int i = 1; i >= 0; --i) {
            p1 = p + p1 * x;
            p = c[i] + p * x;
        }
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

/* spring.cpp:55-87 */
static void cubic_coefficients(int type, double dtby4, double gama, double natscl, double* c) {
    c[0] = c[1] = c[2] = 0.0;
    c[3] = 1.0;
    switch (type) {
        case BDSIM_SPRING_FENE:
            c[0] = gama; c[1] = -(1.0 + dtby4); c[2] = -gama;
            break;
        case BDSIM_SPRING_ILC: {
            const double d = 3.0 + dtby4;
            c[0] = 3.0 * gama / d; c[1] = -(3.0 + 3.0 * dtby4) / d; c[2] = -3.0 * gama / d;
            break;
        }
        case BDSIM_SPRING_WLC: {
            const double d = 2.0 * (1.5 + dtby4);
            c[0] = -3.0 * gama / d;
            c[1] = 3.0 * (1.0 + dtby4 + 2.0 * gama) / d;
            c[2] = -1.5 * (4.0 + 3.0 * dtby4 + 2.0 * gama) / d;
            break;
        }
        case BDSIM_SPRING_FENE_FRAENKEL:
            c[0] = dtby4 * natscl + gama * (1.0 - natscl * natscl);
            c[1] = -1.0 + natscl * natscl - dtby4 + 2.0 * gama * natscl;
            c[2] = -(2.0 * natscl + gama);
            break;
        default: break;
    }
}

/* spring.cpp:90-97 */
static double asymptotic_guess(int type, double dtby4, double gama) {
    switch (type) {
        case BDSIM_SPRING_FENE: return 1.0 - dtby4 / 2.0 / gama;
        case BDSIM_SPRING_ILC: return 1.0 - dtby4 / 3.0 / gama;
        case BDSIM_SPRING_WLC: return 1.0 - sqrt(dtby4 / 6.0 / gama);
        default: return gama / (1.0 + dtby4);
    }
}

/* spring.cpp:101-119 with rtsafe of numerics.hpp:24-73. */
static void wlcb_fd(double dt, double s, double gama, double x, double* f, double* df) {
    *f = 6.0 * x / dt + 0.25 * (1.0 - s) * (1.0 - s) / ((1.0 - x) * (1.0 - x)) -
         0.25 + (x - s) / (1.0 - s) -
         s / 4.0 * (1.0 - s) * (1.0 - s) / ((1.0 + x - 2.0 * s) * (1.0 + x - 2.0 * s)) +
         s / 4.0 + s * (x - s) / (1.0 - s) - 6.0 * gama / dt;
    *df = 6.0 / dt + 1.0 / (1.0 - s) + s / (1.0 - s) +
          (1.0 - s) * (1.0 - s) / (2.0 * pow(1.0 - x, 3.0)) +
          ((1.0 - s) * (1.0 - s) * s) / (2.0 * pow(1.0 - 2.0 * s + x, 3.0));
}

static int solve_wlc_bounded(double dtby4, double gama, double sigma, double* root) {
    if (sigma > 1.0) return BDSIM_STATUS_WLCB_SIGMA;
    const double dt = dtby4 * 4.0;
    const double eps = DBL_EPSILON;
    const double lo = (0.0 < 2.0 * sigma - 1.0 + eps) ? 2.0 * sigma - 1.0 + eps : 0.0;
    const double x1 = lo, x2 = 1.0 - eps, xacc = 1e-15;
    double fl, fh, df;
    wlcb_fd(dt, sigma, gama, x1, &fl, &df);
    wlcb_fd(dt, sigma, gama, x2, &fh, &df);
    if ((fl > 0.0 && fh > 0.0) || (fl < 0.0 && fh < 0.0)) return BDSIM_STATUS_RTSAFE_UNBRACKETED;
    if (fl == 0.0) { *root = x1; return BDSIM_STATUS_OK; }
    if (fh == 0.0) { *root = x2; return BDSIM_STATUS_OK; }
    double xl, xh;
    if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
    double rts = 0.5 * (x1 + x2);
    double dxold = fabs(x2 - x1);
    double dx = dxold;
    double f;
    wlcb_fd(dt, sigma, gama, rts, &f, &df);
    for (int j = 0; j < 10000; ++j) {
        if (((rts - xh) * df - f) * ((rts - xl) * df - f) >= 0.0 ||
            fabs(2.0 * f) > fabs(dxold * df)) {
            dxold = dx;
            dx = 0.5 * (xh - xl);
            rts = xl + dx;
            if (xl == rts) { *root = rts; return BDSIM_STATUS_OK; }
        } else {
            dxold = dx;
            dx = f / df;
            const double temp = rts;
            rts -= dx;
            if (temp == rts) { *root = rts; return BDSIM_STATUS_OK; }
        }
        if (fabs(dx) < xacc) { *root = rts; return BDSIM_STATUS_OK; }
        wlcb_fd(dt, sigma, gama, rts, &f, &df);
        if (f < 0.0) xl = rts; else xh = rts;
    }
    return BDSIM_STATUS_RTSAFE_MAXIT;
}

/* spring.cpp:123-172 */
int bo_solve_implicit_r(int type, double dtby4, double gama, double natscl,
                        double* r_out, double* ff_out) {
    double r;
    switch (type) {
        case BDSIM_SPRING_HOOK:
            *r_out = gama / (1.0 + dtby4); *ff_out = 1.0;
            return BDSIM_STATUS_OK;
        case BDSIM_SPRING_FRAENKEL:
            r = (gama + dtby4 * natscl) / (1.0 + dtby4);
            *r_out = r; *ff_out = 1.0 - natscl / r;
            return BDSIM_STATUS_OK;
        case BDSIM_SPRING_WLC_BOUNDED: {
            const int st = solve_wlc_bounded(dtby4, gama, natscl, &r);
            if (st != BDSIM_STATUS_OK) { *r_out = NAN; *ff_out = NAN; return st; }
            *r_out = r; *ff_out = force_sans_hookean(type, r, natscl);
            return BDSIM_STATUS_OK;
        }
        default: break;
    }
    double c[4];
    cubic_coefficients(type, dtby4, gama, natscl, c);
    if (type == BDSIM_SPRING_FENE_FRAENKEL) {
        const double lo = (natscl - 1.0 < 0.0) ? 0.0 : natscl - 1.0, hi = natscl + 1.0;
        r = find_roots_cubic(c, lo, hi);
        r = polish_poly_root(c, r, 1e-14);
        const double margin = 1e-6;
        /* std::min(std::max(r, lo + margin), hi - margin), NaN-faithful */
        const double t = (r < lo + margin) ? lo + margin : r;
        r = (hi - margin < t) ? hi - margin : t;
    } else {
        r = (gama < 1.0) ? gama / (1.0 + dtby4) : asymptotic_guess(type, dtby4, gama);
        if (!(type == BDSIM_SPRING_WLC && gama > 100.0)) r = polish_poly_root(c, r, 1e-14);
        const double margin = 1e-6;
        const double t = (r < 0.0) ? 0.0 : r;
        r = (1.0 - margin < t) ? 1.0 - margin : t;
    }
    *r_out = r;
    *ff_out = force_sans_hookean(type, r, natscl);
    return BDSIM_STATUS_OK;
}

/* spring.cpp:174-181 */
int bo_solve_connector(const double gamma[3], double dt, int type, double sqrtb,
                       double natural_length, double out[3]) {
    const double q0 = natural_length / sqrtb;
    const double gmag = sqrt(n2(gamma));
    double r, ff;
    const int st = bo_solve_implicit_r(type, dt / 4.0, gmag / sqrtb, q0, &r, &ff);
    if (fabs(r - 1.0) < 1e-6) r = 1.0 - 1e-6;
    const double s = r * sqrtb / gmag;
    for (int d = 0; d < 3; ++d) out[d] = s * gamma[d];
    return st;
}

void bo_spring_force(int N, const double* R, const bdsim_phys* p, double* F) {
    double* bond = (double*)malloc(sizeof(double) * 3 * (N > 1 ? N - 1 : 1));
    double* len = (double*)malloc(sizeof(double) * (N > 1 ? N - 1 : 1));
    double* Fb = (double*)malloc(sizeof(double) * 3 * (N > 1 ? N - 1 : 1));
    chain_geometry(N, R, bond, len);
    connector_forces(N - 1, bond, len, p->spring_type, p->sqrtb,
                     p->natural_length / p->sqrtb, Fb);
    bead_forces_from_connectors(N, Fb, F);
    free(bond); free(len); free(Fb);
}

/* ======================================================================= */
/* bending (bending.cpp:5-43, config.cpp:35-45)                              */
/* ======================================================================= */

static void bend_term(double s, const double* a, double ca, const double* b, double* out) {
    /* out = s * (ca * a - b) */
    for (int d = 0; d < 3; ++d) out[d] = s * (ca * a[d] - b[d]);
}
static void bend_term2(double s, double ca, const double* a, const double* b, double* out) {
    /* out = s * ((-ca) * a + b) */
    for (int d = 0; d < 3; ++d) out[d] = s * ((-ca) * a[d] + b[d]);
}

static void bending_from_geometry(int N, const double* bond, const double* len,
                                  int type, double C, double* F) {
    memset(F, 0, sizeof(double) * 3 * N);
    if (type == BDSIM_BENDING_NONE) return;
    if (N == 2) return;
    double* u = (double*)malloc(sizeof(double) * 3 * (N - 1));
    double* cs = (double*)malloc(sizeof(double) * (N - 2));
    for (int i = 0; i < N - 1; ++i)
        for (int d = 0; d < 3; ++d) u[3 * i + d] = (1.0 / len[i]) * bond[3 * i + d];
    for (int i = 0; i + 1 < N - 1; ++i) {
        const double* a = u + 3 * (i + 1);
        const double* b = u + 3 * i;
        cs[i] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    }
#define U(k) (u + 3 * (k))
    double A[3], B[3], Cc[3], D[3];
    bend_term(C / len[0], U(0), cs[0], U(1), F);                       /* F[0]   */
    bend_term2(C / len[N - 2], cs[N - 3], U(N - 2), U(N - 3), F + 3 * (N - 1));
    if (N == 3) {
        bend_term(1.0 / len[1], U(1), cs[0], U(0), A);
        bend_term2(1.0 / len[0], cs[0], U(0), U(1), B);
        for (int d = 0; d < 3; ++d) F[3 + d] = C * (A[d] + B[d]);
        free(u); free(cs);
        return;
    }
    bend_term(1.0 / len[1], U(1), cs[0], U(0), A);
    bend_term2(1.0 / len[0], cs[0], U(0), U(1), B);
    bend_term(1.0 / len[1], U(1), cs[1], U(2), D);
    for (int d = 0; d < 3; ++d) F[3 + d] = C * (A[d] + B[d] + D[d]);

    bend_term(1.0 / len[N - 2], U(N - 2), cs[N - 3], U(N - 3), A);
    bend_term2(1.0 / len[N - 3], cs[N - 3], U(N - 3), U(N - 2), B);
    bend_term2(1.0 / len[N - 3], cs[N - 4], U(N - 3), U(N - 4), Cc);
    for (int d = 0; d < 3; ++d) F[3 * (N - 2) + d] = C * (A[d] + B[d] + Cc[d]);

    for (int m = 2; m <= N - 3; ++m) {
        bend_term(1.0 / len[m], U(m), cs[m - 1], U(m - 1), A);
        bend_term2(1.0 / len[m - 1], cs[m - 1], U(m - 1), U(m), B);
        bend_term2(1.0 / len[m - 1], cs[m - 2], U(m - 1), U(m - 2), Cc);
        bend_term(1.0 / len[m], U(m), cs[m], U(m + 1), D);
        for (int d = 0; d < 3; ++d) F[3 * m + d] = C * (A[d] + B[d] + Cc[d] + D[d]);
    }
#undef U
    free(u); free(cs);
}

void bo_bending_force(int N, const double* R, const bdsim_phys* p, double* F) {
    double* bond = (double*)malloc(sizeof(double) * 3 * (N > 1 ? N - 1 : 1));
    double* len = (double*)malloc(sizeof(double) * (N > 1 ? N - 1 : 1));
    chain_geometry(N, R, bond, len);
    bending_from_geometry(N, bond, len, p->bend_type, p->bend_stiffness, F);
    free(bond); free(len);
}

/* ======================================================================= */
/* excluded volume (excluded_volume.cpp:14-89)                               */
/* ======================================================================= */

typedef struct ev_consts {
    int type, cd, equil;
    double zsbyds5, pt5bydssq, LJa, LJb, SDKa, SDKb, Rmin, Rcutg, Rcutp, RcutpSDK;
    const double* phi;  /* NULL => zeros */
} ev_consts;

static void ev_setup(const bdsim_phys* p, ev_consts* e) {
    const double z = p->zstar, d = p->dstar;
    e->type = p->ev_type;
    e->cd = p->contour_dist_for_ev;
    e->equil = p->equilibration != 0;
    e->phi = (p->ev_type == BDSIM_EV_SDK && p->n_phi > 0) ? p->phi : NULL;
    e->zsbyds5 = z / pow(d, 5.0);
    e->pt5bydssq = 0.5 / (d * d);
    e->LJa = 12.0 * 4.0 * z * pow(d, 12.0);
    e->LJb = 6.0 * 4.0 * z * pow(d, 6.0);
    e->SDKa = 12.0 * 4.0 * pow(d, 12.0);
    e->SDKb = 6.0 * 4.0 * pow(d, 6.0);
    e->Rmin = p->min_cutoff;
    e->Rcutp = p->max_cutoff;
    e->RcutpSDK = p->max_cutoff;
    e->Rcutg = d * pow(2.0, 1.0 / 6.0);
}

static double ev_factor(const ev_consts* e, double r, double phi_pair) {
    const double SDK_ALPHA = 1.530633312, SDK_BETA = 1.213115524;
    switch (e->type) {
        case BDSIM_EV_GAUSS: return e->zsbyds5 * exp(-r * r * e->pt5bydssq);
        case BDSIM_EV_LJ: {
            const double Rcut = e->equil ? e->Rcutg : e->Rcutp;
            if (r <= Rcut && r >= e->Rmin)
                return e->LJa / pow(r, 12.0 + 2.0) - e->LJb / pow(r, 6.0 + 2.0);
            if (r < e->Rmin)
                return e->LJa / pow(e->Rmin, 12.0 + 2.0) - e->LJb / pow(e->Rmin, 6.0 + 2.0);
            return 0.0;
        }
        case BDSIM_EV_SDK: {
            double fac = 0.0;
            if (r <= e->Rcutg && r >= e->Rmin)
                fac = e->SDKa / pow(r, 12.0 + 2.0) - e->SDKb / pow(r, 6.0 + 2.0);
            else if (r < e->Rmin)
                fac = e->SDKa / pow(e->Rmin, 12.0 + 2.0) - e->SDKb / pow(e->Rmin, 6.0 + 2.0);
            if (!e->equil && r <= e->RcutpSDK && r > e->Rcutg)
                fac += SDK_ALPHA * phi_pair * sin(SDK_ALPHA * r * r + SDK_BETA);
            return fac;
        }
        default: return 0.0;
    }
}

static void ev_force_c(int N, const double* R, const ev_consts* e, double* F) {
    memset(F, 0, sizeof(double) * 3 * N);
    if (e->type == BDSIM_EV_NONE) return;
    for (int nu = 1; nu < N; ++nu) {
        for (int mu = 0; mu < nu; ++mu) {
            if ((nu - mu) < e->cd) continue;
            double q[3];
            for (int d = 0; d < 3; ++d) q[d] = R[3 * nu + d] - R[3 * mu + d];
            double sq = n2(q);
            if (sq < 1.0e-12) sq = 1.0e-12;
            const double r = sqrt(sq);
            const double phi = e->phi ? e->phi[mu * N + nu] : 0.0;
            const double fac = ev_factor(e, r, phi);
            for (int d = 0; d < 3; ++d) {
                const double f = (-fac) * q[d];
                F[3 * mu + d] = F[3 * mu + d] + f;
                F[3 * nu + d] = F[3 * nu + d] - f;
            }
        }
    }
}

void bo_ev_force(int N, const double* R, const bdsim_phys* p, double* F) {
    ev_consts e;
    ev_setup(p, &e);
    ev_force_c(N, R, &e, F);
}

/* ======================================================================= */
/* flow and external forces (flow.cpp:5-26, external_force.cpp:8-52)         */
/* ======================================================================= */

void bo_kappa(const bdsim_flow* f, double t, double K[9]) {
    if (f->constant || f->n_times == 0) { memcpy(K, f->kappa0, sizeof(double) * 9); return; }
    const int n = f->n_times;
    if (n == 1 || t <= f->times[0]) { memcpy(K, f->kappas, sizeof(double) * 9); return; }
    if (t >= f->times[n - 1]) { memcpy(K, f->kappas + 9 * (n - 1), sizeof(double) * 9); return; }
    int i = 0;
    while (i < n - 1 && !(t >= f->times[i] && t < f->times[i + 1])) ++i;
    const double x0 = f->times[i], x1 = f->times[i + 1];
    const double w = (t - x0) / (x1 - x0);
    const double* A = f->kappas + 9 * i;
    const double* B = f->kappas + 9 * (i + 1);
    for (int k = 0; k < 9; ++k) K[k] = A[k] + w * (B[k] - A[k]);
}

static void sample_vec(int n, const double* times, const double* values, double t, double* out) {
    if (n == 0) { out[0] = out[1] = out[2] = 0.0; return; }
    if (n == 1 || t <= times[0]) { memcpy(out, values, sizeof(double) * 3); return; }
    if (t >= times[n - 1]) { memcpy(out, values + 3 * (n - 1), sizeof(double) * 3); return; }
    int i = 0;
    while (i < n - 1 && !(t >= times[i] && t < times[i + 1])) ++i;
    const double w = (t - times[i]) / (times[i + 1] - times[i]);
    for (int d = 0; d < 3; ++d)
        out[d] = values[3 * i + d] + w * (values[3 * (i + 1) + d] - values[3 * i + d]);
}

void bo_external_force(const bdsim_phys* p, int N, double t, double* F) {
    memset(F, 0, sizeof(double) * 3 * N);
    for (int k = 0; k < p->n_external; ++k) {
        const bdsim_bead_force* e = &p->external[k];
        const int i = e->bead < 0 ? N + e->bead : e->bead;
        if (i < 0 || i >= N) continue;
        double v[3];
        if (e->constant) memcpy(v, e->value, sizeof v);
        else sample_vec(e->n_times, e->times, e->values, t, v);
        for (int d = 0; d < 3; ++d) F[3 * i + d] = F[3 * i + d] + v[d];
    }
}

void bo_total_force(int N, const double* R, const bdsim_phys* p, double* F) {
    double* Fe = (double*)malloc(sizeof(double) * 3 * N);
    double* Fb = (double*)malloc(sizeof(double) * 3 * N);
    bo_spring_force(N, R, p, F);
    bo_ev_force(N, R, p, Fe);
    bo_bending_force(N, R, p, Fb);
    /* non_spring_force = ev + bending (model.hpp:56-58), then spring + that */
    for (int i = 0; i < 3 * N; ++i) F[i] = F[i] + (Fe[i] + Fb[i]);
    free(Fe); free(Fb);
}

/* ======================================================================= */
/* hydrodynamics (hydrodynamics.cpp, linalg.cpp hand-rolled branch)          */
/* ======================================================================= */

/* hydrodynamics.cpp:12-54 */
void bo_rpy_matrix(int N, const double* R, double hstar, double* D) {
    const int n = 3 * N;
    memset(D, 0, sizeof(double) * (size_t)n * n);
    for (int mu = 0; mu < N; ++mu)
        for (int i = 0; i < 3; ++i) D[(size_t)(3 * mu + i) * n + 3 * mu + i] = 1.0;
    if (hstar <= 0.0) return;
    const double RPI = sqrt(BO_PI), TRPI = 2.0 * RPI, TPI = 2.0 * BO_PI;
    const double C1 = TPI / 3.0, C3 = 0.75 / RPI * 0.375, C7 = 0.09375 / RPI;
    const double RPI3by4 = RPI * 0.75;
    for (int nu = 1; nu < N; ++nu) {
        for (int mu = 0; mu < nu; ++mu) {
            double b[3];
            for (int d = 0; d < 3; ++d) b[d] = R[3 * nu + d] - R[3 * mu + d];
            double sq = n2(b);
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
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const double val = om2 * b[i] * b[j] + (i == j ? om1 : 0.0);
                    D[(size_t)(3 * mu + i) * n + 3 * nu + j] = val;
                    D[(size_t)(3 * nu + j) * n + 3 * mu + i] = val;
                }
        }
    }
}

/* linalg.cpp:81-89 */
void bo_matvec(int n, const double* M, const double* x, double* y) {
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j < n; ++j) s += M[(size_t)i * n + j] * x[j];
        y[i] = s;
    }
}

/* linalg.cpp:91-107 */
int bo_cholesky_lower(int n, const double* M, double* L, int* pivot, double* value) {
    memset(L, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j <= i; ++j) {
            double sum = M[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) sum -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            if (i == j) {
                if (!(sum > 0.0)) { if (pivot) *pivot = i; if (value) *value = sum; return -1; }
                L[(size_t)i * n + j] = sqrt(sum);
            } else {
                L[(size_t)i * n + j] = sum / L[(size_t)j * n + j];
            }
        }
    }
    return 0;
}

/* linalg.cpp:109-117 */
void bo_lower_matvec(int n, const double* L, const double* x, double* y) {
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int j = 0; j <= i; ++j) s += L[(size_t)i * n + j] * x[j];
        y[i] = s;
    }
}

/* hydrodynamics.cpp:83-87 */
static void default_seed(int n, double* v) {
    for (int i = 0; i < n; ++i) v[i] = sin(0.7654321 * (i + 1)) + 0.5;
}

/* hydrodynamics.cpp:91-101 */
static int sturm_count(int k, const double* a, const double* b, double x) {
    int count = 0;
    double d = a[0] - x;
    if (d < 0.0) ++count;
    for (int i = 1; i < k; ++i) {
        if (d == 0.0) d = 1.0e-300;
        d = (a[i] - x) - b[i - 1] * b[i - 1] / d;
        if (d < 0.0) ++count;
    }
    return count;
}

/* hydrodynamics.cpp:106-136 */
static void tridiag_eigenvector(int m, const double* a, const double* b, double lambda,
                                double* y) {
    const double scale = (fabs(lambda) < 1.0) ? 1.0 : fabs(lambda);
    const double shift = lambda - 1.0e-7 * scale;
    double c[BO_LANCZOS_COLD], dd[BO_LANCZOS_COLD];
    for (int i = 0; i < m; ++i) y[i] = 1.0 / sqrt((double)m);
    for (int it = 0; it < 2; ++it) {
        double beta0 = a[0] - shift;
        if (fabs(beta0) < 1.0e-300) beta0 = 1.0e-300;
        c[0] = (m > 1 ? b[0] / beta0 : 0.0);
        dd[0] = y[0] / beta0;
        for (int i = 1; i < m; ++i) {
            double beta = (a[i] - shift) - b[i - 1] * c[i - 1];
            if (fabs(beta) < 1.0e-300) beta = 1.0e-300;
            c[i] = (i < m - 1 ? b[i] / beta : 0.0);
            dd[i] = (y[i] - b[i - 1] * dd[i - 1]) / beta;
        }
        y[m - 1] = dd[m - 1];
        for (int i = m - 2; i >= 0; --i) y[i] = dd[i] - c[i] * y[i + 1];
        double nrm = 0.0;
        for (int i = 0; i < m; ++i) nrm += y[i] * y[i];
        nrm = sqrt(nrm);
        if (!(nrm > 0.0) || !isfinite(nrm)) {
            for (int i = 0; i < m; ++i) y[i] = 1.0 / sqrt((double)m);
            return;
        }
        for (int i = 0; i < m; ++i) y[i] /= nrm;
    }
}

typedef struct lanczos_out {
    double rmin, rmax;
    double* vmin;  /* [n] */
    double* vmax;  /* [n] */
    int steps;
} lanczos_out;

/* hydrodynamics.cpp:146-220; seed NULL => default seed. Returns mat-vecs done. */
static int lanczos_extremes(int n, const double* D, int m, const double* seed,
                            lanczos_out* out) {
    if (m > n) m = n;
    double* V = (double*)malloc(sizeof(double) * (size_t)m * n);
    double* v = (double*)malloc(sizeof(double) * n);
    double* vprev = (double*)calloc((size_t)n, sizeof(double));
    double* w = (double*)malloc(sizeof(double) * n);
    double alpha[BO_LANCZOS_COLD], beta[BO_LANCZOS_COLD];
    int na = 0, nbeta = 0, mv = 0;
    if (seed) memcpy(v, seed, sizeof(double) * n);
    else default_seed(n, v);
    {
        double nv = sqrt(dotn(n, v, v));
        if (!(nv > 0.0)) { default_seed(n, v); nv = sqrt(dotn(n, v, v)); }
        for (int i = 0; i < n; ++i) v[i] /= nv;
    }
    double b = 0.0;
    for (int j = 0; j < m; ++j) {
        memcpy(V + (size_t)j * n, v, sizeof(double) * n);
        bo_matvec(n, D, v, w);
        ++mv;
        const double a = dotn(n, v, w);
        alpha[na++] = a;
        for (int i = 0; i < n; ++i) w[i] -= a * v[i] + b * vprev[i];
        for (int q = 0; q <= j; ++q) {
            const double* u = V + (size_t)q * n;
            const double d = dotn(n, w, u);
            for (int i = 0; i < n; ++i) w[i] -= d * u[i];
        }
        b = sqrt(dotn(n, w, w));
        if (b < 1.0e-10) break;
        memcpy(vprev, v, sizeof(double) * n);
        for (int i = 0; i < n; ++i) v[i] = w[i] / b;
        beta[nbeta++] = b;
    }
    const int k = na;
    out->steps = k;
    if (k == 1) {
        out->rmin = out->rmax = alpha[0];
        memcpy(out->vmin, V, sizeof(double) * n);
        memcpy(out->vmax, V, sizeof(double) * n);
    } else {
        const double* bb = beta;  /* first k-1 entries */
        double lo = alpha[0], hi = alpha[0];
        for (int i = 0; i < k; ++i) {
            const double rad = (i > 0 ? fabs(bb[i - 1]) : 0.0) + (i < k - 1 ? fabs(bb[i]) : 0.0);
            if (alpha[i] - rad < lo) lo = alpha[i] - rad;
            if (alpha[i] + rad > hi) hi = alpha[i] + rad;
        }
        double a1 = lo, b1 = hi;
        for (int it = 0; it < 100; ++it) {
            const double mid = 0.5 * (a1 + b1);
            if (sturm_count(k, alpha, bb, mid) >= 1) b1 = mid; else a1 = mid;
        }
        double a2 = lo, b2 = hi;
        for (int it = 0; it < 100; ++it) {
            const double mid = 0.5 * (a2 + b2);
            if (sturm_count(k, alpha, bb, mid) >= k) b2 = mid; else a2 = mid;
        }
        out->rmin = 0.5 * (a1 + b1);
        out->rmax = 0.5 * (a2 + b2);
        double ymin[BO_LANCZOS_COLD], ymax[BO_LANCZOS_COLD];
        tridiag_eigenvector(k, alpha, bb, out->rmin, ymin);
        tridiag_eigenvector(k, alpha, bb, out->rmax, ymax);
        for (int i = 0; i < n; ++i) { out->vmin[i] = 0.0; out->vmax[i] = 0.0; }
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < n; ++i) {
                out->vmin[i] += ymin[j] * V[(size_t)j * n + i];
                out->vmax[i] += ymax[j] * V[(size_t)j * n + i];
            }
    }
    free(V); free(v); free(vprev); free(w);
    return mv;
}

/* hydrodynamics.cpp:224-264.  Returns mat-vecs done. */
static int spectral_bounds_c(int n, const double* D, double* seed, int* seed_valid,
                             double* lmin_out, double* lmax_out, int* steps, int* cold) {
    const int can_warm = seed_valid && *seed_valid;
    lanczos_out r;
    r.vmin = (double*)malloc(sizeof(double) * n);
    r.vmax = (double*)malloc(sizeof(double) * n);
    int mv = lanczos_extremes(n, D, can_warm ? BO_LANCZOS_WARM : BO_LANCZOS_COLD,
                              can_warm ? seed : NULL, &r);
    int was_cold = !can_warm;
    if (can_warm && r.steps < 4) {
        mv += lanczos_extremes(n, D, BO_LANCZOS_COLD, NULL, &r);
        was_cold = 1;
    }
    if (seed && seed_valid) {
        double* u = (double*)malloc(sizeof(double) * n);
        double* s = (double*)malloc(sizeof(double) * n);
        default_seed(n, u);
        const double un = sqrt(dotn(n, u, u));
        for (int i = 0; i < n; ++i) s[i] = r.vmin[i] + r.vmax[i] + 0.01 * u[i] / un;
        const double sn = sqrt(dotn(n, s, s));
        if (sn > 0.0 && isfinite(sn)) {
            for (int i = 0; i < n; ++i) s[i] /= sn;
            memcpy(seed, s, sizeof(double) * n);
            *seed_valid = 1;
        } else {
            *seed_valid = 0;
        }
        free(u); free(s);
    }
    double lmin = 0.5 * r.rmin;
    double lmax = 1.02 * r.rmax;
    if (lmin < 1.0e-12) lmin = 1.0e-12;
    if (lmin >= lmax) lmin = lmax * 1.0e-6;
    *lmin_out = lmin;
    *lmax_out = lmax;
    if (steps) *steps = r.steps;
    if (cold) *cold = was_cold;
    free(r.vmin); free(r.vmax);
    return mv;
}

void bo_spectral_bounds(int n, const double* D, double* seed, int* seed_valid,
                        double* lmin, double* lmax, int* steps, int* cold) {
    spectral_bounds_c(n, D, seed, seed_valid, lmin, lmax, steps, cold);
}

/* hydrodynamics.cpp:267-279 */
static void sqrt_chebyshev_coefficients(int L, double da, double db, double* a) {
    double* xks = (double*)malloc(sizeof(double) * (L + 1));
    for (int j = 0; j <= L; ++j) a[j] = 0.0;
    for (int k = 0; k <= L; ++k) xks[k] = cos(BO_PI * (k + 0.5) / (L + 1)) / da - db / da;
    for (int j = 0; j <= L; ++j) {
        for (int k = 0; k <= L; ++k) a[j] += sqrt(xks[k]) * cos(j * (k + 0.5) * BO_PI / (L + 1));
        a[j] *= 2.0 / (L + 1);
    }
    a[0] /= 2.0;
    free(xks);
}

/* hydrodynamics.cpp:284-312.  Returns mat-vecs done (L). */
static int chebyshev_series(int n, const double* D, const double* x0, int L, double lmin,
                            double lmax, double* result) {
    const double da = 2.0 / (lmax - lmin);
    const double db = -(lmax + lmin) / (lmax - lmin);
    double* a = (double*)malloc(sizeof(double) * (L + 1));
    double* tprev = (double*)malloc(sizeof(double) * n);
    double* tcur = (double*)malloc(sizeof(double) * n);
    double* tnext = (double*)malloc(sizeof(double) * n);
    double* Dv = (double*)malloc(sizeof(double) * n);
    sqrt_chebyshev_coefficients(L, da, db, a);
    for (int i = 0; i < n; ++i) result[i] = a[0] * x0[i];
    memcpy(tprev, x0, sizeof(double) * n);
    bo_matvec(n, D, x0, Dv);
    for (int i = 0; i < n; ++i) tcur[i] = da * Dv[i] + db * x0[i];
    for (int i = 0; i < n; ++i) result[i] += a[1] * tcur[i];
    for (int l = 2; l <= L; ++l) {
        bo_matvec(n, D, tcur, Dv);
        for (int i = 0; i < n; ++i) tnext[i] = da * Dv[i] + db * tcur[i];
        for (int i = 0; i < n; ++i) tnext[i] = 2.0 * tnext[i] - tprev[i];
        double* tmp = tprev; tprev = tcur; tcur = tnext; tnext = tmp;
        for (int i = 0; i < n; ++i) result[i] += a[l] * tcur[i];
    }
    free(a); free(tprev); free(tcur); free(tnext); free(Dv);
    return L;
}

typedef struct cheb_stats { int matvecs, lanczos_mv, lanczos_steps, cold, L, attempts, terms; double fd, lmin, lmax; } cheb_stats;

/* hydrodynamics.cpp:317-337 */
static void chebyshev_sqrt_times_c(int n, const double* D, const double* x0, double mult,
                                   double fd_tol, double* seed, int* seed_valid,
                                   double* result, cheb_stats* st) {
    double lmin, lmax;
    int steps = 0, cold = 0;
    const int lmv = spectral_bounds_c(n, D, seed, seed_valid, &lmin, &lmax, &steps, &cold);
    double* Dx0 = (double*)malloc(sizeof(double) * n);
    bo_matvec(n, D, x0, Dx0);
    const double x0Dx0 = dotn(n, x0, Dx0);
    int L = (int)(sqrt(lmax / lmin) * mult + 0.5) + 1;
    if (L > BO_MAXCHEB) L = BO_MAXCHEB;
    if (L < 2) L = 2;
    double fd = 0.0;
    int attempts = 0, terms = 0;
    for (;;) {
        chebyshev_series(n, D, x0, L, lmin, lmax, result);
        ++attempts;
        terms += L;
        fd = fabs(dotn(n, result, result) - x0Dx0) / x0Dx0;
        if (fd <= fd_tol || L >= BO_MAXCHEB) break;
        const int grow = (L / 4 > 1) ? L / 4 : 1;
        L = (L + grow < BO_MAXCHEB) ? L + grow : BO_MAXCHEB;
    }
    free(Dx0);
    st->lanczos_mv = lmv;
    st->lanczos_steps = steps;
    st->cold = cold;
    st->L = L;
    st->attempts = attempts;
    st->terms = terms;
    st->matvecs = lmv + 1 + terms;
    st->fd = fd;
    st->lmin = lmin;
    st->lmax = lmax;
}

void bo_chebyshev_sqrt_times(int n, const double* D, const double* x0, double mult,
                             double fd_tol, double* seed, int* seed_valid,
                             double* out, int* nterms, double* fd_err) {
    cheb_stats st;
    chebyshev_sqrt_times_c(n, D, x0, mult, fd_tol, seed, seed_valid, out, &st);
    if (nterms) *nterms = st.L;
    if (fd_err) *fd_err = st.fd;
}

/* ======================================================================= */
/* the integrator (integrator.cpp:46-177)                                    */
/* ======================================================================= */

typedef struct stepper {
    int N, n;
    const bdsim_phys* p;
    const bdsim_sim* s;
    ev_consts ev;
    double* D;      /* n x n (hstar > 0) */
    double* L;      /* n x n Cholesky factor (Cholesky method / fallback) */
    int hi;         /* hstar > 0 */
    double* seed;   /* spectral cache */
    int seed_valid;
    bdsim_counters* cnt;
    double* tmp;    /* n scratch */
} stepper;

/* Diffusion::apply (mobility.cpp:53-56) */
static void apply_D(stepper* st, const double* F, double* out) {
    if (!st->hi) { memcpy(out, F, sizeof(double) * st->n); return; }
    bo_matvec(st->n, st->D, F, out);
    if (st->cnt) st->cnt->matvecs++;
}

/* model.hpp:56-64: non-spring force + external force at t */
static void explicit_force(stepper* st, const double* R, const double* bond,
                           const double* len, double t, double* Fe) {
    const int N = st->N;
    double* Fev = (double*)malloc(sizeof(double) * 3 * N);
    double* Fbe = (double*)malloc(sizeof(double) * 3 * N);
    double* Fx = (double*)malloc(sizeof(double) * 3 * N);
    ev_force_c(N, R, &st->ev, Fev);
    bending_from_geometry(N, bond, len, st->p->bend_type, st->p->bend_stiffness, Fbe);
    if (st->p->n_external > 0) bo_external_force(st->p, N, t, Fx);
    else memset(Fx, 0, sizeof(double) * 3 * N);
    for (int i = 0; i < 3 * N; ++i) Fe[i] = (Fev[i] + Fbe[i]) + Fx[i];
    free(Fev); free(Fbe); free(Fx);
}

static void flow_velocity(const double K[9], int N, const double* R, double* out) {
    for (int i = 0; i < N; ++i) {
        const double* v = R + 3 * i;
        out[3 * i + 0] = K[0] * v[0] + K[1] * v[1] + K[2] * v[2];
        out[3 * i + 1] = K[3] * v[0] + K[4] * v[1] + K[5] * v[2];
        out[3 * i + 2] = K[6] * v[0] + K[7] * v[1] + K[8] * v[2];
    }
}

int bo_integrate(int N, double* R, const bdsim_phys* phys, const bdsim_sim* sim,
                 bdsim_rng_state* rng, bdsim_counters* cnt, bdsim_step_log* log,
                 int log_cap) {
    const int n = 3 * N;
    const double dt = sim->dt;
    const double sqrt2inv = 1.0 / sqrt(2.0);
    int status = BDSIM_STATUS_OK;
    stepper st;
    memset(&st, 0, sizeof st);
    st.N = N; st.n = n; st.p = phys; st.s = sim; st.cnt = cnt;
    ev_setup(phys, &st.ev);
    st.hi = phys->hstar > 0.0;
    if (cnt) memset(cnt, 0, sizeof *cnt);
    if (st.hi) {
        st.D = (double*)malloc(sizeof(double) * (size_t)n * n);
        st.L = (double*)malloc(sizeof(double) * (size_t)n * n);
        st.seed = (double*)malloc(sizeof(double) * n);
    }
    st.seed_valid = 0;   /* fresh PhysicalModel per call => cold first step */

    double* bond = (double*)malloc(sizeof(double) * 3 * N);
    double* len = (double*)malloc(sizeof(double) * N);
    double* Fb = (double*)malloc(sizeof(double) * 3 * N);
    double* Fs = (double*)malloc(sizeof(double) * 3 * N);
    double* Fe = (double*)malloc(sizeof(double) * 3 * N);
    double* Fep = (double*)malloc(sizeof(double) * 3 * N);
    double* x0 = (double*)malloc(sizeof(double) * 3 * N);
    double* bx0 = (double*)malloc(sizeof(double) * 3 * N);
    double* dW = (double*)malloc(sizeof(double) * 3 * N);
    double* Rp = (double*)malloc(sizeof(double) * 3 * N);
    double* ups = (double*)malloc(sizeof(double) * 3 * N);
    double* v1 = (double*)malloc(sizeof(double) * 3 * N);
    double* v2 = (double*)malloc(sizeof(double) * 3 * N);
    double* v3 = (double*)malloc(sizeof(double) * 3 * N);
    double* DF = (double*)malloc(sizeof(double) * 3 * N);
    double* Rc = (double*)malloc(sizeof(double) * 3 * N);
    double* Rprev = (double*)malloc(sizeof(double) * 3 * N);
    double* dups = (double*)malloc(sizeof(double) * 3 * N);
    double* cbond = (double*)malloc(sizeof(double) * 3 * N);
    double* clen = (double*)malloc(sizeof(double) * N);
    const double q0 = phys->natural_length / phys->sqrtb;
    int step = 0;

    for (double t = sim->time_start; t <= sim->time_end + sim->dt / 2.0; t += sim->dt) {
        bdsim_step_log lg;
        memset(&lg, 0, sizeof lg);
        lg.t = t;
        /* recenter (integrator.cpp:88-93) */
        if (sim->update_center_of_mass) {
            double com[3] = {0, 0, 0};
            for (int i = 0; i < N; ++i)
                for (int d = 0; d < 3; ++d) com[d] = com[d] + R[3 * i + d];
            const double inv = 1.0 / N;
            for (int d = 0; d < 3; ++d) com[d] = inv * com[d];
            for (int i = 0; i < N; ++i)
                for (int d = 0; d < 3; ++d) R[3 * i + d] = R[3 * i + d] - com[d];
        }
        chain_geometry(N, R, bond, len);
        connector_forces(N - 1, bond, len, phys->spring_type, phys->sqrtb, q0, Fb);
        bead_forces_from_connectors(N, Fb, Fs);
        explicit_force(&st, R, bond, len, t, Fe);

        /* Mobility::at + Diffusion::rpy (mobility.cpp:39-51) */
        if (st.hi) {
            bo_rpy_matrix(N, R, phys->hstar, st.D);
            if (phys->hi_method == BDSIM_HI_CHOLESKY) {
                int piv; double val;
                if (bo_cholesky_lower(n, st.D, st.L, &piv, &val) != 0) {
                    status = BDSIM_STATUS_NOT_POSITIVE_DEFINITE;
                    if (cnt) { cnt->fail_index = piv; cnt->fail_value = val; }
                    break;
                }
            }
        }
        /* Diffusion::brownian_step (mobility.cpp:58-88) */
        bo_raw_noise(rng, N, x0);
        const double sqrt_dt = sqrt(dt);
        if (!st.hi) {
            for (int i = 0; i < n; ++i) dW[i] = sqrt_dt * x0[i];
        } else {
            if (phys->hi_method == BDSIM_HI_CHOLESKY) {
                bo_lower_matvec(n, st.L, x0, bx0);
            } else if (phys->hi_method == BDSIM_HI_CHEBYSHEV) {
                cheb_stats cs;
                chebyshev_sqrt_times_c(n, st.D, x0, phys->ncheb_multiplier, phys->fd_err_max,
                                       st.seed, &st.seed_valid, bx0, &cs);
                lg.lmin = cs.lmin; lg.lmax = cs.lmax; lg.fd_err = cs.fd;
                lg.lanczos_cold = cs.cold; lg.lanczos_steps = cs.lanczos_steps;
                lg.cheb_L = cs.L; lg.cheb_attempts = cs.attempts;
                if (cnt) {
                    cnt->matvecs += cs.matvecs;
                    cnt->lanczos_iters += cs.lanczos_mv;
                    if (cs.cold) cnt->lanczos_cold++; else cnt->lanczos_warm++;
                    cnt->cheb_terms += cs.terms;
                    cnt->cheb_attempts += cs.attempts;
                }
                if (!(cs.fd <= phys->fd_err_max)) {
                    int piv; double val;
                    lg.fallback = 1;
                    if (cnt) cnt->fallbacks++;
                    if (bo_cholesky_lower(n, st.D, st.L, &piv, &val) != 0) {
                        status = BDSIM_STATUS_NOT_POSITIVE_DEFINITE;
                        if (cnt) { cnt->fail_index = piv; cnt->fail_value = val; }
                        break;
                    }
                    bo_lower_matvec(n, st.L, x0, bx0);
                }
            } else {
                status = BDSIM_STATUS_EXACT_SQRT;
                break;
            }
            for (int i = 0; i < n; ++i) dW[i] = sqrt_dt * bx0[i];
        }

        double K0[9], K1[9];
        bo_kappa(&phys->flow, t, K0);
        bo_kappa(&phys->flow, t + dt, K1);

        /* predictor (integrator.cpp:96-101) */
        for (int i = 0; i < n; ++i) v1[i] = Fs[i] + Fe[i];
        apply_D(&st, v1, DF);
        flow_velocity(K0, N, R, v2);
        for (int i = 0; i < n; ++i) Rp[i] = R[i] + dt * (v2[i] + 0.25 * DF[i]) + sqrt2inv * dW[i];

        /* upsilon (integrator.cpp:107-119) */
        chain_geometry(N, Rp, cbond, clen);
        explicit_force(&st, Rp, cbond, clen, t + dt, Fep);
        flow_velocity(K1, N, Rp, v3);
        for (int i = 0; i < n; ++i) v3[i] = 0.5 * (v2[i] + v3[i]);   /* flow */
        apply_D(&st, Fs, v2);                                         /* DFs */
        for (int i = 0; i < n; ++i) v1[i] = Fe[i] + Fep[i];
        apply_D(&st, v1, DF);                                         /* DFe */
        for (int i = 0; i < n; ++i)
            ups[i] = R[i] + dt * (v3[i] + 0.125 * v2[i] + 0.125 * DF[i]) + sqrt2inv * dW[i];

        /* solve_connectors (integrator.cpp:125-165) */
        {
            const int nb = N - 1;
            for (int mu = 0; mu < nb; ++mu)
                for (int d = 0; d < 3; ++d) {
                    dups[3 * mu + d] = ups[3 * (mu + 1) + d] - ups[3 * mu + d];
                    cbond[3 * mu + d] = Rp[3 * (mu + 1) + d] - Rp[3 * mu + d];
                }
            double anchor[3] = {Rp[0], Rp[1], Rp[2]};
            memcpy(Rprev, Rp, sizeof(double) * n);
            const int max_iter = 100 * N;
            double increment = 0.0;
            int iter = 0;
            for (; iter <= max_iter; ++iter) {
                for (int mu = 0; mu < nb; ++mu) clen[mu] = clamped_length(cbond + 3 * mu);
                connector_forces(nb, cbond, clen, phys->spring_type, phys->sqrtb, q0, Fb);
                bead_forces_from_connectors(N, Fb, v1);
                apply_D(&st, v1, DF);
                if (cnt) cnt->corrector_iters++;
                lg.corrector_iters++;
                for (int d = 0; d < 3; ++d) Rc[d] = anchor[d];
                for (int mu = 0; mu < nb; ++mu) {
                    double g[3];
                    for (int d = 0; d < 3; ++d)
                        g[d] = dups[3 * mu + d] +
                               (0.125 * dt) * (DF[3 * (mu + 1) + d] - DF[3 * mu + d]) +
                               (0.25 * dt) * Fb[3 * mu + d];
                    const int sst = bo_solve_connector(g, dt, phys->spring_type, phys->sqrtb,
                                                       phys->natural_length, cbond + 3 * mu);
                    if (sst != BDSIM_STATUS_OK) status = sst;
                    for (int d = 0; d < 3; ++d) Rc[3 * (mu + 1) + d] = Rc[3 * mu + d] + cbond[3 * mu + d];
                }
                if (status != BDSIM_STATUS_OK) break;
                double s = 0.0;
                for (int i = 0; i < N; ++i) {
                    double dv[3];
                    for (int d = 0; d < 3; ++d) dv[d] = Rc[3 * i + d] - Rprev[3 * i + d];
                    s += n2(dv);
                }
                increment = sqrt(s) / N;
                if (increment < sim->implicit_loop_tol) break;
                memcpy(Rprev, Rc, sizeof(double) * n);
            }
            if (status != BDSIM_STATUS_OK) break;
            if (iter > max_iter && cnt) cnt->nonconverged++;
            memcpy(R, Rc, sizeof(double) * n);
        }
        if (cnt) cnt->steps++;
        if (log && step < log_cap) log[step] = lg;
        ++step;
    }

    free(bond); free(len); free(Fb); free(Fs); free(Fe); free(Fep); free(x0); free(bx0);
    free(dW); free(Rp); free(ups); free(v1); free(v2); free(v3); free(DF); free(Rc);
    free(Rprev); free(dups); free(cbond); free(clen);
    free(st.D); free(st.L); free(st.seed);
    return status;
}

int bo_integrate_batch(int n_traj, const int32_t* nbeads, double* R, const bdsim_phys* phys,
                       const bdsim_sim* sim, bdsim_rng_state* rng, int32_t* status,
                       bdsim_counters* counters) {
    size_t off = 0;
    for (int i = 0; i < n_traj; ++i) {
        bdsim_phys p = phys[i];
        p.number_of_beads = nbeads[i];
        status[i] = bo_integrate(nbeads[i], R + 3 * off, &p, &sim[i], &rng[i],
                                 counters ? &counters[i] : NULL, NULL, 0);
        off += (size_t)nbeads[i];
    }
    return 0;
}
