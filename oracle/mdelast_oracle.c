/*
 * CPU ORACLE — test infrastructure only (see mdelast_oracle.h).
 *
 * Plain C99, FP64, single thread. Each function cites the reference code it
 * restates. Build: oracle/Makefile (gcc -O2, no -march, like the reference's
 * Release flags in src/CMakeLists.txt:22-25, so no FMA contraction on x86-64).
 */
#include "mdelast_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* src/units.hpp:6-24 */
static const double K_BOLTZ = 1.38064852e-16;

/* Voigt index pairs xx yy zz yz xz xy (md_driver.hpp:23-24) */
static const int VA[6] = {0, 1, 2, 1, 0, 0};
static const int VB[6] = {0, 1, 2, 2, 2, 1};

/* ------------------------------------------------------------------------- */
/* Tables                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    /* cardinal cubic B-spline through n values on x_k = k h, k = 0..n-1 */
    int n;
    double h, inv_h, avg;
    double *beta; /* n + 2 coefficients, beta[k] multiplies B3(t - k), t = x/h + 1 */
} bspline;

struct orc_table {
    int kind;
    double h, inv_h, inv_h2;
    int n;          /* intervals (quintic) */
    double *coef;   /* [n][6] */
    bspline s[3];   /* f, f', f'' (B-spline kind only) */
    double prm[3];  /* analytic kinds: Morse {D, alpha, r0}; LJ {epsilon, sigma, vcut} */
};

/* Quintic Hermite coefficients per interval in t = (x - x_i)/h, formed in long
 * double: the polynomial with p, p', p'' equal to the node data at t = 0 and 1
 * (poly_table.hpp:45-62). */
static void quintic_build(orc_table *t, const double *f, const double *f1, const double *f2) {
    long double const hl = (long double)t->h;
    for (int i = 0; i < t->n; ++i) {
        long double const dy = (long double)f[i + 1] - (long double)f[i];
        long double const D0 = hl * (long double)f1[i], D1 = hl * (long double)f1[i + 1];
        long double const S0 = hl * hl * (long double)f2[i], S1 = hl * hl * (long double)f2[i + 1];
        double *c = t->coef + 6 * (size_t)i;
        c[0] = f[i];
        c[1] = (double)D0;
        c[2] = (double)(0.5L * S0);
        c[3] = (double)(10.L * dy - 6.L * D0 - 4.L * D1 - 0.5L * (3.L * S0 - S1));
        c[4] = (double)(-15.L * dy + 8.L * D0 + 7.L * D1 + 0.5L * (3.L * S0 - 2.L * S1));
        c[5] = (double)(6.L * dy - 3.L * (D0 + D1) - 0.5L * (S0 - S1));
    }
}

/* Interval of x and the local coordinate; the range test precedes the int
 * conversion so that NaN gives interval 0 and a NaN result, and points outside
 * the grid use the end polynomials (poly_table.hpp:105-114). */
static const double *quintic_locate(const orc_table *t, double x, double *tt) {
    double const u = x * t->inv_h;
    int i = 0;
    if (u >= (double)t->n)
        i = t->n - 1;
    else if (u > 0.)
        i = (int)u;
    *tt = u - (double)i;
    return t->coef + 6 * (size_t)i;
}

/* poly_table.hpp:71-97 */
static double quintic_value(const orc_table *t, double x) {
    double s;
    const double *c = quintic_locate(t, x, &s);
    return c[0] + s * (c[1] + s * (c[2] + s * (c[3] + s * (c[4] + s * c[5]))));
}
static double quintic_d1(const orc_table *t, double x) {
    double s;
    const double *c = quintic_locate(t, x, &s);
    return (c[1] + s * (2. * c[2] + s * (3. * c[3] + s * (4. * c[4] + s * 5. * c[5])))) * t->inv_h;
}
static double quintic_d2(const orc_table *t, double x) {
    double s;
    const double *c = quintic_locate(t, x, &s);
    return (2. * c[2] + s * (6. * c[3] + s * (12. * c[4] + s * 20. * c[5]))) * t->inv_h2;
}

/* Uniform cubic B-spline kernel, support [-2, 2]. */
static double b3(double x) {
    double const a = fabs(x);
    if (a < 1.) {
        double const y = 2. - a, z = 1. - a;
        return (y * y * y - 4. * z * z * z) / 6.;
    }
    if (a < 2.) {
        double const y = 2. - a;
        return y * y * y / 6.;
    }
    return 0.;
}

/* Cardinal cubic B-spline interpolant of Boost.Math 1.84
 * (boost/math/interpolators/cardinal_cubic_b_spline.hpp, used at
 * poly_table.hpp:124-134; the library is not vendored): s(x) = avg +
 * sum_k beta_k B3(x/h + 1 - k) interpolating y_k = f_k - avg at the nodes, with
 * the end slopes not given and estimated by the one-sided five-point O(h^4)
 * difference formulas. The (n+2) coefficients solve the tridiagonal system
 * beta_{k-1} + 4 beta_k + beta_{k+1} = 6 y_k, closed by the slope conditions
 * (beta_{k+1} - beta_{k-1}) / (2h) = s'(x_k) at both ends. */
static int bspline_build(bspline *s, double h, int n, const double *f) {
    if (n < 5)
        return -1;
    s->n = n;
    s->h = h;
    s->inv_h = 1. / h;
    double avg = 0.;
    for (int k = 0; k < n; ++k)
        avg += f[k];
    avg /= n;
    s->avg = avg;
    double const a1 = (-25. * f[0] + 48. * f[1] - 36. * f[2] + 16. * f[3] - 3. * f[4]) / (12. * h);
    double const b1 = (25. * f[n - 1] - 48. * f[n - 2] + 36. * f[n - 3] - 16. * f[n - 4] +
                       3. * f[n - 5]) / (12. * h);
    /* Unknowns x_k = beta_k (Kress index k = 0..n-1); Thomas algorithm. */
    double *cp = malloc(sizeof(double) * (size_t)n);
    double *dp = malloc(sizeof(double) * (size_t)n);
    s->beta = malloc(sizeof(double) * (size_t)(n + 2));
    if (!cp || !dp || !s->beta) {
        free(cp);
        free(dp);
        return -2;
    }
    /* row 0: 4 x0 + 2 x1 = 6 y0 + 2 h a1 */
    double b = 4., c = 2., d = 6. * (f[0] - avg) + 2. * h * a1;
    cp[0] = c / b;
    dp[0] = d / b;
    for (int k = 1; k < n; ++k) {
        double const a = (k == n - 1) ? 2. : 1.;
        c = (k == n - 1) ? 0. : 1.;
        d = 6. * (f[k] - avg) - ((k == n - 1) ? 2. * h * b1 : 0.);
        double const m = 4. - a * cp[k - 1];
        cp[k] = c / m;
        dp[k] = (d - a * dp[k - 1]) / m;
    }
    double *x = s->beta + 1;
    x[n - 1] = dp[n - 1];
    for (int k = n - 2; k >= 0; --k)
        x[k] = dp[k] - cp[k] * x[k + 1];
    s->beta[0] = x[1] - 2. * h * a1;
    s->beta[n + 1] = x[n - 2] + 2. * h * b1;
    free(cp);
    free(dp);
    return 0;
}

static double bspline_eval(const bspline *s, double x) {
    if (isnan(x))
        return NAN;
    double const t = x * s->inv_h + 1.;
    double lo = ceil(t - 2.), hi = floor(t + 2.);
    long const last = (long)s->n + 1;
    long kmin = lo < 0. ? 0 : (lo > (double)last ? last + 1 : (long)lo);
    long kmax = hi < 0. ? 0 : (hi > (double)last ? last : (long)hi);
    double z = s->avg;
    for (long k = kmin; k <= kmax; ++k)
        z += s->beta[k] * b3(t - (double)k);
    return z;
}

orc_table *orc_table_new(int kind, double h, int n_nodes, const double *f, const double *f1,
                         const double *f2) {
    /* poly_table.hpp:39-44: >= 2 nodes and h > 0 */
    if (n_nodes < 2 || !(h > 0.) || (kind != ORC_QUINTIC && kind != ORC_BSPLINE))
        return NULL;
    orc_table *t = calloc(1, sizeof *t);
    if (!t)
        return NULL;
    t->kind = kind;
    t->h = h;
    t->inv_h = 1. / h;
    t->inv_h2 = 1. / (h * h);
    t->n = n_nodes - 1;
    t->coef = malloc(sizeof(double) * 6 * (size_t)t->n);
    if (!t->coef) {
        free(t);
        return NULL;
    }
    quintic_build(t, f, f1, f2);
    if (kind == ORC_BSPLINE) {
        const double *src[3] = {f, f1, f2};
        for (int k = 0; k < 3; ++k)
            if (bspline_build(&t->s[k], h, n_nodes, src[k]) != 0) {
                orc_table_free(t);
                return NULL;
            }
    }
    return t;
}

void orc_table_free(orc_table *t) {
    if (!t)
        return;
    for (int k = 0; k < 3; ++k)
        free(t->s[k].beta);
    free(t->coef);
    free(t);
}

/* boost::math::pow<6> (Boost.Math special_functions/pow.hpp, a header-only
 * dependency of the reference, version unpinned): binary powering, x^6 = (x^3)^2
 * with x^3 = (x x) x. */
static double pow6_binary(double x) {
    double const x3 = x * x * x;
    return x3 * x3;
}

orc_table *orc_table_analytic(int kind, const double *param) {
    if ((kind != ORC_MORSE && kind != ORC_LJ) || !param)
        return NULL;
    orc_table *t = calloc(1, sizeof *t);
    if (!t)
        return NULL;
    t->kind = kind;
    t->prm[0] = param[0];
    t->prm[1] = param[1];
    if (kind == ORC_MORSE) {
        t->prm[2] = param[2];
    } else {
        /* LennardJones ctor, potential.hpp:55-59 */
        double const p6 = pow6_binary(param[1] / param[2]), p12 = p6 * p6;
        t->prm[2] = 4. * param[0] * (p12 - p6);
    }
    return t;
}

/* Morse::calculate (potential.hpp:40-49) and LennardJones::calculate (:65-73),
 * inside the cut-off (the callers apply it); which = 0, 1, 2 for V, V', V''. */
static double analytic_eval(const orc_table *t, double r, int which) {
    if (t->kind == ORC_MORSE) {
        double const D = t->prm[0], alpha = t->prm[1], r0 = t->prm[2];
        double const exp1 = exp(-alpha * (r - r0));
        double const exp2 = exp1 * exp1;
        if (which == 0)
            return D * (exp2 - 2. * exp1);
        if (which == 1)
            return -2. * D * alpha * (exp2 - exp1);
        return 2. * alpha * alpha * D * (2 * exp2 - exp1);
    }
    double const eps = t->prm[0], sigma = t->prm[1], vcut = t->prm[2];
    double const v6 = pow6_binary(sigma / r);
    double const v12 = v6 * v6;
    if (which == 0)
        return 4. * eps * (v12 - v6) - vcut;
    if (which == 1)
        return 24. * eps * (v6 - 2. * v12) / r;
    return 24. * eps * (26. * v12 - 7. * v6) / (r * r);
}

/* Table dispatch (poly_table.hpp:164-175), analytic forms (potential.hpp:29-75) */
double orc_table_value(const orc_table *t, double x) {
    if (t->kind == ORC_MORSE || t->kind == ORC_LJ)
        return analytic_eval(t, x, 0);
    return t->kind == ORC_QUINTIC ? quintic_value(t, x) : bspline_eval(&t->s[0], x);
}
double orc_table_d1(const orc_table *t, double x) {
    if (t->kind == ORC_MORSE || t->kind == ORC_LJ)
        return analytic_eval(t, x, 1);
    return t->kind == ORC_QUINTIC ? quintic_d1(t, x) : bspline_eval(&t->s[1], x);
}
double orc_table_d2(const orc_table *t, double x) {
    if (t->kind == ORC_MORSE || t->kind == ORC_LJ)
        return analytic_eval(t, x, 2);
    return t->kind == ORC_QUINTIC ? quintic_d2(t, x) : bspline_eval(&t->s[2], x);
}
const double *orc_table_coef(const orc_table *t, int *n_intervals) {
    if (n_intervals)
        *n_intervals = t->n;
    return t->kind == ORC_QUINTIC ? t->coef : NULL;
}

/* ------------------------------------------------------------------------- */
/* Thermostat chain (thermostat.{hpp,cpp}), barostat (barostat.{hpp,cpp})     */
/* ------------------------------------------------------------------------- */

typedef struct {
    double T, Nf;
    int M;
    double *q, *p_eta, *eta;
} chain;

/* numerics.hpp:11-18: exp(-x) and (1 - exp(-x))/x, 5th-order Taylor for |x| <= 1e-3 */
static void expx_operator(double x, double *expx, double *op) {
    static const double c1 = -1.0 / 2.0, c2 = 1.0 / 6.0, c3 = -1.0 / 24.0, c4 = 1. / 120.,
                        c5 = -1. / 720.;
    double const e = exp(-x);
    *expx = e;
    *op = fabs(x) > 1e-3 ? (1.0 - e) / x : 1.0 + x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * c5))));
}

/* numerics.hpp:22-27 */
static const double WSY7[7] = {
    0.784513610477560,  0.235573213359357,
    -1.17767998417887,  1. - 2. * (0.784513610477560 + 0.235573213359357 - 1.17767998417887),
    -1.17767998417887,  0.235573213359357,
    0.784513610477560};

/* thermostat.cpp:40-66 (Yu et al. 2010, Eq. 8; Allen-Tildesley 3.73, 3.74a) */
static void chain_U4(chain *c, double dt, double e_couple, int outward) {
    double const kT = K_BOLTZ * c->T;
    int const last = c->M - 1;
    for (int s = 0; s < c->M; ++s) {
        int const j = outward ? s : last - s;
        double const G = j != 0 ? c->p_eta[j - 1] * c->p_eta[j - 1] / c->q[j - 1] - kT
                                : 2. * e_couple - c->Nf * kT;
        if (j != last) {
            double e, op;
            expx_operator(dt * c->p_eta[j + 1] / c->q[j + 1], &e, &op);
            c->p_eta[j] = c->p_eta[j] * e + dt * G * op;
        } else {
            c->p_eta[j] += dt * G;
        }
    }
}

/* thermostat.hpp:24-35: scale factor of the coupled momenta, and eta update */
static double chain_U3(chain *c, double dt) {
    double const scale = exp(-dt * c->p_eta[0] / c->q[0]);
    for (int j = 0; j < c->M; ++j)
        c->eta[j] += dt * c->p_eta[j] / c->q[j];
    return scale;
}

/* thermostat.cpp:24-37 */
static double chain_etot(const chain *c) {
    double e = c->Nf * c->eta[0];
    for (int j = 1; j < c->M; ++j)
        e += c->eta[j];
    e *= K_BOLTZ * c->T;
    for (int j = 0; j < c->M; ++j)
        e += 0.5 * (c->p_eta[j] * c->p_eta[j]) / c->q[j];
    return e;
}

static int chain_init(chain *c, double T, double Nf, int M, const double *q, const double *p) {
    if (M < 1 || !q || !p)
        return -1;
    c->T = T;
    c->Nf = Nf;
    c->M = M;
    c->q = malloc(sizeof(double) * (size_t)M);
    c->p_eta = malloc(sizeof(double) * (size_t)M);
    c->eta = calloc((size_t)M, sizeof(double));
    if (!c->q || !c->p_eta || !c->eta)
        return -1;
    memcpy(c->q, q, sizeof(double) * (size_t)M);
    memcpy(c->p_eta, p, sizeof(double) * (size_t)M);
    return 0;
}

static void chain_free(chain *c) {
    free(c->q);
    free(c->p_eta);
    free(c->eta);
}

/* ------------------------------------------------------------------------- */
/* MD state (md_driver.hpp:87-167)                                            */
/* ------------------------------------------------------------------------- */

struct orc_md {
    int N, ntypes, has_eam, neighbors_list;
    double r_skin;
    int *point, *neighbors;
    size_t nb_cap;
    double Nf, alpha;
    double *pos, *coord, *vel, *force; /* N*3 */
    double *rho, *F, *dF, *d2F, *g, *h, *hsq, *epot, *ekin;
    double mom[3];
    double r_cut_sq, r_cutr_skin_sq;
    int npair;
    const orc_table **pair;
    double *pair_rcut, *pair_rmin;
    const orc_table **rho_t, **F_t;
    double *eam_rcut, *eam_rhomax;
    int ensemble, coupling;
    chain th;
    double bar_P, bar_W, bar_eps, bar_peps;
    chain bth;
    double box0[3], box[3];
    double *mass;
    unsigned *types;
    double epot_eam_atom, epot_atom, ekin_atom, etot_atom, ekin_tot;
    double etot_cons_initial, etot_cons, etot_cons_err;
    double virial, bst[6], bet[36];
    double lap, f2, Tk, Tci, P, st[6];
    double dt, time;
    unsigned long long cycle;
    double fatal;
    int sample;
    double sc_vir, sc_bst[6], sc_bet[36], sc_lap;
    long warn_virial, warn_rmin;
    int failed;
    char err[256];
};

static int fail(orc_md *md, int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(md->err, sizeof md->err, fmt, ap);
    va_end(ap);
    md->failed = code;
    return code;
}

/* md_driver.hpp:48-50: round half away from zero via truncation of x +- 0.5 */
static double nint(double x) {
    return (double)(x < 0. ? (int)(x - 0.5) : (int)(x + 0.5));
}

/* md_driver.cpp:175-185: wrap to [-0.5, 0.5] and scale by the box */
static void get_position(const orc_md *md, const double *c, double *p) {
    for (int a = 0; a < 3; ++a)
        p[a] = (c[a] - nint(c[a])) * md->box[a];
}

static double box_volume(const double *b) { return b[0] * (b[1] * b[2]); }

/* EAM / TablePotential lookups with their cut-off rules (eam.hpp:30-70,
 * potential.hpp:91-103) */
static double eam_rho(const orc_md *md, int t, double r) {
    return r > md->eam_rcut[t] ? 0. : orc_table_value(md->rho_t[t], r);
}
static double eam_drho(const orc_md *md, int t, double r) {
    return r > md->eam_rcut[t] ? 0. : orc_table_d1(md->rho_t[t], r);
}
static double eam_d2rho(const orc_md *md, int t, double r) {
    return r > md->eam_rcut[t] ? 0. : orc_table_d2(md->rho_t[t], r);
}
static int eam_F(orc_md *md, int t, double rho, double *F, double *dF, double *d2F) {
    if (rho > md->eam_rhomax[t])
        return fail(md, ORC_ERR_RHOMAX, "FATAL: EAM F rho >rhomax_fatal %g %g", rho,
                    md->eam_rhomax[t]);
    if (F)
        *F = orc_table_value(md->F_t[t], rho);
    if (dF)
        *dF = orc_table_d1(md->F_t[t], rho);
    if (d2F)
        *d2F = orc_table_d2(md->F_t[t], rho);
    return 0;
}
static void pair_calc(orc_md *md, int p, double r, double *V, double *dV, double *d2V) {
    if (r < md->pair_rmin[p])
        md->warn_rmin++;
    if (r > md->pair_rcut[p]) {
        *V = *dV = *d2V = 0.;
        return;
    }
    *V = orc_table_value(md->pair[p], r);
    *dV = orc_table_d1(md->pair[p], r);
    *d2V = orc_table_d2(md->pair[p], r);
}

static void nl_range(const orc_md *md, int i, int *start, int *end) {
    if (md->neighbors_list) {
        *start = md->point[i];
        *end = md->point[i + 1];
    } else {
        *start = i + 1;
        *end = md->N;
    }
}

/* md_driver.cpp:279-298: half list of j > i within r_c * r_skin */
int orc_md_calculate_neighbors(orc_md *md) {
    if (md->failed)
        return ORC_ERR_STATE;
    if (!md->neighbors_list)
        return 0;
    size_t k = 0;
    for (int i = 0; i < md->N; ++i) {
        const double *ci = md->coord + 3 * (size_t)i;
        md->point[i] = (int)k;
        for (int j = i + 1; j < md->N; ++j) {
            const double *cj = md->coord + 3 * (size_t)j;
            double s[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]}, d[3];
            get_position(md, s, d);
            if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > md->r_cutr_skin_sq)
                continue;
            if (k == md->nb_cap) {
                size_t cap = md->nb_cap ? 2 * md->nb_cap : 64;
                int *nb = realloc(md->neighbors, sizeof(int) * cap);
                if (!nb)
                    return fail(md, ORC_ERR_ARG, "out of memory in the neighbour list");
                md->neighbors = nb;
                md->nb_cap = cap;
            }
            md->neighbors[k++] = j;
        }
    }
    md->point[md->N] = (int)k;
    return 0;
}

/* md_driver.cpp:195-277: densities, and when sampled g_i (Voigt) and h_i */
static int calculate_eam(orc_md *md) {
    if (!md->has_eam)
        return 0;
    int const obs = md->sample;
    int const N = md->N;
    memset(md->rho, 0, sizeof(double) * (size_t)N);
    if (obs) {
        memset(md->h, 0, sizeof(double) * 3 * (size_t)N);
        memset(md->g, 0, sizeof(double) * 6 * (size_t)N);
    }
    for (int i = 0; i < N; ++i) {
        const double *ci = md->coord + 3 * (size_t)i;
        int const ti = (int)md->types[i];
        int start, end;
        nl_range(md, i, &start, &end);
        for (int k = start; k < end; ++k) {
            int const j = md->neighbors_list ? md->neighbors[k] : k;
            int const tj = (int)md->types[j];
            const double *cj = md->coord + 3 * (size_t)j;
            double s[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]}, d[3];
            get_position(md, s, d);
            double const r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 > md->r_cut_sq)
                continue;
            double const r = sqrt(r2);
            md->rho[i] += eam_rho(md, tj, r);
            md->rho[j] += eam_rho(md, ti, r);
            if (!obs)
                continue;
            double const wij = eam_drho(md, tj, r) / r;
            double const wji = eam_drho(md, ti, r) / r;
            double *gi = md->g + 6 * (size_t)i, *gj = md->g + 6 * (size_t)j;
            for (int v = 0; v < 6; ++v) {
                gi[v] += wij * d[VA[v]] * d[VB[v]];
                gj[v] += wji * d[VA[v]] * d[VB[v]];
            }
            double *hi = md->h + 3 * (size_t)i, *hj = md->h + 3 * (size_t)j;
            for (int a = 0; a < 3; ++a) {
                hi[a] += wij * d[a];
                hj[a] -= wji * d[a];
            }
        }
    }
    for (int i = 0; i < N; ++i) {
        int const ti = (int)md->types[i];
        int rc = eam_F(md, ti, md->rho[i], &md->F[i], &md->dF[i], obs ? &md->d2F[i] : NULL);
        if (rc)
            return rc;
        if (obs) {
            const double *hi = md->h + 3 * (size_t)i;
            md->hsq[i] = hi[0] * hi[0] + hi[1] * hi[1] + hi[2] * hi[2];
        }
    }
    return 0;
}

/* md_driver.cpp:300-430 (formulas of pass B); also the |term| sums */
int orc_md_calculate_forces(orc_md *md) {
    if (md->failed)
        return ORC_ERR_STATE;
    int rc = calculate_eam(md);
    if (rc)
        return rc;
    int const obs = md->sample;
    int const N = md->N;
    memset(md->force, 0, sizeof(double) * 3 * (size_t)N);
    memset(md->epot, 0, sizeof(double) * (size_t)N);
    md->virial = 0.;
    md->sc_vir = 0.;
    memset(md->bst, 0, sizeof md->bst);
    memset(md->sc_bst, 0, sizeof md->sc_bst);
    if (obs) {
        memset(md->bet, 0, sizeof md->bet);
        memset(md->sc_bet, 0, sizeof md->sc_bet);
        md->lap = 0.;
        md->sc_lap = 0.;
    }
    for (int i = 0; i < N; ++i) {
        double *fi = md->force + 3 * (size_t)i;
        const double *ci = md->coord + 3 * (size_t)i;
        double const dFi = md->has_eam ? md->dF[i] : 0.;
        double const d2Fi = md->has_eam ? md->d2F[i] : 0.;
        int const ti = (int)md->types[i];
        if (md->has_eam)
            md->epot[i] += md->F[i];
        int start, end;
        nl_range(md, i, &start, &end);
        for (int k = start; k < end; ++k) {
            int const j = md->neighbors_list ? md->neighbors[k] : k;
            int const tj = (int)md->types[j];
            const double *cj = md->coord + 3 * (size_t)j;
            double s[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]}, d[3];
            get_position(md, s, d);
            double const r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 > md->r_cut_sq)
                continue;
            int const tmax = ti > tj ? ti : tj, tmin = ti > tj ? tj : ti;
            int const p = tmax * (tmax + 1) / 2 + tmin;
            double const r = sqrt(r2);
            double V, dV, d2V;
            pair_calc(md, p, r, &V, &dV, &d2V);
            double dedr = dV, X = 0.;
            if (obs) {
                X = d2V - dV / r;
                double const t = 2. * (d2V + 2. * dV / r);
                md->lap += t;
                md->sc_lap += fabs(t);
            }
            if (md->has_eam) {
                double const dri = eam_drho(md, ti, r), drj = eam_drho(md, tj, r);
                double const dFj = md->dF[j];
                dedr += dFi * drj + dFj * dri;
                if (obs) {
                    double const d2ri = eam_d2rho(md, ti, r), d2rj = eam_d2rho(md, tj, r);
                    double const t1 = 2. * (dFj * (d2ri + 2. * dri / r) + dFi * (d2rj + 2. * drj / r));
                    double const t2 = d2Fi * (drj * drj) + md->d2F[j] * (dri * dri);
                    md->lap += t1;
                    md->lap += t2;
                    md->sc_lap += fabs(t1) + fabs(t2);
                    X += dFi * (d2rj - drj / r) + dFj * (d2ri - dri / r);
                }
            }
            X /= r * r;
            double const a = -dedr / r;
            for (int c = 0; c < 3; ++c) {
                fi[c] += a * d[c];
                md->force[3 * (size_t)j + c] -= a * d[c];
            }
            md->epot[i] += 0.5 * V;
            md->epot[j] += 0.5 * V;
            md->virial += r * dedr;
            md->sc_vir += fabs(r * dedr);
            for (int v = 0; v < 6; ++v) {
                double const num = d[VA[v]] * d[VB[v]];
                double const tb = dedr * num / r;
                md->bst[v] += tb;
                md->sc_bst[v] += fabs(tb);
                if (!obs)
                    continue;
                for (int w = v; w < 6; ++w) {
                    double const tc = num * d[VA[w]] * d[VB[w]] * X;
                    md->bet[6 * v + w] += tc;
                    md->sc_bet[6 * v + w] += fabs(tc);
                }
            }
        }
        if (md->has_eam && obs) {
            const double *gi = md->g + 6 * (size_t)i;
            for (int v = 0; v < 6; ++v)
                for (int w = v; w < 6; ++w) {
                    double const tc = d2Fi * gi[v] * gi[w];
                    md->bet[6 * v + w] += tc;
                    md->sc_bet[6 * v + w] += fabs(tc);
                }
            double const tl = d2Fi * md->hsq[i];
            md->lap += tl;
            md->sc_lap += fabs(tl);
        }
    }
    md->virial /= -3.;
    double const vtr = -(md->bst[0] + md->bst[1] + md->bst[2]);
    if (fabs(md->virial / (vtr / 3.) - 1.) > 1e-7)
        md->warn_virial++;
    return 0;
}

/* md_driver.cpp:603-609 */
static void set_kinetic_energies(orc_md *md) {
    double sum = 0.;
    for (int i = 0; i < md->N; ++i) {
        const double *v = md->vel + 3 * (size_t)i;
        md->ekin[i] = 0.5 * md->mass[md->types[i]] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        sum += md->ekin[i];
    }
    md->ekin_tot = sum;
    md->ekin_atom = sum / md->N;
}

/* md_driver.cpp:611-628 */
static int set_momentum_tot(orc_md *md) {
    md->mom[0] = md->mom[1] = md->mom[2] = 0.;
    for (int i = 0; i < md->N; ++i) {
        double const m = md->mass[md->types[i]];
        for (int a = 0; a < 3; ++a)
            md->mom[a] += m * md->vel[3 * (size_t)i + a];
    }
    double const sc2 = 2. * md->mass[0] * md->ekin_tot;
    double const p2 = md->mom[0] * md->mom[0] + md->mom[1] * md->mom[1] + md->mom[2] * md->mom[2];
    if (p2 / sc2 > 1e-14)
        return fail(md, ORC_ERR_MOMENTUM, "FATAL: total momentum = (%g, %g, %g) scale=%g",
                    md->mom[0], md->mom[1], md->mom[2], sqrt(sc2));
    return 0;
}

/* md_driver.cpp:630-646 */
static void conf_temperature(orc_md *md) {
    if (!md->sample)
        return;
    double f2 = 0.;
    for (int i = 0; i < 3 * md->N; ++i)
        f2 += md->force[i] * md->force[i];
    md->f2 = f2;
    md->Tci = (f2 != 0. && md->lap != 0.) ? K_BOLTZ * md->lap / f2 : 0.;
}

/* md_driver.cpp:648-730 (without the printout) */
static int calculate_thermodynamics(orc_md *md) {
    int rc = set_momentum_tot(md);
    if (rc)
        return rc;
    int const N = md->N;
    double se = 0., sf = 0.;
    for (int i = 0; i < N; ++i)
        se += md->epot[i];
    for (int i = 0; i < N; ++i)
        sf += md->F[i];
    md->epot_atom = se / N;
    md->epot_eam_atom = md->has_eam ? sf / N : 0.;
    md->etot_atom = md->ekin_atom + md->epot_atom;
    double const old = md->etot_cons;
    md->etot_cons = md->etot_atom * N;
    if (md->ensemble == ORC_NVT || md->ensemble == ORC_NPT)
        md->etot_cons += chain_etot(&md->th);
    double const V = box_volume(md->box);
    if (md->ensemble == ORC_NPT)
        md->etot_cons += V * md->bar_P + (0.5 * (md->bar_peps * md->bar_peps) / md->bar_W +
                                          chain_etot(&md->bth));
    md->etot_cons_err = fabs(md->etot_cons / md->etot_cons_initial - 1.);
    double const err_cur = fabs(md->etot_cons / old - 1.);
    md->Tk = 2. * md->ekin_tot / (K_BOLTZ * md->Nf);
    conf_temperature(md);
    md->P = (md->ekin_tot * 2. / 3. + md->virial) / V;
    for (int v = 0; v < 6; ++v)
        md->st[v] = -md->bst[v];
    for (int i = 0; i < N; ++i) {
        const double *vi = md->vel + 3 * (size_t)i;
        double const m = md->mass[md->types[i]];
        for (int v = 0; v < 6; ++v)
            md->st[v] += m * vi[VA[v]] * vi[VB[v]];
    }
    for (int v = 0; v < 6; ++v) {
        md->st[v] /= V;
        md->bst[v] /= V;
    }
    if (md->sample)
        for (int v = 0; v < 36; ++v)
            md->bet[v] /= V;
    double const ptr = md->st[0] + md->st[1] + md->st[2];
    double const perr = fabs(md->P / (ptr / 3.) - 1.);
    if (perr > 1e-5)
        return fail(md, ORC_ERR_PRESSURE, "FATAL PRESS err=%g %1.16E %1.16E", perr, md->P, ptr / 3.);
    int const bad = (md->etot_cons_err > md->fatal || err_cur > md->fatal) && md->cycle > 0;
    if (bad)
        return fail(md, ORC_ERR_CONSERVATION, "FATAL: bad energy conservation (err %g, %g)",
                    md->etot_cons_err, err_cur);
    return 0;
}

/* U2, md_driver.cpp:433-452 */
static void U2(orc_md *md, double dt) {
    double c = 1., e = 1.;
    if (md->ensemble == ORC_NPT)
        expx_operator(md->alpha * dt * md->bar_peps / md->bar_W, &e, &c);
    for (int i = 0; i < md->N; ++i) {
        double const m = md->mass[md->types[i]];
        for (int a = 0; a < 3; ++a) {
            double *v = md->vel + 3 * (size_t)i + a;
            *v = *v * e + (c * dt) * md->force[3 * (size_t)i + a] / m;
        }
    }
    set_kinetic_energies(md);
    conf_temperature(md);
}

/* U2p, md_driver.cpp:454-459 (Yu et al. 2010, Eq. 31) */
static void U2p(orc_md *md, double dt) {
    double const pv = md->bar_P * box_volume(md->box);
    md->bar_peps += (2. * md->alpha * md->ekin_tot + 3. * md->virial - 3. * pv) * dt;
}

/* U1, md_driver.cpp:462-499 */
static int U1(orc_md *md, double dt) {
    double cf = 1.;
    if (md->ensemble == ORC_NPT) {
        double const x = dt * md->bar_peps / md->bar_W;
        md->bar_eps += x;
        double e;
        expx_operator(x, &e, &cf);
    }
    for (int i = 0; i < md->N; ++i) {
        double *c = md->coord + 3 * (size_t)i;
        for (int a = 0; a < 3; ++a)
            c[a] += (cf * dt) * md->vel[3 * (size_t)i + a] / md->box[a];
        get_position(md, c, md->pos + 3 * (size_t)i);
    }
    if (md->ensemble == ORC_NPT) {
        double const expe = exp(md->bar_eps);
        double const vold = box_volume(md->box);
        for (int a = 0; a < 3; ++a)
            md->box[a] = md->box0[a] * expe;
        if (fabs(box_volume(md->box) / vold - 1.) > 0.05)
            return fail(md, ORC_ERR_VOLUME, "FATAL: volume change too large... expe=%g", expe);
    }
    return 0;
}

/* md_driver.cpp:512-531: the energy the thermostat chain couples to */
static double thermo_couple(orc_md *md) {
    double e = md->ekin_tot;
    if (md->coupling != ORC_KINETIC && md->Tci != 0.) {
        double const econf = (K_BOLTZ / md->Tci) * md->Nf / 2.;
        e = md->coupling == ORC_CONFIGURATIONAL ? econf : 0.5 * (md->ekin_tot + econf);
    }
    return e;
}

/* md_driver.cpp:539-549 */
static void U_NHC_thermo(orc_md *md, double dt) {
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        chain_U4(&md->th, 0.5 * dtw, thermo_couple(md), 0);
        double const s = chain_U3(&md->th, dtw);
        for (int i = 0; i < 3 * md->N; ++i)
            md->vel[i] *= s;
        set_kinetic_energies(md);
        chain_U4(&md->th, 0.5 * dtw, thermo_couple(md), 1);
    }
}

/* md_driver.cpp:551-561 */
static void U_NHC_baro(orc_md *md, double dt) {
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        chain_U4(&md->bth, 0.5 * dtw, 0.5 * (md->bar_peps * md->bar_peps) / md->bar_W, 0);
        md->bar_peps *= chain_U3(&md->bth, dtw);
        chain_U4(&md->bth, 0.5 * dtw, 0.5 * (md->bar_peps * md->bar_peps) / md->bar_W, 1);
    }
}

/* md_driver.cpp:563-601 */
int orc_md_integrate(orc_md *md) {
    if (md->failed)
        return ORC_ERR_STATE;
    int const npt = md->ensemble == ORC_NPT;
    int const thermo = md->ensemble == ORC_NVT || npt;
    double const h = 0.5 * md->dt;
    int rc;
    if (npt)
        U_NHC_baro(md, h);
    if (thermo)
        U_NHC_thermo(md, h);
    if (npt)
        U2p(md, h);
    U2(md, h);
    if ((rc = U1(md, md->dt)))
        return rc;
    if ((rc = orc_md_calculate_forces(md)))
        return rc;
    U2(md, h);
    if (npt)
        U2p(md, h);
    if (thermo)
        U_NHC_thermo(md, h);
    if (npt)
        U_NHC_baro(md, h);
    if ((rc = calculate_thermodynamics(md)))
        return rc;
    md->time += md->dt;
    md->cycle++;
    return 0;
}

void orc_md_set_sample(orc_md *md, int sample) { md->sample = sample ? 1 : 0; }

void orc_md_free(orc_md *md) {
    if (!md)
        return;
    free(md->point);
    free(md->neighbors);
    double *arrs[] = {md->pos, md->coord, md->vel, md->force, md->rho, md->F, md->dF,
                      md->d2F, md->g, md->h, md->hsq, md->epot, md->ekin, md->pair_rcut,
                      md->pair_rmin, md->eam_rcut, md->eam_rhomax, md->mass};
    for (size_t k = 0; k < sizeof arrs / sizeof arrs[0]; ++k)
        free(arrs[k]);
    free((void *)md->pair);
    free((void *)md->rho_t);
    free((void *)md->F_t);
    free(md->types);
    chain_free(&md->th);
    chain_free(&md->bth);
    free(md);
}

static double *dalloc(size_t n) { return calloc(n ? n : 1, sizeof(double)); }

/* MDDriver ctor, md_driver.cpp:16-160 */
orc_md *orc_md_new(const orc_config *cfg, int *err, char *msg, int msglen) {
    orc_md *md = calloc(1, sizeof *md);
    *err = ORC_ERR_ARG;
    if (!md)
        return NULL;
    int const N = cfg->N, T = cfg->ntypes;
    md->N = N;
    md->ntypes = T;
    md->has_eam = cfg->has_eam;
    md->neighbors_list = cfg->neighbors_list;
    md->r_skin = cfg->r_skin;
    md->Nf = cfg->Nf;
    md->alpha = 1. + 3. / cfg->Nf;
    md->ensemble = cfg->ensemble;
    md->coupling = cfg->coupling;
    md->dt = cfg->dt;
    md->fatal = cfg->fatal_conservation;
    md->sample = 1;
    md->npair = T * (T + 1) / 2;
    size_t const n3 = 3 * (size_t)N;
    md->pos = dalloc(n3);
    md->coord = dalloc(n3);
    md->vel = dalloc(n3);
    md->force = dalloc(n3);
    md->rho = dalloc(N);
    md->F = dalloc(N);
    md->dF = dalloc(N);
    md->d2F = dalloc(N);
    md->g = dalloc(6 * (size_t)N);
    md->h = dalloc(n3);
    md->hsq = dalloc(N);
    md->epot = dalloc(N);
    md->ekin = dalloc(N);
    md->mass = dalloc(T);
    md->types = calloc((size_t)N + 1, sizeof(unsigned));
    md->pair = calloc((size_t)md->npair, sizeof(orc_table *));
    md->pair_rcut = dalloc(md->npair);
    md->pair_rmin = dalloc(md->npair);
    md->rho_t = calloc((size_t)T, sizeof(orc_table *));
    md->F_t = calloc((size_t)T, sizeof(orc_table *));
    md->eam_rcut = dalloc(T);
    md->eam_rhomax = dalloc(T);
    if (md->neighbors_list) {
        md->point = calloc((size_t)N + 1, sizeof(int));
        md->nb_cap = 40 * (size_t)N;
        md->neighbors = calloc(md->nb_cap ? md->nb_cap : 1, sizeof(int));
    }
    if (chain_init(&md->th, cfg->th_T, cfg->th_Nf, cfg->th_M, cfg->th_q, cfg->th_p_eta) ||
        chain_init(&md->bth, cfg->bth_T, cfg->bth_Nf, cfg->bth_M, cfg->bth_q, cfg->bth_p_eta) ||
        !md->types || !md->mass || !md->pair || !md->rho_t || !md->F_t ||
        (md->neighbors_list && (!md->point || !md->neighbors)) || N < 1 || T < 1) {
        snprintf(msg, (size_t)msglen, "bad configuration or out of memory");
        orc_md_free(md);
        return NULL;
    }
    md->bar_P = cfg->bar_P;
    md->bar_W = cfg->bar_W;
    md->bar_peps = cfg->bar_p_eps;
    md->bar_eps = 0.;
    memcpy(md->mass, cfg->mass, sizeof(double) * (size_t)T);
    for (int p = 0; p < md->npair; ++p) {
        md->pair[p] = cfg->pair[p];
        md->pair_rcut[p] = cfg->pair_rcut[p];
        md->pair_rmin[p] = cfg->pair_rmin[p];
    }
    if (md->has_eam)
        for (int t = 0; t < T; ++t) {
            md->rho_t[t] = cfg->rho_tab[t];
            md->F_t[t] = cfg->F_tab[t];
            md->eam_rcut[t] = cfg->eam_rcut[t];
            md->eam_rhomax[t] = cfg->eam_rhomax[t];
        }
    md->r_cut_sq = md->pair_rcut[0] * md->pair_rcut[0];
    md->r_cutr_skin_sq = (md->pair_rcut[0] * md->r_skin) * (md->pair_rcut[0] * md->r_skin);
    for (int a = 0; a < 3; ++a)
        md->box0[a] = md->box[a] = cfg->box[a];
    memcpy(md->pos, cfg->pos, sizeof(double) * n3);
    memcpy(md->vel, cfg->vel, sizeof(double) * n3);
    for (int i = 0; i < N; ++i) {
        const double *p = cfg->pos + 3 * (size_t)i;
        if (fabs(p[0]) >= 0.5 * md->box[0] || fabs(p[1]) >= 0.5 * md->box[1] ||
            fabs(p[2]) >= 0.5 * md->box[2]) {
            *err = ORC_ERR_POSITION;
            snprintf(msg, (size_t)msglen, "bad initial position");
            orc_md_free(md);
            return NULL;
        }
        double *c = md->coord + 3 * (size_t)i;
        for (int a = 0; a < 3; ++a) {
            double const s = p[a] / md->box[a];
            c[a] = s - nint(s);
        }
        if (cfg->types[i] >= (unsigned)T) {
            *err = ORC_ERR_TYPE;
            snprintf(msg, (size_t)msglen, "bad type");
            orc_md_free(md);
            return NULL;
        }
        md->types[i] = cfg->types[i];
    }
    if (md->ensemble != ORC_NVE && md->ensemble != ORC_NVT && md->ensemble != ORC_NPT) {
        *err = ORC_ERR_ENSEMBLE;
        snprintf(msg, (size_t)msglen, "bad ensamble %d", md->ensemble);
        orc_md_free(md);
        return NULL;
    }
    int rc = orc_md_calculate_neighbors(md);
    if (!rc)
        rc = orc_md_calculate_forces(md);
    if (!rc) {
        set_kinetic_energies(md);
        rc = calculate_thermodynamics(md);
    }
    if (rc) {
        *err = rc;
        snprintf(msg, (size_t)msglen, "%s", md->err);
        orc_md_free(md);
        return NULL;
    }
    md->etot_cons_initial = md->etot_cons;
    *err = 0;
    return md;
}

const char *orc_md_error(const orc_md *md) { return md->err; }

void orc_md_scalars(const orc_md *md, double *o) {
    memset(o, 0, sizeof(double) * ORC_NSCALARS);
    o[ORC_S_EPOT_EAM_ATOM] = md->epot_eam_atom;
    o[ORC_S_EPOT_ATOM] = md->epot_atom;
    o[ORC_S_EKIN_ATOM] = md->ekin_atom;
    o[ORC_S_ETOT_ATOM] = md->etot_atom;
    o[ORC_S_EKIN_TOT] = md->ekin_tot;
    o[ORC_S_ETOT_CONS_INITIAL] = md->etot_cons_initial;
    o[ORC_S_ETOT_CONS] = md->etot_cons;
    o[ORC_S_ETOT_CONS_ERR] = md->etot_cons_err;
    o[ORC_S_VIRIAL] = md->virial;
    o[ORC_S_LAPLACIAN] = md->lap;
    o[ORC_S_FORCE2] = md->f2;
    o[ORC_S_TKIN] = md->Tk;
    o[ORC_S_TCONF_INV] = md->Tci;
    o[ORC_S_PRESSURE] = md->P;
    o[ORC_S_TIME] = md->time;
    o[ORC_S_CYCLE] = (double)md->cycle;
    for (int a = 0; a < 3; ++a) {
        o[ORC_S_BOX0 + a] = md->box[a];
        o[ORC_S_MOM0 + a] = md->mom[a];
    }
    o[ORC_S_BAR_EPS] = md->bar_eps;
    o[ORC_S_BAR_PEPS] = md->bar_peps;
    for (int v = 0; v < 6; ++v) {
        o[ORC_S_BORN_STRESS + v] = md->bst[v];
        o[ORC_S_STRESS + v] = md->st[v];
    }
    for (int v = 0; v < 36; ++v)
        o[ORC_S_BORN_ELAST + v] = md->bet[v];
}

void orc_md_scales(const orc_md *md, double *o) {
    o[0] = md->sc_vir;
    for (int v = 0; v < 6; ++v)
        o[1 + v] = md->sc_bst[v];
    for (int v = 0; v < 36; ++v)
        o[7 + v] = md->sc_bet[v];
    o[43] = md->sc_lap;
}

int orc_md_atoms(const orc_md *md, int what, double *out) {
    size_t const N = (size_t)md->N;
    const double *src = NULL;
    size_t n = N;
    switch (what) {
    case ORC_A_POSITIONS: src = md->pos; n = 3 * N; break;
    case ORC_A_COORDINATES: src = md->coord; n = 3 * N; break;
    case ORC_A_VELOCITIES: src = md->vel; n = 3 * N; break;
    case ORC_A_FORCES: src = md->force; n = 3 * N; break;
    case ORC_A_RHO: src = md->rho; break;
    case ORC_A_EPOT: src = md->epot; break;
    case ORC_A_EKIN: src = md->ekin; break;
    case ORC_A_F: src = md->F; break;
    case ORC_A_DF: src = md->dF; break;
    case ORC_A_D2F: src = md->d2F; break;
    default: return ORC_ERR_ARG;
    }
    memcpy(out, src, sizeof(double) * n);
    return 0;
}

int orc_md_chain(const orc_md *md, int which, double *p_eta, double *eta) {
    const chain *c = which ? &md->bth : &md->th;
    memcpy(p_eta, c->p_eta, sizeof(double) * (size_t)c->M);
    memcpy(eta, c->eta, sizeof(double) * (size_t)c->M);
    return 0;
}

long orc_md_nlist_size(const orc_md *md) {
    return md->neighbors_list ? md->point[md->N] : -1;
}

long orc_md_warnings(const orc_md *md, int which) {
    return which == 0 ? md->warn_virial : md->warn_rmin;
}

/* M_PI of <math.h> (not defined under -std=c99) */
static const double PI_D = 3.14159265358979323846;

/* md_driver.cpp:732-863: pair correlations over all pairs i < j whose
 * minimum-image distance is below rmaxg = rg[nbin-1] + drg/2 (md_driver.cpp:41),
 * counted in bin floor(r / drg); mode 0 total, 1 total and per species pair,
 * 2 per species pair only; the array a mode does not compute is not touched. */
int orc_md_pair_correlation(orc_md *md, const double *rg, int nbin, int mode, double *gr,
                            double *gr_partial) {
    if (nbin < 2 || mode < 0 || mode > 2)
        return fail(md, 1, "pair correlation: bad arguments");
    double const drg = rg[1] - rg[0], rmaxg = rg[nbin - 1] + 0.5 * drg;
    int const S = md->ntypes, npair = S * (S + 1) / 2;
    int const tot = mode != 2, part = mode != 0;
    if (tot)
        for (int k = 0; k < nbin; ++k)
            gr[k] = 0.;
    if (part)
        for (int k = 0; k < npair * nbin; ++k)
            gr_partial[k] = 0.;
    for (int i = 0; i < md->N; ++i) {
        const double *ci = md->coord + 3 * (size_t)i;
        for (int j = i + 1; j < md->N; ++j) {
            const double *cj = md->coord + 3 * (size_t)j;
            double const s[3] = {ci[0] - cj[0], ci[1] - cj[1], ci[2] - cj[2]};
            double d[3];
            get_position(md, s, d);
            double const dr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (dr < rmaxg) {
                int const k = (int)floor(dr / drg);
                if (k >= nbin)
                    continue;
                if (tot)
                    gr[k] += 1.;
                if (part) {
                    unsigned const a = md->types[i], b = md->types[j];
                    unsigned const tmax = a > b ? a : b, tmin = a > b ? b : a;
                    gr_partial[(size_t)(tmax * (tmax + 1) / 2 + tmin) * nbin + k] += 1.;
                }
            }
        }
    }
    double const V = box_volume(md->box), Np = (double)md->N;
    if (tot)
        for (int k = 0; k < nbin; ++k)
            gr[k] *= V / (2. * PI_D * Np * Np * rg[k] * rg[k] * drg);
    if (part) {
        double *cnt = calloc((size_t)S, sizeof(double));
        for (int i = 0; i < md->N; ++i)
            cnt[md->types[i]] += 1.;
        int q = 0;
        for (int b = 0; b < S; ++b)
            for (int a = 0; a <= b; ++a, ++q) {
                double const Na = cnt[a], Nb = cnt[b];
                double pc = 0.;
                if (a == b)
                    pc = Na > 1 ? Na * (Na - 1.) / 2. : 0.;
                else
                    pc = Na * Nb;
                double *g = gr_partial + (size_t)q * nbin;
                if (pc > 0.) {
                    double const pref = V / (4. * PI_D * pc * drg);
                    for (int k = 0; k < nbin; ++k)
                        g[k] *= pref / (rg[k] * rg[k]);
                } else {
                    for (int k = 0; k < nbin; ++k)
                        g[k] = 0.;
                }
            }
        free(cnt);
    }
    return 0;
}
