/*
 * CPU ORACLE — test infrastructure only.
 *
 * Single-threaded FP64 restatement of the reference's per-step EAM path
 * (mdelast v1.2, src/md_driver.cpp, src/poly_table.hpp, src/eam.hpp,
 * src/potential.hpp, src/thermostat.{hpp,cpp}, src/barostat.{hpp,cpp},
 * src/numerics.hpp). It is the checker for the HIP library and the timed
 * CPU baseline; the product never links, loads or calls it.
 *
 * Differences from the reference are deliberate and limited to error
 * handling: every printf+exit(1) FATAL path returns an error code and leaves
 * a message in orc_md_error(); soft warnings are counted, not printed.
 */
#ifndef MDELAST_ORACLE_H_
#define MDELAST_ORACLE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_QUINTIC = 0, ORC_BSPLINE = 1, ORC_MORSE = 2, ORC_LJ = 3 };
enum { ORC_NVE = 0, ORC_NVT = 1, ORC_NPT = 2, ORC_NPT_FLEX = 3 };
enum { ORC_KINETIC = 0, ORC_CONFIGURATIONAL = 1, ORC_AVERAGE = 2 };

enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,         /* invalid argument / table */
    ORC_ERR_POSITION = 2,    /* bad initial position (md_driver.cpp:91-97) */
    ORC_ERR_TYPE = 3,        /* bad type (md_driver.cpp:101-104) */
    ORC_ERR_ENSEMBLE = 4,    /* bad ensemble (md_driver.cpp:141-144) */
    ORC_ERR_RHOMAX = 5,      /* rho > rho_max in F, F', F'' (eam.hpp:48-70) */
    ORC_ERR_VOLUME = 6,      /* NPT volume change > 5 % (md_driver.cpp:493-497) */
    ORC_ERR_MOMENTUM = 7,    /* net momentum (md_driver.cpp:620-627) */
    ORC_ERR_PRESSURE = 8,    /* pressure vs stress trace (md_driver.cpp:693-698) */
    ORC_ERR_CONSERVATION = 9,/* energy conservation (md_driver.cpp:700-725) */
    ORC_ERR_STATE = 10       /* the driver already failed */
};

typedef struct orc_table orc_table;

/* Quintic: QuinticTable ctor (poly_table.hpp:35-63). B-spline: three cardinal
 * cubic B-splines through f, f1, f2 (poly_table.hpp:120-135). */
orc_table *orc_table_new(int kind, double h, int n_nodes, const double *f, const double *f1,
                         const double *f2);
/* Analytic pair forms behind the same evaluation calls: potential::Morse
 * (potential.hpp:29-51), param {D, alpha, r0}; potential::LennardJones
 * (:53-75), param {epsilon, sigma, rcut}. No cut-off inside: pair_calc applies it. */
orc_table *orc_table_analytic(int kind, const double *param);
void orc_table_free(orc_table *t);
double orc_table_value(const orc_table *t, double x);
double orc_table_d1(const orc_table *t, double x);
double orc_table_d2(const orc_table *t, double x);
/* Quintic coefficients [n_intervals][6]; NULL for B-spline tables. */
const double *orc_table_coef(const orc_table *t, int *n_intervals);

/* One MD configuration; arrays are borrowed for the duration of orc_md_new. */
typedef struct {
    int N, ntypes;
    const double *pos;      /* N*3, CGS, |p| < L/2 */
    const double *vel;      /* N*3 */
    const unsigned *types;  /* N */
    const double *mass;     /* ntypes */
    double box[3];          /* diagonal of h */
    int ensemble, coupling;
    /* thermostat (Thermostat ctor, thermostat.cpp:11-22) */
    double th_T, th_Nf;
    int th_M;
    const double *th_q, *th_p_eta;
    /* barostat (Barostat ctor, barostat.cpp:6-10) and its chain */
    double bar_P, bar_W, bar_p_eps;
    double bth_T, bth_Nf;
    int bth_M;
    const double *bth_q, *bth_p_eta;
    double Nf, dt;
    int neighbors_list;
    double r_skin, fatal_conservation;
    /* pair tables, ntypes(ntypes+1)/2, slot j(j+1)/2+i */
    const orc_table *const *pair;
    const double *pair_rcut, *pair_rmin;
    /* EAM per type; has_eam = 0 for a pair potential only */
    int has_eam;
    const orc_table *const *rho_tab;
    const orc_table *const *F_tab;
    const double *eam_rcut, *eam_rhomax;
} orc_config;

typedef struct orc_md orc_md;

/* MDDriver ctor (md_driver.cpp:16-160). On error returns NULL and sets *err. */
orc_md *orc_md_new(const orc_config *cfg, int *err, char *msg, int msglen);
void orc_md_free(orc_md *md);
const char *orc_md_error(const orc_md *md);

int orc_md_integrate(orc_md *md);              /* md_driver.cpp:563-601 */
int orc_md_calculate_neighbors(orc_md *md);    /* md_driver.cpp:279-298 */
int orc_md_calculate_forces(orc_md *md);       /* md_driver.cpp:300-430 (with :195-277) */
void orc_md_set_sample(orc_md *md, int sample);

/* Scalar state, in this order (orc_md_scalars fills ORC_NSCALARS doubles). */
enum {
    ORC_S_EPOT_EAM_ATOM, ORC_S_EPOT_ATOM, ORC_S_EKIN_ATOM, ORC_S_ETOT_ATOM, ORC_S_EKIN_TOT,
    ORC_S_ETOT_CONS_INITIAL, ORC_S_ETOT_CONS, ORC_S_ETOT_CONS_ERR, ORC_S_VIRIAL,
    ORC_S_LAPLACIAN, ORC_S_FORCE2, ORC_S_TKIN, ORC_S_TCONF_INV, ORC_S_PRESSURE, ORC_S_TIME,
    ORC_S_CYCLE, ORC_S_BOX0, ORC_S_BOX1, ORC_S_BOX2, ORC_S_BAR_EPS, ORC_S_BAR_PEPS,
    ORC_S_MOM0, ORC_S_MOM1, ORC_S_MOM2,
    ORC_S_BORN_STRESS = 24,      /* 6 */
    ORC_S_STRESS = 30,           /* 6 */
    ORC_S_BORN_ELAST = 36,       /* 36, row-major, upper triangle filled */
    ORC_NSCALARS = 72
};
void orc_md_scalars(const orc_md *md, double *out);

/* Sums of |pair contribution| of the last force evaluation, the scales of the
 * parity metric: [0] virial, [1..6] born stress, [7..42] born elasticity (6x6),
 * [43] Laplacian. Not divided by the volume. */
enum { ORC_NSCALES = 44 };
void orc_md_scales(const orc_md *md, double *out);

/* Per-atom arrays (N*3 or N doubles). */
enum {
    ORC_A_POSITIONS, ORC_A_COORDINATES, ORC_A_VELOCITIES, ORC_A_FORCES, ORC_A_RHO,
    ORC_A_EPOT, ORC_A_EKIN, ORC_A_F, ORC_A_DF, ORC_A_D2F
};
int orc_md_atoms(const orc_md *md, int what, double *out);
/* Thermostat (which=0) or barostat-chain (which=1) p_eta and eta, M values each. */
int orc_md_chain(const orc_md *md, int which, double *p_eta, double *eta);
/* Neighbour-list size (point[N]) and soft-warning counters. */
long orc_md_nlist_size(const orc_md *md);
/* md_driver.cpp:732-863 (mode 0 gr, 1 gr and gr_partial, 2 gr_partial) */
int orc_md_pair_correlation(orc_md *md, const double *rg, int nbin, int mode, double *gr,
                            double *gr_partial);
long orc_md_warnings(const orc_md *md, int which);  /* 0 virial self-check, 1 r < r_min */

#ifdef __cplusplus
}
#endif
#endif
