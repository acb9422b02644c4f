/*
 * mdelast-mi355x: C-ABI of the MI355X-native EAM step (libmdelast_hip.so).
 *
 * This is the drop-in boundary that replaces the pybind11 module
 * `md_driver_module` of mdelast v1.2 (src/md_driver_module.cpp:16-224) for the
 * per-step EAM path. Plain C types only: pointers, sizes, doubles, ints. All
 * quantities are CGS, as at the reference boundary (src/units.hpp:6-24).
 *
 * Every function returns an mdl_status; on failure mdl_last_error() holds a
 * message for the calling thread. No function calls exit(): the FATAL paths
 * of the reference (printf + exit(1)) become MDL_ERR_* codes.
 *
 * Threading: one driver handle must not be used from two threads at once.
 * Each driver owns one HIP stream on the current device at creation.
 */
#ifndef MDELAST_H_
#define MDELAST_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mdl_status;

enum {
    MDL_OK = 0,
    MDL_ERR_ARG = 1,          /* invalid argument */
    MDL_ERR_POSITION = 2,     /* initial |x_a| >= L_a/2 (md_driver.cpp:91-97) */
    MDL_ERR_TYPE = 3,         /* type >= ntypes (md_driver.cpp:101-104) */
    MDL_ERR_ENSEMBLE = 4,     /* unknown/unsupported ensemble (md_driver.cpp:141-144) */
    MDL_ERR_RHOMAX = 5,       /* rho > rho_max in F(rho) (eam.hpp:48-70) */
    MDL_ERR_VOLUME = 6,       /* NPT volume change > 5 % in one step (md_driver.cpp:493-497) */
    MDL_ERR_MOMENTUM = 7,     /* |P|^2 / (2 m_0 K) > 1e-14 (md_driver.cpp:620-627) */
    MDL_ERR_PRESSURE = 8,     /* |P / (tr sigma / 3) - 1| > 1e-5 (md_driver.cpp:693-698) */
    MDL_ERR_CONSERVATION = 9, /* E_cons error > fatal_conservation (md_driver.cpp:700-725) */
    MDL_ERR_STATE = 10,       /* the driver failed before; it accepts no more steps */
    MDL_ERR_DEVICE = 11,      /* HIP runtime error (message holds the HIP error string) */
    MDL_ERR_UNSUPPORTED = 12, /* a configuration the GPU path does not implement */
    MDL_ERR_NEIGHBORS = 13,   /* neighbour list or cell grid overflow */
    MDL_ERR_TABLE = 14        /* invalid table (poly_table.hpp:39-44) */
};

/* Table interpolation, poly_table.hpp:140-152; analytic pair forms,
 * potential.hpp:29-75 (Morse, LennardJones) */
enum { MDL_QUINTIC = 0, MDL_BSPLINE = 1, MDL_MORSE = 2, MDL_LJ = 3 };
/* MDDriver::Ensamble and ThermostatCoupling, md_driver.hpp:20-21 */
enum { MDL_NVE = 0, MDL_NVT = 1, MDL_NPT = 2, MDL_NPT_FLEX = 3 };
enum { MDL_KINETIC = 0, MDL_CONFIGURATIONAL = 1, MDL_AVERAGE = 2 };
/* Arithmetic of the pair/table passes */
enum { MDL_FP64 = 0, MDL_MIXED = 1 };

/* Message of the last failed call on this thread ("" if none). */
const char *mdl_last_error(void);
/* Name and compute units of the current HIP device; MDL_ERR_DEVICE without one. */
mdl_status mdl_device_info(char *name, int32_t name_len, int32_t *compute_units);

/* ---------------------------------------------------------------------------
 * Tables. Replaces potential::Table / QuinticTable / BSplineTable
 * (poly_table.hpp:33-176), the tabulation inside TablePotential
 * (potential.hpp:77-104) and EAM (eam.hpp:10-71).
 * ------------------------------------------------------------------------- */
typedef struct mdl_table mdl_table;

/* Nodes x_k = k h, k = 0..n_nodes-1, with f, f', f'' at the nodes.
 * MDL_QUINTIC: per interval the quintic Hermite polynomial (poly_table.hpp:35-63),
 * coefficients formed in long double. MDL_BSPLINE: three cardinal cubic
 * B-splines through f, f', f'' (poly_table.hpp:120-135), n_nodes >= 5.
 * Host-only: no device memory is touched until a driver or eval uses it. */
mdl_status mdl_table_create(int32_t kind, double h, int64_t n_nodes, const double *f,
                            const double *f1, const double *f2, mdl_table **out);
/* An analytic pair potential as a table handle. Replaces potential::Morse and
 * potential::LennardJones (potential.hpp:29-75, bound at md_driver_module.cpp:45-56).
 * MDL_MORSE: param = {D, alpha, r0}, V = D (e^2 - 2 e), e = exp(-alpha (r - r0)).
 * MDL_LJ: param = {epsilon, sigma, rcut}, V = 4 epsilon ((s/r)^12 - (s/r)^6) - vcut
 * with vcut the unshifted value at rcut. Neither cuts at rcut here: the driver
 * takes the cut-off from mdl_config.pair_rcut, as for tables. A driver accepts
 * these handles as its pair potentials (all of one form) when has_eam = 0, in
 * MDL_FP64; mdl_table_eval evaluates them; mdl_table_coefficients returns one
 * record of stride 4, the device parameters (Morse {D, alpha, r0, 0}, LJ
 * {epsilon, sigma, vcut, 0}). */
mdl_status mdl_table_create_analytic(int32_t kind, const double *param, mdl_table **out);
mdl_status mdl_table_destroy(mdl_table *t);
/* Host copy of the per-interval polynomial coefficients: quintic [n][6]
 * (value polynomial in t); B-spline [n][12] (cubics of f, f', f'' in t, 4 each).
 * Pass out = NULL to query *n_intervals and *stride. */
mdl_status mdl_table_coefficients(const mdl_table *t, double *out, int64_t *n_intervals,
                                  int32_t *stride);
/* Evaluate value, d1, d2 at n points on the GPU (poly_table.hpp:71-97, 164-175);
 * any output pointer may be NULL. Host arrays in and out. */
mdl_status mdl_table_eval(mdl_table *t, int64_t n, const double *x, double *value, double *d1,
                          double *d2);

/* ---------------------------------------------------------------------------
 * Driver. Replaces MDDriver (md_driver.hpp:18-168): ctor md_driver.cpp:16-160,
 * integrate() :563-601, calculate_neighbors() :279-298, and the attribute
 * getters bound at md_driver_module.cpp:103-140.
 * ------------------------------------------------------------------------- */
typedef struct mdl_driver mdl_driver;

typedef struct {
    int64_t N;                /* atoms */
    int32_t ntypes;           /* species = len(mass) */
    const double *positions;  /* N*3, |x_a| < L_a/2 */
    const double *velocities; /* N*3 */
    const uint32_t *types;    /* N */
    const double *mass;       /* ntypes, per type */
    double box[3];            /* diagonal box (md_driver.cpp:60) */
    int32_t ensemble;         /* MDL_NVE / NVT / NPT */
    int32_t coupling;         /* MDL_KINETIC / CONFIGURATIONAL / AVERAGE */
    /* Thermostat(T, q, p_eta, Nf) (thermostat.hpp:17-18) */
    double th_T, th_Nf;
    int32_t th_M;
    const double *th_q, *th_p_eta;
    /* Barostat(P, w, p_epsilon, Thermostat) (barostat.hpp:12-13) */
    double bar_P, bar_W, bar_p_eps;
    double bth_T, bth_Nf;
    int32_t bth_M;
    const double *bth_q, *bth_p_eta;
    double Nf, dt;
    int32_t neighbors_list;   /* reference list on/off; the GPU always keeps an exact list */
    double r_skin;            /* list radius factor (>= 1) */
    double fatal_conservation;
    /* pair potentials, ntypes(ntypes+1)/2 in slot j(j+1)/2 + i, i <= j */
    const mdl_table *const *pair;
    const double *pair_rcut, *pair_rmin;
    int32_t has_eam;
    /* EAM per type: rho(r) and F(rho) tables, cut-off and rho_max */
    const mdl_table *const *rho_tab;
    const mdl_table *const *F_tab;
    const double *eam_rcut, *eam_rhomax;
    int32_t precision;        /* MDL_FP64 or MDL_MIXED */
    int32_t lanes_per_atom;   /* 0 = autotuned; else 4, 8, 16, 32 or 64 */
} mdl_config;

/* Copies the state to the device, builds the list, computes forces and the
 * thermodynamics of the initial state (md_driver.cpp:16-160). */
mdl_status mdl_driver_create(const mdl_config *cfg, mdl_driver **out);
mdl_status mdl_driver_destroy(mdl_driver *d);

/* One call = nsteps consecutive integrate() (md_driver.cpp:563-601) with the
 * current sample_observables flag, then a synchronisation and error check. */
mdl_status mdl_driver_integrate(mdl_driver *d, int64_t nsteps);
/* Production loop of MolecularDynamics.run (molecular_dynamics.py:1240-1255)
 * without the Python: nsteps steps; step k samples iff (cycle+1) % sample_stride
 * == 0; a list rebuild is forced before a step whose cycle % rebuild_every == 0
 * (0 = never forced). If records != NULL, each sampled step appends
 * MDL_NRECORD doubles (the MDL_S_* layout below, first MDL_NRECORD entries).
 * *n_records receives the count. */
mdl_status mdl_driver_run(mdl_driver *d, int64_t nsteps, int64_t sample_stride,
                          int64_t rebuild_every, double *records, int64_t *n_records);
/* mdl_driver_run on n drivers at once: the launches of step k of every driver,
 * each on its own stream, then those of step k + 1, so the systems advance
 * side by side on the device (independent replicas of uneven size,
 * composition and temperature). records[i] and n_records[i] as for
 * mdl_driver_run (records or records[i] may be NULL). No counterpart in the
 * reference, whose replicas run as separate processes. */
mdl_status mdl_drivers_run(mdl_driver *const *drivers, int32_t n, int64_t nsteps,
                           int64_t sample_stride, int64_t rebuild_every, double *const *records,
                           int64_t *n_records);
/* mdl_drivers_run with every replica in the same launches: each pass of a
 * step is one launch over the blocks of all n replicas (the pair passes one
 * launch per group of equal interpolation kind and lanes per atom). Each
 * replica keeps its own grid, partial sums and their order, so its results
 * are bit for bit those of mdl_driver_run. The list check runs as one block
 * per replica, and the whole step is one chain on the first driver's stream.
 * The drivers must use mixed precision and share device and cycle; at most
 * 4 * 256 atom blocks per replica (MDL_ERR_ARG before any driver is touched
 * otherwise). A requested rebuild is taken first by one ordinary step of
 * every driver. No counterpart in the reference (batched small replicas). */
mdl_status mdl_replicas_run(mdl_driver *const *drivers, int32_t n, int64_t nsteps,
                            int64_t sample_stride, int64_t rebuild_every, double *const *records,
                            int64_t *n_records);
/* mdl_driver_run and mdl_drivers_run replay chunks of `steps` consecutive
 * steps as one captured HIP graph (0, the default: launch step by step). A
 * chunk is used where its first cycle is a multiple of `steps`, `steps` is a
 * multiple of sample_stride and of rebuild_every (when > 0), no rebuild was
 * requested and the run's final step is not in it; other steps launch as
 * usual. The arithmetic and its order are those of the step-by-step launches.
 * No counterpart in the reference (a launch-overhead measure for small
 * systems and many replicas). */
mdl_status mdl_driver_set_graph_steps(mdl_driver *d, int32_t steps);
/* Pair correlation of the current configuration: MDDriver::
 * calculate_pair_correlation (mode 0: gr), calculate_pair_correlations (mode 1:
 * gr and gr_partial) and calculate_partial_pair_correlation (mode 2: gr_partial)
 * of md_driver.cpp:732-863. Every pair i < j whose minimum-image distance r is
 * below rg[nbin-1] + drg/2 (drg = rg[1] - rg[0], rg cell-centred as the
 * driver's rg) counts in bin floor(r / drg); the counts are normalised as
 * there. gr: nbin doubles; gr_partial: S(S+1)/2 rows of nbin (species pair
 * (a <= b) in row b(b+1)/2 + a), host memory; the array a mode does not
 * compute is not touched. Synchronous. */
mdl_status mdl_driver_pair_correlation(mdl_driver *d, const double *rg, int32_t nbin,
                                       int32_t mode, double *gr, double *gr_partial);
/* mdl_driver_pair_correlation without any allocation: the same results, bit for
 * bit, with the bin counters in a per-device scratch array of the library
 * (MDL_GR_SCRATCH_COUNTS 64-bit counters, 4 MiB, held on each GPU the call
 * runs on until the process ends) and the counts read back through a fixed
 * host buffer. No counterpart in the reference: it is for callers that cannot
 * allocate (loops under a no-allocation rule, code between captured work).
 * (1 + S(S+1)/2) * nbin above MDL_GR_SCRATCH_COUNTS gives MDL_ERR_ARG. Calls on
 * one GPU share the scratch; the library serialises its own calls, and other
 * code overlapping them on that GPU would race. Synchronous. */
#define MDL_GR_SCRATCH_COUNTS 524288
mdl_status mdl_driver_pair_correlation_noalloc(mdl_driver *d, const double *rg, int32_t nbin,
                                               int32_t mode, double *gr, double *gr_partial);
/* Force a rebuild before the next force evaluation (md_driver.cpp:279-298). */
mdl_status mdl_driver_calculate_neighbors(mdl_driver *d);
/* MDDriver::sample_observables (md_driver.hpp:163-167) */
mdl_status mdl_driver_set_sample_observables(mdl_driver *d, int32_t flag);
mdl_status mdl_driver_get_sample_observables(mdl_driver *d, int32_t *flag);

/* Scalar attributes, MDL_NSCALARS doubles in this order. */
enum {
    MDL_S_EPOT_EAM_ATOM = 0, MDL_S_EPOT_ATOM, MDL_S_EKIN_ATOM, MDL_S_ETOT_ATOM, MDL_S_EKIN_TOT,
    MDL_S_ETOT_CONS_INITIAL, MDL_S_ETOT_CONS, MDL_S_ETOT_CONS_ERR, MDL_S_VIRIAL,
    MDL_S_LAPLACIAN, MDL_S_FORCE2, MDL_S_TKIN, MDL_S_TCONF_INV, MDL_S_PRESSURE, MDL_S_TIME,
    MDL_S_CYCLE, MDL_S_BOX0, MDL_S_BOX1, MDL_S_BOX2, MDL_S_BAR_EPS, MDL_S_BAR_PEPS,
    MDL_S_MOM0, MDL_S_MOM1, MDL_S_MOM2,
    MDL_S_BORN_STRESS = 24,   /* 6, Voigt xx yy zz yz xz xy, divided by V */
    MDL_S_STRESS = 30,        /* 6 */
    MDL_S_BORN_ELAST = 36,    /* 36, row-major 6x6, upper triangle filled, lower 0 */
    MDL_NRECORD = 72,
    MDL_S_BOX_INITIAL0 = 72, MDL_S_BOX_INITIAL1, MDL_S_BOX_INITIAL2,
    MDL_S_NLIST_BUILDS = 75,  /* list builds since creation */
    MDL_S_WARN_VIRIAL = 76,   /* steps whose virial self-check failed (md_driver.cpp:422-429) */
    MDL_S_WARN_RMIN = 77,     /* pair distances below r_min seen (potential.hpp:93-95) */
    MDL_S_NLIST_ENTRIES = 78, /* stored (full) list entries */
    MDL_S_NHC_GENERAL = 79,   /* chain propagations with a factor in the general form */
    MDL_S_NMID_BUILDS = 80,   /* filterings of the mid list (pairs within r_c + 0.3 A) */
    MDL_S_LANES_PER_ATOM = 81,/* lanes per atom of the pair passes (given or autotuned) */
    MDL_NSCALARS = 82
};
mdl_status mdl_driver_get_scalars(mdl_driver *d, double *out);

/* Per-atom arrays (N*3 or N doubles). POSITIONS are wrapped real positions as
 * the reference keeps them (md_driver.cpp:485), COORDINATES the unwrapped
 * scaled ones. */
enum {
    MDL_A_POSITIONS = 0, MDL_A_COORDINATES, MDL_A_VELOCITIES, MDL_A_FORCES, MDL_A_RHO,
    MDL_A_EPOT, MDL_A_EKIN,
    MDL_A_PAIRS_RC = 7,       /* N: pairs within r_c of each atom at the last density pass
                                 (the pairs the force pass evaluates; EAM drivers only) */
    MDL_A_PAIRS_MID = 8       /* N: entries of each atom's mid list (pairs within r_c + 0.3 A,
                                 mixed: 0.15 A, at its last filtering; the density pass's
                                 scan length, and in mixed precision the force pass's) */
};
mdl_status mdl_driver_get_atoms(mdl_driver *d, int32_t what, double *out);
/* p_eta and eta of the thermostat (which = 0) or barostat chain (which = 1). */
mdl_status mdl_driver_get_chain(mdl_driver *d, int32_t which, double *p_eta, double *eta);

/* Average device time per launch of the two pair kernels over the last
 * mdl_driver_run, measured with HIP events on the driver's stream:
 * out[0] density pass, out[1] force pass, out[2] whole step (ms). */
mdl_status mdl_driver_kernel_times(mdl_driver *d, double *out);

/* exp(-x) and (1 - exp(-x))/x at n points on the GPU, as the integrator takes
 * them (expx_operator, numerics.hpp:11-27; thermostat.cpp:40-66). form 0: the
 * barostat's form; 1: the chain sweeps' wide form (any x); 2: their series form
 * (|x| <= 1/4). Host arrays in and out. */
mdl_status mdl_expx_eval(int32_t form, int64_t n, const double *x, double *expx, double *op);

#ifdef __cplusplus
}
#endif
#endif /* MDELAST_H_ */
