// Internal definitions of libmdelast_hip: host table objects, the device state
// record, and the device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mdelast.h"

namespace mdl {

// ----------------------------------------------------------------------------
// Units (src/units.hpp:6-24) and fixed sizes
// ----------------------------------------------------------------------------
constexpr double K_BOLTZ = 1.38064852e-16;
constexpr int MAX_TYPES = 8;                         // species per driver
constexpr int MAX_PAIRS = MAX_TYPES * (MAX_TYPES + 1) / 2;
constexpr int MAX_CHAIN = 64;                        // Nose-Hoover chain length
constexpr int BLOCK = 256;                           // threads per block of the atom kernels

// Per-block partial sums written by the atom kernels and reduced by the
// scalar kernel in a fixed order (deterministic).
enum PartialA { PA_EF = 0, PA_N = 1 };
enum PartialB {
    PB_EPOT = 0, PB_VIR = 1, PB_SB = 2 /*6*/, PB_CB = 8 /*15 monomials*/, PB_LAP = 23,
    PB_WRMIN = 24, PB_N = 25
};
enum PartialK {
    PK_K = 0, PK_MVV = 1 /*6*/, PK_MOM = 7 /*3*/, PK_F2 = 10, PK_CBF = 11 /*21*/, PK_LAPF = 32,
    PK_N = 33
};
// The 15 distinct degree-4 monomials of the pair vector d in the pair Born term
enum Monomial {
    M_X4, M_Y4, M_Z4, M_X3Y, M_X3Z, M_Y3X, M_Y3Z, M_Z3X, M_Z3Y, M_X2Y2, M_X2Z2, M_Y2Z2,
    M_X2YZ, M_Y2XZ, M_Z2XY
};

// Error bits raised on the device (mirrors MDL_ERR_*).
enum DevErr : int {
    DE_RHOMAX = 1 << 0,
    DE_VOLUME = 1 << 1,
    DE_MOMENTUM = 1 << 2,
    DE_PRESSURE = 1 << 3,
    DE_CONSERVATION = 1 << 4,
    DE_NBR_OVERFLOW = 1 << 5,
    DE_CELL_GRID = 1 << 6,
    DE_BARRIER = 1 << 7,
    DE_CELL_OVERFLOW = 1 << 8,   // one cell holds more atoms than the build stages
};

// ----------------------------------------------------------------------------
// Host table object (the C handle mdl_table)
// ----------------------------------------------------------------------------
struct HostTable {
    int kind = MDL_QUINTIC;
    double h = 0, inv_h = 0, inv_h2 = 0;
    int64_t n = 0;                  // intervals
    int stride = 6;                 // doubles per interval record (6 quintic, 12 B-spline)
    std::vector<double> rec;        // [n][stride]
    // device copy for mdl_table_eval (lazily created)
    double *d_rec = nullptr;
    int device = -1;
};

// ----------------------------------------------------------------------------
// Device-resident scalar state of one driver (single copy, updated by the
// one-thread scalar kernels; the host reads it back for attributes)
// ----------------------------------------------------------------------------
struct Chain {
    double T, Nf;
    int M;
    double q[MAX_CHAIN], iq[MAX_CHAIN], p_eta[MAX_CHAIN], eta[MAX_CHAIN];   // iq = 1/q
};

struct DevState {
    // configuration
    int N, ntypes, ensemble, coupling, has_eam, sample;
    double Nf, alpha, dt, fatal;
    double mass[MAX_TYPES];
    // box
    double L[3], L0[3], L_drift[3], L_build[3];
    // thermostat, barostat
    Chain th, bth;
    double bar_P, bar_W, eps, peps;
    // velocity bookkeeping: stored v times vscale is the physical velocity
    double vscale;
    // coefficients of the current step (written by k_pre)
    double kick_s, kick_e, kick_c, drift_cf, drift_L[3];
    // thermodynamics (the MDL_S_* record)
    double rec[MDL_NSCALARS];
    double unit_E, unit_L;   // mixed mode: eV -> erg and eV/A^2 -> erg/cm^2 (1 in FP64)
    int mixed, pad_mixed;
    double ekin_tot, virial, lap, f2, Tci;
    double etot_cons_initial, etot_cons, etot_cons_err;
    double time;
    unsigned long long cycle;
    // neighbour-list control. The mid list holds the pairs of the full list
    // within r_mid; pass A scans it instead of the full list while no pair
    // outside it can have entered r_c, and refilters it from the full list
    // (mid_build) after every full build and whenever that could have happened.
    double r_mid, L_mid[3];
    int mid_build, pad_mid;
    long long nmid_builds;
    double r_list, skin, rc;
    double umax2;            // max squared displacement since the last build (k_kick_drift)
    int force_rebuild, do_rebuild;
    long long nbuilds, nbr_needed, nlist_entries;
    int ncell[3];
    // errors and warnings
    int err;
    double err_val[4];
    long long warn_virial, warn_rmin;
    long long nhc_general;   // chain propagations with a factor from the general form
    int nhc_form[2];         // thermostat, barostat chain: the sweep form its last
                             // propagation needed (SW_SERIES, SW_TOP, SW_WIDE)
    int init_done;
};

// ----------------------------------------------------------------------------
// Device table lookups. Records of all r-functions (pair slots first, then
// rho per type) are interleaved per interval: rec[(i*nf + f)*S + c], so the
// functions a pair needs sit next to each other.
// ----------------------------------------------------------------------------
struct RTables {
    const double *rec;
    int nf, n;              // functions per interval, intervals
    double inv_h, inv_h2;
    double rc_f[MAX_PAIRS + MAX_TYPES];   // per-function cut-off (0 beyond)
    double rmin_p[MAX_PAIRS];
};
struct FTables {
    const double *rec;      // [n][ntypes][S]
    int n;
    double inv_h, inv_h2;
    double rhomax[MAX_TYPES];
};

// Interval and local coordinate, range test before the conversion
// (poly_table.hpp:105-114): NaN -> interval 0, outside -> end polynomial.
__device__ __forceinline__ int tab_locate(double x, double inv_h, int n, double &t) {
    double const u = x * inv_h;
    int i = 0;
    if (u >= (double)n)
        i = n - 1;
    else if (u > 0.)
        i = (int)u;
    t = u - (double)i;
    return i;
}

template <int KIND> struct Poly;

// Quintic: value/d1/d2 of one polynomial (poly_table.hpp:71-97)
template <> struct Poly<MDL_QUINTIC> {
    static constexpr int S = 6;
    double c[6];
    __device__ __forceinline__ void load(const double *p) {
        const double2 *q = reinterpret_cast<const double2 *>(p);
        double2 a = q[0], b = q[1], e = q[2];
        c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y; c[4] = e.x; c[5] = e.y;
    }
    __device__ __forceinline__ double v(double t) const {
        return c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
    }
    __device__ __forceinline__ double d1(double t, double ih) const {
        return (c[1] + t * (2. * c[2] + t * (3. * c[3] + t * (4. * c[4] + t * 5. * c[5])))) * ih;
    }
    __device__ __forceinline__ double d2(double t, double ih2) const {
        return (2. * c[2] + t * (6. * c[3] + t * (12. * c[4] + t * 20. * c[5]))) * ih2;
    }
};

// B-spline kind: three independent cubics (f, f', f'') per interval.
template <> struct Poly<MDL_BSPLINE> {
    static constexpr int S = 12;
    double c[12];
    __device__ __forceinline__ void load(const double *p) {
        const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double2 a = q[k];
            c[2 * k] = a.x;
            c[2 * k + 1] = a.y;
        }
    }
    __device__ __forceinline__ double v(double t) const {
        return c[0] + t * (c[1] + t * (c[2] + t * c[3]));
    }
    __device__ __forceinline__ double d1(double t, double) const {
        return c[4] + t * (c[5] + t * (c[6] + t * c[7]));
    }
    __device__ __forceinline__ double d2(double t, double) const {
        return c[8] + t * (c[9] + t * (c[10] + t * c[11]));
    }
};

// Analytic pair forms. The device record of one pair slot is 4 doubles:
// Morse {D, alpha, r0, 0}, Lennard-Jones {epsilon, sigma, vcut, 0}.
constexpr int PAIR_PARAMS = 4;
constexpr bool analytic_kind(int kind) { return kind == MDL_MORSE || kind == MDL_LJ; }

// (s)^6 by binary powering, as boost::math::pow<6> forms it: ((s s) s)^2
__host__ __device__ __forceinline__ double pow6_binary(double s) {
    double const s3 = s * s * s;
    return s3 * s3;
}

// V, dV/dr, d2V/dr2 of the analytic forms inside the cut-off: Morse
// potential.hpp:40-49, Lennard-Jones potential.hpp:65-73 (same products, in
// the same order).
template <int KIND>
__host__ __device__ __forceinline__ void pair_analytic(const double *p, double r, double &V,
                                                       double &dV, double &d2V) {
    if constexpr (KIND == MDL_MORSE) {
        double const D = p[0], alpha = p[1], r0 = p[2];
        double const e1 = exp(-alpha * (r - r0));
        double const e2 = e1 * e1;
        V = D * (e2 - 2. * e1);
        dV = -2. * D * alpha * (e2 - e1);
        d2V = 2. * alpha * alpha * D * (2. * e2 - e1);
    } else {
        double const eps = p[0], sigma = p[1], vcut = p[2];
        double const v6 = pow6_binary(sigma / r);
        double const v12 = v6 * v6;
        V = 4. * eps * (v12 - v6) - vcut;
        dV = 24. * eps * (v6 - 2. * v12) / r;
        d2V = 24. * eps * (26. * v12 - 7. * v6) / (r * r);
    }
}

// Round half away from zero (md_driver.hpp:48-50): trunc(x -+ 0.5). The half
// is selected first, so one add and one trunc run instead of two of each
// (x + (-0.5) is x - 0.5 exactly).
__device__ __forceinline__ double nint_away(double x) {
    double const h = x < 0. ? -0.5 : 0.5;
    return trunc(x + h);
}

// Sums over aligned groups of W lanes (W a power of two, up to the wave) of NV
// components per lane, by recursive halving over the lane bits W/2, W/4, ...,
// 1: at each step a lane keeps one half of its current array and sends the
// other half to its partner across that bit, which keeps the other one
// (halves of ceil(A/2), the odd one padded with zero). About NV exchanges in
// all, where a butterfly per component takes NV log2(W). Every component is
// summed over the same tree as by that butterfly (pairs across bit W/2 first,
// then W/4, ...; the additions commute exactly), so the sums are its bits. On
// return, each lane holds the group's sum of component `idx`, or idx < 0 if
// its slot is padding; every component sits in exactly one lane of the group.
constexpr int halved(int n, int steps) {
    for (int k = 0; k < steps; ++k)
        n = (n + 1) / 2;
    return n;
}

template <int W, int NV, int S, typename T>
__device__ __forceinline__ void halving_steps(T (&v)[NV], int lw, int &lo, int &hi) {
    if constexpr ((W >> (S + 1)) >= 1) {
        constexpr int A = halved(NV, S), H = (A + 1) / 2, OFF = (W / 2) >> S;
        bool const up = (lw & OFF) != 0;
#pragma unroll
        for (int i = 0; i < H; ++i) {
            T const a = v[i];
            T const b = H + i < A ? v[H + i] : T(0);
            T const keep = up ? b : a;
            v[i] = keep + __shfl_xor(up ? a : b, OFF, W);
        }
        int const mid = lo + H;
        if (up)
            lo = mid;
        else
            hi = min(hi, mid);
        halving_steps<W, NV, S + 1>(v, lw, lo, hi);
    }
}

template <int W, int NV, typename T>
__device__ __forceinline__ T group_sum_scatter(T (&v)[NV], int &idx) {
    static_assert(W >= 1 && W <= 64 && (W & (W - 1)) == 0, "group of 2^k lanes");
    static_assert(NV >= 1 && NV <= W, "one component per lane at most");
    int lo = 0, hi = NV;
    halving_steps<W, NV, 0>(v, (int)threadIdx.x & (W - 1), lo, hi);
    idx = lo < hi ? lo : -1;
    return v[0];
}

}  // namespace mdl

// Error reporting shared by the translation units
namespace mdl {
void set_error(const char *fmt, ...);
mdl_status hip_fail(hipError_t e, const char *what);
}  // namespace mdl

#define MDL_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            return ::mdl::hip_fail(e_, #call);                                  \
    } while (0)
