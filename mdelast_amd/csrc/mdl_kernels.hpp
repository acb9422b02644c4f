// Device kernels of one MD step. Included by mdl_driver.hip only.
//
// Step = k_pre, or the previous step's k_finish (scalars: NHC, barostat,
//        kick/drift coefficients)
//      -> k_kick_drift (first half kick + drift, displacement since the list build)
//      -> k_list (rebuild decision; the rebuild itself when one is due)
//      -> k_density (pass A) -> k_force (pass B) -> k_kick2 (second half kick)
//      -> k_finish (reductions, NHC, thermodynamics, checks; then the next
//         step's scalar first half)
// which is MDDriver::integrate (md_driver.cpp:563-601) with the velocity
// scalings of the Nose-Hoover chain folded into one factor per half step.
#pragma once

#include "mdl_internal.hpp"

namespace mdl {

struct AtomArrays {
    double4 *pos;      // unwrapped scaled coordinates c (md_driver.hpp:97), type in .w
    int4 *qpos;        // the same as 32-bit fixed point (quantize), type in .w: the
                       // mixed passes' pair geometry
    double *vx, *vy, *vz;
    double *fx, *fy, *fz;
    int *type;
    double4 *cref;     // coordinates at the last list build
    double4 *mref;     // coordinates at the last build of the mid list
    double2 *fprime;   // F'(rho_i), F''(rho_i)
    double2 *geam;     // sampled steps: g_i (6, Voigt) and |h_i|^2, 4 double2 per atom
    float2 *fprime_f;  // mixed mode: F', F'' in eV A^3, eV A^6 for the pair gathers
    double *Femb;      // F(rho_i)
    double *rho;       // rho_i
    double *epot;      // per-atom potential energy
};

struct ListArrays {
    int *nbr;          // [N][stride] full list (both i->j and j->i)
    int *nnbr;         // [N]
    int *cnbr;         // [N][stride] pairs inside r_c, compacted by pass A
    int *ccnt;         // [N]
    int *mnbr;         // [N][stride] pairs inside r_mid (r_c < r_mid <= r_list), list order
    int *mcnt;         // [N]
    int stride;
    int *cell_of, *cell_count, *cell_start, *cell_fill, *cell_atoms;
    int ncells;
};

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// Partial records are stored component-major, part[k * PART_STRIDE + block],
// so the final reduction reads them coalesced.
constexpr int PART_STRIDE = 2048;
// dmax_part[DMAX_MID + b]: block b's largest squared displacement since the
// mid list was built (dmax_part[b]: since the full list was built)
constexpr int DMAX_MID = 1024;

// The block coordinates a kernel body works with: the hardware's, or one
// system's share of a launch that covers several systems.
struct Blk {
    int id, n;
};
__device__ __forceinline__ Blk hw_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }

template <int NV>
__device__ __forceinline__ double wave_sum_scatter(double (&v)[NV], int &idx) {
    return group_sum_scatter<64, NV>(v, idx);
}

// Block sums of NV components (wave sums, then the waves in order), stored as
// out[k * PART_STRIDE + bid]. acc is consumed.
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&acc)[NV], double *out, int bid) {
    int idx;
    double const w = wave_sum_scatter<NV>(acc, idx);
    __shared__ double sm[BLOCK / 64][NV];
    int const wave = threadIdx.x >> 6;
    if (idx >= 0)
        sm[wave][idx] = w;
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.;
#pragma unroll
        for (int k = 0; k < BLOCK / 64; ++k)
            s += sm[k][threadIdx.x];
        out[(size_t)threadIdx.x * PART_STRIDE + bid] = s;
    }
}

template <int NV>
__device__ __forceinline__ void block_sum_store(double (&acc)[NV], double *out) {
    block_sum_store<NV>(acc, out, (int)blockIdx.x);
}

// block_sum_store without the components [S0, S1): a step that does not sample
// leaves the sampled-only partials (Born and Laplacian terms) unreduced and
// their slots untouched; post_step reads them only on sampled steps. The other
// components are reduced exactly as by block_sum_store (the same tree per
// component, whatever its place in the array).
template <int NV, int S0, int S1>
__device__ __forceinline__ void block_sum_store_skip(double (&acc)[NV], double *out, int bid) {
    static_assert(0 <= S0 && S0 <= S1 && S1 <= NV && NV - (S1 - S0) >= 1, "skip range");
    constexpr int NK = NV - (S1 - S0);
    double v[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k)
        v[k] = acc[k < S0 ? k : k + (S1 - S0)];
    int idx;
    double const w = wave_sum_scatter<NK>(v, idx);
    __shared__ double sm[BLOCK / 64][NK];
    int const wave = threadIdx.x >> 6;
    if (idx >= 0)
        sm[wave][idx] = w;
    __syncthreads();
    if (threadIdx.x < NK) {
        double s = 0.;
#pragma unroll
        for (int k = 0; k < BLOCK / 64; ++k)
            s += sm[k][threadIdx.x];
        int const c = threadIdx.x < S0 ? threadIdx.x : threadIdx.x + (S1 - S0);
        out[(size_t)c * PART_STRIDE + bid] = s;
    }
}

// The partial sums of a pass: all components on sampled steps, the others
// without [S0, S1).
template <bool SAMPLE, int NV, int S0, int S1>
__device__ __forceinline__ void pass_sum_store(double (&acc)[NV], double *out, int bid) {
    if constexpr (SAMPLE)
        block_sum_store<NV>(acc, out, bid);
    else
        block_sum_store_skip<NV, S0, S1>(acc, out, bid);
}

// XCD-aware slice of [0, nitems): blocks are dispatched round-robin over the
// 8 XCDs, so block b runs on XCD b % 8; give each XCD one contiguous chunk so
// its L2 holds one spatial region of the neighbour gathers.
struct XcdRange {
    int first, end, step;
    __device__ __forceinline__ XcdRange(int nitems) : XcdRange(nitems, hw_blk()) {}
    __device__ __forceinline__ XcdRange(int nitems, Blk bk) {
        int const nper = bk.n >> 3;
        int const xcd = bk.id & 7, b = bk.id >> 3;
        int const chunk = (nitems + 7) >> 3;
        first = xcd * chunk + b;
        end = min(nitems, (xcd + 1) * chunk);
        step = nper;
    }
};

template <int LPA>
__device__ __forceinline__ unsigned long long group_mask() {
    int const l64 = threadIdx.x & 63;
    unsigned long long const m = LPA == 64 ? ~0ull : ((1ull << LPA) - 1ull);
    return m << (l64 & ~(LPA - 1));
}

// A scaled coordinate as 32-bit fixed point modulo the box: c 2^32 rounded to
// an integer, modulo 2^32. The wrapping difference of two atoms' values is then
// their minimum-image separation in units of L 2^-32 (a quantum of 2.5e-8 A in
// a 100 A box), with no rounding step.
__device__ __forceinline__ int quantize(double c) {
    return (int)(unsigned)(unsigned long long)__double2ll_rn(c * 4294967296.0);
}
__device__ __forceinline__ int4 quantize(const double4 &c) {
    return make_int4(quantize(c.x), quantize(c.y), quantize(c.z), (int)c.w);
}

// Pair vector d = x_i - x_j from fixed-point positions: the wrapping integer
// difference (the minimum image, exact), one rounding to float and one product
// by S = L 2^-32 (A). Exactly antisymmetric in (i, j), so a pair's two visits
// see the same r and opposite vectors.
__device__ __forceinline__ void qmin_image(const int4 &qi, const int4 &qj, const float *S,
                                           float &x, float &y, float &z) {
    x = (float)(int)((unsigned)qi.x - (unsigned)qj.x) * S[0];
    y = (float)(int)((unsigned)qi.y - (unsigned)qj.y) * S[1];
    z = (float)(int)((unsigned)qi.z - (unsigned)qj.z) * S[2];
}

// Minimum-image pair vector d = (s - nint(s)) L of s = c_i - c_j
// (md_driver.cpp:175-185, applied to the difference as at :222 and :333)
__device__ __forceinline__ void min_image(const double4 &ci, const double4 &cj, const double *L,
                                          double &dx, double &dy, double &dz) {
    double const sx = ci.x - cj.x, sy = ci.y - cj.y, sz = ci.z - cj.z;
    dx = (sx - nint_away(sx)) * L[0];
    dy = (sy - nint_away(sy)) * L[1];
    dz = (sz - nint_away(sz)) * L[2];
}

// exp(-x) and op(x) = (1 - exp(-x))/x of numerics.hpp:11-18. The reference
// takes exp from the library and op from a 5th-order Taylor form for
// |x| <= 1e-3, else from the quotient. The chain arguments are mostly small
// (|x| ~ dt/tau): for |x| <= 1/4, op comes from its Taylor series to degree 12
// (truncation below 2e-19 relative) and exp(-x) = 1 - x op(x) from one more
// FMA (within rounding of the library exp; the quotient form of op loses up to
// 1/|x| ulp to cancellation and agrees to that).
constexpr double EXPX_SMALL = 0.25;

// Both polynomials below are evaluated by Estrin's scheme: the chain sweeps
// run them on one lane, one after another, so their cost is the length of the
// dependent chain (4 FMA levels after the powers instead of 10-12 in Horner form).
__device__ __forceinline__ double op_series(double x) {
    // sum_{m=0}^{12} (-x)^m / (m+1)!
    double const y = -x, y2 = y * y, y4 = y2 * y2, y8 = y4 * y4;
    double const q0 = fma(y, 0.5, 1.), q1 = fma(y, 1. / 24., 1. / 6.),
                 q2 = fma(y, 1. / 720., 1. / 120.), q3 = fma(y, 1. / 40320., 1. / 5040.),
                 q4 = fma(y, 1. / 3628800., 1. / 362880.),
                 q5 = fma(y, 1. / 479001600., 1. / 39916800.);
    double const s0 = fma(q1, y2, q0), s1 = fma(q3, y2, q2), s2 = fma(q5, y2, q4);
    double const t0 = fma(s1, y4, s0), t1 = fma(1. / 6227020800., y4, s2);
    return fma(t1, y8, t0);
}

// exp(-x) for any x: y = -x = k ln2 + r with |r| <= ln2/2 (Cody-Waite: ln2
// split into a 32-bit head, so k*head is exact, and a tail), e^r = 1 + r p(r)
// with p the Taylor series of (e^r - 1)/r to degree 12 (truncation below 5e-18
// relative), then scaled by 2^k. Within 1 ulp of the correctly rounded value
// over the whole range; overflow gives inf and underflow 0 through the scaling.
__device__ __forceinline__ double exp_neg_general(double x) {
    double const y = -x;
    double const k = rint(y * 1.4426950408889634);   // y / ln 2
    double r = fma(k, -0x1.62e42ff000000p-1, y);     // ln2 head (32 significant bits)
    r = fma(k, 0x1.718432a1b0e26p-35, r);            // -(ln2 tail)
    // p(r) = sum_{m=0}^{12} r^m / (m+1)!
    double const r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    double const q0 = fma(r, 0.5, 1.), q1 = fma(r, 1. / 24., 1. / 6.),
                 q2 = fma(r, 1. / 720., 1. / 120.), q3 = fma(r, 1. / 40320., 1. / 5040.),
                 q4 = fma(r, 1. / 3628800., 1. / 362880.),
                 q5 = fma(r, 1. / 479001600., 1. / 39916800.);
    double const s0 = fma(q1, r2, q0), s1 = fma(q3, r2, q2), s2 = fma(q5, r2, q4);
    double const t0 = fma(s1, r4, s0), t1 = fma(1. / 6227020800., r4, s2);
    double const p = fma(t1, r8, t0);
    return ldexp(fma(r, p, 1.), (int)k);
}

// 1/x from the hardware estimate and two Newton steps (within 1 ulp for
// normal x). It does not depend on exp(-x), so it runs beside that chain, and
// (1 - exp(-x)) * (1/x) replaces the quotient that would follow it (one
// rounding more).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.);
    r = fma(r, e, r);
    e = fma(-x, r, 1.);
    return fma(r, e, r);
}

__device__ __forceinline__ double exp_neg(double x) {
    return fabs(x) <= EXPX_SMALL ? fma(-x, op_series(x), 1.) : exp_neg_general(x);
}

__device__ __forceinline__ void expx_operator(double x, double &expx, double &op) {
    if (fabs(x) <= EXPX_SMALL) {
        op = op_series(x);
        expx = fma(-x, op, 1.);
    } else {
        double const ix = rcp_nr(x);
        expx = exp_neg_general(x);
        op = (1.0 - expx) * ix;
    }
}

__constant__ const double WSY7[7] = {
    0.784513610477560, 0.235573213359357, -1.17767998417887,
    1. - 2. * (0.784513610477560 + 0.235573213359357 - 1.17767998417887),
    -1.17767998417887, 0.235573213359357, 0.784513610477560};

// The chain sweeps are one long sequence of dependent scalar FP64 operations
// run by a single lane, so their cost is their instruction count: 1/q_j is
// precomputed (Chain::iq) and p/q becomes p * (1/q) (one rounding more than
// the reference's division, far below the parity tolerances).

// thermostat.cpp:40-66
__device__ void chain_U4(Chain &c, double dt, double e_couple, bool outward) {
    double const kT = K_BOLTZ * c.T;
    int const last = c.M - 1;
    for (int s = 0; s < c.M; ++s) {
        int const j = outward ? s : last - s;
        double const G = j != 0 ? c.p_eta[j - 1] * c.p_eta[j - 1] * c.iq[j - 1] - kT
                                : 2. * e_couple - c.Nf * kT;
        if (j != last) {
            double e, op;
            expx_operator(dt * c.p_eta[j + 1] * c.iq[j + 1], e, op);
            c.p_eta[j] = c.p_eta[j] * e + dt * G * op;
        } else {
            c.p_eta[j] += dt * G;
        }
    }
}

// thermostat.hpp:24-35: returns the momentum scale factor
__device__ double chain_U3(Chain &c, double dt) {
    double const s = exp_neg(dt * c.p_eta[0] * c.iq[0]);
    for (int j = 0; j < c.M; ++j)
        c.eta[j] += dt * c.p_eta[j] * c.iq[j];
    return s;
}

// thermostat.cpp:24-37
__device__ double chain_etot(const Chain &c) {
    double e = c.Nf * c.eta[0];
    for (int j = 1; j < c.M; ++j)
        e += c.eta[j];
    e *= K_BOLTZ * c.T;
    for (int j = 0; j < c.M; ++j)
        e += 0.5 * (c.p_eta[j] * c.p_eta[j]) * c.iq[j];
    return e;
}

__device__ __forceinline__ double box_volume(const double *L) { return L[0] * (L[1] * L[2]); }

// md_driver.cpp:512-531
__device__ double thermo_couple(const DevState &s) {
    double e = s.ekin_tot;
    if (s.coupling != MDL_KINETIC && s.Tci != 0.) {
        double const econf = (K_BOLTZ / s.Tci) * s.Nf / 2.;
        e = s.coupling == MDL_CONFIGURATIONAL ? econf : 0.5 * (s.ekin_tot + econf);
    }
    return e;
}

// A chain of compile-time length M held in registers (no LDS round trip per
// element).
template <int M> struct RegChain {
    double p[M], iq[M], eta[M];
    __device__ __forceinline__ void load(const Chain &c) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            p[j] = c.p_eta[j];
            iq[j] = c.iq[j];
            eta[j] = c.eta[j];
        }
    }
    __device__ __forceinline__ void store(Chain &c) const {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            c.p_eta[j] = p[j];
            c.eta[j] = eta[j];
        }
    }
};

// exp(-x) and op(x) for any x with no branch: y = -x = k ln2 + r as in
// exp_neg_general, e^r = 1 + r p(r). For k = 0 (|x| < ln2/2) op(x) is p(r)
// itself (r = -x exactly, p is the series of (e^r - 1)/r); otherwise op(x) =
// -expm1(-x)/x with expm1(y) = 2^k r p(r) + (2^k - 1), which does not cancel
// for |y| >= ln2/2, times the Newton reciprocal of x. Within a few ulp of
// exp(-x) and op(x) everywhere (x = 0 included: op = 1).
__device__ __forceinline__ void expx_any(double x, double &expx, double &op) {
    double const y = -x;
    double const k = rint(y * 1.4426950408889634);
    double r = fma(k, -0x1.62e42ff000000p-1, y);
    r = fma(k, 0x1.718432a1b0e26p-35, r);
    double const r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    double const q0 = fma(r, 0.5, 1.), q1 = fma(r, 1. / 24., 1. / 6.),
                 q2 = fma(r, 1. / 720., 1. / 120.), q3 = fma(r, 1. / 40320., 1. / 5040.),
                 q4 = fma(r, 1. / 3628800., 1. / 362880.),
                 q5 = fma(r, 1. / 479001600., 1. / 39916800.);
    double const s0 = fma(q1, r2, q0), s1 = fma(q3, r2, q2), s2 = fma(q5, r2, q4);
    double const t0 = fma(s1, r4, s0), t1 = fma(1. / 6227020800., r4, s2);
    double const p = fma(t1, r8, t0);
    int const ki = (int)k;
    expx = ldexp(fma(r, p, 1.), ki);
    double const em1 = ldexp(r * p, ki) + (ldexp(1., ki) - 1.);
    double const o = -em1 * rcp_nr(x);
    op = k == 0. ? p : o;
}

// The sweeps of one U_NHC propagation (7 Suzuki-Yoshida weights;
// md_driver.cpp:539-561) of a chain of compile-time length M in registers. A
// lone lane's worth of work (the wave runs it with identical values), so its
// cost is its instruction count and its dependent chain: the inward sweep's
// G_j read p_{j-1} before its update and are formed first. The outward sweep's
// factor of element j has the argument dt p_{j+1}/q_{j+1} with p_{j+1} as the
// inward sweep left it (U3 changes no p, and the outward sweep reads p_{j+1}
// before updating it), which is the argument of the inward sweep's factor of
// element j: the inward factors are kept and used again, bit for bit the
// values the outward sweep would form (thermostat.cpp:40-66 forms them twice).
// dt*(p^2/q - kT) is formed as
// p^2*(dt/q) - dt*kT (thermostat.cpp:40-66 otherwise, element by element).
// BARO: the coupling is p_eps^2/(2W) and U3 scales p_eps (held in kin); else
// the coupling is cw*K + c0 (md_driver.cpp:512-531) and U3 scales K by f^2 and
// the running velocity factor by f.
// FORM says which factors come from expx_any (valid for any x); the others
// come from the series form (op_series), valid for |x| <= EXPX_SMALL only:
// SW_SERIES none, SW_TOP the factor of the top element (argument p_{M-1}, the
// chain's end, whose momentum a hard-driven chain leaves largest), SW_WIDE all.
// Returns which arguments exceeded EXPX_SMALL: bit 0 the top element's, bit 1
// any other's; no branch depends on the values.
enum SweepForm { SW_SERIES = 0, SW_TOP = 1, SW_WIDE = 2 };
template <int M, bool BARO, int FORM>
__device__ __forceinline__ int nhc_sweeps(double (&p)[M], double (&eta)[M], const double (&iq)[M],
                                          double Nf, double kT, double dt, double &kin,
                                          double &scale, double cw, double c0, double inv_W) {
    bool over_top = false, over_rest = false;
    auto factor = [&over_top, &over_rest](double x, double &e, double &o, bool top) {
        bool const out = !(fabs(x) <= EXPX_SMALL);
        if (top)
            over_top |= out;
        else
            over_rest |= out;
        if (FORM == SW_WIDE || (FORM == SW_TOP && top)) {
            expx_any(x, e, o);
        } else {
            o = op_series(x);
            e = fma(-x, o, 1.);
        }
    };
#pragma unroll 1
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt, dt4 = 0.5 * dtw;
        double c[M];
#pragma unroll
        for (int j = 0; j < M; ++j)
            c[j] = dt4 * iq[j];
        double const dtkT = dt4 * kT;
        // U4, inward (its factors E, O kept for the outward sweep)
        double E[M], O[M];
        double tG[M];
        {
            double const e = BARO ? 0.5 * (kin * kin) * inv_W : cw * kin + c0;
            tG[0] = dt4 * (2. * e - Nf * kT);
        }
#pragma unroll
        for (int j = 1; j < M; ++j)
            tG[j] = fma(p[j - 1] * p[j - 1], c[j - 1], -dtkT);
        p[M - 1] += tG[M - 1];
#pragma unroll
        for (int j = M - 2; j >= 0; --j) {
            factor(c[j + 1] * p[j + 1], E[j], O[j], j == M - 2);
            p[j] = fma(p[j], E[j], tG[j] * O[j]);
        }
        // U3
        {
            double f, unused;
            factor((2. * c[0]) * p[0], f, unused, false);
#pragma unroll
            for (int j = 0; j < M; ++j)
                eta[j] = fma(2. * c[j], p[j], eta[j]);
            if (BARO) {
                kin *= f;
            } else {
                scale *= f;
                kin *= f * f;
            }
        }
        // U4, outward
        double const e = BARO ? 0.5 * (kin * kin) * inv_W : cw * kin + c0;
        double const tG0 = dt4 * (2. * e - Nf * kT);
        if (M > 1)
            p[0] = fma(p[0], E[0], tG0 * O[0]);
        else
            p[0] += tG0;
#pragma unroll
        for (int j = 1; j < M; ++j) {
            double const t = fma(p[j - 1] * p[j - 1], c[j - 1], -dtkT);
            if (j + 1 < M)
                p[j] = fma(p[j], E[j], t * O[j]);
            else
                p[j] += t;
        }
    }
    return (over_top ? 1 : 0) | (over_rest ? 2 : 0);
}

// One U_NHC propagation. `form` is the sweep form the chain's previous
// propagation needed (the series form where its arguments stayed within
// EXPX_SMALL). That form runs first; should an argument it takes in the series
// form leave the range after all, the chain is restored and the wide sweeps run
// instead. One uniform branch per propagation, none per factor. `form` is
// updated for the next propagation. Returns whether any factor came from
// expx_any.
template <int M, bool BARO>
__device__ __forceinline__ bool nhc_prop(double (&p)[M], double (&eta)[M], const double (&iq)[M],
                                         double Nf, double kT, double dt, double &kin,
                                         double &scale, double cw, double c0, double inv_W,
                                         int &form) {
    auto next = [](int over) { return (over & 2) ? SW_WIDE : (over & 1) ? SW_TOP : SW_SERIES; };
    if (form != SW_WIDE) {
        double p0[M], eta0[M];
#pragma unroll
        for (int j = 0; j < M; ++j) {
            p0[j] = p[j];
            eta0[j] = eta[j];
        }
        double const kin0 = kin, scale0 = scale;
        int const tried = form;
        int const over = __builtin_amdgcn_readfirstlane(
            tried == SW_SERIES
                ? nhc_sweeps<M, BARO, SW_SERIES>(p, eta, iq, Nf, kT, dt, kin, scale, cw, c0, inv_W)
                : nhc_sweeps<M, BARO, SW_TOP>(p, eta, iq, Nf, kT, dt, kin, scale, cw, c0, inv_W));
        // the series factors were all in range: the result stands
        if ((over & (tried == SW_SERIES ? 3 : 2)) == 0) {
            form = next(over);
            return tried != SW_SERIES;
        }
#pragma unroll
        for (int j = 0; j < M; ++j) {
            p[j] = p0[j];
            eta[j] = eta0[j];
        }
        kin = kin0;
        scale = scale0;
    }
    form = next(__builtin_amdgcn_readfirstlane(
        nhc_sweeps<M, BARO, SW_WIDE>(p, eta, iq, Nf, kT, dt, kin, scale, cw, c0, inv_W)));
    return true;
}

// U_NHC_thermo_propagator (md_driver.cpp:539-549) with U3 acting on a running
// scale factor: K' = s^2 K replaces the O(N) recomputation (md_driver.cpp:504).
template <int M> __device__ __forceinline__ double nhc_thermo_m(DevState &s, double dt) {
    RegChain<M> c;
    c.load(s.th);
    double const kT = K_BOLTZ * s.th.T;
    double cw = 1., c0 = 0.;   // thermo_couple as cw K + c0
    if (s.coupling != MDL_KINETIC && s.Tci != 0.) {
        double const econf = (K_BOLTZ / s.Tci) * s.Nf / 2.;
        cw = s.coupling == MDL_CONFIGURATIONAL ? 0. : 0.5;
        c0 = s.coupling == MDL_CONFIGURATIONAL ? econf : 0.5 * econf;
    }
    double kin = s.ekin_tot, scale = 1.;
    if (nhc_prop<M, false>(c.p, c.eta, c.iq, s.th.Nf, kT, dt, kin, scale, cw, c0, 0.,
                           s.nhc_form[0]))
        s.nhc_general++;
    s.ekin_tot = kin;
    c.store(s.th);
    return scale;
}

__device__ double nhc_thermo_mem(DevState &s, double dt) {
    double scale = 1.;
    s.nhc_general++;
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        chain_U4(s.th, 0.5 * dtw, thermo_couple(s), false);
        double const f = chain_U3(s.th, dtw);
        scale *= f;
        s.ekin_tot *= f * f;
        chain_U4(s.th, 0.5 * dtw, thermo_couple(s), true);
    }
    return scale;
}

// U_NHC_baro_propagator (md_driver.cpp:551-561)
template <int M> __device__ __forceinline__ void nhc_baro_m(DevState &s, double dt) {
    RegChain<M> c;
    c.load(s.bth);
    double const kT = K_BOLTZ * s.bth.T;
    double peps = s.peps, unused = 1.;
    if (nhc_prop<M, true>(c.p, c.eta, c.iq, s.bth.Nf, kT, dt, peps, unused, 0., 0.,
                          1. / s.bar_W, s.nhc_form[1]))
        s.nhc_general++;
    s.peps = peps;
    c.store(s.bth);
}

__device__ void nhc_baro_mem(DevState &s, double dt) {
    s.nhc_general++;
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        chain_U4(s.bth, 0.5 * dtw, 0.5 * (s.peps * s.peps) / s.bar_W, false);
        s.peps *= chain_U3(s.bth, dtw);
        chain_U4(s.bth, 0.5 * dtw, 0.5 * (s.peps * s.peps) / s.bar_W, true);
    }
}

#define MDL_CHAIN_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

__device__ __forceinline__ double nhc_thermo(DevState &s, double dt) {
    switch (s.th.M) {
#define MDL_CASE(m) case m: return nhc_thermo_m<m>(s, dt);
        MDL_CHAIN_CASES(MDL_CASE)
#undef MDL_CASE
    default: return nhc_thermo_mem(s, dt);
    }
}

__device__ __forceinline__ void nhc_baro(DevState &s, double dt) {
    switch (s.bth.M) {
#define MDL_CASE(m) case m: nhc_baro_m<m>(s, dt); return;
        MDL_CHAIN_CASES(MDL_CASE)
#undef MDL_CASE
    default: nhc_baro_mem(s, dt);
    }
}

// U2p (md_driver.cpp:454-459)
__device__ void U2p(DevState &s, double dt) {
    double const pv = s.bar_P * box_volume(s.L);
    s.peps += (2. * s.alpha * s.ekin_tot + 3. * s.virial - 3. * pv) * dt;
}

// The scalar kernels work on an LDS copy of the state: the chain updates are a
// long serial sequence of dependent reads and writes, which must not each pay
// a global-memory round trip.
static_assert(sizeof(DevState) % 8 == 0, "DevState is copied in 8-byte words");
// Every word of the copy is loaded before any is stored, so the copy costs one
// memory latency rather than one per loop trip. Blocks of BLOCK threads.
constexpr int STATE_WORDS = (int)(sizeof(DevState) / 8);
constexpr int STATE_PER_THREAD = (STATE_WORDS + BLOCK - 1) / BLOCK;
__device__ __forceinline__ void state_load(DevState &sh, const DevState *gs) {
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(gs);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(&sh);
    unsigned long long r[STATE_PER_THREAD];
#pragma unroll
    for (int u = 0; u < STATE_PER_THREAD; ++u) {
        int const k = threadIdx.x + u * BLOCK;
        r[u] = k < STATE_WORDS ? src[k] : 0ull;
    }
#pragma unroll
    for (int u = 0; u < STATE_PER_THREAD; ++u) {
        int const k = threadIdx.x + u * BLOCK;
        if (k < STATE_WORDS)
            dst[k] = r[u];
    }
    __syncthreads();
}
__device__ __forceinline__ void state_store(DevState *gs, const DevState &sh) {
    __syncthreads();
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(&sh);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(gs);
#pragma unroll
    for (int u = 0; u < STATE_PER_THREAD; ++u) {
        int const k = threadIdx.x + u * BLOCK;
        if (k < STATE_WORDS)
            dst[k] = src[k];
    }
}

// ---------------------------------------------------------------------------
// k_pre: first half of integrate() up to the drift (md_driver.cpp:567-581)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pre_step(DevState &s, int sample);

__global__ __launch_bounds__(BLOCK) void k_pre(DevState *__restrict__ gs, int sample) {
    if (gs->err)
        return;
    __shared__ DevState s;
    state_load(s, gs);
    // the whole first wave runs the scalar step with identical values (one
    // active lane instead measured slower)
    if (threadIdx.x < 64) {
        pre_step(s, sample);
    }
    state_store(gs, s);
}

__device__ __forceinline__ void pre_step(DevState &s, int sample) {
    s.sample = sample;
    bool const npt = s.ensemble == MDL_NPT;
    bool const thermo = s.ensemble == MDL_NVT || npt;
    double const h = 0.5 * s.dt;
    double scale = 1.;
    if (npt)
        nhc_baro(s, h);
    if (thermo)
        scale = nhc_thermo(s, h);
    if (npt)
        U2p(s, h);
    double e = 1., c = 1.;
    if (npt)
        expx_operator(s.alpha * h * s.peps / s.bar_W, e, c);
    s.kick_s = s.vscale * scale;
    s.vscale = 1.;
    s.kick_e = e;
    s.kick_c = c * h;
    double cf = 1.;
    if (npt) {
        double const x = s.dt * s.peps / s.bar_W;
        s.eps += x;
        double ex;
        expx_operator(x, ex, cf);
    }
    s.drift_cf = cf * s.dt;
    for (int a = 0; a < 3; ++a) {
        s.drift_L[a] = s.L[a];
        s.L_drift[a] = s.L[a];
    }
    if (npt) {
        double const expe = exp(s.eps);
        double const vold = box_volume(s.L);
        for (int a = 0; a < 3; ++a)
            s.L[a] = s.L0[a] * expe;
        if (fabs(box_volume(s.L) / vold - 1.) > 0.05) {
            s.err |= DE_VOLUME;
            s.err_val[0] = expe;
            s.err_val[1] = vold;
            s.err_val[2] = box_volume(s.L);
        }
    }
}

// First U2 (md_driver.cpp:433-452) and U1 drift (md_driver.cpp:462-499).
// Each block leaves the largest squared displacement of its atoms since the
// last list build in dmax_part[blockIdx.x].
__device__ __forceinline__ void kick_drift_body(AtomArrays a, int N, DevState *__restrict__ gs,
                                                double *__restrict__ dmax_part, Blk bk) {
    if (gs->err)
        return;
    double const s = gs->kick_s, e = gs->kick_e, ch = gs->kick_c, cf = gs->drift_cf;
    double const Ld[3] = {gs->drift_L[0], gs->drift_L[1], gs->drift_L[2]};
    double const Ln[3] = {gs->L[0], gs->L[1], gs->L[2]};
    double dmax = 0., dmid = 0.;
    for (int i = bk.id * BLOCK + threadIdx.x; i < N; i += bk.n * BLOCK) {
        double const m = gs->mass[a.type[i]];
        double vx = a.vx[i] * s, vy = a.vy[i] * s, vz = a.vz[i] * s;
        vx = vx * e + ch * a.fx[i] / m;
        vy = vy * e + ch * a.fy[i] / m;
        vz = vz * e + ch * a.fz[i] / m;
        a.vx[i] = vx;
        a.vy[i] = vy;
        a.vz[i] = vz;
        double4 c = a.pos[i];
        c.x += cf * vx / Ld[0];
        c.y += cf * vy / Ld[1];
        c.z += cf * vz / Ld[2];
        a.pos[i] = c;
        a.qpos[i] = quantize(c);
        double4 const r = a.cref[i];
        double const ux = (c.x - r.x) * Ln[0], uy = (c.y - r.y) * Ln[1], uz = (c.z - r.z) * Ln[2];
        dmax = fmax(dmax, ux * ux + uy * uy + uz * uz);
        double4 const q = a.mref[i];
        double const vx2 = (c.x - q.x) * Ln[0], vy2 = (c.y - q.y) * Ln[1], vz2 = (c.z - q.z) * Ln[2];
        dmid = fmax(dmid, vx2 * vx2 + vy2 * vy2 + vz2 * vz2);
    }
    for (int off = 32; off > 0; off >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
        dmid = fmax(dmid, __shfl_xor(dmid, off, 64));
    }
    __shared__ double sm[2][BLOCK / 64];
    if ((threadIdx.x & 63) == 0) {
        sm[0][threadIdx.x >> 6] = dmax;
        sm[1][threadIdx.x >> 6] = dmid;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = sm[0][0], mm = sm[1][0];
        for (int w = 1; w < BLOCK / 64; ++w) {
            m = fmax(m, sm[0][w]);
            mm = fmax(mm, sm[1][w]);
        }
        dmax_part[bk.id] = m;
        dmax_part[DMAX_MID + bk.id] = mm;
    }
}

__global__ __launch_bounds__(BLOCK) void k_kick_drift(AtomArrays a, int N,
                                                     DevState *__restrict__ gs,
                                                     double *__restrict__ dmax_part) {
    kick_drift_body(a, N, gs, dmax_part, hw_blk());
}

// ---------------------------------------------------------------------------
// Neighbour list (full list, both i->j and j->i, radius r_list = r_c r_skin).
//
// A build bins the atoms into cells, scans the counts, scatters, sorts each
// cell's members by atom index (so neither the list order nor the summation
// order of the passes depends on atomic arrival order), then builds the lists
// one cell per block: the members of the 27 surrounding cells are staged in
// LDS as Cartesian FP32 coordinates relative to the cell centre, with the
// periodic image fixed by the cell offset, and lane groups test them against
// each member of the cell. The FP32 test accepts up to r_list (1 + 1e-5),
// more than its rounding error, so the list holds every pair within r_list (a
// few beyond it change nothing: the passes select pairs by r <= r_c in FP64).
// Boxes with fewer than 3 cells per direction test every atom pair in FP64.
//
// Builds the host asks for (calculate_neighbors(), the run schedule, the
// constructor) run as a chain of kernels, k_rb_*. Every other step runs
// k_list: it decides whether a pair outside the list could now be inside r_c
// - for a list built with radius r_list at box L_b, such a pair is at least
// lambda r_list - 2 u_max apart now, lambda = min L/L_b, u_max the largest
// displacement since the build - and on most steps every block returns there.
// When a build is due it runs the same phases in that one kernel, separated
// by grid barriers (its grid is no larger than what the device holds at once).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wrap_scaled(double c) { return c - nint_away(c); }

__device__ __forceinline__ int cell_coord(double c, int n) {
    int k = (int)floor((wrap_scaled(c) + 0.5) * n);
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

constexpr int CELL_MAX = 1024;          // members of one cell the sort stages in LDS
constexpr int STAGE_CAP = 2048;         // candidates staged in LDS at once
constexpr int LIST_LANES = 8;           // lanes per atom of the list build
constexpr long long BARRIER_TICKS = 100000000ll;   // 1 s of the 100 MHz real-time counter

struct ListSmem {
    float4 stage[STAGE_CAP];            // also the sort buffer (CELL_MAX ints)
    int nb_start[27], nb_off[28];
    double nb_shift[27][3];
    int sc[BLOCK];
    unsigned long long entries;
};
static_assert(CELL_MAX * sizeof(int) <= sizeof(float4) * STAGE_CAP, "sort buffer in stage");

// Start of a build (one thread): list bookkeeping and the cell-size check.
__device__ __forceinline__ void rb_begin(DevState *gs, int use_cells) {
    gs->force_rebuild = 0;
    gs->do_rebuild = 1;
    gs->nbuilds++;
    gs->mid_build = 1;   // the mid list is filtered from the new list in pass A
    gs->nmid_builds++;
    for (int c = 0; c < 3; ++c) {
        gs->L_build[c] = gs->L[c];
        gs->L_mid[c] = gs->L[c];
    }
    if (use_cells)
        for (int c = 0; c < 3; ++c)
            if (gs->L[c] / gs->ncell[c] < gs->r_list) {
                atomicOr(&gs->err, (int)DE_CELL_GRID);
                gs->err_val[0] = gs->L[c];
            }
}

// Lanes of a wave holding the same key add their count with one atomic; each
// active lane gets base[key] before the add plus its rank among those lanes.
// Called by every lane of the wave.
__device__ __forceinline__ int wave_add(int *base, int key, bool active) {
    int const l64 = threadIdx.x & 63;
    unsigned long long const below = (l64 == 0) ? 0ull : (~0ull >> (64 - l64));
    unsigned long long pending = __ballot(active);
    int res = 0;
    while (pending) {
        int const leader = __ffsll((long long)pending) - 1;
        int const lkey = __shfl(key, leader, 64);
        unsigned long long const same = __ballot(active && key == lkey) & pending;
        int old = 0;
        if (l64 == leader)
            old = atomicAdd(&base[lkey], __popcll(same));
        old = __shfl(old, leader, 64);
        if ((same >> l64) & 1ull)
            res = old + __popcll(same & below);
        pending &= ~same;
    }
    return res;
}

// Atoms of block slots [first, N) step `step` (multiples of BLOCK).
__device__ __forceinline__ void rb_bin(const AtomArrays &a, ListArrays &l, int N,
                                       const DevState *gs, int first, int step) {
    int const nx = gs->ncell[0], ny = gs->ncell[1], nz = gs->ncell[2];
    for (int b = first; b < N; b += step) {
        int const i = b + threadIdx.x;
        bool const active = i < N;
        int cid = 0;
        if (active) {
            double4 const c = a.pos[i];
            cid = (cell_coord(c.z, nz) * ny + cell_coord(c.y, ny)) * nx + cell_coord(c.x, nx);
            l.cell_of[i] = cid;
        }
        wave_add(l.cell_count, cid, active);
    }
}

// One block: exclusive scan of the counts into cell_start; zero counts, fills.
__device__ __forceinline__ void rb_scan(ListArrays &l, int *sc) {
    int const per = (l.ncells + BLOCK - 1) / BLOCK;
    int const b0 = min(l.ncells, (int)threadIdx.x * per), e0 = min(l.ncells, b0 + per);
    int s = 0;
    for (int k = b0; k < e0; ++k)
        s += l.cell_count[k];
    sc[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {
        int const t = threadIdx.x >= off ? sc[threadIdx.x - off] : 0;
        __syncthreads();
        sc[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sc[threadIdx.x] - s;
    for (int k = b0; k < e0; ++k) {
        int const v = l.cell_count[k];
        l.cell_start[k] = run;
        run += v;
        l.cell_count[k] = 0;
        l.cell_fill[k] = 0;
    }
    if (threadIdx.x == BLOCK - 1)
        l.cell_start[l.ncells] = sc[BLOCK - 1];
}

__device__ __forceinline__ void rb_scatter(ListArrays &l, int N, int first, int step) {
    for (int b = first; b < N; b += step) {
        int const i = b + threadIdx.x;
        bool const active = i < N;
        int const cid = active ? l.cell_of[i] : 0;
        int const slot = wave_add(l.cell_fill, cid, active);
        if (active)
            l.cell_atoms[l.cell_start[cid] + slot] = i;
    }
}

// One block: members of cell c in ascending atom order (each placed at its rank).
__device__ __forceinline__ void rb_sort_cell(ListArrays &l, int c, int *v, DevState *gs) {
    int const b = l.cell_start[c], n = l.cell_start[c + 1] - b;
    if (n > CELL_MAX) {
        if (threadIdx.x == 0) {
            atomicOr(&gs->err, (int)DE_CELL_OVERFLOW);
            gs->err_val[1] = (double)n;
        }
        return;   // uniform in the block
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += BLOCK)
        v[k] = l.cell_atoms[b + k];
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += BLOCK) {
        int const x = v[k];
        int rank = 0;
        for (int q = 0; q < n; ++q)
            rank += v[q] < x;
        l.cell_atoms[b + rank] = x;
    }
}

// Per-atom end of a list build; entries are summed per block in *blk.
__device__ __forceinline__ void list_finish(ListArrays &l, AtomArrays &a, int i, int cnt,
                                            DevState *gs, unsigned long long *blk) {
    l.nnbr[i] = min(cnt, l.stride);
    a.cref[i] = a.pos[i];
    atomicAdd(blk, (unsigned long long)min(cnt, l.stride));
    if (cnt > l.stride) {
        atomicOr(&gs->err, (int)DE_NBR_OVERFLOW);
        atomicMax(reinterpret_cast<unsigned long long *>(&gs->nbr_needed),
                  (unsigned long long)cnt);
    }
}

// Ordered append of the lanes' hits: within a lane group the entries keep the
// candidate order (ballot prefix).
__device__ __forceinline__ void list_append(bool in, int j, unsigned long long gmask,
                                            unsigned long long below, int stride, int *row,
                                            int &cnt) {
    unsigned long long const bal = __ballot(in) & gmask;
    if (in) {
        int const slot = cnt + __popcll(bal & below);
        if (slot < stride)
            row[slot] = j;
    }
    cnt += __popcll(bal);
}

// One block: the lists of the members of cell c.
__device__ __forceinline__ void rb_build_cell(AtomArrays &a, ListArrays &l, int c,
                                              DevState *gs, const double *L, double r_list,
                                              ListSmem &sm) {
    int const nx = gs->ncell[0], ny = gs->ncell[1], nz = gs->ncell[2];
    int const lane = threadIdx.x % LIST_LANES;
    int const l64 = threadIdx.x & 63;
    unsigned long long const gmask = group_mask<LIST_LANES>();
    unsigned long long const below = (l64 == 0) ? 0ull : (~0ull >> (64 - l64));
    float const rl2f = (float)(r_list * r_list * (1. + 1e-5));
    int const cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    double const scx = (cx + 0.5) / nx - 0.5, scy = (cy + 0.5) / ny - 0.5,
                 scz = (cz + 0.5) / nz - 0.5;
    __syncthreads();   // the previous cell's staging is consumed
    if (threadIdx.x < 27) {
        int const q = threadIdx.x;
        int x = cx + q % 3 - 1, y = cy + (q / 3) % 3 - 1, z = cz + q / 9 - 1;
        double shx = 0., shy = 0., shz = 0.;
        if (x < 0) { x += nx; shx = -1.; } else if (x >= nx) { x -= nx; shx = 1.; }
        if (y < 0) { y += ny; shy = -1.; } else if (y >= ny) { y -= ny; shy = 1.; }
        if (z < 0) { z += nz; shz = -1.; } else if (z >= nz) { z -= nz; shz = 1.; }
        int const cell = (z * ny + y) * nx + x;
        sm.nb_start[q] = l.cell_start[cell];
        sm.nb_off[q + 1] = l.cell_start[cell + 1] - l.cell_start[cell];
        sm.nb_shift[q][0] = shx;
        sm.nb_shift[q][1] = shy;
        sm.nb_shift[q][2] = shz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sm.nb_off[0] = 0;
        for (int q = 0; q < 27; ++q)
            sm.nb_off[q + 1] += sm.nb_off[q];
    }
    __syncthreads();
    int const T = sm.nb_off[27];
    int const cb = l.cell_start[c], mc = l.cell_start[c + 1] - cb;
    int const nchunks = (T + STAGE_CAP - 1) / STAGE_CAP;
    for (int r0 = 0; r0 < mc; r0 += BLOCK / LIST_LANES) {
        int const k = r0 + threadIdx.x / LIST_LANES;
        bool const active = k < mc;
        int const i = active ? l.cell_atoms[cb + k] : -1;
        float xi = 0.f, yi = 0.f, zi = 0.f;
        if (active) {
            double4 const p = a.pos[i];
            xi = (float)((wrap_scaled(p.x) - scx) * L[0]);
            yi = (float)((wrap_scaled(p.y) - scy) * L[1]);
            zi = (float)((wrap_scaled(p.z) - scz) * L[2]);
        }
        int *row = l.nbr + (size_t)(active ? i : 0) * l.stride;
        int cnt = 0;
        for (int ch = 0; ch < nchunks; ++ch) {
            int const t0 = ch * STAGE_CAP, tn = min(T - t0, STAGE_CAP);
            if (nchunks > 1 || r0 == 0) {
                __syncthreads();
                for (int t = threadIdx.x; t < tn; t += BLOCK) {
                    int const tt = t0 + t;
                    int q = 0;
                    while (sm.nb_off[q + 1] <= tt)
                        ++q;
                    int const j = l.cell_atoms[sm.nb_start[q] + (tt - sm.nb_off[q])];
                    double4 const p = a.pos[j];
                    sm.stage[t] = make_float4(
                        (float)((wrap_scaled(p.x) + sm.nb_shift[q][0] - scx) * L[0]),
                        (float)((wrap_scaled(p.y) + sm.nb_shift[q][1] - scy) * L[1]),
                        (float)((wrap_scaled(p.z) + sm.nb_shift[q][2] - scz) * L[2]),
                        __int_as_float(j));
                }
                __syncthreads();
            }
            for (int u0 = 0; u0 < tn; u0 += LIST_LANES) {
                int const u = u0 + lane;
                bool in = false;
                int j = 0;
                if (active && u < tn) {
                    float4 const p = sm.stage[u];
                    j = __float_as_int(p.w);
                    float const dx = p.x - xi, dy = p.y - yi, dz = p.z - zi;
                    in = j != i && dx * dx + dy * dy + dz * dz <= rl2f;
                }
                list_append(in, j, gmask, below, l.stride, row, cnt);
            }
        }
        if (active && lane == 0)
            list_finish(l, a, i, cnt, gs, &sm.entries);
    }
}

// Lane groups over atoms [first group, N) step `groups`: every other atom,
// FP64 minimum image (boxes with fewer than 3 cells per direction).
__device__ __forceinline__ void rb_build_brute(AtomArrays &a, ListArrays &l, int N,
                                               DevState *gs, const double *L, double r_list,
                                               int group0, int groups,
                                               unsigned long long *blk) {
    int const lane = threadIdx.x % LIST_LANES;
    int const l64 = threadIdx.x & 63;
    unsigned long long const gmask = group_mask<LIST_LANES>();
    unsigned long long const below = (l64 == 0) ? 0ull : (~0ull >> (64 - l64));
    double const rl2 = r_list * r_list;
    for (int i = group0; i < N; i += groups) {
        double4 const ci = a.pos[i];
        int *row = l.nbr + (size_t)i * l.stride;
        int cnt = 0;
        for (int k0 = 0; k0 < N; k0 += LIST_LANES) {
            int const j = k0 + lane;
            bool in = false;
            if (j < N && j != i) {
                double dx, dy, dz;
                min_image(ci, a.pos[j], L, dx, dy, dz);
                in = dx * dx + dy * dy + dz * dz <= rl2;
            }
            list_append(in, j, gmask, below, l.stride, row, cnt);
        }
        if (lane == 0)
            list_finish(l, a, i, cnt, gs, blk);
    }
}

__device__ __forceinline__ void rb_flush_entries(DevState *gs, unsigned long long &blk) {
    __syncthreads();
    if (threadIdx.x == 0 && blk)
        atomicAdd(reinterpret_cast<unsigned long long *>(&gs->nlist_entries), blk);
}

// --- the chain of a requested build ----------------------------------------
// k_rb_begin (brute builds) or k_rb_bin (cell builds, block 0 begins the
// build) -> k_rb_scan -> k_rb_scatter -> k_rb_sort -> k_rb_build.
__global__ __launch_bounds__(64) void k_rb_begin(DevState *__restrict__ gs) {
    if (gs->err || threadIdx.x != 0)
        return;
    rb_begin(gs, 0);
    gs->nlist_entries = 0;
}

__global__ __launch_bounds__(BLOCK) void k_rb_bin(AtomArrays a, ListArrays l, int N,
                                                 DevState *__restrict__ gs) {
    if (gs->err)
        return;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        rb_begin(gs, 1);
        gs->nlist_entries = 0;
    }
    rb_bin(a, l, N, gs, blockIdx.x * BLOCK, gridDim.x * BLOCK);
}

__global__ __launch_bounds__(BLOCK) void k_rb_scan(ListArrays l, const DevState *__restrict__ gs) {
    if (gs->err)
        return;
    __shared__ int sc[BLOCK];
    rb_scan(l, sc);
}

__global__ __launch_bounds__(BLOCK) void k_rb_scatter(ListArrays l, int N,
                                                     const DevState *__restrict__ gs) {
    if (gs->err)
        return;
    rb_scatter(l, N, blockIdx.x * BLOCK, gridDim.x * BLOCK);
}

__global__ __launch_bounds__(BLOCK) void k_rb_sort(ListArrays l, DevState *__restrict__ gs) {
    if (gs->err)
        return;
    __shared__ int v[CELL_MAX];
    rb_sort_cell(l, blockIdx.x, v, gs);
}

__global__ __launch_bounds__(BLOCK) void k_rb_build(AtomArrays a, ListArrays l, int N,
                                                   DevState *__restrict__ gs, int use_cells) {
    if (gs->err)
        return;
    __shared__ ListSmem sm;
    if (threadIdx.x == 0)
        sm.entries = 0ull;
    double const L[3] = {gs->L[0], gs->L[1], gs->L[2]};
    double const r_list = gs->r_list;
    if (use_cells)
        rb_build_cell(a, l, blockIdx.x, gs, L, r_list, sm);
    else
        rb_build_brute(a, l, N, gs, L, r_list, (blockIdx.x * BLOCK + threadIdx.x) / LIST_LANES,
                       gridDim.x * BLOCK / LIST_LANES, &sm.entries);
    rb_flush_entries(gs, sm.entries);
}

// --- the per-step check, with the build in one kernel when one is due ------
// Generation-counting grid barrier over gridDim.x resident blocks. Waiting
// blocks poll with relaxed loads (an acquiring poll would invalidate the
// caches on every iteration) and fence once when released. The wait is
// bounded: past BARRIER_TICKS the block reports DE_BARRIER and goes on.
__device__ bool grid_barrier(unsigned *bar, DevState *gs, int nblk) {
    __syncthreads();
    __shared__ int ok;
    if (threadIdx.x == 0) {
        ok = 1;
        __threadfence();
        unsigned const gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned const arrived =
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == (unsigned)nblk - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long const t0 = wall_clock64();
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(8);
                if (wall_clock64() - t0 > BARRIER_TICKS) {
                    atomicOr(&gs->err, (int)DE_BARRIER);
                    ok = 0;
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
    return ok != 0;
}

__device__ __forceinline__ void list_body(AtomArrays a, ListArrays l, int N,
                                          DevState *__restrict__ gs, unsigned *__restrict__ bar,
                                          const double *__restrict__ dmax_part, int nparts,
                                          int use_cells, Blk bk) {
    if (gs->err)
        return;
    // the decision, from the same inputs in every block (u_max^2 from
    // k_kick_drift's per-block maxima, nparts <= 4 BLOCK)
    __shared__ double red[2][BLOCK / 64];
    double m = 0., mm = 0.;
    {
        double v[4], w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int const b = threadIdx.x + u * BLOCK;
            v[u] = b < nparts ? dmax_part[b] : 0.;
            w[u] = b < nparts ? dmax_part[DMAX_MID + b] : 0.;
        }
        m = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
        mm = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = fmax(m, __shfl_xor(m, off, 64));
        mm = fmax(mm, __shfl_xor(mm, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = m;
        red[1][threadIdx.x >> 6] = mm;
    }
    __syncthreads();
    m = red[0][0];
    mm = red[1][0];
    for (int w = 1; w < BLOCK / 64; ++w) {
        m = fmax(m, red[0][w]);
        mm = fmax(mm, red[1][w]);
    }
    double lam = 1e300, lam_mid = 1e300;
    for (int c = 0; c < 3; ++c) {
        lam = fmin(lam, gs->L[c] / gs->L_build[c]);
        lam_mid = fmin(lam_mid, gs->L[c] / gs->L_mid[c]);
    }
    bool const need = gs->force_rebuild || !(lam * gs->r_list - 2. * sqrt(m) > gs->rc);
    if (!need) {
        // the mid list: a pair outside it is at least lambda r_mid - 2 u apart,
        // u the largest displacement since its build; block 0 records the
        // decision (every block reads the same inputs, and only pass A uses it)
        if (bk.id == 0 && threadIdx.x == 0) {
            bool const mid = !(lam_mid * gs->r_mid - 2. * sqrt(mm) > gs->rc);
            gs->mid_build = mid ? 1 : 0;
            if (mid) {
                gs->nmid_builds++;
                for (int c = 0; c < 3; ++c)
                    gs->L_mid[c] = gs->L[c];
            }
        }
        return;
    }

    __shared__ ListSmem sm;
    if (threadIdx.x == 0) {
        sm.entries = 0ull;
        if (bk.id == 0)
            gs->nlist_entries = 0;   // not a decision input; summed only after barrier 1
    }
    bool ok = grid_barrier(bar, gs, bk.n);   // every block has read the decision inputs
    if (bk.id == 0 && threadIdx.x == 0)
        rb_begin(gs, use_cells);
    double const L[3] = {gs->L[0], gs->L[1], gs->L[2]};
    double const r_list = gs->r_list;
    int const nthreads = bk.n * BLOCK;
    if (!use_cells) {
        if (ok)
            rb_build_brute(a, l, N, gs, L, r_list, (bk.id * BLOCK + threadIdx.x) / LIST_LANES,
                           nthreads / LIST_LANES, &sm.entries);
    } else {
        rb_bin(a, l, N, gs, bk.id * BLOCK, nthreads);
        ok = grid_barrier(bar, gs, bk.n) && ok;
        if (bk.id == 0)
            rb_scan(l, sm.sc);
        ok = grid_barrier(bar, gs, bk.n) && ok;
        rb_scatter(l, N, bk.id * BLOCK, nthreads);
        ok = grid_barrier(bar, gs, bk.n) && ok;
        for (int c = bk.id; c < l.ncells; c += bk.n)
            rb_sort_cell(l, c, reinterpret_cast<int *>(sm.stage), gs);
        ok = grid_barrier(bar, gs, bk.n) && ok;
        if (ok)
            for (int c = bk.id; c < l.ncells; c += bk.n)
                rb_build_cell(a, l, c, gs, L, r_list, sm);
    }
    rb_flush_entries(gs, sm.entries);
}

__global__ __launch_bounds__(BLOCK) void k_list(AtomArrays a, ListArrays l, int N,
                                               DevState *__restrict__ gs,
                                               unsigned *__restrict__ bar,
                                               const double *__restrict__ dmax_part, int nparts,
                                               int use_cells) {
    list_body(a, l, N, gs, bar, dmax_part, nparts, use_cells, hw_blk());
}

// ---------------------------------------------------------------------------
// Pass A: densities (md_driver.cpp:195-277) over the full skin list. Per atom
// rho_i, and when sampled g_i (Voigt) and h_i (md_driver.cpp:245-264); then
// F, F', F'' (md_driver.cpp:267-276). The pass also compacts the pairs inside
// r_c into the short list that pass B walks (positions do not change between
// the passes, so pass B's r^2 test would select exactly these pairs). Within
// a lane group the compacted order is the candidate order (ballot prefix).
// ---------------------------------------------------------------------------

template <int KIND, int LPA, bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_density(AtomArrays a, ListArrays l, RTables rt,
                                                  FTables ft, int N, int npair, int ntypes,
                                                  double rc2, double rc_rho,
                                                  DevState *__restrict__ gs,
                                                  double *__restrict__ part) {
    if (gs->err)
        return;
    using P = Poly<KIND>;
    constexpr int APB = BLOCK / LPA;
    int const lane = threadIdx.x % LPA, slot = threadIdx.x / LPA;
    int const l64 = threadIdx.x & 63;
    unsigned long long const gmask = group_mask<LPA>();
    unsigned long long const below = (l64 == 0) ? 0ull : (~0ull >> (64 - l64));
    double const L[3] = {gs->L[0], gs->L[1], gs->L[2]};
    // scan the mid list, or the full list while the mid list is refiltered
    bool const mid_build = gs->mid_build != 0;
    double const rm2 = gs->r_mid * gs->r_mid;
    double acc[PA_N];
#pragma unroll
    for (int k = 0; k < PA_N; ++k)
        acc[k] = 0.;
    XcdRange const R((N + APB - 1) / APB);
    for (int grp = R.first; grp < R.end; grp += R.step) {
        int const i = grp * APB + slot;
        bool const active = i < N;
        double rho = 0., g0 = 0., g1 = 0., g2 = 0., g3 = 0., g4 = 0., g5 = 0.;
        double hx = 0., hy = 0., hz = 0.;
        double4 ci = make_double4(0., 0., 0., 0.);
        int cnt = 0;
        if (active) {
            ci = a.pos[i];
            cnt = mid_build ? l.nnbr[i] : l.mcnt[i];
        }
        size_t const row = (size_t)(active ? i : 0) * l.stride;
        const int *nb = (mid_build ? l.nbr : l.mnbr) + row;
        int *cb = l.cnbr + row;
        int *mb = l.mnbr + row;
        int nin = 0, nmid = 0;   // compacted entries so far (uniform in the group)
        for (int k0 = 0; k0 < cnt; k0 += LPA) {
            int const k = k0 + lane;
            bool in = false, inm = false;
            int j = 0;
            double dx = 0., dy = 0., dz = 0., r2 = 0.;
            double4 cj;
            if (k < cnt) {
                j = nb[k];
                cj = a.pos[j];
                min_image(ci, cj, L, dx, dy, dz);
                r2 = dx * dx + dy * dy + dz * dz;
                in = !(r2 > rc2);
                inm = !(r2 > rm2);
            }
            if (mid_build) {
                unsigned long long const balm = __ballot(inm) & gmask;
                if (inm)
                    mb[nmid + __popcll(balm & below)] = j;
                nmid += __popcll(balm);
            }
            unsigned long long const bal = __ballot(in) & gmask;
            if (in) {
                cb[nin + __popcll(bal & below)] = j;
                double const r = sqrt(r2);
                if (!(r > rc_rho)) {   // eam.hpp:30-44: rho and derivatives are 0 beyond rcut
                    int const tj = (int)cj.w;
                    double t;
                    int const it = tab_locate(r, rt.inv_h, rt.n, t);
                    P p;
                    p.load(rt.rec + ((size_t)it * rt.nf + npair + tj) * P::S);
                    rho += p.v(t);
                    if (SAMPLE) {
                        double const w = p.d1(t, rt.inv_h) / r;
                        g0 += w * dx * dx;
                        g1 += w * dy * dy;
                        g2 += w * dz * dz;
                        g3 += w * dy * dz;
                        g4 += w * dx * dz;
                        g5 += w * dx * dy;
                        hx += w * dx;
                        hy += w * dy;
                        hz += w * dz;
                    }
                }
            }
            nin += __popcll(bal);
        }
#pragma unroll
        for (int off = LPA / 2; off > 0; off >>= 1)
            rho += __shfl_xor(rho, off, LPA);
        // sampled steps: g (6) and h (3) of the atom, each summed into one lane
        // of its group (groups of 16 lanes or more), which stores its g
        // component; |h|^2 is summed from the h lanes' squares
        double h2 = 0.;
        if constexpr (SAMPLE && LPA >= 16) {
            double gs9[9] = {g0, g1, g2, g3, g4, g5, hx, hy, hz};
            int c;
            double const v = group_sum_scatter<LPA, 9>(gs9, c);
            h2 = c >= 6 ? v * v : 0.;
#pragma unroll
            for (int off = LPA / 2; off > 0; off >>= 1)
                h2 += __shfl_xor(h2, off, LPA);
            if (active && c >= 0 && c < 6)
                reinterpret_cast<double *>(a.geam + 4 * (size_t)i)[c] = v;
        } else if constexpr (SAMPLE) {
#pragma unroll
            for (int off = LPA / 2; off > 0; off >>= 1) {
                g0 += __shfl_xor(g0, off, LPA);
                g1 += __shfl_xor(g1, off, LPA);
                g2 += __shfl_xor(g2, off, LPA);
                g3 += __shfl_xor(g3, off, LPA);
                g4 += __shfl_xor(g4, off, LPA);
                g5 += __shfl_xor(g5, off, LPA);
                hx += __shfl_xor(hx, off, LPA);
                hy += __shfl_xor(hy, off, LPA);
                hz += __shfl_xor(hz, off, LPA);
            }
            h2 = hx * hx + hy * hy + hz * hz;
        }
        if (active && lane == 0) {
            l.ccnt[i] = nin;
            if (mid_build) {
                l.mcnt[i] = nmid;
                a.mref[i] = ci;
            }
            int const ti = (int)ci.w;
            if (rho > ft.rhomax[ti]) {
                atomicOr(&gs->err, (int)DE_RHOMAX);
                gs->err_val[0] = rho;
                gs->err_val[1] = ft.rhomax[ti];
            }
            double t;
            int const it = tab_locate(rho, ft.inv_h, ft.n, t);
            P p;
            p.load(ft.rec + ((size_t)it * ntypes + ti) * P::S);
            double const F = p.v(t), dF = p.d1(t, ft.inv_h);
            double const d2F = SAMPLE ? p.d2(t, ft.inv_h2) : 0.;
            a.Femb[i] = F;
            a.rho[i] = rho;
            a.fprime[i] = make_double2(dF, d2F);
            acc[PA_EF] += F;
            if (SAMPLE) {
                double2 *gp = a.geam + 4 * (size_t)i;
                if constexpr (LPA < 16) {
                    gp[0] = make_double2(g0, g1);
                    gp[1] = make_double2(g2, g3);
                    gp[2] = make_double2(g4, g5);
                }
                gp[3] = make_double2(h2, 0.);
            }
        }
    }
    block_sum_store<PA_N>(acc, part);
}

// ---------------------------------------------------------------------------
// Pass B: forces, energies, virial, Born stress; when sampled the Born
// elasticity and the Laplacian pair terms (md_driver.cpp:300-430), over the
// compacted in-cut-off list. Each pair is visited from both atoms (full list):
// forces and energies are gathered without atomics; the pair sums carry a
// factor 2, removed in post_step. The pair Born term X d_a d_b d_c d_d is fully
// symmetric in its four indices, so its 21 Voigt entries are 15 distinct
// monomials of d, accumulated here and expanded in post_step.
// ---------------------------------------------------------------------------
template <int KIND, int LPA, bool SAMPLE, bool EAM>
__global__ __launch_bounds__(BLOCK) void k_force(AtomArrays a, ListArrays l, RTables rt, int N,
                                                int npair, double rc2, double rc_pair,
                                                double rc_rho, double rmin,
                                                DevState *__restrict__ gs,
                                                double *__restrict__ part) {
    if (gs->err)
        return;
    // analytic pair forms (Morse, Lennard-Jones) evaluate in place of the
    // pair table; they run without EAM
    constexpr bool ANALYTIC = analytic_kind(KIND);
    static_assert(!(ANALYTIC && EAM), "analytic pair forms run without EAM tables");
    using P = Poly<ANALYTIC ? MDL_QUINTIC : KIND>;
    constexpr int APB = BLOCK / LPA;
    int const lane = threadIdx.x % LPA, slot = threadIdx.x / LPA;
    double const L[3] = {gs->L[0], gs->L[1], gs->L[2]};
    double acc[PB_N];
#pragma unroll
    for (int k = 0; k < PB_N; ++k)
        acc[k] = 0.;
    int nrmin = 0;
    XcdRange const R((N + APB - 1) / APB);
    for (int grp = R.first; grp < R.end; grp += R.step) {
        int const i = grp * APB + slot;
        bool const active = i < N;
        double fx = 0., fy = 0., fz = 0., e = 0.;
        if (active) {
            double4 const ci = a.pos[i];
            int const ti = (int)ci.w;
            double2 const Fi = EAM ? a.fprime[i] : make_double2(0., 0.);
            int const cnt = EAM ? l.ccnt[i] : l.nnbr[i];
            const int *nb = (EAM ? l.cnbr : l.nbr) + (size_t)i * l.stride;
            for (int k = lane; k < cnt; k += LPA) {
                int const j = nb[k];
                double4 const cj = a.pos[j];
                double dx, dy, dz;
                min_image(ci, cj, L, dx, dy, dz);
                double const r2 = dx * dx + dy * dy + dz * dz;
                if (!EAM && r2 > rc2)   // the compacted list is inside r_c already
                    continue;
                int const tj = (int)cj.w;
                int const tmax = max(ti, tj), tmin = min(ti, tj);
                int const pslot = tmax * (tmax + 1) / 2 + tmin;
                double const r = sqrt(r2);
                nrmin += r < rmin;
                double t = 0.;
                const double *base = rt.rec;
                if constexpr (!ANALYTIC) {
                    int const it = tab_locate(r, rt.inv_h, rt.n, t);
                    base += (size_t)it * rt.nf * P::S;
                }
                double V = 0., dV = 0., d2V = 0.;
                if (!(r > rc_pair)) {   // potential.hpp:36-37, 62-63, 97-103
                    if constexpr (ANALYTIC) {
                        pair_analytic<KIND>(rt.rec + pslot * PAIR_PARAMS, r, V, dV, d2V);
                    } else {
                        P pp;
                        pp.load(base + pslot * P::S);
                        V = pp.v(t);
                        dV = pp.d1(t, rt.inv_h);
                        if (SAMPLE)
                            d2V = pp.d2(t, rt.inv_h2);
                    }
                }
                double const ir = 1. / r;
                double dedr = dV, X = 0.;
                if (SAMPLE) {
                    X = d2V - dV * ir;
                    acc[PB_LAP] += 2. * (d2V + 2. * dV * ir);
                }
                if (EAM && !(r > rc_rho)) {
                    P pi, pj;
                    pi.load(base + (npair + ti) * P::S);
                    pj.load(base + (npair + tj) * P::S);
                    double const dri = pi.d1(t, rt.inv_h), drj = pj.d1(t, rt.inv_h);
                    double2 const Fj = a.fprime[j];
                    dedr += Fi.x * drj + Fj.x * dri;
                    if (SAMPLE) {
                        double const d2ri = pi.d2(t, rt.inv_h2), d2rj = pj.d2(t, rt.inv_h2);
                        acc[PB_LAP] += 2. * (Fj.x * (d2ri + 2. * dri * ir) +
                                             Fi.x * (d2rj + 2. * drj * ir));
                        acc[PB_LAP] += Fi.y * (drj * drj) + Fj.y * (dri * dri);
                        X += Fi.x * (d2rj - drj * ir) + Fj.x * (d2ri - dri * ir);
                    }
                }
                double const s = -dedr * ir;
                fx += s * dx;
                fy += s * dy;
                fz += s * dz;
                e += 0.5 * V;
                acc[PB_VIR] += r * dedr;
                double const xx = dx * dx, yy = dy * dy, zz = dz * dz;
                double const yz = dy * dz, xz = dx * dz, xy = dx * dy;
                double const w = dedr * ir;
                acc[PB_SB + 0] += w * xx;
                acc[PB_SB + 1] += w * yy;
                acc[PB_SB + 2] += w * zz;
                acc[PB_SB + 3] += w * yz;
                acc[PB_SB + 4] += w * xz;
                acc[PB_SB + 5] += w * xy;
                if (SAMPLE) {
                    X *= ir * ir;
                    double const Xxx = X * xx, Xyy = X * yy, Xzz = X * zz;
                    acc[PB_CB + M_X4] += Xxx * xx;
                    acc[PB_CB + M_Y4] += Xyy * yy;
                    acc[PB_CB + M_Z4] += Xzz * zz;
                    acc[PB_CB + M_X3Y] += Xxx * xy;
                    acc[PB_CB + M_X3Z] += Xxx * xz;
                    acc[PB_CB + M_Y3X] += Xyy * xy;
                    acc[PB_CB + M_Y3Z] += Xyy * yz;
                    acc[PB_CB + M_Z3X] += Xzz * xz;
                    acc[PB_CB + M_Z3Y] += Xzz * yz;
                    acc[PB_CB + M_X2Y2] += Xxx * yy;
                    acc[PB_CB + M_X2Z2] += Xxx * zz;
                    acc[PB_CB + M_Y2Z2] += Xyy * zz;
                    acc[PB_CB + M_X2YZ] += Xxx * yz;
                    acc[PB_CB + M_Y2XZ] += Xyy * xz;
                    acc[PB_CB + M_Z2XY] += Xzz * xy;
                }
            }
        }
#pragma unroll
        for (int off = LPA / 2; off > 0; off >>= 1) {
            fx += __shfl_xor(fx, off, LPA);
            fy += __shfl_xor(fy, off, LPA);
            fz += __shfl_xor(fz, off, LPA);
            e += __shfl_xor(e, off, LPA);
        }
        if (active && lane == 0) {
            a.fx[i] = fx;
            a.fy[i] = fy;
            a.fz[i] = fz;
            double const ei = (EAM ? a.Femb[i] : 0.) + e;
            a.epot[i] = ei;
            acc[PB_EPOT] += ei;
        }
    }
    acc[PB_WRMIN] = (double)nrmin;
    pass_sum_store<SAMPLE, PB_N, PB_CB, PB_LAP + 1>(acc, part, (int)blockIdx.x);
}

// Second U2 kick (md_driver.cpp:433-452) and the per-atom sums the
// thermodynamics need: K, sum m v_a v_b, momentum, sum |f|^2; when sampled
// also the embedding part of the Born tensor and Laplacian, F''_i g_i g_i
// and F''_i |h_i|^2 (md_driver.cpp:410-419), from pass A's per-atom g, |h|^2.
template <bool SAMPLE>
__device__ __forceinline__ void kick2_atom(const AtomArrays &a, int i,
                                           const DevState *__restrict__ gs, double e, double ch,
                                           bool init, int has_eam, double (&acc)[PK_N]) {
    double const m = gs->mass[a.type[i]];
    double const fx = a.fx[i], fy = a.fy[i], fz = a.fz[i];
    double vx = a.vx[i], vy = a.vy[i], vz = a.vz[i];
    if (!init) {
        vx = vx * e + ch * fx / m;
        vy = vy * e + ch * fy / m;
        vz = vz * e + ch * fz / m;
        a.vx[i] = vx;
        a.vy[i] = vy;
        a.vz[i] = vz;
    }
    acc[PK_K] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    acc[PK_MVV + 0] += m * vx * vx;
    acc[PK_MVV + 1] += m * vy * vy;
    acc[PK_MVV + 2] += m * vz * vz;
    acc[PK_MVV + 3] += m * vy * vz;
    acc[PK_MVV + 4] += m * vx * vz;
    acc[PK_MVV + 5] += m * vx * vy;
    acc[PK_MOM + 0] += m * vx;
    acc[PK_MOM + 1] += m * vy;
    acc[PK_MOM + 2] += m * vz;
    acc[PK_F2] += fx * fx + fy * fy + fz * fz;
    if (SAMPLE && has_eam) {
        double const d2F = a.fprime[i].y;
        const double2 *gp = a.geam + 4 * (size_t)i;
        double2 const ga = gp[0], gb = gp[1], gc = gp[2], gd = gp[3];
        double const g[6] = {ga.x, ga.y, gb.x, gb.y, gc.x, gc.y};
        int q = 0;
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            double const fg = d2F * g[v];
#pragma unroll
            for (int w = v; w < 6; ++w)
                acc[PK_CBF + q++] += fg * g[w];
        }
        acc[PK_LAPF] += d2F * gd.x;
    }
}

template <bool SAMPLE>
__device__ __forceinline__ void kick2_body(AtomArrays a, int N, const DevState *__restrict__ gs,
                                           int init, int has_eam, double *__restrict__ part,
                                           Blk bk) {
    if (gs->err)
        return;
    double const e = init ? 1. : gs->kick_e, ch = init ? 0. : gs->kick_c;
    double acc[PK_N];
#pragma unroll
    for (int k = 0; k < PK_N; ++k)
        acc[k] = 0.;
    for (int i = bk.id * BLOCK + threadIdx.x; i < N; i += bk.n * BLOCK)
        kick2_atom<SAMPLE>(a, i, gs, e, ch, init != 0, has_eam, acc);
    pass_sum_store<SAMPLE, PK_N, PK_CBF, PK_LAPF + 1>(acc, part, bk.id);
}

template <bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_kick2(AtomArrays a, int N,
                                                const DevState *__restrict__ gs, int init,
                                                int has_eam, double *__restrict__ part) {
    kick2_body<SAMPLE>(a, N, gs, init, has_eam, part, hw_blk());
}

// Deterministic sums of the per-block partial records, one block per
// component (thread t loads blocks t, t+BLOCK, ..., all summed in fixed
// trees); the last block to finish then runs the scalar part of the step:
// post_step (mode 0 = constructor, md_driver.cpp:114-118; mode 1 = end of
// integrate(), md_driver.cpp:583-600) and, when next_sample >= 0, the first
// half of the next step (pre_step), which needs nothing in between.
// Components are numbered A (PA_N), then B (PB_N), then K (PK_N).
constexpr int NTOTAL = PA_N + PB_N + PK_N;

__device__ __forceinline__ void post_step(DevState &s, const double *tA, const double *tB, const double *tK,
                          int mode, double *__restrict__ rec_out);

// Component k of the step's totals: the sum of its per-block partials in a
// fixed order (the value is returned to thread 0).
__device__ __forceinline__ double finish_sum(const DevState *__restrict__ gs,
                                             const double *__restrict__ pA, int nA,
                                             const double *__restrict__ pB, int nB,
                                             const double *__restrict__ pK, int nK, int k) {
    const double *p;
    int nb;
    if (k < PA_N) {
        p = pA;
        nb = gs->has_eam ? nA : 0;
    } else if (k < PA_N + PB_N) {
        p = pB;
        k -= PA_N;
        nb = nB;
    } else {
        p = pK;
        k -= PA_N + PB_N;
        nb = nK;
    }
    // nb <= PART_STRIDE: every partial of a thread is loaded before the adds
    constexpr int PER = PART_STRIDE / BLOCK;
    double v[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        int const b = threadIdx.x + u * BLOCK;
        v[u] = b < nb ? p[(size_t)k * PART_STRIDE + b] : 0.;
    }
#pragma unroll
    for (int w = 1; w < PER; w <<= 1)
#pragma unroll
        for (int u = 0; u + w < PER; u += 2 * w)
            v[u] += v[u + w];
    double sum = v[0];
    for (int off = 32; off > 0; off >>= 1)
        sum += __shfl_xor(sum, off, 64);
    __shared__ double sm[BLOCK / 64];
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = sum;
    __syncthreads();
    double t = 0.;
    if (threadIdx.x == 0)
        for (int w = 0; w < BLOCK / 64; ++w)
            t += sm[w];
    return t;
}

// The scalar end of the step from the totals (one block; the first wave
// works); ticket, unless null, is reset for the next step.
__device__ __forceinline__ void finish_tail(DevState *__restrict__ gs,
                                            const double *__restrict__ totals,
                                            unsigned *__restrict__ ticket, int mode,
                                            double *__restrict__ rec_out, int next_sample) {
    __shared__ double tot[NTOTAL];
    for (int c = threadIdx.x; c < NTOTAL; c += BLOCK)
        tot[c] = __hip_atomic_load(&totals[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __shared__ DevState s;
    state_load(s, gs);
    bool const copy = rec_out && s.sample != 0;   // this step's record, as post_step samples
    // the whole first wave runs the scalar step with identical values
    if (threadIdx.x < 64) {
        post_step(s, tot, tot + PA_N, tot + PA_N + PB_N, mode, rec_out);
        if (next_sample >= 0 && !s.err)
            pre_step(s, next_sample);
        if (ticket && threadIdx.x == 0)
            *ticket = 0u;
    }
    __syncthreads();
    // the sampled record out, one element per thread (pre_step leaves it as is)
    if (copy)
        for (int k = threadIdx.x; k < MDL_NRECORD; k += BLOCK)
            rec_out[k] = s.rec[k];
    state_store(gs, s);
}

// One block per component (blockIdx.x); the block that completes the set
// runs the tail.
__device__ __forceinline__ void finish_body(DevState *__restrict__ gs,
                                            const double *__restrict__ pA, int nA,
                                            const double *__restrict__ pB, int nB,
                                            const double *__restrict__ pK, int nK,
                                            double *__restrict__ totals,
                                            unsigned *__restrict__ ticket, int mode,
                                            double *__restrict__ rec_out, int next_sample) {
    if (gs->err)
        return;
    double const t = finish_sum(gs, pA, nA, pB, nB, pK, nK, (int)blockIdx.x);
    __shared__ int last;
    if (threadIdx.x == 0) {
        totals[blockIdx.x] = t;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last)
        return;
    __threadfence();
    finish_tail(gs, totals, ticket, mode, rec_out, next_sample);
}

__global__ __launch_bounds__(BLOCK) void k_finish(DevState *__restrict__ gs,
                                                 const double *__restrict__ pA, int nA,
                                                 const double *__restrict__ pB, int nB,
                                                 const double *__restrict__ pK, int nK,
                                                 double *__restrict__ totals,
                                                 unsigned *__restrict__ ticket, int mode,
                                                 double *__restrict__ rec_out, int next_sample) {
    finish_body(gs, pA, nA, pB, nB, pK, nK, totals, ticket, mode, rec_out, next_sample);
}

__device__ __forceinline__ void post_step(DevState &s, const double *tA, const double *tB, const double *tK,
                          int mode, double *__restrict__ rec_out) {
    bool const sample = s.sample != 0;
    bool const npt = s.ensemble == MDL_NPT;
    bool const thermo = s.ensemble == MDL_NVT || npt;
    double *rec = s.rec;
    int const N = s.N;

    // pair sums were taken over the full list: halve them (md_driver.cpp:389-407).
    // unit_E / unit_L convert the mixed path's eV and eV/A^2 sums to CGS (1 in FP64).
    double const uE = s.unit_E, uL = s.unit_L;
    double sB[6], CB[21];
    for (int v = 0; v < 6; ++v)
        sB[v] = 0.5 * tB[PB_SB + v] * uE;
    // mixed mode: the virial is the Born-stress trace, so P and tr(sigma)/3 agree
    // to FP64 round-off as they do in the reference (md_driver.cpp:693-698)
    s.virial = s.mixed ? -(sB[0] + sB[1] + sB[2]) / 3. : -(0.5 * tB[PB_VIR]) / 3.;
    if (sample) {
        // Voigt (v <= w) -> monomial of the pair term, then the embedding term
        const int mono[21] = {M_X4, M_X2Y2, M_X2Z2, M_X2YZ, M_X3Z, M_X3Y,
                              M_Y4, M_Y2Z2, M_Y3Z, M_Y2XZ, M_Y3X,
                              M_Z4, M_Z3Y, M_Z3X, M_Z2XY,
                              M_Y2Z2, M_Z2XY, M_Y2XZ,
                              M_X2Z2, M_X2YZ,
                              M_X2Y2};
        for (int q = 0; q < 21; ++q)
            CB[q] = (0.5 * tB[PB_CB + mono[q]] + tK[PK_CBF + q]) * uE;
        s.lap = (0.5 * tB[PB_LAP] + tK[PK_LAPF]) * uL;
    }
    s.warn_rmin += (long long)(0.5 * tB[PB_WRMIN] + 0.25);
    {   // virial self-check (md_driver.cpp:422-429), a warning only
        double const vtr = -(sB[0] + sB[1] + sB[2]);
        if (fabs(s.virial / (vtr / 3.) - 1.) > 1e-7)
            s.warn_virial++;
    }
    double const K = tK[PK_K];
    s.ekin_tot = K;
    // calculate_configurational_temperature (md_driver.cpp:630-646)
    if (sample) {
        s.f2 = tK[PK_F2];
        s.Tci = (s.f2 != 0. && s.lap != 0.) ? K_BOLTZ * s.lap / s.f2 : 0.;
    }
    double scale = 1.;
    if (mode == 1) {
        double const h = 0.5 * s.dt;
        if (npt)
            U2p(s, h);
        if (thermo)
            scale = nhc_thermo(s, h);
        if (npt)
            nhc_baro(s, h);
        s.vscale = scale;
    }
    // calculate_thermodynamics (md_driver.cpp:648-730) on the scaled velocities
    // the per-atom and per-volume records as products with one reciprocal each
    // (within an ulp of the quotients)
    double const iN = 1. / N;
    double const ekin_atom = s.ekin_tot * iN;
    double mom[3];
    for (int c = 0; c < 3; ++c)
        mom[c] = scale * tK[PK_MOM + c];
    double const p2 = mom[0] * mom[0] + mom[1] * mom[1] + mom[2] * mom[2];
    double const sc2 = 2. * s.mass[0] * s.ekin_tot;
    if (p2 / sc2 > 1e-14) {
        s.err |= DE_MOMENTUM;
        s.err_val[0] = mom[0];
        s.err_val[1] = mom[1];
        s.err_val[2] = mom[2];
        s.err_val[3] = sqrt(sc2);
    }
    double const epot_atom = tB[PB_EPOT] * uE * iN;
    double const epot_eam_atom = s.has_eam ? tA[PA_EF] * uE * iN : 0.;
    double const etot_atom = ekin_atom + epot_atom;
    double const old = s.etot_cons;
    double ec = etot_atom * N;
    double const V = box_volume(s.L), iV = 1. / V;
    if (thermo)
        ec += chain_etot(s.th);
    if (npt)
        ec += V * s.bar_P + (0.5 * (s.peps * s.peps) / s.bar_W + chain_etot(s.bth));
    s.etot_cons = ec;
    s.etot_cons_err = fabs(ec / s.etot_cons_initial - 1.);
    double const err_cur = fabs(ec / old - 1.);
    double const Tk = 2. * s.ekin_tot / (K_BOLTZ * s.Nf);
    double const P = (s.ekin_tot * 2. / 3. + s.virial) * iV;
    double st[6];
    double const s2 = scale * scale;
    for (int v = 0; v < 6; ++v)
        st[v] = (-sB[v] + s2 * tK[PK_MVV + v]) * iV;
    double const ptr = st[0] + st[1] + st[2];
    double const perr = fabs(P / (ptr / 3.) - 1.);
    if (perr > 1e-5) {
        s.err |= DE_PRESSURE;
        s.err_val[0] = perr;
        s.err_val[1] = P;
        s.err_val[2] = ptr / 3.;
    }
    if ((s.etot_cons_err > s.fatal || err_cur > s.fatal) && s.cycle > 0) {
        s.err |= DE_CONSERVATION;
        s.err_val[0] = s.etot_cons_err;
        s.err_val[1] = err_cur;
    }
    if (mode == 0)
        s.etot_cons_initial = ec;
    else {
        s.time += s.dt;
        s.cycle += 1;
    }
    rec[MDL_S_EPOT_EAM_ATOM] = epot_eam_atom;
    rec[MDL_S_EPOT_ATOM] = epot_atom;
    rec[MDL_S_EKIN_ATOM] = ekin_atom;
    rec[MDL_S_ETOT_ATOM] = etot_atom;
    rec[MDL_S_EKIN_TOT] = s.ekin_tot;
    rec[MDL_S_ETOT_CONS_INITIAL] = s.etot_cons_initial;
    rec[MDL_S_ETOT_CONS] = s.etot_cons;
    rec[MDL_S_ETOT_CONS_ERR] = s.etot_cons_err;
    rec[MDL_S_VIRIAL] = s.virial;
    rec[MDL_S_LAPLACIAN] = s.lap;
    rec[MDL_S_FORCE2] = s.f2;
    rec[MDL_S_TKIN] = Tk;
    rec[MDL_S_TCONF_INV] = s.Tci;
    rec[MDL_S_PRESSURE] = P;
    rec[MDL_S_TIME] = s.time;
    rec[MDL_S_CYCLE] = (double)s.cycle;
    for (int c = 0; c < 3; ++c) {
        rec[MDL_S_BOX0 + c] = s.L[c];
        rec[MDL_S_MOM0 + c] = mom[c];
        rec[MDL_S_BOX_INITIAL0 + c] = s.L0[c];
    }
    rec[MDL_S_BAR_EPS] = s.eps;
    rec[MDL_S_BAR_PEPS] = s.peps;
    for (int v = 0; v < 6; ++v) {
        rec[MDL_S_BORN_STRESS + v] = sB[v] * iV;
        rec[MDL_S_STRESS + v] = st[v];
    }
    if (sample) {
        int q = 0;
        for (int v = 0; v < 6; ++v)
            for (int w = 0; w < 6; ++w)
                rec[MDL_S_BORN_ELAST + 6 * v + w] = w >= v ? CB[q++] * iV : 0.;
    }
    rec[MDL_S_NLIST_BUILDS] = (double)s.nbuilds;
    rec[MDL_S_WARN_VIRIAL] = (double)s.warn_virial;
    rec[MDL_S_WARN_RMIN] = (double)s.warn_rmin;
    rec[MDL_S_NLIST_ENTRIES] = (double)s.nlist_entries;
    rec[MDL_S_NHC_GENERAL] = (double)s.nhc_general;
}

// The fixed-point positions from the FP64 ones, after the host wrote those.
__global__ __launch_bounds__(BLOCK) void k_quantize(AtomArrays a, int N) {
    int const i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < N)
        a.qpos[i] = quantize(a.pos[i]);
}

// Host-visible per-atom arrays (md_driver_module.cpp:103-109)
__global__ void k_materialize(AtomArrays a, int N, const DevState *__restrict__ gs, int what,
                              double *__restrict__ out) {
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    double const vs = gs->vscale;
    switch (what) {
    case MDL_A_POSITIONS: {
        double4 const c = a.pos[i];
        out[3 * i + 0] = (c.x - nint_away(c.x)) * gs->L_drift[0];
        out[3 * i + 1] = (c.y - nint_away(c.y)) * gs->L_drift[1];
        out[3 * i + 2] = (c.z - nint_away(c.z)) * gs->L_drift[2];
        break;
    }
    case MDL_A_COORDINATES: {
        double4 const c = a.pos[i];
        out[3 * i + 0] = c.x;
        out[3 * i + 1] = c.y;
        out[3 * i + 2] = c.z;
        break;
    }
    case MDL_A_VELOCITIES:
        out[3 * i + 0] = a.vx[i] * vs;
        out[3 * i + 1] = a.vy[i] * vs;
        out[3 * i + 2] = a.vz[i] * vs;
        break;
    case MDL_A_FORCES:
        out[3 * i + 0] = a.fx[i];
        out[3 * i + 1] = a.fy[i];
        out[3 * i + 2] = a.fz[i];
        break;
    case MDL_A_RHO:
        out[i] = a.rho[i];
        break;
    case MDL_A_EPOT:
        out[i] = a.epot[i];
        break;
    case MDL_A_EKIN: {
        double const vx = a.vx[i] * vs, vy = a.vy[i] * vs, vz = a.vz[i] * vs;
        out[i] = 0.5 * gs->mass[a.type[i]] * (vx * vx + vy * vy + vz * vz);
        break;
    }
    default:
        break;
    }
}

}  // namespace mdl
