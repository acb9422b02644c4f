// Device kernels of one MD step. Included by mdl_driver.hip only.
//
// Step = k_pre (scalars: NHC, barostat, kick/drift coefficients)
//      -> k_kick_drift (first half kick + drift, displacement since the list build)
//      -> k_decide -> [list build kernels, each a no-op unless a rebuild is due]
//      -> k_density (pass A) -> k_force (pass B) -> k_kick2 (second half kick)
//      -> k_post (reductions, NHC, thermodynamics, checks)
// which is MDDriver::integrate (md_driver.cpp:563-601) with the velocity
// scalings of the Nose-Hoover chain folded into one factor per half step.
#pragma once

#include "mdl_internal.hpp"

namespace mdl {

struct AtomArrays {
    double4 *pos;      // unwrapped scaled coordinates c (md_driver.hpp:97), type in .w
    double *vx, *vy, *vz;
    double *fx, *fy, *fz;
    int *type;
    double4 *cref;     // coordinates at the last list build
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

template <int NV>
__device__ __forceinline__ void block_sum_store(double (&acc)[NV], double *out) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    __shared__ double sm[BLOCK / 64][NV];
    int const wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    if (l64 == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            sm[wave][k] = acc[k];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.;
#pragma unroll
        for (int w = 0; w < BLOCK / 64; ++w)
            s += sm[w][threadIdx.x];
        out[(size_t)threadIdx.x * PART_STRIDE + blockIdx.x] = s;
    }
}

// XCD-aware slice of [0, nitems): blocks are dispatched round-robin over the
// 8 XCDs, so block b runs on XCD b % 8; give each XCD one contiguous chunk so
// its L2 holds one spatial region of the neighbour gathers.
struct XcdRange {
    int first, end, step;
    __device__ __forceinline__ XcdRange(int nitems) {
        int const nper = gridDim.x >> 3;
        int const xcd = blockIdx.x & 7, b = blockIdx.x >> 3;
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

// Minimum-image pair vector d = (s - nint(s)) L of s = c_i - c_j
// (md_driver.cpp:175-185, applied to the difference as at :222 and :333)
__device__ __forceinline__ void min_image(const double4 &ci, const double4 &cj, const double *L,
                                          double &dx, double &dy, double &dz) {
    double const sx = ci.x - cj.x, sy = ci.y - cj.y, sz = ci.z - cj.z;
    dx = (sx - nint_away(sx)) * L[0];
    dy = (sy - nint_away(sy)) * L[1];
    dz = (sz - nint_away(sz)) * L[2];
}

__device__ __forceinline__ void expx_operator(double x, double &expx, double &op) {
    // numerics.hpp:11-18
    const double c1 = -1.0 / 2.0, c2 = 1.0 / 6.0, c3 = -1.0 / 24.0, c4 = 1. / 120., c5 = -1. / 720.;
    expx = exp(-x);
    op = fabs(x) > 1e-3 ? (1.0 - expx) / x : 1.0 + x * (c1 + x * (c2 + x * (c3 + x * (c4 + x * c5))));
}

__constant__ const double WSY7[7] = {
    0.784513610477560, 0.235573213359357, -1.17767998417887,
    1. - 2. * (0.784513610477560 + 0.235573213359357 - 1.17767998417887),
    -1.17767998417887, 0.235573213359357, 0.784513610477560};

// thermostat.cpp:40-66
__device__ void chain_U4(Chain &c, double dt, double e_couple, bool outward) {
    double const kT = K_BOLTZ * c.T;
    int const last = c.M - 1;
    for (int s = 0; s < c.M; ++s) {
        int const j = outward ? s : last - s;
        double const G = j != 0 ? c.p_eta[j - 1] * c.p_eta[j - 1] / c.q[j - 1] - kT
                                : 2. * e_couple - c.Nf * kT;
        if (j != last) {
            double e, op;
            expx_operator(dt * c.p_eta[j + 1] / c.q[j + 1], e, op);
            c.p_eta[j] = c.p_eta[j] * e + dt * G * op;
        } else {
            c.p_eta[j] += dt * G;
        }
    }
}

// thermostat.hpp:24-35: returns the momentum scale factor
__device__ double chain_U3(Chain &c, double dt) {
    double const s = exp(-dt * c.p_eta[0] / c.q[0]);
    for (int j = 0; j < c.M; ++j)
        c.eta[j] += dt * c.p_eta[j] / c.q[j];
    return s;
}

// thermostat.cpp:24-37
__device__ double chain_etot(const Chain &c) {
    double e = c.Nf * c.eta[0];
    for (int j = 1; j < c.M; ++j)
        e += c.eta[j];
    e *= K_BOLTZ * c.T;
    for (int j = 0; j < c.M; ++j)
        e += 0.5 * (c.p_eta[j] * c.p_eta[j]) / c.q[j];
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

// U4 and U3 of a chain of compile-time length M held in registers (the
// chain sweep is a long chain of dependent FP64 operations; keeping p_eta,
// eta and q out of memory shortens it). Same arithmetic as chain_U4/chain_U3.
template <int M>
__device__ __forceinline__ void reg_U4(double (&p)[M], const double (&q)[M], double Nf,
                                       double kT, double dt, double e_couple, bool outward) {
#pragma unroll
    for (int s = 0; s < M; ++s) {
        int const j = outward ? s : M - 1 - s;
        double const G = j != 0 ? p[j - 1] * p[j - 1] / q[j - 1] - kT : 2. * e_couple - Nf * kT;
        if (j != M - 1) {
            double e, op;
            expx_operator(dt * p[j + 1] / q[j + 1], e, op);
            p[j] = p[j] * e + dt * G * op;
        } else {
            p[j] += dt * G;
        }
    }
}

template <int M>
__device__ __forceinline__ double reg_U3(const double (&p)[M], const double (&q)[M],
                                         double (&eta)[M], double dt) {
    double const f = exp(-dt * p[0] / q[0]);
#pragma unroll
    for (int j = 0; j < M; ++j)
        eta[j] += dt * p[j] / q[j];
    return f;
}

template <int M> struct RegChain {
    double p[M], q[M], eta[M];
    __device__ __forceinline__ void load(const Chain &c) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            p[j] = c.p_eta[j];
            q[j] = c.q[j];
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

// U_NHC_thermo_propagator (md_driver.cpp:539-549) with U3 acting on a running
// scale factor: K' = s^2 K replaces the O(N) recomputation (md_driver.cpp:504).
template <int M> __device__ double nhc_thermo_m(DevState &s, double dt) {
    RegChain<M> c;
    c.load(s.th);
    double const kT = K_BOLTZ * s.th.T;
    double scale = 1.;
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        reg_U4<M>(c.p, c.q, s.th.Nf, kT, 0.5 * dtw, thermo_couple(s), false);
        double const f = reg_U3<M>(c.p, c.q, c.eta, dtw);
        scale *= f;
        s.ekin_tot *= f * f;
        reg_U4<M>(c.p, c.q, s.th.Nf, kT, 0.5 * dtw, thermo_couple(s), true);
    }
    c.store(s.th);
    return scale;
}

__device__ double nhc_thermo_mem(DevState &s, double dt) {
    double scale = 1.;
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
template <int M> __device__ void nhc_baro_m(DevState &s, double dt) {
    RegChain<M> c;
    c.load(s.bth);
    double const kT = K_BOLTZ * s.bth.T;
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        reg_U4<M>(c.p, c.q, s.bth.Nf, kT, 0.5 * dtw, 0.5 * (s.peps * s.peps) / s.bar_W, false);
        s.peps *= reg_U3<M>(c.p, c.q, c.eta, dtw);
        reg_U4<M>(c.p, c.q, s.bth.Nf, kT, 0.5 * dtw, 0.5 * (s.peps * s.peps) / s.bar_W, true);
    }
    c.store(s.bth);
}

__device__ void nhc_baro_mem(DevState &s, double dt) {
    for (int k = 0; k < 7; ++k) {
        double const dtw = WSY7[k] * dt;
        chain_U4(s.bth, 0.5 * dtw, 0.5 * (s.peps * s.peps) / s.bar_W, false);
        s.peps *= chain_U3(s.bth, dtw);
        chain_U4(s.bth, 0.5 * dtw, 0.5 * (s.peps * s.peps) / s.bar_W, true);
    }
}

#define MDL_CHAIN_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) \
    X(13) X(14) X(15) X(16)

__device__ double nhc_thermo(DevState &s, double dt) {
    switch (s.th.M) {
#define MDL_CASE(m) case m: return nhc_thermo_m<m>(s, dt);
        MDL_CHAIN_CASES(MDL_CASE)
#undef MDL_CASE
    default: return nhc_thermo_mem(s, dt);
    }
}

__device__ void nhc_baro(DevState &s, double dt) {
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
__device__ __forceinline__ void state_load(DevState &sh, const DevState *gs) {
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(gs);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(&sh);
    for (int k = threadIdx.x; k < (int)(sizeof(DevState) / 8); k += blockDim.x)
        dst[k] = src[k];
    __syncthreads();
}
__device__ __forceinline__ void state_store(DevState *gs, const DevState &sh) {
    __syncthreads();
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(&sh);
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(gs);
    for (int k = threadIdx.x; k < (int)(sizeof(DevState) / 8); k += blockDim.x)
        dst[k] = src[k];
}

// ---------------------------------------------------------------------------
// k_pre: first half of integrate() up to the drift (md_driver.cpp:567-581)
// ---------------------------------------------------------------------------
__device__ void pre_step(DevState &s, int sample);

__global__ __launch_bounds__(64) void k_pre(DevState *__restrict__ gs, int sample, int force) {
    if (gs->err)
        return;
    __shared__ DevState s;
    state_load(s, gs);
    if (threadIdx.x == 0) {
        pre_step(s, sample);
        if (force)
            s.force_rebuild = 1;
    }
    state_store(gs, s);
}

__device__ void pre_step(DevState &s, int sample) {
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
__global__ __launch_bounds__(BLOCK) void k_kick_drift(AtomArrays a, int N,
                                                     DevState *__restrict__ gs,
                                                     double *__restrict__ dmax_part,
                                                     unsigned *__restrict__ ticket) {
    if (gs->err)
        return;
    double const s = gs->kick_s, e = gs->kick_e, ch = gs->kick_c, cf = gs->drift_cf;
    double const Ld[3] = {gs->drift_L[0], gs->drift_L[1], gs->drift_L[2]};
    double const Ln[3] = {gs->L[0], gs->L[1], gs->L[2]};
    double dmax = 0.;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < N; i += gridDim.x * BLOCK) {
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
        double4 const r = a.cref[i];
        double const ux = (c.x - r.x) * Ln[0], uy = (c.y - r.y) * Ln[1], uz = (c.z - r.z) * Ln[2];
        dmax = fmax(dmax, ux * ux + uy * uy + uz * uz);
    }
    for (int off = 32; off > 0; off >>= 1)
        dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
    __shared__ double sm[BLOCK / 64];
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = dmax;
    __syncthreads();
    __shared__ int last;
    if (threadIdx.x == 0) {
        double m = sm[0];
        for (int w = 1; w < BLOCK / 64; ++w)
            m = fmax(m, sm[w]);
        dmax_part[blockIdx.x] = m;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last)
        return;
    // the last block to finish reduces the per-block maxima
    __threadfence();
    if (threadIdx.x < 64) {
        double m = 0.;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += 64)
            m = fmax(m, __hip_atomic_load(&dmax_part[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        for (int off = 32; off > 0; off >>= 1)
            m = fmax(m, __shfl_xor(m, off, 64));
        if (threadIdx.x == 0) {
            gs->umax2 = m;
            *ticket = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// Neighbour list: one cooperative kernel per step.
//
// Phase 0 (every block, no communication): decide whether the list must be
// rebuilt - forced by calculate_neighbors(), or a pair outside the list could
// now be inside r_c. For a list built with radius r_list at box L_b, a pair
// not in it is at least lambda r_list - 2 u_max apart now, lambda = min L/L_b,
// u_max the largest displacement since the build. Every block reads the same
// inputs, so all take the same branch; on most steps they all return here and
// the kernel costs one launch.
// On a rebuild, phases separated by grid barriers: bin atoms into cells,
// scan the counts (block 0), scatter, sort each cell's members by index (so
// the list order, and the FP summation order of the passes, do not depend on
// atomic arrival order), and build each atom's full list with lane groups.
// The grid is sized so that all blocks are resident (cooperative launch); the
// barrier spin is bounded and reports DE_BARRIER instead of waiting forever.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(double c, int n) {
    double const w = c - nint_away(c);   // [-0.5, 0.5]
    int k = (int)floor((w + 0.5) * n);
    return k < 0 ? 0 : (k >= n ? n - 1 : k);
}

constexpr int CELL_MAX = 1024;          // members of one cell the sort stages in LDS
constexpr int LIST_LANES = 8;           // lanes per atom of the list build
constexpr long long BARRIER_SPINS = 1ll << 24;

// Sense-counting grid barrier over gridDim.x resident blocks.
__device__ bool grid_barrier(unsigned *bar, DevState *gs) {
    __threadfence();
    __syncthreads();
    __shared__ int ok;
    if (threadIdx.x == 0) {
        unsigned const gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned const arrived =
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        ok = 1;
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long long spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > BARRIER_SPINS) {
                    atomicOr(&gs->err, (int)DE_BARRIER);
                    ok = 0;
                    break;
                }
            }
        }
    }
    __syncthreads();
    __threadfence();
    return ok != 0;
}

__device__ __forceinline__ void list_finish(ListArrays &l, AtomArrays &a, int i, int cnt,
                                            DevState *gs) {
    l.nnbr[i] = min(cnt, l.stride);
    a.cref[i] = a.pos[i];
    atomicAdd(reinterpret_cast<unsigned long long *>(&gs->nlist_entries),
              (unsigned long long)min(cnt, l.stride));
    if (cnt > l.stride) {
        atomicOr(&gs->err, (int)DE_NBR_OVERFLOW);
        atomicMax(reinterpret_cast<unsigned long long *>(&gs->nbr_needed),
                  (unsigned long long)cnt);
    }
}

// Append the members of [b, e) of cell_atoms (or atoms b..e-1 when
// cell_atoms is null) within r_list of atom i, LIST_LANES lanes per atom, in
// member order.
__device__ __forceinline__ void list_scan_range(const AtomArrays &a, const ListArrays &l,
                                                const int *members, int b, int e, int i,
                                                const double4 &ci, const double *L, double rl2,
                                                int lane, unsigned long long gmask,
                                                unsigned long long below, int *row, int &cnt) {
    for (int k0 = b; k0 < e; k0 += LIST_LANES) {
        int const k = k0 + lane;
        bool in = false;
        int j = 0;
        if (k < e) {
            j = members ? members[k] : k;
            if (j != i) {
                double dx, dy, dz;
                min_image(ci, a.pos[j], L, dx, dy, dz);
                in = dx * dx + dy * dy + dz * dz <= rl2;
            }
        }
        unsigned long long const bal = __ballot(in) & gmask;
        if (in) {
            int const slot = cnt + __popcll(bal & below);
            if (slot < l.stride)
                row[slot] = j;
        }
        cnt += __popcll(bal);
    }
}

__global__ __launch_bounds__(BLOCK) void k_list(AtomArrays a, ListArrays l, int N,
                                               DevState *__restrict__ gs,
                                               unsigned *__restrict__ bar, int use_cells) {
    if (gs->err)
        return;
    // phase 0: the decision, taken identically by every block from the same
    // inputs (u_max^2 was reduced by k_kick_drift)
    double lam = 1e300;
    for (int c = 0; c < 3; ++c)
        lam = fmin(lam, gs->L[c] / gs->L_build[c]);
    bool const need = gs->force_rebuild || !(lam * gs->r_list - 2. * sqrt(gs->umax2) > gs->rc);
    if (!need)
        return;
    bool ok = grid_barrier(bar, gs);   // everyone has read the decision inputs
    int const nthreads = gridDim.x * BLOCK;
    int const gtid = blockIdx.x * BLOCK + threadIdx.x;
    double const L[3] = {gs->L[0], gs->L[1], gs->L[2]};
    double const rl2 = gs->r_list * gs->r_list;
    int const nx = gs->ncell[0], ny = gs->ncell[1], nz = gs->ncell[2];
    if (gtid == 0) {
        gs->force_rebuild = 0;
        gs->do_rebuild = 1;
        gs->nbuilds++;
        gs->nlist_entries = 0;
        for (int c = 0; c < 3; ++c)
            gs->L_build[c] = gs->L[c];
        if (use_cells)
            for (int c = 0; c < 3; ++c)
                if (gs->L[c] / gs->ncell[c] < gs->r_list) {
                    atomicOr(&gs->err, (int)DE_CELL_GRID);
                    gs->err_val[0] = gs->L[c];
                }
    }
    int const lane = threadIdx.x % LIST_LANES;
    int const l64 = threadIdx.x & 63;
    unsigned long long const gmask = group_mask<LIST_LANES>();
    unsigned long long const below = (l64 == 0) ? 0ull : (~0ull >> (64 - l64));
    int const groups = nthreads / LIST_LANES;
    if (use_cells) {
        // phase 1: bin
        for (int i = gtid; i < N; i += nthreads) {
            double4 const c = a.pos[i];
            int const cid = (cell_coord(c.z, nz) * ny + cell_coord(c.y, ny)) * nx +
                            cell_coord(c.x, nx);
            l.cell_of[i] = cid;
            atomicAdd(&l.cell_count[cid], 1);
        }
        ok = grid_barrier(bar, gs) && ok;
        // phase 2: exclusive scan of the counts (block 0); zero counts and fills
        if (blockIdx.x == 0) {
            __shared__ int sm[BLOCK];
            __shared__ int carry;
            if (threadIdx.x == 0)
                carry = 0;
            __syncthreads();
            for (int base = 0; base < l.ncells; base += BLOCK) {
                int const k = base + threadIdx.x;
                int const v = k < l.ncells ? l.cell_count[k] : 0;
                sm[threadIdx.x] = v;
                __syncthreads();
                for (int off = 1; off < BLOCK; off <<= 1) {
                    int const t = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
                    __syncthreads();
                    sm[threadIdx.x] += t;
                    __syncthreads();
                }
                if (k < l.ncells) {
                    l.cell_start[k] = carry + sm[threadIdx.x] - v;
                    l.cell_count[k] = 0;
                    l.cell_fill[k] = 0;
                }
                __syncthreads();
                if (threadIdx.x == BLOCK - 1)
                    carry += sm[BLOCK - 1];
                __syncthreads();
            }
            if (threadIdx.x == 0)
                l.cell_start[l.ncells] = carry;
        }
        ok = grid_barrier(bar, gs) && ok;
        // phase 3: scatter
        for (int i = gtid; i < N; i += nthreads) {
            int const cid = l.cell_of[i];
            int const slot = atomicAdd(&l.cell_fill[cid], 1);
            l.cell_atoms[l.cell_start[cid] + slot] = i;
        }
        ok = grid_barrier(bar, gs) && ok;
        // phase 4: one block per cell places each member at its rank
        {
            __shared__ int v[CELL_MAX];
            for (int c = blockIdx.x; c < l.ncells; c += gridDim.x) {
                int const b = l.cell_start[c], n = l.cell_start[c + 1] - b;
                if (n > CELL_MAX) {
                    if (threadIdx.x == 0) {
                        atomicOr(&gs->err, (int)DE_CELL_GRID);
                        gs->err_val[0] = (double)n;
                    }
                    continue;   // uniform in the block
                }
                for (int k = threadIdx.x; k < n; k += BLOCK)
                    v[k] = l.cell_atoms[b + k];
                __syncthreads();
                for (int k = threadIdx.x; k < n; k += BLOCK) {
                    int const x = v[k];
                    int rank = 0;
                    for (int m = 0; m < n; ++m)
                        rank += v[m] < x;
                    l.cell_atoms[b + rank] = x;
                }
                __syncthreads();
            }
        }
        ok = grid_barrier(bar, gs) && ok;
        if (!ok)
            return;
        // phase 5: lists from the 27 surrounding cells, in (cell, member) order
        for (int i = gtid / LIST_LANES; i < N; i += groups) {
            double4 const ci = a.pos[i];
            int const cx = cell_coord(ci.x, nx), cy = cell_coord(ci.y, ny), cz = cell_coord(ci.z, nz);
            int *row = l.nbr + (size_t)i * l.stride;
            int cnt = 0;
            for (int q = 0; q < 27; ++q) {
                int const dz = q / 9 - 1, dy = (q / 3) % 3 - 1, dx = q % 3 - 1;
                int const z = (cz + dz + nz) % nz, y = (cy + dy + ny) % ny, x = (cx + dx + nx) % nx;
                int const cell = (z * ny + y) * nx + x;
                list_scan_range(a, l, l.cell_atoms, l.cell_start[cell], l.cell_start[cell + 1], i,
                                ci, L, rl2, lane, gmask, below, row, cnt);
            }
            if (lane == 0)
                list_finish(l, a, i, cnt, gs);
        }
    } else {
        if (!ok)
            return;
        // small boxes (< 3 cells per direction): test every other atom
        for (int i = gtid / LIST_LANES; i < N; i += groups) {
            double4 const ci = a.pos[i];
            int *row = l.nbr + (size_t)i * l.stride;
            int cnt = 0;
            list_scan_range(a, l, nullptr, 0, N, i, ci, L, rl2, lane, gmask, below, row, cnt);
            if (lane == 0)
                list_finish(l, a, i, cnt, gs);
        }
    }
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
            cnt = l.nnbr[i];
        }
        const int *nb = l.nbr + (size_t)(active ? i : 0) * l.stride;
        int *cb = l.cnbr + (size_t)(active ? i : 0) * l.stride;
        int nin = 0;   // compacted entries so far (uniform in the group)
        for (int k0 = 0; k0 < cnt; k0 += LPA) {
            int const k = k0 + lane;
            bool in = false;
            int j = 0;
            double dx = 0., dy = 0., dz = 0., r2 = 0.;
            double4 cj;
            if (k < cnt) {
                j = nb[k];
                cj = a.pos[j];
                min_image(ci, cj, L, dx, dy, dz);
                r2 = dx * dx + dy * dy + dz * dz;
                in = !(r2 > rc2);
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
        for (int off = LPA / 2; off > 0; off >>= 1) {
            rho += __shfl_xor(rho, off, LPA);
            if (SAMPLE) {
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
        }
        if (active && lane == 0) {
            l.ccnt[i] = nin;
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
                gp[0] = make_double2(g0, g1);
                gp[1] = make_double2(g2, g3);
                gp[2] = make_double2(g4, g5);
                gp[3] = make_double2(hx * hx + hy * hy + hz * hz, 0.);
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
// factor 2, removed in k_post. The pair Born term X d_a d_b d_c d_d is fully
// symmetric in its four indices, so its 21 Voigt entries are 15 distinct
// monomials of d, accumulated here and expanded in k_post.
// ---------------------------------------------------------------------------
template <int KIND, int LPA, bool SAMPLE, bool EAM>
__global__ __launch_bounds__(BLOCK) void k_force(AtomArrays a, ListArrays l, RTables rt, int N,
                                                int npair, double rc2, double rc_pair,
                                                double rc_rho, double rmin,
                                                DevState *__restrict__ gs,
                                                double *__restrict__ part) {
    if (gs->err)
        return;
    using P = Poly<KIND>;
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
                double t;
                int const it = tab_locate(r, rt.inv_h, rt.n, t);
                const double *base = rt.rec + (size_t)it * rt.nf * P::S;
                double V = 0., dV = 0., d2V = 0.;
                if (!(r > rc_pair)) {   // potential.hpp:97-103
                    P pp;
                    pp.load(base + pslot * P::S);
                    V = pp.v(t);
                    dV = pp.d1(t, rt.inv_h);
                    if (SAMPLE)
                        d2V = pp.d2(t, rt.inv_h2);
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
    block_sum_store<PB_N>(acc, part);
}

// Second U2 kick (md_driver.cpp:433-452) and the per-atom sums the
// thermodynamics need: K, sum m v_a v_b, momentum, sum |f|^2; when sampled
// also the embedding part of the Born tensor and Laplacian, F''_i g_i g_i
// and F''_i |h_i|^2 (md_driver.cpp:410-419), from pass A's per-atom g, |h|^2.
template <bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_kick2(AtomArrays a, int N,
                                                const DevState *__restrict__ gs, int init,
                                                int has_eam, double *__restrict__ part) {
    if (gs->err)
        return;
    double const e = init ? 1. : gs->kick_e, ch = init ? 0. : gs->kick_c;
    double acc[PK_N];
#pragma unroll
    for (int k = 0; k < PK_N; ++k)
        acc[k] = 0.;
    for (int i = blockIdx.x * BLOCK + threadIdx.x; i < N; i += gridDim.x * BLOCK) {
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
    block_sum_store<PK_N>(acc, part);
}

// Deterministic sums of the per-block partial records, one block per
// component (thread t adds blocks t, t+BLOCK, ... in order, then a fixed
// tree); the last block to finish then runs the scalar part of the step:
// post_step (mode 0 = constructor, md_driver.cpp:114-118; mode 1 = end of
// integrate(), md_driver.cpp:583-600) and, when next_sample >= 0, the first
// half of the next step (pre_step), which needs nothing in between.
// Components are numbered A (PA_N), then B (PB_N), then K (PK_N).
constexpr int NTOTAL = PA_N + PB_N + PK_N;

__device__ void post_step(DevState &s, const double *tA, const double *tB, const double *tK,
                          int mode, double *__restrict__ rec_out);

__global__ __launch_bounds__(BLOCK) void k_finish(DevState *__restrict__ gs,
                                                 const double *__restrict__ pA, int nA,
                                                 const double *__restrict__ pB, int nB,
                                                 const double *__restrict__ pK, int nK,
                                                 double *__restrict__ totals,
                                                 unsigned *__restrict__ ticket, int mode,
                                                 double *__restrict__ rec_out, int next_sample,
                                                 int next_force) {
    if (gs->err)
        return;
    int k = blockIdx.x;
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
    double sum = 0.;
    for (int b = threadIdx.x; b < nb; b += BLOCK)
        sum += p[(size_t)k * PART_STRIDE + b];
    for (int off = 32; off > 0; off >>= 1)
        sum += __shfl_xor(sum, off, 64);
    __shared__ double sm[BLOCK / 64];
    __shared__ int last;
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < BLOCK / 64; ++w)
            t += sm[w];
        totals[blockIdx.x] = t;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last)
        return;
    __threadfence();
    __shared__ double tot[NTOTAL];
    for (int c = threadIdx.x; c < NTOTAL; c += BLOCK)
        tot[c] = __hip_atomic_load(&totals[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __shared__ DevState s;
    state_load(s, gs);
    if (threadIdx.x == 0) {
        post_step(s, tot, tot + PA_N, tot + PA_N + PB_N, mode, rec_out);
        if (next_sample >= 0 && !s.err) {
            pre_step(s, next_sample);
            if (next_force)
                s.force_rebuild = 1;
        }
        *ticket = 0u;
    }
    state_store(gs, s);
}

__device__ void post_step(DevState &s, const double *tA, const double *tB, const double *tK,
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
    double const ekin_atom = s.ekin_tot / N;
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
    double const epot_atom = tB[PB_EPOT] * uE / N;
    double const epot_eam_atom = s.has_eam ? tA[PA_EF] * uE / N : 0.;
    double const etot_atom = ekin_atom + epot_atom;
    double const old = s.etot_cons;
    double ec = etot_atom * N;
    double const V = box_volume(s.L);
    if (thermo)
        ec += chain_etot(s.th);
    if (npt)
        ec += V * s.bar_P + (0.5 * (s.peps * s.peps) / s.bar_W + chain_etot(s.bth));
    s.etot_cons = ec;
    s.etot_cons_err = fabs(ec / s.etot_cons_initial - 1.);
    double const err_cur = fabs(ec / old - 1.);
    double const Tk = 2. * s.ekin_tot / (K_BOLTZ * s.Nf);
    double const P = (s.ekin_tot * 2. / 3. + s.virial) / V;
    double st[6];
    double const s2 = scale * scale;
    for (int v = 0; v < 6; ++v)
        st[v] = (-sB[v] + s2 * tK[PK_MVV + v]) / V;
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
        rec[MDL_S_BORN_STRESS + v] = sB[v] / V;
        rec[MDL_S_STRESS + v] = st[v];
    }
    if (sample) {
        int q = 0;
        for (int v = 0; v < 6; ++v)
            for (int w = 0; w < 6; ++w)
                rec[MDL_S_BORN_ELAST + 6 * v + w] = w >= v ? CB[q++] / V : 0.;
    }
    rec[MDL_S_NLIST_BUILDS] = (double)s.nbuilds;
    rec[MDL_S_WARN_VIRIAL] = (double)s.warn_virial;
    rec[MDL_S_WARN_RMIN] = (double)s.warn_rmin;
    rec[MDL_S_NLIST_ENTRIES] = (double)s.nlist_entries;
    if (rec_out && sample)
        for (int k = 0; k < MDL_NRECORD; ++k)
            rec_out[k] = rec[k];
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
