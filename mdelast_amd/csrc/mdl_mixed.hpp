// Mixed-precision pair passes (MDL_MIXED). Included by mdl_driver.hip after
// mdl_kernels.hpp.
//
// What stays FP64: positions, the pair vector d and r^2 (so the set of pairs
// inside r_c and the antisymmetry d_ji = -d_ij are those of the FP64 path),
// the per-atom force and energy accumulators, rho_i, and every block-level
// reduction. What is FP32: the table polynomials (records converted to eV,
// Angstrom units, since CGS values over- and underflow FP32), the pair
// algebra of dE/dr, X and the Laplacian, and the per-lane partial sums of the
// virial/Born/Laplacian terms. dE/dr is formed without contraction so that the
// pair force of (i, j) and (j, i) is bit-identical: the FP64 force sums then
// cancel pairwise and the net momentum stays at FP64 round-off.
#pragma once

#include "mdl_kernels.hpp"

namespace mdl {

// s_waitcnt vmcnt(0) (every vector memory load of the wave has landed, also
// those written to LDS) and s_waitcnt lgkmcnt(0) (every LDS access has
// returned), in the gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4,
// lgkmcnt 11:8; a field at its maximum does not wait.
__device__ __forceinline__ void mdl_wait_vmem() {
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ void mdl_wait_lds() {
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
}

// Table records in FP32, Angstrom units. Quintic records are padded to 8 floats
// (32 B) so each loads as one float4 + one float2. `prec` holds the force
// pass's fused records, one per interval and species pair (PairRecF).
struct RTablesF {
    const float *rec;
    const float *prec;
    int nf, n;
    float inv_h, inv_h2;   // 1/A, 1/A^2
};
struct FTablesF {
    const float *rec;      // [n][ntypes][S]
    int n;
    double inv_h;          // A^3 (the argument is rho in 1/A^3)
    float inv_hf, inv_h2f;
    double rhomax[MAX_TYPES];   // 1/A^3
};

template <int KIND> struct PolyF;
template <> struct PolyF<MDL_QUINTIC> {
    static constexpr int S = 8;
    float c[6];
    __device__ __forceinline__ void load(const float *p) {
        float4 const a = *reinterpret_cast<const float4 *>(p);
        float2 const b = *reinterpret_cast<const float2 *>(p + 4);
        c[0] = a.x; c[1] = a.y; c[2] = a.z; c[3] = a.w; c[4] = b.x; c[5] = b.y;
    }
    __device__ __forceinline__ float v(float t) const {
        return c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
    }
    __device__ __forceinline__ float d1(float t, float ih) const {
        return (c[1] + t * (2.f * c[2] + t * (3.f * c[3] + t * (4.f * c[4] + t * 5.f * c[5])))) * ih;
    }
    __device__ __forceinline__ float d2(float t, float ih2) const {
        return (2.f * c[2] + t * (6.f * c[3] + t * (12.f * c[4] + t * 20.f * c[5]))) * ih2;
    }
};
template <> struct PolyF<MDL_BSPLINE> {
    static constexpr int S = 12;
    float c[12];
    __device__ __forceinline__ void load(const float *p) {
        const float4 *q = reinterpret_cast<const float4 *>(p);
        float4 const a = q[0], b = q[1], e = q[2];
        c[0] = a.x; c[1] = a.y; c[2] = a.z; c[3] = a.w;
        c[4] = b.x; c[5] = b.y; c[6] = b.z; c[7] = b.w;
        c[8] = e.x; c[9] = e.y; c[10] = e.z; c[11] = e.w;
    }
    __device__ __forceinline__ float v(float t) const { return c[0] + t * (c[1] + t * (c[2] + t * c[3])); }
    __device__ __forceinline__ float d1(float t, float) const { return c[4] + t * (c[5] + t * (c[6] + t * c[7])); }
    __device__ __forceinline__ float d2(float t, float) const { return c[8] + t * (c[9] + t * (c[10] + t * c[11])); }
};

// The force pass's table record of one interval and one species pair (a <= b),
// aligned to its size: phi_ab, then the derivative parts of rho_a and rho_b.
// One record is one cache line or half of one, so a pair's three lookups touch
// one line instead of three.
//   quintic (16 floats, 64 B): phi c0..c5 | rho_a c1..c5 | rho_b c1..c5
//   B-spline (32 floats, 128 B): phi value, d1, d2 cubics (12) | rho_a d1, d2
//     cubics (8) | rho_b d1, d2 cubics (8) | 4 pad
// The derivative polynomials are the PolyF ones, term for term.
template <int KIND> struct PairRecF;
template <> struct PairRecF<MDL_QUINTIC> {
    static constexpr int NF = 16;
    float c[16];
    __device__ __forceinline__ void load(const float *p) {
        const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float4 const v = q[k];
            c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w;
        }
    }
    // from an LDS record whose chunk k sits in slot (k + rot) mod 4
    __device__ __forceinline__ void load_rot(const float4 *p, int rot) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float4 const v = p[(k + rot) & 3];
            c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w;
        }
    }
    __device__ __forceinline__ float v(float t) const {
        return c[0] + t * (c[1] + t * (c[2] + t * (c[3] + t * (c[4] + t * c[5]))));
    }
    __device__ __forceinline__ float d1(float t, float ih) const {
        return (c[1] + t * (2.f * c[2] + t * (3.f * c[3] + t * (4.f * c[4] + t * 5.f * c[5])))) * ih;
    }
    __device__ __forceinline__ float d2(float t, float ih2) const {
        return (2.f * c[2] + t * (6.f * c[3] + t * (12.f * c[4] + t * 20.f * c[5]))) * ih2;
    }
    // rho of the lower (O = 6) or higher (O = 11) species: c1..c5 at O
    template <int o> __device__ __forceinline__ float rd1(float t, float ih) const {
        return (c[o] + t * (2.f * c[o + 1] + t * (3.f * c[o + 2] + t * (4.f * c[o + 3] +
                                                                        t * 5.f * c[o + 4])))) * ih;
    }
    template <int o> __device__ __forceinline__ float rd2(float t, float ih2) const {
        return (2.f * c[o + 1] + t * (6.f * c[o + 2] + t * (12.f * c[o + 3] +
                                                             t * 20.f * c[o + 4]))) * ih2;
    }
    static constexpr int RHO_LO = 6, RHO_HI = 11;
};
template <> struct PairRecF<MDL_BSPLINE> {
    static constexpr int NF = 32;
    float c[28];
    __device__ __forceinline__ void load(const float *p) {
        const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            float4 const v = q[k];
            c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w;
        }
    }
    // from an LDS record whose chunk k sits in slot (k + rot) mod 8 (chunk 7 is padding)
    __device__ __forceinline__ void load_rot(const float4 *p, int rot) {
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            float4 const v = p[(k + rot) & 7];
            c[4 * k] = v.x; c[4 * k + 1] = v.y; c[4 * k + 2] = v.z; c[4 * k + 3] = v.w;
        }
    }
    __device__ __forceinline__ float v(float t) const { return c[0] + t * (c[1] + t * (c[2] + t * c[3])); }
    __device__ __forceinline__ float d1(float t, float) const { return c[4] + t * (c[5] + t * (c[6] + t * c[7])); }
    __device__ __forceinline__ float d2(float t, float) const { return c[8] + t * (c[9] + t * (c[10] + t * c[11])); }
    template <int o> __device__ __forceinline__ float rd1(float t, float) const {
        return c[o] + t * (c[o + 1] + t * (c[o + 2] + t * c[o + 3]));
    }
    template <int o> __device__ __forceinline__ float rd2(float t, float) const {
        return c[o + 4] + t * (c[o + 5] + t * (c[o + 6] + t * c[o + 7]));
    }
    static constexpr int RHO_LO = 12, RHO_HI = 20;
};

// Interval and local coordinate as tab_locate (range test before the
// conversion, NaN -> interval 0, outside -> end polynomial), without branches:
// clamping u to [0, n-1] first gives the same interval. nm1 = n - 1.
__device__ __forceinline__ int tab_locate_f(float x, float inv_h, float nm1, float &t) {
    float const u = x * inv_h;
    int const i = (int)fminf(fmaxf(u, 0.f), nm1);
    t = u - (float)i;
    return i;
}

// Squared pair distance from the fixed-point pair vector, in one fixed form
// (explicit FMAs), so that both passes take the same r_c decision for a pair.
__device__ __forceinline__ float pair_r2(float x, float y, float z) {
    return fmaf(z, z, fmaf(y, y, x * x));
}

// The pair algebra that must be symmetric in (i, j): products and sums in a
// fixed order with no fused operations.
__device__ __forceinline__ float sym_dedr(float dphi, float Fi, float drj, float Fj, float dri) {
#pragma clang fp contract(off)
    float const a = Fi * drj;
    float const b = Fj * dri;
    return dphi + (a + b);
}

// ---------------------------------------------------------------------------
// Pass A, mixed: as k_density (FP64 rho), with the pair geometry from the
// fixed-point positions and the rho(r) and F(rho) polynomials and the g/h sums
// in FP32. It scans the mid list (the full list on the steps that refilter it)
// and, unlike the FP64 pass, writes no compacted r_c list: pass B scans the mid
// list too and takes the same r_c decision (pair_r2 against rc2f, identical bits
// in both passes and for both visits of a pair).
// ---------------------------------------------------------------------------
template <int KIND, int LPA, bool SAMPLE>
__device__ __forceinline__ void density_mx_body(AtomArrays a, ListArrays l, RTablesF rt,
                                                FTablesF ft, int N, int npair, int ntypes,
                                                double rc2, DevState *__restrict__ gs,
                                                double *__restrict__ part, Blk bk) {
    if (gs->err)
        return;
    using P = PolyF<KIND>;
    constexpr int APB = BLOCK / LPA;
    int const lane = threadIdx.x % LPA, slot = threadIdx.x / LPA;
    int const l64 = threadIdx.x & 63;
    unsigned long long const gmask = group_mask<LPA>();
    unsigned long long const below = (l64 == 0) ? 0ull : (~0ull >> (64 - l64));
    float const S[3] = {(float)(gs->L[0] * 1e8) * 0x1p-32f, (float)(gs->L[1] * 1e8) * 0x1p-32f,
                        (float)(gs->L[2] * 1e8) * 0x1p-32f};   // A per fixed-point unit
    // scan the mid list, or the full list while the mid list is refiltered
    bool const mid_build = gs->mid_build != 0;
    // r_mid widened by 2e-6 relative: the float r of a pair errs by far less,
    // so every pair the exact test keeps is kept
    float const rm2 = (float)((gs->r_mid * 1e8) * (gs->r_mid * 1e8) * (1. + 4e-6));   // A^2
    float const rc2f = (float)rc2;
    int const rowS = rt.nf * P::S;           // floats per interval (fits int)
    float const nm1 = (float)(rt.n - 1);
    double acc[PA_N];
#pragma unroll
    for (int k = 0; k < PA_N; ++k)
        acc[k] = 0.;
    XcdRange const R((N + APB - 1) / APB, bk);
    for (int grp = R.first; grp < R.end; grp += R.step) {
        int const i = grp * APB + slot;
        bool const active = i < N;
        double rho = 0.;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, g4 = 0.f, g5 = 0.f;
        float hx = 0.f, hy = 0.f, hz = 0.f;
        int4 qi = make_int4(0, 0, 0, 0);
        int cnt = 0;
        if (active) {
            qi = a.qpos[i];
            cnt = mid_build ? l.nnbr[i] : l.mcnt[i];
        }
        int const ia = active ? i : 0;
        const int *nb = (mid_build ? l.nbr : l.mnbr) + (size_t)ia * l.stride;
        int *mb = l.mnbr + (size_t)ia * l.stride;
        int nmid = 0;
        // Software pipeline: the position of the next candidate and the index
        // of the one after are loaded while this candidate is processed (the
        // table record of this candidate is requested before them, so waiting
        // for it does not wait for them). Lanes past the end use atom ia itself
        // as a harmless dummy; every contribution is selected by `in`.
        int j0 = lane < cnt ? nb[lane] : ia;
        int j1 = LPA + lane < cnt ? nb[LPA + lane] : ia;
        int4 c0 = a.qpos[j0];
        for (int k0 = 0; k0 < cnt; k0 += LPA) {
            int const k = k0 + lane;
            int const j = j0;
            int4 const cj = c0;
            float fx, fy, fz;
            qmin_image(qi, cj, S, fx, fy, fz);
            float const r2 = pair_r2(fx, fy, fz);
            bool const in = k < cnt && !(r2 > rc2f);
            if (mid_build) {
                bool const inm = k < cnt && !(r2 > rm2);
                unsigned long long const balm = __ballot(inm) & gmask;
                if (inm)
                    mb[nmid + __popcll(balm & below)] = j;
                nmid += __popcll(balm);
            }
            float const r = __builtin_amdgcn_sqrtf(r2);
            int const tj = cj.w;
            float t;
            int const it = tab_locate_f(r, rt.inv_h, nm1, t);
            // the candidates beyond r_c (about one in seven of a mid list, half
            // of the full list when it is refiltered) request no table record
            P p;
            if (in)
                p.load(rt.rec + (it * rowS + (npair + tj) * P::S));
            c0 = a.qpos[j1];
            j0 = j1;
            j1 = k + 2 * LPA < cnt ? nb[k + 2 * LPA] : ia;
            rho += in ? (double)p.v(t) : 0.;
            if (SAMPLE) {
                float const w = in ? p.d1(t, rt.inv_h) * __builtin_amdgcn_rcpf(r) : 0.f;
                float const wx = w * fx, wy = w * fy, wz = w * fz;
                g0 += wx * fx;
                g1 += wy * fy;
                g2 += wz * fz;
                g3 += wy * fz;
                g4 += wx * fz;
                g5 += wx * fy;
                hx += wx;
                hy += wy;
                hz += wz;
            }
        }
#pragma unroll
        for (int off = LPA / 2; off > 0; off >>= 1)
            rho += __shfl_xor(rho, off, LPA);
        // sampled steps: g (6) and h (3) of the atom, each summed into one lane
        // of its group (groups of 16 lanes or more), which stores its g
        // component; |h|^2 is summed from the h lanes' squares
        float gh = 0.f, h2 = 0.f;
        int c = -1;
        if constexpr (SAMPLE && LPA >= 16) {
            float gs9[9] = {g0, g1, g2, g3, g4, g5, hx, hy, hz};
            gh = group_sum_scatter<LPA, 9>(gs9, c);
            h2 = c >= 6 ? gh * gh : 0.f;
#pragma unroll
            for (int off = LPA / 2; off > 0; off >>= 1)
                h2 += __shfl_xor(h2, off, LPA);
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
        if constexpr (SAMPLE && LPA >= 16) {
            if (active && c >= 0 && c < 6)
                reinterpret_cast<double *>(a.geam + 4 * (size_t)i)[c] = (double)gh;
        }
        if (active && lane == 0) {
            if (mid_build) {
                l.mcnt[i] = nmid;
                a.mref[i] = a.pos[i];
            }
            int const ti = qi.w;
            // selected, not indexed: a dynamic index into the table struct
            // would put a copy of it in scratch memory
            double rmax = ft.rhomax[0];
#pragma unroll
            for (int t = 1; t < MAX_TYPES; ++t)
                rmax = ti == t ? ft.rhomax[t] : rmax;
            if (rho > rmax) {
                atomicOr(&gs->err, (int)DE_RHOMAX);
                gs->err_val[0] = rho * 1e24;
                gs->err_val[1] = rmax * 1e24;
            }
            double const u = rho * ft.inv_h;
            int it = 0;
            if (u >= (double)ft.n)
                it = ft.n - 1;
            else if (u > 0.)
                it = (int)u;
            float const t = (float)(u - (double)it);
            P p;
            p.load(ft.rec + ((size_t)it * ntypes + ti) * P::S);
            float const F = p.v(t), dF = p.d1(t, ft.inv_hf);
            float const d2F = SAMPLE ? p.d2(t, ft.inv_h2f) : 0.f;
            a.Femb[i] = (double)F;
            a.rho[i] = rho * 1e24;                     // CGS for the attribute
            a.fprime[i] = make_double2((double)dF, (double)d2F);
            a.fprime_f[i] = make_float2(dF, d2F);
            acc[PA_EF] += (double)F;
            if (SAMPLE) {
                double2 *gp = a.geam + 4 * (size_t)i;
                if constexpr (LPA < 16) {
                    gp[0] = make_double2(g0, g1);
                    gp[1] = make_double2(g2, g3);
                    gp[2] = make_double2(g4, g5);
                }
                gp[3] = make_double2((double)h2, 0.);
            }
        }
    }
    block_sum_store<PA_N>(acc, part, bk.id);
}

template <int KIND, int LPA, bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_density_mx(AtomArrays a, ListArrays l, RTablesF rt,
                                                     FTablesF ft, int N, int npair, int ntypes,
                                                     double rc2, DevState *__restrict__ gs,
                                                     double *__restrict__ part) {
    density_mx_body<KIND, LPA, SAMPLE>(a, l, rt, ft, N, npair, ntypes, rc2, gs, part, hw_blk());
}

// ---------------------------------------------------------------------------
// Pass B, mixed: over the mid list, the pairs within r_c by pass A's decision;
// pair geometry from the fixed-point positions; forces (FP64 sums of FP32 pair
// terms, symmetric in the pair), energies,
// Born stress, and when sampled the Born monomials and Laplacian (FP32 per
// lane). The virial is taken as the trace of the Born stress in k_post.
// Outputs per atom in CGS; block sums in eV and Angstrom (scaled in k_post).
// ---------------------------------------------------------------------------
template <int KIND, int LPA, bool SAMPLE>
__device__ __forceinline__ void force_mx_body(AtomArrays a, ListArrays l, RTablesF rt, int N,
                                              int npair, float rmin, float rc2f,
                                              DevState *__restrict__ gs,
                                              double *__restrict__ part, Blk bk) {
    if (gs->err)
        return;
    using Rec = PairRecF<KIND>;
    constexpr int APB = BLOCK / LPA;
    constexpr double F_CGS = 1.6021773e-12 / 1e-8;   // eV/A -> dyn
    constexpr double E_CGS = 1.6021773e-12;          // eV -> erg
    int const lane = threadIdx.x % LPA, slot = threadIdx.x / LPA;
    float const S[3] = {(float)(gs->L[0] * 1e8) * 0x1p-32f, (float)(gs->L[1] * 1e8) * 0x1p-32f,
                        (float)(gs->L[2] * 1e8) * 0x1p-32f};   // A per fixed-point unit
    float const nm1 = (float)(rt.n - 1);
    float sb0 = 0.f, sb1 = 0.f, sb2 = 0.f, sb3 = 0.f, sb4 = 0.f, sb5 = 0.f;
    float cb[15];
#pragma unroll
    for (int k = 0; k < 15; ++k)
        cb[k] = 0.f;
    float lap = 0.f;
    double epot = 0.;
    int nrmin = 0;
    // Cooperative record loads straight into LDS: in round q of CH, lane l
    // fetches one 16-byte chunk of the record of lane p = q * (64 / CH) + l / CH,
    // so the CH lanes that share a record share its cache line in one load
    // instruction; the load writes lane l's chunk to stage + 16 l (the
    // global-to-LDS form of gfx950), which is record p's 16 CH bytes. Within a
    // record the chunk c sits in slot (c + rot(p)) mod CH, rot(p) = p CH / 16, so
    // when each lane then reads its own record chunk by chunk, the 16 lanes of
    // a 16-byte read fall into distinct banks.
    // One stage per wave, filled by one round of cooperative loads per pair.
    // The waits are explicit: the compiler does not order every LDS read after
    // the global-to-LDS loads that fill it.
    constexpr int CH = Rec::NF / 4;
    __shared__ float4 stage_all[BLOCK / 64][64 * CH];
    // the wave's stage, its index made wave-uniform (the LDS base of the
    // global-to-LDS loads is a scalar register)
    float4 *const stage = stage_all[__builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64))];
    int const l64 = threadIdx.x & 63;
    int const rot = l64 * CH / 16;
    const float4 *const prec4 = reinterpret_cast<const float4 *>(rt.prec);
    XcdRange const R((N + APB - 1) / APB, bk);
    for (int grp = R.first; grp < R.end; grp += R.step) {
        int const i = grp * APB + slot;
        bool const active = i < N;
        double fx = 0., fy = 0., fz = 0., e = 0.;
        // every lane of the wave takes part in every load round, so the trip
        // count is the wave's longest list; lanes past their own list (and
        // lanes of no atom) carry atom ia itself as a harmless dummy
        int const ia = active ? i : 0;
        int const cnt = active ? l.mcnt[i] : 0;
        int wmax = cnt;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            wmax = max(wmax, __shfl_xor(wmax, off, 64));
        {
            int4 const ci = a.qpos[ia];
            int const ti = ci.w;
            float2 const Fi = a.fprime_f[ia];
            const int *nb = l.mnbr + (size_t)ia * l.stride;
            struct Geo {
                float x, y, z, r, t;
                int tj;
                bool act;
            };
            // pair vector, interval and record of one pair
            auto geometry = [&](const int4 &cj, bool act, Geo &g) -> int {
                qmin_image(ci, cj, S, g.x, g.y, g.z);
                float const r2 = pair_r2(g.x, g.y, g.z);
                g.r = __builtin_amdgcn_sqrtf(r2);
                g.tj = cj.w;
                g.act = act && !(r2 > rc2f);
                int const tmax = max(ti, g.tj), tmin = min(ti, g.tj);
                int const pslot = ((tmax * (tmax + 1)) >> 1) + tmin;
                int const it = tab_locate_f(g.r, rt.inv_h, nm1, g.t);
                return act ? it * npair + pslot : 0;
            };
            // the wave's records into stage st, round q from the lanes q * (64 / CH) ..
            auto request = [&](float4 *st, int rec) {
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    int const src = q * (64 / CH) + l64 / CH;
                    int const rsrc = __shfl(rec, src, 64);
                    int const chunk = (l64 % CH - src * CH / 16) & (CH - 1);
                    // a 32-bit offset from the table base (the record index
                    // times CH stays far below 2^32)
                    __builtin_amdgcn_global_load_lds(
                        (const void __attribute__((address_space(1))) *)(
                            prec4 + (unsigned)(rsrc * CH + chunk)),
                        (void __attribute__((address_space(3))) *)(st + q * 64), 16, 0, 0);
                }
            };
            auto physics = [&](const Geo &g, const Rec &pp, float2 Fj) {
                float const x = g.x, y = g.y, z = g.z, r = g.r, t = g.t;
                float const xx = x * x, yy = y * y, zz = z * z;
                nrmin += r < rmin;
                // rho of species i and j: the record's lower or higher one
                bool const ilo = ti <= g.tj;
                float const V = pp.v(t), dV = pp.d1(t, rt.inv_h);
                float const dlo = pp.template rd1<Rec::RHO_LO>(t, rt.inv_h);
                float const dhi = pp.template rd1<Rec::RHO_HI>(t, rt.inv_h);
                float const dri = ilo ? dlo : dhi, drj = ilo ? dhi : dlo;
                float const ir = __builtin_amdgcn_rcpf(r);
                float const dedr = sym_dedr(dV, Fi.x, drj, Fj.x, dri);
                float const s = -dedr * ir;
                fx += (double)s * (double)x;   // exact products of two floats
                fy += (double)s * (double)y;
                fz += (double)s * (double)z;
                e += (double)(0.5f * V);
                float const w = dedr * ir;
                sb0 += w * xx;
                sb1 += w * yy;
                sb2 += w * zz;
                sb3 += w * (y * z);
                sb4 += w * (x * z);
                sb5 += w * (x * y);
                if (SAMPLE) {
                    float const d2V = pp.d2(t, rt.inv_h2);
                    float const d2lo = pp.template rd2<Rec::RHO_LO>(t, rt.inv_h2);
                    float const d2hi = pp.template rd2<Rec::RHO_HI>(t, rt.inv_h2);
                    float const d2ri = ilo ? d2lo : d2hi, d2rj = ilo ? d2hi : d2lo;
                    float X = d2V - dV * ir;
                    lap += 2.f * (d2V + 2.f * dV * ir);
                    lap += 2.f * (Fj.x * (d2ri + 2.f * dri * ir) + Fi.x * (d2rj + 2.f * drj * ir));
                    lap += Fi.y * (drj * drj) + Fj.y * (dri * dri);
                    X += Fi.x * (d2rj - drj * ir) + Fj.x * (d2ri - dri * ir);
                    X *= ir * ir;
                    float const Xxx = X * xx, Xyy = X * yy, Xzz = X * zz;
                    float const yz = y * z, xz = x * z, xy = x * y;
                    cb[M_X4] += Xxx * xx;
                    cb[M_Y4] += Xyy * yy;
                    cb[M_Z4] += Xzz * zz;
                    cb[M_X3Y] += Xxx * xy;
                    cb[M_X3Z] += Xxx * xz;
                    cb[M_Y3X] += Xyy * xy;
                    cb[M_Y3Z] += Xyy * yz;
                    cb[M_Z3X] += Xzz * xz;
                    cb[M_Z3Y] += Xzz * yz;
                    cb[M_X2Y2] += Xxx * yy;
                    cb[M_X2Z2] += Xxx * zz;
                    cb[M_Y2Z2] += Xyy * zz;
                    cb[M_X2YZ] += Xxx * yz;
                    cb[M_Y2XZ] += Xyy * xz;
                    cb[M_Z2XY] += Xzz * xy;
                }
            };
            // Software pipeline: the next pair's position and the index after it
            // are requested after this pair's table record.
            int j0 = lane < cnt ? nb[lane] : ia;
            int j1 = LPA + lane < cnt ? nb[LPA + lane] : ia;
            int4 c0 = a.qpos[j0];
            for (int k = lane; k - lane < wmax; k += LPA) {
                Geo g;
                int const rec = geometry(c0, k < cnt, g);
                int const j = j0;
                mdl_wait_lds();   // the stage was read out for the pair before
                request(stage, rec);
                float2 const Fj = a.fprime_f[j];
                c0 = a.qpos[j1];
                j0 = j1;
                j1 = k + 2 * LPA < cnt ? nb[k + 2 * LPA] : ia;
                mdl_wait_vmem();
                Rec pp;
                pp.load_rot(stage + l64 * CH, rot);
                if (g.act)
                    physics(g, pp, Fj);
            }
        }
        // the atom's sums: fx, fy, fz and e each land in one lane of its group
        double fs[4] = {fx, fy, fz, e};
        int c;
        double const v = group_sum_scatter<LPA, 4>(fs, c);
        if (active && c >= 0) {
            if (c == 0) {
                a.fx[i] = v * F_CGS;
            } else if (c == 1) {
                a.fy[i] = v * F_CGS;
            } else if (c == 2) {
                a.fz[i] = v * F_CGS;
            } else {
                double const ei = a.Femb[i] + v;   // eV
                a.epot[i] = ei * E_CGS;
                epot += ei;
            }
        }
    }
    double acc[PB_N];
#pragma unroll
    for (int k = 0; k < PB_N; ++k)
        acc[k] = 0.;
    acc[PB_EPOT] = epot;
    acc[PB_SB + 0] = sb0;
    acc[PB_SB + 1] = sb1;
    acc[PB_SB + 2] = sb2;
    acc[PB_SB + 3] = sb3;
    acc[PB_SB + 4] = sb4;
    acc[PB_SB + 5] = sb5;
#pragma unroll
    for (int k = 0; k < 15; ++k)
        acc[PB_CB + k] = cb[k];
    acc[PB_LAP] = lap;
    acc[PB_WRMIN] = (double)nrmin;
    pass_sum_store<SAMPLE, PB_N, PB_CB, PB_LAP + 1>(acc, part, bk.id);
}

// Four waves per SIMD: the force pass hides its gathers better with 128 VGPRs
// (and a few spilled) than with three waves at 148 (measured on the 108,000-atom
// step: 95 against 98 us); the density pass is fastest at its natural four.
template <int KIND, int LPA, bool SAMPLE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void k_force_mx(AtomArrays a, ListArrays l, RTablesF rt, int N, int npair, float rmin,
                float rc2f, DevState *__restrict__ gs, double *__restrict__ part) {
    force_mx_body<KIND, LPA, SAMPLE>(a, l, rt, N, npair, rmin, rc2f, gs, part, hw_blk());
}

}  // namespace mdl
