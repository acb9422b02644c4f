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

// Table records in FP32, Angstrom units. Quintic records are padded to 8 floats
// (32 B) so each loads as one float4 + one float2.
struct RTablesF {
    const float *rec;
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

// Interval and local coordinate as tab_locate (range test before the
// conversion, NaN -> interval 0, outside -> end polynomial), without branches:
// clamping u to [0, n-1] first gives the same interval. nm1 = n - 1.
__device__ __forceinline__ int tab_locate_f(float x, float inv_h, float nm1, float &t) {
    float const u = x * inv_h;
    int const i = (int)fminf(fmaxf(u, 0.f), nm1);
    t = u - (float)i;
    return i;
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
// Pass A, mixed: as k_density (FP64 geometry, compaction, FP64 rho), with the
// rho(r) and F(rho) polynomials and the g/h sums in FP32.
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
    double const L[3] = {gs->L[0] * 1e8, gs->L[1] * 1e8, gs->L[2] * 1e8};   // A
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
        double4 ci = make_double4(0., 0., 0., 0.);
        int cnt = 0;
        if (active) {
            ci = a.pos[i];
            cnt = l.nnbr[i];
        }
        int const ia = active ? i : 0;
        const int *nb = l.nbr + (size_t)ia * l.stride;
        int *cb = l.cnbr + (size_t)ia * l.stride;
        int nin = 0;
        // Software pipeline: the position of the next candidate and the index
        // of the one after are loaded while this candidate is processed (the
        // table record of this candidate is requested before them, so waiting
        // for it does not wait for them). Lanes past the end use atom ia itself
        // as a harmless dummy; every contribution is selected by `in`.
        int j0 = lane < cnt ? nb[lane] : ia;
        int j1 = LPA + lane < cnt ? nb[LPA + lane] : ia;
        double4 c0 = a.pos[j0];
        for (int k0 = 0; k0 < cnt; k0 += LPA) {
            int const k = k0 + lane;
            int const j = j0;
            double4 const cj = c0;
            double dx, dy, dz;
            min_image(ci, cj, L, dx, dy, dz);
            bool const in = k < cnt && !(dx * dx + dy * dy + dz * dz > rc2);
            float const fx = (float)dx, fy = (float)dy, fz = (float)dz;
            float const r = __builtin_amdgcn_sqrtf(fx * fx + fy * fy + fz * fz);
            int const tj = (int)cj.w;
            float t;
            int const it = tab_locate_f(r, rt.inv_h, nm1, t);
            P p;
            p.load(rt.rec + (it * rowS + (npair + tj) * P::S));
            c0 = a.pos[j1];
            j0 = j1;
            j1 = k + 2 * LPA < cnt ? nb[k + 2 * LPA] : ia;
            unsigned long long const bal = __ballot(in) & gmask;
            if (in)
                cb[nin + __popcll(bal & below)] = j;
            nin += __popcll(bal);
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
                gp[0] = make_double2(g0, g1);
                gp[1] = make_double2(g2, g3);
                gp[2] = make_double2(g4, g5);
                gp[3] = make_double2((double)(hx * hx + hy * hy + hz * hz), 0.);
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
// Pass B, mixed: forces (FP64 sums of symmetric FP32 pair terms), energies,
// Born stress, and when sampled the Born monomials and Laplacian (FP32 per
// lane). The virial is taken as the trace of the Born stress in k_post.
// Outputs per atom in CGS; block sums in eV and Angstrom (scaled in k_post).
// ---------------------------------------------------------------------------
template <int KIND, int LPA, bool SAMPLE>
__device__ __forceinline__ void force_mx_body(AtomArrays a, ListArrays l, RTablesF rt, int N,
                                              int npair, float rmin, DevState *__restrict__ gs,
                                              double *__restrict__ part, Blk bk) {
    if (gs->err)
        return;
    using P = PolyF<KIND>;
    constexpr int APB = BLOCK / LPA;
    constexpr double F_CGS = 1.6021773e-12 / 1e-8;   // eV/A -> dyn
    constexpr double E_CGS = 1.6021773e-12;          // eV -> erg
    int const lane = threadIdx.x % LPA, slot = threadIdx.x / LPA;
    double const L[3] = {gs->L[0] * 1e8, gs->L[1] * 1e8, gs->L[2] * 1e8};
    int const rowS = rt.nf * P::S;           // floats per interval (fits int)
    float const nm1 = (float)(rt.n - 1);
    float sb0 = 0.f, sb1 = 0.f, sb2 = 0.f, sb3 = 0.f, sb4 = 0.f, sb5 = 0.f;
    float cb[15];
#pragma unroll
    for (int k = 0; k < 15; ++k)
        cb[k] = 0.f;
    float lap = 0.f;
    double epot = 0.;
    int nrmin = 0;
    XcdRange const R((N + APB - 1) / APB, bk);
    for (int grp = R.first; grp < R.end; grp += R.step) {
        int const i = grp * APB + slot;
        bool const active = i < N;
        double fx = 0., fy = 0., fz = 0., e = 0.;
        if (active) {
            double4 const ci = a.pos[i];
            int const ti = (int)ci.w;
            float2 const Fi = a.fprime_f[i];
            int const cnt = l.ccnt[i];
            const int *nb = l.cnbr + (size_t)i * l.stride;
            // Software pipeline as in pass A: the next pair's position and the
            // index after it are requested after this pair's table records.
            int j0 = lane < cnt ? nb[lane] : i;
            int j1 = LPA + lane < cnt ? nb[LPA + lane] : i;
            double4 c0 = a.pos[j0];
            for (int k = lane; k < cnt; k += LPA) {
                int const j = j0;
                double4 const cj = c0;
                double dx, dy, dz;
                min_image(ci, cj, L, dx, dy, dz);
                float const x = (float)dx, y = (float)dy, z = (float)dz;
                float const xx = x * x, yy = y * y, zz = z * z;
                float const r = __builtin_amdgcn_sqrtf(xx + yy + zz);
                int const tj = (int)cj.w;
                int const tmax = max(ti, tj), tmin = min(ti, tj);
                int const pslot = ((tmax * (tmax + 1)) >> 1) + tmin;
                nrmin += r < rmin;
                float t;
                int const it = tab_locate_f(r, rt.inv_h, nm1, t);
                const float *base = rt.rec + it * rowS;
                P pp, pi, pj;
                pp.load(base + pslot * P::S);
                pi.load(base + (npair + ti) * P::S);
                pj.load(base + (npair + tj) * P::S);
                float2 const Fj = a.fprime_f[j];
                c0 = a.pos[j1];
                j0 = j1;
                j1 = k + 2 * LPA < cnt ? nb[k + 2 * LPA] : i;
                float const V = pp.v(t), dV = pp.d1(t, rt.inv_h);
                float const dri = pi.d1(t, rt.inv_h), drj = pj.d1(t, rt.inv_h);
                float const ir = __builtin_amdgcn_rcpf(r);
                float const dedr = sym_dedr(dV, Fi.x, drj, Fj.x, dri);
                float const s = -dedr * ir;
                fx += (double)s * dx;
                fy += (double)s * dy;
                fz += (double)s * dz;
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
                    float const d2ri = pi.d2(t, rt.inv_h2), d2rj = pj.d2(t, rt.inv_h2);
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
            a.fx[i] = fx * F_CGS;
            a.fy[i] = fy * F_CGS;
            a.fz[i] = fz * F_CGS;
            double const ei = a.Femb[i] + e;   // eV
            a.epot[i] = ei * E_CGS;
            epot += ei;
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

template <int KIND, int LPA, bool SAMPLE>
__global__ __launch_bounds__(BLOCK) void k_force_mx(AtomArrays a, ListArrays l, RTablesF rt,
                                                   int N, int npair, float rmin,
                                                   DevState *__restrict__ gs,
                                                   double *__restrict__ part) {
    force_mx_body<KIND, LPA, SAMPLE>(a, l, rt, N, npair, rmin, gs, part, hw_blk());
}

}  // namespace mdl
