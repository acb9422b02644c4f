// Learned Stein control variate: device kernels (C-ABI in include/mdelast_cv.h).
//
// A per-atom network g(x) maps local radial-basis descriptors D_i to a 3-vector
// per estimator head; the control variate of a head is
//   s(x) = sum_i [ div_i g_i + beta g_i . F_i ],   E_NVT[s] = 0,
// where div_i g_i = sum_a d g_{i,a} / d x_{i,a} needs, for each atom, the
// derivatives of the outputs along the three coordinates of that atom. They
// are carried through the network in forward mode: every layer maps inputs X
// and three tangents T_a (X' along x_{i,a}) to Y = act(X W + b) and
// Y'_a = act'(X W + b) (T_a W). The descriptor kernel supplies the first
// tangents, J_a = d D_i / d x_{i,a}. The dense products run on MFMA
// (bf16 operands, fp32 accumulation, 16x16x16 tiles); the backward pass of a
// layer is three more products of the same kinds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "mdl_internal.hpp"
#include "mdl_cv.hpp"
#include "../../include/mdelast_cv.h"

namespace {

using mdl::set_error;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float x) {
    // round to nearest even (inputs are finite)
    unsigned const u = __float_as_uint(x);
    unsigned const r = u + 0x7fffu + ((u >> 16) & 1u);
    return (short)(r >> 16);
}

// ---------------------------------------------------------------------------
// Descriptors by brute force over all pairs (the reference form for positions
// handed over by the caller; the driver's own frames use its neighbour list).
// One wave per two atoms: lanes 0-31 take atom 2w, lanes 32-63 atom 2w+1,
// lane k%32 owns radial function k for both neighbour species.
// Positions in A (FP64), minimum image in FP64, basis in FP32.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_cv_descriptors(const double *__restrict__ pos,
                                                        const int *__restrict__ types, int N,
                                                        double Lx, double Ly, double Lz,
                                                        mdl::CvBasis basis, float *__restrict__ D,
                                                        float *__restrict__ J,
                                                        float *__restrict__ L) {
    int const lane = threadIdx.x & 63, k = lane & 31;
    int const i = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 2 + (lane >> 5);
    bool const active = i < N;
    int const ia = active ? i : 0;
    double const xi = pos[3 * ia], yi = pos[3 * ia + 1], zi = pos[3 * ia + 2];
    mdl::CvAcc<> acc;
    for (int j = 0; j < N; ++j) {
        double sx = xi - pos[3 * j], sy = yi - pos[3 * j + 1], sz = zi - pos[3 * j + 2];
        sx -= Lx * rint(sx / Lx);
        sy -= Ly * rint(sy / Ly);
        sz -= Lz * rint(sz / Lz);
        float const x = (float)sx, y = (float)sy, z = (float)sz;
        float const r2 = x * x + y * y + z * z;
        if (j == ia || !(r2 < basis.rc2))
            continue;   // the same for the 32 lanes of an atom
        acc.add(basis, k, types[j], x, y, z, sqrtf(r2));
    }
    if (active)
        acc.store(i, k, N, D, J, L);
}

// ---------------------------------------------------------------------------
// C[R][N] = A[R][K] B + bias, B = W[K][N] (or W^T when transB, W[N][K]).
// One wave per 16x16 tile of C; K, N multiples of 16; rows past R are zero.
// MFMA 16x16x16 bf16 lane layout: A row l%16, k 4(l/16)+i; B col l%16,
// k 4(l/16)+i; D col l%16, row 4(l/16)+r.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_nn(const float *__restrict__ A,
                                                 const float *__restrict__ W, int transB,
                                                 const float *__restrict__ bias,
                                                 float *__restrict__ C, int R, int K, int N) {
    int const lane = threadIdx.x & 63;
    int const tile = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int const ntn = N / 16;
    int const tm = tile / ntn, tn = tile % ntn;
    int const r0 = tm * 16, n0 = tn * 16;
    if (r0 >= R)
        return;   // whole wave
    int const row = r0 + (lane & 15), col = n0 + (lane & 15), kq = 4 * (lane >> 4);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int const kk = k0 + kq + q;
            a[q] = bf16_rne(row < R ? A[(size_t)row * K + kk] : 0.f);
            b[q] = bf16_rne(transB ? W[(size_t)col * K + kk] : W[(size_t)kk * N + col]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int const rr = r0 + kq + q;
        if (rr < R)
            C[(size_t)rr * N + col] = acc[q] + (bias ? bias[col] : 0.f);
    }
}

// The same product with the B operand staged once per block in LDS as bf16,
// transposed to [N][K + 4] (so a lane's four k values are one 8-byte read and
// the 16 columns of a tile fall in distinct banks), and each wave's A
// fragments (16 rows x K) held in registers across all N/16 column tiles.
// Waves stride over 16-row tiles; KT = K / 16.
template <int KT>
__global__ __launch_bounds__(256) void k_gemm_lds(const float *__restrict__ A,
                                                  const float *__restrict__ W, int transB,
                                                  const float *__restrict__ bias,
                                                  float *__restrict__ C, int R, int N) {
    extern __shared__ short wl[];
    constexpr int K = KT * 16, KP = K + 4;
    for (int e = threadIdx.x; e < K * N; e += blockDim.x) {
        int k, n;
        if (transB) {
            n = e / K;
            k = e - n * K;
        } else {
            k = e / N;
            n = e - k * N;
        }
        wl[n * KP + k] = bf16_rne(W[e]);
    }
    __syncthreads();
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const kq = 4 * (lane >> 4), c16 = lane & 15;
    for (int r0 = (blockIdx.x * 4 + wave) * 16; r0 < R; r0 += gridDim.x * 64) {
        int const row = r0 + c16;
        bf16x4 af[KT];
#pragma unroll
        for (int t = 0; t < KT; ++t)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                af[t][q] = bf16_rne(row < R ? A[(size_t)row * K + 16 * t + kq + q] : 0.f);
        for (int n0 = 0; n0 < N; n0 += 16) {
            int const col = n0 + c16;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                bf16x4 const b = *reinterpret_cast<const bf16x4 *>(&wl[col * KP + 16 * t + kq]);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[t], b, acc, 0, 0, 0);
            }
            float const bb = bias ? bias[col] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int const rr = r0 + kq + q;
                if (rr < R)
                    C[(size_t)rr * N + col] = acc[q] + bb;
            }
        }
    }
}

// part[c][K][N] = sum over rows r of chunk c of A[r][K]^T B[r][N]; chunks of
// `rows` rows (a multiple of 16). One wave per (chunk, 16x16 tile of C).
__global__ __launch_bounds__(256) void k_gemm_tn(const float *__restrict__ A,
                                                 const float *__restrict__ B,
                                                 float *__restrict__ part, int R, int K, int N,
                                                 int rows, int nchunk) {
    int const lane = threadIdx.x & 63;
    int const w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int const ntk = K / 16, ntn = N / 16;
    int const per = ntk * ntn;
    int const c = w / per, t = w % per;
    if (c >= nchunk)
        return;
    int const k0 = (t / ntn) * 16, n0 = (t % ntn) * 16;
    int const kq = 4 * (lane >> 4);
    int const rb = c * rows, re = min(R, rb + rows);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int s0 = rb; s0 < re; s0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int const rr = s0 + kq + q;
            bool const ok = rr < re;
            a[q] = bf16_rne(ok ? A[(size_t)rr * K + k0 + (lane & 15)] : 0.f);
            b[q] = bf16_rne(ok ? B[(size_t)rr * N + n0 + (lane & 15)] : 0.f);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float *out = part + (size_t)c * K * N;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        out[(size_t)(k0 + kq + q) * N + n0 + (lane & 15)] = acc[q];
}

// out[e] = sum_c part[c][e], fixed order (deterministic)
__global__ void k_sum_chunks(const float *__restrict__ part, float *__restrict__ out, int nchunk,
                             int n) {
    int const e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n)
        return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c)
        s += part[(size_t)c * n + e];
    out[e] = s;
}

// part[c][n] = sum of X[r][n] over the rows r of chunk c (CS_ROWS rows per
// block); a block takes 64 columns (blockIdx.y) with 4 row lanes, then adds
// the lanes in a fixed order.
constexpr int CS_ROWS = 1024;

__global__ __launch_bounds__(256) void k_colsum_part(const float *__restrict__ X,
                                                     float *__restrict__ part, int R, int N) {
    __shared__ float sm[256];
    int const n = blockIdx.y * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
    int const r0 = blockIdx.x * CS_ROWS, r1 = min(R, r0 + CS_ROWS);
    float s = 0.f;
    if (n < N)
        for (int r = r0 + g; r < r1; r += 4)
            s += X[(size_t)r * N + n];
    sm[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 64 && n < N)
        part[(size_t)blockIdx.x * N + n] = (sm[threadIdx.x] + sm[64 + threadIdx.x]) +
                                           (sm[128 + threadIdx.x] + sm[192 + threadIdx.x]);
}

// Activation with tangents. act 0: tanh; 1: identity.
// Y = act(Z); Yt[a] = act'(Z) Zt[a], a = 0..2 (Zt, Yt are [3][M][N])
__global__ void k_act_fwd(const float *__restrict__ Z, const float *__restrict__ Zt,
                          float *__restrict__ Y, float *__restrict__ Yt, int64_t MN, int act) {
    int64_t const e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MN)
        return;
    float const z = Z[e];
    float y = z, s = 1.f;
    if (act == 0) {
        y = tanhf(z);
        s = 1.f - y * y;
    }
    Y[e] = y;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        Yt[a * MN + e] = s * Zt[a * MN + e];
}

// Backward of k_act_fwd: gZ = gY act'(Z) + sum_a gYt[a] Zt[a] act''(Z);
// gZt[a] = gYt[a] act'(Z).
__global__ void k_act_bwd(const float *__restrict__ Z, const float *__restrict__ Zt,
                          const float *__restrict__ gY, const float *__restrict__ gYt,
                          float *__restrict__ gZ, float *__restrict__ gZt, int64_t MN, int act) {
    int64_t const e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= MN)
        return;
    float s = 1.f, s2 = 0.f;
    if (act == 0) {
        float const y = tanhf(Z[e]);
        s = 1.f - y * y;
        s2 = -2.f * y * s;
    }
    float g = gY[e] * s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float const gt = gYt[a * MN + e];
        g += gt * Zt[a * MN + e] * s2;
        gZt[a * MN + e] = gt * s;
    }
    gZ[e] = g;
}

// ---------------------------------------------------------------------------
// Fields from their weights. For atom m and field f (weights w[m][f][k] and
// their tangents wt[a][m][f][k], k = 0..63):
//   c_a   = sum_k w_k J[a][m][k]                      (the field, 3-vector)
//   div   = sum_k (w_k L[m][k] + sum_a wt_a,k J[a][m][k])
//   a     = div + beta c . F[m],  cF = c . F[m],  cG = c . G[m]
// One wave per atom, lane = k; the k-sums are wave reductions in a fixed
// order. Memory-bound: w and wt are read once.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_field_contract(
    const float *__restrict__ w, const float *__restrict__ wt, const float *__restrict__ J,
    const float *__restrict__ L, const float *__restrict__ F, const float *__restrict__ G,
    float beta, int M, int nfield, float *__restrict__ a_out, float *__restrict__ cf_out,
    float *__restrict__ cg_out) {
    int const k = threadIdx.x & 63;
    int const m = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (m >= M)
        return;   // whole wave
    size_t const NFK = (size_t)nfield * 64;
    size_t const jm = (size_t)m * 64 + k;
    size_t const jplane = (size_t)M * 64;
    float const jx = J[jm], jy = J[jplane + jm], jz = J[2 * jplane + jm], l = L[jm];
    float const fx = F[3 * m], fy = F[3 * m + 1], fz = F[3 * m + 2];
    float const gx = G[3 * m], gy = G[3 * m + 1], gz = G[3 * m + 2];
    size_t const wplane = (size_t)M * NFK;
    for (int f = 0; f < nfield; ++f) {
        size_t const o = (size_t)m * NFK + (size_t)f * 64 + k;
        float const wk = w[o];
        float const d = wk * l + (wt[o] * jx + wt[wplane + o] * jy + wt[2 * wplane + o] * jz);
        float const cx = wave_sum(wk * jx), cy = wave_sum(wk * jy), cz = wave_sum(wk * jz);
        float const dv = wave_sum(d);
        if (k == 0) {
            float const cF = cx * fx + cy * fy + cz * fz;
            a_out[(size_t)m * nfield + f] = dv + beta * cF;
            cf_out[(size_t)m * nfield + f] = cF;
            cg_out[(size_t)m * nfield + f] = cx * gx + cy * gy + cz * gz;
        }
    }
}

// Its backward with respect to the weights (J, L, F, G are inputs):
//   gw[m][f][k]     = ga (L_k + beta JF_k) + gcF JF_k + gcG JG_k
//   gwt[a][m][f][k] = ga J[a][m][k]
// with JF_k = sum_a J[a][m][k] F_a, JG_k = sum_a J[a][m][k] G_a.
__global__ __launch_bounds__(256) void k_field_contract_bwd(
    const float *__restrict__ J, const float *__restrict__ L, const float *__restrict__ F,
    const float *__restrict__ G, float beta, const float *__restrict__ ga,
    const float *__restrict__ gcf, const float *__restrict__ gcg, int M, int nfield,
    float *__restrict__ gw, float *__restrict__ gwt) {
    int const k = threadIdx.x & 63;
    int const m = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (m >= M)
        return;
    size_t const NFK = (size_t)nfield * 64;
    size_t const jm = (size_t)m * 64 + k;
    size_t const jplane = (size_t)M * 64;
    float const jx = J[jm], jy = J[jplane + jm], jz = J[2 * jplane + jm], l = L[jm];
    float const JF = jx * F[3 * m] + jy * F[3 * m + 1] + jz * F[3 * m + 2];
    float const JG = jx * G[3 * m] + jy * G[3 * m + 1] + jz * G[3 * m + 2];
    size_t const wplane = (size_t)M * NFK;
    for (int f = 0; f < nfield; ++f) {
        size_t const o = (size_t)m * NFK + (size_t)f * 64 + k;
        size_t const q = (size_t)m * nfield + f;
        float const a = ga[q];
        gw[o] = a * (l + beta * JF) + gcf[q] * JF + gcg[q] * JG;
        gwt[o] = a * jx;
        gwt[wplane + o] = a * jy;
        gwt[2 * wplane + o] = a * jz;
    }
}

// ---------------------------------------------------------------------------
// Stress proxy h_v(x) = off_v + sum over ordered pairs (i, j), r_ij < rc, of
// psi_{t_i + t_j}(r_ij) n_a n_b (Voigt v = (a, b), n = d / r, d = x_i - x_j),
// psi_p(r) = fc(r) sum_k c[p][k] g_k(r), g_k = exp(-(r - mu_k)^2 / (2 w^2)).
// Its Stein fields need, per atom m, grad_m h_v = 2 sum_j T_v(d_mj) with
//   T_v,c = psi' n_a n_b n_c + (psi / r)(d_ac n_b + d_bc n_a - 2 n_a n_b n_c),
// and the Laplacian lap_m h_v = 2 sum_j [(psi'' + 2 psi'/r - 6 psi/r^2) n_a n_b
// + 2 (psi / r^2) d_ab] (the Laplacian in d of psi(r) n_a n_b; both atoms of
// a pair see it). Brute force over all pairs, FP64 throughout (the Stein sums
// need the derivatives of exactly the function evaluated).
// ---------------------------------------------------------------------------
__constant__ int SP_VA[6] = {0, 1, 2, 1, 0, 0};
__constant__ int SP_VB[6] = {0, 1, 2, 2, 2, 1};

struct SpBasis {
    double mu0, dmu, iw2, pirc, rc2;
    double q;   // exp(-dmu^2 / w^2): the ratio of successive Gaussian ratios
};

// The basis sums of one frame: Z[p][k][v] = sum over ordered pairs of species
// pair p = t_i + t_j of fc(r) g_k(r) n_a n_b. One wave per atom at a time: the
// lanes that find a neighbour within rc (64 candidates at a time) write its
// geometry (n, r, fc and the species pair) to LDS once, in index order; the
// half-waves then take alternate pairs from there (lane & 31 = k), 18 sums per
// lane; per block, the sums over its waves and halves in a fixed order.
constexpr int SB_NBR = 128;   // neighbours buffered per wave

__global__ __launch_bounds__(256) void k_cv_stress_basis(const double *__restrict__ pos,
                                                         const int *__restrict__ types, int N,
                                                         double Lx, double Ly, double Lz,
                                                         SpBasis b,
                                                         const int *__restrict__ plist,
                                                         const int *__restrict__ pcount,
                                                         int width, double *__restrict__ part) {
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6, k = lane & 31, half = lane >> 5;
    int const frame = blockIdx.y;
    const double *P = pos + (size_t)frame * N * 3;
    // the neighbour geometry of the pair loop and the final per-lane sums share
    // the block's LDS
    __shared__ union {
        struct {
            double geo[4][SB_NBR][5];
            int pt[4][SB_NBR];
        } nb;
        double red[4][64][18];
    } sh;
    auto &geo = sh.nb.geo;
    auto &gpt = sh.nb.pt;
    unsigned long long const below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    double acc[3][6];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int v = 0; v < 6; ++v)
            acc[p][v] = 0.0;
    double const mu = b.mu0 + b.dmu * k;
    double const iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
    for (int i = blockIdx.x * 4 + wave; i < N; i += gridDim.x * 4) {
        double const xi = P[3 * i], yi = P[3 * i + 1], zi = P[3 * i + 2];
        int const ti = types[i];
        int cnt = 0;
        // candidates: every atom, or the atom's captured pair list in list order
        const int *pl = nullptr;
        int nc = N;
        if (plist) {
            size_t const o = (size_t)frame * N + i;
            pl = plist + o * width;
            nc = min(pcount[o], width);
        }
        for (int j0 = 0; j0 < nc; j0 += 64) {
            int const t = j0 + lane;
            int const j = pl ? (t < nc ? pl[t] : i) : t;
            double dx = 0., dy = 0., dz = 0., r2 = 1e300;
            if (t < nc && j != i) {
                dx = xi - P[3 * j];
                dy = yi - P[3 * j + 1];
                dz = zi - P[3 * j + 2];
                dx -= Lx * rint(dx * iLx);
                dy -= Ly * rint(dy * iLy);
                dz -= Lz * rint(dz * iLz);
                r2 = dx * dx + dy * dy + dz * dz;
            }
            bool const hit = r2 < b.rc2;
            unsigned long long const m = __ballot(hit);
            if (hit) {
                int const s = cnt + __popcll(m & below);
                double const r = sqrt(r2), ir = 1.0 / r;
                geo[wave][s][0] = dx * ir;
                geo[wave][s][1] = dy * ir;
                geo[wave][s][2] = dz * ir;
                geo[wave][s][3] = r;
                geo[wave][s][4] = 0.5 * (cos(b.pirc * r) + 1.0);
                gpt[wave][s] = ti + types[j];
            }
            cnt += __popcll(m);
            if (cnt <= SB_NBR - 64 && j0 + 64 < nc)
                continue;
            // this wave's LDS writes before any lane reads them
            __builtin_amdgcn_s_waitcnt(0xC07F);
            __builtin_amdgcn_wave_barrier();
            for (int t = half; t < cnt; t += 2) {
                double const n[3] = {geo[wave][t][0], geo[wave][t][1], geo[wave][t][2]};
                double const u = geo[wave][t][3] - mu;
                double const g = exp(-0.5 * u * u * b.iw2) * geo[wave][t][4];
                int const pt = gpt[wave][t];
#pragma unroll
                for (int v = 0; v < 6; ++v) {
                    double const w = g * n[SP_VA[v]] * n[SP_VB[v]];
                    acc[0][v] += pt == 0 ? w : 0.0;
                    acc[1][v] += pt == 1 ? w : 0.0;
                    acc[2][v] += pt == 2 ? w : 0.0;
                }
            }
            cnt = 0;
            // the batch is read out before the next one is written
            __builtin_amdgcn_s_waitcnt(0xC07F);
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();   // every wave is done with the neighbour geometry
    auto &red = sh.red;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int v = 0; v < 6; ++v)
            red[wave][lane][p * 6 + v] = acc[p][v];
    __syncthreads();
    double *out = part + ((size_t)frame * gridDim.x + blockIdx.x) * 576;
    for (int q = threadIdx.x; q < 576; q += 256) {
        int const p = q / 192, kk = (q / 6) % 32, v = q % 6;
        double t = 0.0;
        for (int w = 0; w < 4; ++w)
            t += red[w][kk][p * 6 + v] + red[w][kk + 32][p * 6 + v];
        out[q] = t;   // [p][k][v]
    }
}

// The proxy's per-atom terms with fitted coefficients c [3][32]. One wave per
// atom i. The wave screens all atoms 64 at a time and collects the neighbours
// within rc in LDS, in index order (batches of up to SP_NBR); then each lane
// takes one neighbour and evaluates psi, psi' and psi'' of its pair over the 32
// basis functions itself, so nothing is summed across lanes per pair, and adds
// the pair's terms to the atom's 30 sums: q < 6 the share of h_v, 6 <= q < 12
// the Laplacian of h_v, 12 + 3 v + c the gradient component d h_v / d x_{i,c}.
// The wave sums them (recursive halving, one sum per lane).
constexpr int SP_NBR = 256;   // neighbours buffered per wave
constexpr int SP_FIELD = 39;  // per-frame field sums of k_cv_stress_fields

__device__ __forceinline__ void sp_min_image(const double *P, int i, int j, double Lx, double Ly,
                                             double Lz, double iLx, double iLy, double iLz,
                                             double &dx, double &dy, double &dz) {
    dx = P[3 * i] - P[3 * j];
    dy = P[3 * i + 1] - P[3 * j + 1];
    dz = P[3 * i + 2] - P[3 * j + 2];
    dx -= Lx * rint(dx * iLx);
    dy -= Ly * rint(dy * iLy);
    dz -= Lz * rint(dz * iLz);
}

// sum_k c[k] g_k(r) (times 1, -q_k and q_k^2 - 1/w^2, q_k = (r - mu_k)/w^2:
// psi's three radial sums). RECUR: the Gaussians from four anchors (k = 0, 8,
// 16, 24, each an exp) by their ratio g_{k+1}/g_k = exp((r - mu_k) dmu / w^2 -
// dmu^2 / (2 w^2)), whose own ratio is the constant b.q; within 7 steps of an
// anchor the product stays within a few ulp. The host takes this form only
// where no anchor can underflow; else one exp per basis function.
template <bool RECUR>
__device__ __forceinline__ void sp_radial(double r, const double *__restrict__ c, const SpBasis &b,
                                          double &a0, double &a1, double &a2) {
    a0 = a1 = a2 = 0.0;
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 8) {
        double g = 0.0, R = 0.0;
        if (RECUR) {
            double const u0 = r - (b.mu0 + b.dmu * k0);
            g = exp(-0.5 * u0 * u0 * b.iw2);
            R = exp(b.iw2 * (u0 * b.dmu - 0.5 * b.dmu * b.dmu));
        }
#pragma unroll
        for (int k = k0; k < k0 + 8; ++k) {
            double const u = r - (b.mu0 + b.dmu * k), qv = u * b.iw2;
            double const gk = RECUR ? g : exp(-0.5 * u * qv);
            double const cg = c[k] * gk;
            a0 += cg;
            a1 += -cg * qv;
            a2 += cg * (qv * qv - b.iw2);
            if (RECUR) {
                g *= R;
                R *= b.q;
            }
        }
    }
}

// The 30 per-atom sums of atom i in `acc` (each lane its own pairs' share).
// Candidates: every atom (screened 64 at a time, the neighbours collected in
// LDS in index order), or with `pl` the atom's captured pair list (pl[0..pc),
// in list order, re-checked against rc).
template <bool RECUR>
__device__ __forceinline__ void sp_atom_terms(const double *__restrict__ P,
                                              const int *__restrict__ types, int N, int i,
                                              double Lx, double Ly, double Lz, const SpBasis &b,
                                              const double *cs, int *nb, const int *pl, int pc,
                                              double (&acc)[30]) {
    int const lane = threadIdx.x & 63;
    double const iLx = 1.0 / Lx, iLy = 1.0 / Ly, iLz = 1.0 / Lz;
    int const ti = types[i];
    unsigned long long const below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int q = 0; q < 30; ++q)
        acc[q] = 0.0;
    // the terms of neighbour jj (within rc) into this lane's sums
    auto pair = [&](int jj, double ex, double ey, double ez) {
        double const r = sqrt(ex * ex + ey * ey + ez * ez), ir = 1.0 / r;
        double a0, a1, a2;
        sp_radial<RECUR>(r, cs + 32 * (ti + types[jj]), b, a0, a1, a2);
        double const cr = cos(b.pirc * r), sr = sin(b.pirc * r);
        double const fc = 0.5 * (cr + 1.0), dfc = -0.5 * b.pirc * sr,
                     ddfc = -0.5 * b.pirc * b.pirc * cr;
        double const psi = fc * a0, dpsi = dfc * a0 + fc * a1,
                     ddpsi = ddfc * a0 + 2.0 * dfc * a1 + fc * a2;
        double const n[3] = {ex * ir, ey * ir, ez * ir};
        double const pr = psi * ir;
        double const lapc = ddpsi + 2.0 * dpsi * ir - 6.0 * pr * ir;
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            int const va = v < 3 ? v : (v == 3 ? 1 : 0), vb = v < 3 ? v : (v == 5 ? 1 : 2);
            double const nab = n[va] * n[vb];
            acc[v] += psi * nab;
            acc[6 + v] += 2.0 * (lapc * nab + (va == vb ? 2.0 * pr * ir : 0.0));
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
                acc[12 + 3 * v + ax] +=
                    2.0 * (dpsi * nab * n[ax] +
                           pr * ((va == ax ? n[vb] : 0.0) + (vb == ax ? n[va] : 0.0) -
                                 2.0 * nab * n[ax]));
        }
    };
    if (pl) {
        for (int t = lane; t < pc; t += 64) {
            int const jj = pl[t];
            double ex, ey, ez;
            sp_min_image(P, i, jj, Lx, Ly, Lz, iLx, iLy, iLz, ex, ey, ez);
            if (jj != i && ex * ex + ey * ey + ez * ez < b.rc2)
                pair(jj, ex, ey, ez);
        }
        return;
    }
    int cnt = 0;
    for (int j0 = 0; j0 < N; j0 += 64) {
        int const j = j0 + lane;
        bool hit = false;
        if (j < N && j != i) {
            double dx, dy, dz;
            sp_min_image(P, i, j, Lx, Ly, Lz, iLx, iLy, iLz, dx, dy, dz);
            hit = dx * dx + dy * dy + dz * dz < b.rc2;
        }
        unsigned long long const m = __ballot(hit);
        if (hit)
            nb[cnt + __popcll(m & below)] = j;
        cnt += __popcll(m);
        if (cnt <= SP_NBR - 64 && j0 + 64 < N)
            continue;
        // this wave's LDS writes before any lane reads them
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
        for (int t0 = 0; t0 < cnt; t0 += 64) {
            int const t = t0 + lane;
            if (t >= cnt)
                continue;
            int const jj = nb[t];
            double ex, ey, ez;
            sp_min_image(P, i, jj, Lx, Ly, Lz, iLx, iLy, iLz, ex, ey, ez);
            pair(jj, ex, ey, ez);
        }
        cnt = 0;
        // the batch is read out before the next one is written
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
    }
}

// Per atom: hl[i] = the first 12 sums, grad[i][v][c] the other 18.
template <bool RECUR>
__global__ __launch_bounds__(256) void k_cv_stress_proxy(const double *__restrict__ pos,
                                                         const int *__restrict__ types, int N,
                                                         double Lx, double Ly, double Lz,
                                                         SpBasis b, const double *__restrict__ coef,
                                                         double *__restrict__ grad,
                                                         double *__restrict__ hl) {
    __shared__ double cs[96];
    __shared__ int nbr[4][SP_NBR];
    if (threadIdx.x < 96)
        cs[threadIdx.x] = coef[threadIdx.x];
    __syncthreads();
    int const wave = threadIdx.x >> 6;
    int const i = blockIdx.x * 4 + wave;
    if (i >= N)
        return;   // a whole wave
    int const frame = blockIdx.y;
    double acc[30];
    sp_atom_terms<RECUR>(pos + (size_t)frame * N * 3, types, N, i, Lx, Ly, Lz, b, cs, nbr[wave],
                         nullptr, 0, acc);
    int q;
    double const sum = mdl::group_sum_scatter<64, 30>(acc, q);
    if (q >= 0) {
        size_t const o = (size_t)frame * N + i;
        if (q < 12)
            hl[o * 12 + q] = sum;
        else
            grad[o * 18 + (q - 12)] = sum;
    }
}

// The per-frame sums the variate's stress fields need, without the per-atom
// arrays: over the frame's atoms m, the share of h_v (6), lap_m h_v (6),
// grad_m h_v . F_m (6, F = fscale times the forces given) and
// grad_m h_a . grad_m h_b (21, a <= b, row by row). Each wave takes atoms in
// turn; lane q keeps the running sum of quantity q; per block, the waves in
// order: part[frame][block][39].
template <bool RECUR>
__global__ __launch_bounds__(256) void k_cv_stress_fields(const double *__restrict__ pos,
                                                          const double *__restrict__ F,
                                                          double fscale,
                                                          const int *__restrict__ types, int N,
                                                          double Lx, double Ly, double Lz,
                                                          SpBasis b,
                                                          const double *__restrict__ coef,
                                                          const int *__restrict__ plist,
                                                          const int *__restrict__ pcount,
                                                          int width,
                                                          double *__restrict__ part) {
    __shared__ double cs[96];
    __shared__ int nbr[4][SP_NBR];
    __shared__ double sv[4][32];
    __shared__ double red[4][SP_FIELD];
    if (threadIdx.x < 96)
        cs[threadIdx.x] = coef[threadIdx.x];
    __syncthreads();
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int const frame = blockIdx.y;
    const double *P = pos + (size_t)frame * N * 3;
    const double *Ff = F + (size_t)frame * N * 3;
    // lane's quantity: 18 + index of the pair (a, b), a <= b, for lanes 18..38
    int pa = 0, pb = 0;
    if (lane >= 18 && lane < SP_FIELD) {
        int t = lane - 18;
        while (t >= 6 - pa) {
            t -= 6 - pa;
            ++pa;
        }
        pb = pa + t;
    }
    double run = 0.0;
    for (int i = blockIdx.x * 4 + wave; i < N; i += gridDim.x * 4) {
        double acc[30];
        const int *pl = nullptr;
        int pc = 0;
        if (plist) {
            size_t const o = (size_t)frame * N + i;
            pl = plist + o * width;
            pc = min(pcount[o], width);
        }
        sp_atom_terms<RECUR>(P, types, N, i, Lx, Ly, Lz, b, cs, nbr[wave], pl, pc, acc);
        int q;
        double const sum = mdl::group_sum_scatter<64, 30>(acc, q);
        if (q >= 0)
            sv[wave][q] = sum;
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
        const double *g = sv[wave] + 12;   // g[3 v + c] = d h_v / d x_{i,c}
        double val = 0.0;
        if (lane < 12) {
            val = sv[wave][lane];
        } else if (lane < 18) {
            int const v = lane - 12;
            val = fscale * (g[3 * v] * Ff[3 * i] + g[3 * v + 1] * Ff[3 * i + 1] +
                            g[3 * v + 2] * Ff[3 * i + 2]);
        } else if (lane < SP_FIELD) {
            val = g[3 * pa] * g[3 * pb] + g[3 * pa + 1] * g[3 * pb + 1] +
                  g[3 * pa + 2] * g[3 * pb + 2];
        }
        run += val;
        // the atom's sums are read out before the next atom's are written
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_wave_barrier();
    }
    if (lane < SP_FIELD)
        red[wave][lane] = run;
    __syncthreads();
    if (threadIdx.x < SP_FIELD)
        part[((size_t)frame * gridDim.x + blockIdx.x) * SP_FIELD + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
            red[3][threadIdx.x];
}

mdl_status launch_ok(const char *what) {
    hipError_t const e = hipGetLastError();
    if (e != hipSuccess)
        return mdl::hip_fail(e, what);
    return MDL_OK;
}

}  // namespace

extern "C" {

mdl_status mdl_cv_descriptors(const double *pos, const int32_t *types, int32_t N,
                              const double *box, float rc, float mu0, float dmu, float width,
                              float *D, float *J, float *L, void *stream) {
    if (!pos || !types || !box || !D || !J || N < 1 || !(rc > 0.f) || !(width > 0.f)) {
        set_error("mdl_cv_descriptors: bad arguments");
        return MDL_ERR_ARG;
    }
    // the minimum image is exact (and the descriptors smooth and periodic) only
    // while the cut-off sphere fits in the box
    if (!(2.0 * rc <= box[0] && 2.0 * rc <= box[1] && 2.0 * rc <= box[2])) {
        set_error("mdl_cv_descriptors: the cut-off exceeds half the box");
        return MDL_ERR_ARG;
    }
    int const waves = (N + 1) / 2;
    int const blocks = (waves * 64 + 255) / 256;
    k_cv_descriptors<<<blocks, 256, 0, (hipStream_t)stream>>>(
        pos, types, N, box[0], box[1], box[2], mdl::cv_basis(rc, mu0, dmu, width), D, J, L);
    return launch_ok("mdl_cv_descriptors");
}

static bool sp_args(const char *who, const double *pos, const int32_t *types, int32_t N,
                    int32_t frames, const double *box, double rc, double width) {
    if (!pos || !types || !box || N < 2 || frames < 1 || frames > 65535 || !(rc > 0.) ||
        !(width > 0.)) {
        set_error("%s: bad arguments", who);
        return false;
    }
    if (!(2.0 * rc <= box[0] && 2.0 * rc <= box[1] && 2.0 * rc <= box[2])) {
        set_error("%s: the cut-off exceeds half the box", who);
        return false;
    }
    return true;
}

static SpBasis sp_basis(double rc, double mu0, double dmu, double width) {
    SpBasis b;
    b.mu0 = mu0;
    b.dmu = dmu;
    b.iw2 = 1.0 / (width * width);
    b.pirc = 3.14159265358979323846 / rc;
    b.rc2 = rc * rc;
    b.q = std::exp(-dmu * dmu * b.iw2);
    return b;
}

// The Gaussians may come from anchors and ratios when, for any r in [0, rc],
// no anchor underflows and no ratio overflows: |r - mu_k0| <= max(rc - mu0,
// mu0 + 24 dmu) over the anchors k0 = 0, 8, 16, 24 (e^-700 is a normal double).
static bool sp_recur(const SpBasis &b, double rc) {
    double const u = std::max(std::fabs(rc - b.mu0), std::fabs(b.mu0 + 24.0 * b.dmu));
    return b.dmu > 0.0 && 0.5 * u * u * b.iw2 < 700.0 && u * b.dmu * b.iw2 < 700.0;
}

mdl_status mdl_cv_stress_basis_sums(const double *pos, const int32_t *types, int32_t N,
                                    int32_t frames, const double *box, double rc, double mu0,
                                    double dmu, double width, const int32_t *pairs,
                                    const int32_t *pair_count, int32_t pair_width, double *part,
                                    int32_t nblocks, void *stream) {
    if (!sp_args("mdl_cv_stress_basis_sums", pos, types, N, frames, box, rc, width))
        return MDL_ERR_ARG;
    if (!part || nblocks < 1 || nblocks > 65535 || (pairs && (!pair_count || pair_width < 1))) {
        set_error("mdl_cv_stress_basis_sums: bad arguments");
        return MDL_ERR_ARG;
    }
    k_cv_stress_basis<<<dim3(nblocks, frames), 256, 0, (hipStream_t)stream>>>(
        pos, types, N, box[0], box[1], box[2], sp_basis(rc, mu0, dmu, width), pairs, pair_count,
        pair_width, part);
    return launch_ok("mdl_cv_stress_basis_sums");
}

mdl_status mdl_cv_stress_proxy(const double *pos, const int32_t *types, int32_t N,
                               int32_t frames, const double *box, double rc, double mu0,
                               double dmu, double width, const double *coef, double *grad,
                               double *hl, void *stream) {
    if (!sp_args("mdl_cv_stress_proxy", pos, types, N, frames, box, rc, width))
        return MDL_ERR_ARG;
    if (!coef || !grad || !hl) {
        set_error("mdl_cv_stress_proxy: bad arguments");
        return MDL_ERR_ARG;
    }
    SpBasis const b = sp_basis(rc, mu0, dmu, width);
    if (sp_recur(b, rc))
        k_cv_stress_proxy<true><<<dim3((N + 3) / 4, frames), 256, 0, (hipStream_t)stream>>>(
            pos, types, N, box[0], box[1], box[2], b, coef, grad, hl);
    else
        k_cv_stress_proxy<false><<<dim3((N + 3) / 4, frames), 256, 0, (hipStream_t)stream>>>(
            pos, types, N, box[0], box[1], box[2], b, coef, grad, hl);
    return launch_ok("mdl_cv_stress_proxy");
}

mdl_status mdl_cv_stress_fields(const double *pos, const double *F, double fscale,
                                const int32_t *types, int32_t N, int32_t frames,
                                const double *box, double rc, double mu0, double dmu,
                                double width, const double *coef, const int32_t *pairs,
                                const int32_t *pair_count, int32_t pair_width, double *part,
                                int32_t nblocks, void *stream) {
    if (!sp_args("mdl_cv_stress_fields", pos, types, N, frames, box, rc, width))
        return MDL_ERR_ARG;
    if (!F || !coef || !part || nblocks < 1 || nblocks > 65535 ||
        (pairs && (!pair_count || pair_width < 1))) {
        set_error("mdl_cv_stress_fields: bad arguments");
        return MDL_ERR_ARG;
    }
    SpBasis const b = sp_basis(rc, mu0, dmu, width);
    if (sp_recur(b, rc))
        k_cv_stress_fields<true><<<dim3(nblocks, frames), 256, 0, (hipStream_t)stream>>>(
            pos, F, fscale, types, N, box[0], box[1], box[2], b, coef, pairs, pair_count,
            pair_width, part);
    else
        k_cv_stress_fields<false><<<dim3(nblocks, frames), 256, 0, (hipStream_t)stream>>>(
            pos, F, fscale, types, N, box[0], box[1], box[2], b, coef, pairs, pair_count,
            pair_width, part);
    return launch_ok("mdl_cv_stress_fields");
}

mdl_status mdl_cv_gemm(const float *A, const float *W, int32_t transB, const float *bias,
                       float *C, int64_t R, int32_t K, int32_t N, void *stream) {
    if (!A || !W || !C || R < 0 || K < 16 || N < 16 || K % 16 || N % 16 || R > (1ll << 31) - 16) {
        set_error("mdl_cv_gemm: K and N must be positive multiples of 16");
        return MDL_ERR_ARG;
    }
    if (R == 0)
        return MDL_OK;
    hipStream_t const st = (hipStream_t)stream;
    // staged B: N (K + 4) bf16 within 64 KB of LDS
    if ((int64_t)N * (K + 4) <= 32768 && (K == 16 || K == 32 || K == 64 || K == 128 ||
                                           K == 256 || K == 384)) {
        unsigned const blocks = (unsigned)std::min<int64_t>((R + 63) / 64, 512);
        size_t const lds = (size_t)N * (K + 4) * sizeof(short);
        switch (K / 16) {
        case 1: k_gemm_lds<1><<<blocks, 256, lds, st>>>(A, W, transB, bias, C, (int)R, N); break;
        case 2: k_gemm_lds<2><<<blocks, 256, lds, st>>>(A, W, transB, bias, C, (int)R, N); break;
        case 4: k_gemm_lds<4><<<blocks, 256, lds, st>>>(A, W, transB, bias, C, (int)R, N); break;
        case 8: k_gemm_lds<8><<<blocks, 256, lds, st>>>(A, W, transB, bias, C, (int)R, N); break;
        case 16: k_gemm_lds<16><<<blocks, 256, lds, st>>>(A, W, transB, bias, C, (int)R, N); break;
        default: k_gemm_lds<24><<<blocks, 256, lds, st>>>(A, W, transB, bias, C, (int)R, N); break;
        }
        return launch_ok("mdl_cv_gemm");
    }
    int64_t const tiles = ((R + 15) / 16) * (N / 16);
    int64_t const blocks = (tiles * 64 + 255) / 256;
    k_gemm_nn<<<(unsigned)blocks, 256, 0, st>>>(A, W, transB, bias, C, (int)R, K, N);
    return launch_ok("mdl_cv_gemm");
}

mdl_status mdl_cv_gemm_tn(const float *A, const float *B, float *C, float *work, int64_t R,
                          int32_t K, int32_t N, int32_t rows_per_chunk, void *stream) {
    if (!A || !B || !C || !work || R < 1 || K % 16 || N % 16 || K < 16 || N < 16 ||
        rows_per_chunk < 16 || rows_per_chunk % 16) {
        set_error("mdl_cv_gemm_tn: bad arguments");
        return MDL_ERR_ARG;
    }
    int const nchunk = (int)((R + rows_per_chunk - 1) / rows_per_chunk);
    int64_t const waves = (int64_t)nchunk * (K / 16) * (N / 16);
    k_gemm_tn<<<(unsigned)((waves * 64 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        A, B, work, (int)R, K, N, rows_per_chunk, nchunk);
    if (mdl_status s = launch_ok("mdl_cv_gemm_tn"))
        return s;
    int const n = K * N;
    k_sum_chunks<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(work, C, nchunk, n);
    return launch_ok("mdl_cv_gemm_tn reduce");
}

mdl_status mdl_cv_colsum(const float *X, float *out, float *work, int64_t R, int32_t N,
                         void *stream) {
    if (!X || !out || !work || R < 1 || R > (1ll << 31) - 1 || N < 1) {
        set_error("mdl_cv_colsum: bad arguments");
        return MDL_ERR_ARG;
    }
    int const nchunk = (int)((R + CS_ROWS - 1) / CS_ROWS);
    k_colsum_part<<<dim3(nchunk, (N + 63) / 64), 256, 0, (hipStream_t)stream>>>(X, work, (int)R,
                                                                               N);
    if (mdl_status s = launch_ok("mdl_cv_colsum"))
        return s;
    k_sum_chunks<<<(N + 255) / 256, 256, 0, (hipStream_t)stream>>>(work, out, nchunk, N);
    return launch_ok("mdl_cv_colsum reduce");
}

mdl_status mdl_cv_act_forward(const float *Z, const float *Zt, float *Y, float *Yt, int64_t MN,
                              int32_t act, void *stream) {
    if (!Z || !Zt || !Y || !Yt || MN < 0 || (act != 0 && act != 1)) {
        set_error("mdl_cv_act_forward: bad arguments");
        return MDL_ERR_ARG;
    }
    if (MN == 0)
        return MDL_OK;
    k_act_fwd<<<(unsigned)((MN + 255) / 256), 256, 0, (hipStream_t)stream>>>(Z, Zt, Y, Yt, MN,
                                                                             act);
    return launch_ok("mdl_cv_act_forward");
}

mdl_status mdl_cv_act_backward(const float *Z, const float *Zt, const float *gY, const float *gYt,
                               float *gZ, float *gZt, int64_t MN, int32_t act, void *stream) {
    if (!Z || !Zt || !gY || !gYt || !gZ || !gZt || MN < 0 || (act != 0 && act != 1)) {
        set_error("mdl_cv_act_backward: bad arguments");
        return MDL_ERR_ARG;
    }
    if (MN == 0)
        return MDL_OK;
    k_act_bwd<<<(unsigned)((MN + 255) / 256), 256, 0, (hipStream_t)stream>>>(Z, Zt, gY, gYt, gZ,
                                                                             gZt, MN, act);
    return launch_ok("mdl_cv_act_backward");
}

mdl_status mdl_cv_field_contract(const float *w, const float *wt, const float *J, const float *L,
                                 const float *F, const float *G, float beta, int64_t M,
                                 int32_t nfield, float *a, float *cF, float *cG, void *stream) {
    if (!w || !wt || !J || !L || !F || !G || !a || !cF || !cG || M < 0 ||
        M > (1ll << 31) / 64 - 1 || nfield < 1) {
        set_error("mdl_cv_field_contract: bad arguments");
        return MDL_ERR_ARG;
    }
    if (M == 0)
        return MDL_OK;
    k_field_contract<<<(unsigned)((M * 64 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        w, wt, J, L, F, G, beta, (int)M, nfield, a, cF, cG);
    return launch_ok("mdl_cv_field_contract");
}

mdl_status mdl_cv_field_contract_backward(const float *J, const float *L, const float *F,
                                          const float *G, float beta, const float *ga,
                                          const float *gcF, const float *gcG, int64_t M,
                                          int32_t nfield, float *gw, float *gwt, void *stream) {
    if (!J || !L || !F || !G || !ga || !gcF || !gcG || !gw || !gwt || M < 0 ||
        M > (1ll << 31) / 64 - 1 || nfield < 1) {
        set_error("mdl_cv_field_contract_backward: bad arguments");
        return MDL_ERR_ARG;
    }
    if (M == 0)
        return MDL_OK;
    k_field_contract_bwd<<<(unsigned)((M * 64 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        J, L, F, G, beta, ga, gcF, gcG, (int)M, nfield, gw, gwt);
    return launch_ok("mdl_cv_field_contract_backward");
}

}  // extern "C"
