// Radial basis of the control variate's descriptors (device side), shared by
// the stand-alone descriptor kernel (mdl_cv.hip) and the one that reads the
// driver's neighbour list (mdl_driver.hip).
//
//   phi_k(r) = exp(-(r - mu_k)^2 / (2 w^2)) fc(r),  fc(r) = (cos(pi r / rc) + 1) / 2
//
// For one pair at distance r and separation vector d = x_i - x_j, atom i's
// descriptor D_ik gains phi_k(r); its derivative along x_{i,a} gains
// phi_k'(r) d_a / r; and the Laplacian sum_a d^2 D_ik / d x_{i,a}^2 gains
// phi_k''(r) + 2 phi_k'(r) / r (the Laplacian of a radial function in 3-D),
// of which axis a contributes (phi'' - phi'/r)(d_a/r)^2 + phi'/r.
#pragma once

#include <hip/hip_runtime.h>

namespace mdl {

constexpr int CV_NB = 32;   // radial functions per neighbour species

struct CvBasis {
    float mu0, dmu, iw2, pirc, rc2;
};

__host__ __device__ inline CvBasis cv_basis(float rc, float mu0, float dmu, float width) {
    CvBasis b;
    b.mu0 = mu0;
    b.dmu = dmu;
    b.iw2 = 1.f / (width * width);
    b.pirc = 3.14159265358979f / rc;
    b.rc2 = rc * rc;
    return b;
}

// phi_k, phi_k' / r and phi_k'' of basis function k at r
__device__ __forceinline__ void cv_phi(const CvBasis &b, int k, float r, float &phi, float &w,
                                       float &d2) {
    float const u = r - (b.mu0 + b.dmu * (float)k);
    float const g = __expf(-0.5f * u * u * b.iw2);
    float sn, c;
    __sincosf(b.pirc * r, &sn, &c);   // argument in [0, pi]
    float const fc = 0.5f * (c + 1.f);
    float const dfc = -0.5f * b.pirc * sn;
    float const d2fc = -0.5f * b.pirc * b.pirc * c;
    float const ui = u * b.iw2;
    phi = g * fc;
    float const d1 = g * (dfc - ui * fc);
    d2 = g * ((ui * ui - b.iw2) * fc - 2.f * ui * dfc + d2fc);
    w = d1 * __builtin_amdgcn_rcpf(r);
}

// One half-wave (32 lanes, lane = basis index) accumulates the pairs of one
// atom; store its 64 descriptors, 3 x 64 derivatives, 64 Laplacians and, with
// AXES, the Laplacian's three axis terms d^2 D / d x_a^2 =
// (phi'' - phi'/r)(d_a/r)^2 + phi'/r (their sum is phi'' + 2 phi'/r, which is
// all that is accumulated without AXES).
template <bool AXES = false> struct CvAcc {
    float d[2], jx[2], jy[2], jz[2], lx[2], ly[2], lz[2];

    __device__ __forceinline__ CvAcc() {
#pragma unroll
        for (int s = 0; s < 2; ++s)
            d[s] = jx[s] = jy[s] = jz[s] = lx[s] = ly[s] = lz[s] = 0.f;
    }

    __device__ __forceinline__ void add(const CvBasis &b, int k, int species, float x, float y,
                                        float z, float r) {
        float phi, w, d2;
        cv_phi(b, k, r, phi, w, d2);
        int const s = species != 0;
        // two accumulators selected without dynamic register indexing
        float const m0 = s ? 0.f : 1.f, m1 = 1.f - m0;
        d[0] += m0 * phi;
        d[1] += m1 * phi;
        jx[0] += m0 * w * x;
        jx[1] += m1 * w * x;
        jy[0] += m0 * w * y;
        jy[1] += m1 * w * y;
        jz[0] += m0 * w * z;
        jz[1] += m1 * w * z;
        if (AXES) {
            float const ir2 = __builtin_amdgcn_rcpf(r * r);
            float const q = (d2 - w) * ir2;
            float const ax = fmaf(q, x * x, w), ay = fmaf(q, y * y, w), az = fmaf(q, z * z, w);
            lx[0] += m0 * ax;
            lx[1] += m1 * ax;
            ly[0] += m0 * ay;
            ly[1] += m1 * ay;
            lz[0] += m0 * az;
            lz[1] += m1 * az;
        } else {
            float const l = fmaf(2.f, w, d2);   // phi'' + 2 phi'/r
            lx[0] += m0 * l;
            lx[1] += m1 * l;
        }
    }

    __device__ __forceinline__ void store(int i, int k, int N, float *D, float *J, float *L,
                                          float *La = nullptr) const {
        int const F = 2 * CV_NB;
        size_t const plane = (size_t)N * F;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            size_t const o = (size_t)i * F + s * CV_NB + k;
            D[o] = d[s];
            J[o] = jx[s];
            J[plane + o] = jy[s];
            J[2 * plane + o] = jz[s];
            if (L)
                L[o] = AXES ? lx[s] + ly[s] + lz[s] : lx[s];
            if (AXES && La) {
                La[o] = lx[s];
                La[plane + o] = ly[s];
                La[2 * plane + o] = lz[s];
            }
        }
    }
};

}  // namespace mdl
