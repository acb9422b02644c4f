/*
 * mdelast-mi355x: device kernels of the learned Stein control variate
 * (libmdelast_hip.so). No counterpart in the reference: the control variate
 * is new work on top of its stress-fluctuation estimator
 * (molecular_dynamics.py:825-1087), whose plain form it must not change in
 * expectation: E_NVT[div g - beta g . grad U] = 0 for any smooth periodic g.
 *
 * All pointers are device pointers (float32 unless stated); `stream` is a
 * hipStream_t (NULL = the null stream). Calls are asynchronous; errors of the
 * launch are reported, and mdl_last_error() (mdelast.h) holds the message.
 */
#ifndef MDELAST_CV_H_
#define MDELAST_CV_H_

#include "mdelast.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Radial-basis descriptors of N atoms of two species (types 0/1), their
 * derivatives along each atom's own coordinates and their Laplacians there:
 *   D[i][t*32 + k]    = sum_{j of type t, r_ij < rc} phi_k(r_ij)
 *   J[a][i][t*32 + k] = d D[i][t*32 + k] / d x_{i,a}
 *   L[i][t*32 + k]    = sum_a d^2 D[i][t*32 + k] / d x_{i,a}^2
 * phi_k(r) = exp(-(r - mu0 - k dmu)^2 / (2 width^2)) (cos(pi r / rc) + 1) / 2,
 * k = 0..31. pos: N*3 FP64 in A (minimum image in the box), box: 3 FP64 on the
 * HOST, D: [N][64], J: [3][N][64], L: [N][64] or NULL. MDL_ERR_ARG unless
 * 2 rc <= each box edge. Brute force over all pairs. */
mdl_status mdl_cv_descriptors(const double *pos, const int32_t *types, int32_t N,
                              const double *box, float rc, float mu0, float dmu, float width,
                              float *D, float *J, float *L, void *stream);

/* The same descriptors of the driver's current configuration, from its
 * neighbour list (lengths in A; the driver's species index is the type),
 * unless NULL the Laplacian's axis terms La[a][i][q] = d^2 D[i][q] / d x_{i,a}^2
 * ([3][N][64]), and, unless NULL, its positions and forces (device [N][3] FP64, the layout and
 * CGS units of mdl_driver_get_atoms; positions need a step taken). rc must not
 * exceed the list's guaranteed cut-off and the driver must have at most two
 * species. With a stream (hipStream_t) the kernels run there, ordered after
 * the driver's queued work, and the driver's later work waits for them (no
 * host synchronisation); with NULL they run on the driver's stream and the
 * call returns when they are done. */
mdl_status mdl_driver_cv_descriptors(mdl_driver *d, float rc, float mu0, float dmu, float width,
                                     float *D, float *J, float *L, float *La, double *pos_out,
                                     double *force_out, void *stream);

/* The same inputs, taken inside mdl_driver_run / mdl_drivers_run at every
 * sampled step (on the driver's stream, right after the step) into `slots`
 * frame slots used in turn from slot 0: D [slots][N][64], J [slots][3][N][64],
 * L [slots][N][64] (FP32), pos_out and force_out [slots][N][3] (FP64, CGS; all
 * device memory that must outlive the runs). A call restarts at slot 0;
 * slots = 0 stops the capture. While it is on, steps launch one by one (no
 * captured graphs) and mdl_replicas_run refuses the driver. Work queued on the
 * driver's stream (mdl_driver_stream) after a run reads the slots in order. */
mdl_status mdl_driver_set_cv_capture(mdl_driver *d, float rc, float mu0, float dmu, float width,
                                     float *D, float *J, float *L, double *pos_out,
                                     double *force_out, int32_t slots);

/* With a capture set, also the candidate pairs of each captured step: nbr_out
 * [slots][N][width] and count_out [slots][N] (int32 device) receive, per atom,
 * a list in the driver's order that holds all its pairs within the capture's
 * cut-off (FP64 drivers: the compacted pairs within r_c; mixed drivers: the mid
 * list, pairs within r_c + 0.15 A; either: the full list when rc exceeds r_c),
 * and its length; readers test r. A count above width means the row was cut.
 * NULL stops it; a new mdl_driver_set_cv_capture clears it. No counterpart in
 * the reference. */
mdl_status mdl_driver_set_cv_capture_pairs(mdl_driver *d, int32_t *nbr_out, int32_t *count_out,
                                           int32_t width);

/* The driver's stream (a hipStream_t) into *out. */
mdl_status mdl_driver_stream(mdl_driver *d, void **out);

/* Device-to-device copy of a per-atom array of the driver (MDL_A_*, the
 * layout and units of mdl_driver_get_atoms) into device memory `out`;
 * returns when done. */
mdl_status mdl_driver_copy_atoms(mdl_driver *d, int32_t what, double *out);

/* C[R][N] = A[R][K] B + bias on MFMA (bf16 operands rounded to nearest even,
 * fp32 accumulation); B = W[K][N], or W^T with W[N][K] when transB != 0.
 * K and N positive multiples of 16; bias may be NULL. */
mdl_status mdl_cv_gemm(const float *A, const float *W, int32_t transB, const float *bias,
                       float *C, int64_t R, int32_t K, int32_t N, void *stream);

/* C[K][N] = sum_r A[r][K]^T B[r][N] (bf16 MFMA, fp32), summed per chunk of
 * rows_per_chunk rows (a multiple of 16) into work[ceil(R/rows)][K][N] and then
 * over chunks in a fixed order. */
mdl_status mdl_cv_gemm_tn(const float *A, const float *B, float *C, float *work, int64_t R,
                          int32_t K, int32_t N, int32_t rows_per_chunk, void *stream);

/* out[n] = sum_r X[r][n] in a fixed order (per chunk of 1024 rows into
 * work[ceil(R/1024)][N], then over chunks). */
mdl_status mdl_cv_colsum(const float *X, float *out, float *work, int64_t R, int32_t N,
                         void *stream);

/* Activation with the three tangents: Y = act(Z), Yt[a] = act'(Z) Zt[a]
 * (Z, Y: MN floats; Zt, Yt: [3][MN]). act 0 = tanh, 1 = identity. */
mdl_status mdl_cv_act_forward(const float *Z, const float *Zt, float *Y, float *Yt, int64_t MN,
                              int32_t act, void *stream);
/* Its backward: gZ = gY act'(Z) + sum_a gYt[a] Zt[a] act''(Z), gZt[a] = gYt[a] act'(Z). */
mdl_status mdl_cv_act_backward(const float *Z, const float *Zt, const float *gY, const float *gYt,
                               float *gZ, float *gZt, int64_t MN, int32_t act, void *stream);

/* Fields from their weights, per atom m and field f (w: [M][nfield][64],
 * wt: [3][M][nfield][64], J: [3][M][64], L: [M][64], F and G: [M][3]):
 *   c = sum_k w_k J[:][m][k],  div = sum_k (w_k L[m][k] + sum_a wt_a,k J[a][m][k]),
 *   a[m][f] = div + beta c . F[m],  cF[m][f] = c . F[m],  cG[m][f] = c . G[m]. */
mdl_status mdl_cv_field_contract(const float *w, const float *wt, const float *J, const float *L,
                                 const float *F, const float *G, float beta, int64_t M,
                                 int32_t nfield, float *a, float *cF, float *cG, void *stream);
/* Its backward with respect to w and wt, given ga, gcF, gcG ([M][nfield]). */
mdl_status mdl_cv_field_contract_backward(const float *J, const float *L, const float *F,
                                          const float *G, float beta, const float *ga,
                                          const float *gcF, const float *gcG, int64_t M,
                                          int32_t nfield, float *gw, float *gwt, void *stream);

/* Stress proxy of the zero-variance fields (no counterpart in the reference).
 * For `frames` frames of N atoms of two species (types 0/1; the same types for
 * every frame), positions pos [frames][N][3] FP64 in A, box: 3 FP64 on the
 * HOST (2 rc <= each edge), FP64 throughout. The pairs are found among all
 * atoms, or, for mdl_cv_stress_basis_sums and mdl_cv_stress_fields with
 * `pairs` non-NULL, taken from per-frame pair lists as
 * mdl_driver_set_cv_capture_pairs writes them (pairs [frames][N][pair_width],
 * pair_count [frames][N], int32 device; each pair re-checked against rc; a
 * row's first min(count, pair_width) entries):
 *
 * mdl_cv_stress_basis_sums: part[frame][block][p][k][v] (blocks = nblocks) =
 *   the share of block `block` of Z[p][k][v] = sum over ordered pairs (i, j),
 *   r_ij < rc, t_i + t_j = p, of fc(r) g_k(r) n_a n_b, with n = (x_i - x_j)/r,
 *   Voigt v = (a, b) in the order xx, yy, zz, yz, xz, xy,
 *   g_k(r) = exp(-(r - mu0 - k dmu)^2 / (2 width^2)), fc = (cos(pi r / rc) + 1)/2,
 *   k = 0..31. Summing over blocks gives Z (576 values per frame).
 *
 * mdl_cv_stress_proxy: with psi_p(r) = fc(r) sum_k coef[p][k] g_k(r) (coef:
 *   [3][32] FP64 device), the proxy h_v = sum over ordered pairs of
 *   psi_{t_i+t_j}(r) n_a n_b has per atom i: hl[frame][i][0..5] = its share
 *   sum_j psi(r_ij) n_a n_b, hl[frame][i][6..11] = the Laplacian of h_v along
 *   atom i's coordinates, grad[frame][i][v][c] = d h_v / d x_{i,c}. */
mdl_status mdl_cv_stress_basis_sums(const double *pos, const int32_t *types, int32_t N,
                                    int32_t frames, const double *box, double rc, double mu0,
                                    double dmu, double width, const int32_t *pairs,
                                    const int32_t *pair_count, int32_t pair_width, double *part,
                                    int32_t nblocks, void *stream);
mdl_status mdl_cv_stress_proxy(const double *pos, const int32_t *types, int32_t N,
                               int32_t frames, const double *box, double rc, double mu0,
                               double dmu, double width, const double *coef, double *grad,
                               double *hl, void *stream);

/* mdl_cv_stress_fields: the per-frame sums of the proxy's zero-variance fields
 * without the per-atom arrays, over the frame's atoms m: part[frame][block][q]
 * (blocks = nblocks; summing over blocks gives the frame's 39 values) with
 * q < 6 the share of h_v, 6 <= q < 12 the Laplacian lap_m h_v, 12 <= q < 18
 * grad_m h_v . (fscale F_m) (F [frames][N][3] FP64 device), and 18 + the index
 * of (a, b), a <= b, in row order of grad_m h_a . grad_m h_b. */
mdl_status mdl_cv_stress_fields(const double *pos, const double *F, double fscale,
                                const int32_t *types, int32_t N, int32_t frames,
                                const double *box, double rc, double mu0, double dmu,
                                double width, const double *coef, const int32_t *pairs,
                                const int32_t *pair_count, int32_t pair_width, double *part,
                                int32_t nblocks, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MDELAST_CV_H_ */
