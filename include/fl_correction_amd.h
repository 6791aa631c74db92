/*
 * fl_correction_amd.h -- C ABI of the MI355X-native XRF self-absorption hot path.
 *
 * Plain pointers and sizes, no framework types.  Every array argument is a
 * device pointer to contiguous row-major float32 data on the current HIP
 * device unless stated otherwise; `stream` is a hipStream_t (NULL = the
 * default stream).  All calls are asynchronous with respect to the host and
 * ordered on `stream`.  Return value: FL_OK or an FL_ERR_* code, with a
 * message from fl_last_error().  Outputs are written, never accumulated into,
 * unless stated otherwise.  Index arithmetic is 64-bit throughout.
 *
 * Shape letters: A angles, B slices, H rows (beam axis), W columns (detector
 * bins), R references summed into one emission line, K elements.
 *
 * Citations name the reference interface each entry point replaces
 * (paths relative to the gmysage/FL_correction root).
 */
#ifndef FL_CORRECTION_AMD_H
#define FL_CORRECTION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FL_OK 0
#define FL_ERR_INVALID_ARG 1
#define FL_ERR_HIP 2
#define FL_ERR_UNSUPPORTED 3
#define FL_ERR_WORKSPACE 4

#if defined(__GNUC__)
#define FL_API __attribute__((visibility("default")))
#else
#define FL_API
#endif

/* Message for the last failing call on this host thread ("" if none). */
FL_API const char *fl_last_error(void);
/* ABI version, (major << 16) | minor. */
FL_API int fl_version(void);

/* ---------------------------------------------------------------------------
 * Emission projector and its transpose.
 * ------------------------------------------------------------------------- */

/* Pf[a,b,w] = sum_h atten[a,b,h,w] * sum_r em_cs[a,r] * bilin_clamp(C[r,b], p_a(h,w)).
 * Replaces forward_emission_batch (FL/cuda_lib/atten_cuda.cpp:142-182;
 * kernel FL/cuda_lib/atten_cuda_kernel.cu:298-337).  atten (A,B,H,W),
 * C (R,B,H,W), em_cs (A,R), theta (A) radians, Pf (A,B,W).  Deterministic:
 * every column sum is formed in a fixed order, without atomics. */
FL_API int fl_forward_emission_batch(const float *atten, const float *C, const float *em_cs,
                                     const float *theta, float *Pf, int64_t A, int64_t B, int64_t R,
                                     int64_t H, int64_t W, void *stream);

/* out[r,b] = sum_a em_cs[a,r] * S_a^T( atten[a,b] * ratio[a,b,:] ), S_a^T the exact
 * transpose of the bilinear sampling above.  Replaces backward_emission_batch
 * (atten_cuda.cpp:188-214; kernel atten_cuda_kernel.cu:390-446).  ratio (A,B,W),
 * out (R,B,H,W).  ratio == NULL means ratio == 1 (the MLEM sensitivity).
 * Gather-by-tile: no global atomics; per-tile sums in a fixed order. */
FL_API int fl_backward_emission_batch(const float *atten, const float *ratio, const float *em_cs,
                               const float *theta, float *out, int64_t A, int64_t B, int64_t R,
                               int64_t H, int64_t W, void *stream);

/* Single-slice forms: the batch forms with B = 1.  Replace forward_emission
 * (atten_cuda.cpp:82-108) and backward_emission (atten_cuda.cpp:113-135):
 * atten (A,H,W), C (R,H,W), Pf (A,W), ratio (A,W), out (R,H,W). */
FL_API int fl_forward_emission(const float *atten, const float *C, const float *em_cs,
                               const float *theta, float *Pf, int64_t A, int64_t R, int64_t H,
                               int64_t W, void *stream);
FL_API int fl_backward_emission(const float *atten, const float *ratio, const float *em_cs,
                         const float *theta, float *out, int64_t A, int64_t R, int64_t H,
                         int64_t W, void *stream);


/* ---------------------------------------------------------------------------
 * MLEM / One-Step-Late Huber MAP-EM.
 * ------------------------------------------------------------------------- */

/* 4-neighbour Huber gradient in (H,W): grad = sum_nb clip(C - C_nb, -delta, delta).
 * The reference keeps it internal: huber_grad_batch (atten_cuda.cpp:229-238;
 * kernel atten_cuda_kernel.cu:564-616).  C, grad (R,B,H,W). */
FL_API int fl_huber_grad_batch(const float *C, float *grad, int64_t R, int64_t B, int64_t H,
                        int64_t W, float delta, void *stream);

/* C_out = n_iter iterations of mlem_cuda_batch (atten_cuda.cpp:283-345):
 *   sens = max(B^T 1, 1e-6);  Pf = max(F C, 1e-6);  ratio = I/Pf;  back = B^T ratio;
 *   C = max(0, C*back / (sens + beta*huber(C)))   (beta > 0)   or   C*back/sens.
 * C_init (R,B,H,W) is not modified; I (A,B,W).  `workspace` must hold
 * fl_mlem_workspace_bytes(...) bytes; the call allocates nothing.
 * Replaces mlem_cuda_batch and, with B = 1, mlem_cuda (atten_cuda.cpp:244-281). */
FL_API int fl_mlem_batch(const float *C_init, const float *atten, const float *em_cs,
                  const float *theta, const float *I, float *C_out, int64_t A, int64_t B,
                  int64_t R, int64_t H, int64_t W, int64_t n_iter, double beta, double delta,
                  void *workspace, size_t workspace_bytes, void *stream);
FL_API size_t fl_mlem_workspace_bytes(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W);

/* fl_mlem_batch with no workspace argument and no allocation of any kind: the scratch is
 * a static device array of FL_MLEM_NOALLOC_SCRATCH_BYTES in the library (each GPU the
 * library runs on holds one until the process ends).  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Problems whose
 * fl_mlem_workspace_bytes exceeds the array fail with FL_ERR_WORKSPACE.  All calls on one
 * GPU share that GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race and must be serialised by the caller.  The stream must belong to the
 * current device. */
#define FL_MLEM_NOALLOC_SCRATCH_BYTES (64u << 20)
FL_API int fl_mlem_batch_noalloc(const float *C_init, const float *atten, const float *em_cs,
                                 const float *theta, const float *I, float *C_out, int64_t A,
                                 int64_t B, int64_t R, int64_t H, int64_t W, int64_t n_iter,
                                 double beta, double delta, void *stream);

/* ---------------------------------------------------------------------------
 * Slice-major path.  The back-projection reads the attenuation as a slice-major
 * copy atten_sm (A, H, W, B): the B slices of one detector point are contiguous,
 * so one sample position and its four weights serve all of them at once.  The
 * results equal the reference-layout back-projector's to rounding (the sums run
 * in another fixed order).  One reference per line (R = 1; em_cs (A)); ratio_sm
 * is the MLEM ratio laid out (A, W, B); H*W*B < 2^29.  The back-projector needs
 * fl_backproject_sm_scratch_bytes of device scratch (per-block partial sums).
 * No reference counterpart for the layout: these compute backward_emission_batch
 * (FL/cuda_lib/atten_cuda.cpp:188-214) and the MLEM stages of mlem_cuda_batch
 * (atten_cuda.cpp:308-341) on it.
 * ------------------------------------------------------------------------- */

/* dst[n, l, b] = src[n, b, l] for n < N, b < B, l < L: atten (A,B,H,W) -> atten_sm
 * with N = A, L = H*W. */
FL_API int fl_slice_major(const float *src, float *dst, int64_t N, int64_t B, int64_t L, void *stream);
FL_API size_t fl_backproject_sm_scratch_bytes(int64_t A, int64_t B, int64_t H, int64_t W);
/* out (B,H,W) (=, or += with accumulate) = B^T ratio on atten_sm; ratio_sm NULL means 1. */
FL_API int fl_backward_emission_sm(const float *atten_sm, const float *ratio_sm, const float *em_cs,
                                   const float *theta, float *out, int64_t A, int64_t B, int64_t H,
                                   int64_t W, int accumulate, void *scratch, size_t scratch_bytes,
                                   void *stream);
/* ratio_sm (A,W,B) = I / max(F C, 1e-6): the forward on atten (A,B,H,W), C (B,H,W), I (A,B,W). */
FL_API int fl_mlem_ratio_sm(const float *atten, const float *C, const float *em_cs,
                            const float *theta, const float *I, float *ratio_sm, int64_t A, int64_t B,
                            int64_t H, int64_t W, void *stream);
/* sens (B,H,W) = max(B^T 1, 1e-6) on atten_sm. */
FL_API int fl_mlem_sensitivity_sm(const float *atten_sm, const float *em_cs, const float *theta,
                                  float *sens, int64_t A, int64_t B, int64_t H, int64_t W,
                                  void *scratch, size_t scratch_bytes, void *stream);
/* C_next = max(0, C*(B^T ratio)/(sens + beta*huber(C))), or C*(B^T ratio)/sens when
 * beta <= 0, on atten_sm and ratio_sm; C, sens, C_next (B,H,W); C_next must not alias C. */
FL_API int fl_mlem_update_sm(const float *atten_sm, const float *ratio_sm, const float *em_cs,
                             const float *theta, const float *C, const float *sens, float *C_next,
                             int64_t A, int64_t B, int64_t H, int64_t W, double beta, double delta,
                             void *scratch, size_t scratch_bytes, void *stream);
/* Workspace with which fl_mlem_batch takes the slice-major path (0 if it does not
 * apply: R != 1 or H*W*B too large).  fl_mlem_batch takes it when R = 1, B >= 32 and
 * workspace_bytes is at least this (it then makes the slice-major copy itself); with
 * fl_mlem_workspace_bytes it keeps the reference-layout back-projector. */
FL_API size_t fl_mlem_workspace_bytes_sm(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W);

/* The three stages of fl_mlem_batch, for callers that drive the iteration
 * themselves (and time its kernels):
 *   fl_mlem_sensitivity: sens = max(B^T 1, 1e-6)                  (atten_cuda.cpp:308-314)
 *   fl_mlem_ratio:       ratio = I / max(F C, 1e-6), ratio (A,B,W)  (atten_cuda.cpp:323-328)
 *   fl_mlem_update:      C_next = max(0, C*(B^T ratio)/(sens + beta*huber(C))), or
 *                        C*(B^T ratio)/sens when beta <= 0         (atten_cuda.cpp:330-341);
 *                        C_next must not alias C. */
FL_API int fl_mlem_sensitivity(const float *atten, const float *em_cs, const float *theta,
                               float *sens, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W,
                               void *stream);
FL_API int fl_mlem_ratio(const float *atten, const float *C, const float *em_cs,
                         const float *theta, const float *I, float *ratio, int64_t A, int64_t B,
                         int64_t R, int64_t H, int64_t W, void *stream);
FL_API int fl_mlem_update(const float *atten, const float *ratio, const float *em_cs,
                          const float *theta, const float *C, const float *sens, float *C_next,
                          int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, double beta,
                          double delta, void *stream);

/* ---------------------------------------------------------------------------
 * Batched MLEM: many independent problems of uneven size, each pass of an
 * iteration one persistent launch over all of them.
 *
 * The reference has no batched form: these calls replace a host loop of
 * mlem_cuda_batch calls (FL/cuda_lib/atten_cuda.cpp:283-345), one per sample
 * and emission line, each behind torch_mlem_recon_batch (FL/torch_recon.py:50-98).
 * The iteration is mlem_cuda_batch's, per problem, with R = 1.
 * ------------------------------------------------------------------------- */

/* One problem: one emission line of one sample.  atten (A,B,H,W), em_cs (A),
 * theta (A) radians, I (A,B,W), C (B,H,W): the starting estimate, overwritten
 * with the result by fl_mlem_multi_run.  Tile shapes (0 = default):
 * fwd_angles, angles per forward workgroup, 4 or 8 (default 8); bwd_rows, image
 * rows per back-projector tile, 15 or 31 (default 31).  The results do not
 * depend on fwd_angles; bwd_rows changes only the order of the per-voxel sums. */
typedef struct fl_mlem_problem {
    const float *atten;
    const float *em_cs;
    const float *theta;
    const float *I;
    float *C;
    int64_t A, B, H, W;
    int32_t fwd_angles;
    int32_t bwd_rows;
} fl_mlem_problem;

/* Host-side plan: job tables (in the workspace), persistent grid sizes and,
 * once run in graph mode, the instantiated hipGraph of two iterations. */
typedef struct fl_mlem_multi fl_mlem_multi;

/* Workspace for n problems: job tables, and per problem the sensitivity, a
 * second estimate (B,H,W) and the ratio (A,B,W).  0 for invalid arguments.
 * Host-only arithmetic. */
FL_API size_t fl_mlem_multi_workspace_bytes(const fl_mlem_problem *problems, int64_t n);

/* Workspace of the slice-major mode, which fl_mlem_multi_create takes when its
 * workspace is at least this large: problems with the same theta array (the
 * same pointer) and the same A, H, W -- the emission lines of one sample -- are
 * packed along the slice axis into one problem of sum(B) slices, and per pack
 * the workspace holds a slice-major copy of the attenuation (A, H, W, sum B),
 * the ratio (A, W, sum B), the sensitivity, two estimates and the
 * back-projector's scratch.  Slices are independent, so the results equal
 * each problem's own to rounding.  0 when the mode does not apply (a pack
 * with H*W*sum(B) >= 2^29 or A >= 65536) or for invalid arguments.  Host-only
 * arithmetic.  No counterpart in the reference, which runs one problem per call. */
FL_API size_t fl_mlem_multi_workspace_bytes_sm(const fl_mlem_problem *problems, int64_t n);

/* Validate the problems (shapes, tile shapes, disjoint C), build and upload the
 * job tables into `workspace` (device memory, 256-byte aligned, of at least
 * fl_mlem_multi_workspace_bytes bytes; with at least
 * fl_mlem_multi_workspace_bytes_sm bytes the slice-major mode) and enqueue the
 * sensitivity pass sens = max(B^T 1, 1e-6) on `stream` (in the slice-major
 * mode after the slice-major copies).  The problem arrays (not the struct
 * array) must stay valid while the plan is used.  Allocates no device memory. */
FL_API int fl_mlem_multi_create(const fl_mlem_problem *problems, int64_t n, double beta,
                                double delta, void *workspace, size_t workspace_bytes,
                                void *stream, fl_mlem_multi **plan);

/* n_iter MLEM (beta <= 0) or OSL-Huber (beta > 0) iterations of every problem,
 * continuing from its C and leaving the result in C.  use_graph != 0: pairs of
 * iterations are replays of one hipGraph, captured on `stream` (which must not
 * be the NULL stream) at the first such call; an odd last iteration runs
 * eagerly.  Calls on one plan are ordered on their streams; they must not
 * overlap. */
FL_API int fl_mlem_multi_run(fl_mlem_multi *plan, int64_t n_iter, int use_graph, void *stream);

/* Release the plan (host memory and its graph) once its work has finished.
 * The workspace stays the caller's. */
FL_API int fl_mlem_multi_destroy(fl_mlem_multi *plan);

/* ---------------------------------------------------------------------------
 * Attenuation maps.
 *
 * Frames: the original frame is (B, H, W) = (slice, row, col), incident beam
 * along rows toward row 0, detector at +col (side 0) or -col (side 1).  The
 * mask frame (Z, X, Y) is the frame atten_kernel works in: for side 0,
 * X = col and Y = H-1-row (torch_rot90_4D 'clock', FL/torch_functions.py:143);
 * for side 1, X = W-1-col and Y = row ('c-clock', :145).
 * ------------------------------------------------------------------------- */

/* out[k,z,x,y] = sum_{dz,dx,dy} mask[dz,dx,dy] * mu[k, z+dz-Mz/2, x+dx, y+dy-My/2],
 * taps outside the volume skipped.  Replaces cal_atten_cuda (FL/cuda_lib/atten_cuda.cpp:12-28;
 * kernel atten_cuda_kernel.cu:10-57).  mu, out (K,Z,X,Y); mask (Mz,Mx,My).  Any Y
 * (the reference needs Y <= 1024). */
FL_API int fl_cal_atten(const float *mu, const float *mask, float *out, int64_t K, int64_t Z,
                        int64_t X, int64_t Y, int64_t Mz, int64_t Mx, int64_t My, void *stream);

/* In-plane rotation of N planes (H,W): F.affine_grid([[c,-s,0],[s,c,0]]) then
 * F.grid_sample (bilinear, zero padding, align_corners=False).  Replaces
 * torch_rotate_image_nd with its defaults (FL/torch_image_util.py:13-114).
 * cos_t, sin_t: the float32 matrix entries (the reference rounds math.cos(theta)). */
FL_API int fl_rotate(const float *x, float *out, int64_t N, int64_t H, int64_t W, float cos_t,
                     float sin_t, void *stream);

/* Incident beam: out[e,b,H-1,c] = 1, out[e,b,r,c] = prod_{m=r+1}^{H-1} (1 - frac[e,b,m,c]*coef[e]).
 * Replaces torch_xray_atten_3D (FL/torch_functions.py:83-106) with mu_x from
 * torch_extract_3D_mu_xray (:8-52).  coef: device (K,) floats, or NULL when frac
 * already holds mu_x.  Element strides (in floats) let frac/out be slices of
 * larger arrays; B*H*W for dense (K,B,H,W). */
FL_API int fl_incident_atten(const float *frac, const float *coef, float *out, int64_t K, int64_t B,
                             int64_t H, int64_t W, int64_t frac_element_stride,
                             int64_t out_element_stride, void *stream);

/* mu at each fluorescence line, written in the mask frame:
 *   mu_rot[e] = maskframe( sum_j frac[j] * coef[j*K + e] ),  frac (Kc,B,H,W), coef device (Kc,K),
 *   mu_rot (K, B, W, H).  torch_extract_3D_mu_xrf (FL/torch_functions.py:55-81) followed by
 * torch_rot90_4D (:142-145). */
FL_API int fl_mu_xrf_rot(const float *frac, const float *coef, float *mu_rot, int64_t Kc,
                         int64_t K, int64_t B, int64_t H, int64_t W, int side, void *stream);

/* Direction-sampled detector cone in the mask frame (cone along +X):
 *   T[k,z,x,y] = sum_d weight[d] * sum_{t>=1, x+t<X} bilin0(mu[k,.,x+t,.], z+t*pz[d]/q[d], y+t*py[d]/q[d])
 * bilin0: bilinear with zero outside the volume.  Equals fl_cal_atten with the
 * mask that splats these rays (every tap of such a mask lies on an integer voxel).
 * Direction arrays are HOST memory; reduced denominators must be <= 32.
 * No reference counterpart: the reference only has the dense mask form. */
FL_API int fl_cone_atten(const float *mu_rot, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y,
                         const int32_t *pz, const int32_t *py, const int32_t *q,
                         const float *weight, int64_t ndir, void *stream);

/* out[e,b,r,c] = exp(-T_rot[e, b, maskframe(r,c)]) * (incident ? incident[e,b,r,c] : 1).
 * The tail of torch_xrf_atten_3D (rot90 back and exp, FL/torch_functions.py:170-174)
 * fused with the per-element product of cal_total_atten_at_angles (FL/core.py:99).
 * T_rot (K,B,W,H) dense; incident/out element strides in floats. */
FL_API int fl_xrf_combine(const float *T_rot, const float *incident, float *out, int64_t K,
                          int64_t B, int64_t H, int64_t W, int64_t inc_element_stride,
                          int64_t out_element_stride, int side, void *stream);

/* Transpose of fl_cal_atten: out = M^T (*) g, i.e. the correlation with the
 * flipped mask.  No reference counterpart (the reference differentiates nothing
 * through the fluorescence path); used by the adjoint of the attenuated model. */
FL_API int fl_cal_atten_adjoint(const float *g, const float *mask, float *out, int64_t K,
                                int64_t Z, int64_t X, int64_t Y, int64_t Mz, int64_t Mx,
                                int64_t My, void *stream);

/* Transpose of fl_cone_atten (the march run backwards with negated offsets).
 * No reference counterpart. */
FL_API int fl_cone_atten_adjoint(const float *g_rot, float *dmu_rot, int64_t K, int64_t Z,
                                 int64_t X, int64_t Y, const int32_t *pz, const int32_t *py,
                                 const int32_t *q, const float *weight, int64_t ndir,
                                 void *stream);

/* fl_cone_atten (adjoint = 0) or fl_cone_atten_adjoint (adjoint = 1) as one sweep
 * over blocks of Q planes for all directions at once: T is written once instead of
 * read and rewritten per direction.  Every direction's denominator q[d] must divide
 * one Q in {4, 8, 16} (Q = max q[d]); at most 64 directions; the results equal the
 * per-direction form to rounding (the ray sums are split every Q steps instead of
 * every reduced q[d]).  Returns FL_ERR_UNSUPPORTED for other direction sets.  The
 * workspace (fl_cone_sweep_workspace_bytes, 0 when unsupported) holds each
 * direction's running ray sums of one block, twice.  Direction arrays are HOST
 * memory.  No reference counterpart (the reference has only the dense mask form). */
FL_API size_t fl_cone_sweep_workspace_bytes(int64_t K, int64_t Z, int64_t X, int64_t Y,
                                            const int32_t *pz, const int32_t *py, const int32_t *q,
                                            int64_t ndir);
FL_API int fl_cone_sweep(const float *mu_rot, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y,
                         const int32_t *pz, const int32_t *py, const int32_t *q, const float *weight,
                         int64_t ndir, int adjoint, void *workspace, size_t workspace_bytes,
                         void *stream);

/* ---------------------------------------------------------------------------
 * Adjoint of the attenuated emission model.  The reference differentiates the
 * projector with respect to C only, holding the attenuation fixed
 * (ForwardEmissionBatchFunction, FL/cuda_lib/autograd.py:40-74, backward =
 * backward_emission_batch); these add d/d(atten) and the chain through the
 * attenuation's dependence on the concentrations.
 * ------------------------------------------------------------------------- */

/* out += B^T(...): fl_backward_emission_batch accumulating into `out`. */
FL_API int fl_backward_emission_accumulate(const float *atten, const float *ratio,
                                           const float *em_cs, const float *theta, float *out,
                                           int64_t A, int64_t B, int64_t R, int64_t H, int64_t W,
                                           void *stream);

/* d Pf / d atten contracted with G:  out[a,b,h,w] = G[a,b,w] * sum_r em_cs[a,r] * S_a(C[r,b])(h,w),
 * S_a the projector's clamp-bilinear sample.  C (R,B,H,W), G (A,B,W); out planes for
 * angle a start at a*out_angle_stride floats. */
FL_API int fl_sample_grad(const float *C, const float *em_cs, const float *theta, const float *G,
                          float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W,
                          int64_t out_angle_stride, void *stream);

/* Per line e: Axrf = exp(-T), dT_rot = -gA*Axrf*Ax (mask frame), gx = gA*Axrf,
 * and (if A_tot != NULL) A_tot = Axrf*Ax.  T_rot, dT_rot (K,B,W,H); Ax, gA, gx (K,B,H,W). */
FL_API int fl_xrf_grad_prep(const float *T_rot, const float *Ax, const float *gA, float *dT_rot,
                            float *gx, float *A_tot, int64_t A_tot_element_stride, int64_t K,
                            int64_t B, int64_t H, int64_t W, int side, void *stream);

/* Adjoint of the incident product: dmux[m] = -Ax[m]*U[m], U[0] = 0,
 * U[m+1] = U[m]*(1 - coef_x[e]*F[m]) + gx[m]  (division-free).  F, Ax, gx, dmux (K,B,H,W). */
FL_API int fl_incident_grad(const float *F, const float *coef_x, const float *Ax, const float *gx,
                            float *dmux, int64_t K, int64_t B, int64_t H, int64_t W, void *stream);

/* dF[j] = sum_e coef[j*K+e] * dmu_rot[e] (mask frame to original) + coef_x[j]*dmux[j];
 * dmux may be NULL.  dmu_rot (K,B,W,H); dmux (K,B,H,W); dF (Kc,B,H,W); K <= 8. */
FL_API int fl_mu_adjoint(const float *dmu_rot, const float *dmux, const float *coef,
                         const float *coef_x, float *dF, int64_t Kc, int64_t K, int64_t B,
                         int64_t H, int64_t W, int side, void *stream);

/* Transpose of fl_rotate: dC (+)= Rot^T(dF) over N planes, by gather (no atomics). */
FL_API int fl_rotate_adjoint(const float *dF, float *dC, int64_t N, int64_t H, int64_t W,
                             float cos_t, float sin_t, int accumulate, void *stream);

/* ---------------------------------------------------------------------------
 * cal_frac (FL/core.py:159-211) and its adaptive threshold (FL/image_util.py:337-375).
 * `f64` selects the element type of img/out/thresholds: 0 float32, 1 float64.
 * ------------------------------------------------------------------------- */

/* out[p] = sum_i img[i, p] (in order, in the element type) + 1e-12; img (n, npix)
 * (np.sum(img, axis=0) + 1e-12, core.py:171). */
FL_API int fl_frac_sum(const void *img, void *out, int64_t n, int64_t npix, int f64, void *stream);

/* adaptive_threshold (image_util.py:337-375) of n_img images (H, W) each: mask =
 * img >= the lowest three-class multi-Otsu threshold (skimage threshold_multiotsu,
 * 256 bins), then binary_fill_holes with a 5x5 structure if fill_hole, then
 * `dilation` dilations and `erosion` erosions with the 4-neighbour cross.  mask
 * (n_img, H, W) float64 0/1.  status[i] = 1 where the threshold fails (fewer than
 * three distinct histogram bins, or a non-finite value): that image's mask is 0, as
 * the reference's except branch sets it; else 0 and thresholds[i] holds the
 * threshold.  H * ceil(W/32) <= 18432. */
FL_API int fl_adaptive_mask(const void *img, double *mask, int64_t n_img, int64_t H, int64_t W,
                            int fill_hole, int dilation, int erosion, int f64, int *status,
                            void *thresholds, void *stream);

/* tmp[i, p] = rm_abnormal(img[i, p] / img_sum[p]) in float64 (NaN, inf, < 0 -> 0);
 * core.py:194 with image_util.py:396-401.  single != 0 rounds the quotient to float32
 * (the reference's float32 arithmetic when its mask falls back to 1). */
FL_API int fl_frac_ratio(const void *img, const double *img_sum, double *tmp, int64_t n,
                         int64_t npix, int f64, int single, void *stream);

/* frac[i, p] = (clip_i(tmp[i, p]) * mask[p]) * ratio[p] (core.py:195-201): clip_i to
 * [lohi[2i], lohi[2i+1]] when lohi[2i] <= lohi[2i+1], else no clipping; mask or ratio
 * NULL means 1; single != 0 rounds each product to float32. */
FL_API int fl_frac_final(const double *tmp, const double *mask, const double *ratio,
                         const double *lohi, double *frac, int64_t n, int64_t npix, int single,
                         void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FL_CORRECTION_AMD_H */
