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
 * C (R,B,H,W), em_cs (A,R), theta (A) radians, Pf (A,B,W).
 * `workspace` holds ceil(H/64)*A*B*W floats of partial sums
 * (fl_projector_workspace_bytes); the call allocates nothing. */
FL_API int fl_forward_emission_batch(const float *atten, const float *C, const float *em_cs,
                              const float *theta, float *Pf, int64_t A, int64_t B, int64_t R,
                              int64_t H, int64_t W, void *workspace, size_t workspace_bytes,
                              void *stream);

/* out[r,b] = sum_a em_cs[a,r] * S_a^T( atten[a,b] * ratio[a,b,:] ), S_a^T the exact
 * transpose of the bilinear sampling above.  Replaces backward_emission_batch
 * (atten_cuda.cpp:188-214; kernel atten_cuda_kernel.cu:390-446).  ratio (A,B,W),
 * out (R,B,H,W).  ratio == NULL means ratio == 1 (the MLEM sensitivity). */
FL_API int fl_backward_emission_batch(const float *atten, const float *ratio, const float *em_cs,
                               const float *theta, float *out, int64_t A, int64_t B, int64_t R,
                               int64_t H, int64_t W, void *stream);

/* Single-slice forms: the batch forms with B = 1.  Replace forward_emission
 * (atten_cuda.cpp:82-108) and backward_emission (atten_cuda.cpp:113-135):
 * atten (A,H,W), C (R,H,W), Pf (A,W), ratio (A,W), out (R,H,W). */
FL_API int fl_forward_emission(const float *atten, const float *C, const float *em_cs,
                        const float *theta, float *Pf, int64_t A, int64_t R, int64_t H,
                        int64_t W, void *workspace, size_t workspace_bytes, void *stream);
FL_API int fl_backward_emission(const float *atten, const float *ratio, const float *em_cs,
                         const float *theta, float *out, int64_t A, int64_t R, int64_t H,
                         int64_t W, void *stream);

/* Bytes of workspace fl_forward_emission[_batch] needs. */
FL_API size_t fl_projector_workspace_bytes(int64_t A, int64_t B, int64_t H, int64_t W);

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

#ifdef __cplusplus
}
#endif
#endif /* FL_CORRECTION_AMD_H */
