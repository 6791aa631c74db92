/*
 * fl_oracle.c -- CPU oracle for the XRF self-absorption hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (fl_correction_amd/) links,
 * loads or calls this file; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do, and only to check results.
 *
 * Two families of functions:
 *   *_f32  "same-order" FP32 restatements: every loop runs in the order the
 *          reference kernel's loops run, every product and sum is associated
 *          the way the reference source writes it.  Where the reference sums
 *          with float atomics (non-deterministic order) the serial loop order
 *          of the launch grid is used.  These stand in for "the reference
 *          path" (the reference's CUDA extension cannot be built here).
 *   *_f64  the same mathematics in double precision: the error yardstick.
 *
 * Compile with -ffp-contract=off so that no multiply-add is fused.
 * OpenMP parallel loops only ever split over independent outputs, so results
 * do not depend on the thread count.
 *
 * Reference files restated (paths relative to the reference root):
 *   FL/cuda_lib/atten_cuda_kernel.cu   atten_kernel :10-57, forward_emission_kernel
 *        :91-142, backward_emission_kernel :177-231, bilinear_sample :271-296,
 *        forward_emission_batch_kernel :298-337, backward_emission_batch_kernel
 *        :390-446, huber_grad_kernel :495-534, huber_grad_batch_kernel :564-616
 *   FL/cuda_lib/atten_cuda.cpp         mlem_cuda :244-281, mlem_cuda_batch :283-345
 *   FL/torch_functions.py              torch_xray_atten_3D :83-106 (incident product)
 *   FL/torch_image_util.py             torch_rotate_image_nd :13-114 (grid_sample rotation)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IDX4(a, b, c, d, B, C, D) ((((int64_t)(a) * (B) + (b)) * (C) + (c)) * (D) + (d))

/* ------------------------------------------------------------------------- */
/* A1: detector-mask correlation, FL/cuda_lib/atten_cuda_kernel.cu:10-57      */
/* out[k,z,x,y] = sum_{dz,dx,dy} mu[k, z+dz-Mz/2, x+dx, y+dy-My/2]*mask[dz,dx,dy] */
/* out-of-range taps skipped (:32,:36,:40); loop order dz, dx, dy (:30-51)   */
/* ------------------------------------------------------------------------- */
void or_atten_f32(const float *mu, const float *mask, float *out, int K, int Z, int X,
                  int Y, int Mz, int Mx, int My)
{
    const int hz = Mz / 2, hy = My / 2;
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < K; ++k)
        for (int sli = 0; sli < Z; ++sli)
            for (int row = 0; row < X; ++row)
                for (int col = 0; col < Y; ++col) {
                    float acc = 0.0f;
                    for (int dz = 0; dz < Mz; ++dz) {
                        int z = sli + dz - hz;
                        if (z < 0 || z >= Z) continue;
                        for (int dx = 0; dx < Mx; ++dx) {
                            int x = row + dx;
                            if (x < 0 || x >= X) continue;
                            for (int dy = 0; dy < My; ++dy) {
                                int y = col + dy - hy;
                                if (y < 0 || y >= Y) continue;
                                acc += mu[IDX4(k, z, x, y, Z, X, Y)] *
                                       mask[((int64_t)dz * Mx + dx) * My + dy];
                            }
                        }
                    }
                    out[IDX4(k, sli, row, col, Z, X, Y)] = acc;
                }
}

void or_atten_f64(const double *mu, const double *mask, double *out, int K, int Z, int X,
                  int Y, int Mz, int Mx, int My)
{
    const int hz = Mz / 2, hy = My / 2;
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < K; ++k)
        for (int sli = 0; sli < Z; ++sli)
            for (int row = 0; row < X; ++row)
                for (int col = 0; col < Y; ++col) {
                    double acc = 0.0;
                    for (int dz = 0; dz < Mz; ++dz) {
                        int z = sli + dz - hz;
                        if (z < 0 || z >= Z) continue;
                        for (int dx = 0; dx < Mx; ++dx) {
                            int x = row + dx;
                            if (x >= X) break;
                            for (int dy = 0; dy < My; ++dy) {
                                int y = col + dy - hy;
                                if (y < 0 || y >= Y) continue;
                                acc += mu[IDX4(k, z, x, y, Z, X, Y)] *
                                       mask[((int64_t)dz * Mx + dx) * My + dy];
                            }
                        }
                    }
                    out[IDX4(k, sli, row, col, Z, X, Y)] = acc;
                }
}

/* ------------------------------------------------------------------------- */
/* F1 single slice: forward_emission_kernel, atten_cuda_kernel.cu:91-142      */
/* ------------------------------------------------------------------------- */
void or_forward_single_f32(const float *atten, const float *C, const float *em_cs,
                           const float *theta, float *Pf, int n_angle, int H, int W,
                           int n_ref)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < n_angle; ++i)
        for (int q = 0; q < W; ++q) {
            float cx = (W - 1) * 0.5f;
            float cy = (H - 1) * 0.5f;
            float th = theta[i];
            float c = cosf(-th);
            float s = sinf(-th);
            float acc = 0.0f;
            for (int row = 0; row < H; ++row) {
                float x = c * (row - cy) - s * (q - cx) + cy;
                float y = s * (row - cy) + c * (q - cx) + cx;
                x = fminf(fmaxf(x, 0.f), H - 1.001f);
                y = fminf(fmaxf(y, 0.f), W - 1.001f);
                int x0 = (int)x;
                int y0 = (int)y;
                float dx = x - x0;
                float dy = y - y0;
                float val = 0.0f;
                for (int j = 0; j < n_ref; ++j) {
                    const float *Cj = C + (int64_t)j * H * W;
                    float interp = Cj[x0 * W + y0] * (1 - dx) * (1 - dy) +
                                   Cj[x0 * W + y0 + 1] * (1 - dx) * dy +
                                   Cj[(x0 + 1) * W + y0] * dx * (1 - dy) +
                                   Cj[(x0 + 1) * W + y0 + 1] * dx * dy;
                    val += em_cs[i * n_ref + j] * interp;
                }
                acc += atten[(int64_t)i * H * W + row * W + q] * val;
            }
            Pf[(int64_t)i * W + q] = acc;
        }
}

/* B1 single slice: backward_emission_kernel, atten_cuda_kernel.cu:177-231.
 * The reference scatters with atomicAdd; here the scatter runs serially in
 * launch-grid order (i, q, row_det, j).  out is zeroed first (:254-255). */
void or_backward_single_f32(const float *atten, const float *ratio, const float *em_cs,
                            const float *theta, float *out, int n_angle, int H, int W,
                            int n_ref)
{
    memset(out, 0, sizeof(float) * (size_t)n_ref * H * W);
    for (int i = 0; i < n_angle; ++i)
        for (int q = 0; q < W; ++q) {
            float cx = (W - 1) * 0.5f;
            float cy = (H - 1) * 0.5f;
            float th = theta[i];
            float c = cosf(th);
            float s = sinf(th);
            float r = ratio[(int64_t)i * W + q];
            for (int row_det = 0; row_det < H; ++row_det) {
                float col_img = c * (q - cx) - s * (row_det - cy) + cx;
                float row_img = s * (q - cx) + c * (row_det - cy) + cy;
                col_img = fminf(fmaxf(col_img, 0.f), W - 1.001f);
                row_img = fminf(fmaxf(row_img, 0.f), H - 1.001f);
                int x0 = (int)col_img;
                int y0 = (int)row_img;
                float dx = col_img - x0;
                float dy = row_img - y0;
                float a = atten[(int64_t)i * H * W + row_det * W + q];
                for (int j = 0; j < n_ref; ++j) {
                    float w = em_cs[i * n_ref + j] * a * r;
                    float *o = out + (int64_t)j * H * W;
                    o[y0 * W + x0] += w * (1 - dx) * (1 - dy);
                    o[y0 * W + (x0 + 1)] += w * dx * (1 - dy);
                    o[(y0 + 1) * W + x0] += w * (1 - dx) * dy;
                    o[(y0 + 1) * W + (x0 + 1)] += w * dx * dy;
                }
            }
        }
}

/* bilinear_sample, atten_cuda_kernel.cu:271-296 (called with y=row, x=col) */
static inline float bilinear_sample_f32(const float *img, int H, int W, float y, float x)
{
    x = fminf(fmaxf(x, 0.f), W - 1.001f);
    y = fminf(fmaxf(y, 0.f), H - 1.001f);
    int x0 = (int)floorf(x);
    int y0 = (int)floorf(y);
    int x1 = x0 + 1;
    int y1 = y0 + 1;
    float dx = x - x0;
    float dy = y - y0;
    float v00 = img[y0 * W + x0];
    float v01 = (x1 < W) ? img[y0 * W + x1] : v00;
    float v10 = (y1 < H) ? img[y1 * W + x0] : v00;
    float v11 = (x1 < W && y1 < H) ? img[y1 * W + x1] : v00;
    return (1 - dx) * (1 - dy) * v00 + dx * (1 - dy) * v01 + (1 - dx) * dy * v10 +
           dx * dy * v11;
}

/* F1 batch: forward_emission_batch_kernel, atten_cuda_kernel.cu:298-337.
 * Per (a,b,w) the reference atomically adds one partial per h; here the h
 * loop runs in increasing order (the deterministic reading of :312-335). */
void or_forward_batch_f32(const float *atten, const float *C, const float *em_cs,
                          const float *theta, float *Pf, int A, int B, int R, int H, int W)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < A; ++a)
        for (int b = 0; b < B; ++b) {
            float th = theta[a];
            float c = cosf(-th);
            float s = sinf(-th);
            float cx = (W - 1) * 0.5f;
            float cy = (H - 1) * 0.5f;
            for (int w = 0; w < W; ++w) {
                float acc = 0.0f;
                for (int h = 0; h < H; ++h) {
                    float x_rot = c * (h - cy) - s * (w - cx) + cy;
                    float y_rot = s * (h - cy) + c * (w - cx) + cx;
                    float sum_val = 0.0f;
                    for (int r = 0; r < R; ++r) {
                        const float *Cr = C + ((int64_t)r * B + b) * H * W;
                        float val = bilinear_sample_f32(Cr, H, W, x_rot, y_rot);
                        float cs = em_cs[a * R + r];
                        sum_val += cs * val * atten[IDX4(a, b, h, w, B, H, W)];
                    }
                    acc += sum_val;
                }
                Pf[((int64_t)a * B + b) * W + w] = acc;
            }
        }
}

/* B1 batch: backward_emission_batch_kernel, atten_cuda_kernel.cu:390-446.
 * Slices are independent, so the loop is split over b; inside a slice the
 * scatter runs in launch-grid order (a, q, row_det, j). */
void or_backward_batch_f32(const float *atten, const float *ratio, const float *em_cs,
                           const float *theta, float *out, int A, int B, int R, int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; ++b) {
        for (int j = 0; j < R; ++j)
            memset(out + ((int64_t)j * B + b) * H * W, 0, sizeof(float) * (size_t)H * W);
        for (int a = 0; a < A; ++a) {
            float cx = (W - 1) * 0.5f;
            float cy = (H - 1) * 0.5f;
            float th = theta[a];
            float c = cosf(th);
            float s = sinf(th);
            for (int q = 0; q < W; ++q) {
                float r = ratio[((int64_t)a * B + b) * W + q];
                for (int row_det = 0; row_det < H; ++row_det) {
                    float col_img = c * (q - cx) - s * (row_det - cy) + cx;
                    float row_img = s * (q - cx) + c * (row_det - cy) + cy;
                    col_img = fminf(fmaxf(col_img, 0.f), W - 1.001f);
                    row_img = fminf(fmaxf(row_img, 0.f), H - 1.001f);
                    int x0 = (int)col_img;
                    int y0 = (int)row_img;
                    float dx = col_img - x0;
                    float dy = row_img - y0;
                    float a_val = atten[IDX4(a, b, row_det, q, B, H, W)];
                    for (int jj = 0; jj < R; ++jj) {
                        float w = em_cs[a * R + jj] * a_val * r;
                        float *o = out + IDX4(jj, b, y0, x0, B, H, W);
                        o[0] += w * (1 - dx) * (1 - dy);
                        o[1] += w * dx * (1 - dy);
                        o[W] += w * (1 - dx) * dy;
                        o[W + 1] += w * dx * dy;
                    }
                }
            }
        }
    }
}

/* ---- FP64 yardsticks of F1/B1 (same sample map, double arithmetic) ---- */
static inline void sample_point_f64(double th, int h, int w, int H, int W, int *x0, int *y0,
                                    double *dx, double *dy)
{
    /* x = row coordinate, y = column coordinate of the sample (.cu:320-321) */
    double c = cos(th), s = sin(th);
    double cx = (W - 1) * 0.5, cy = (H - 1) * 0.5;
    double xr = c * (h - cy) + s * (w - cx) + cy;
    double yr = -s * (h - cy) + c * (w - cx) + cx;
    xr = fmin(fmax(xr, 0.0), H - 1.001);
    yr = fmin(fmax(yr, 0.0), W - 1.001);
    *x0 = (int)floor(xr);
    *y0 = (int)floor(yr);
    *dx = xr - *x0;
    *dy = yr - *y0;
}

void or_forward_batch_f64(const double *atten, const double *C, const double *em_cs,
                          const double *theta, double *Pf, int A, int B, int R, int H, int W)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < A; ++a)
        for (int b = 0; b < B; ++b)
            for (int w = 0; w < W; ++w) {
                double acc = 0.0;
                for (int h = 0; h < H; ++h) {
                    int x0, y0;
                    double dx, dy;
                    sample_point_f64(theta[a], h, w, H, W, &x0, &y0, &dx, &dy);
                    double v = 0.0;
                    for (int r = 0; r < R; ++r) {
                        const double *Cr = C + ((int64_t)r * B + b) * H * W;
                        double val = (1 - dx) * (1 - dy) * Cr[x0 * W + y0] +
                                     (1 - dx) * dy * Cr[x0 * W + y0 + 1] +
                                     dx * (1 - dy) * Cr[(x0 + 1) * W + y0] +
                                     dx * dy * Cr[(x0 + 1) * W + y0 + 1];
                        v += em_cs[a * R + r] * val;
                    }
                    acc += v * atten[IDX4(a, b, h, w, B, H, W)];
                }
                Pf[((int64_t)a * B + b) * W + w] = acc;
            }
}

void or_backward_batch_f64(const double *atten, const double *ratio, const double *em_cs,
                           const double *theta, double *out, int A, int B, int R, int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int b = 0; b < B; ++b) {
        for (int j = 0; j < R; ++j)
            memset(out + ((int64_t)j * B + b) * H * W, 0, sizeof(double) * (size_t)H * W);
        for (int a = 0; a < A; ++a)
            for (int q = 0; q < W; ++q)
                for (int h = 0; h < H; ++h) {
                    int x0, y0;
                    double dx, dy;
                    sample_point_f64(theta[a], h, q, H, W, &x0, &y0, &dx, &dy);
                    double base = atten[IDX4(a, b, h, q, B, H, W)] *
                                  ratio[((int64_t)a * B + b) * W + q];
                    for (int j = 0; j < R; ++j) {
                        double w = em_cs[a * R + j] * base;
                        double *o = out + IDX4(j, b, x0, y0, B, H, W);
                        o[0] += w * (1 - dx) * (1 - dy);
                        o[1] += w * (1 - dx) * dy;
                        o[W] += w * dx * (1 - dy);
                        o[W + 1] += w * dx * dy;
                    }
                }
    }
}

/* ------------------------------------------------------------------------- */
/* U1: huber_grad_batch_kernel, atten_cuda_kernel.cu:564-616 (also the       */
/* single-slice kernel :495-534 with B=1).  Neighbour order up, down, left,  */
/* right (:596-597).                                                          */
/* ------------------------------------------------------------------------- */
void or_huber_batch_f32(const float *C, float *grad, int R, int B, int H, int W, float delta)
{
    static const int dr[4] = {-1, 1, 0, 0};
    static const int dc[4] = {0, 0, -1, 1};
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < R; ++j)
        for (int b = 0; b < B; ++b)
            for (int r = 0; r < H; ++r)
                for (int c = 0; c < W; ++c) {
                    int64_t base = IDX4(j, b, r, c, B, H, W);
                    float Cp = C[base];
                    float g = 0.0f;
                    for (int k = 0; k < 4; ++k) {
                        int rr = r + dr[k], cc = c + dc[k];
                        if (rr < 0 || rr >= H || cc < 0 || cc >= W) continue;
                        float diff = Cp - C[IDX4(j, b, rr, cc, B, H, W)];
                        if (fabsf(diff) <= delta)
                            g += diff;
                        else
                            g += delta * copysignf(1.0f, diff);
                    }
                    grad[base] = g;
                }
}

void or_huber_batch_f64(const double *C, double *grad, int R, int B, int H, int W,
                        double delta)
{
    static const int dr[4] = {-1, 1, 0, 0};
    static const int dc[4] = {0, 0, -1, 1};
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < R; ++j)
        for (int b = 0; b < B; ++b)
            for (int r = 0; r < H; ++r)
                for (int c = 0; c < W; ++c) {
                    int64_t base = IDX4(j, b, r, c, B, H, W);
                    double g = 0.0;
                    for (int k = 0; k < 4; ++k) {
                        int rr = r + dr[k], cc = c + dc[k];
                        if (rr < 0 || rr >= H || cc < 0 || cc >= W) continue;
                        double diff = C[base] - C[IDX4(j, b, rr, cc, B, H, W)];
                        g += fabs(diff) <= delta ? diff : delta * copysign(1.0, diff);
                    }
                    grad[base] = g;
                }
}

/* ------------------------------------------------------------------------- */
/* U2: mlem_cuda_batch, atten_cuda.cpp:283-345.                               */
/*   sens = max(B^T 1, 1e-6); per iteration: Pf = max(F C, 1e-6);            */
/*   ratio = I / Pf; back = B^T ratio;                                        */
/*   C = C*back / (sens + beta*huber(C))  (beta > 0)   or  C*back/sens;      */
/*   C = max(C, 0)  (NaN propagates, as torch clamp_min does).               */
/* ------------------------------------------------------------------------- */
void or_mlem_batch_f32(const float *C_init, const float *atten, const float *em_cs,
                       const float *theta, const float *I, float *C_out, int R, int A, int B,
                       int H, int W, int n_iter, double beta, double delta)
{
    const int64_t nC = (int64_t)R * B * H * W;
    const int64_t nP = (int64_t)A * B * W;
    float *sens = (float *)malloc(sizeof(float) * nC);
    float *back = (float *)malloc(sizeof(float) * nC);
    float *grad = (float *)calloc(nC, sizeof(float));
    float *Pf = (float *)malloc(sizeof(float) * nP);
    float *ones = (float *)malloc(sizeof(float) * nP);
    const float eps = (float)1e-6;
    const float fbeta = (float)beta;
    memcpy(C_out, C_init, sizeof(float) * nC);
    for (int64_t i = 0; i < nP; ++i) ones[i] = 1.0f;
    or_backward_batch_f32(atten, ones, em_cs, theta, sens, A, B, R, H, W);
    for (int64_t i = 0; i < nC; ++i)
        if (sens[i] < eps) sens[i] = eps;
    for (int it = 0; it < n_iter; ++it) {
        or_forward_batch_f32(atten, C_out, em_cs, theta, Pf, A, B, R, H, W);
        for (int64_t i = 0; i < nP; ++i) {
            float p = Pf[i] < eps ? eps : Pf[i];
            Pf[i] = I[i] / p; /* ratio */
        }
        or_backward_batch_f32(atten, Pf, em_cs, theta, back, A, B, R, H, W);
        if (beta > 0.0) or_huber_batch_f32(C_out, grad, R, B, H, W, (float)delta);
        for (int64_t i = 0; i < nC; ++i) {
            float num = C_out[i] * back[i];
            float den = beta > 0.0 ? sens[i] + fbeta * grad[i] : sens[i];
            float v = num / den;
            C_out[i] = v < 0.0f ? 0.0f : v;
        }
    }
    free(sens);
    free(back);
    free(grad);
    free(Pf);
    free(ones);
}

void or_mlem_batch_f64(const double *C_init, const double *atten, const double *em_cs,
                       const double *theta, const double *I, double *C_out, int R, int A,
                       int B, int H, int W, int n_iter, double beta, double delta)
{
    const int64_t nC = (int64_t)R * B * H * W;
    const int64_t nP = (int64_t)A * B * W;
    double *sens = (double *)malloc(sizeof(double) * nC);
    double *back = (double *)malloc(sizeof(double) * nC);
    double *grad = (double *)calloc(nC, sizeof(double));
    double *Pf = (double *)malloc(sizeof(double) * nP);
    double *ones = (double *)malloc(sizeof(double) * nP);
    memcpy(C_out, C_init, sizeof(double) * nC);
    for (int64_t i = 0; i < nP; ++i) ones[i] = 1.0;
    or_backward_batch_f64(atten, ones, em_cs, theta, sens, A, B, R, H, W);
    for (int64_t i = 0; i < nC; ++i)
        if (sens[i] < 1e-6) sens[i] = 1e-6;
    for (int it = 0; it < n_iter; ++it) {
        or_forward_batch_f64(atten, C_out, em_cs, theta, Pf, A, B, R, H, W);
        for (int64_t i = 0; i < nP; ++i) Pf[i] = I[i] / (Pf[i] < 1e-6 ? 1e-6 : Pf[i]);
        or_backward_batch_f64(atten, Pf, em_cs, theta, back, A, B, R, H, W);
        if (beta > 0.0) or_huber_batch_f64(C_out, grad, R, B, H, W, delta);
        for (int64_t i = 0; i < nC; ++i) {
            double den = beta > 0.0 ? sens[i] + beta * grad[i] : sens[i];
            double v = C_out[i] * back[i] / den;
            C_out[i] = v < 0.0 ? 0.0 : v;
        }
    }
    free(sens);
    free(back);
    free(grad);
    free(Pf);
    free(ones);
}

/* ------------------------------------------------------------------------- */
/* A4: incident-beam attenuation, FL/torch_functions.py:95-102.               */
/*   out[n,r,c] = prod_{m=r+1}^{H-1} (1 - mux[n,m,c]),  out[n,H-1,c] = 1      */
/* FP32 order: torch.cumprod over the flipped rows, i.e. the running product */
/* starts at row H-1 and multiplies downward.                                 */
/* ------------------------------------------------------------------------- */
void or_incident_f32(const float *mux, float *out, int N, int H, int W)
{
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n)
        for (int c = 0; c < W; ++c) {
            const float *m = mux + (int64_t)n * H * W;
            float *o = out + (int64_t)n * H * W;
            o[(int64_t)(H - 1) * W + c] = 1.0f;
            float p = 1.0f;
            for (int r = H - 2; r >= 0; --r) {
                float f = 1.0f - m[(int64_t)(r + 1) * W + c];
                p = (r == H - 2) ? f : p * f;
                o[(int64_t)r * W + c] = p;
            }
        }
}

void or_incident_f64(const double *mux, double *out, int N, int H, int W)
{
    for (int n = 0; n < N; ++n)
        for (int c = 0; c < W; ++c) {
            const double *m = mux + (int64_t)n * H * W;
            double *o = out + (int64_t)n * H * W;
            double p = 1.0;
            o[(int64_t)(H - 1) * W + c] = 1.0;
            for (int r = H - 2; r >= 0; --r) {
                p *= 1.0 - m[(int64_t)(r + 1) * W + c];
                o[(int64_t)r * W + c] = p;
            }
        }
}

/* ------------------------------------------------------------------------- */
/* A5: in-plane rotation, FL/torch_image_util.py:13-114 with enable_crop,     */
/* align_corners=False, bilinear, zero padding: affine_grid with              */
/* [[cos,-sin,0],[sin,cos,0]] followed by grid_sample.  Double arithmetic.    */
/* ------------------------------------------------------------------------- */
void or_rotate_f64(const double *x, double theta, double *out, int N, int H, int W)
{
    const double c = cos(theta), s = sin(theta);
#pragma omp parallel for schedule(static)
    for (int n = 0; n < N; ++n)
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                double xn = (2.0 * j + 1.0) / W - 1.0;
                double yn = (2.0 * i + 1.0) / H - 1.0;
                double gx = c * xn - s * yn;
                double gy = s * xn + c * yn;
                double ix = ((gx + 1.0) * W - 1.0) * 0.5;
                double iy = ((gy + 1.0) * H - 1.0) * 0.5;
                double fx0 = floor(ix), fy0 = floor(iy);
                int x0 = (int)fx0, y0 = (int)fy0;
                double tx = ix - fx0, ty = iy - fy0;
                const double *img = x + (int64_t)n * H * W;
                double v = 0.0;
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx) {
                        int yy = y0 + dy, xx = x0 + dx;
                        if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                        double wgt = (dx ? tx : 1.0 - tx) * (dy ? ty : 1.0 - ty);
                        v += wgt * img[(int64_t)yy * W + xx];
                    }
                out[(int64_t)n * H * W + (int64_t)i * W + j] = v;
            }
}

/* ------------------------------------------------------------------------- */
/* Direction-sampled detector cone (the definition the product's cone march  */
/* must reproduce).  Frame of atten_kernel: volume (Z, X, Y), cone toward +X. */
/*   out[z,x,y] = sum_d w_d sum_{t=1}^{X-1-x} bilin0(mu[., x+t, .],           */
/*                                   z + t*pz_d/q_d, y + t*py_d/q_d)          */
/* bilin0 = bilinear interpolation with zero outside the (Z, Y) plane.        */
/* Offsets are exact rationals: floor and remainder in integer arithmetic.   */
/* This is atten_kernel (.cu:10-57) applied to the ray-splat mask of the      */
/* same directions (every tap of the mask lands on an integer voxel).        */
/* ------------------------------------------------------------------------- */
static inline int64_t floordiv64(int64_t a, int64_t b)
{
    int64_t q = a / b, r = a % b;
    return (r != 0 && ((r < 0) != (b < 0))) ? q - 1 : q;
}

/* Outputs z0 <= z < z1, y0 <= y < y1 only (all of mu is read): out is          */
/* (z1-z0, X, y1-y0).                                                          */
void or_cone_march_box_f64(const double *mu, double *out, int Z, int X, int Y, int z0s, int z1s,
                           int y0s, int y1s, const int *pz, const int *py, const int *qd,
                           const double *wgt, int ndir)
{
    const int Yo = y1s - y0s;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int z = z0s; z < z1s; ++z)
        for (int x = 0; x < X; ++x)
            for (int y = y0s; y < y1s; ++y) {
                double acc = 0.0;
                for (int d = 0; d < ndir; ++d) {
                    double sd = 0.0;
                    for (int t = 1; x + t < X; ++t) {
                        int64_t nz = (int64_t)t * pz[d], ny = (int64_t)t * py[d];
                        int64_t z0 = floordiv64(nz, qd[d]), y0 = floordiv64(ny, qd[d]);
                        double fz = (double)(nz - z0 * qd[d]) / qd[d];
                        double fy = (double)(ny - y0 * qd[d]) / qd[d];
                        for (int a = 0; a < 2; ++a)
                            for (int b = 0; b < 2; ++b) {
                                int64_t zz = z + z0 + a, yy = y + y0 + b;
                                if (zz < 0 || zz >= Z || yy < 0 || yy >= Y) continue;
                                double wz = a ? fz : 1.0 - fz, wy = b ? fy : 1.0 - fy;
                                if (wz == 0.0 || wy == 0.0) continue;
                                sd += wz * wy * mu[((int64_t)zz * X + (x + t)) * Y + yy];
                            }
                    }
                    acc += wgt[d] * sd;
                }
                out[((int64_t)(z - z0s) * X + x) * Yo + (y - y0s)] = acc;
            }
}

void or_cone_march_f64(const double *mu, double *out, int Z, int X, int Y, const int *pz,
                       const int *py, const int *qd, const double *wgt, int ndir)
{
    or_cone_march_box_f64(mu, out, Z, X, Y, 0, Z, 0, Y, pz, py, qd, wgt, ndir);
}

/* ------------------------------------------------------------------------- */
/* Legacy CPU path, FL/legacy/FL_correction_core.py, FP64 (timed as the      */
/* legacy-faithful CPU baseline, and a second statement of the projector).   */
/*                                                                           */
/* or_legacy_H_f64: generate_H_jit :2203-2279, the dense system matrix of    */
/* one slice: row i*W + q (angle i, detector column q), column r*W + c.      */
/* Detector point (p = row, q = column) samples the image at                 */
/*   t0 = cos(-th)(p - cx) - sin(-th)(q - cy) + cx   (row)                   */
/*   t1 = sin(-th)(p - cx) + cos(-th)(q - cy) + cy   (column)                */
/* with cx = (W-1)/2, cy = (H-1)/2 paired as the reference pairs them        */
/* (:2234-2235, :2252-2253; the same as the projector's for square slices).  */
/* Points outside [0, H-1] x [0, W-1] are skipped (:2255-2256; the GPU       */
/* projector clamps them instead); inside ones add att[p,q] times the        */
/* bilinear weights of the floor/ceil corners, in the order ul, dl, ur, dr   */
/* (:2258-2276).  Loops: angles, then columns, then rows (:2239, :2247-2248).*/
void or_legacy_H_f64(const double *theta, const double *atten, double *Hm, int A, int H, int W)
{
    const int64_t npix = (int64_t)H * W;
    const double cx = (W - 1) / 2.0, cy = (H - 1) / 2.0;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A; ++i) {
        const double T00 = cos(-theta[i]), T01 = -sin(-theta[i]);
        const double T10 = sin(-theta[i]), T11 = cos(-theta[i]);
        const double *att = atten + (int64_t)i * npix;
        double *Hi = Hm + (int64_t)i * W * npix;
        memset(Hi, 0, sizeof(double) * (size_t)(W * npix));
        for (int q = 0; q < W; ++q)
            for (int p = 0; p < H; ++p) {
                const double t0 = T00 * (p - cx) + T01 * (q - cy) + cx;
                const double t1 = T10 * (p - cx) + T11 * (q - cy) + cy;
                if (t0 > H - 1 || t0 < 0 || t1 > W - 1 || t1 < 0) continue;
                const double rf = t0 - floor(t0), cf = t1 - floor(t1);
                const int64_t ru = (int64_t)floor(t0), rd = (int64_t)ceil(t0);
                const int64_t cl = (int64_t)floor(t1), cr = (int64_t)ceil(t1);
                const double a = att[(int64_t)p * W + q];
                double *row = Hi + (int64_t)q * npix;
                row[ru * W + cl] += a * (1 - rf) * (1 - cf);
                row[rd * W + cl] += a * rf * (1 - cf);
                row[ru * W + cr] += a * (1 - rf) * cf;
                row[rd * W + cr] += a * rf * cf;
            }
    }
}

/* or_legacy_mlem_f64: mlem_matrix :99-119 on one slice.  A = C0 (float32,   */
/* :104), y clipped at 0 (:107), b_sum = 1^T p (:106, row_sum :248-252); per  */
/* iteration Pf = p A, Pf < 1e-6 -> 1 (:111-112, not a clamp to 1e-6),       */
/* a_sum = (y/Pf)^T p, A = A * a_sum / b_sum, NaN -> 0 (:113-117): pixels     */
/* that no row samples (b_sum = 0) become 0.  No Huber term.  Products are   */
/* summed in index order (the reference's BLAS order is not fixed).          */
void or_legacy_mlem_f64(const double *Hm, const double *y_in, const float *C0, double *out, int64_t nrow,
                        int64_t npix, int n_iter)
{
    double *bsum = calloc((size_t)npix, sizeof(double));
    double *asum = malloc(sizeof(double) * (size_t)npix);
    double *t2 = malloc(sizeof(double) * (size_t)nrow);
    for (int64_t c = 0; c < npix; ++c) out[c] = (double)C0[c];
    if (!bsum || !asum || !t2) {  /* out of memory: NaN marks the result invalid */
        for (int64_t c = 0; c < npix; ++c) out[c] = NAN;
        free(bsum);
        free(asum);
        free(t2);
        return;
    }
    const int64_t cb = 256;  /* column blocks: each thread owns whole columns */
#pragma omp parallel for schedule(static)
    for (int64_t c0 = 0; c0 < npix; c0 += cb) {
        const int64_t c1 = c0 + cb < npix ? c0 + cb : npix;
        for (int64_t r = 0; r < nrow; ++r)
            for (int64_t c = c0; c < c1; ++c) bsum[c] += Hm[r * npix + c];
    }
    for (int it = 0; it < n_iter; ++it) {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < nrow; ++r) {
            double pf = 0.0;
            for (int64_t c = 0; c < npix; ++c) pf += Hm[r * npix + c] * out[c];
            if (pf < 1e-6) pf = 1.0;
            const double y = y_in[r] < 0 ? 0.0 : y_in[r];
            t2[r] = y / pf;
        }
#pragma omp parallel for schedule(static)
        for (int64_t c0 = 0; c0 < npix; c0 += cb) {
            const int64_t c1 = c0 + cb < npix ? c0 + cb : npix;
            for (int64_t c = c0; c < c1; ++c) asum[c] = 0.0;
            for (int64_t r = 0; r < nrow; ++r)
                for (int64_t c = c0; c < c1; ++c) asum[c] += t2[r] * Hm[r * npix + c];
            for (int64_t c = c0; c < c1; ++c) {
                const double v = out[c] * (asum[c] / bsum[c]);
                out[c] = isnan(v) ? 0.0 : v;
            }
        }
    }
    free(bsum);
    free(asum);
    free(t2);
}
