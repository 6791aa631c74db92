// Adjoint of the attenuated emission model for CDNA4.
//
// Forward model, per angle a and emission line e (K lines, frac = C):
//   F_j   = Rot_a(C_j)                                   rotation (zeros, align_corners=False)
//   A_x,e = prod_{m>r} (1 - coef_x[e] F_e[m])             incident beam, per column
//   T_e   = Cone(sum_j coef[j,e] F_j)                     fluorescence path (mask frame)
//   A_e   = exp(-T_e) * A_x,e
//   Pf_e  = em[a,e] * sum_h A_e * S_a(C_e)                rotated clamp-bilinear projector
// Given G = dL/dPf, dL/dC collects
//   (1) the projector term  em * S_a^T(A_e * G)           (back-projection kernel)
//   (2) the attenuation term, through
//       gA   = em * G * S_a(C_e)                          sample_grad_kernel
//       dT   = -gA * A_e,  gx = gA * exp(-T_e)            xrf_grad_prep_kernel
//       dmu  = Cone^T(dT)                                 cone kernel, reverse mode
//       dmux[m] = -A_x[m] * U[m],  U[m+1] = U[m] f[m] + gx[m]   (division-free)
//                                                         incident_grad_kernel
//       dF_j = sum_e coef[j,e] dmu_e + coef_x[j] dmux_j   mu_adjoint_kernel
//       dC_j += Rot_a^T(dF_j)                              rotate_adjoint_kernel
// Every kernel is a gather or a per-column scan: no atomics.
#include "common.hpp"

#include <algorithm>

namespace fl {

// ---------------------------------------------------------------------------
// out[a,b,h,w] = G[a,b,w] * sum_r em[a,r] * S_a(C[r,b])(h,w)
// (d Pf / d atten for forward_emission_batch; the gA term above).  The sample
// arithmetic is the forward projector's, so this is its exact derivative.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sample_grad_kernel(const float *__restrict__ C, const float *__restrict__ em,
                                                          const float *__restrict__ theta, const float *__restrict__ G,
                                                          float *__restrict__ out, int A, int B, int R, int H, int W,
                                                          int64_t out_astride)
{
    const int w = blockIdx.x * 64 + (threadIdx.x & 63);
    const int h = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int ab = blockIdx.z;
    const int a = ab / B, b = ab % B;
    if (w >= W || h >= H) return;
    const float th = finite_or_zero(theta[a]);
    const float c = cosf(-th), s = sinf(-th);
    const float cx = (W - 1) * 0.5f, cy = (H - 1) * 0.5f;
    const float wc = (float)w - cx, hc = (float)h - cy;
    const float x = clampf(fmaf(c, hc, fmaf(-s, wc, cy)), 0.f, H - 1.001f);
    const float y = clampf(fmaf(s, hc, fmaf(c, wc, cx)), 0.f, W - 1.001f);
    const int x0 = (int)x, y0 = (int)y;
    const float fx = x - (float)x0, fy = y - (float)y0;
    const int64_t HW = (int64_t)H * W;
    const int64_t li = (int64_t)x0 * W + y0;
    float val = 0.0f;
    for (int r = 0; r < R; ++r) {
        const float *Cr = C + ((int64_t)r * B + b) * HW;
        const float v00 = Cr[li], v01 = Cr[li + 1], v10 = Cr[li + W], v11 = Cr[li + W + 1];
        const float top = fmaf(fy, v01 - v00, v00);
        const float bot = fmaf(fy, v11 - v10, v10);
        val = fmaf(em[(int64_t)a * R + r], fmaf(fx, bot - top, top), val);
    }
    out[(int64_t)a * out_astride + (int64_t)b * HW + (int64_t)h * W + w] = val * G[((int64_t)a * B + b) * W + w];
}

// Offset of original-frame voxel (r, c) in one (X, Y) plane of the mask frame (side 0: X = c, Y = H-1-r;
// side 1: X = W-1-c, Y = r).
__device__ __forceinline__ int64_t mframe(int r, int c, int H, int W, int side)
{
    return side == 0 ? (int64_t)c * H + (H - 1 - r) : (int64_t)(W - 1 - c) * H + r;
}

constexpr int kAR = 16, kAC = 64, kAMaxLines = 8;

// ---------------------------------------------------------------------------
// Per element line e (blockIdx.z = e*B + b), a 16x64 tile:
//   Axrf = exp(-T),  dT = -gA * Axrf * Ax  (mask frame),  gx = gA * Axrf,
//   A_tot = Axrf * Ax (optional, for the projector term)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void xrf_grad_prep_kernel(const float *__restrict__ T_rot,
                                                            const float *__restrict__ Ax,
                                                            const float *__restrict__ gA, float *__restrict__ dT_rot,
                                                            float *__restrict__ gx, float *__restrict__ A_tot,
                                                            int64_t A_tot_estride, int B, int H, int W, int side)
{
    __shared__ float tile[kAR][kAC + 1];
    const int c0 = blockIdx.x * kAC, r0 = blockIdx.y * kAR;
    const int b = blockIdx.z % B, e = blockIdx.z / B;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t HW = (int64_t)H * W;
    const int64_t eb = ((int64_t)e * B + b) * HW;
    // transposed read of T: lanes run along the mask-frame Y (contiguous)
    const int q4 = lane >> 4, rr = lane & 15;
    for (int cc = wave * 4 + q4; cc < kAC; cc += 16) {
        const int c = c0 + cc, r = r0 + rr;
        if (c < W && r < H) tile[rr][cc] = T_rot[eb + mframe(r, c, H, W, side)];
    }
    __syncthreads();
    for (int r_ = wave; r_ < kAR; r_ += 4) {
        const int r = r0 + r_, c = c0 + lane;
        if (r >= H || c >= W) continue;
        const int64_t o = eb + (int64_t)r * W + c;
        const float axrf = expf(-tile[r_][lane]);
        const float ax = Ax[o];
        const float g = gA[o];
        const float at = axrf * ax;
        gx[o] = g * axrf;
        if (A_tot) A_tot[(int64_t)e * A_tot_estride + (int64_t)b * HW + (int64_t)r * W + c] = at;
        tile[r_][lane] = -g * at;
    }
    __syncthreads();
    for (int cc = wave * 4 + q4; cc < kAC; cc += 16) {
        const int c = c0 + cc, r = r0 + rr;
        if (c < W && r < H) dT_rot[eb + mframe(r, c, H, W, side)] = tile[rr][cc];
    }
}

// ---------------------------------------------------------------------------
// Incident-beam adjoint, one thread per (e, b, column):
//   f[m] = 1 - coef_x[e] F[e,b,m,c];  dmux[m] = -Ax[m] * U[m];  U[0] = 0, U[m+1] = U[m] f[m] + gx[m]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void incident_grad_kernel(const float *__restrict__ F,
                                                            const float *__restrict__ coef_x,
                                                            const float *__restrict__ Ax,
                                                            const float *__restrict__ gx, float *__restrict__ dmux,
                                                            int B, int H, int W)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y, e = blockIdx.z;
    if (c >= W) return;
    const int64_t HW = (int64_t)H * W;
    const int64_t base = ((int64_t)e * B + b) * HW + c;
    const float k = coef_x[e];
    float U = 0.0f;
    for (int m = 0; m < H; ++m) {
        const int64_t o = base + (int64_t)m * W;
        dmux[o] = -Ax[o] * U;
        const float f = 1.0f - F[o] * k;
        U = fmaf(U, f, gx[o]);
    }
}

// ---------------------------------------------------------------------------
// dF[j] = sum_e coef[j*K + e] * dmu_rot[e] (mask frame -> original) + coef_x[j] * dmux[j]
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void mu_adjoint_kernel(const float *__restrict__ dmu_rot,
                                                         const float *__restrict__ dmux,
                                                         const float *__restrict__ coef,
                                                         const float *__restrict__ coef_x, float *__restrict__ dF,
                                                         int Kc, int K, int B, int H, int W, int side)
{
    __shared__ float tile[kAMaxLines][kAR][kAC + 1];
    const int c0 = blockIdx.x * kAC, r0 = blockIdx.y * kAR;
    const int b = blockIdx.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t HW = (int64_t)H * W, BHW = (int64_t)B * HW;
    const int q4 = lane >> 4, rr = lane & 15;
    for (int e = 0; e < K; ++e)
        for (int cc = wave * 4 + q4; cc < kAC; cc += 16) {
            const int c = c0 + cc, r = r0 + rr;
            if (c < W && r < H) tile[e][rr][cc] = dmu_rot[e * BHW + b * HW + mframe(r, c, H, W, side)];
        }
    __syncthreads();
    for (int r_ = wave; r_ < kAR; r_ += 4) {
        const int r = r0 + r_, c = c0 + lane;
        if (r >= H || c >= W) continue;
        for (int j = 0; j < Kc; ++j) {
            float v = 0.0f;
            for (int e = 0; e < K; ++e) v = fmaf(coef[j * K + e], tile[e][r_][lane], v);
            const int64_t o = j * BHW + b * HW + (int64_t)r * W + c;
            if (dmux && j < K) v = fmaf(coef_x[j], dmux[o], v);
            dF[o] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// dC[n] (+)= Rot^T(dF[n]): gather.  Input pixel (y, x) receives from the output
// pixels whose sample point has floor in {y-1, y} x {x-1, x}; they lie within
// the rotated 2x2 box around its preimage.  Each candidate's sample point is
// recomputed with the forward arithmetic (rot_sample), so the weights are those
// of the forward kernel exactly.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rotate_adjoint_kernel(const float *__restrict__ dF, float *__restrict__ dC,
                                                             int64_t N, int H, int W, float c, float s, float rj,
                                                             float ri, int accumulate)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    // preimage of the pixel centre in output pixel coordinates
    const float gx = (float)(2 * x + 1) / (float)W - 1.0f;
    const float gy = (float)(2 * y + 1) / (float)H - 1.0f;
    const float xn = c * gx + s * gy;
    const float yn = -s * gx + c * gy;
    const float jc = ((xn + 1.0f) * W - 1.0f) * 0.5f;
    const float ic = ((yn + 1.0f) * H - 1.0f) * 0.5f;
    const int jlo = max(0, (int)ceilf(jc - rj)), jhi = min(W - 1, (int)floorf(jc + rj));
    const int ilo = max(0, (int)ceilf(ic - ri)), ihi = min(H - 1, (int)floorf(ic + ri));
    // candidate weights (independent of the plane n)
    constexpr int kMaxCand = 25;
    int idx[kMaxCand];
    float wt[kMaxCand];
    int nc = 0;
    for (int i = ilo; i <= ihi; ++i)
        for (int j = jlo; j <= jhi; ++j) {
            int y0, x0;
            float fy, fx;
            rot_sample(i, j, H, W, c, s, y0, x0, fy, fx);
            const float wy = y0 == y ? 1.0f - fy : (y0 + 1 == y ? fy : 0.0f);
            const float wx = x0 == x ? 1.0f - fx : (x0 + 1 == x ? fx : 0.0f);
            const float wgt = wy * wx;
            if (wgt != 0.0f && nc < kMaxCand) {
                idx[nc] = i * W + j;
                wt[nc] = wgt;
                ++nc;
            }
        }
    const int64_t HW = (int64_t)H * W;
    for (int64_t n = blockIdx.z; n < N; n += gridDim.z) {
        const float *src = dF + n * HW;
        float v = 0.0f;
        for (int k = 0; k < nc; ++k) v = fmaf(wt[k], src[idx[k]], v);
        float *o = dC + n * HW + (int64_t)y * W + x;
        *o = accumulate ? *o + v : v;
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
int launch_sample_grad(const float *C, const float *em, const float *theta, const float *G, float *out, int64_t A,
                       int64_t B, int64_t R, int64_t H, int64_t W, int64_t out_astride, hipStream_t s)
{
    FL_REQUIRE(A * B <= 65535, "sample_grad: A*B must be <= 65535");
    dim3 grid((unsigned)((W + 63) / 64), (unsigned)((H + 3) / 4), (unsigned)(A * B));
    hipLaunchKernelGGL(sample_grad_kernel, grid, dim3(256), 0, s, C, em, theta, G, out, (int)A, (int)B, (int)R,
                       (int)H, (int)W, out_astride);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_xrf_grad_prep(const float *T_rot, const float *Ax, const float *gA, float *dT_rot, float *gx,
                         float *A_tot, int64_t A_tot_estride, int64_t K, int64_t B, int64_t H, int64_t W, int side,
                         hipStream_t s)
{
    FL_REQUIRE(K * B <= 65535, "xrf_grad_prep: K*B must be <= 65535");
    dim3 grid((unsigned)((W + kAC - 1) / kAC), (unsigned)((H + kAR - 1) / kAR), (unsigned)(K * B));
    hipLaunchKernelGGL(xrf_grad_prep_kernel, grid, dim3(256), 0, s, T_rot, Ax, gA, dT_rot, gx, A_tot, A_tot_estride,
                       (int)B, (int)H, (int)W, side);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_incident_grad(const float *F, const float *coef_x, const float *Ax, const float *gx, float *dmux,
                         int64_t K, int64_t B, int64_t H, int64_t W, hipStream_t s)
{
    dim3 grid((unsigned)((W + 255) / 256), (unsigned)B, (unsigned)K);
    hipLaunchKernelGGL(incident_grad_kernel, grid, dim3(256), 0, s, F, coef_x, Ax, gx, dmux, (int)B, (int)H, (int)W);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_mu_adjoint(const float *dmu_rot, const float *dmux, const float *coef, const float *coef_x, float *dF,
                      int64_t Kc, int64_t K, int64_t B, int64_t H, int64_t W, int side, hipStream_t s)
{
    FL_REQUIRE(K <= kAMaxLines, "mu_adjoint: at most 8 emission lines");
    dim3 grid((unsigned)((W + kAC - 1) / kAC), (unsigned)((H + kAR - 1) / kAR), (unsigned)B);
    hipLaunchKernelGGL(mu_adjoint_kernel, grid, dim3(256), 0, s, dmu_rot, dmux, coef, coef_x, dF, (int)Kc, (int)K,
                       (int)B, (int)H, (int)W, side);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_rotate_adjoint(const float *dF, float *dC, int64_t N, int64_t H, int64_t W, float c, float s_,
                          int accumulate, hipStream_t s)
{
    // half-widths of the candidate box around the preimage, in output pixels, plus a margin
    const float ac = std::fabs(c), as = std::fabs(s_);
    const float rj = ac + as * (float)W / (float)H + 0.01f;
    const float ri = ac + as * (float)H / (float)W + 0.01f;
    FL_REQUIRE((2 * rj + 1) * (2 * ri + 1) <= 25.0f, "rotate_adjoint: aspect ratio too large");
    // The candidate weights depend only on the pixel: each block finds them once
    // and applies them to N / grid.z planes.  grid.z is as small as keeps about
    // kBlocks blocks in flight (enough waves to hide the loads' latency).
    constexpr int64_t kBlocks = 16384;
    const int64_t nxy = (int64_t)((W + 63) / 64) * ((H + 3) / 4);
    const int64_t gz = std::max<int64_t>(1, std::min<int64_t>(N, (kBlocks + nxy - 1) / nxy));
    dim3 grid((unsigned)((W + 63) / 64), (unsigned)((H + 3) / 4), (unsigned)std::min<int64_t>(gz, 65535));
    hipLaunchKernelGGL(rotate_adjoint_kernel, grid, dim3(256), 0, s, dF, dC, N, (int)H, (int)W, c, s_, rj, ri,
                       accumulate);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

}  // namespace fl
