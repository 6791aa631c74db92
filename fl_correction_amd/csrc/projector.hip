// Rotated, attenuation-weighted emission projector (forward) and its exact
// transpose (back-projection) for CDNA4.
//
// Geometry shared with the reference (FL/cuda_lib/atten_cuda_kernel.cu):
//   detector point (a, h, w) samples the image at
//     row x = c*(h-cy) - s*(w-cx) + cy,  col y = s*(h-cy) + c*(w-cx) + cx,
//   c = cos(-theta_a), s = sin(-theta_a), cx = (W-1)/2, cy = (H-1)/2   (.cu:313-321)
//   clamped to [0, H-1.001] x [0, W-1.001] and read bilinearly        (.cu:271-296)
//   Pf[a,b,w] = sum_h atten[a,b,h,w] * sum_r em[a,r] * C_r(x, y)       (.cu:298-337)
//   out[r,b]  = transpose of the same sampling, weighted by em*atten*ratio (.cu:390-446)
//
// Forward: one workgroup owns a 64x64 detector tile of one slice and a group of
// angles.  The image-space footprint of the tile (the rotated tile's bounding
// box, unioned over as many consecutive angles as fit) is staged in LDS once;
// every bilinear tap is then an LDS read.  Each of the 4 waves sums 16 rows per
// lane (lane = detector column) in registers; the 4 wave partials are summed in
// a fixed order and written as one partial per 64-row tile.  A second small
// kernel sums the ceil(H/64) partials in order.  No atomics: deterministic.
//
// Back-projection: one wave owns a 32x64 image tile of one slice and keeps its
// accumulator (plus a one-voxel halo) privately in LDS.  For every angle it
// walks the detector points whose clamped sample lands in the tile, in 8x8
// blocks (empty blocks skipped by a wave ballot), and splats the four bilinear
// weights into its own accumulator.  No global atomics; nothing is shared
// between waves, so the result does not depend on wave scheduling.  The MLEM
// update can be fused into the write-out.
#include "common.hpp"

#include <algorithm>
#include <climits>
#include <vector>

namespace fl {

// ---------------------------------------------------------------------------
// Forward projector
// ---------------------------------------------------------------------------
constexpr int kFTileW = 64;
constexpr int kFTileH = 64;
constexpr int kFWaves = 4;
constexpr int kFRowsPerWave = kFTileH / kFWaves;
// Upper bound of one angle's staged footprint for a 64x64 tile: the rotated
// tile spans at most 63*sqrt(2) < 89.1 rows/cols, plus the 0.5 margins and the
// bilinear neighbour: <= 92 rows, pitch <= 93.
constexpr int kFMaxFootprint = 93 * 93;

struct FwdArgs {
    const float *atten;
    const float *C;
    const float *em;
    const float *theta;
    float *part;  // (nht, A, B, W)
    int A, B, R, H, W;
    int r0, nref;
    int accumulate;
    int nht, ntw, n_groups, group;
    int lds_floats;
};

struct Box {
    int rlo, rhi, clo, chi;
};

// Footprint of detector rows [h0,h1) x cols [w0,w1) at one angle, in clamped
// image coordinates, widened by 0.5 and by the bilinear neighbour.
__device__ __forceinline__ Box fwd_footprint(float c, float s, int h0, int h1, int w0, int w1,
                                             float cx, float cy, float xmax, float ymax)
{
    const float hs[2] = {(float)h0 - cy, (float)(h1 - 1) - cy};
    const float ws[2] = {(float)w0 - cx, (float)(w1 - 1) - cx};
    float xlo = 3.0e38f, xhi = -3.0e38f, ylo = 3.0e38f, yhi = -3.0e38f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float x = fmaf(c, hs[i], fmaf(-s, ws[j], cy));
            const float y = fmaf(s, hs[i], fmaf(c, ws[j], cx));
            xlo = fminf(xlo, x);
            xhi = fmaxf(xhi, x);
            ylo = fminf(ylo, y);
            yhi = fmaxf(yhi, y);
        }
    Box b;
    b.rlo = (int)clampf(xlo - 0.5f, 0.f, xmax);
    b.rhi = (int)clampf(xhi + 0.5f, 0.f, xmax) + 1;
    b.clo = (int)clampf(ylo - 0.5f, 0.f, ymax);
    b.chi = (int)clampf(yhi + 0.5f, 0.f, ymax) + 1;
    return b;
}

__global__ __launch_bounds__(256) void fwd_tile_kernel(FwdArgs p)
{
    extern __shared__ float lds[];
    float *red = lds + p.lds_floats;  // [2][kFWaves][64]

    uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const int tw = t % p.ntw;
    t /= p.ntw;
    const int th = t % p.nht;
    t /= p.nht;
    const int g = t % p.n_groups;
    const int b = t / p.n_groups;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int w0 = tw * kFTileW, w1 = min(w0 + kFTileW, p.W);
    const int h0 = th * kFTileH, h1 = min(h0 + kFTileH, p.H);
    const int w = w0 + lane;
    const bool wvalid = w < p.W;
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float xmax = p.H - 1.001f, ymax = p.W - 1.001f;
    const int64_t HW = (int64_t)p.H * p.W;
    const int a_begin = g * p.group, a_end = min(a_begin + p.group, p.A);
    // Lanes past the last column sample the last valid column (their atten is 0).
    const float wc = (float)min(w, w1 - 1) - cx;
    const int hw0 = h0 + wave * kFRowsPerWave;

    int a = a_begin;
    while (a < a_end) {
        // Union footprint of the longest run of angles that fits the LDS budget.
        Box u;
        int a_hi = a;
        {
            const float th0 = finite_or_zero(p.theta[a]);
            u = fwd_footprint(cosf(-th0), sinf(-th0), h0, h1, w0, w1, cx, cy, xmax, ymax);
            ++a_hi;
            while (a_hi < a_end) {
                const float tha = finite_or_zero(p.theta[a_hi]);
                Box nb = fwd_footprint(cosf(-tha), sinf(-tha), h0, h1, w0, w1, cx, cy, xmax, ymax);
                Box m{min(u.rlo, nb.rlo), max(u.rhi, nb.rhi), min(u.clo, nb.clo), max(u.chi, nb.chi)};
                const int pitch = (m.chi - m.clo + 1) | 1;
                if ((m.rhi - m.rlo + 1) * pitch * p.nref > p.lds_floats) break;
                u = m;
                ++a_hi;
            }
        }
        const int rlo = uniform_int(u.rlo), clo = uniform_int(u.clo);
        const int rows = uniform_int(u.rhi) - rlo + 1;
        const int cols = uniform_int(u.chi) - clo + 1;
        const int pitch = cols | 1;
        const int fsz = rows * pitch;

        __syncthreads();  // previous footprint and partial buffers are free
        for (int rr = 0; rr < p.nref; ++rr) {
            const float *Cs = p.C + ((int64_t)(p.r0 + rr) * p.B + b) * HW;
            float *dst = lds + rr * fsz;
            for (int r = wave; r < rows; r += kFWaves) {
                const float *src = Cs + (int64_t)(rlo + r) * p.W + clo;
                for (int cc = lane; cc < cols; cc += 64) dst[r * pitch + cc] = src[cc];
            }
        }
        __syncthreads();

        for (int aa = a; aa < a_hi; ++aa) {
            const float tha = finite_or_zero(p.theta[aa]);
            const float c = cosf(-tha), s = sinf(-tha);
            const float bx = fmaf(-s, wc, cy);
            const float by = fmaf(c, wc, cx);
            const float *ap = p.atten + ((int64_t)aa * p.B + b) * HW + (int64_t)hw0 * p.W + w;
            float av[kFRowsPerWave];
#pragma unroll
            for (int i = 0; i < kFRowsPerWave; ++i)
                av[i] = (wvalid && hw0 + i < h1) ? ap[(int64_t)i * p.W] : 0.0f;
            // Rows past the image edge sample the last valid row (times atten 0),
            // so every LDS index stays inside the staged footprint.

            float acc = 0.0f;
            if (p.nref == 1) {
                const float e = p.em[(int64_t)aa * p.R + p.r0];
#pragma unroll
                for (int i = 0; i < kFRowsPerWave; ++i) {
                    const float hc = (float)min(hw0 + i, h1 - 1) - cy;
                    const float x = clampf(fmaf(c, hc, bx), 0.f, xmax);
                    const float y = clampf(fmaf(s, hc, by), 0.f, ymax);
                    const int x0 = (int)x, y0 = (int)y;
                    const float fx = x - (float)x0, fy = y - (float)y0;
                    const int li = (x0 - rlo) * pitch + (y0 - clo);
                    const float v00 = lds[li], v01 = lds[li + 1];
                    const float v10 = lds[li + pitch], v11 = lds[li + pitch + 1];
                    const float top = fmaf(fy, v01 - v00, v00);
                    const float bot = fmaf(fy, v11 - v10, v10);
                    const float val = fmaf(fx, bot - top, top);
                    acc = fmaf(val, av[i], acc);
                }
                acc *= e;
            } else {
#pragma unroll
                for (int i = 0; i < kFRowsPerWave; ++i) {
                    const float hc = (float)min(hw0 + i, h1 - 1) - cy;
                    const float x = clampf(fmaf(c, hc, bx), 0.f, xmax);
                    const float y = clampf(fmaf(s, hc, by), 0.f, ymax);
                    const int x0 = (int)x, y0 = (int)y;
                    const float fx = x - (float)x0, fy = y - (float)y0;
                    const int li = (x0 - rlo) * pitch + (y0 - clo);
                    float sum = 0.0f;
                    for (int rr = 0; rr < p.nref; ++rr) {
                        const float *f = lds + rr * fsz;
                        const float v00 = f[li], v01 = f[li + 1];
                        const float v10 = f[li + pitch], v11 = f[li + pitch + 1];
                        const float top = fmaf(fy, v01 - v00, v00);
                        const float bot = fmaf(fy, v11 - v10, v10);
                        sum = fmaf(p.em[(int64_t)aa * p.R + p.r0 + rr], fmaf(fx, bot - top, top), sum);
                    }
                    acc = fmaf(sum, av[i], acc);
                }
            }
            float *rb = red + (aa & 1) * (kFWaves * 64);
            rb[wave * 64 + lane] = acc;
            __syncthreads();
            if (wave == 0 && wvalid) {
                const float tot = ((rb[lane] + rb[64 + lane]) + rb[128 + lane]) + rb[192 + lane];
                float *o = p.part + (((int64_t)th * p.A + aa) * p.B + b) * p.W + w;
                *o = p.accumulate ? *o + tot : tot;
            }
        }
        a = a_hi;
    }
}

// Pf[a,b,w] = sum_t part[t,a,b,w] in order t = 0..nht-1.
// mode 0: write Pf.  mode 1: write ratio = I / max(Pf, 1e-6) (atten_cuda.cpp:326-328).
__global__ __launch_bounds__(256) void fwd_reduce_kernel(const float *part, const float *I, float *out,
                                                         int64_t n, int nht, int mode)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = part[i];
    for (int t = 1; t < nht; ++t) s += part[(int64_t)t * n + i];
    if (mode == 1) {
        const float pf = s < 1e-6f ? 1e-6f : s;
        s = I[i] / pf;
    }
    out[i] = s;
}

// ---------------------------------------------------------------------------
// Back-projection (exact transpose) and the fused MLEM update
// ---------------------------------------------------------------------------
constexpr int kBTileI = 32;  // image rows per wave tile
constexpr int kBTileJ = 64;  // image cols per wave tile
constexpr int kBWaves = 4;   // one slice per wave
constexpr int kBPitch = kBTileJ + 3;  // halo of one on each side, odd pitch
constexpr int kBAcc = (kBTileI + 2) * kBPitch;

enum BwdMode { kBwdWrite = 0, kBwdSens = 1, kBwdUpdate = 2 };

struct BwdArgs {
    const float *atten;
    const float *ratio;  // (A,B,W) or null (= 1)
    const float *em;
    const float *theta;
    float *out;  // (R,B,H,W)
    int A, B, R, H, W;
    int r0, nref;
    int nti, ntj;
    int mode;
    const float *C;     // update: current estimate (R,B,H,W)
    const float *sens;  // update: sensitivity (R,B,H,W)
    float beta, delta;
};

__device__ __forceinline__ float huber_at(const float *Cs, int i, int j, int H, int W, float delta)
{
    // FL/cuda_lib/atten_cuda_kernel.cu:593-613: neighbours up, down, left, right
    const float Cp = Cs[(int64_t)i * W + j];
    float g = 0.0f;
    const int di[4] = {-1, 1, 0, 0};
    const int dj[4] = {0, 0, -1, 1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ii = i + di[k], jj = j + dj[k];
        if (ii < 0 || ii >= H || jj < 0 || jj >= W) continue;
        const float diff = Cp - Cs[(int64_t)ii * W + jj];
        g += fabsf(diff) <= delta ? diff : delta * copysignf(1.0f, diff);
    }
    return g;
}

__global__ __launch_bounds__(256) void bwd_tile_kernel(BwdArgs p)
{
    extern __shared__ float lds[];
    uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const int tj = t % p.ntj;
    t /= p.ntj;
    const int ti = t % p.nti;
    const int bg = t / p.nti;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int b = bg * kBWaves + wave;
    const bool active = b < p.B;
    const int I0 = ti * kBTileI, J0 = tj * kBTileJ;
    float *acc = lds + wave * p.nref * kBAcc;
    for (int k = lane; k < p.nref * kBAcc; k += 64) acc[k] = 0.0f;

    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float xmax = p.H - 1.001f, ymax = p.W - 1.001f;
    const int64_t HW = (int64_t)p.H * p.W;
    constexpr float BIG = 1.0e6f;
    // Clamped sample rows landing in [I0-1, I0+T): unclamped bounds, open where
    // the clamp folds the outside onto the tile.
    const float lox = I0 <= 1 ? -BIG : (float)(I0 - 1);
    const float hix = I0 + kBTileI >= p.H - 1 ? BIG : (float)(I0 + kBTileI);
    const float loy = J0 <= 1 ? -BIG : (float)(J0 - 1);
    const float hiy = J0 + kBTileJ >= p.W - 1 ? BIG : (float)(J0 + kBTileJ);
    const int lh = lane >> 3, lw = lane & 7;

    if (active) {
        for (int a = 0; a < p.A; ++a) {
            const float tha = finite_or_zero(p.theta[a]);
            const float c = cosf(tha), s = sinf(tha);
            // Preimage bounding box: invert the rotation at the region corners.
            float hmin = 3.0e38f, hmax = -3.0e38f, qmin = 3.0e38f, qmax = -3.0e38f;
            const float Xs[2] = {lox - cy, hix - cy};
            const float Ys[2] = {loy - cx, hiy - cx};
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const float hc = c * Xs[i] - s * Ys[j];
                    const float qc = s * Xs[i] + c * Ys[j];
                    hmin = fminf(hmin, hc);
                    hmax = fmaxf(hmax, hc);
                    qmin = fminf(qmin, qc);
                    qmax = fmaxf(qmax, qc);
                }
            const int hlo = uniform_int((int)fmaxf(floorf(hmin + cy) - 1.f, 0.f));
            const int hhi = uniform_int((int)fminf(ceilf(hmax + cy) + 1.f, (float)(p.H - 1)));
            const int qlo = uniform_int((int)fmaxf(floorf(qmin + cx) - 1.f, 0.f));
            const int qhi = uniform_int((int)fminf(ceilf(qmax + cx) + 1.f, (float)(p.W - 1)));
            const float *ap = p.atten + ((int64_t)a * p.B + b) * HW;
            const float *rp = p.ratio ? p.ratio + ((int64_t)a * p.B + b) * p.W : nullptr;
            const float *ep = p.em + (int64_t)a * p.R + p.r0;

            for (int hb = hlo; hb <= hhi; hb += 8) {
                const int h = hb + lh;
                const float hc = (float)h - cy;
                for (int qb = qlo; qb <= qhi; qb += 8) {
                    const int q = qb + lw;
                    const float qc = (float)q - cx;
                    // backward_emission_batch_kernel (.cu:416-421)
                    const float col_u = c * qc - s * hc + cx;
                    const float row_u = s * qc + c * hc + cy;
                    const bool in = h <= hhi && q <= qhi && row_u >= lox && row_u < hix &&
                                    col_u >= loy && col_u < hiy;
                    if (__builtin_amdgcn_ballot_w64(in) == 0) continue;
                    if (in) {
                        const float rc = clampf(row_u, 0.f, xmax);
                        const float cc = clampf(col_u, 0.f, ymax);
                        const int y0 = (int)rc, x0 = (int)cc;
                        const float dy = rc - (float)y0, dx = cc - (float)x0;
                        float v = ap[(int64_t)h * p.W + q];
                        if (rp) v *= rp[q];
                        const int li = (y0 - I0 + 1) * kBPitch + (x0 - J0 + 1);
                        const float a0 = 1.0f - dy, a1 = dy;
                        const float b0 = 1.0f - dx, b1 = dx;
                        for (int rr = 0; rr < p.nref; ++rr) {
                            const float wv = ep[rr] * v;
                            float *A_ = acc + rr * kBAcc + li;
                            const float wt = wv * a0, wb = wv * a1;
                            atomicAdd(A_, wt * b0);
                            atomicAdd(A_ + 1, wt * b1);
                            atomicAdd(A_ + kBPitch, wb * b0);
                            atomicAdd(A_ + kBPitch + 1, wb * b1);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (!active) return;

    // Write-out of the tile interior (lane = column).
    const int j = J0 + lane;
    for (int rr = 0; rr < p.nref; ++rr) {
        const int r = p.r0 + rr;
        const int64_t sb = ((int64_t)r * p.B + b) * HW;
        const float *A_ = acc + rr * kBAcc;
        for (int ii = 0; ii < kBTileI; ++ii) {
            const int i = I0 + ii;
            if (i >= p.H || j >= p.W || lane >= kBTileJ) continue;
            const float back = A_[(ii + 1) * kBPitch + lane + 1];
            const int64_t o = sb + (int64_t)i * p.W + j;
            if (p.mode == kBwdWrite) {
                p.out[o] = back;
            } else if (p.mode == kBwdSens) {
                p.out[o] = back < 1e-6f ? 1e-6f : back;  // sens.clamp_min_(1e-6)
            } else {
                // C.mul_(back).div_(sens + beta*grad); C.clamp_min_(0)  (atten_cuda.cpp:334-341)
                const float Cv = p.C[o];
                float den = p.sens[o];
                if (p.beta > 0.0f) den = den + p.beta * huber_at(p.C + sb, i, j, p.H, p.W, p.delta);
                const float v = (Cv * back) / den;
                p.out[o] = v < 0.0f ? 0.0f : v;
            }
        }
    }
}

__global__ __launch_bounds__(256) void huber_kernel(const float *C, float *grad, int64_t n_slices, int H,
                                                    int W, float delta)
{
    const int64_t HW = (int64_t)H * W;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_slices * HW) return;
    const int64_t sl = idx / HW;
    const int rem = (int)(idx - sl * HW);
    grad[idx] = huber_at(C + sl * HW, rem / W, rem % W, H, W, delta);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static int fwd_lds_floats() { return 10240; }  // 40 KB of footprint per workgroup

int launch_forward(const float *atten, const float *C, const float *em, const float *theta,
                   float *part, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W,
                   hipStream_t stream, int group)
{
    const int lds_floats = fwd_lds_floats();
    const int ref_chunk = std::max(1, lds_floats / kFMaxFootprint);
    FwdArgs p{};
    p.atten = atten;
    p.C = C;
    p.em = em;
    p.theta = theta;
    p.part = part;
    p.A = (int)A;
    p.B = (int)B;
    p.R = (int)R;
    p.H = (int)H;
    p.W = (int)W;
    p.nht = (int)((H + kFTileH - 1) / kFTileH);
    p.ntw = (int)((W + kFTileW - 1) / kFTileW);
    p.group = group;
    p.n_groups = (int)((A + group - 1) / group);
    p.lds_floats = lds_floats;
    const int64_t nwg = (int64_t)p.ntw * p.nht * p.n_groups * B;
    FL_REQUIRE(nwg < (int64_t)UINT32_MAX, "forward projector: grid too large");
    const size_t shmem = sizeof(float) * (lds_floats + 2 * kFWaves * 64);
    for (int64_t r0 = 0; r0 < R; r0 += ref_chunk) {
        p.r0 = (int)r0;
        p.nref = (int)std::min<int64_t>(ref_chunk, R - r0);
        p.accumulate = r0 > 0;
        hipLaunchKernelGGL(fwd_tile_kernel, dim3((uint32_t)nwg), dim3(256), shmem, stream, p);
        FL_LAUNCH_CHECK();
    }
    return FL_OK;
}

int launch_forward_reduce(const float *part, const float *I, float *out, int64_t A, int64_t B,
                          int64_t H, int64_t W, int mode, hipStream_t stream)
{
    const int64_t n = A * B * W;
    const int nht = (int)((H + kFTileH - 1) / kFTileH);
    if (n == 0) return FL_OK;
    hipLaunchKernelGGL(fwd_reduce_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, part,
                       I, out, n, nht, mode);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_backward(const float *atten, const float *ratio, const float *em, const float *theta,
                    float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode,
                    const float *C, const float *sens, float beta, float delta, hipStream_t stream)
{
    const int ref_chunk = 2;
    BwdArgs p{};
    p.atten = atten;
    p.ratio = ratio;
    p.em = em;
    p.theta = theta;
    p.out = out;
    p.A = (int)A;
    p.B = (int)B;
    p.R = (int)R;
    p.H = (int)H;
    p.W = (int)W;
    p.nti = (int)((H + kBTileI - 1) / kBTileI);
    p.ntj = (int)((W + kBTileJ - 1) / kBTileJ);
    p.mode = mode;
    p.C = C;
    p.sens = sens;
    p.beta = beta;
    p.delta = delta;
    const int64_t nwg = (int64_t)p.nti * p.ntj * ((B + kBWaves - 1) / kBWaves);
    FL_REQUIRE(nwg < (int64_t)UINT32_MAX, "back-projector: grid too large");
    static bool attr_set = false;
    if (!attr_set) {
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)bwd_tile_kernel,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(sizeof(float) * kBWaves * ref_chunk * kBAcc)));
        attr_set = true;
    }
    for (int64_t r0 = 0; r0 < R; r0 += ref_chunk) {
        p.r0 = (int)r0;
        p.nref = (int)std::min<int64_t>(ref_chunk, R - r0);
        const size_t shmem = sizeof(float) * (size_t)kBWaves * p.nref * kBAcc;
        hipLaunchKernelGGL(bwd_tile_kernel, dim3((uint32_t)nwg), dim3(256), shmem, stream, p);
        FL_LAUNCH_CHECK();
    }
    return FL_OK;
}

int launch_huber(const float *C, float *grad, int64_t n_slices, int64_t H, int64_t W, float delta,
                 hipStream_t stream)
{
    const int64_t n = n_slices * H * W;
    if (n == 0) return FL_OK;
    hipLaunchKernelGGL(huber_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, C, grad,
                       n_slices, (int)H, (int)W, delta);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

size_t forward_part_floats(int64_t A, int64_t B, int64_t H, int64_t W)
{
    return (size_t)((H + kFTileH - 1) / kFTileH) * A * B * W;
}

}  // namespace fl
