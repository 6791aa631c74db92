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
// Forward: one workgroup owns a 64-column detector strip of one slice for a
// group of 8 angles and keeps the 8 column sums of every lane (lane = detector
// column) in registers while it walks the strip's full height in 32-row chunks.
// For each chunk, the image footprint of the chunk (the rotated 64x32 tile's
// bounding box, unioned over as many of the 8 angles as fit) is staged in LDS,
// so every bilinear tap is an LDS read.  The 4 waves (8 rows each) are summed in
// a fixed order at the end and written out directly, or turned into the MLEM
// ratio I / max(Pf, 1e-6).  No atomics, no partial-sum buffer: deterministic.
//
// Back-projection: one workgroup owns a 31x63 image tile of one slice (the
// clamped samples that touch it fill a 32x64 region); each of its 4 waves keeps
// a private LDS accumulator (tile plus a one-voxel halo).  For every angle the
// waves take turns over the detector rows of the tile's preimage: each row's
// column interval comes from the rotated region's edges, and the row's points
// are one coalesced segment of atten (lane = column; two rows per instruction
// when rows are at most 32 long).  Four segments are loaded together, then the
// bilinear weights of the points that land in the tile are splatted into the
// wave's own accumulator.  No global atomics; the four accumulators are summed in
// a fixed order at the end, which is also where the MLEM update is applied.
#include "common.hpp"

#include <algorithm>
#include <climits>

namespace fl {

// ---------------------------------------------------------------------------
// Forward projector
// ---------------------------------------------------------------------------
constexpr int kFW = 64;                   // detector columns per workgroup (one per lane)
constexpr int kFH = 32;                   // detector rows per chunk
constexpr int kFWaves = 4;
constexpr int kFRows = kFH / kFWaves;     // rows per wave per chunk
constexpr int kFG = 8;                    // angles per workgroup (one accumulator each)
constexpr int kFLds = 11264;              // footprint floats staged per workgroup (44 KB)
// One angle's footprint for a 64x32 tile: the sample points span at most
// hypot(31, 63) < 70.3 rows (and cols); with the 0.5 margins and the bilinear
// neighbour that is <= 74 rows and a pitch <= 75.
constexpr int kFMaxFootprint = 74 * 75;
constexpr int kFRefChunk = kFLds / kFMaxFootprint;  // references staged together (2)

enum FwdMode { kFwdWrite = 0, kFwdAccumulate = 1, kFwdRatio = 2, kFwdAccumulateRatio = 3 };

struct FwdArgs {
    const float *atten;
    const float *C;
    const float *em;
    const float *theta;
    const float *I;  // ratio modes
    float *out;      // (A, B, W): Pf or ratio
    int A, B, R, H, W;
    int r0, nref;
    int mode;
    int ntw, n_groups;
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

__global__ __launch_bounds__(256, 2) void fwd_strip_kernel(FwdArgs p)
{
    extern __shared__ float lds[];
    __shared__ float trig[2 * kFG];
    __shared__ float emv[kFG * kFRefChunk];

    uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const int tw = t % p.ntw;
    t /= p.ntw;
    const int g = t % p.n_groups;
    const int b = t / p.n_groups;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int w0 = tw * kFW, w1 = min(w0 + kFW, p.W);
    const int w = w0 + lane;
    const bool wvalid = w < p.W;
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float xmax = p.H - 1.001f, ymax = p.W - 1.001f;
    const int64_t HW = (int64_t)p.H * p.W;
    const int a0 = g * kFG;
    const int na = min(kFG, p.A - a0);
    // Lanes past the last column sample the last valid column (their atten is 0),
    // so every LDS index stays inside the staged footprint.
    const float wc = (float)min(w, w1 - 1) - cx;

    if (threadIdx.x < kFG) {
        const int k = threadIdx.x;
        const float th = k < na ? finite_or_zero(p.theta[a0 + k]) : 0.0f;
        trig[2 * k] = cosf(-th);
        trig[2 * k + 1] = sinf(-th);
    }
    if (threadIdx.x < kFG * kFRefChunk) {
        const int k = threadIdx.x / kFRefChunk, rr = threadIdx.x % kFRefChunk;
        emv[threadIdx.x] = (k < na && rr < p.nref) ? p.em[(int64_t)(a0 + k) * p.R + p.r0 + rr] : 0.0f;
    }
    __syncthreads();

    float acc[kFG];
#pragma unroll
    for (int k = 0; k < kFG; ++k) acc[k] = 0.0f;

    for (int h0 = 0; h0 < p.H; h0 += kFH) {
        const int h1 = min(h0 + kFH, p.H);
        const int hr0 = h0 + wave * kFRows;
        // This chunk's atten rows for all angles of the group, issued before the
        // footprint staging so that their latency overlaps it.
        float av[kFG][kFRows];
#pragma unroll
        for (int k = 0; k < kFG; ++k) {
            const float *ap = p.atten + ((int64_t)(a0 + min(k, na - 1)) * p.B + b) * HW + (int64_t)hr0 * p.W + w;
#pragma unroll
            for (int i = 0; i < kFRows; ++i)
                av[k][i] = (k < na && wvalid && hr0 + i < h1) ? __builtin_nontemporal_load(ap + (int64_t)i * p.W)
                                                              : 0.0f;
        }
        int gi = 0;
        while (gi < na) {
            // Longest run of angles gi..gj-1 whose union footprint fits the budget.
            Box u = fwd_footprint(trig[2 * gi], trig[2 * gi + 1], h0, h1, w0, w1, cx, cy, xmax, ymax);
            int gj = gi + 1;
            while (gj < na) {
                Box nb = fwd_footprint(trig[2 * gj], trig[2 * gj + 1], h0, h1, w0, w1, cx, cy, xmax, ymax);
                Box m{min(u.rlo, nb.rlo), max(u.rhi, nb.rhi), min(u.clo, nb.clo), max(u.chi, nb.chi)};
                if ((m.rhi - m.rlo + 1) * ((m.chi - m.clo + 1) | 1) * p.nref > kFLds) break;
                u = m;
                ++gj;
            }
            const int rlo = uniform_int(u.rlo), clo = uniform_int(u.clo);
            const int rows = uniform_int(u.rhi) - rlo + 1;
            const int cols = uniform_int(u.chi) - clo + 1;
            const int pitch = cols | 1;
            const int fsz = rows * pitch;
            gj = uniform_int(gj);

            __syncthreads();  // everyone is done with the previous footprint
            // Staging: the footprint as a flat index range; eight loads in flight per thread.
            for (int rr = 0; rr < p.nref; ++rr) {
                const float *Cs = p.C + ((int64_t)(p.r0 + rr) * p.B + b) * HW + (int64_t)rlo * p.W + clo;
                float *dst = lds + rr * fsz;
                const int n = rows * cols;
                for (int base = threadIdx.x; base < n; base += 8 * 256) {
                    float v[8];
                    int off[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int idx = base + e * 256;
                        const int r = idx / cols, cc = idx - r * cols;
                        off[e] = r * pitch + cc;
                        v[e] = idx < n ? Cs[(int64_t)r * p.W + cc] : 0.0f;
                    }
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        if (base + e * 256 < n) dst[off[e]] = v[e];
                }
            }
            __syncthreads();

#pragma unroll
            for (int k = 0; k < kFG; ++k) {
                if (k < gi || k >= gj) continue;  // wave-uniform
                const float c = trig[2 * k], s = trig[2 * k + 1];
                const float bx = fmaf(-s, wc, cy);
                const float by = fmaf(c, wc, cx);
                float sum = 0.0f;
#pragma unroll
                for (int i = 0; i < kFRows; ++i) {
                    // Rows past the image edge sample the last valid row (times atten 0).
                    const float hc = (float)min(hr0 + i, h1 - 1) - cy;
                    const float x = clampf(fmaf(c, hc, bx), 0.f, xmax);
                    const float y = clampf(fmaf(s, hc, by), 0.f, ymax);
                    const int x0 = (int)x, y0 = (int)y;
                    const float fx = x - (float)x0, fy = y - (float)y0;
                    const int li = (x0 - rlo) * pitch + (y0 - clo);
                    if (p.nref == 1) {
                        const float v00 = lds[li], v01 = lds[li + 1];
                        const float v10 = lds[li + pitch], v11 = lds[li + pitch + 1];
                        const float top = fmaf(fy, v01 - v00, v00);
                        const float bot = fmaf(fy, v11 - v10, v10);
                        sum = fmaf(fmaf(fx, bot - top, top), av[k][i], sum);
                    } else {
                        float val = 0.0f;
                        for (int rr = 0; rr < p.nref; ++rr) {
                            const float *f = lds + rr * fsz;
                            const float v00 = f[li], v01 = f[li + 1];
                            const float v10 = f[li + pitch], v11 = f[li + pitch + 1];
                            const float top = fmaf(fy, v01 - v00, v00);
                            const float bot = fmaf(fy, v11 - v10, v10);
                            val = fmaf(emv[k * kFRefChunk + rr], fmaf(fx, bot - top, top), val);
                        }
                        sum = fmaf(val, av[k][i], sum);
                    }
                }
                acc[k] += sum;
            }
            gi = gj;
        }
    }

    // Fixed-order reduction of the four waves' partial sums.
    __syncthreads();
    float *red = lds;  // [kFG][kFWaves][64]
#pragma unroll
    for (int k = 0; k < kFG; ++k) red[(k * kFWaves + wave) * 64 + lane] = acc[k];
    __syncthreads();
    for (int idx = threadIdx.x; idx < na * 64; idx += 256) {
        const int k = idx >> 6, l = idx & 63;
        const int wo = w0 + l;
        if (wo >= p.W) continue;
        const float *rk = red + k * kFWaves * 64 + l;
        float tot = ((rk[0] + rk[64]) + rk[128]) + rk[192];
        if (p.nref == 1) tot *= emv[k * kFRefChunk];
        const int64_t o = ((int64_t)(a0 + k) * p.B + b) * p.W + wo;
        if (p.mode == kFwdAccumulate || p.mode == kFwdAccumulateRatio) tot += p.out[o];
        if (p.mode >= kFwdRatio) {
            // Pf.clamp_min_(1e-6); ratio = I / Pf   (atten_cuda.cpp:326-328)
            const float pf = tot < 1e-6f ? 1e-6f : tot;
            tot = p.I[o] / pf;
        }
        p.out[o] = tot;
    }
}

// ---------------------------------------------------------------------------
// Back-projection (exact transpose) and the fused MLEM update
// ---------------------------------------------------------------------------
constexpr int kBI = 31;               // image rows per tile (clamped-sample region: 32 rows)
constexpr int kBJ = 63;               // image cols per tile (region: 64 cols = one wave)
constexpr int kBWaves = 4;            // private accumulators per workgroup
constexpr int kBPitch = kBJ + 2;      // 65: region cols plus the right neighbour (odd)
constexpr int kBAcc = (kBI + 2) * kBPitch;
constexpr int kBRefChunk = 2;
constexpr int kBUnroll = 4;           // row segments loaded together

enum BwdMode { kBwdWrite = 0, kBwdSens = 1, kBwdUpdate = 2, kBwdAccumulate = 3 };

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

// Range of qc with lo <= slope*qc + base < hi, intersected into [qmin, qmax];
// widened by one on each side (the exact test is made per point).
__device__ __forceinline__ void slope_interval(float slope, float lo, float hi, float base, float &qmin,
                                               float &qmax)
{
    const float L = lo - base, U = hi - base;  // need L <= slope*qc < U
    if (slope > 1e-6f) {
        qmin = fmaxf(qmin, L / slope - 1.0f);
        qmax = fminf(qmax, U / slope + 1.0f);
    } else if (slope < -1e-6f) {
        qmin = fmaxf(qmin, U / slope - 1.0f);
        qmax = fminf(qmax, L / slope + 1.0f);
    } else if (!(L <= 1e-3f && U > -1e-3f)) {
        qmin = 1.0f;
        qmax = 0.0f;  // empty
    }
}

__global__ __launch_bounds__(256) void bwd_tile_kernel(BwdArgs p)
{
    extern __shared__ float lds[];
    uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
    const int tj = t % p.ntj;
    t /= p.ntj;
    const int ti = t % p.nti;
    const int b = t / p.nti;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int I0 = ti * kBI, J0 = tj * kBJ;
    float *acc = lds + wave * p.nref * kBAcc;
    for (int k = lane; k < p.nref * kBAcc; k += 64) acc[k] = 0.0f;

    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float xmax = p.H - 1.001f, ymax = p.W - 1.001f;
    const int64_t HW = (int64_t)p.H * p.W;
    constexpr float BIG = 1.0e6f;
    // Unclamped sample rows whose clamped value lands in [I0-1, I0+kBI): open on
    // a side where the clamp folds everything beyond onto the tile.
    const float lox = I0 <= 1 ? -BIG : (float)(I0 - 1);
    const float hix = I0 + kBI >= p.H - 1 ? BIG : (float)(I0 + kBI);
    const float loy = J0 <= 1 ? -BIG : (float)(J0 - 1);
    const float hiy = J0 + kBJ >= p.W - 1 ? BIG : (float)(J0 + kBJ);
    const bool border = lox < 0.0f || hix > (float)p.H || loy < 0.0f || hiy > (float)p.W;

    for (int a = 0; a < p.A; ++a) {
        const float tha = finite_or_zero(p.theta[a]);
        const float c = cosf(tha), s = sinf(tha);
        // backward_emission_batch_kernel (.cu:416-417):
        //   col = c*(q-cx) - s*(h-cy) + cx,  row = s*(q-cx) + c*(h-cy) + cy
        // Preimage rows: invert the rotation at the region's corners.
        float hmin = 3.0e38f, hmax = -3.0e38f;
        {
            const float Xs[2] = {lox - cy, hix - cy};
            const float Ys[2] = {loy - cx, hiy - cx};
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const float hc = c * Xs[i] - s * Ys[j];
                    hmin = fminf(hmin, hc);
                    hmax = fmaxf(hmax, hc);
                }
        }
        const int hlo = uniform_int((int)clampf(floorf(hmin + cy) - 1.f, 0.f, (float)p.H));
        const int hhi = uniform_int((int)clampf(ceilf(hmax + cy) + 1.f, -1.f, (float)(p.H - 1)));
        const float *ap = p.atten + ((int64_t)a * p.B + b) * HW;
        const float *rp = p.ratio ? p.ratio + ((int64_t)a * p.B + b) * p.W : nullptr;
        float ew[kBRefChunk];
#pragma unroll
        for (int rr = 0; rr < kBRefChunk; ++rr) ew[rr] = rr < p.nref ? p.em[(int64_t)a * p.R + p.r0 + rr] : 0.0f;

        // This wave's rows: h = hlo + wave + 4k.  Lane k computes row k's column interval.
        for (int kb = 0; hlo + wave + 4 * kb <= hhi; kb += 64) {
            const int hk = hlo + wave + 4 * (kb + lane);
            float qmin = -cx - 1.0f, qmax = (float)(p.W - 1) - cx + 1.0f;
            {
                const float hc = (float)hk - cy;
                slope_interval(s, lox, hix, fmaf(c, hc, cy), qmin, qmax);
                slope_interval(c, loy, hiy, fmaf(-s, hc, cx), qmin, qmax);
            }
            const int qlk = (int)clampf(floorf(qmin + cx), 0.f, (float)p.W);
            const int qhk = hk <= hhi ? (int)clampf(ceilf(qmax + cx), -1.f, (float)(p.W - 1)) : -1;
            const int lenk = max(qhk - qlk + 1, 0);
            const int nrows = min(64, (hhi - (hlo + wave + 4 * kb)) / 4 + 1);
            // longest row of the batch decides the lane layout: two rows per
            // instruction (32 lanes each) when every row fits in 32 lanes
            int maxlen = lenk;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, off));
            maxlen = uniform_int(maxlen);
            if (maxlen == 0) continue;
            const bool half = maxlen <= 32;
            const int rows_per = half ? 2 : 1;
            const int nch = half ? 1 : (maxlen + 63) / 64;
            const int lsub = half ? (lane & 31) : lane;
            const int lrow = half ? (lane >> 5) : 0;
            const int nitems = ((nrows + rows_per - 1) / rows_per) * nch;

            for (int it0 = 0; it0 < nitems; it0 += kBUnroll) {
                float av[kBUnroll], rv[kBUnroll], X[kBUnroll], Y[kBUnroll];
                bool in[kBUnroll];
#pragma unroll
                for (int u = 0; u < kBUnroll; ++u) {
                    const int it = it0 + u;
                    const int rowi = (it / nch) * rows_per + lrow;  // row index within the batch
                    const int ch = it % nch;
                    const int rsel = min(rowi, 63);
                    const int ql = __shfl(qlk, rsel), qh = __shfl(qhk, rsel);
                    const int h = hlo + wave + 4 * (kb + rowi);
                    const int q = ql + ch * 64 + lsub;
                    const float hc = (float)h - cy, qc = (float)q - cx;
                    X[u] = fmaf(s, qc, fmaf(c, hc, cy));
                    Y[u] = fmaf(c, qc, fmaf(-s, hc, cx));
                    in[u] = it < nitems && rowi < nrows && q <= qh && X[u] >= lox && X[u] < hix && Y[u] >= loy &&
                            Y[u] < hiy;
                    av[u] = in[u] ? ap[(int64_t)h * p.W + q] : 0.0f;
                    rv[u] = (in[u] && rp) ? rp[q] : 1.0f;
                }
                // Splat.  Float LDS atomics are slow on this hardware, so the accumulator is
                // updated with plain read-add-write, ordered so that no two lanes touch
                // one word in the same instruction: along a detector row two points two
                // apart differ by (2s, 2c), at least 1.41 in one coordinate, so their
                // floor cells differ -- even-column lanes first, then odd ones, one
                // corner at a time.  Segments of different rows are >= 4 rows apart and
                // cannot meet.  Samples moved by the clamp (border tiles only) can
                // collapse onto one cell; they go through atomics in a third phase.
                int li[kBUnroll];
                float wq[kBUnroll][4];
                bool fast[kBUnroll], slow[kBUnroll], odd[kBUnroll];
#pragma unroll
                for (int u = 0; u < kBUnroll; ++u) {
                    const float Xc = clampf(X[u], 0.f, xmax), Yc = clampf(Y[u], 0.f, ymax);
                    const bool clamped = border && (Xc != X[u] || Yc != Y[u]);
                    const int y0 = (int)Xc, x0 = (int)Yc;
                    const float dy = Xc - (float)y0, dx = Yc - (float)x0;
                    const float v = av[u] * rv[u];
                    li[u] = in[u] ? (y0 - I0 + 1) * kBPitch + (x0 - J0 + 1) : 0;
                    const float wt = v * (1.0f - dy), wb = v * dy;
                    wq[u][0] = wt * (1.0f - dx);
                    wq[u][1] = wt * dx;
                    wq[u][2] = wb * (1.0f - dx);
                    wq[u][3] = wb * dx;
                    fast[u] = in[u] && !clamped;
                    slow[u] = in[u] && clamped;
                    odd[u] = (lsub & 1) != 0;
                }
                const int offs[4] = {0, 1, kBPitch, kBPitch + 1};
#pragma unroll
                for (int rr = 0; rr < kBRefChunk; ++rr) {
                    if (rr >= p.nref) break;
                    float *A_ = acc + rr * kBAcc;
#pragma unroll
                    for (int ph = 0; ph < 2; ++ph) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            float cur[kBUnroll];
#pragma unroll
                            for (int u = 0; u < kBUnroll; ++u)
                                cur[u] = (fast[u] && odd[u] == (ph == 1)) ? A_[li[u] + offs[k]] : 0.0f;
#pragma unroll
                            for (int u = 0; u < kBUnroll; ++u)
                                if (fast[u] && odd[u] == (ph == 1)) A_[li[u] + offs[k]] = fmaf(ew[rr], wq[u][k], cur[u]);
                            __asm__ volatile("" ::: "memory");
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kBUnroll; ++u)
                        if (slow[u])
#pragma unroll
                            for (int k = 0; k < 4; ++k) atomicAdd(A_ + li[u] + offs[k], ew[rr] * wq[u][k]);
                    __asm__ volatile("" ::: "memory");
                }
            }
        }
    }
    __syncthreads();

    // Sum the four private accumulators in a fixed order and write the tile (lane = column).
    const int j = J0 + lane;
    for (int rr = 0; rr < p.nref; ++rr) {
        const int r = p.r0 + rr;
        const int64_t sb = ((int64_t)r * p.B + b) * HW;
        for (int ii = wave; ii < kBI; ii += kBWaves) {
            const int i = I0 + ii;
            if (i >= p.H || j >= p.W || lane >= kBJ) continue;
            const int li = rr * kBAcc + (ii + 1) * kBPitch + lane + 1;
            const int st = p.nref * kBAcc;
            const float back = ((lds[li] + lds[st + li]) + lds[2 * st + li]) + lds[3 * st + li];
            const int64_t o = sb + (int64_t)i * p.W + j;
            if (p.mode == kBwdWrite) {
                p.out[o] = back;
            } else if (p.mode == kBwdAccumulate) {
                p.out[o] += back;
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

// mode: kFwdWrite (Pf) or kFwdRatio (ratio = I / max(Pf, 1e-6))
int launch_forward(const float *atten, const float *C, const float *em, const float *theta, const float *I,
                   float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode, hipStream_t stream)
{
    static bool attr_set = false;
    const size_t shmem = sizeof(float) * kFLds;
    if (!attr_set) {
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)fwd_strip_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)shmem));
        attr_set = true;
    }
    FwdArgs p{};
    p.atten = atten;
    p.C = C;
    p.em = em;
    p.theta = theta;
    p.I = I;
    p.out = out;
    p.A = (int)A;
    p.B = (int)B;
    p.R = (int)R;
    p.H = (int)H;
    p.W = (int)W;
    p.ntw = (int)((W + kFW - 1) / kFW);
    p.n_groups = (int)((A + kFG - 1) / kFG);
    const int64_t nwg = (int64_t)p.ntw * p.n_groups * B;
    FL_REQUIRE(nwg < (int64_t)UINT32_MAX, "forward projector: grid too large");
    for (int64_t r0 = 0; r0 < R; r0 += kFRefChunk) {
        p.r0 = (int)r0;
        p.nref = (int)std::min<int64_t>(kFRefChunk, R - r0);
        const bool first = r0 == 0, last = r0 + kFRefChunk >= R;
        const bool ratio = mode == kFwdRatio;
        p.mode = first ? (last && ratio ? kFwdRatio : kFwdWrite)
                       : (last && ratio ? kFwdAccumulateRatio : kFwdAccumulate);
        hipLaunchKernelGGL(fwd_strip_kernel, dim3((uint32_t)nwg), dim3(256), shmem, stream, p);
        FL_LAUNCH_CHECK();
    }
    return FL_OK;
}

int launch_backward(const float *atten, const float *ratio, const float *em, const float *theta,
                    float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode,
                    const float *C, const float *sens, float beta, float delta, hipStream_t stream)
{
    static bool attr_set = false;
    if (!attr_set) {
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)bwd_tile_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(sizeof(float) * kBWaves * kBRefChunk * kBAcc)));
        attr_set = true;
    }
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
    p.nti = (int)((H + kBI - 1) / kBI);
    p.ntj = (int)((W + kBJ - 1) / kBJ);
    p.mode = mode;
    p.C = C;
    p.sens = sens;
    p.beta = beta;
    p.delta = delta;
    const int64_t nwg = (int64_t)p.nti * p.ntj * B;
    FL_REQUIRE(nwg < (int64_t)UINT32_MAX, "back-projector: grid too large");
    for (int64_t r0 = 0; r0 < R; r0 += kBRefChunk) {
        p.r0 = (int)r0;
        p.nref = (int)std::min<int64_t>(kBRefChunk, R - r0);
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

}  // namespace fl
