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
// base cells of the samples that touch it fill a 32x64 region); each of its 4
// waves keeps a private LDS accumulator (tile plus a one-voxel halo) and takes
// the angles a = wave (mod 4).  Per angle a wave enumerates exactly the
// detector points whose sample lands in the tile (rows of the tile's preimage,
// each row's column interval) and splats their bilinear weights; the samples
// the reference clamps onto the image border are gathered per border cell.  No
// global atomics; the four accumulators are summed in a fixed order at the
// end, which is also where the MLEM update is applied.
#include "projector.hpp"

#include <algorithm>
#include <climits>

namespace fl {

// ---------------------------------------------------------------------------
// Forward projector
// ---------------------------------------------------------------------------

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

// One workgroup's tile t: a strip of kFW detector columns, an angle group of G
// angles and a slice.  Slices go in runs of kFSliceRun, the slice fastest
// within a run, then the strip, then the angle group; slices past the last
// whole run follow with the strip fastest.  Workgroups that run together then
// share their strip's angles and footprints (their samples' positions are the
// same), a run's few slices of C stay in L2 while its strips and angle groups
// go by, and the slice-major output (A, W, B) is written in whole 32-byte
// sectors.  At 256^3 x 360: 5.17 vs 5.66 ms with the plain strip-fastest order.
constexpr int kFSliceRun = 8;
template <int NR, int G>
__device__ __forceinline__ void fwd_strip_tile(const FwdArgs &p, uint32_t t)
{
    extern __shared__ float lds[];
    __shared__ float trig[2 * G];
    __shared__ float emv[G * kFRefChunk];

    int tw, g, b;
    {
        const uint32_t per_slice = (uint32_t)p.ntw * (uint32_t)p.n_groups;
        const uint32_t full = (uint32_t)p.B / kFSliceRun * kFSliceRun;
        if (t < per_slice * full) {
            const uint32_t lo = t % kFSliceRun;
            t /= kFSliceRun;
            tw = (int)(t % (uint32_t)p.ntw);
            t /= (uint32_t)p.ntw;
            g = (int)(t % (uint32_t)p.n_groups);
            b = (int)((t / (uint32_t)p.n_groups) * kFSliceRun + lo);
        } else {
            t -= per_slice * full;
            tw = (int)(t % (uint32_t)p.ntw);
            t /= (uint32_t)p.ntw;
            g = (int)(t % (uint32_t)p.n_groups);
            b = (int)(full + t / (uint32_t)p.n_groups);
        }
    }

    const int lane = threadIdx.x & 63;
    const int wave = uniform_int(threadIdx.x >> 6);
    const int w0 = tw * kFW, w1 = min(w0 + kFW, p.W);
    const int w = w0 + lane;
    const bool wvalid = w < p.W;
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float xmax = p.H - 1.001f, ymax = p.W - 1.001f;
    const int64_t HW = (int64_t)p.H * p.W;
    const int a0 = g * G;
    const int na = min(G, p.A - a0);
    // Lanes past the last column sample the last valid column (their atten is 0),
    // so every LDS index stays inside the staged footprint.
    const float wc = (float)min(w, w1 - 1) - cx;

    if (threadIdx.x < G) {
        const int k = threadIdx.x;
        const float th = k < na ? finite_or_zero(p.theta[a0 + k]) : 0.0f;
        trig[2 * k] = cosf(-th);
        trig[2 * k + 1] = sinf(-th);
    }
    if (threadIdx.x < G * kFRefChunk) {
        const int k = threadIdx.x / kFRefChunk, rr = threadIdx.x % kFRefChunk;
        emv[threadIdx.x] = (k < na && rr < NR) ? p.em[(int64_t)(a0 + k) * p.R + p.r0 + rr] : 0.0f;
    }
    __syncthreads();

    float acc[G];
#pragma unroll
    for (int k = 0; k < G; ++k) acc[k] = 0.0f;

    for (int h0 = 0; h0 < p.H; h0 += kFH) {
        const int h1 = min(h0 + kFH, p.H);
        const int hr0 = h0 + wave * kFRows;
        // This chunk's atten rows for all angles of the group, issued before the
        // footprint staging so that their latency overlaps it.
        // Bounds-checked buffer loads: rows past the image and lanes past the
        // last column (offset moved out of range) read 0; the row offset is a
        // scalar operand, so no vector address arithmetic per load.
        float av[G][kFRows];
        const uint32_t voff = wvalid ? (uint32_t)(hr0 * p.W + w) * 4u : 0x80000000u;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const __amdgpu_buffer_rsrc_t ra =
                make_rsrc(p.atten + ((int64_t)(a0 + min(k, na - 1)) * p.B + b) * HW, k < na ? (int)HW : 0);
#pragma unroll
            for (int i = 0; i < kFRows; ++i) av[k][i] = load_f32_stream(ra, voff, (uint32_t)(i * p.W) * 4u);
        }
        // Rows past the image edge sample the last valid row (times atten 0).
        float hcs[kFRows];
#pragma unroll
        for (int i = 0; i < kFRows; ++i) hcs[i] = (float)min(hr0 + i, h1 - 1) - cy;
        // Footprint of every angle of the group, lane k computing angle k's box.
        Box mine{0, 0, 0, 0};
        {
            const int k = min(lane, G - 1);
            mine = fwd_footprint(trig[2 * k], trig[2 * k + 1], h0, h1, w0, w1, cx, cy, xmax, ymax);
        }
        int gi = 0;
        while (gi < na) {
            // Longest run of angles gi..gj-1 whose union footprint fits the budget
            // and the staged width.
            int rlo = lane_value(mine.rlo, gi), rhi = lane_value(mine.rhi, gi);
            int clo = lane_value(mine.clo, gi), chi = lane_value(mine.chi, gi);
            int gj = gi + 1;
            while (gj < na) {
                const int mrlo = min(rlo, lane_value(mine.rlo, gj)), mrhi = max(rhi, lane_value(mine.rhi, gj));
                const int mclo = min(clo, lane_value(mine.clo, gj)), mchi = max(chi, lane_value(mine.chi, gj));
                if ((mrhi - mrlo + 1) * ((mchi - mclo + 1) | 1) > fwd_lds_ref(NR) || mchi - mclo + 1 > kFStageCols)
                    break;
                rlo = mrlo;
                rhi = mrhi;
                clo = mclo;
                chi = mchi;
                ++gj;
            }
            rlo = uniform_int(rlo);
            clo = uniform_int(clo);
            const int rows = uniform_int(rhi) - rlo + 1;
            const int cols = uniform_int(chi) - clo + 1;
            const int pitch = cols | 1;
            const int fsz = rows * pitch;
            gj = uniform_int(gj);

            __syncthreads();  // everyone is done with the previous footprint
            // Staging: wave w takes footprint rows w, w+4, ..., lane l the columns
            // l and l + 64 (cols <= kFStageCols); eight loads in flight per lane.
#pragma unroll
            for (int rr = 0; rr < NR; ++rr) {
                const float *Cs = p.C + ((int64_t)(p.r0 + rr) * p.B + b) * HW + (int64_t)rlo * p.W + clo;
                float *dst = lds + rr * fsz;
                for (int r0 = wave; r0 < rows; r0 += 4 * kFWaves) {
                    float v[4][2];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            const int r = r0 + e * kFWaves, cc = lane + 64 * k;
                            v[e][k] = (r < rows && cc < cols) ? Cs[r * p.W + cc] : 0.0f;
                        }
#pragma unroll
                    for (int e = 0; e < 4; ++e)
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            const int r = r0 + e * kFWaves, cc = lane + 64 * k;
                            if (r < rows && cc < cols) dst[r * pitch + cc] = v[e][k];
                        }
                }
            }
            __syncthreads();

            // Footprint cell (x0, y0) is LDS word x0 * pitch + y0 - fbase.
            const char *fb = reinterpret_cast<const char *>(lds) - 4 * (rlo * pitch + clo);
            const uint32_t pitch4 = 4u * (uint32_t)pitch;
#pragma unroll
            for (int k = 0; k < G; ++k) {
                if (k < gi || k >= gj) continue;  // wave-uniform
                // sample position of this lane's column at row h: (c*(h-cy) + bx, s*(h-cy) + by)
                const float c = trig[2 * k], s = trig[2 * k + 1];
                const float bx = fmaf(-s, wc, cy), by = fmaf(c, wc, cx);
                float sum = 0.0f;
#pragma unroll
                for (int i = 0; i < kFRows; ++i) {
                    const float x = clampf(fmaf(c, hcs[i], bx), 0.f, xmax);
                    const float y = clampf(fmaf(s, hcs[i], by), 0.f, ymax);
                    // positions are >= 0: (int) is floor, fract the exact remainder
                    const int x0 = (int)x, y0 = (int)y;
                    const float fx = __builtin_amdgcn_fractf(x), fy = __builtin_amdgcn_fractf(y);
                    const float *f =
                        reinterpret_cast<const float *>(fb + (__umul24((uint32_t)x0, pitch4) + 4u * (uint32_t)y0));
                    if constexpr (NR == 1) {
                        const float v00 = f[0], v01 = f[1];
                        const float v10 = f[pitch], v11 = f[pitch + 1];
                        const float top = fmaf(fy, v01 - v00, v00);
                        const float bot = fmaf(fy, v11 - v10, v10);
                        sum = fmaf(fmaf(fx, bot - top, top), av[k][i], sum);
                    } else {
                        float val = 0.0f;
#pragma unroll
                        for (int rr = 0; rr < NR; ++rr) {
                            const float *fr = f + rr * fsz;
                            const float v00 = fr[0], v01 = fr[1];
                            const float v10 = fr[pitch], v11 = fr[pitch + 1];
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
    float *red = lds;  // [G][kFWaves][64]
#pragma unroll
    for (int k = 0; k < G; ++k) red[(k * kFWaves + wave) * 64 + lane] = acc[k];
    __syncthreads();
    for (int idx = threadIdx.x; idx < na * 64; idx += 256) {
        const int k = idx >> 6, l = idx & 63;
        const int wo = w0 + l;
        if (wo >= p.W) continue;
        const float *rk = red + k * kFWaves * 64 + l;
        float tot = ((rk[0] + rk[64]) + rk[128]) + rk[192];
        if (p.nref == 1) tot *= emv[k * kFRefChunk];
        // (A, B, W), or slice-major (A, W, B) for the slice-major back-projector
        const int64_t o = p.out_sm ? ((int64_t)(a0 + k) * p.W + wo) * (p.out_pitch ? p.out_pitch : p.B) + b
                                   : ((int64_t)(a0 + k) * p.B + b) * p.W + wo;
        const int64_t oi = ((int64_t)(a0 + k) * p.B + b) * p.W + wo;
        if (p.mode == kFwdAccumulate || p.mode == kFwdAccumulateRatio) tot += p.out[o];
        if (p.mode >= kFwdRatio) {
            // Pf.clamp_min_(1e-6); ratio = I / Pf   (atten_cuda.cpp:326-328)
            const float pf = tot < 1e-6f ? 1e-6f : tot;
            tot = p.I[oi] / pf;
            if (p.fold_em) tot *= emv[k * kFRefChunk];
        }
        p.out[o] = tot;
    }
}

template <int NR, int G>
__global__ __launch_bounds__(256, 2) void fwd_strip_kernel(FwdArgs p)
{
    fwd_strip_tile<NR, G>(p, xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// Back-projection (exact transpose) and the fused MLEM update
// ---------------------------------------------------------------------------

// Range of qc with lo <= fmaf(slope, qc, base) < hi, intersected into [qmin, qmax].
// The real-arithmetic bounds are widened by 0.01 plus the distance the rounding
// of the fma (at most tol in value) can move them, so the range always contains
// every qc that passes the exact test; the caller tightens it point by point.
__device__ __forceinline__ void slope_interval(float slope, float lo, float hi, float base, float tol,
                                               float &qmin, float &qmax)
{
    const float L = lo - base, U = hi - base;
    const float as = fabsf(slope);
    if (as > 0.0f) {
        // approximate quotients: their error (below 1e-4 here) is far inside the
        // margin m
        const float r = __frcp_rn(slope);
        const float m = 0.01f + tol * fabsf(r);
        const float a = (slope > 0.0f ? L : U) * r, z = (slope > 0.0f ? U : L) * r;
        qmin = fmaxf(qmin, a - m);
        qmax = fminf(qmax, z + m);
    } else if (!(L <= 0.0f && U > 0.0f)) {
        qmin = 1.0f;
        qmax = 0.0f;  // empty
    }
}

__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Per-wave view of the back-projector's state for one tile.
struct BwdTile {
    int I0, J0, b, bi;         // bi: tile height
    int edges;                 // image edges whose border cells lie in the tile (kEdge*)
    float lox, hix, loy, hiy;  // region of unclamped sample positions
    float tol;                 // rounding allowance of one sample coordinate
    float *lds;                // dynamic LDS: accumulators (cell-major, NR values per cell), sinks
    int cbase;                 // image cell (y0, x0) -> accumulator cell y0 * kBPitch + x0 + cbase
    uint32_t *mask;            // row starts of the current window, two words per item
    int4 *rows;                // compacted rows of the current batch
    int lane, sink;            // sink: this lane's sink cell
};

// Sample stream of one wave.  Angles a = wave, wave+4, ...; for each angle the
// detector rows of the tile's preimage in two classes, h = hlo + cls + 2k,
// kBRows rows per batch (lane k holds row k).  Each row's in-region samples form one
// column interval, padded to an even count; a batch's samples, row after row,
// are numbered f = 0..T-1, cut into windows of kBWin and items of 128: lane l
// of item n holds the pair f = 128n+2l, 128n+2l+1 (points 0 and 1), always of
// one row.  A splat instruction uses one point per lane, all of one parity, so
// its samples lie >= 2 apart on one row or on rows >= 2 apart: >= 1.41 apart in
// one coordinate, so their base cells differ, and a plain read-add-write, one
// corner at a time, cannot lose an update (float LDS atomics are far slower on
// this hardware).  Only unclamped samples are enumerated; the clamped ones are
// gathered per border cell (bwd_edge).
struct BwdCursor {
    int a;
    int cls, rb, hlo, hhi;
    int T, nne, win0, nit, n, kb;
    int P;    // this lane's row: first sample number (batch setup)
    bool ne;  // this lane's row is not empty
    float c, s;
    __amdgpu_buffer_rsrc_t ra, rr;
    float ew[kBRefChunk];
};

// One item pair (256 samples, four per lane) between its loads and its splat.
struct BwdPair {
    float X[kBUnroll][2], Y[kBUnroll][2], av[kBUnroll][2], rv[kBUnroll][2];
    bool val[kBUnroll][2];
    float ew[kBRefChunk];
};

// c, s: cosf and sinf of the angle's theta.
template <int NR>
__device__ __forceinline__ void bwd_begin_angle(const BwdArgs &p, const BwdTile &t, BwdCursor &cur, float c, float s)
{
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    // backward_emission_batch_kernel (.cu:416-417):
    //   col = c*(q-cx) - s*(h-cy) + cx,  row = s*(q-cx) + c*(h-cy) + cy
    // Preimage rows: invert the rotation at the region's corners.
    float hmin = 3.0e38f, hmax = -3.0e38f;
    const float Xs[2] = {t.lox - cy, t.hix - cy};
    const float Ys[2] = {t.loy - cx, t.hiy - cx};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float hc = c * Xs[i] - s * Ys[j];
            hmin = fminf(hmin, hc);
            hmax = fmaxf(hmax, hc);
        }
    cur.hlo = uniform_int((int)clampf(floorf(hmin + cy) - 1.f, 0.f, (float)p.H));
    cur.hhi = uniform_int((int)clampf(ceilf(hmax + cy) + 1.f, -1.f, (float)(p.H - 1)));
    cur.c = c;
    cur.s = s;
    const int64_t sl = (int64_t)cur.a * p.B + t.b;
    cur.ra = make_rsrc(p.atten + sl * p.H * p.W, p.H * p.W);
    cur.rr = make_rsrc(p.ratio ? p.ratio + sl * p.W : p.atten, p.ratio ? p.W : 0);
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) cur.ew[rr] = scalar_read(p.em, (int64_t)cur.a * p.R + p.r0 + rr);
}

// Column intervals of the batch's rows, their prefix sums and the compacted
// row table.
__device__ __forceinline__ void bwd_begin_batch(const BwdArgs &p, const BwdTile &t, BwdCursor &cur)
{
    const int lane = t.lane;
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float c = cur.c, s = cur.s;
    const int h = cur.rb + 2 * lane;
    int ql = 0, len = 0;
    if (lane < kBRows && h <= cur.hhi) {
        const float hc = (float)h - cy;
        const float bx = fmaf(c, hc, cy), by = fmaf(-s, hc, cx);
        auto in_region = [&](int q) {
            const float qc = (float)q - cx;
            const float X = fmaf(s, qc, bx), Y = fmaf(c, qc, by);
            return X >= t.lox && X < t.hix && Y >= t.loy && Y < t.hiy;
        };
        float qmin = -cx, qmax = (float)(p.W - 1) - cx;
        slope_interval(s, t.lox, t.hix, bx, t.tol, qmin, qmax);
        slope_interval(c, t.loy, t.hiy, by, t.tol, qmin, qmax);
        int qa = (int)clampf(floorf(qmin + cx), 0.f, (float)p.W);
        int qb = (int)clampf(ceilf(qmax + cx), -1.f, (float)(p.W - 1));
        // Tighten to the exact set: with the same arithmetic as the splat, a
        // sample passes this test iff its clamped base cell lies in
        // [I0-1, I0+BI-1] x [J0-1, J0+kBJ-1] (BI the tile height).
        while (qa <= qb && !in_region(qa)) ++qa;
        while (qa <= qb && !in_region(qb)) --qb;
        ql = qa;
        len = max(qb - qa + 1, 0);
    }
    const int lp = len + (len & 1);
    int P = wave_inclusive_scan(lp);  // inclusive prefix sum over the batch
    cur.T = lane_value(P, 63);
    P -= lp;
    const bool ne = len > 0;
    const uint64_t nem = __ballot(ne);
    cur.nne = __popcll(nem);
    // Non-empty rows, compacted: (row * W, first sample - first column, last sample, row - cy).
    if (ne) t.rows[lanes_below(nem)] = make_int4(h * p.W, P - ql, P + len - 1, __float_as_int((float)h - cy));
    cur.P = P;
    cur.ne = ne;
}

// Row starts (even) of the window at win0, one bit per lane pair: item n's
// lanes 0-31 / 32-63 in words 2n / 2n+1.
__device__ __forceinline__ void bwd_begin_window(const BwdTile &t, BwdCursor &cur)
{
    t.mask[t.lane] = 0u;
    if (cur.ne && cur.P >= cur.win0 && cur.P < cur.win0 + kBWin) {
        const int o = (cur.P - cur.win0) >> 1;
        atomicOr(&t.mask[o >> 5], 1u << (o & 31));
    }
    cur.kb = __popcll(__ballot(cur.ne && cur.P < cur.win0)) - 1;  // row of the sample before win0
    cur.nit = min(kBItems, (cur.T - cur.win0 + 127) >> 7);
    cur.n = 0;
}

// Positions of the current item pair's samples, and their atten and ratio
// loads (bounds-checked: an offset past the slice reads 0, so lanes without a
// sample need no branch; their values go to the sink).
template <int NR, bool RATIO>
__device__ __forceinline__ void bwd_gather(const BwdArgs &p, const BwdTile &t, BwdCursor &cur, BwdPair &g)
{
    const int lane = t.lane;
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float c = cur.c, s = cur.s;
#pragma unroll
    for (int u = 0; u < kBUnroll; ++u) {
        const uint2 m = *reinterpret_cast<const uint2 *>(&t.mask[(cur.n + u) * 2]);
        const uint64_t E = (uint64_t)m.x | ((uint64_t)m.y << 32);
        const int ci = cur.kb + lanes_below(E) + (int)((E >> lane) & 1);
        cur.kb += __popcll(E);
        const int f0 = cur.win0 + (cur.n + u) * 128 + 2 * lane;
        const int4 rw = t.rows[min(max(ci, 0), cur.nne - 1)];
        const int q = f0 - rw.y;
        const float qc = (float)q - cx, hc = __int_as_float(rw.w);
        const float bx = fmaf(c, hc, cy), by = fmaf(-s, hc, cx);
        g.X[u][0] = fmaf(s, qc, bx);
        g.Y[u][0] = fmaf(c, qc, by);
        g.X[u][1] = fmaf(s, qc + 1.0f, bx);
        g.Y[u][1] = fmaf(c, qc + 1.0f, by);
        const uint32_t oa = (uint32_t)(rw.x + q) * 4u;
        g.av[u][0] = load_f32(cur.ra, oa);
        g.av[u][1] = load_f32(cur.ra, oa + 4u);
        if (RATIO) {
            const uint32_t orat = (uint32_t)q * 4u;
            g.rv[u][0] = load_f32(cur.rr, orat);
            g.rv[u][1] = load_f32(cur.rr, orat + 4u);
        } else {
            g.rv[u][0] = g.rv[u][1] = 1.0f;
        }
        g.val[u][0] = f0 <= rw.z;
        g.val[u][1] = f0 < rw.z;
    }
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) g.ew[rr] = cur.ew[rr];
}

// Splat one item pair into the wave's accumulator.  Every sample here is
// unclamped (the clamped ones are summed per border cell by bwd_edge).
template <int NR>
__device__ __forceinline__ void bwd_splat(const BwdTile &t, const BwdPair &g)
{
    float *lds = t.lds;
    int cell[kBUnroll][2];
    float wq[kBUnroll][2][4];
#pragma unroll
    for (int u = 0; u < kBUnroll; ++u)
#pragma unroll
        for (int pt = 0; pt < 2; ++pt) {
            // sample positions are >= 0 here, so (int) is floor and fract is the
            // exact remainder
            const float Xc = g.X[u][pt], Yc = g.Y[u][pt];
            const int y0 = (int)Xc, x0 = (int)Yc;
            const float dy = __builtin_amdgcn_fractf(Xc), dx = __builtin_amdgcn_fractf(Yc);
            cell[u][pt] = g.val[u][pt] ? y0 * kBPitch + x0 + t.cbase : t.sink;
            // w*(1-dx)*(1-dy), w*dx*(1-dy), w*(1-dx)*dy, w*dx*dy  (.cu:434-443),
            // formed as differences: w*(1-dx) = w - w*dx
            const float v = g.av[u][pt] * g.rv[u][pt];
            const float w1 = v * dx, w0 = v - w1;
            const float w01 = w0 * dy, w11 = w1 * dy;
            wq[u][pt][0] = w0 - w01;
            wq[u][pt][1] = w1 - w11;
            wq[u][pt][2] = w01;
            wq[u][pt][3] = w11;
        }
    constexpr int offs[4] = {0, NR, kBPitch * NR, (kBPitch + 1) * NR};
#pragma unroll
    for (int pt = 0; pt < 2; ++pt) {
        float *pa[kBUnroll];
#pragma unroll
        for (int u = 0; u < kBUnroll; ++u) pa[u] = lds + cell[u][pt] * NR;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if constexpr (NR == 1) {
                float cur[kBUnroll];
#pragma unroll
                for (int u = 0; u < kBUnroll; ++u) cur[u] = pa[u][offs[k]];
#pragma unroll
                for (int u = 0; u < kBUnroll; ++u) pa[u][offs[k]] = fmaf(g.ew[0], wq[u][pt][k], cur[u]);
            } else {
                float2 cur[kBUnroll];
#pragma unroll
                for (int u = 0; u < kBUnroll; ++u) cur[u] = *reinterpret_cast<const float2 *>(pa[u] + offs[k]);
#pragma unroll
                for (int u = 0; u < kBUnroll; ++u) {
                    cur[u].x = fmaf(g.ew[0], wq[u][pt][k], cur[u].x);
                    cur[u].y = fmaf(g.ew[1], wq[u][pt][k], cur[u].y);
                    *reinterpret_cast<float2 *>(pa[u] + offs[k]) = cur[u];
                }
            }
            // keep corner k's writes before corner k+1's reads
            __asm__ volatile("" ::: "memory");
        }
    }
}

// Clamped samples.  The reference moves a sample position outside
// [0, H-1.001] x [0, W-1.001] onto the border before splatting it
// (.cu:421-424), so every such sample lands on a border base cell: row 0 (top,
// X < 0), row H-2 (bottom, X > H-1.001), column 0 (left, Y < 0, X inside) or
// column W-2 (right, Y > W-1.001, X inside); a sample clamped in both
// coordinates counts as top or bottom.  Many of them share a cell within one
// splat instruction, so they are gathered per border cell instead: in a tile
// that touches an edge, lane k owns one of the tile's border base cells and,
// per angle, walks the strip of detector points that clamp into it, testing
// each candidate with the splat's own arithmetic, and sums its weights in
// registers; the four corner sums are then added to the accumulator one
// corner at a time (the lanes' cells differ, so no two lanes collide).
enum : int { kEdgeTop = 1, kEdgeBottom = 2, kEdgeLeft = 4, kEdgeRight = 8 };

// A buffer offset past every slice (H*W*4 < 2^31 bytes): the load returns 0.
constexpr uint32_t kOOB = 0x80000000u;

// One bound on the minor coordinate n as a linear function of the major one:
// n >= u*m + v (lower) or n <= u*m + v (upper); inactive bounds are +-1e30.
struct LinBound {
    float u, v;
};

// Half-plane a*m + b*n + g >= 0 as a bound on n.  A half-plane that hardly
// depends on n (|b| below 1/16) becomes a condition on m alone, relaxed by the
// largest |b*n| over the box so that it never loses a point.
__device__ __forceinline__ void half_plane(float a, float b, float g, float nabs, LinBound &lo, LinBound &hi,
                                           float &mlo, float &mhi)
{
    constexpr float kBig = 1.0e30f;
    lo = {0.0f, -kBig};
    hi = {0.0f, kBig};
    if (b >= 0.0625f) {
        const float r = 1.0f / b;
        lo = {-a * r, -g * r};
    } else if (b <= -0.0625f) {
        const float r = -1.0f / b;
        hi = {a * r, g * r};
    } else {
        const float gg = g + fabsf(b) * nabs;
        if (a > 0.0f)
            mlo = fmaxf(mlo, -gg / a);
        else if (a < 0.0f)
            mhi = fminf(mhi, gg / -a);
        else if (gg < 0.0f)
            mhi = mlo - 1.0f;  // empty
    }
}

// Narrow [mlo, mhi] to where lower bound L can lie below upper bound U.
__device__ __forceinline__ void pair_range(const LinBound &L, const LinBound &U, float &mlo, float &mhi)
{
    const float du = L.u - U.u, dv = U.v - L.v;  // need du*m <= dv
    if (du > 0.0f)
        mhi = fminf(mhi, dv / du);
    else if (du < 0.0f)
        mlo = fmaxf(mlo, dv / du);
    else if (dv < 0.0f)
        mhi = mlo - 1.0f;
}

// One edge candidate pair: detector points (m, n) and (m, n + 1) of the walk.
struct EdgePair {
    float av[2], rv[2];  // atten and ratio (0 where the point is not taken)
    float fr[2];         // fraction along the edge: dx (top, bottom) or dy (left, right)
};

template <int NR, bool RATIO, int EDGE>
__device__ __forceinline__ void bwd_edge(const BwdArgs &p, const BwdTile &t, const BwdCursor &cur)
{
    constexpr bool xedge = EDGE == kEdgeTop || EDGE == kEdgeBottom;  // clamped in X (rows 0 / H-2)
    constexpr bool low_edge = EDGE == kEdgeTop || EDGE == kEdgeLeft;
    const float cx = (p.W - 1) * 0.5f, cy = (p.H - 1) * 0.5f;
    const float xmax = p.H - 1.001f, ymax = p.W - 1.001f;
    const float c = cur.c, s = cur.s;
    // the lane's border base cell: column k of row 0 / H-2, or row k of column 0 / W-2
    const int k = (xedge ? t.J0 : t.I0) - 1 + t.lane;
    const int kmax = xedge ? p.W - 2 : p.H - 2;
    const int nlanes = xedge ? kBJ + 1 : t.bi + 1;
    const bool active = t.lane < nlanes && k >= 0 && k <= kmax;
    // Walk the detector axis along the strip (major m), the other one per step
    // (minor n): the strip coordinate P (Y for top/bottom, X for left/right)
    // changes by at least 0.7 per minor step, so a step of a strip one cell wide
    // holds at most 2 points (the corner cells' strips are open on one side).
    //   X - cy = c*hc + s*qc,  Y - cx = -s*hc + c*qc   (hc = h - cy, qc = q - cx)
    const bool major_h = xedge ? fabsf(c) >= fabsf(s) : fabsf(c) < fabsf(s);
    const float cm = major_h ? cy : cx, cn = major_h ? cx : cy;
    const int Mmax = major_h ? p.H - 1 : p.W - 1, Nmax = major_h ? p.W - 1 : p.H - 1;
    const float aX = major_h ? c : s, bX = major_h ? s : c;
    const float aY = major_h ? -s : c, bY = major_h ? c : -s;
    const float aP = xedge ? aY : aX, bP = xedge ? bY : bX, P0 = xedge ? cx : cy;
    const float aQ = xedge ? aX : aY, bQ = xedge ? bX : bY, Q0 = xedge ? cy : cx;
    constexpr float kBig = 1.0e30f;
    // Every bound below is relaxed by tol (coordinate units), far more than the
    // rounding of the splat's arithmetic: the walk visits a superset of the
    // strip, and each point is then tested exactly.
    const float tol = 2.0f * t.tol + 0.01f;
    // strip kk <= P < kk+1 (top/bottom: column 0 also takes Y < 0, column W-2
    // also Y > W-1.001): n between slo(mc) and shi(mc)
    const float rP = 1.0f / bP;
    const auto strip = [&](int kk, LinBound &lo, LinBound &hi) {
        const float plo = (xedge && kk == 0) ? -kBig : (float)kk - tol;
        const float phi = (xedge && kk == kmax) ? kBig : (float)(kk + 1) + tol;
        const float vA = (plo - P0) * rP, vB = (phi - P0) * rP;
        lo = {-aP * rP, fminf(vA, vB)};
        hi = {-aP * rP, fmaxf(vA, vB)};
    };
    // depth beyond the edge, dA*mc + dB*nc + dG >= 0: Q < 0 (top, left) or
    // Q > H-1.001 / W-1.001 (bottom, right)
    const float qlim = low_edge ? 0.0f : (xedge ? xmax : ymax);
    const float dA = low_edge ? -aQ : aQ, dB = low_edge ? -bQ : bQ;
    const float dG = low_edge ? tol - Q0 : Q0 - qlim + tol;
    float mlo = -cm, mhi = (float)Mmax - cm;
    LinBound dlo, dhi;
    half_plane(dA, dB, dG, fmaxf(cn, (float)Nmax - cn), dlo, dhi, mlo, mhi);
    // steps m0..m1 of this lane's own strip
    int m0, m1;
    {
        LinBound slo, shi;
        strip(k, slo, shi);
        const LinBound L[3] = {{0.0f, -cn}, slo, dlo}, U[3] = {{0.0f, (float)Nmax - cn}, shi, dhi};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) pair_range(L[i], U[j], mlo, mhi);
        m0 = (int)fmaxf(ceilf(mlo + cm - 0.01f), 0.0f);
        m1 = active ? (int)fminf(floorf(mhi + cm + 0.01f), (float)Mmax) : -1;
    }
    const int len = max(m1 - m0 + 1, 0);

    // Balance: the strips are cut into chunks of at most Lc steps, at most 64
    // chunks in all, and lane j walks chunk j.  Chunk table: the non-empty
    // strips compacted (cell, m0, m1, first chunk); the start bits of the strips'
    // first chunks give each lane its strip.
    const int total = lane_value(wave_inclusive_scan(len), 63);
    if (total == 0) return;
    const uint64_t nem = __ballot(len > 0);
    const int spare = max(64 - __popcll(nem), 1);
    const int Lc = max((total + spare - 1) / spare, 1);
    const int nch = (len + Lc - 1) / Lc;
    const int cP = wave_inclusive_scan(nch) - nch;  // first chunk of this lane's strip
    const int nchunks = lane_value(cP + nch, 63);
    int4 *tab = t.rows;
    if (t.lane < 2) t.mask[t.lane] = 0u;
    __asm__ volatile("" ::: "memory");
    if (len > 0) {
        tab[lanes_below(nem)] = make_int4(k, m0, m1, cP);
        atomicOr(&t.mask[cP >> 5], 1u << (cP & 31));
    }
    __asm__ volatile("" ::: "memory");
    const uint2 mw = *reinterpret_cast<const uint2 *>(t.mask);
    const uint64_t starts = (uint64_t)mw.x | ((uint64_t)mw.y << 32);
    const int ord = __popcll(starts & (t.lane == 63 ? ~0ull : (2ull << t.lane) - 1ull)) - 1;
    const int4 ce = tab[max(ord, 0)];
    const int ck = ce.x;                              // the chunk's cell
    int m = ce.y + (t.lane - ce.w) * Lc;              // its first step
    const int mend = t.lane < nchunks ? min(m + Lc - 1, ce.z) : -1;
    LinBound slo, shi;
    strip(ck, slo, shi);

    // minor range of step mm
    const auto nrange = [&](int mm, int &na, int &nb) {
        const float mc = (float)mm - cm;
        const float lo = fmaxf(fmaf(slo.u, mc, slo.v), fmaf(dlo.u, mc, dlo.v));
        const float hi = fminf(fmaf(shi.u, mc, shi.v), fmaf(dhi.u, mc, dhi.v));
        na = (int)fmaxf(ceilf(lo + cn - 0.01f), 0.0f);
        nb = (int)fminf(floorf(hi + cn + 0.01f), (float)Nmax);
    };
    // candidates (mm, nn) and (mm, nn+1): the splat's arithmetic (bwd_gather),
    // then the exact test; the loads of points not taken read 0
    const auto gather = [&](int mm, int nn, int nb, EdgePair &e) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int h = major_h ? mm : nn + i, q = major_h ? nn + i : mm;
            const float hc = (float)h - cy, qc = (float)q - cx;
            const float X = fmaf(s, qc, fmaf(c, hc, cy)), Y = fmaf(c, qc, fmaf(-s, hc, cx));
            bool in;
            if constexpr (xedge) {
                const float Yc = clampf(Y, 0.f, ymax);
                in = (low_edge ? X < 0.0f : X > xmax) && (int)Yc == ck;
                e.fr[i] = __builtin_amdgcn_fractf(Yc);
            } else {
                in = X >= 0.0f && X <= xmax && (low_edge ? Y < 0.0f : Y > ymax) && (int)X == ck;
                e.fr[i] = __builtin_amdgcn_fractf(X);
            }
            in = in && nn + i <= nb;
            e.av[i] = load_f32(cur.ra, in ? (__umul24((uint32_t)h, (uint32_t)p.W) + (uint32_t)q) * 4u : kOOB);
            e.rv[i] = RATIO ? load_f32(cur.rr, in ? (uint32_t)q * 4u : kOOB) : 1.0f;
        }
    };

    // Sums over the chunk: A[0] of ew*w*(1-f), A[1] of ew*w*f, f the fraction
    // along the edge; the other coordinate's fraction is the same for every
    // point of an edge (0 at top/left, that of H-1.001 / W-1.001 at bottom/right)
    // and is applied when the sums are added.
    float A[2][NR];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) A[j][rr] = 0.0f;
    const auto add = [&](const EdgePair &e) {
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            // w*(1-f), w*f: the splat's differences (.cu:434-443)
            const float v = e.av[i] * e.rv[i];
            const float w1 = v * e.fr[i], w0 = v - w1;
#pragma unroll
            for (int rr = 0; rr < NR; ++rr) {
                A[0][rr] = fmaf(cur.ew[rr], w0, A[0][rr]);
                A[1][rr] = fmaf(cur.ew[rr], w1, A[1][rr]);
            }
        }
    };
    // One step per iteration: its points n0 and n0+1, the next step's loads in
    // flight (two pair buffers used in turn, so no register copy waits on a
    // load).  A step holds more than 2 points only where a strip reaches into a
    // corner region (top/bottom column 0 or W-2); those extra points are taken
    // on the spot.
    const auto extra = [&](int mm, int na, int nb) {
        for (int nn = na + 2; nn <= nb; nn += 2) {
            EdgePair x;
            gather(mm, nn, nb, x);
            add(x);
        }
    };
    int na, nb;
    EdgePair e0, e1;
    if (m <= mend) {
        nrange(m, na, nb);
        gather(m, na, nb, e0);
    }
    while (m <= mend) {
        extra(m, na, nb);
        ++m;
        const bool more = m <= mend;
        if (more) {
            nrange(m, na, nb);
            gather(m, na, nb, e1);
        }
        add(e0);
        if (!more) break;
        extra(m, na, nb);
        ++m;
        const bool more2 = m <= mend;
        if (more2) {
            nrange(m, na, nb);
            gather(m, na, nb, e0);
        }
        add(e1);
        if (!more2) break;
    }

    // Chunk sums in the table's place (it has been read), then each cell's lane
    // adds its chunks in order.
    float *part = reinterpret_cast<float *>(t.rows);
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) part[(t.lane * 2 + j) * NR + rr] = A[j][rr];
    __asm__ volatile("" ::: "memory");
    float S[2][NR];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) S[j][rr] = 0.0f;
    for (int i = 0; i < nch; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int rr = 0; rr < NR; ++rr) S[j][rr] += part[((cP + i) * 2 + j) * NR + rr];
    __asm__ volatile("" ::: "memory");

    // Add to the accumulator, one corner at a time (corner j of base cell
    // (y0, x0) is (y0 + j/2, x0 + j%2); the lanes' base cells differ).
    const int y0 = EDGE == kEdgeTop ? 0 : EDGE == kEdgeBottom ? p.H - 2 : k;
    const int x0 = EDGE == kEdgeLeft ? 0 : EDGE == kEdgeRight ? p.W - 2 : k;
    float *pa = t.lds + (len > 0 ? y0 * kBPitch + x0 + t.cbase : t.sink) * NR;
    constexpr int offs[4] = {0, NR, kBPitch * NR, (kBPitch + 1) * NR};
    // the fixed fraction: dy at the bottom edge, dx at the right edge
    const float D = __builtin_amdgcn_fractf(EDGE == kEdgeBottom ? xmax : ymax);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        // along-edge index ja and across-edge index jb of corner j
        const int ja = xedge ? (j & 1) : (j >> 1), jb = xedge ? (j >> 1) : (j & 1);
        if (low_edge && jb == 1) continue;  // the fixed fraction is 0
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) {
            const float a = S[ja][rr];
            const float part_j = low_edge ? a : (jb == 0 ? a - a * D : a * D);
            pa[offs[j] + rr] += part_j;
        }
        __asm__ volatile("" ::: "memory");
    }
}

// The wave's whole stream: angles a = wave, wave+4, ...; per angle the two row
// classes, their batches and the batches' windows.  Inside a window the loads
// of the next item pair are issued before the current pair is splatted (two
// pair buffers, used in turn, so no register copy waits on a load in flight).
// RATIO: ratio is given (false: ratio = 1, the sensitivity).  Tiles on the
// image border also gather their border cells' clamped samples per angle.
template <int NR, bool RATIO>
__device__ __forceinline__ void bwd_stream(const BwdArgs &p, const BwdTile &t, int wave)
{
    BwdCursor cur;
    float cs_c = 0.0f, cs_s = 0.0f;  // lane l: angle a0 + 4l of the current block of 64
    for (cur.a = wave; cur.a < p.A; cur.a += kBWaves) {
        const int j = ((cur.a - wave) / kBWaves) & 63;  // wave-uniform
        if (j == 0) {
            const int aa = cur.a + kBWaves * t.lane;
            const float th = aa < p.A ? finite_or_zero(p.theta[aa]) : 0.0f;
            cs_c = cosf(th);
            cs_s = sinf(th);
        }
        bwd_begin_angle<NR>(p, t, cur, __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cs_c), j)),
                            __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cs_s), j)));
        if (t.edges & kEdgeTop) bwd_edge<NR, RATIO, kEdgeTop>(p, t, cur);
        if (t.edges & kEdgeBottom) bwd_edge<NR, RATIO, kEdgeBottom>(p, t, cur);
        if (t.edges & kEdgeLeft) bwd_edge<NR, RATIO, kEdgeLeft>(p, t, cur);
        if (t.edges & kEdgeRight) bwd_edge<NR, RATIO, kEdgeRight>(p, t, cur);
        for (cur.cls = 0; cur.cls < 2; ++cur.cls) {
            for (cur.rb = cur.hlo + cur.cls; cur.rb <= cur.hhi; cur.rb += 2 * kBRows) {
                bwd_begin_batch(p, t, cur);
                for (cur.win0 = 0; cur.win0 < cur.T; cur.win0 += kBWin) {
                    bwd_begin_window(t, cur);
                    BwdPair g0, g1;
                    bwd_gather<NR, RATIO>(p, t, cur, g0);
                    while (true) {
                        cur.n += kBUnroll;
                        const bool more = cur.n < cur.nit;
                        if (more) bwd_gather<NR, RATIO>(p, t, cur, g1);
                        bwd_splat<NR>(t, g0);
                        if (!more) break;
                        cur.n += kBUnroll;
                        const bool more2 = cur.n < cur.nit;
                        if (more2) bwd_gather<NR, RATIO>(p, t, cur, g0);
                        bwd_splat<NR>(t, g1);
                        if (!more2) break;
                    }
                }
            }
        }
    }
}

// Per-wave tables of the back-projector in static LDS.  One instance per
// kernel: every tile height a kernel instantiates shares it.
struct BwdTables {
    __align__(16) uint32_t mask[kBWaves][kBItems * 2];  // row starts of the current window
    int4 row[kBWaves][64];  // compacted rows of the current batch (kBRows); edge chunk table, partial sums
};

__device__ __forceinline__ BwdTables &bwd_tables()
{
    __shared__ BwdTables tab;
    return tab;
}

// One workgroup's tile id: image tile (id / ntj) % nti, id % ntj of slice
// id / (nti * ntj), BI image rows high.
template <int NR, int BI>
__device__ __forceinline__ void bwd_tile(const BwdArgs &p, uint32_t id)
{
    constexpr int BACC = bwd_acc_cells(BI);
    extern __shared__ float lds[];
    BwdTables &tab = bwd_tables();

    const int tj = id % p.ntj;
    id /= p.ntj;
    const int ti = id % p.nti;
    const int b = id / p.nti;
    const int lane = threadIdx.x & 63;
    const int wave = uniform_int(threadIdx.x >> 6);  // wave-uniform: keeps per-wave state in SGPRs
    constexpr int nacc = NR * BACC;

    BwdTile t;
    t.I0 = ti * BI;
    t.J0 = tj * kBJ;
    t.b = b;
    t.bi = BI;
    // Unclamped sample positions (0 <= X <= H-1.001, 0 <= Y <= W-1.001) whose
    // base cell lies in [I0-1, I0+BI-1] x [J0-1, J0+kBJ-1]; the upper limits are
    // the floats just above H-1.001 and W-1.001 (X < that iff X <= H-1.001).
    const float xup = __int_as_float(__float_as_int(p.H - 1.001f) + 1);
    const float yup = __int_as_float(__float_as_int(p.W - 1.001f) + 1);
    t.lox = (float)max(t.I0 - 1, 0);
    t.hix = fminf((float)(t.I0 + BI), xup);
    t.loy = (float)max(t.J0 - 1, 0);
    t.hiy = fminf((float)(t.J0 + kBJ), yup);
    // Border base cells in the tile: row 0 / H-2, column 0 / W-2.
    const auto holds = [](int lo, int hi, int v) { return lo <= v && v <= hi; };
    t.edges = (holds(t.I0 - 1, t.I0 + BI - 1, 0) ? kEdgeTop : 0) |
              (holds(t.I0 - 1, t.I0 + BI - 1, p.H - 2) ? kEdgeBottom : 0) |
              (holds(t.J0 - 1, t.J0 + kBJ - 1, 0) ? kEdgeLeft : 0) |
              (holds(t.J0 - 1, t.J0 + kBJ - 1, p.W - 2) ? kEdgeRight : 0);
    // Largest rounding error of one sample coordinate near an active bound.
    t.tol = 2.4e-7f * (float)(2 * max(p.H, p.W) + 4);
    t.lds = lds;
    t.cbase = wave * BACC + (1 - t.I0) * kBPitch + 1 - t.J0;
    t.mask = tab.mask[wave];
    t.rows = tab.row[wave];
    t.lane = lane;
    // Sink words (after all accumulators): lanes with nothing to add in a step
    // read-modify-write there, each lane its own NR words.
    t.sink = kBWaves * BACC + lane;
    for (int k = lane; k < nacc; k += 64) lds[wave * nacc + k] = 0.0f;

    // Wave w takes angles w, w+4, ... into its own accumulator.
    if (p.ratio)
        bwd_stream<NR, true>(p, t, wave);
    else
        bwd_stream<NR, false>(p, t, wave);
    __syncthreads();

    // Sum the four private accumulators in a fixed order and write the tile (lane = column).
    const int I0 = t.I0, J0 = t.J0;
    const int64_t HW = (int64_t)p.H * p.W;
    const int j = J0 + lane;
    for (int rr = 0; rr < NR; ++rr) {
        const int r = p.r0 + rr;
        const int64_t sb = ((int64_t)r * p.B + b) * HW;
        for (int ii = wave; ii < BI; ii += kBWaves) {
            const int i = I0 + ii;
            if (i >= p.H || j >= p.W || lane >= kBJ) continue;
            const int li = ((ii + 1) * kBPitch + lane + 1) * NR + rr;
            const float back = ((lds[li] + lds[nacc + li]) + lds[2 * nacc + li]) + lds[3 * nacc + li];
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

template <int NR, int BI>
__global__ __launch_bounds__(256) void bwd_tile_kernel(BwdArgs p)
{
    bwd_tile<NR, BI>(p, xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------
// Grouped persistent launches.  The grid holds as many workgroups as are
// resident at once; worker w (hardware id remapped so that each XCD holds one
// contiguous range of workers) takes logical tiles w, w + grid, ...: in every
// round each XCD works on one contiguous run of tiles, which share slices of
// C and atten in that XCD's L2.  Jobs are ordered by cost per tile, largest
// first, so that the rounds stay balanced.  A barrier separates consecutive
// tiles of one workgroup (the next tile's LDS writes wait for the previous
// tile's cross-wave reads).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256, 4) void fwd_group_kernel(const FwdArgs *jobs, const uint32_t *begin, int n_jobs,
                                                           uint32_t n_tiles)
{
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x);
    for (uint32_t t = w; t < n_tiles; t += gridDim.x) {
        const int j = find_job(begin, n_jobs, t);
        const FwdArgs p = scalar_record(jobs + j);
        const uint32_t lt = t - scalar_read_u32(begin, j);
        if (t != w) __syncthreads();
        if (p.group_angles == 4)
            fwd_strip_tile<1, 4>(p, lt);
        else
            fwd_strip_tile<1, 8>(p, lt);
    }
}

__global__ __launch_bounds__(256) void bwd_group_kernel(const BwdArgs *jobs, const uint32_t *begin, int n_jobs,
                                                        uint32_t n_tiles)
{
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x);
    for (uint32_t t = w; t < n_tiles; t += gridDim.x) {
        const int j = find_job(begin, n_jobs, t);
        const BwdArgs p = scalar_record(jobs + j);
        const uint32_t lt = t - scalar_read_u32(begin, j);
        if (t != w) __syncthreads();
        if (p.tile_rows == 15)
            bwd_tile<1, 15>(p, lt);
        else
            bwd_tile<1, 31>(p, lt);
    }
}

__global__ __launch_bounds__(256) void copy_group_kernel(const CopyJob *jobs, const uint32_t *begin, int n_jobs,
                                                         uint32_t n_tiles)
{
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int j = find_job(begin, n_jobs, t);
        const CopyJob c = scalar_record(jobs + j);
        const int64_t i0 = (int64_t)(t - scalar_read_u32(begin, j)) * kCopyTile;
        const int64_t i1 = i0 + kCopyTile < c.n ? i0 + kCopyTile : c.n;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) c.dst[i] = c.src[i];
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
                   float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode, hipStream_t stream,
                   int out_sm)
{
    static bool attr_set = false;
    const size_t shmem1 = fwd_lds_bytes(1), shmem2 = fwd_lds_bytes(2);
    if (!attr_set) {
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)fwd_strip_kernel<1, kFG>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)shmem1));
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)fwd_strip_kernel<2, kFG>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)shmem2));
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
    p.out_sm = out_sm;
    FL_REQUIRE(!out_sm || R <= kFRefChunk, "slice-major output: at most two references");
    const int64_t nwg = (int64_t)p.ntw * p.n_groups * B;
    FL_REQUIRE(nwg < (int64_t)UINT32_MAX, "forward projector: grid too large");
    for (int64_t r0 = 0; r0 < R; r0 += kFRefChunk) {
        p.r0 = (int)r0;
        p.nref = (int)std::min<int64_t>(kFRefChunk, R - r0);
        const bool first = r0 == 0, last = r0 + kFRefChunk >= R;
        const bool ratio = mode == kFwdRatio;
        p.mode = first ? (last && ratio ? kFwdRatio : kFwdWrite)
                       : (last && ratio ? kFwdAccumulateRatio : kFwdAccumulate);
        const dim3 grid((uint32_t)nwg), block(256);
        if (p.nref == 2)
            fwd_strip_kernel<2, kFG><<<grid, block, shmem2, stream>>>(p);
        else
            fwd_strip_kernel<1, kFG><<<grid, block, shmem1, stream>>>(p);
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
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)bwd_tile_kernel<1, kBI>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)bwd_lds_bytes(1)));
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)bwd_tile_kernel<2, kBI>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)bwd_lds_bytes(2)));
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
        if (p.nref == 2)
            bwd_tile_kernel<2, kBI><<<dim3((uint32_t)nwg), dim3(256), bwd_lds_bytes(2), stream>>>(p);
        else
            bwd_tile_kernel<1, kBI><<<dim3((uint32_t)nwg), dim3(256), bwd_lds_bytes(1), stream>>>(p);
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

// Workgroups of `kernel` resident on the current device at once.
static int resident_grid(const void *kernel, size_t lds, int *grid)
{
    int dev = 0, cus = 0, per_cu = 0;
    FL_HIP_CHECK(hipGetDevice(&dev));
    FL_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    FL_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds));
    *grid = std::max(1, cus) * std::max(1, per_cu);
    return FL_OK;
}

static const size_t kFwdGroupLds = fwd_lds_bytes(1);
static const size_t kBwdGroupLds = bwd_lds_bytes(1, 31);  // the larger of the two tile heights

int fwd_group_grid(int *grid)
{
    FL_HIP_CHECK(hipFuncSetAttribute((const void *)fwd_group_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kFwdGroupLds));
    return resident_grid((const void *)fwd_group_kernel, kFwdGroupLds, grid);
}

int bwd_group_grid(int *grid)
{
    FL_HIP_CHECK(hipFuncSetAttribute((const void *)bwd_group_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)kBwdGroupLds));
    return resident_grid((const void *)bwd_group_kernel, kBwdGroupLds, grid);
}

int copy_group_grid(int *grid) { return resident_grid((const void *)copy_group_kernel, 0, grid); }

int launch_fwd_group(const FwdArgs *jobs, const uint32_t *begin, int n_jobs, uint32_t n_tiles, int grid,
                     hipStream_t s)
{
    if (n_tiles == 0) return FL_OK;
    fwd_group_kernel<<<dim3((uint32_t)grid), dim3(256), kFwdGroupLds, s>>>(jobs, begin, n_jobs, n_tiles);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_bwd_group(const BwdArgs *jobs, const uint32_t *begin, int n_jobs, uint32_t n_tiles, int grid,
                     hipStream_t s)
{
    if (n_tiles == 0) return FL_OK;
    bwd_group_kernel<<<dim3((uint32_t)grid), dim3(256), kBwdGroupLds, s>>>(jobs, begin, n_jobs, n_tiles);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_copy_group(const CopyJob *jobs, const uint32_t *begin, int n_jobs, uint32_t n_tiles, int grid,
                      hipStream_t s)
{
    if (n_tiles == 0) return FL_OK;
    copy_group_kernel<<<dim3((uint32_t)grid), dim3(256), 0, s>>>(jobs, begin, n_jobs, n_tiles);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

}  // namespace fl
