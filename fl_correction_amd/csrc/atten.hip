// Attenuation maps for CDNA4: the incident-beam product, the fluorescence-path
// line integral over the detector cone, the in-plane rotation of the volume and
// the frame changes between them.
//
// Frames.  The original frame is (B, H, W) = (slice, row, col); the incident
// beam runs along rows toward row 0 (FL/torch_functions.py:95-102) and the
// detector sits at +col.  The mask frame of atten_kernel (FL/cuda_lib/
// atten_cuda_kernel.cu:10-57) is (Z, X, Y) = (slice, col, H-1-row): the
// reference reaches it with torch_rot90_4D(..., 'clock') (torch_functions.py:143)
// and the cone extends along +X.  Here mu is written straight into that frame
// by the assembly kernel and read back, transposed, by the combine kernel.
//
// Kernels
//   mask_corr_kernel   out = M (*) mu, any dense mask, zero taps skipped
//   rotate_kernel      affine_grid + grid_sample rotation (zeros, align_corners=False)
//   incident_kernel    prod_{m>r} (1 - frac*coef) down each column
//   mu_rot_kernel      mu_e = sum_j frac_j*coef[j,e], written in the mask frame
//   cone_kernel        one sampled cone direction: exact rational-step recursion
//   combine_kernel     exp(-T) (optionally times the incident product), back to (B,H,W)
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace fl {

// ---------------------------------------------------------------------------
// Dense mask correlation (atten_kernel, .cu:10-57):
//   out[k,z,x,y] = sum_{dz,dx,dy} mask[dz,dx,dy] * mu[k, z+dz-hz, x+dx-hx, y+dy-hy]
// with (hz, hx, hy) = (Mz/2, 0, My/2) for the reference op; the transpose is the
// same correlation with the mask flipped and (Mz-1-hz, Mx-1, My-1-hy).
// One thread computes kXB outputs along x for one (k, z, y); lanes run along y,
// so every mu load is coalesced and each loaded value feeds kXB multiply-adds.
// Mask values are wave-uniform (scalar loads); zero taps are skipped.
// ---------------------------------------------------------------------------
constexpr int kXB = 8;

__global__ __launch_bounds__(256) void mask_corr_kernel(const float *__restrict__ mu, const float *__restrict__ mask,
                                                        float *__restrict__ out, int K, int Z, int X, int Y, int Mz,
                                                        int Mx, int My, int hz, int hx, int hy, int flip, int nxb)
{
    const int y = blockIdx.x * 64 + (threadIdx.x & 63);
    const int xb = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int kz = blockIdx.z;
    const int k = kz / Z, z = kz % Z;
    if (xb >= nxb) return;
    const int x0 = xb * kXB;
    const int64_t XY = (int64_t)X * Y;
    const float *muk = mu + (int64_t)k * Z * XY;
    float acc[kXB];
#pragma unroll
    for (int j = 0; j < kXB; ++j) acc[j] = 0.0f;

    for (int dz = 0; dz < Mz; ++dz) {
        const int zz = z + dz - hz;
        if (zz < 0 || zz >= Z) continue;
        const float *plane = muk + (int64_t)zz * XY;
        for (int dy = 0; dy < My; ++dy) {
            const int yy = y + dy - hy;
            const bool yok = yy >= 0 && yy < Y;
            // mask[dz, :, dy] with stride My, or the flipped mask M[Mz-1-dz, Mx-1-dx, My-1-dy]
            const float *mcol = flip ? mask + ((int64_t)(Mz - 1 - dz) * Mx + (Mx - 1)) * My + (My - 1 - dy)
                                     : mask + (int64_t)dz * Mx * My + dy;
            const int64_t mstride = flip ? -(int64_t)My : (int64_t)My;
            // input rows x0 - hx + m feed outputs x0 + j through tap dx = m - j
            const int mfirst = max(0, hx - x0), mlast = min(kXB + Mx - 1, X - x0 + hx);
            for (int m = mfirst; m < mlast; ++m) {
                const float v = yok ? plane[(int64_t)(x0 - hx + m) * Y + yy] : 0.0f;
#pragma unroll
                for (int j = 0; j < kXB; ++j) {
                    const int dx = m - j;
                    if (dx >= 0 && dx < Mx) {
                        const float mv = mcol[dx * mstride];
                        if (mv != 0.0f) acc[j] = fmaf(mv, v, acc[j]);
                    }
                }
            }
        }
    }
    if (y >= Y) return;
    float *o = out + (int64_t)k * Z * XY + (int64_t)z * XY + y;
#pragma unroll
    for (int j = 0; j < kXB; ++j)
        if (x0 + j < X) o[(int64_t)(x0 + j) * Y] = acc[j];
}

// ---------------------------------------------------------------------------
// Rotation: F.affine_grid([[c,-s,0],[s,c,0]]) + F.grid_sample(bilinear, zeros,
// align_corners=False), as torch_rotate_image_nd (FL/torch_image_util.py:13-114).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rotate_kernel(const float *__restrict__ x, float *__restrict__ out, int64_t N,
                                                     int H, int W, float c, float s)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (j >= W || i >= H) return;
    int y0, x0;
    float d, e;
    rot_sample(i, j, H, W, c, s, y0, x0, d, e);
    const float w_ = 1.0f - e, n_ = 1.0f - d;
    const bool x0ok = x0 >= 0 && x0 < W, x1ok = x0 + 1 >= 0 && x0 + 1 < W;
    const bool y0ok = y0 >= 0 && y0 < H, y1ok = y0 + 1 >= 0 && y0 + 1 < H;
    for (int64_t n = blockIdx.z; n < N; n += gridDim.z) {
    const float *img = x + n * H * W;
    float v = 0.0f;
    if (y0ok && x0ok) v += img[(int64_t)y0 * W + x0] * (w_ * n_);
    if (y0ok && x1ok) v += img[(int64_t)y0 * W + x0 + 1] * (e * n_);
    if (y1ok && x0ok) v += img[(int64_t)(y0 + 1) * W + x0] * (w_ * d);
    if (y1ok && x1ok) v += img[(int64_t)(y0 + 1) * W + x0 + 1] * (e * d);
    out[n * H * W + (int64_t)i * W + j] = v;
    }
}

// ---------------------------------------------------------------------------
// Incident beam (torch_xray_atten_3D, torch_functions.py:92-102):
//   A[e,b,H-1,c] = 1,  A[e,b,r,c] = prod_{m=r+1}^{H-1} (1 - frac[e,b,m,c]*coef[e])
// coef == nullptr: frac already holds mu_x.  The running product starts at row
// H-1 and multiplies toward row 0 (the order of cumprod over the flipped rows).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void incident_kernel(const float *__restrict__ frac, const float *__restrict__ coef,
                                                       float *__restrict__ out, int64_t frac_estride,
                                                       int64_t out_estride, int K, int B, int H, int W)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y, e = blockIdx.z;
    if (c >= W) return;
    const int64_t HW = (int64_t)H * W;
    const float *f = frac + e * frac_estride + b * HW + c;
    float *o = out + e * out_estride + b * HW + c;
    const float k = coef ? coef[e] : 1.0f;
    o[(int64_t)(H - 1) * W] = 1.0f;
    float p = 1.0f;
    for (int r = H - 2; r >= 0; --r) {
        const float mu = coef ? f[(int64_t)(r + 1) * W] * k : f[(int64_t)(r + 1) * W];
        const float t = 1.0f - mu;
        p = (r == H - 2) ? t : p * t;
        o[(int64_t)r * W] = p;
    }
}

// ---------------------------------------------------------------------------
// Fluorescence-line attenuation coefficients in the mask frame
// (torch_extract_3D_mu_xrf, torch_functions.py:71-81, then rot90 'clock'):
//   mu_rot[e, b, c, H-1-r] = sum_j frac[j,b,r,c] * coef[j*K + e]
// Tile 32 rows x 64 cols through LDS so both the read and the write coalesce.
// ---------------------------------------------------------------------------
constexpr int kTR = 32, kTC = 64, kMaxLines = 4;

// Offset of original-frame voxel (r, c) inside one (X, Y) plane of the mask frame.
// side 0, detector at +col ('r', rot90 'clock'):      X = c,       Y = H-1-r
// side 1, detector at -col ('l', rot90 'c-clock'):    X = W-1-c,   Y = r
__device__ __forceinline__ int64_t mask_frame_offset(int r, int c, int H, int W, int side)
{
    return side == 0 ? (int64_t)c * H + (H - 1 - r) : (int64_t)(W - 1 - c) * H + r;
}

__global__ __launch_bounds__(256) void mu_rot_kernel(const float *__restrict__ frac, const float *__restrict__ coef,
                                                     float *__restrict__ mu_rot, int Kc, int K, int B, int H, int W,
                                                     int side)
{
    __shared__ float tile[kMaxLines][kTR][kTC + 1];
    const int c0 = blockIdx.x * kTC, r0 = blockIdx.y * kTR;
    const int b = blockIdx.z % B, echunk = blockIdx.z / B;
    const int e0 = echunk * kMaxLines, ne = min(kMaxLines, K - e0);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t HW = (int64_t)H * W, BHW = (int64_t)B * HW;
    for (int rr = wave; rr < kTR; rr += 4) {
        const int r = r0 + rr, c = c0 + lane;
        float acc[kMaxLines];
#pragma unroll
        for (int e = 0; e < kMaxLines; ++e) acc[e] = 0.0f;
        if (r < H && c < W) {
            for (int j = 0; j < Kc; ++j) {
                const float fj = frac[j * BHW + b * HW + (int64_t)r * W + c];
#pragma unroll
                for (int e = 0; e < kMaxLines; ++e)
                    if (e < ne) acc[e] = acc[e] + fj * coef[j * K + e0 + e];
            }
        }
#pragma unroll
        for (int e = 0; e < kMaxLines; ++e)
            if (e < ne) tile[e][rr][lane] = acc[e];
    }
    __syncthreads();
    // write: for column c, the mask-frame row index is y = H-1-r (contiguous in mu_rot)
    const int half = lane >> 5, rr = lane & 31;
    for (int e = 0; e < ne; ++e)
        for (int cc = wave * 2 + half; cc < kTC; cc += 8) {
            const int c = c0 + cc, r = r0 + rr;
            if (c < W && r < H)
                mu_rot[(int64_t)(e0 + e) * BHW + b * HW + mask_frame_offset(r, c, H, W, side)] = tile[e][rr][cc];
        }
}

// ---------------------------------------------------------------------------
// One direction of the sampled detector cone, mask frame (Z, X, Y), cone along +X.
//   J(v) = sum_{t>=1} bilin0(mu[., x+t, .], z + t*pz/q, y + t*py/q)
// With D = (pz, q, py) integer, J(v) = sum_{t=1}^{q} bilin0(...) + J(v + D): every
// voxel lies on one lattice chain v0 + m*D, and each thread walks one chain from
// its far end, so the recursion is exact (no re-interpolation of J).  The q-step
// stencil (<= 4q taps with exact rational weights) comes from the host.
//   T[v] (+)= weight * J(v)
// reverse = 1 runs the transpose (the adjoint used by the gradient): taps with
// negated offsets, chains walked from the near end, J(v) = S(v) + J(v - D).
// ---------------------------------------------------------------------------
//
// One thread per chain: lanes take neighbouring chains along y, each wave
// one z.  Every tap of the q-step stencil sits at a constant offset from the
// chain point (dz*X*Y + t*Y + dy floats), so a step issues all taps as
// independent bounds-checked buffer loads from one per-lane base (no per-tap
// address arithmetic, no branches), summed in four partial sums.  When every
// lane's stencil lies inside in y, the two taps of a bilinear sample that are
// neighbours in y form one slot: each lane loads the lower one, and the upper
// one is the next lane's lower one (the lanes' chains are neighbours in y at the
// same z and x), moved across by a wave shift.  A wave therefore owns 63 chains:
// lane 63 walks the next chain only to supply lane 62's upper taps and writes
// nothing (the next block's lane 0 owns that chain).  That halves the vector
// loads (2.5% less time at 256^3 x 4 lines: load issue is not the main
// limit of the march).  Lanes whose stencil lies wholly outside the
// volume in y get a base past the buffer (their loads read 0).  A wave with a
// lane whose stencil straddles a y edge, or whose stencil leaves the volume in
// z or x, checks every tap instead; such steps
// occur only where chains cross the volume's faces.  Steps where the wave's
// stencils and chain points all lie outside are skipped (nothing to add or
// write).  The stencil table is padded to a power of two with zero-weight taps
// at offset 0.
// ---------------------------------------------------------------------------
constexpr int kConeMaxTaps = 128;
// Base byte offset that stays at or past every buffer end (buffers hold < 2^31
// bytes) for every tap offset (|offset| < 2^29 bytes), without wrapping.
constexpr uint32_t kConeOOB = 0xC0000000u;

struct ConeArgs {
    int pz, py, q;
    float weight;
    int accumulate;
    int reverse;
    int zlo, nz, ylo, ny;           // chain start ranges (mask frame)
    int dzmin, dzmax, dymin, dymax;  // stencil extent
    int tmin, tmax;
    int ezlo, ezhi, eylo, eyhi;      // stencil extent including the chain point itself
    int off[kConeMaxTaps];           // tap byte offsets
    int8_t t[kConeMaxTaps];
    int8_t dz[kConeMaxTaps];
    int8_t dy[kConeMaxTaps];
    float w[kConeMaxTaps];
    int soff[kConeMaxTaps];          // slots: byte offset of the lower tap in y
    float slo[kConeMaxTaps], shi[kConeMaxTaps];  // weights of the lower and upper tap (0: none)
};

template <int NT, int NS>
__global__ __launch_bounds__(256) void cone_kernel(const float *__restrict__ mu, float *__restrict__ T, int Z, int X,
                                                   int Y, ConeArgs p)
{
    // The tap table, staged once per workgroup and read back every step (one
    // broadcast LDS read per tap): kept in registers across the walk it would
    // not fit the scalar registers and be spilled.
    __shared__ int4 tab[NT];   // (byte offset, weight bits, dy, dz | t << 16)
    __shared__ int4 stab[NS];  // slots: (byte offset, lower weight bits, upper weight bits, 0)
    if ((int)threadIdx.x < NT)
        tab[threadIdx.x] = make_int4(p.off[threadIdx.x], __float_as_int(p.w[threadIdx.x]), p.dy[threadIdx.x],
                                     (p.dz[threadIdx.x] & 0xffff) | (p.t[threadIdx.x] << 16));
    if ((int)threadIdx.x < NS)
        stab[threadIdx.x] =
            make_int4(p.soff[threadIdx.x], __float_as_int(p.slo[threadIdx.x]), __float_as_int(p.shi[threadIdx.x]), 0);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // Block order: the chain plane x0 fastest, then the y block, the z group and
    // the line.  The q chains of one (y, z) tile read the same planes of mu, and
    // neighbouring z groups share rows: remapped so that each XCD receives a
    // contiguous run of this order, they run together and share its L2.
    const uint32_t nyb = (uint32_t)(p.ny + 62) / 63u, nzb = (uint32_t)(p.nz + 3) / 4u;
    uint32_t id = xcd_remap(blockIdx.x, gridDim.x);
    const int x0 = (int)(id % (uint32_t)p.q);
    id /= (uint32_t)p.q;
    const int yb = (int)(id % nyb);
    id /= nyb;
    const int zb = (int)(id % nzb);
    const int k = (int)(id / nzb);
    const int zi = zb * 4 + uniform_int(threadIdx.x >> 6);
    if (zi >= p.nz) return;  // the whole wave
    const int yi = yb * 63 + lane;
    const bool act = yi < p.ny;
    const int64_t ZXY = (int64_t)Z * X * Y;
    const float *muk = mu + k * ZXY;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(muk, (int)ZXY);
    float *Tk = T + k * ZXY;
    const __amdgpu_buffer_rsrc_t rt = make_rsrc(Tk, (int)ZXY);
    const int z0 = p.zlo + zi;
    const int yw0 = p.ylo + yb * 63;  // lane 0's chain start
    const int M = (X - x0 + p.q - 1) / p.q;        // chain points with x < X
    float J = 0.0f;
    for (int it = 0; it < M; ++it) {
        const int m = p.reverse ? it : M - 1 - it;
        const int x = x0 + m * p.q;
        const int z = z0 + m * p.pz;
        const int yw = yw0 + m * p.py;
        const int y = yw + lane;
        if (z + p.ezhi < 0 || z + p.ezlo >= Z || yw + 63 + p.eyhi < 0 || yw + p.eylo >= Y) continue;
        const bool yin = act && y + p.dymin >= 0 && y + p.dymax < Y;
        const bool yout = !act || y + p.dymax < 0 || y + p.dymin >= Y;
        const bool zx_in = z + p.dzmin >= 0 && z + p.dzmax < Z && x + p.tmin >= 0 && x + p.tmax < X;
        // chain point as a float index; wraps harmlessly where it is not used
        const uint32_t base = ((uint32_t)z * (uint32_t)X + (uint32_t)x) * (uint32_t)Y + (uint32_t)y;
        // the running sum of earlier directions, loaded with the taps
        const bool wr = act & (lane < 63) & ((uint32_t)z < (uint32_t)Z) & ((uint32_t)y < (uint32_t)Y);
        const float tprev = p.accumulate ? load_f32(rt, wr ? base * 4u : kConeOOB) : 0.0f;
        float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        __asm__ volatile("" ::: "memory");  // re-read the tap table at every step
        if (zx_in && __all(yin)) {
            const uint32_t vb = base * 4u;
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                const int4 sl = stab[i];
                const float lo = load_f32(rs, vb + (uint32_t)sl.x);
                // the next lane's lower tap is this lane's upper one (wave_shl:1;
                // lane 63 gets 0 and is not written)
                const float hi = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(lo), 0x130, 0xf, 0xf, false));
                s[i & 3] = fmaf(__int_as_float(sl.y), lo, fmaf(__int_as_float(sl.z), hi, s[i & 3]));
            }
        } else if (zx_in && __all(yin || yout)) {
            const uint32_t vb = yin ? base * 4u : kConeOOB;
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const int4 tp = tab[i];
                s[i & 3] = fmaf(__int_as_float(tp.y), load_f32(rs, vb + (uint32_t)tp.x), s[i & 3]);
            }
        } else if (zx_in) {
            // Some lane's stencil crosses a y face (the taps are inside in z and x):
            // check the column of every tap, lane by lane.
            const uint32_t b4 = base * 4u;
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const int4 tp = tab[i];
                const bool ok = act & ((uint32_t)(y + tp.z) < (uint32_t)Y);
                s[i & 3] = fmaf(__int_as_float(tp.y), load_f32(rs, ok ? b4 + (uint32_t)tp.x : kConeOOB), s[i & 3]);
            }
        } else {
#pragma unroll
            for (int i = 0; i < NT; ++i) {
                const int4 tp = tab[i];
                const int zz = z + (int16_t)(tp.w & 0xffff), xx = x + (tp.w >> 16), yy = y + tp.z;
                const bool ok = act & ((uint32_t)zz < (uint32_t)Z) & ((uint32_t)xx < (uint32_t)X) &
                                ((uint32_t)yy < (uint32_t)Y);
                const uint32_t v = ok ? base * 4u + (uint32_t)tp.x : kConeOOB;
                s[i & 3] = fmaf(__int_as_float(tp.y), load_f32(rs, v), s[i & 3]);
            }
        }
        J += (s[0] + s[1]) + (s[2] + s[3]);
        if (wr) Tk[base] = p.accumulate ? fmaf(p.weight, J, tprev) : p.weight * J;
    }
}

// ---------------------------------------------------------------------------
// out[e,b,r,c] = exp(-T_rot[e,b,c,H-1-r]) * (incident ? incident[e,b,r,c] : 1)
// (torch_functions.py:170-174, then core.py:99 for the product), via an LDS tile.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void combine_kernel(const float *__restrict__ T_rot,
                                                      const float *__restrict__ incident, float *__restrict__ out,
                                                      int64_t inc_estride, int64_t out_estride, int B, int H, int W,
                                                      int side)
{
    __shared__ float tile[kTR][kTC + 1];
    const int c0 = blockIdx.x * kTC, r0 = blockIdx.y * kTR;
    const int b = blockIdx.z % B, e = blockIdx.z / B;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t HW = (int64_t)H * W;
    const float *Te = T_rot + (int64_t)e * B * HW + b * HW;
    const int half = lane >> 5, rr = lane & 31;
    for (int cc = wave * 2 + half; cc < kTC; cc += 8) {
        const int c = c0 + cc, r = r0 + rr;
        if (c < W && r < H) tile[rr][cc] = Te[mask_frame_offset(r, c, H, W, side)];
    }
    __syncthreads();
    for (int r_ = wave; r_ < kTR; r_ += 4) {
        const int r = r0 + r_, c = c0 + lane;
        if (r >= H || c >= W) continue;
        float v = expf(-tile[r_][lane]);
        if (incident) v *= incident[e * inc_estride + b * HW + (int64_t)r * W + c];
        out[e * out_estride + b * HW + (int64_t)r * W + c] = v;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int64_t floordiv(int64_t a, int64_t b)
{
    int64_t q = a / b, r = a % b;
    return (r != 0 && ((r < 0) != (b < 0))) ? q - 1 : q;
}

int launch_mask_corr(const float *mu, const float *mask, float *out, int64_t K, int64_t Z, int64_t X, int64_t Y,
                     int64_t Mz, int64_t Mx, int64_t My, int64_t hz, int64_t hx, int64_t hy, int flip, hipStream_t s)
{
    const int nxb = (int)((X + kXB - 1) / kXB);
    dim3 grid((unsigned)((Y + 63) / 64), (unsigned)((nxb + 3) / 4), (unsigned)(K * Z));
    FL_REQUIRE(K * Z <= 65535, "mask correlation: K*Z must be <= 65535");
    hipLaunchKernelGGL(mask_corr_kernel, grid, dim3(256), 0, s, mu, mask, out, (int)K, (int)Z, (int)X, (int)Y,
                       (int)Mz, (int)Mx, (int)My, (int)hz, (int)hx, (int)hy, flip, nxb);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_rotate(const float *x, float *out, int64_t N, int64_t H, int64_t W, float c, float s_, hipStream_t s)
{
    dim3 grid((unsigned)((W + 63) / 64), (unsigned)((H + 3) / 4), (unsigned)std::min<int64_t>(N, 65535));
    hipLaunchKernelGGL(rotate_kernel, grid, dim3(256), 0, s, x, out, N, (int)H, (int)W, c, s_);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_incident(const float *frac, const float *coef, float *out, int64_t frac_estride, int64_t out_estride,
                    int64_t K, int64_t B, int64_t H, int64_t W, hipStream_t s)
{
    dim3 grid((unsigned)((W + 255) / 256), (unsigned)B, (unsigned)K);
    hipLaunchKernelGGL(incident_kernel, grid, dim3(256), 0, s, frac, coef, out, frac_estride, out_estride, (int)K,
                       (int)B, (int)H, (int)W);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_mu_rot(const float *frac, const float *coef, float *mu_rot, int64_t Kc, int64_t K, int64_t B, int64_t H,
                  int64_t W, int side, hipStream_t s)
{
    const int64_t echunks = (K + kMaxLines - 1) / kMaxLines;
    dim3 grid((unsigned)((W + kTC - 1) / kTC), (unsigned)((H + kTR - 1) / kTR), (unsigned)(B * echunks));
    hipLaunchKernelGGL(mu_rot_kernel, grid, dim3(256), 0, s, frac, coef, mu_rot, (int)Kc, (int)K, (int)B, (int)H,
                       (int)W, side);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_combine(const float *T_rot, const float *incident, float *out, int64_t inc_estride, int64_t out_estride,
                   int64_t K, int64_t B, int64_t H, int64_t W, int side, hipStream_t s)
{
    dim3 grid((unsigned)((W + kTC - 1) / kTC), (unsigned)((H + kTR - 1) / kTR), (unsigned)(B * K));
    hipLaunchKernelGGL(combine_kernel, grid, dim3(256), 0, s, T_rot, incident, out, inc_estride, out_estride, (int)B,
                       (int)H, (int)W, side);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

// Builds the q-step stencil of direction (pz, py, q) and launches one chain walk.
int launch_cone_direction(const float *mu, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y, int pz, int py,
                          int q, float weight, int accumulate, int reverse, hipStream_t s)
{
    FL_REQUIRE(q >= 1, "cone direction: q must be >= 1");
    // reduce the rational slopes so the chain step is as short as possible
    int g = std::gcd(std::gcd(std::abs(pz), std::abs(py)), q);
    if (g > 1) {
        pz /= g;
        py /= g;
        q /= g;
    }
    FL_REQUIRE(4 * q <= kConeMaxTaps, "cone direction: reduced denominator too large (max 32)");
    FL_REQUIRE(std::abs(pz) <= 64 * q && std::abs(py) <= 64 * q, "cone direction: slope too steep");
    FL_REQUIRE(Z * X * Y < ((int64_t)1 << 29), "cone: one line's volume must hold fewer than 2^29 voxels");
    ConeArgs p{};
    p.pz = pz;
    p.py = py;
    p.q = q;
    p.weight = weight;
    p.accumulate = accumulate;
    p.reverse = reverse;
    const int sg = reverse ? -1 : 1;
    int n = 0, ns = 0;
    p.dzmin = p.dymin = p.tmin = INT32_MAX;
    p.dzmax = p.dymax = p.tmax = INT32_MIN;
    for (int t = 1; t <= q; ++t) {
        const int64_t nz = (int64_t)t * pz, ny = (int64_t)t * py;
        const int64_t z0 = floordiv(nz, q), y0 = floordiv(ny, q);
        const double fz = (double)(nz - z0 * q) / q, fy = (double)(ny - y0 * q) / q;
        for (int a = 0; a < 2; ++a) {
            const int n0 = n;  // this row's taps: b = 0 always (1 - fy > 0), b = 1 if fy > 0
            for (int b = 0; b < 2; ++b) {
                const double wz = a ? fz : 1.0 - fz, wy = b ? fy : 1.0 - fy;
                if (wz == 0.0 || wy == 0.0) continue;
                FL_REQUIRE(n < kConeMaxTaps && std::abs(z0 + a) < 127 && std::abs(y0 + b) < 127,
                           "cone direction: stencil too large");
                const int tt = sg * t, dz = (int)(sg * (z0 + a)), dy = (int)(sg * (y0 + b));
                const int64_t off = ((int64_t)dz * X * Y + (int64_t)tt * Y + dy) * 4;
                FL_REQUIRE(std::llabs(off) < ((int64_t)1 << 29), "cone direction: stencil reach too large");
                p.t[n] = (int8_t)tt;
                p.dz[n] = (int8_t)dz;
                p.dy[n] = (int8_t)dy;
                p.off[n] = (int)off;
                p.w[n] = (float)(wz * wy);
                p.dzmin = std::min(p.dzmin, dz);
                p.dzmax = std::max(p.dzmax, dz);
                p.dymin = std::min(p.dymin, dy);
                p.dymax = std::max(p.dymax, dy);
                p.tmin = std::min(p.tmin, tt);
                p.tmax = std::max(p.tmax, tt);
                ++n;
            }
            if (n == n0) continue;
            // one slot per row: its lower tap in y (b = 1 lies at +1 in y, or at
            // -1 when reverse) and the tap right above it, if any
            if (n - n0 == 2) {
                const bool lo0 = p.off[n0] < p.off[n0 + 1];
                p.soff[ns] = lo0 ? p.off[n0] : p.off[n0 + 1];
                p.slo[ns] = lo0 ? p.w[n0] : p.w[n0 + 1];
                p.shi[ns] = lo0 ? p.w[n0 + 1] : p.w[n0];
            } else {
                p.soff[ns] = p.off[n0];
                p.slo[ns] = p.w[n0];
                p.shi[ns] = 0.0f;
            }
            ++ns;
        }
    }
    p.ezlo = std::min(p.dzmin, 0);
    p.ezhi = std::max(p.dzmax, 0);
    p.eylo = std::min(p.dymin, 0);
    p.eyhi = std::max(p.dymax, 0);
    int NT = 4;
    while (NT < n) NT *= 2;
    // slots are between n/2 and n, so NT/2 or NT of them (padded alike)
    const bool half = ns <= NT / 2;
    // the padding taps and slots (weight 0, offset 0) read the chain point itself
    const int64_t Mmax = (X + q - 1) / q;
    // chain starts in plane x0 in [0,q): z0 = z - m*pz for z in [0,Z), m in [0,Mmax);
    // one extra voxel each side for the bilinear reach
    const int64_t zlo = -(Mmax - 1) * std::max(pz, 0) - 1, zhi = Z - 1 - (Mmax - 1) * std::min(pz, 0) + 1;
    const int64_t ylo = -(Mmax - 1) * std::max(py, 0) - 1, yhi = Y - 1 - (Mmax - 1) * std::min(py, 0) + 1;
    p.zlo = (int)zlo;
    p.nz = (int)(zhi - zlo + 1);
    p.ylo = (int)ylo;
    p.ny = (int)(yhi - ylo + 1);
    const int64_t nblocks = (int64_t)((p.ny + 62) / 63) * ((p.nz + 3) / 4) * K * q;
    FL_REQUIRE(nblocks < ((int64_t)1 << 31), "cone direction: grid too large");
    dim3 grid((unsigned)nblocks);
#define FL_CONE_CASE(nt)                                                                              \
    case nt:                                                                                          \
        if (half)                                                                                     \
            cone_kernel<nt, (nt > 4 ? nt / 2 : 4)><<<grid, dim3(256), 0, s>>>(mu, T, (int)Z, (int)X, (int)Y, p); \
        else                                                                                          \
            cone_kernel<nt, nt><<<grid, dim3(256), 0, s>>>(mu, T, (int)Z, (int)X, (int)Y, p);        \
        break;
    switch (NT) {
        FL_CONE_CASE(4)
        FL_CONE_CASE(8)
        FL_CONE_CASE(16)
        FL_CONE_CASE(32)
        FL_CONE_CASE(64)
        FL_CONE_CASE(128)
    }
#undef FL_CONE_CASE
    FL_LAUNCH_CHECK();
    return FL_OK;
}

}  // namespace fl
