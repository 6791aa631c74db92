// Shared definitions for the MI355X (gfx950) kernels of the XRF self-absorption
// hot path.  Everything here is device-side HIP written for CDNA4: 64-lane
// wavefronts, 160 KB LDS per CU, 8 XCDs with private L2.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/fl_correction_amd.h"

namespace fl {

constexpr int kWave = 64;
constexpr int kNumXcd = 8;

// Thread-local message behind fl_last_error().
void set_error(const std::string &msg);

// Remap the hardware workgroup id so that the workgroups dispatched to one XCD
// (the dispatcher deals them round-robin: id % 8) cover one contiguous range of
// logical ids.  Logical ids are laid out so that neighbours share input data
// (the same slice of C, the same image tile), which then stays in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t id, uint32_t n)
{
    const uint32_t x = id % kNumXcd, l = id / kNumXcd;
    const uint32_t q = n / kNumXcd, r = n % kNumXcd;
    return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// The next float above a positive finite x (x < next_up(x) holds exactly
// for the values <= x).
__device__ __forceinline__ float next_up(float x) { return __int_as_float(__float_as_int(x) + 1); }

__device__ __forceinline__ float clampf(float v, float lo, float hi)
{
    return __builtin_amdgcn_fmed3f(v, lo, hi);
}

// Non-finite angles are read as 0: every sample coordinate then stays finite and
// every index stays inside the arrays.
__device__ __forceinline__ float finite_or_zero(float v) { return isfinite(v) ? v : 0.0f; }

// Sample point of output pixel (i, j) for torch's affine_grid([[c,-s,0],[s,c,0]])
// + grid_sample(align_corners=False) (FL/torch_image_util.py:77-104): the
// input position (iy, ix) split into its integer corner (y0, x0) and fractions.
// Shared by the rotation and its transpose so both use identical arithmetic.
__device__ __forceinline__ void rot_sample(int i, int j, int H, int W, float c, float s, int &y0, int &x0,
                                           float &fy, float &fx)
{
    const float xn = (float)(2 * j + 1) / (float)W - 1.0f;
    const float yn = (float)(2 * i + 1) / (float)H - 1.0f;
    const float gx = c * xn - s * yn;
    const float gy = s * xn + c * yn;
    const float ix = ((gx + 1.0f) * W - 1.0f) * 0.5f;
    const float iy = ((gy + 1.0f) * H - 1.0f) * 0.5f;
    const float fx0 = floorf(ix), fy0 = floorf(iy);
    x0 = (int)fx0;
    y0 = (int)fy0;
    fx = ix - fx0;
    fy = iy - fy0;
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane l's value v, for a wave-uniform l (v_readlane: no LDS round trip).
__device__ __forceinline__ int lane_value(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Minimum / maximum over the 64 lanes of a wave (wave-uniform result).
__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d));
    return uniform_int(v);
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d));
    return uniform_int(v);
}

// Inclusive prefix sum over the 64 lanes of a wave, in registers: Hillis-Steele
// shifts by 1, 2, 4, 8 within each row of 16 lanes, then row 0's total into
// row 1 and row 2's into row 3 (row_bcast:15), then rows 0-1's total into rows
// 2-3 (row_bcast:31).  Lanes without a source add 0.
__device__ __forceinline__ int wave_inclusive_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2 and 3
    return v;
}

// Raw buffer resource over n floats at p (CDNA: word 3 = 0x00020000, 32-bit
// data format).  Loads through it are range-checked by the hardware: a byte
// offset at or past 4*n reads 0 instead of touching memory.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float *p, int n)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, n * 4, 0x00020000);
}

// Read of a wave-uniform element through the scalar data cache (s_load: counted
// apart from the vector loads, so it does not wait for loads in flight).  Only
// for inputs that the kernel does not write.
__device__ __forceinline__ float scalar_read(const float *p, int64_t i)
{
    return ((__attribute__((address_space(4))) const float *)p)[i];
}

__device__ __forceinline__ uint32_t scalar_read_u32(const uint32_t *p, int64_t i)
{
    return ((__attribute__((address_space(4))) const uint32_t *)p)[i];
}

// A wave-uniform record (kernel argument block, job table entry) through the
// scalar data cache.
template <class T>
__device__ __forceinline__ T scalar_record(const T *p)
{
    static_assert(sizeof(T) % 4 == 0, "records are read in 32-bit words");
    T r;
    uint32_t *d = reinterpret_cast<uint32_t *>(&r);
    const auto *s = (__attribute__((address_space(4))) const uint32_t *)p;
#pragma unroll
    for (size_t i = 0; i < sizeof(T) / 4; ++i) d[i] = s[i];
    return r;
}

// Job of logical tile t in a grouped launch: the last j with begin[j] <= t
// (begin ascending, begin[0] = 0).  Wave-uniform.
__device__ __forceinline__ int find_job(const uint32_t *begin, int n_jobs, uint32_t t)
{
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (scalar_read_u32(begin, mid) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One float at byte offset off (0 past the end).
__device__ __forceinline__ float load_f32(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// Same, at byte offset voff + soff (soff wave-uniform), for data read once:
// marked non-temporal so that it does not evict reused lines from the caches.
__device__ __forceinline__ float load_f32_stream(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 2));
}

__device__ __forceinline__ float uniform_float(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

}  // namespace fl

#define FL_HIP_CHECK(expr)                                                            \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            fl::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));         \
            return FL_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

#define FL_LAUNCH_CHECK()                                                             \
    do {                                                                              \
        hipError_t _e = hipGetLastError();                                            \
        if (_e != hipSuccess) {                                                       \
            fl::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));    \
            return FL_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

#define FL_REQUIRE(cond, msg)                                                         \
    do {                                                                              \
        if (!(cond)) {                                                                \
            fl::set_error(msg);                                                       \
            return FL_ERR_INVALID_ARG;                                                \
        }                                                                             \
    } while (0)
