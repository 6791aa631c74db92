// cal_frac on the GPU (FL/core.py:159-211): the element sum, the adaptive
// threshold mask (FL/image_util.py:337-375) and the per-element fractions.
//
// frac_sum_kernel     img_sum[p] = ((img[0,p] + img[1,p]) + ...) + 1e-12, in the
//                     input's type (numpy's sum over axis 0 of a C-ordered stack
//                     adds the planes in order)
// adaptive_mask_kernel  one workgroup per image (a 2-D slice):
//   1. three-class multi-Otsu threshold (skimage.filters.threshold_multiotsu,
//      nbins = 256): numpy's histogram over [min, max] (float edges from
//      np.linspace in the input's type, the bin index ((x - min) / (max - min))
//      * 256 truncated and corrected by one against the edges), bin
//      probabilities count / total rounded to float32, the between-class
//      variance sum_k S_k^2 / P_k of every split (i1, i2) from float32
//      cumulative moments, the first maximum in (i1, i2) order; the mask is
//      img >= the centre of bin i1 (np.digitize ... < 1 is img < threshold);
//   2. binary_fill_holes with a 5 x 5 structure: background pixels reached from
//      outside the image (border value 1) through background within Chebyshev
//      distance 2, propagated to convergence; the rest is foreground;
//   3. `dilation` binary dilations, then `erosion` binary erosions, with the
//      4-neighbour cross (border value 0).
//   The image's bits live in LDS, 32 pixels of a row per word.  A histogram with
//   fewer than three nonzero bins, or a non-finite image, fails as
//   threshold_multiotsu raises: the reference's except branch zeroes that
//   image's mask.
// frac_ratio_kernel   tmp = rm_abnormal(img[i] / img_sum) in double (the
//                     reference's img_sum is float64 once multiplied by the
//                     float64 mask)
// frac_final_kernel   frac = (clip(tmp, lo, hi) * mask) * ratio
// With `single` both round every quotient and product to float32: the
// reference's arithmetic when its mask falls back to the integer 1 and a
// float32 stack stays float32 (float32 operands computed in double and rounded
// once to float32 give the correctly rounded float32 result).
#include "common.hpp"

#include <cmath>

namespace fl {

template <typename T>
__global__ __launch_bounds__(256) void frac_sum_kernel(const T *__restrict__ img, T *__restrict__ out, int64_t n,
                                                       int64_t npix)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    T s = img[p];
    for (int64_t i = 1; i < n; ++i) s = s + img[i * npix + p];
    out[p] = s + (T)1e-12;
}

constexpr int kOtsuBins = 256;
constexpr int kMaskThreads = 1024;

template <typename T>
__device__ __forceinline__ bool finite_val(T v)
{
    return v == v && v - v == (T)0;  // not NaN and not +-inf
}

// Between-class variance of bins [i, j] from the cumulative moments (float32).
__device__ __forceinline__ float otsu_var(const float *P, const float *S, int i, int j)
{
    const float z = i == 0 ? P[j] : P[j] - P[i - 1];
    const float f = i == 0 ? S[j] : S[j] - S[i - 1];
    return z > 0.0f ? f * f / z : 0.0f;
}

template <typename T>
__global__ __launch_bounds__(kMaskThreads) void adaptive_mask_kernel(const T *__restrict__ img,
                                                                     double *__restrict__ mask, int H, int W,
                                                                     int fill_hole, int dilation, int erosion,
                                                                     int *__restrict__ status,
                                                                     T *__restrict__ thresholds)
{
    extern __shared__ uint32_t bits[];  // two bit planes of H rows x Wd words
    __shared__ int hist[kOtsuBins];
    __shared__ T edges[kOtsuBins + 1];
    __shared__ float P[kOtsuBins], S[kOtsuBins];
    __shared__ T red_lo[kMaskThreads / 64], red_hi[kMaskThreads / 64];
    __shared__ float best_v[kMaskThreads / 64];
    __shared__ int best_i[kMaskThreads / 64];
    __shared__ int changed, fail, thr_idx;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t npix = (int64_t)H * W;
    const T *im = img + (int64_t)blockIdx.x * npix;
    double *mk = mask + (int64_t)blockIdx.x * npix;
    const int Wd = (W + 31) / 32;
    const int nw = H * Wd;
    uint32_t *A = bits, *Bp = bits + nw;

    // ---- 1. min / max -------------------------------------------------------
    T lo = im[0], hi = im[0];
    bool fin = true;
    for (int64_t p = tid; p < npix; p += kMaskThreads) {
        const T v = im[p];
        fin = fin && finite_val(v);
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        const T ol = __shfl_xor(lo, d), oh = __shfl_xor(hi, d);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    const bool wfin = __all(fin);
    if (tid < kOtsuBins) hist[tid] = 0;
    if (tid == 0) fail = 0;
    __syncthreads();
    if (lane == 0) {
        red_lo[wave] = lo;
        red_hi[wave] = hi;
        if (!wfin) fail = 1;
    }
    __syncthreads();
    lo = red_lo[0];
    hi = red_hi[0];
    for (int w = 1; w < kMaskThreads / 64; ++w) {
        lo = red_lo[w] < lo ? red_lo[w] : lo;
        hi = red_hi[w] > hi ? red_hi[w] : hi;
    }
    if (fail) {  // np.histogram: "autodetected range ... is not finite"
        for (int64_t p = tid; p < npix; p += kMaskThreads) mk[p] = 0.0;
        if (tid == 0) status[blockIdx.x] = 1;
        return;
    }
    // np.histogram's outer edges: an empty range widens by 0.5 each way
    T first = lo, last = hi;
    if (first == last) {
        first = first - (T)0.5;
        last = last + (T)0.5;
    }
    // ---- 2. histogram (np.histogram, 256 equal bins) -------------------------
    // edges: np.linspace(first, last, 257) in T: arange * step + first, the last
    // edge set to `last` (a zero step takes arange / 256 * delta)
    if (tid <= kOtsuBins) {
        const T delta = last - first;
        const T step = delta / (T)kOtsuBins;
        T y = step == (T)0 ? ((T)tid / (T)kOtsuBins) * delta : (T)tid * step;
        y = y + first;
        edges[tid] = tid == kOtsuBins ? last : y;
    }
    __syncthreads();
    {
        const T denom = last - first;
        for (int64_t p = tid; p < npix; p += kMaskThreads) {
            const T v = im[p];
            int idx = (int)(((v - first) / denom) * (T)kOtsuBins);
            if (idx == kOtsuBins) idx -= 1;
            if (v < edges[idx]) idx -= 1;
            if (idx != kOtsuBins - 1 && v >= edges[idx + 1]) idx += 1;
            atomicAdd(&hist[idx], 1);  // integer counts: the order does not matter
        }
    }
    __syncthreads();
    // ---- 3. three-class split --------------------------------------------------
    if (tid == 0) {
        int nz = 0;
        for (int i = 0; i < kOtsuBins; ++i) {
            // normalised histogram (float64 division), then .astype(float32)
            P[i] = (float)((double)hist[i] / (double)npix);
            nz += P[i] != 0.0f;
        }
        // cumulative moments in bin order
        float z = P[0], f = 0.0f * P[0];
        float p0 = P[0];
        S[0] = f;
        for (int i = 1; i < kOtsuBins; ++i) {
            const float pi = P[i];
            z = z + pi;
            f = f + (float)i * pi;
            P[i] = z;
            S[i] = f;
        }
        P[0] = p0;
        thr_idx = -1;
        if (nz < 3) {
            fail = 1;  // "The input image has only ... different values"
        } else if (nz == 3) {
            // the first two nonzero bins (np.flatnonzero(prob)[:-1])
            int k = 0;
            for (int i = 0; i < kOtsuBins && k == 0; ++i)
                if (hist[i] != 0) {
                    thr_idx = i;
                    k = 1;
                }
        }
        changed = nz;
    }
    __syncthreads();
    if (fail) {
        for (int64_t p = tid; p < npix; p += kMaskThreads) mk[p] = 0.0;
        if (tid == 0) status[blockIdx.x] = 1;
        return;
    }
    if (changed > 3) {
        // thread i1 scans i2 = i1+1 .. 254, keeping its first maximum; the
        // first maximum over all threads in (i1, i2) order wins
        float bv = 0.0f;
        int bi = -1;
        if (tid < kOtsuBins - 2) {
            const int i1 = tid;
            const float v0 = otsu_var(P, S, 0, i1);
            for (int i2 = i1 + 1; i2 < kOtsuBins - 1; ++i2) {
                const float sg = (v0 + otsu_var(P, S, i1 + 1, i2)) + otsu_var(P, S, i2 + 1, kOtsuBins - 1);
                if (sg > bv) {
                    bv = sg;
                    bi = i1;
                }
            }
        }
        // wave and block reductions: larger value, then smaller i1
        for (int d = 32; d >= 1; d >>= 1) {
            const float ov = __shfl_xor(bv, d);
            const int oi = __shfl_xor(bi, d);
            const bool take = oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi));
            bv = take ? ov : bv;
            bi = take ? oi : bi;
        }
        if (lane == 0) {
            best_v[wave] = bv;
            best_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            float v = best_v[0];
            int i = best_i[0];
            for (int w = 1; w < kMaskThreads / 64; ++w) {
                const bool take = best_i[w] >= 0 && (i < 0 || best_v[w] > v || (best_v[w] == v && best_i[w] < i));
                v = take ? best_v[w] : v;
                i = take ? best_i[w] : i;
            }
            thr_idx = i;
        }
        __syncthreads();
    }
    if (thr_idx < 0) {  // no split with a positive variance
        for (int64_t p = tid; p < npix; p += kMaskThreads) mk[p] = 0.0;
        if (tid == 0) status[blockIdx.x] = 1;
        return;
    }
    // bin centre (bin_edges[:-1] + bin_edges[1:]) / 2
    const T t0 = (edges[thr_idx] + edges[thr_idx + 1]) / (T)2;
    if (tid == 0) {
        thresholds[blockIdx.x] = t0;
        status[blockIdx.x] = 0;
    }

    // ---- 4. the mask's bits: A = foreground (img >= t0) ------------------------
    for (int w = tid; w < nw; w += kMaskThreads) {
        const int y = w / Wd, x0 = (w % Wd) * 32;
        uint32_t m = 0;
        for (int b = 0; b < 32 && x0 + b < W; ++b) m |= (uint32_t)(im[(int64_t)y * W + x0 + b] >= t0) << b;
        A[w] = m;
    }
    __syncthreads();
    // bits of a row past the image's last column
    const int tail = W - (Wd - 1) * 32;
    const uint32_t tail_mask = tail == 32 ? 0xffffffffu : ((1u << tail) - 1u);

    if (fill_hole) {
        // Bp = background reached from outside; padding bits count as outside
        for (int w = tid; w < nw; w += kMaskThreads) Bp[w] = (w % Wd == Wd - 1) ? ~tail_mask : 0u;
        for (;;) {
            __syncthreads();
            if (tid == 0) changed = 0;
            __syncthreads();
            bool any = false;
            for (int w = tid; w < nw; w += kMaskThreads) {
                const int y = w / Wd, c = w % Wd;
                const uint32_t bg = ~A[w] & (c == Wd - 1 ? tail_mask : 0xffffffffu);
                const uint32_t old = Bp[w];
                if ((bg & ~old) == 0) continue;
                // OR of the reached bits within two rows and two columns
                uint32_t acc = 0;
                for (int dy = -2; dy <= 2; ++dy) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= H) {
                        acc = 0xffffffffu;  // outside the image: border value 1
                        break;
                    }
                    const uint32_t *row = Bp + (int64_t)yy * Wd;
                    const uint32_t mid = row[c];
                    const uint32_t lft = c > 0 ? row[c - 1] : 0xffffffffu;  // column -1, -2: outside
                    const uint32_t rgt = c < Wd - 1 ? row[c + 1] : 0xffffffffu;
                    // pixel x + k for k = -2..2: shift the row window
                    const uint32_t r1 = (mid << 1) | (lft >> 31);   // x-1 reached
                    const uint32_t r2 = (mid << 2) | (lft >> 30);   // x-2
                    const uint32_t l1 = (mid >> 1) | (rgt << 31);   // x+1
                    const uint32_t l2 = (mid >> 2) | (rgt << 30);   // x+2
                    acc |= mid | r1 | r2 | l1 | l2;
                }
                const uint32_t nv = old | (bg & acc);
                if (nv != old) {
                    Bp[w] = nv;
                    any = true;
                }
            }
            if (any) changed = 1;
            __syncthreads();
            if (!changed) break;
        }
        // filled = not reached (within the image)
        for (int w = tid; w < nw; w += kMaskThreads) A[w] = ~Bp[w] & (w % Wd == Wd - 1 ? tail_mask : 0xffffffffu);
        __syncthreads();
    }
    // ---- 5. dilations with the cross, outside = 0 ------------------------------
    for (int it = 0; it < dilation; ++it) {
        for (int w = tid; w < nw; w += kMaskThreads) {
            const int y = w / Wd, c = w % Wd;
            const uint32_t mid = A[w];
            const uint32_t lft = c > 0 ? A[w - 1] : 0u, rgt = c < Wd - 1 ? A[w + 1] : 0u;
            const uint32_t up = y > 0 ? A[w - Wd] : 0u, dn = y < H - 1 ? A[w + Wd] : 0u;
            const uint32_t v = mid | up | dn | (mid << 1) | (lft >> 31) | (mid >> 1) | (rgt << 31);
            Bp[w] = v & (c == Wd - 1 ? tail_mask : 0xffffffffu);
        }
        __syncthreads();
        for (int w = tid; w < nw; w += kMaskThreads) A[w] = Bp[w];
        __syncthreads();
    }
    // ---- 6. erosions with the cross, outside = 0 -------------------------------
    for (int it = 0; it < erosion; ++it) {
        for (int w = tid; w < nw; w += kMaskThreads) {
            const int y = w / Wd, c = w % Wd;
            const uint32_t mid = A[w];
            const uint32_t lft = c > 0 ? A[w - 1] : 0u, rgt = c < Wd - 1 ? A[w + 1] : 0u;
            const uint32_t up = y > 0 ? A[w - Wd] : 0u, dn = y < H - 1 ? A[w + Wd] : 0u;
            const uint32_t v = mid & up & dn & ((mid << 1) | (lft >> 31)) & ((mid >> 1) | (rgt << 31));
            Bp[w] = v & (c == Wd - 1 ? tail_mask : 0xffffffffu);
        }
        __syncthreads();
        for (int w = tid; w < nw; w += kMaskThreads) A[w] = Bp[w];
        __syncthreads();
    }
    for (int64_t p = tid; p < npix; p += kMaskThreads) {
        const int y = (int)(p / W), x = (int)(p % W);
        mk[p] = (double)((A[y * Wd + x / 32] >> (x % 32)) & 1u);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void frac_ratio_kernel(const T *__restrict__ img, const double *__restrict__ img_sum,
                                                         double *__restrict__ tmp, int64_t n, int64_t npix,
                                                         int single)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n * npix) return;
    double v = (double)img[q] / img_sum[q % npix];
    if (single) v = (double)(float)v;
    // rm_abnormal: NaN, inf and negatives to 0
    tmp[q] = (v == v && v - v == 0.0 && v >= 0.0) ? v : 0.0;
}

__global__ __launch_bounds__(256) void frac_final_kernel(const double *__restrict__ tmp, const double *__restrict__ mask,
                                                         const double *__restrict__ ratio,
                                                         const double *__restrict__ lohi, double *__restrict__ frac,
                                                         int64_t n, int64_t npix, int mask_scalar, int ratio_scalar,
                                                         int single)
{
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= n * npix) return;
    const int64_t i = q / npix, p = q % npix;
    double v = tmp[q];
    const double lo = lohi[2 * i], hi = lohi[2 * i + 1];
    if (lo <= hi) {  // clipping applies (more than 100 positive values)
        if (v >= hi) v = hi;
        if (v <= lo) v = lo;
    }
    const double m = mask_scalar ? 1.0 : mask[p];
    const double r = ratio_scalar ? 1.0 : ratio[p];
    double out = v * m;
    if (single) out = (double)(float)out;
    out = out * r;
    if (single) out = (double)(float)out;
    frac[q] = out;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
size_t adaptive_mask_lds_bytes(int64_t H, int64_t W) { return (size_t)(2 * H * ((W + 31) / 32)) * sizeof(uint32_t); }

template <typename T>
static int launch_frac_sum_t(const void *img, void *out, int64_t n, int64_t npix, hipStream_t s)
{
    frac_sum_kernel<T><<<dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s>>>((const T *)img, (T *)out, n, npix);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_frac_sum(const void *img, void *out, int64_t n, int64_t npix, int f64, hipStream_t s)
{
    FL_REQUIRE(n >= 1 && npix >= 1 && (npix + 255) / 256 < (int64_t)UINT32_MAX, "frac sum: bad shape");
    return f64 ? launch_frac_sum_t<double>(img, out, n, npix, s) : launch_frac_sum_t<float>(img, out, n, npix, s);
}

int launch_adaptive_mask(const void *img, double *mask, int64_t n_img, int64_t H, int64_t W, int fill_hole,
                         int dilation, int erosion, int f64, int *status, void *thresholds, hipStream_t s)
{
    FL_REQUIRE(n_img >= 1 && H >= 1 && W >= 1 && n_img < (int64_t)UINT32_MAX, "adaptive mask: bad shape");
    FL_REQUIRE(dilation >= 0 && erosion >= 0, "adaptive mask: dilation and erosion must be >= 0");
    const size_t lds = adaptive_mask_lds_bytes(H, W);
    FL_REQUIRE(lds <= 144 * 1024, "adaptive mask: image too large for one workgroup's LDS (H * ceil(W/32) <= 18432)");
    if (f64) {
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)adaptive_mask_kernel<double>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        adaptive_mask_kernel<double><<<dim3((unsigned)n_img), dim3(kMaskThreads), lds, s>>>(
            (const double *)img, mask, (int)H, (int)W, fill_hole, dilation, erosion, status, (double *)thresholds);
    } else {
        FL_HIP_CHECK(hipFuncSetAttribute((const void *)adaptive_mask_kernel<float>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        adaptive_mask_kernel<float><<<dim3((unsigned)n_img), dim3(kMaskThreads), lds, s>>>(
            (const float *)img, mask, (int)H, (int)W, fill_hole, dilation, erosion, status, (float *)thresholds);
    }
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_frac_ratio(const void *img, const double *img_sum, double *tmp, int64_t n, int64_t npix, int f64,
                      int single, hipStream_t s)
{
    const int64_t total = n * npix;
    FL_REQUIRE(total >= 1 && (total + 255) / 256 < (int64_t)UINT32_MAX, "frac ratio: bad shape");
    if (f64)
        frac_ratio_kernel<double><<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
            (const double *)img, img_sum, tmp, n, npix, single);
    else
        frac_ratio_kernel<float><<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
            (const float *)img, img_sum, tmp, n, npix, single);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

int launch_frac_final(const double *tmp, const double *mask, const double *ratio, const double *lohi, double *frac,
                      int64_t n, int64_t npix, int single, hipStream_t s)
{
    const int64_t total = n * npix;
    FL_REQUIRE(total >= 1 && (total + 255) / 256 < (int64_t)UINT32_MAX, "frac final: bad shape");
    frac_final_kernel<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
        tmp, mask, ratio, lohi, frac, n, npix, mask == nullptr, ratio == nullptr, single);
    FL_LAUNCH_CHECK();
    return FL_OK;
}

}  // namespace fl
