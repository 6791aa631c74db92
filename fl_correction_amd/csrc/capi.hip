// C ABI entry points (declared in include/fl_correction_amd.h).  Each checks
// its arguments on the host, then enqueues kernels on the caller's stream.
// No entry point allocates device memory: scratch comes from the caller's
// workspace argument, or for fl_mlem_batch_noalloc from a static device array.
#include "projector.hpp"

#include <mutex>
#include <string>

namespace fl {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

int launch_forward(const float *atten, const float *C, const float *em, const float *theta, const float *I,
                   float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode, hipStream_t stream,
                   int out_sm = 0);
int launch_backward(const float *atten, const float *ratio, const float *em, const float *theta,
                    float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode,
                    const float *C, const float *sens, float beta, float delta, hipStream_t stream);
int launch_huber(const float *C, float *grad, int64_t n_slices, int64_t H, int64_t W, float delta,
                 hipStream_t stream);
int launch_mask_corr(const float *mu, const float *mask, float *out, int64_t K, int64_t Z, int64_t X, int64_t Y,
                     int64_t Mz, int64_t Mx, int64_t My, int64_t hz, int64_t hx, int64_t hy, int flip, hipStream_t s);
int launch_sample_grad(const float *C, const float *em, const float *theta, const float *G, float *out, int64_t A,
                       int64_t B, int64_t R, int64_t H, int64_t W, int64_t out_astride, hipStream_t s);
int launch_xrf_grad_prep(const float *T_rot, const float *Ax, const float *gA, float *dT_rot, float *gx,
                         float *A_tot, int64_t A_tot_estride, int64_t K, int64_t B, int64_t H, int64_t W, int side,
                         hipStream_t s);
int launch_incident_grad(const float *F, const float *coef_x, const float *Ax, const float *gx, float *dmux,
                         int64_t K, int64_t B, int64_t H, int64_t W, hipStream_t s);
int launch_mu_adjoint(const float *dmu_rot, const float *dmux, const float *coef, const float *coef_x, float *dF,
                      int64_t Kc, int64_t K, int64_t B, int64_t H, int64_t W, int side, hipStream_t s);
int launch_rotate_adjoint(const float *dF, float *dC, int64_t N, int64_t H, int64_t W, float c, float s_,
                          int accumulate, hipStream_t s);
int launch_rotate(const float *x, float *out, int64_t N, int64_t H, int64_t W, float c, float s_, hipStream_t s);
int launch_incident(const float *frac, const float *coef, float *out, int64_t frac_estride, int64_t out_estride,
                    int64_t K, int64_t B, int64_t H, int64_t W, hipStream_t s);
int launch_mu_rot(const float *frac, const float *coef, float *mu_rot, int64_t Kc, int64_t K, int64_t B, int64_t H,
                  int64_t W, int side, hipStream_t s);
int launch_combine(const float *T_rot, const float *incident, float *out, int64_t inc_estride, int64_t out_estride,
                   int64_t K, int64_t B, int64_t H, int64_t W, int side, hipStream_t s);
int launch_cone_direction(const float *mu, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y, int pz, int py,
                          int q, float weight, int accumulate, int reverse, hipStream_t s);
size_t cone_sweep_workspace_bytes(int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz, const int32_t *py,
                                  const int32_t *q, int64_t ndir);
int launch_cone_sweep(const float *mu, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz,
                      const int32_t *py, const int32_t *q, const float *weight, int64_t ndir, int reverse,
                      void *workspace, size_t workspace_bytes, hipStream_t s);
int launch_frac_sum(const void *img, void *out, int64_t n, int64_t npix, int f64, hipStream_t s);
int launch_adaptive_mask(const void *img, double *mask, int64_t n_img, int64_t H, int64_t W, int fill_hole,
                         int dilation, int erosion, int f64, int *status, void *thresholds, hipStream_t s);
int launch_frac_ratio(const void *img, const double *img_sum, double *tmp, int64_t n, int64_t npix, int f64,
                      int single, hipStream_t s);
int launch_frac_final(const double *tmp, const double *mask, const double *ratio, const double *lohi, double *frac,
                      int64_t n, int64_t npix, int single, hipStream_t s);

static int check_dims(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W)
{
    FL_REQUIRE(A >= 0 && B >= 0 && R >= 1 && H >= 2 && W >= 2,
               "shape: need A >= 0, B >= 0, R >= 1, H >= 2, W >= 2");
    FL_REQUIRE(H < (1 << 20) && W < (1 << 20) && A < (1 << 24) && B < (1 << 24) && R < 65536,
               "shape: dimension too large");
    // one slice is addressed with 32-bit byte offsets
    FL_REQUIRE(H * W < (int64_t)1 << 29, "shape: H*W must be below 2^29");
    return FL_OK;
}

// ---------------------------------------------------------------------------
// fl_mlem_batch without a workspace: its scratch is a file-scope device array.
// Every GPU the library runs on holds its own copy of the array (a module
// global) until the process ends.  In host code `scratch` is only a handle:
// its device address on the current device comes from hipGetSymbolAddress,
// looked up once per device and kept in g_scratch_addr.
// ---------------------------------------------------------------------------

static __device__ __align__(256) unsigned char scratch[FL_MLEM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_addr[kMaxDevices] = {};

// Device address of this device's `scratch` (the current device must be `dev`).
static int scratch_address(int dev, void **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(err));
            return FL_ERR_HIP;
        }
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return FL_OK;
}

}  // namespace fl

using namespace fl;

// Slices at which fl_mlem_batch takes the slice-major path (64 slices per
// wavefront there: with fewer, most lanes would idle).
constexpr int64_t kSmMinSlices = 32;

extern "C" {

const char *fl_last_error(void) { return g_last_error.c_str(); }

int fl_version(void) { return (1 << 16) | 0; }

int fl_forward_emission_batch(const float *atten, const float *C, const float *em_cs,
                              const float *theta, float *Pf, int64_t A, int64_t B, int64_t R,
                              int64_t H, int64_t W, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    if (A == 0 || B == 0) return FL_OK;
    FL_REQUIRE(atten && C && em_cs && theta && Pf, "null pointer argument");
    return launch_forward(atten, C, em_cs, theta, nullptr, Pf, A, B, R, H, W, 0, (hipStream_t)stream);
}

int fl_backward_emission_batch(const float *atten, const float *ratio, const float *em_cs,
                               const float *theta, float *out, int64_t A, int64_t B, int64_t R,
                               int64_t H, int64_t W, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    if (B == 0) return FL_OK;
    FL_REQUIRE(out && (A == 0 || (atten && em_cs && theta)), "null pointer argument");
    hipStream_t s = (hipStream_t)stream;
    if (A == 0) {
        FL_HIP_CHECK(hipMemsetAsync(out, 0, sizeof(float) * (size_t)(R * B * H * W), s));
        return FL_OK;
    }
    return launch_backward(atten, ratio, em_cs, theta, out, A, B, R, H, W, 0, nullptr, nullptr, 0.f,
                           0.f, s);
}

int fl_forward_emission(const float *atten, const float *C, const float *em_cs,
                        const float *theta, float *Pf, int64_t A, int64_t R, int64_t H,
                        int64_t W, void *stream)
{
    return fl_forward_emission_batch(atten, C, em_cs, theta, Pf, A, 1, R, H, W, stream);
}

int fl_backward_emission(const float *atten, const float *ratio, const float *em_cs,
                         const float *theta, float *out, int64_t A, int64_t R, int64_t H,
                         int64_t W, void *stream)
{
    return fl_backward_emission_batch(atten, ratio, em_cs, theta, out, A, 1, R, H, W, stream);
}

int fl_huber_grad_batch(const float *C, float *grad, int64_t R, int64_t B, int64_t H, int64_t W,
                        float delta, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(0, B, R, H, W)) return e;
    FL_REQUIRE(C && grad, "null pointer argument");
    return launch_huber(C, grad, R * B, H, W, delta, (hipStream_t)stream);
}

int fl_mlem_sensitivity(const float *atten, const float *em_cs, const float *theta, float *sens,
                        int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    FL_REQUIRE(A >= 1, "sensitivity needs at least one angle");
    if (B == 0) return FL_OK;
    FL_REQUIRE(atten && em_cs && theta && sens, "null pointer argument");
    // sens = max(B^T 1, 1e-6)   (atten_cuda.cpp:308-314)
    return launch_backward(atten, nullptr, em_cs, theta, sens, A, B, R, H, W, 1, nullptr, nullptr, 0.f,
                           0.f, (hipStream_t)stream);
}

int fl_mlem_ratio(const float *atten, const float *C, const float *em_cs, const float *theta,
                  const float *I, float *ratio, int64_t A, int64_t B, int64_t R, int64_t H,
                  int64_t W, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    if (A == 0 || B == 0) return FL_OK;
    FL_REQUIRE(atten && C && em_cs && theta && I && ratio, "null pointer argument");
    // Pf = forward(C); ratio = I / max(Pf, 1e-6)   (atten_cuda.cpp:323-328), one kernel
    return launch_forward(atten, C, em_cs, theta, I, ratio, A, B, R, H, W, 2, (hipStream_t)stream);
}

int fl_mlem_update(const float *atten, const float *ratio, const float *em_cs, const float *theta,
                   const float *C, const float *sens, float *C_next, int64_t A, int64_t B,
                   int64_t R, int64_t H, int64_t W, double beta, double delta, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    FL_REQUIRE(A >= 1, "update needs at least one angle");
    if (B == 0) return FL_OK;
    FL_REQUIRE(atten && ratio && em_cs && theta && C && sens && C_next, "null pointer argument");
    FL_REQUIRE(C_next != C, "C_next must not alias C (neighbours of C are read by other workgroups)");
    // back = B^T ratio; C_next = max(0, C*back/(sens + beta*huber(C)))   (atten_cuda.cpp:330-341)
    return launch_backward(atten, ratio, em_cs, theta, C_next, A, B, R, H, W, 2, C, sens, (float)beta,
                           (float)delta, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Slice-major path: the back-projector on atten_sm (A, H, W, B)
// ---------------------------------------------------------------------------

int fl_slice_major(const float *src, float *dst, int64_t N, int64_t B, int64_t L, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(N >= 0 && B >= 0 && L >= 0, "shape: negative size");
    if (N == 0 || B == 0 || L == 0) return FL_OK;
    FL_REQUIRE(src && dst && src != dst, "null or aliased pointer argument");
    return launch_sm_transpose(src, dst, N, B, L, (hipStream_t)stream);
}

size_t fl_backproject_sm_scratch_bytes(int64_t A, int64_t B, int64_t H, int64_t W)
{
    if (A < 1 || B < 1 || H < 2 || W < 2) return 0;
    return sm_scratch_bytes(A, B, H, W);
}

static int check_sm(int64_t A, int64_t B, int64_t H, int64_t W)
{
    if (int e = check_dims(A, B, 1, H, W)) return e;
    FL_REQUIRE(A >= 1, "at least one angle");
    FL_REQUIRE(H * W * B < ((int64_t)1 << 29), "slice-major path: H*W*B must be below 2^29");
    return FL_OK;
}

int fl_mlem_ratio_sm(const float *atten, const float *C, const float *em_cs, const float *theta, const float *I,
                     float *ratio_sm, int64_t A, int64_t B, int64_t H, int64_t W, void *stream)
{
    g_last_error.clear();
    if (int e = check_sm(A, B, H, W)) return e;
    if (B == 0) return FL_OK;
    FL_REQUIRE(atten && C && em_cs && theta && I && ratio_sm, "null pointer argument");
    return launch_forward(atten, C, em_cs, theta, I, ratio_sm, A, B, 1, H, W, 2, (hipStream_t)stream, 1);
}

int fl_backward_emission_sm(const float *atten_sm, const float *ratio_sm, const float *em_cs, const float *theta,
                            float *out, int64_t A, int64_t B, int64_t H, int64_t W, int accumulate, void *scratch,
                            size_t scratch_bytes, void *stream)
{
    g_last_error.clear();
    if (int e = check_sm(A, B, H, W)) return e;
    if (B == 0) return FL_OK;
    FL_REQUIRE(atten_sm && em_cs && theta && out, "null pointer argument");
    return launch_sm_backward(atten_sm, ratio_sm, em_cs, theta, out, A, B, H, W,
                              accumulate ? kBwdAccumulate : kBwdWrite, nullptr, nullptr, 0.f, 0.f, scratch,
                              scratch_bytes, (hipStream_t)stream);
}

int fl_mlem_sensitivity_sm(const float *atten_sm, const float *em_cs, const float *theta, float *sens, int64_t A,
                           int64_t B, int64_t H, int64_t W, void *scratch, size_t scratch_bytes, void *stream)
{
    g_last_error.clear();
    if (int e = check_sm(A, B, H, W)) return e;
    if (B == 0) return FL_OK;
    FL_REQUIRE(atten_sm && em_cs && theta && sens, "null pointer argument");
    return launch_sm_backward(atten_sm, nullptr, em_cs, theta, sens, A, B, H, W, kBwdSens, nullptr, nullptr, 0.f,
                              0.f, scratch, scratch_bytes, (hipStream_t)stream);
}

int fl_mlem_update_sm(const float *atten_sm, const float *ratio_sm, const float *em_cs, const float *theta,
                      const float *C, const float *sens, float *C_next, int64_t A, int64_t B, int64_t H, int64_t W,
                      double beta, double delta, void *scratch, size_t scratch_bytes, void *stream)
{
    g_last_error.clear();
    if (int e = check_sm(A, B, H, W)) return e;
    if (B == 0) return FL_OK;
    FL_REQUIRE(atten_sm && ratio_sm && em_cs && theta && C && sens && C_next, "null pointer argument");
    FL_REQUIRE(C_next != C, "C_next must not alias C (neighbours of C are read by other workgroups)");
    return launch_sm_backward(atten_sm, ratio_sm, em_cs, theta, C_next, A, B, H, W, kBwdUpdate, C, sens,
                              (float)beta, (float)delta, scratch, scratch_bytes, (hipStream_t)stream);
}

size_t fl_mlem_workspace_bytes_sm(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W)
{
    if (R != 1 || A < 1 || B < 1 || H < 2 || W < 2 || H * W * B >= ((int64_t)1 << 29)) return 0;
    auto al = [](size_t f) { return (f * sizeof(float) + 255) / 256 * 256; };
    const size_t nC = (size_t)(B * H * W), nP = (size_t)(A * B * W);
    // sens, ping-pong estimate, ratio (A, W, B), atten_sm (A, H, W, B), back-projector scratch
    return al(nC) + al(nC) + al(nP) + al((size_t)A * nC) + sm_scratch_bytes(A, B, H, W);
}

size_t fl_mlem_workspace_bytes(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W)
{
    const size_t nC = (size_t)(R * B * H * W);
    const size_t nP = (size_t)(A * B * W);
    // sens, ping-pong estimate, ratio; each region 256-byte aligned
    auto al = [](size_t f) { return (f * sizeof(float) + 255) / 256 * 256; };
    return al(nC) + al(nC) + al(nP);
}

int fl_mlem_batch(const float *C_init, const float *atten, const float *em_cs,
                  const float *theta, const float *I, float *C_out, int64_t A, int64_t B,
                  int64_t R, int64_t H, int64_t W, int64_t n_iter, double beta, double delta,
                  void *workspace, size_t workspace_bytes, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    FL_REQUIRE(n_iter >= 0, "n_iter must be >= 0");
    FL_REQUIRE(A >= 1, "mlem needs at least one angle");
    if (B == 0) return FL_OK;
    FL_REQUIRE(C_init && atten && em_cs && theta && I && C_out, "null pointer argument");
    const size_t need = fl_mlem_workspace_bytes(A, B, R, H, W);
    if (workspace_bytes < need || !workspace) {
        set_error("workspace too small: need " + std::to_string(need) + " bytes");
        return FL_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    const size_t nC = (size_t)(R * B * H * W);
    auto al = [](size_t f) { return (f * sizeof(float) + 255) / 256 * 256; };
    char *ws = (char *)workspace;
    float *sens = (float *)ws;
    float *other = (float *)(ws + al(nC));
    float *ratio = (float *)(ws + 2 * al(nC));
    const size_t need_sm = fl_mlem_workspace_bytes_sm(A, B, R, H, W);
    if (need_sm && workspace_bytes >= need_sm && B >= kSmMinSlices) {
        // slice-major path: a slice-major copy of atten once, then every
        // back-projection on it (the forward keeps the reference's layout)
        const size_t nP = (size_t)(A * B * W);
        float *atten_sm = (float *)(ws + 2 * al(nC) + al(nP));
        void *scr = ws + 2 * al(nC) + al(nP) + al((size_t)A * nC);
        const size_t scr_bytes = sm_scratch_bytes(A, B, H, W);
        if (int e = fl_slice_major(atten, atten_sm, A, B, H * W, stream)) return e;
        if (int e = fl_mlem_sensitivity_sm(atten_sm, em_cs, theta, sens, A, B, H, W, scr, scr_bytes, stream)) return e;
        FL_HIP_CHECK(hipMemcpyAsync(C_out, C_init, nC * sizeof(float), hipMemcpyDeviceToDevice, s));
        float *cur = C_out, *nxt = other;
        for (int64_t it = 0; it < n_iter; ++it) {
            if (int e = fl_mlem_ratio_sm(atten, cur, em_cs, theta, I, ratio, A, B, H, W, stream)) return e;
            if (int e = fl_mlem_update_sm(atten_sm, ratio, em_cs, theta, cur, sens, nxt, A, B, H, W, beta, delta,
                                          scr, scr_bytes, stream))
                return e;
            float *tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
        if (cur != C_out)
            FL_HIP_CHECK(hipMemcpyAsync(C_out, cur, nC * sizeof(float), hipMemcpyDeviceToDevice, s));
        return FL_OK;
    }

    if (int e = fl_mlem_sensitivity(atten, em_cs, theta, sens, A, B, R, H, W, stream)) return e;
    // C = C_init.clone()  (atten_cuda.cpp:303); iterate between C_out and `other`
    FL_HIP_CHECK(hipMemcpyAsync(C_out, C_init, nC * sizeof(float), hipMemcpyDeviceToDevice, s));
    float *cur = C_out, *nxt = other;
    for (int64_t it = 0; it < n_iter; ++it) {
        if (int e = fl_mlem_ratio(atten, cur, em_cs, theta, I, ratio, A, B, R, H, W, stream)) return e;
        if (int e = fl_mlem_update(atten, ratio, em_cs, theta, cur, sens, nxt, A, B, R, H, W, beta,
                                   delta, stream))
            return e;
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (cur != C_out)
        FL_HIP_CHECK(hipMemcpyAsync(C_out, cur, nC * sizeof(float), hipMemcpyDeviceToDevice, s));
    return FL_OK;
}

int fl_mlem_batch_noalloc(const float *C_init, const float *atten, const float *em_cs,
                          const float *theta, const float *I, float *C_out, int64_t A, int64_t B,
                          int64_t R, int64_t H, int64_t W, int64_t n_iter, double beta, double delta,
                          void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    const size_t need = fl_mlem_workspace_bytes(A, B, R, H, W);
    if (need > (size_t)FL_MLEM_NOALLOC_SCRATCH_BYTES) {
        set_error("fl_mlem_batch_noalloc: the problem needs " + std::to_string(need) +
                  " bytes of scratch, the static scratch holds " +
                  std::to_string((size_t)FL_MLEM_NOALLOC_SCRATCH_BYTES) +
                  "; call fl_mlem_batch with a workspace instead");
        return FL_ERR_WORKSPACE;
    }
    int dev = -1;
    FL_HIP_CHECK(hipGetDevice(&dev));
    FL_REQUIRE(dev >= 0 && dev < kMaxDevices, "fl_mlem_batch_noalloc: device ordinal out of range");
    if (stream) {
        int sdev = -1;
        FL_HIP_CHECK(hipStreamGetDevice((hipStream_t)stream, &sdev));
        FL_REQUIRE(sdev == dev, "fl_mlem_batch_noalloc: the stream belongs to another device than the current one");
    }
    void *ws = nullptr;
    if (int e = scratch_address(dev, &ws)) return e;
    // every region of the scratch is written before it is read (sens by the sensitivity
    // pass, the ratio by each forward pass, the second estimate by each update), so the
    // previous call's contents need no clearing
    // exactly the workspace fl_mlem_workspace_bytes asks for: the same kernels and
    // results as fl_mlem_batch called with that workspace
    return fl_mlem_batch(C_init, atten, em_cs, theta, I, C_out, A, B, R, H, W, n_iter, beta, delta, ws, need, stream);
}

// ---------------------------------------------------------------------------
// Attenuation maps
// ---------------------------------------------------------------------------

int fl_cal_atten(const float *mu, const float *mask, float *out, int64_t K, int64_t Z, int64_t X, int64_t Y,
                 int64_t Mz, int64_t Mx, int64_t My, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 0 && Z >= 0 && X >= 0 && Y >= 0 && Mz >= 1 && Mx >= 1 && My >= 1, "shape: negative size");
    FL_REQUIRE(X < (1 << 24) && Y < (1 << 24) && Z < (1 << 24), "shape: dimension too large");
    if (K == 0 || Z == 0 || X == 0 || Y == 0) return FL_OK;
    FL_REQUIRE(mu && mask && out, "null pointer argument");
    FL_REQUIRE(mu != out, "out must not alias mu");
    return launch_mask_corr(mu, mask, out, K, Z, X, Y, Mz, Mx, My, Mz / 2, 0, My / 2, 0, (hipStream_t)stream);
}

int fl_cal_atten_adjoint(const float *g, const float *mask, float *out, int64_t K, int64_t Z, int64_t X, int64_t Y,
                         int64_t Mz, int64_t Mx, int64_t My, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 0 && Z >= 0 && X >= 0 && Y >= 0 && Mz >= 1 && Mx >= 1 && My >= 1, "shape: negative size");
    FL_REQUIRE(X < (1 << 24) && Y < (1 << 24) && Z < (1 << 24), "shape: dimension too large");
    if (K == 0 || Z == 0 || X == 0 || Y == 0) return FL_OK;
    FL_REQUIRE(g && mask && out && g != out, "null or aliased pointer argument");
    // transpose of the correlation: flipped mask, centres Mz-1-Mz/2, Mx-1, My-1-My/2
    return launch_mask_corr(g, mask, out, K, Z, X, Y, Mz, Mx, My, Mz - 1 - Mz / 2, Mx - 1, My - 1 - My / 2, 1,
                            (hipStream_t)stream);
}

int fl_rotate(const float *x, float *out, int64_t N, int64_t H, int64_t W, float cos_t, float sin_t, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(N >= 0 && H >= 1 && W >= 1 && H < (1 << 20) && W < (1 << 20), "shape: bad size");
    FL_REQUIRE(N < (1 << 24), "shape: too many planes");
    if (N == 0) return FL_OK;
    FL_REQUIRE(x && out && x != out, "null or aliased pointer argument");
    return launch_rotate(x, out, N, H, W, cos_t, sin_t, (hipStream_t)stream);
}

int fl_incident_atten(const float *frac, const float *coef, float *out, int64_t K, int64_t B, int64_t H, int64_t W,
                      int64_t frac_element_stride, int64_t out_element_stride, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 0 && B >= 0 && H >= 1 && W >= 1 && K < 65536 && B < 65536, "shape: bad size");
    if (K == 0 || B == 0) return FL_OK;
    FL_REQUIRE(frac && out, "null pointer argument");
    return launch_incident(frac, coef, out, frac_element_stride, out_element_stride, K, B, H, W,
                           (hipStream_t)stream);
}

int fl_mu_xrf_rot(const float *frac, const float *coef, float *mu_rot, int64_t Kc, int64_t K, int64_t B, int64_t H,
                  int64_t W, int side, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(Kc >= 1 && K >= 1 && B >= 0 && H >= 1 && W >= 1, "shape: bad size");
    FL_REQUIRE(side == 0 || side == 1, "side must be 0 (detector at +col) or 1 (at -col)");
    FL_REQUIRE(B * ((K + 3) / 4) < 65536 && H < (1 << 20), "shape: too large");
    if (B == 0) return FL_OK;
    FL_REQUIRE(frac && coef && mu_rot, "null pointer argument");
    return launch_mu_rot(frac, coef, mu_rot, Kc, K, B, H, W, side, (hipStream_t)stream);
}

int fl_cone_atten(const float *mu_rot, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz,
                  const int32_t *py, const int32_t *q, const float *weight, int64_t ndir, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 1 && Z >= 1 && X >= 1 && Y >= 1 && ndir >= 1, "shape: bad size");
    FL_REQUIRE(Z < (1 << 20) && X < (1 << 20) && Y < (1 << 20) && K < 2048, "shape: too large");
    FL_REQUIRE(mu_rot && T && pz && py && q && weight, "null pointer argument");
    FL_REQUIRE(mu_rot != T, "T must not alias mu");
    for (int64_t d = 0; d < ndir; ++d)
        if (int e = launch_cone_direction(mu_rot, T, K, Z, X, Y, pz[d], py[d], q[d], weight[d], d > 0, 0,
                                          (hipStream_t)stream))
            return e;
    return FL_OK;
}

int fl_cone_atten_adjoint(const float *g_rot, float *dmu_rot, int64_t K, int64_t Z, int64_t X, int64_t Y,
                          const int32_t *pz, const int32_t *py, const int32_t *q, const float *weight, int64_t ndir,
                          void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 1 && Z >= 1 && X >= 1 && Y >= 1 && ndir >= 1, "shape: bad size");
    FL_REQUIRE(Z < (1 << 20) && X < (1 << 20) && Y < (1 << 20) && K < 2048, "shape: too large");
    FL_REQUIRE(g_rot && dmu_rot && pz && py && q && weight, "null pointer argument");
    FL_REQUIRE(g_rot != dmu_rot, "output must not alias input");
    for (int64_t d = 0; d < ndir; ++d)
        if (int e = launch_cone_direction(g_rot, dmu_rot, K, Z, X, Y, pz[d], py[d], q[d], weight[d], d > 0, 1,
                                          (hipStream_t)stream))
            return e;
    return FL_OK;
}

// ---------------------------------------------------------------------------
// Adjoint of the attenuated forward model
// ---------------------------------------------------------------------------

int fl_backward_emission_accumulate(const float *atten, const float *ratio, const float *em_cs,
                                    const float *theta, float *out, int64_t A, int64_t B, int64_t R, int64_t H,
                                    int64_t W, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    if (A == 0 || B == 0) return FL_OK;
    FL_REQUIRE(atten && ratio && em_cs && theta && out, "null pointer argument");
    return launch_backward(atten, ratio, em_cs, theta, out, A, B, R, H, W, 3, nullptr, nullptr, 0.f, 0.f,
                           (hipStream_t)stream);
}

int fl_sample_grad(const float *C, const float *em_cs, const float *theta, const float *G, float *out, int64_t A,
                   int64_t B, int64_t R, int64_t H, int64_t W, int64_t out_angle_stride, void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    if (A == 0 || B == 0) return FL_OK;
    FL_REQUIRE(C && em_cs && theta && G && out, "null pointer argument");
    return launch_sample_grad(C, em_cs, theta, G, out, A, B, R, H, W, out_angle_stride, (hipStream_t)stream);
}

int fl_xrf_grad_prep(const float *T_rot, const float *Ax, const float *gA, float *dT_rot, float *gx, float *A_tot,
                     int64_t A_tot_element_stride, int64_t K, int64_t B, int64_t H, int64_t W, int side,
                     void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 1 && B >= 0 && H >= 1 && W >= 1 && (side == 0 || side == 1), "bad arguments");
    if (B == 0) return FL_OK;
    FL_REQUIRE(T_rot && Ax && gA && dT_rot && gx, "null pointer argument");
    return launch_xrf_grad_prep(T_rot, Ax, gA, dT_rot, gx, A_tot, A_tot_element_stride, K, B, H, W, side,
                                (hipStream_t)stream);
}

int fl_incident_grad(const float *F, const float *coef_x, const float *Ax, const float *gx, float *dmux, int64_t K,
                     int64_t B, int64_t H, int64_t W, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 1 && B >= 0 && H >= 1 && W >= 1 && K < 65536 && B < 65536, "bad arguments");
    if (B == 0) return FL_OK;
    FL_REQUIRE(F && coef_x && Ax && gx && dmux, "null pointer argument");
    return launch_incident_grad(F, coef_x, Ax, gx, dmux, K, B, H, W, (hipStream_t)stream);
}

int fl_mu_adjoint(const float *dmu_rot, const float *dmux, const float *coef, const float *coef_x, float *dF,
                  int64_t Kc, int64_t K, int64_t B, int64_t H, int64_t W, int side, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(Kc >= 1 && K >= 1 && B >= 0 && B < 65536 && H >= 1 && W >= 1 && (side == 0 || side == 1),
               "bad arguments");
    if (B == 0) return FL_OK;
    FL_REQUIRE(dmu_rot && coef && dF && (!dmux || coef_x), "null pointer argument");
    return launch_mu_adjoint(dmu_rot, dmux, coef, coef_x, dF, Kc, K, B, H, W, side, (hipStream_t)stream);
}

int fl_rotate_adjoint(const float *dF, float *dC, int64_t N, int64_t H, int64_t W, float cos_t, float sin_t,
                      int accumulate, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(N >= 0 && H >= 1 && W >= 1 && H < (1 << 20) && W < (1 << 20), "shape: bad size");
    if (N == 0) return FL_OK;
    FL_REQUIRE(dF && dC && dF != dC, "null or aliased pointer argument");
    return launch_rotate_adjoint(dF, dC, N, H, W, cos_t, sin_t, accumulate, (hipStream_t)stream);
}

int fl_xrf_combine(const float *T_rot, const float *incident, float *out, int64_t K, int64_t B, int64_t H,
                   int64_t W, int64_t inc_element_stride, int64_t out_element_stride, int side, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 1 && B >= 0 && H >= 1 && W >= 1 && B * K < 65536, "shape: bad size");
    FL_REQUIRE(side == 0 || side == 1, "side must be 0 or 1");
    if (B == 0) return FL_OK;
    FL_REQUIRE(T_rot && out, "null pointer argument");
    return launch_combine(T_rot, incident, out, inc_element_stride, out_element_stride, K, B, H, W, side,
                          (hipStream_t)stream);
}

int fl_frac_sum(const void *img, void *out, int64_t n, int64_t npix, int f64, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(n >= 1 && npix >= 0, "shape: need n >= 1, npix >= 0");
    if (npix == 0) return FL_OK;
    FL_REQUIRE(img && out, "null pointer argument");
    return launch_frac_sum(img, out, n, npix, f64 != 0, (hipStream_t)stream);
}

int fl_adaptive_mask(const void *img, double *mask, int64_t n_img, int64_t H, int64_t W, int fill_hole,
                     int dilation, int erosion, int f64, int *status, void *thresholds, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(n_img >= 0 && H >= 1 && W >= 1 && H < (1 << 20) && W < (1 << 20), "shape: bad size");
    FL_REQUIRE(dilation >= 0 && erosion >= 0, "dilation and erosion must be >= 0");
    if (n_img == 0) return FL_OK;
    FL_REQUIRE(img && mask && status && thresholds, "null pointer argument");
    return launch_adaptive_mask(img, mask, n_img, H, W, fill_hole != 0, dilation, erosion, f64 != 0, status,
                                thresholds, (hipStream_t)stream);
}

int fl_frac_ratio(const void *img, const double *img_sum, double *tmp, int64_t n, int64_t npix, int f64,
                  int single, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(n >= 0 && npix >= 0, "shape: bad size");
    if (n == 0 || npix == 0) return FL_OK;
    FL_REQUIRE(img && img_sum && tmp, "null pointer argument");
    return launch_frac_ratio(img, img_sum, tmp, n, npix, f64 != 0, single != 0, (hipStream_t)stream);
}

int fl_frac_final(const double *tmp, const double *mask, const double *ratio, const double *lohi, double *frac,
                  int64_t n, int64_t npix, int single, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(n >= 0 && npix >= 0, "shape: bad size");
    if (n == 0 || npix == 0) return FL_OK;
    FL_REQUIRE(tmp && lohi && frac, "null pointer argument");
    return launch_frac_final(tmp, mask, ratio, lohi, frac, n, npix, single != 0, (hipStream_t)stream);
}

size_t fl_cone_sweep_workspace_bytes(int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz,
                                     const int32_t *py, const int32_t *q, int64_t ndir)
{
    if (K < 1 || Z < 1 || X < 1 || Y < 1 || ndir < 1 || !pz || !py || !q) return 0;
    return cone_sweep_workspace_bytes(K, Z, X, Y, pz, py, q, ndir);
}

int fl_cone_sweep(const float *mu_rot, float *T, int64_t K, int64_t Z, int64_t X, int64_t Y, const int32_t *pz,
                  const int32_t *py, const int32_t *q, const float *weight, int64_t ndir, int adjoint,
                  void *workspace, size_t workspace_bytes, void *stream)
{
    g_last_error.clear();
    FL_REQUIRE(K >= 1 && Z >= 1 && X >= 1 && Y >= 1 && ndir >= 1, "shape: bad size");
    FL_REQUIRE(Z < (1 << 20) && X < (1 << 20) && Y < (1 << 20) && K < 2048, "shape: too large");
    FL_REQUIRE(mu_rot && T && pz && py && q && weight, "null pointer argument");
    FL_REQUIRE(mu_rot != T, "T must not alias mu");
    return launch_cone_sweep(mu_rot, T, K, Z, X, Y, pz, py, q, weight, ndir, adjoint != 0, workspace,
                             workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"
