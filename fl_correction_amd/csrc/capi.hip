// C ABI entry points (declared in include/fl_correction_amd.h).  Each checks
// its arguments on the host, then enqueues kernels on the caller's stream.
// No entry point allocates device memory: scratch comes from the caller's
// workspace argument.
#include "common.hpp"

#include <string>

namespace fl {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

int launch_forward(const float *atten, const float *C, const float *em, const float *theta,
                   float *part, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W,
                   hipStream_t stream, int group);
int launch_forward_reduce(const float *part, const float *I, float *out, int64_t A, int64_t B,
                          int64_t H, int64_t W, int mode, hipStream_t stream);
int launch_backward(const float *atten, const float *ratio, const float *em, const float *theta,
                    float *out, int64_t A, int64_t B, int64_t R, int64_t H, int64_t W, int mode,
                    const float *C, const float *sens, float beta, float delta, hipStream_t stream);
int launch_huber(const float *C, float *grad, int64_t n_slices, int64_t H, int64_t W, float delta,
                 hipStream_t stream);
size_t forward_part_floats(int64_t A, int64_t B, int64_t H, int64_t W);

static int check_dims(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W)
{
    FL_REQUIRE(A >= 0 && B >= 0 && R >= 1 && H >= 2 && W >= 2,
               "shape: need A >= 0, B >= 0, R >= 1, H >= 2, W >= 2");
    FL_REQUIRE(H < (1 << 20) && W < (1 << 20) && A < (1 << 24) && B < (1 << 24) && R < 65536,
               "shape: dimension too large");
    return FL_OK;
}

}  // namespace fl

using namespace fl;

extern "C" {

const char *fl_last_error(void) { return g_last_error.c_str(); }

int fl_version(void) { return (1 << 16) | 0; }

size_t fl_projector_workspace_bytes(int64_t A, int64_t B, int64_t H, int64_t W)
{
    return sizeof(float) * forward_part_floats(A, B, H, W);
}

int fl_forward_emission_batch(const float *atten, const float *C, const float *em_cs,
                              const float *theta, float *Pf, int64_t A, int64_t B, int64_t R,
                              int64_t H, int64_t W, void *workspace, size_t workspace_bytes,
                              void *stream)
{
    g_last_error.clear();
    if (int e = check_dims(A, B, R, H, W)) return e;
    if (A == 0 || B == 0) return FL_OK;
    FL_REQUIRE(atten && C && em_cs && theta && Pf, "null pointer argument");
    const size_t need = fl_projector_workspace_bytes(A, B, H, W);
    if (workspace_bytes < need || !workspace) {
        set_error("workspace too small: need " + std::to_string(need) + " bytes");
        return FL_ERR_WORKSPACE;
    }
    hipStream_t s = (hipStream_t)stream;
    float *part = (float *)workspace;
    if (int e = launch_forward(atten, C, em_cs, theta, part, A, B, R, H, W, s, 8)) return e;
    return launch_forward_reduce(part, nullptr, Pf, A, B, H, W, 0, s);
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
                        int64_t W, void *workspace, size_t workspace_bytes, void *stream)
{
    return fl_forward_emission_batch(atten, C, em_cs, theta, Pf, A, 1, R, H, W, workspace,
                                     workspace_bytes, stream);
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

size_t fl_mlem_workspace_bytes(int64_t A, int64_t B, int64_t R, int64_t H, int64_t W)
{
    const size_t nC = (size_t)(R * B * H * W);
    const size_t nP = (size_t)(A * B * W);
    // sens, ping-pong estimate, forward partials, ratio; each region 256-byte aligned
    auto al = [](size_t f) { return (f * sizeof(float) + 255) / 256 * 256; };
    return al(nC) + al(nC) + al(forward_part_floats(A, B, H, W)) + al(nP);
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
    const size_t nP = (size_t)(A * B * W);
    auto al = [](size_t f) { return (f * sizeof(float) + 255) / 256 * 256; };
    char *ws = (char *)workspace;
    float *sens = (float *)ws;
    float *other = (float *)(ws + al(nC));
    float *part = (float *)(ws + 2 * al(nC));
    float *ratio = (float *)(ws + 2 * al(nC) + al(forward_part_floats(A, B, H, W)));
    (void)nP;

    // sens = max(B^T 1, 1e-6)   (atten_cuda.cpp:308-314)
    if (int e = launch_backward(atten, nullptr, em_cs, theta, sens, A, B, R, H, W, 1, nullptr,
                                nullptr, 0.f, 0.f, s))
        return e;
    // C = C_init.clone()  (:303); iterate between C_out and `other`
    FL_HIP_CHECK(hipMemcpyAsync(C_out, C_init, nC * sizeof(float), hipMemcpyDeviceToDevice, s));
    float *cur = C_out, *nxt = other;
    for (int64_t it = 0; it < n_iter; ++it) {
        if (int e = launch_forward(atten, cur, em_cs, theta, part, A, B, R, H, W, s, 8)) return e;
        if (int e = launch_forward_reduce(part, I, ratio, A, B, H, W, 1, s)) return e;
        if (int e = launch_backward(atten, ratio, em_cs, theta, nxt, A, B, R, H, W, 2, cur, sens,
                                    (float)beta, (float)delta, s))
            return e;
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (cur != C_out)
        FL_HIP_CHECK(hipMemcpyAsync(C_out, cur, nC * sizeof(float), hipMemcpyDeviceToDevice, s));
    return FL_OK;
}

}  // extern "C"
