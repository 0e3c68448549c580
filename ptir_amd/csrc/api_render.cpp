// C ABI of the render-side image kernels (declared in include/ptir_amd.h).
#include "../../include/ptir_amd.h"

#include <hip/hip_runtime.h>

#include <string>

#include "internal.hpp"

namespace {

ptir_status render_fail(ptir_status st, const std::string& msg) {
    ptir::set_thread_error(msg);
    return st;
}

ptir_status hip_status(hipError_t e, const char* where) {
    if (e == hipSuccess) return PTIR_OK;
    return render_fail(PTIR_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

bool bad_shape(int32_t B, int32_t H, int32_t W) { return B < 0 || H < 0 || W < 0 || (int64_t)B * H * W > (1ll << 31); }

}  // namespace

extern "C" {

ptir_status ptir_camera_rays(const float* pixel_coords, const float* intrinsics, const float* jitter,
                             int32_t jitter_per_pixel, const float* ray_ori, int32_t S, int32_t B, int32_t H,
                             int32_t W, float* out_ori, float* out_dir, void* stream) {
    if (bad_shape(B, H, W) || S <= 0 || (int64_t)S * B * H * W > (1ll << 31))
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "camera rays: bad sizes");
    if ((int64_t)B * H * W == 0) return PTIR_OK;
    if (!pixel_coords || !intrinsics || !jitter || !ray_ori || !out_ori || !out_dir)
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "camera rays: null pointer");
    return hip_status(ptir::camera_rays(pixel_coords, intrinsics, jitter, jitter_per_pixel != 0, ray_ori, S, B, H, W,
                                        out_ori, out_dir, static_cast<hipStream_t>(stream)),
                      "camera rays");
}

ptir_status ptir_resolve_forward(const ptir_resolve* r, void* stream) {
    if (!r || bad_shape(r->B, r->H, r->W) || r->iterations < 0 || r->iterations > 16)
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "resolve: bad sizes or iteration count");
    if (!r->opacity || !r->dist || !r->pbr || !r->direct || !r->indirect || !r->light || !r->normals ||
        !r->out_dist || !r->out_normals || !r->image || !r->work)
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "resolve: null pointer");
    const ptir::ResolveIO io{r->B, r->H, r->W, r->iterations, r->opacity, r->dist, r->pbr, r->direct, r->indirect,
                             r->light, r->normals, r->out_dist, r->out_normals, r->image, r->work};
    return hip_status(ptir::resolve_forward(io, static_cast<hipStream_t>(stream)), "resolve forward");
}

ptir_status ptir_resolve_backward(const ptir_resolve* r, const float* grad_dist, const float* grad_normals,
                                  float* d_opacity, float* d_dist, float* d_pbr, float* d_direct, float* d_indirect,
                                  float* d_light, float* d_normals, void* stream) {
    if (!r || bad_shape(r->B, r->H, r->W) || r->iterations < 0 || r->iterations > 16)
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "resolve: bad sizes or iteration count");
    if (!r->opacity || !r->dist || !r->pbr || !r->direct || !r->indirect || !r->normals || !r->image || !r->work ||
        !grad_dist || !grad_normals || !d_opacity || !d_dist || !d_pbr || !d_direct || !d_indirect || !d_light ||
        !d_normals)
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "resolve backward: null pointer");
    const ptir::ResolveIO io{r->B, r->H, r->W, r->iterations, r->opacity, r->dist, r->pbr, r->direct, r->indirect,
                             r->light, r->normals, nullptr, nullptr, r->image, r->work};
    const ptir::ResolveGrads g{grad_dist, grad_normals, d_opacity, d_dist, d_pbr, d_direct, d_indirect, d_light,
                               d_normals};
    return hip_status(ptir::resolve_backward(io, g, static_cast<hipStream_t>(stream)), "resolve backward");
}

ptir_status ptir_adam_step(const ptir_adam_group* groups, int32_t num_groups, float beta1, float beta2, float eps,
                           int32_t bias_correction, int32_t* step, void* stream) {
    if (num_groups < 0 || num_groups > 64 || (num_groups && !groups) || !step)
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "adam: bad group list or null step counter");
    if (!(beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps >= 0.f))
        return render_fail(PTIR_ERR_INVALID_ARGUMENT, "adam: betas must be in [0,1) and eps >= 0");
    ptir::AdamGroup g[64];
    for (int k = 0; k < num_groups; ++k) {
        const ptir_adam_group& a = groups[k];
        if (a.rows && a.cols && (!a.param || !a.grad || !a.exp_avg || !a.exp_avg_sq))
            return render_fail(PTIR_ERR_INVALID_ARGUMENT, "adam: null tensor in group " + std::to_string(k));
        g[k] = ptir::AdamGroup{a.param, a.grad, a.exp_avg, a.exp_avg_sq, a.visibility, a.rows, a.cols, a.lr};
    }
    return hip_status(ptir::adam_step(g, num_groups, beta1, beta2, eps, bias_correction != 0, step,
                                      static_cast<hipStream_t>(stream)),
                      "adam step");
}

}  // extern "C"
