// Adam update of the optimised tensors, on the device (the step after the backward).
//
// The reference's PTIR inversion optimises per-Gaussian albedo, roughness, metallic and the
// environment map with torch.optim.Adam, or with SelectiveAdam, which updates only the rows
// of visible Gaussians and applies no bias correction
// (threedgrut/optimizers/optimizers.cu:49-80, threedgrut/optimizers/__init__.py:88-140,
// configs/base_ptir.yaml:30-38).  One kernel per tensor; the step count lives on the device
// (read by every kernel, incremented by a last one-thread kernel) so the whole update can be
// captured in a graph and replayed.
#include "internal.hpp"

namespace ptir {
namespace {

constexpr int kAdamBlock = 256;

// torch.optim.Adam (single tensor, no weight decay, no amsgrad):
//   m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g^2
//   p -= (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps)
// SelectiveAdam: p -= lr m / (sqrt(v) + eps), rows with visibility == 0 untouched.
__global__ void __launch_bounds__(kAdamBlock) k_adam(float* __restrict__ p, const float* __restrict__ g,
                                                     float* __restrict__ m, float* __restrict__ v,
                                                     const uint32_t* __restrict__ vis, uint64_t rows, uint32_t cols,
                                                     float lr, float b1, float b2, float eps, int bias_correction,
                                                     const int32_t* __restrict__ step) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    if (vis && vis[i / cols] == 0u) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    if (bias_correction) {
        const float t = float(*step + 1);
        const float bc1 = 1.0f - powf(b1, t), bc2 = 1.0f - powf(b2, t);
        p[i] -= (lr / bc1) * mi / (sqrtf(vi) / sqrtf(bc2) + eps);
    } else {
        p[i] -= lr * mi / (sqrtf(vi) + eps);
    }
}

__global__ void k_step_increment(int32_t* step) { *step += 1; }

}  // namespace

hipError_t adam_step(const AdamGroup* groups, int n, float b1, float b2, float eps, int bias_correction, int32_t* step,
                     hipStream_t s) {
    for (int k = 0; k < n; ++k) {
        const AdamGroup& a = groups[k];
        const uint64_t total = a.rows * a.cols;
        if (total == 0) continue;
        k_adam<<<unsigned((total + kAdamBlock - 1) / kAdamBlock), kAdamBlock, 0, s>>>(
            a.param, a.grad, a.exp_avg, a.exp_avg_sq, a.visibility, a.rows, a.cols, a.lr, b1, b2, eps, bias_correction,
            step);
    }
    k_step_increment<<<1, 1, 0, s>>>(step);
    return hipGetLastError();
}

}  // namespace ptir
