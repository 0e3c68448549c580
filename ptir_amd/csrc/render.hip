// Caller-side image work of a render, on the device: camera rays for several samples per
// pixel, and the resolve of the averaged tracer outputs into the render outputs.
//
// The reference does both with torch ops on the trainer side
// (threedgptir_tracer/tracer.py:505-526,646-706; threedgrut/model/filters.py:6-69).  Here
//   * k_camera_rays writes one ray per (sample, view, pixel): origin = the view's ray origin,
//     direction = normalize(((px - 0.5 + jx - cx) / fx, (py - 0.5 + jy - cy) / fy, 1)), with
//     the jitter either shared by all pixels of a sample or per pixel;
//   * the resolve stage: depth / opacity (non-finite -> 0), pbr, direct and indirect times
//     opacity, a 3x3 B-spline a-trous filter on pbr/direct/indirect/light (replicate
//     borders, dilation 2^i at iteration i), and unit normals.  The 3x3 kernel
//     [1 2 1]^T [1 2 1] / 16 is separable, so each iteration is a horizontal and a vertical
//     pass over a 12-channel image [B,H,W,12] (pbr, direct, indirect, light).  The backward
//     runs the transposed passes and differentiates the per-pixel maps.
#include "internal.hpp"

namespace ptir {
namespace {

constexpr int kRenderBlock = 256;
constexpr int kFiltC = 12;  // pbr, direct, indirect, light

__global__ void __launch_bounds__(kRenderBlock) k_camera_rays(const float* __restrict__ pix, const float* __restrict__ K,
                                                              const float* __restrict__ jit, int jit_per_pixel,
                                                              const float* __restrict__ ori, int S, int B, int H,
                                                              int W, float* __restrict__ out_o,
                                                              float* __restrict__ out_d) {
    const size_t n = size_t(S) * B * H * W;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t hw = size_t(H) * W;
    const size_t per_sample = size_t(B) * hw;
    const int s = int(i / per_sample);
    const size_t rest = i - size_t(s) * per_sample;   // (b, y, x) of the base batch
    const size_t yx = rest % hw;
    const float px = pix[2 * rest] - 0.5f, py = pix[2 * rest + 1] - 0.5f;
    const float* j = jit + 2 * (jit_per_pixel ? size_t(s) * hw + yx : size_t(s));
    const float fx = K[0], fy = K[1], cx = K[2], cy = K[3];
    const float dx = (px + j[0] - cx) / fx, dy = (py + j[1] - cy) / fy;
    const float inv = 1.0f / fmaxf(sqrtf(dx * dx + dy * dy + 1.0f), 1e-12f);
    out_d[3 * i] = dx * inv;
    out_d[3 * i + 1] = dy * inv;
    out_d[3 * i + 2] = inv;
    out_o[3 * i] = ori[3 * rest];
    out_o[3 * i + 1] = ori[3 * rest + 1];
    out_o[3 * i + 2] = ori[3 * rest + 2];
}

__device__ __forceinline__ float finite_or_zero(float v) { return isfinite(v) ? v : 0.f; }

// Per pixel: normalized depth, unit normal, and the 12 channels to filter.
__global__ void __launch_bounds__(kRenderBlock) k_resolve_prep(size_t npix, const float* __restrict__ op,
                                                               const float* __restrict__ dist,
                                                               const float* __restrict__ pbr,
                                                               const float* __restrict__ dir,
                                                               const float* __restrict__ ind,
                                                               const float* __restrict__ light,
                                                               const float* __restrict__ nrm,
                                                               float* __restrict__ out_dist,
                                                               float* __restrict__ out_nrm, float* __restrict__ img) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float o = op[p];
    out_dist[p] = finite_or_zero(dist[p] / o);
    const float nx = nrm[3 * p], ny = nrm[3 * p + 1], nz = nrm[3 * p + 2];
    const float inv = 1.0f / fmaxf(sqrtf(nx * nx + ny * ny + nz * nz), 1e-12f);
    out_nrm[3 * p] = nx * inv;
    out_nrm[3 * p + 1] = ny * inv;
    out_nrm[3 * p + 2] = nz * inv;
    float* c = img + kFiltC * p;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = pbr[3 * p + k] * o;
        c[3 + k] = dir[3 * p + k] * o;
        c[6 + k] = ind[3 * p + k] * o;
        c[9 + k] = light[3 * p + k];
    }
}

// One 1D pass of the [1 2 1]/4 kernel at dilation d along x (axis 0) or y (axis 1) with
// replicate borders: out[i] = (in[c(i-d)] + 2 in[i] + in[c(i+d)]) / 4.
__global__ void __launch_bounds__(kRenderBlock) k_filter_pass(const float* __restrict__ in, float* __restrict__ out,
                                                              int B, int H, int W, int d, int axis) {
    const size_t n = size_t(B) * H * W * kFiltC;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = int(i % kFiltC);
    const size_t p = i / kFiltC;
    const int x = int(p % W), y = int((p / W) % H);
    const size_t row = p - size_t(x) - size_t(y) * W;  // pixel index of (b, 0, 0)
    int a, b;
    size_t pa, pb;
    if (axis == 0) {
        a = max(x - d, 0);
        b = min(x + d, W - 1);
        pa = row + size_t(y) * W + a;
        pb = row + size_t(y) * W + b;
    } else {
        a = max(y - d, 0);
        b = min(y + d, H - 1);
        pa = row + size_t(a) * W + x;
        pb = row + size_t(b) * W + x;
    }
    out[i] = 0.25f * in[kFiltC * pa + c] + 0.5f * in[i] + 0.25f * in[kFiltC * pb + c];
}

// Transpose of k_filter_pass: gin[j] = sum over outputs i that read j of their weight times
// gout[i].  Output j itself reads j with 1/2; outputs j+d and j-d read j with 1/4 each when
// they exist; at a border the clamped reads of the border output land on j as well.
__global__ void __launch_bounds__(kRenderBlock) k_filter_pass_t(const float* __restrict__ gout,
                                                                float* __restrict__ gin, int B, int H, int W, int d,
                                                                int axis) {
    const size_t n = size_t(B) * H * W * kFiltC;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = int(i % kFiltC);
    const size_t p = i / kFiltC;
    const int x = int(p % W), y = int((p / W) % H);
    const size_t row = p - size_t(x) - size_t(y) * W;
    const int j = axis == 0 ? x : y;
    const int len = axis == 0 ? W : H;
    auto at = [&](int k) {  // gout at position k along the axis
        const size_t q = axis == 0 ? row + size_t(y) * W + k : row + size_t(k) * W + x;
        return gout[kFiltC * q + c];
    };
    float g = 0.5f * at(j);
    // outputs k with clamp(k - d) == j: k = j + d, plus k in [0, d - 1] when j == 0
    if (j + d < len) g += 0.25f * at(j + d);
    if (j == 0)
        for (int k = 0; k < min(d, len); ++k) g += 0.25f * at(k);
    // outputs k with clamp(k + d) == j: k = j - d, plus k in [len - d, len - 1] when j == len - 1
    if (j - d >= 0) g += 0.25f * at(j - d);
    if (j == len - 1)
        for (int k = max(len - d, 0); k < len; ++k) g += 0.25f * at(k);
    gin[i] = g;
}

__global__ void __launch_bounds__(kRenderBlock) k_resolve_prep_bwd(
    size_t npix, const float* __restrict__ op, const float* __restrict__ dist, const float* __restrict__ pbr,
    const float* __restrict__ dir, const float* __restrict__ ind, const float* __restrict__ nrm,
    const float* __restrict__ g_dist, const float* __restrict__ g_nrm, const float* __restrict__ g_img,
    float* __restrict__ d_op, float* __restrict__ d_dist, float* __restrict__ d_pbr, float* __restrict__ d_dir,
    float* __restrict__ d_ind, float* __restrict__ d_light, float* __restrict__ d_nrm) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const float o = op[p];
    const float* gc = g_img + kFiltC * p;
    float gop = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        d_pbr[3 * p + k] = gc[k] * o;
        d_dir[3 * p + k] = gc[3 + k] * o;
        d_ind[3 * p + k] = gc[6 + k] * o;
        d_light[3 * p + k] = gc[9 + k];
        gop += gc[k] * pbr[3 * p + k] + gc[3 + k] * dir[3 * p + k] + gc[6 + k] * ind[3 * p + k];
    }
    // depth / opacity where the quotient is finite (nan_to_num sends the rest to 0)
    const float q = dist[p] / o;
    if (isfinite(q)) {
        d_dist[p] = g_dist[p] / o;
        gop -= g_dist[p] * q / o;
    } else {
        d_dist[p] = 0.f;
    }
    d_op[p] = gop;
    // x / max(|x|, 1e-12)
    const float nx = nrm[3 * p], ny = nrm[3 * p + 1], nz = nrm[3 * p + 2];
    const float len = sqrtf(nx * nx + ny * ny + nz * nz);
    const float gx = g_nrm[3 * p], gy = g_nrm[3 * p + 1], gz = g_nrm[3 * p + 2];
    if (len > 1e-12f) {
        const float inv = 1.0f / len;
        const float ux = nx * inv, uy = ny * inv, uz = nz * inv;
        const float dot = ux * gx + uy * gy + uz * gz;
        d_nrm[3 * p] = (gx - ux * dot) * inv;
        d_nrm[3 * p + 1] = (gy - uy * dot) * inv;
        d_nrm[3 * p + 2] = (gz - uz * dot) * inv;
    } else {
        d_nrm[3 * p] = gx * 1e12f;
        d_nrm[3 * p + 1] = gy * 1e12f;
        d_nrm[3 * p + 2] = gz * 1e12f;
    }
}

inline unsigned grid_of(size_t n) { return unsigned((n + kRenderBlock - 1) / kRenderBlock); }

}  // namespace

hipError_t camera_rays(const float* pix, const float* K, const float* jit, int jit_per_pixel, const float* ori, int S,
                       int B, int H, int W, float* out_o, float* out_d, hipStream_t s) {
    const size_t n = size_t(S) * B * H * W;
    if (n == 0) return hipSuccess;
    k_camera_rays<<<grid_of(n), kRenderBlock, 0, s>>>(pix, K, jit, jit_per_pixel, ori, S, B, H, W, out_o, out_d);
    return hipGetLastError();
}

// img and tmp are [B,H,W,12] work images; the filtered result ends in img.
static void filter(float* img, float* tmp, int B, int H, int W, int iterations, hipStream_t s) {
    const size_t n = size_t(B) * H * W * kFiltC;
    for (int it = 0; it < iterations; ++it) {
        const int d = 1 << it;
        k_filter_pass<<<grid_of(n), kRenderBlock, 0, s>>>(img, tmp, B, H, W, d, 0);
        k_filter_pass<<<grid_of(n), kRenderBlock, 0, s>>>(tmp, img, B, H, W, d, 1);
    }
}

hipError_t resolve_forward(const ResolveIO& io, hipStream_t s) {
    const size_t npix = size_t(io.B) * io.H * io.W;
    if (npix == 0) return hipSuccess;
    k_resolve_prep<<<grid_of(npix), kRenderBlock, 0, s>>>(npix, io.op, io.dist, io.pbr, io.dir, io.ind, io.light,
                                                          io.nrm, io.out_dist, io.out_nrm, io.img);
    filter(io.img, io.tmp, io.B, io.H, io.W, io.iterations, s);
    return hipGetLastError();
}

hipError_t resolve_backward(const ResolveIO& io, const ResolveGrads& g, hipStream_t s) {
    const size_t npix = size_t(io.B) * io.H * io.W;
    if (npix == 0) return hipSuccess;
    const size_t n = npix * kFiltC;
    // transposed passes in reverse order: vertical then horizontal, last iteration first
    for (int it = io.iterations - 1; it >= 0; --it) {
        const int d = 1 << it;
        k_filter_pass_t<<<grid_of(n), kRenderBlock, 0, s>>>(io.img, io.tmp, io.B, io.H, io.W, d, 1);
        k_filter_pass_t<<<grid_of(n), kRenderBlock, 0, s>>>(io.tmp, io.img, io.B, io.H, io.W, d, 0);
    }
    k_resolve_prep_bwd<<<grid_of(npix), kRenderBlock, 0, s>>>(npix, io.op, io.dist, io.pbr, io.dir, io.ind, io.nrm,
                                                              g.g_dist, g.g_nrm, io.img, g.d_op, g.d_dist, g.d_pbr,
                                                              g.d_dir, g.d_ind, g.d_light, g.d_nrm);
    return hipGetLastError();
}

}  // namespace ptir
