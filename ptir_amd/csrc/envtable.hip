// Environment importance table built on the device, for CDF-inversion sampling.
//
// The reference builds a Vose alias table on the host every 32 training steps
// (threedgrut/model/aliastable.py:244-297): the environment is area-resampled to the table
// grid, each cell weighted by its Rec.709 luminance times its equirect solid angle, and the
// per-steradian pdf is w / (sum(w) * solid_angle).  This file computes the same weights and
// pdf on the GPU, in FP64, and replaces the alias construction (a sequential algorithm) by
// two levels of cumulative distributions built with FP64 prefix sums:
//   plane 0  [h, w]  conditional CDF of each row, inclusive: c[y][x] = sum_{x'<=x} w / rowsum
//   plane 1  [h, w]  column 0: marginal CDF of the rows, inclusive: m[y] = sum_{y'<=y} rowsum / total
//                    (columns 1..w-1 are zero)
//   plane 2  [h, w]  pdf per steradian (the same plane the alias layout has)
// The last entry of every CDF is exactly 1.  The tracer inverts them with two binary
// searches (first entry > u) and then places the sample uniformly in the chosen cell,
// drawing the same four random numbers as the alias sampler.
//
// Kernels: k_env_weights (one thread per cell: area average, luminance, solid angle),
// k_env_rowsum (one block per row), k_env_marginal (one block), k_env_finish (one block
// per row: conditional CDF and pdf).  Scratch: FP64 cell weights, row sums and the total.
#include "internal.hpp"

namespace ptir {
namespace {

constexpr int kEnvBlock = 256;

__device__ __forceinline__ double solid_angle_of_row(int y, int h, int w) {
    const double pi = 3.14159265358979323846;
    const double e0 = (double(y) / double(h) - 0.5) * pi;
    const double e1 = (double(y + 1) / double(h) - 0.5) * pi;
    const double sa = (2.0 * pi / double(w)) * (sin(e1) - sin(e0));
    return sa > 0.0 ? sa : 0.0;
}

// Area resampling of F.interpolate(mode="area") (adaptive average pooling): output cell i of
// n covers source rows [floor(i*N/n), ceil((i+1)*N/n)).
__global__ void __launch_bounds__(kEnvBlock) k_env_weights(const float4* __restrict__ env, int H, int W, int h, int w,
                                                           double* __restrict__ wts) {
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= h * w) return;
    const int y = cell / w, x = cell - y * w;
    const int y0 = int((long long)y * H / h), y1 = int(((long long)(y + 1) * H + h - 1) / h);
    const int x0 = int((long long)x * W / w), x1 = int(((long long)(x + 1) * W + w - 1) / w);
    double r = 0.0, g = 0.0, b = 0.0;
    for (int yy = y0; yy < y1; ++yy)
        for (int xx = x0; xx < x1; ++xx) {
            const float4 t = env[size_t(yy) * W + xx];
            r += double(t.x);
            g += double(t.y);
            b += double(t.z);
        }
    const double inv = 1.0 / double((y1 - y0) * (x1 - x0));
    r *= inv;
    g *= inv;
    b *= inv;
    // non-finite or non-positive channels carry no light (aliastable.py _environment_luminance)
    auto clean = [](double v) { return (isfinite(v) && v > 0.0) ? v : 0.0; };
    const double lum = 0.2126 * clean(r) + 0.7152 * clean(g) + 0.0722 * clean(b);
    wts[cell] = lum * solid_angle_of_row(y, h, w);
}

// Block-wide inclusive scan of one double per thread (Hillis-Steele in LDS).
__device__ __forceinline__ double block_inclusive_scan(double v, double* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kEnvBlock; off <<= 1) {
        const double add = threadIdx.x >= unsigned(off) ? lds[threadIdx.x - off] : 0.0;
        __syncthreads();
        lds[threadIdx.x] += add;
        __syncthreads();
    }
    const double r = lds[threadIdx.x];
    __syncthreads();
    return r;
}

// Each thread owns a contiguous span of `per` entries of a row of n; returns the exclusive
// prefix of the thread's span (FP64) and the row total in *total.
__device__ __forceinline__ double span_prefix(const double* v, int n, int per, double* lds, double* total) {
    const int a = int(threadIdx.x) * per, e = min(n, a + per);
    double s = 0.0;
    for (int i = a; i < e; ++i) s += v[i];
    const double inc = block_inclusive_scan(s, lds);
    if (threadIdx.x == kEnvBlock - 1) lds[0] = inc;
    __syncthreads();
    *total = lds[0];
    __syncthreads();
    return inc - s;
}

__global__ void __launch_bounds__(kEnvBlock) k_env_rowsum(const double* __restrict__ wts, int h, int w,
                                                          double* __restrict__ rowsum) {
    __shared__ double lds[kEnvBlock];
    const int y = blockIdx.x;
    const int per = (w + kEnvBlock - 1) / kEnvBlock;
    double tot;
    (void)span_prefix(wts + size_t(y) * w, w, per, lds, &tot);
    if (threadIdx.x == 0) rowsum[y] = tot;
}

// Total, the fallback to pure solid-angle weights when the environment carries no finite
// positive light (aliastable.py:277-280), and the marginal CDF of the rows.
__global__ void __launch_bounds__(kEnvBlock) k_env_marginal(double* __restrict__ rowsum, int h, int w,
                                                            double* __restrict__ info, float* __restrict__ table) {
    __shared__ double lds[kEnvBlock];
    const int per = (h + kEnvBlock - 1) / kEnvBlock;
    double total;
    double pre = span_prefix(rowsum, h, per, lds, &total);
    const bool fallback = !(isfinite(total) && total > 0.0);
    if (fallback) {
        // every cell weighs its solid angle: a row weighs w * its cell solid angle
        const int a = int(threadIdx.x) * per, e = min(h, a + per);
        for (int i = a; i < e; ++i) rowsum[i] = double(w) * solid_angle_of_row(i, h, w);
        __syncthreads();
        pre = span_prefix(rowsum, h, per, lds, &total);
    }
    const size_t plane = size_t(h) * w;
    const int a = int(threadIdx.x) * per, e = min(h, a + per);
    double run = pre;
    for (int i = a; i < e; ++i) {
        run += rowsum[i];
        const float m = (i == h - 1) ? 1.0f : float(run / total);
        float* row = table + plane + size_t(i) * w;
        row[0] = m;
    }
    if (threadIdx.x == 0) {
        info[0] = total;
        info[1] = fallback ? 1.0 : 0.0;
    }
}

__global__ void __launch_bounds__(kEnvBlock) k_env_finish(const double* __restrict__ wts,
                                                          const double* __restrict__ rowsum,
                                                          const double* __restrict__ info, int h, int w,
                                                          float* __restrict__ table) {
    __shared__ double lds[kEnvBlock];
    const int y = blockIdx.x;
    const bool fallback = info[1] != 0.0;
    const double total = info[0];
    const double sa = solid_angle_of_row(y, h, w);
    const int per = (w + kEnvBlock - 1) / kEnvBlock;
    const double* row = wts + size_t(y) * w;
    // in the fallback every weight is the cell's solid angle; scan those instead
    const int a = int(threadIdx.x) * per, e = min(w, a + per);
    double s = 0.0;
    for (int i = a; i < e; ++i) s += fallback ? sa : row[i];
    const double inc = block_inclusive_scan(s, lds);
    const double rs = rowsum[y];
    const size_t plane = size_t(h) * w;
    float* cdf = table + size_t(y) * w;
    float* marg = table + plane + size_t(y) * w;
    float* pdf = table + 2 * plane + size_t(y) * w;
    double run = inc - s;
    for (int i = a; i < e; ++i) {
        const double wi = fallback ? sa : row[i];
        run += wi;
        float c;
        if (i == w - 1) c = 1.0f;
        else if (rs > 0.0) c = float(run / rs);
        else c = float(double(i + 1) / double(w));  // a row without mass is never chosen
        cdf[i] = c;
        if (i > 0) marg[i] = 0.0f;
        pdf[i] = sa > 0.0 ? float(wi / (total * sa)) : 0.0f;
    }
}

}  // namespace

uint64_t env_table_scratch_bytes(int h, int w) {
    return 8ull * (uint64_t(h) * uint64_t(w) + uint64_t(h) + 2ull);
}

hipError_t env_table_build(const float* env, int H, int W, int h, int w, float* table, void* scratch, hipStream_t s) {
    double* wts = static_cast<double*>(scratch);
    double* rowsum = wts + size_t(h) * w;
    double* info = rowsum + h;
    const int cells = h * w;
    k_env_weights<<<(cells + kEnvBlock - 1) / kEnvBlock, kEnvBlock, 0, s>>>(reinterpret_cast<const float4*>(env), H, W,
                                                                            h, w, wts);
    k_env_rowsum<<<h, kEnvBlock, 0, s>>>(wts, h, w, rowsum);
    k_env_marginal<<<1, kEnvBlock, 0, s>>>(rowsum, h, w, info, table);
    k_env_finish<<<h, kEnvBlock, 0, s>>>(wts, rowsum, info, h, w, table);
    return hipGetLastError();
}

}  // namespace ptir
