// Kernels of the 3DGRT primary-only tracer (threedgrt_tracer/src/kernels/cuda/referenceOptix.cu,
// referenceBwdOptix.cu); the device code is in grt_path.hpp.
#include "grt_path.hpp"

#include <hipcub/hipcub.hpp>

namespace ptir {

template <int K, int SD>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(waves_per_eu(K, SD))))
k_grt_fwd(DeviceScene S, grt::FwdOut out, const float4* __restrict__ nodes, const float4* __restrict__ wnodes,
          const float4* __restrict__ grec) {
    S.nodes = nodes;
    S.wnodes = wnodes;
    S.grec = grec;
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    __shared__ uint32_t wst[(kTraceBlock / 64) * kStackDeep];
    const Lane L{lds + threadIdx.x};
    uint32_t* wstack = wst + (threadIdx.x / 64) * kStackDeep;
    const PixelId px = pixel_of(S);
    if (!px.ok) return;
    grt::fwd_path<K, SD>(S, out, L, S.wave_traversal ? wstack : nullptr, px.pix);
}

template <int K, int SD>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(waves_per_eu(K, SD))))
k_grt_bwd(DeviceScene S, grt::BwdIO io, const float4* __restrict__ nodes, const float4* __restrict__ wnodes,
          const float4* __restrict__ grec) {
    S.nodes = nodes;
    S.wnodes = wnodes;
    S.grec = grec;
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    __shared__ uint32_t wst[(kTraceBlock / 64) * kStackDeep];
    const Lane L{lds + threadIdx.x};
    uint32_t* wstack = wst + (threadIdx.x / 64) * kStackDeep;
    const PixelId px = pixel_of(S);
    if (!px.ok) return;
    grt::bwd_path<K, SD>(S, io, L, S.wave_traversal ? wstack : nullptr, px.pix);
}

// The hit records' fold (after k_grt_bwd): per-Gaussian record counts, their prefix sum, the
// records' slots listed by Gaussian, then one lane per Gaussian sums its records.
__global__ void __launch_bounds__(256) k_grt_hist(const uint32_t* __restrict__ key, const uint32_t* __restrict__ cnt,
                                                  uint32_t paths, uint32_t* __restrict__ gcount) {
    const size_t pix = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (pix >= paths) return;
    const uint32_t n = cnt[pix];
    for (uint32_t j = 0; j < n; ++j) atomicAdd(&gcount[key[size_t(j) * paths + pix]], 1u);
}

__global__ void __launch_bounds__(256) k_grt_scatter(const uint32_t* __restrict__ key, const uint32_t* __restrict__ cnt,
                                                     uint32_t paths, const uint32_t* __restrict__ goff,
                                                     uint32_t* __restrict__ gcursor, uint32_t* __restrict__ bucket) {
    const size_t pix = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (pix >= paths) return;
    const uint32_t n = cnt[pix];
    for (uint32_t j = 0; j < n; ++j) {
        const size_t s = size_t(j) * paths + pix;
        const uint32_t id = key[s];
        bucket[goff[id] + atomicAdd(&gcursor[id], 1u)] = uint32_t(s);
    }
}

__global__ void __launch_bounds__(256) k_grt_sum(DeviceScene S, grt::BwdIO io, const uint32_t* __restrict__ gcount,
                                                 const uint32_t* __restrict__ goff, const uint32_t* __restrict__ bucket,
                                                 uint32_t n) {
    const uint32_t id = blockIdx.x * 64u + threadIdx.x / 4u;  // four lanes per Gaussian
    if (id >= n) return;
    grt::sum_records(S, io, id, bucket + goff[id], gcount[id], threadIdx.x & 3u);
}

hipError_t grt_rec_reserve(GrtRecBuffers& r, size_t paths, size_t n) {
    hipError_t e = hipSuccess;
    if (paths > r.paths_cap) {
        (void)hipFree(r.rec);
        (void)hipFree(r.key);
        (void)hipFree(r.cnt);
        (void)hipFree(r.bucket);
        r.rec = nullptr;
        r.key = r.cnt = r.bucket = nullptr;
        r.paths_cap = 0;
        const size_t slots = paths * grt::kRecPerPixel;
        if ((e = hipMalloc(&r.rec, slots * 4 * sizeof(float4))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.key, slots * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.cnt, paths * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.bucket, slots * sizeof(uint32_t))) != hipSuccess) return e;
        r.paths_cap = paths;
    }
    if (n > r.n_cap) {
        (void)hipFree(r.gcount);
        (void)hipFree(r.gcursor);
        (void)hipFree(r.goff);
        (void)hipFree(r.scan_tmp);
        r.gcount = r.gcursor = r.goff = nullptr;
        r.scan_tmp = nullptr;
        r.n_cap = 0;
        size_t tmp = 0;
        if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, r.gcount, r.goff, int(n))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.gcount, n * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.gcursor, n * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.goff, n * sizeof(uint32_t))) != hipSuccess) return e;
        if ((e = hipMalloc(&r.scan_tmp, tmp > 0 ? tmp : 1)) != hipSuccess) return e;
        r.scan_bytes = tmp;
        r.n_cap = n;
    }
    return hipSuccess;
}

void grt_rec_free(GrtRecBuffers& r) {
    (void)hipFree(r.rec);
    (void)hipFree(r.key);
    (void)hipFree(r.cnt);
    (void)hipFree(r.bucket);
    (void)hipFree(r.gcount);
    (void)hipFree(r.gcursor);
    (void)hipFree(r.goff);
    (void)hipFree(r.scan_tmp);
    r = GrtRecBuffers{};
}

namespace {
struct GrtFwdLaunch {
    const DeviceScene& S;
    const grt::FwdOut& out;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_grt_fwd<K, SD><<<S.num_blocks, kTraceBlock, 0, s>>>(S, out, S.nodes, S.wnodes, S.grec);
    }
};
struct GrtBwdLaunch {
    const DeviceScene& S;
    const grt::BwdIO& io;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_grt_bwd<K, SD><<<S.num_blocks, kTraceBlock, 0, s>>>(S, io, S.nodes, S.wnodes, S.grec);
    }
};
}  // namespace

hipError_t launch_grt_fwd(const DeviceScene& S0, const grt::FwdOut& out, int kbuf, bool short_stack, hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    return dispatch(kbuf, short_stack, GrtFwdLaunch{S, out, s});
}

hipError_t launch_grt_bwd(const DeviceScene& S0, const grt::BwdIO& io0, int kbuf, bool short_stack,
                          const GrtRecBuffers* rec, size_t n, hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    grt::BwdIO io = io0;
    const size_t paths = size_t(S.B) * S.H * S.W;
    io.rec = nullptr;
    io.rec_key = io.rec_cnt = nullptr;
    io.paths = uint32_t(paths);
    const bool records = rec && n > 0 && rec->paths_cap >= paths && rec->n_cap >= n;
    hipError_t e = hipSuccess;
    if (records) {
        io.rec = rec->rec;
        io.rec_key = rec->key;
        io.rec_cnt = rec->cnt;
        if ((e = fill_zero(rec->cnt, paths * sizeof(uint32_t), s)) != hipSuccess) return e;
    }
    if ((e = dispatch(kbuf, short_stack, GrtBwdLaunch{S, io, s})) != hipSuccess) return e;
    if (!records) return hipSuccess;
    const unsigned pb = unsigned((paths + 255) / 256), nb = unsigned((n + 63) / 64);
    if ((e = fill_zero(rec->gcount, n * sizeof(uint32_t), s)) != hipSuccess) return e;
    if ((e = fill_zero(rec->gcursor, n * sizeof(uint32_t), s)) != hipSuccess) return e;
    k_grt_hist<<<pb, 256, 0, s>>>(rec->key, rec->cnt, uint32_t(paths), rec->gcount);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t tmp = rec->scan_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(rec->scan_tmp, tmp, rec->gcount, rec->goff, int(n), s)) != hipSuccess)
        return e;
    k_grt_scatter<<<pb, 256, 0, s>>>(rec->key, rec->cnt, uint32_t(paths), rec->goff, rec->gcursor, rec->bucket);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    k_grt_sum<<<nb, 256, 0, s>>>(S, io, rec->gcount, rec->goff, rec->bucket, uint32_t(n));
    return hipGetLastError();
}

}  // namespace ptir
