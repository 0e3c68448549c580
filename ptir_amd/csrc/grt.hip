// Kernels of the 3DGRT primary-only tracer (threedgrt_tracer/src/kernels/cuda/referenceOptix.cu,
// referenceBwdOptix.cu); the device code is in grt_path.hpp.
#include "grt_path.hpp"

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

hipError_t launch_grt_bwd(const DeviceScene& S0, const grt::BwdIO& io, int kbuf, bool short_stack, hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    return dispatch(kbuf, short_stack, GrtBwdLaunch{S, io, s});
}

}  // namespace ptir
