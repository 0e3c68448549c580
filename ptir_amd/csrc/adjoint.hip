// Exact-adjoint kernel and its launch; the device code is in adjoint_path.hpp.
#include "adjoint_path.hpp"

namespace ptir {

template <int K, int SD>
__global__ void __launch_bounds__(kTraceBlock) k_trace_adjoint(DeviceScene S, AdjIO io) {
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    const Lane L{lds + threadIdx.x};
    const PixelId px = pixel_of(S);
    if (!px.ok) return;
    trace_adjoint_path<K, SD>(S, io, L, px);
}

struct AdjLaunch {
    const DeviceScene& S;
    const AdjIO& io;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_trace_adjoint<K, SD><<<S.num_blocks, kTraceBlock, 0, s>>>(S, io);
    }
};

uint64_t adjoint_scratch_bytes(uint64_t paths, uint32_t max_bounces) {
    return adjoint_scratch_words(paths, max_bounces) * sizeof(float4);
}

hipError_t launch_trace_adjoint(const DeviceScene& S0, const AdjIO& io, int kbuf, bool short_stack, hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    return dispatch(kbuf, short_stack, AdjLaunch{S, io, s});
}

}  // namespace ptir
