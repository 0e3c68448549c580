// Host-side interfaces between the C ABI (api.cpp) and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace ptir {

struct KernelOpts {
    int kernel_degree;
    float min_response;
    int density_clamping;
};

// Persistent device buffers of one hierarchy.
struct BvhBuffers {
    uint32_t n = 0, cap = 0;
    float4* aabb = nullptr;          // [2N] per-Gaussian world box (original order)
    unsigned long long* keys = nullptr;   // [N]
    unsigned long long* keys_sorted = nullptr;
    uint32_t* idx = nullptr;         // [N]
    uint32_t* leaf_id = nullptr;     // [N] leaf slot -> original index
    float4* nodes = nullptr;         // [(N-1)*4]
    int32_t* parent = nullptr;       // [2N-1]
    uint32_t* flags = nullptr;       // [N-1]
    int32_t* bounds_i = nullptr;     // [6] order-preserving int encoding of the scene box
    float* bounds = nullptr;         // [6]
    int32_t* depth = nullptr;        // [1]
    void* sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    // per-trace records (leaf order)
    float4* grec = nullptr;          // [4N]
    float4* srec = nullptr;          // [2N]
    float4* prec = nullptr;          // [N]
    float* dmat_leaf = nullptr;      // [5N]
    unsigned long long* counters = nullptr;  // [kCounterCopies * kCounterStride]
    int32_t* error_flag = nullptr;   // [1]
    int32_t* tape_state = nullptr;   // [1]
};

hipError_t bvh_reserve(BvhBuffers& b, uint32_t n);
void bvh_free(BvhBuffers& b);
hipError_t bvh_build(BvhBuffers& b, const KernelOpts& o, const float* pos, const float* rot, const float* scl,
                     const float* dns, uint32_t n, bool rebuild, hipStream_t s);
hipError_t bvh_depth(BvhBuffers& b, hipStream_t s, int32_t* host_depth);
hipError_t prep_records(BvhBuffers& b, const KernelOpts& o, const float* density, const float* material,
                        const float* snormal, hipStream_t s);
hipError_t scatter_material_grad(const BvhBuffers& b, float* dmat, hipStream_t s);
// Zero a device buffer with a kernel (safe inside captured graphs, unlike hipMemsetAsync).
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);

hipError_t launch_trace_fwd(const DeviceScene& S, const FwdOut& out, int kbuf, bool short_stack, hipStream_t s);
hipError_t launch_trace_bwd(const DeviceScene& S, const BwdIO& io, int kbuf, bool short_stack, hipStream_t s);

constexpr int kTraceBlock = 128;  // 2 waves: a 16x8 pixel tile
constexpr int kStackDeep = 64;    // traversal stack entries per lane (LDS): any accepted hierarchy
constexpr int kStackShort = 32;   // hierarchies of depth <= 32 run with this one (more waves fit)

// bytes of a tape with `segs` segments per path and `cap` hits per path
inline uint64_t tape_bytes(uint64_t paths, uint32_t segs, uint32_t cap) {
    const uint64_t rec = uint64_t(kTapePathWords) + uint64_t(kTapeSegWords) * segs;
    return 4ull * (kTapeHeaderWords + paths * rec) + 8ull * paths * cap + 8ull * kTapeChunk * tape_pool_chunks(paths);
}

}  // namespace ptir
