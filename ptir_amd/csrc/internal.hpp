// Host-side interfaces between the C ABI (api.cpp) and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>

#include "device.hpp"
#include "ploc.hpp"
#include "wide.hpp"

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
    float4* nodes = nullptr;         // [(N-1)*4] binary nodes (LBVH or PLOC)
    float4* wide = nullptr;          // [(N-1)*8] 4-wide nodes collapsed from them (the traversal's)
    uint32_t* wflag = nullptr;       // [N-1] binary node is a 4-wide node
    uint32_t* wpre = nullptr;        // [N-1] exclusive prefix sum of wflag
    int32_t* wmap = nullptr;         // [N-1] binary node -> 4-wide node or -1
    int32_t* wneed = nullptr;        // [3] stack entries the 4-wide walks need, 4-wide node count, binary depth
    void* wscan_tmp = nullptr;
    size_t wscan_tmp_bytes = 0;
    int32_t* parent = nullptr;       // [2N-1]
    uint32_t* flags = nullptr;       // [N-1]
    int32_t* bounds_i = nullptr;     // [6] order-preserving int encoding of the scene box
    float* bounds = nullptr;         // [6]
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
    // PLOC builder work buffers (allocated on its first build)
    uint32_t pcap = 0;
    int32_t* pc[2] = {nullptr, nullptr};  // [N] cluster links, ping-pong
    float4* pb[2] = {nullptr, nullptr};   // [2N] cluster boxes, ping-pong
    int32_t* pnn = nullptr;               // [N]
    unsigned long long* pval = nullptr;   // [N]
    unsigned long long* pscan = nullptr;  // [N]
    ploc::State* pstate = nullptr;        // [2], ping-pong
    int32_t* pbad = nullptr;              // [1] structure check result
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
};

hipError_t bvh_reserve(BvhBuffers& b, uint32_t n);
void bvh_free(BvhBuffers& b);
hipError_t bvh_build(BvhBuffers& b, const KernelOpts& o, const float* pos, const float* rot, const float* scl,
                     const float* dns, uint32_t n, bool rebuild, int builder, hipStream_t s);
// Traversal stack entries the built 4-wide hierarchy needs, its node count and the depth of
// the binary one (synchronises the stream).
hipError_t bvh_stack_need(BvhBuffers& b, hipStream_t s, int32_t* host_need, int32_t* host_count, int32_t* host_depth);
hipError_t prep_records(BvhBuffers& b, const KernelOpts& o, const float* density, const float* material,
                        const float* snormal, hipStream_t s);
hipError_t scatter_material_grad(const BvhBuffers& b, float* dmat, hipStream_t s);
// Zero a device buffer with a kernel (safe inside captured graphs, unlike hipMemsetAsync).
hipError_t fill_zero(void* p, size_t bytes, hipStream_t s);

// Environment importance table (CDF layout) built on the device; scratch of
// env_table_scratch_bytes(h, w) bytes, 8-byte aligned.
uint64_t env_table_scratch_bytes(int h, int w);
hipError_t env_table_build(const float* env, int H, int W, int h, int w, float* table, void* scratch, hipStream_t s);
void set_thread_error(const std::string& msg);

// Render-side image work (render.hip).
hipError_t camera_rays(const float* pix, const float* K, const float* jit, int jit_per_pixel, const float* ori, int S,
                       int B, int H, int W, float* out_o, float* out_d, hipStream_t s);
struct ResolveIO {
    int B, H, W, iterations;                                   // iterations of the a-trous filter (0 = none)
    const float *op, *dist, *pbr, *dir, *ind, *light, *nrm;    // [B,H,W,1|3] averaged tracer outputs
    float *out_dist, *out_nrm;                                 // [B,H,W,1], [B,H,W,3]
    float *img, *tmp;                                          // [B,H,W,12] work images
};
struct ResolveGrads {
    const float *g_dist, *g_nrm;                               // incoming gradients of out_dist, out_nrm
    float *d_op, *d_dist, *d_pbr, *d_dir, *d_ind, *d_light, *d_nrm;
};
hipError_t resolve_forward(const ResolveIO& io, hipStream_t s);

// Adam on optimised tensors (adam.hip).
struct AdamGroup {
    float* param;
    const float* grad;
    float *exp_avg, *exp_avg_sq;
    const uint32_t* visibility;  // [rows] nonzero = update the row, or null
    uint64_t rows;
    uint32_t cols;
    float lr;
};
hipError_t adam_step(const AdamGroup* groups, int n, float b1, float b2, float eps, int bias_correction, int32_t* step,
                     hipStream_t s);
hipError_t resolve_backward(const ResolveIO& io, const ResolveGrads& g, hipStream_t s);

hipError_t launch_trace_fwd(const DeviceScene& S, const FwdOut& out, int kbuf, bool short_stack, hipStream_t s);
// 3DGRT primary-only tracer (grt.hip)
hipError_t launch_grt_fwd(const DeviceScene& S, const grt::FwdOut& out, int kbuf, bool short_stack, hipStream_t s);
// Hit records of the 3DGRT backward and the per-Gaussian lists that fold them (grt.hip): the
// tracer keeps them between calls and grows them with the launch (records: 64 B per slot,
// kRecPerPixel slots per pixel).
struct GrtRecBuffers {
    float4* rec = nullptr;        // [paths * kRecPerPixel * 4]
    uint32_t* key = nullptr;      // [paths * kRecPerPixel] Gaussian of each slot
    uint32_t* cnt = nullptr;      // [paths] records per pixel
    uint32_t* bucket = nullptr;   // [paths * kRecPerPixel] slots grouped by Gaussian
    uint32_t* gcount = nullptr;   // [N] records per Gaussian
    uint32_t* gcursor = nullptr;  // [N]
    uint32_t* goff = nullptr;     // [N] exclusive prefix sum of gcount
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    size_t paths_cap = 0, n_cap = 0;
};
hipError_t grt_rec_reserve(GrtRecBuffers& r, size_t paths, size_t n);
void grt_rec_free(GrtRecBuffers& r);
// rec: the record buffers reserved for this launch (then io's record fields are set from
// them), or null (every hit adds its gradients by atomics).
hipError_t launch_grt_bwd(const DeviceScene& S, const grt::BwdIO& io, int kbuf, bool short_stack, const GrtRecBuffers* rec,
                          size_t n, hipStream_t s);
hipError_t launch_trace_bwd(const DeviceScene& S, const BwdIO& io, int kbuf, bool short_stack, hipStream_t s);
// Wavefront schedule of the forward (wavefront.hip); `work` holds wavefront_bytes(paths, max_bounces).
uint64_t wavefront_bytes(uint64_t paths, uint32_t max_bounces);
hipError_t launch_trace_fwd_wavefront(const DeviceScene& S, const FwdOut& out, void* work, int kbuf, bool short_stack,
                                      hipStream_t s);
// Exact adjoint (adjoint.hip); io.scratch holds adjoint_scratch_bytes(paths, max_bounces) bytes.
uint64_t adjoint_scratch_bytes(uint64_t paths, uint32_t max_bounces);
hipError_t launch_trace_adjoint(const DeviceScene& S, const AdjIO& io, int kbuf, bool short_stack, hipStream_t s);

constexpr int kTraceBlock = 128;  // 2 waves: a 16x8 pixel tile
constexpr int kStackDeep = 64;    // traversal stack entries per lane (LDS): any accepted hierarchy
// hierarchies whose walks need at most 36 stack entries run with this one: 36 words per lane
// of LDS (with the wave stacks, 18.9 KB per 128-lane block) still fit 8 blocks = four waves per
// SIMD in 160 KB of LDS
constexpr int kStackShort = 36;

// bytes of a tape with `segs` segments per path and `cap` hits per path
inline uint64_t tape_bytes(uint64_t paths, uint32_t segs, uint32_t cap) {
    const uint64_t rec = uint64_t(kTapePathWords) + uint64_t(kTapeSegWords) * segs;
    return 4ull * (kTapeHeaderWords + paths * rec) + 8ull * paths * cap + 8ull * kTapeChunk * tape_pool_chunks(paths);
}

// ---------------------------------------------------------------- neural radiance cache
namespace nrc {
constexpr int kLevels = 16;            // hash-grid levels, 2 features each
constexpr int kLog2Table = 19;
constexpr int kTableSize = 1 << kLog2Table;  // entries per level
constexpr int kHidden = 4;             // hidden layers of width 64
constexpr int kOut = 3;
constexpr int kOutOffset = kHidden * 4096;   // output layer [16 rows (3 used), 64] after the hidden ones
constexpr int kParams = kOutOffset + 16 * 64;
constexpr int kFrags = 70;             // packed bf16 operand fragments (forward 34, backward 36)
constexpr float kZMax = 15.f;          // exp output: exponent clamp
constexpr float kLossEps = 0.01f;      // relative L2: (p - t)^2 / (p^2 + eps)
// query record, float32 x kRecFloats, written by the tracer at the first secondary vertex
constexpr int kRecFloats = 20;
constexpr int kRecFlag = 0, kRecPos = 1, kRecDir = 4, kRecNrm = 7, kRecRough = 10, kRecAlb = 11, kRecThr = 14,
              kRecTarget = 17;

struct AdamParams {
    float lr, beta1, beta2, eps;
};
}  // namespace nrc

// resolution of hash-grid level l: round(16 * 128^(l/15)), 16 ... 2048
inline int nrc_level_res(int l) {
    return int(floor(16.0 * exp(double(l) * log(2048.0 / 16.0) / double(nrc::kLevels - 1)) + 0.5));
}
inline uint64_t nrc_table_floats() { return uint64_t(nrc::kLevels) * nrc::kTableSize * 2; }

struct NrcBuffers {
    uint64_t max_batch = 0;
    float *w = nullptr, *dw = nullptr, *mw = nullptr, *vw = nullptr;               // [kParams] FP32 master + Adam
    float *table = nullptr, *dtable = nullptr, *mt = nullptr, *vt = nullptr;       // [kLevels, kTableSize, 2]
    void* frags = nullptr;          // [kFrags, 64] x 16 B packed bf16 operands
    void* act = nullptr;            // [5, max_batch, 64] bf16 activations of a training batch
    float* z = nullptr;             // [max_batch, 4] output pre-activations
    float* dx = nullptr;            // [max_batch, 32] hash-feature gradients
    float* dw_part = nullptr;       // [tiles, kParams] per-tile weight gradients
    float* dw_chunks = nullptr;     // [ceil(tiles / 32), kParams] sums over blocks of 32 tiles
    float* loss = nullptr;          // [1]
    uint32_t* step = nullptr;       // [1] Adam step counter
};

hipError_t nrc_alloc(NrcBuffers& b, uint64_t max_batch, uint32_t seed, hipStream_t s);
void nrc_free(NrcBuffers& b);
hipError_t nrc_pack(NrcBuffers& b, hipStream_t s);
hipError_t nrc_query(const NrcBuffers& b, const float* rec, uint64_t n, const float* box, float* pbr, float* comp,
                     float* pred, hipStream_t s);
hipError_t nrc_backprop(NrcBuffers& b, const float* rec, uint64_t n, const float* box, hipStream_t s);
hipError_t nrc_step(NrcBuffers& b, const nrc::AdamParams& h, hipStream_t s);

}  // namespace ptir
