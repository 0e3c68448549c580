// Device-side types and helpers of the gfx950 path tracer.
//
// Scene data lives in HBM in "leaf order" (the order of the LBVH leaves, i.e. Morton
// order of the Gaussian centres) so that rays that traverse neighbouring leaves touch
// neighbouring records:
//   GaussRec  64 B  rows of A = diag(1/s) R^T with mu in .w, then (sigma, kscale, id, -)
//   ShadeRec  32 B  (albedo.xyz, roughness) (metallic, shading normal.xyz)
//   PrimRec   16 B  (R s).xyz   used only by the primary-ray normal output
// A per-lane gather of one record is one 64 B segment, which is why the records are
// array-of-structures: lanes of a wave follow different BVH leaves, so a
// structure-of-arrays layout would touch 16 separate cache lines per candidate.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptir {

// ---------------------------------------------------------------- float3 math
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float a, float b, float c) { return V3{a, b, c}; }
__device__ __forceinline__ V3 v3s(float a) { return V3{a, a, a}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
__device__ __forceinline__ V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ float vmax(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ V3 xyz(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ V3 safe_normalize(V3 v) {
    const float l = dot(v, v);
    return l > 0.f ? v * rsqrtf(l) : v;
}
// Duff et al., "Building an Orthonormal Basis, Revisited" (JCGT 2017).
__device__ __forceinline__ void onb(V3 n, V3& b1, V3& b2) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    b1 = v3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = v3(b, sign + n.y * n.y * a, -n.y);
}

// ---------------------------------------------------------------- constants
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kSelfOccOffset = 2e-2f;
constexpr float kMaxSelfOcc = 1e-1f - 2e-2f;
constexpr float kSelfOccDot = 0.1f;
constexpr float kBrdfEps = 1e-6f;
constexpr float kMinRough = 0.05f;
constexpr float kRoughGradEps = 1e-3f;
constexpr float kEnvDirEps = 1e-3f;
constexpr float kInfDist = 1e20f;
constexpr float kEpsT = 1e-9f;
constexpr float kAlphaMin = 1.0f / 255.0f;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// ---------------------------------------------------------------- PCG32 (O'Neill, XSH-RR)
struct Rng {
    uint64_t state, inc;
    __device__ __forceinline__ static uint64_t launch_seed(uint32_t x, uint32_t y, uint32_t z, uint32_t f) {
        uint64_t s = 1469598103934665603ull;
        s ^= uint64_t(x) + 1099511628211ull * s;
        s ^= uint64_t(y) + 1099511628211ull * s;
        s ^= uint64_t(z) + 1099511628211ull * s;
        s ^= uint64_t(f) + 1099511628211ull * s;
        return s ? s : 1ull;
    }
    __device__ __forceinline__ void seed(uint64_t s) {
        state = 0;
        inc = 1ull;
        next_u32();
        state += s;
        next_u32();
    }
    __device__ __forceinline__ uint32_t next_u32() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    __device__ __forceinline__ float next_1d() { return float(next_u32() & 0x00FFFFFFu) * (1.0f / 16777216.0f); }
};

// ---------------------------------------------------------------- scene description
struct DeviceScene {
    const float4* grec;       // [N*4]   GaussRec, leaf order
    const float4* srec;       // [N*2]   ShadeRec, leaf order
    const float4* prec;       // [N]     PrimRec, leaf order
    const float4* nodes;      // [(N-1)*4] binary internal nodes: child boxes + child links
    const float4* wnodes;     // 4-wide nodes collapsed from them (wide.hpp)
    const float* bounds;      // [6] scene AABB lo.xyz hi.xyz (device)
    const float* sh;          // [N, sh_stride] original order
    const float4* env;        // [env_h, env_w]
    const float* tab;         // [3, tab_h, tab_w]
    const float* ray_to_world;  // 12 floats
    const float* ray_ori;     // [B,H,W,3]
    const float* ray_dir;
    uint32_t n;
    int32_t root;             // child link of the root (>=0 internal, <0 leaf ~slot), N>0
    int32_t sh_stride, sph_degree;
    int32_t env_w, env_h, tab_w, tab_h;
    int32_t tab_kind;         // 0: alias table (prob, alias, pdf); 1: CDF table (conditional, marginal, pdf)
    int32_t kernel_degree;
    float min_response, min_transmittance;
    uint32_t max_bounces, render_opts, frame;
    uint32_t cache_mode;     // 0 none, 1 end paths at the first secondary vertex with a cache query, 2 record targets
    int32_t B, H, W;
    int32_t enable_normals, enable_hitcounts, enable_mis, enable_metallic, has_snormal;
    int32_t wave_traversal;  // primary rays walk the hierarchy per wave (wave_collect)
    int32_t wide_walks;      // bit 0: wave walks use the 4-wide nodes; bit 1: per-lane walks do
    int32_t tiles_x, tiles_y, num_blocks;  // launch tiling
    // ragged batch: num_images > 0 images of their own size, table rows (H, W, first tile,
    // first path); B, H, W are then (1, 1, total paths)
    const int4* images;
    int32_t num_images;
    uint32_t rr_start;       // Russian roulette from this path vertex on (0 = off)
    unsigned long long* counters;          // work counter copies (see count()) or null
    int32_t* error_flag;
};

// ---------------------------------------------------------------- path tape
// The forward can record, per path, what the backward replay would otherwise recompute
// by re-tracing: every segment's accepted hits (leaf slot, weight alpha*T in processing
// order, cut at the flush bound of pathTracer.cuh:326-329), the segment's transmittance,
// flush bound, last hit and interaction, the NEE record and the RNG state after NEE.
// Layout (32-bit words): header[kTapeHeaderWords], then per path a record of
// kTapePathWords + kTapeSegWords * max_bounces words, then per path `cap` hit entries of
// 2 words shared by the path's segments in order (a segment record keeps its offset),
// then a pool of `pool_chunks` chunks of kTapeChunk entries: a path whose hits outgrow its
// own `cap` entries takes one chunk (atomic counter in header word kTapePoolNext) and
// continues there.  Entry i of a path is own[i] for i < cap, else chunk[i - cap].
constexpr uint32_t kTapeMagic = 0x50544150u;  // "PTAP"
constexpr uint32_t kTapeVersion = 2u;
constexpr int kTapeHeaderWords = 64;
constexpr int kTapePoolNext = 16;        // header word: chunks taken so far (zeroed before the forward)
constexpr int kTapePathWords = 16;
constexpr int kTapeSegWords = 16;
constexpr uint32_t kTapeChunk = 128;     // hit entries per pool chunk
constexpr uint32_t kTapePathsPerChunk = 64;
constexpr uint32_t kNoChunk = 0xffffffffu;
constexpr uint32_t kTapeOverflow = 1u;   // path flag: replay this path by re-tracing
constexpr uint32_t kSegIValid = 1u << 16;
constexpr uint32_t kSegValid = 1u << 17;

__host__ __device__ inline uint32_t tape_pool_chunks(uint64_t paths) {
    return uint32_t((paths + kTapePathsPerChunk - 1) / kTapePathsPerChunk);
}

struct Tape {
    uint32_t* base;     // null: no tape
    uint32_t cap;       // own hit entries per path
    uint32_t segs;      // segments per path (= effective max bounces)
    __device__ __forceinline__ uint32_t rec_words() const { return kTapePathWords + kTapeSegWords * segs; }
    __device__ __forceinline__ uint32_t* path(size_t pix) const {
        return base + kTapeHeaderWords + pix * rec_words();
    }
    __device__ __forceinline__ uint2* hits(size_t pix, size_t paths) const {
        uint2* h = reinterpret_cast<uint2*>(base + kTapeHeaderWords + paths * rec_words());
        return h + pix * cap;
    }
    __device__ __forceinline__ uint2* pool(size_t paths) const {
        return reinterpret_cast<uint2*>(base + kTapeHeaderWords + paths * rec_words()) + paths * cap;
    }
    // entry i of a path whose own entries start at `own` (chunk: its pool chunk or kNoChunk)
    __device__ __forceinline__ uint2* entry(uint2* own, uint2* pl, uint32_t chunk, uint32_t i) const {
        return i < cap ? own + i : pl + size_t(chunk) * kTapeChunk + (i - cap);
    }
};

struct FwdOut {
    float *rad, *dns, *hit, *hit2, *dist, *nrm, *snrm, *mat, *hits, *pbr, *light, *comp;
    uint32_t* vis;  // [N] original order
    Tape tape;
    float* cache;   // [B,H,W,20] radiance-cache records (null: none)
};

struct BwdIO {
    const float *pbr, *light, *pbr_grad, *light_grad, *mat_grad;
    float* dmat_leaf;  // [N,5] leaf order (scattered to the original order afterwards)
    float* denv;       // [env_h, env_w, 4]
    Tape tape;
    int32_t* tape_state;  // [1]: 0 = tape valid, 1 = header mismatch (replay every path)
    int32_t replay_mode;  // 0: all paths, 1: only paths the tape kernel could not handle
};

// Exact adjoint: incoming gradients of the forward outputs (null = zero), the trace's particle
// density, the parameter gradients (zeroed by the caller, original order) and the per-path
// vertex records.
struct AdjIO {
    const float *pbr_grad, *light_grad, *mat_grad;                        // [..,3], [..,3], [..,5]
    const float *rad_grad, *dns_grad, *hit_grad, *snrm_grad, *comp_grad;  // [..,3], [..,1], [..,2], [..,3], [..,6]
    const float* density;                                                 // [N,12]
    float *d_density, *d_material, *d_sh, *d_snormal, *d_env;
    float4* scratch;
};

// 3DGRT primary-only tracer (grt_path.hpp): forward outputs, and the backward's forward
// outputs, output gradients, raw Gaussian parameters and gradient sinks.
namespace grt {
// Hit tape of the 3DGRT forward: per pixel, the accepted hits its backward processes, in
// order, so that the backward need not walk the hierarchy again.  Words: a header of
// kTapeHeader words, the per-pixel counts [P] (bit 31: the backward re-traces the pixel), the
// leaf slots [cap * P] and the distances [cap * P] (entry j of pixel p at j * P + p).
constexpr uint32_t kTapeHeader = 16;
constexpr uint32_t kTapeMagic = 0x54545247u;  // "GRTT"
constexpr uint32_t kTapeRetrace = 0x80000000u;
struct TapeView {
    uint32_t* base;
    uint32_t cap;
    size_t P;
    __host__ __device__ uint32_t* header() const { return base; }
    __host__ __device__ uint32_t* count() const { return base + kTapeHeader; }
    __host__ __device__ uint32_t* slot() const { return base + kTapeHeader + P; }
    __host__ __device__ float* dist() const { return reinterpret_cast<float*>(base + kTapeHeader + P + size_t(cap) * P); }
};
__host__ __device__ inline uint64_t tape_bytes(uint64_t P, uint32_t cap) {
    return 4ull * (kTapeHeader + P + 2ull * cap * P);
}
struct FwdOut {
    float *rad, *dns, *hit, *hit2, *dist, *nrm, *snrm, *hits;  // [P,3], [P], [P,2], [P], [P], [P,3], [P,3], [P]
    uint32_t* vis;                                              // [N]
    uint32_t* tape;                                             // hit tape, or null
    uint32_t tape_cap;                                          // hits per pixel it holds
};
struct BwdIO {
    const float *rad, *dns, *hit, *hit2, *dist, *snrm;                               // forward outputs
    const float *rad_grad, *dns_grad, *hit_grad, *hit2_grad, *dist_grad, *snrm_grad;  // hit_grad: [P] (depth)
    const float* density;  // [N,12] position, density, quaternion (w,x,y,z), scale, pad
    const float* snormal;  // [N,3] or null
    float *d_density, *d_sh, *d_snormal;  // [N,12], [N, sh_stride], [N,3] (zeroed by the caller)
    // Hit records (null: every hit adds its gradients by atomics).  Record j of pixel pix is
    // slot j * paths + pix: rec[4 * slot .. 4 * slot + 3] and its Gaussian rec_key[slot];
    // rec_cnt[pix] counts the pixel's records (at most kRecPerPixel).
    float4* rec;
    uint32_t* rec_key;
    uint32_t* rec_cnt;
    uint32_t paths;
    const uint32_t* tape;  // the forward's hit tape, or null (every pixel re-traces)
    uint32_t tape_cap;
};
// Records per pixel of the 3DGRT backward; hits beyond them add their gradients by atomics.
constexpr uint32_t kRecPerPixel = 32;
}  // namespace grt

enum CounterId { C_PATHS = 0, C_SEGMENTS, C_CANDIDATES, C_ACCEPTED, C_ACCEPTED_PRIMARY, C_SHADOW_CANDIDATES,
                 C_ENV_LOOKUPS, C_MAT_ATOMICS, C_ENV_ATOMICS, C_TRAVERSALS, C_NODES, C_LEAF_TESTS,
                 C_COUNT };

// Work counters live in kCounterCopies copies of kCounterStride entries (one 128-byte
// line each); a block adds into copy blockIdx % kCounterCopies, and the host sums the
// copies.  One work unit per calling lane, summed over the wave: one atomic per wave, and
// spread over many lines, so a counting launch runs about as fast as a plain one.
constexpr int kCounterStride = 16;
constexpr int kCounterCopies = 512;
static_assert(C_COUNT <= kCounterStride, "counter copy too small");

__device__ __forceinline__ void count(const DeviceScene& S, int id) {
    if (S.counters) {
        const uint64_t m = __ballot(1);
        if (__lane_id() == __ffsll((long long)m) - 1)
            atomicAdd(S.counters + (blockIdx.x % kCounterCopies) * kCounterStride + id, (unsigned long long)__popcll(m));
    }
}

// generalized Gaussian response exp(-4.5 (r/3)^(2n)) of squared distance d2 = r^2
__device__ __forceinline__ float response(int degree, float d2) {
    switch (degree) {
    case 8: { const float q = d2 * d2; return expf(-0.000685871056241f * q * q); }
    case 5: return expf(-0.0185185185185f * d2 * d2 * sqrtf(d2));
    case 4: return expf(-0.0555555555556f * d2 * d2);
    case 3: return expf(-0.166666666667f * d2 * sqrtf(d2));
    case 1: return expf(-1.5f * sqrtf(d2));
    case 0: return fmaxf(1.f - 0.329630334487f * sqrtf(d2), 0.f);
    default: return expf(-0.5f * d2);
    }
}

}  // namespace ptir
