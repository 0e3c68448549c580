// LBVH over the Gaussians' instance boxes, built on the GPU.
//
// The reference bounds each Gaussian by the unit cube under M = R diag(kscl)
// (threedgptir_tracer/src/particlePrimitives.cu:543-610, kernelScale :27-51) and lets
// OptiX build an instance BVH over them (src/optixTracer.cpp:816-904).  Here:
//   1. k_instances: world AABB of every instance cube and the scene box (order-
//      preserving integer atomics);
//   2. k_morton + radix sort: 63-bit Morton codes of the centres;
//   3. k_karras: binary radix tree (Karras, HPG 2012), duplicate codes split by index;
//   4. k_refit: bottom-up box fit with one arrival counter per node (or PLOC, ploc.hpp,
//      in place of 3 and 4);
//   5. the 4-wide hierarchy the traversal walks (wide.hpp): even-depth nodes flagged and
//      numbered by a prefix sum, filled from their grandchildren; a refit only refills;
//   6. k_wide_need, k_depth: traversal stack entries the 4-wide and the binary walks need,
//      checked against the traversal stack.
// Binary node layout (64 B): {lo_l.xyz, link_l} {hi_l.xyz, link_r} {lo_r.xyz, -} {hi_r.xyz, -};
// link >= 0 is an internal node, link < 0 is leaf slot ~link.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "internal.hpp"

namespace ptir {

__device__ __forceinline__ int32_t f2o(float f) {
    const int32_t b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float o2f(int32_t o) { return __int_as_float(o >= 0 ? o : o ^ 0x7FFFFFFF); }

// kernelScale of particlePrimitives.cu:27-51
__device__ __forceinline__ float kernel_scale(float density, const KernelOpts& o) {
    const float modulation = o.density_clamping ? density : 1.0f;
    const float minR = fminf(o.min_response / modulation, 0.97f);
    const float deg = float(o.kernel_degree);
    if (o.kernel_degree == 0) return ((1.0f - minR) / 3.0f) / -0.329630334487f;
    const float a = -4.5f / powf(3.0f, deg);
    return powf(logf(minR) / a, 1.0f / deg);
}

// standard rotation matrix (column-vector convention) of a (w,x,y,z) quaternion
__device__ __forceinline__ void quat_to_rot(float4 q, float m[3][3]) {
    const float w = q.x, x = q.y, y = q.z, z = q.w;
    m[0][0] = 1.f - 2.f * (y * y + z * z); m[0][1] = 2.f * (x * y - w * z); m[0][2] = 2.f * (x * z + w * y);
    m[1][0] = 2.f * (x * y + w * z); m[1][1] = 1.f - 2.f * (x * x + z * z); m[1][2] = 2.f * (y * z - w * x);
    m[2][0] = 2.f * (x * z - w * y); m[2][1] = 2.f * (y * z + w * x); m[2][2] = 1.f - 2.f * (x * x + y * y);
}

__global__ void k_init_bounds(int32_t* bi) {
    if (threadIdx.x < 3) bi[threadIdx.x] = f2o(__int_as_float(0x7f800000));
    else if (threadIdx.x < 6) bi[threadIdx.x] = f2o(-__int_as_float(0x7f800000));
}

__global__ void __launch_bounds__(256) k_instances(const float* __restrict__ pos, const float* __restrict__ rot,
                                                   const float* __restrict__ scl, const float* __restrict__ dns,
                                                   uint32_t n, KernelOpts o, float4* __restrict__ aabb,
                                                   int32_t* __restrict__ bounds_i) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float lo[3] = {__int_as_float(0x7f800000), __int_as_float(0x7f800000), __int_as_float(0x7f800000)};
    float hi[3] = {-lo[0], -lo[1], -lo[2]};
    if (i < n) {
        float m[3][3];
        quat_to_rot(make_float4(rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]), m);
        const float ks = kernel_scale(dns[i], o);
        const float k[3] = {ks * scl[3 * i], ks * scl[3 * i + 1], ks * scl[3 * i + 2]};
        float bl[3], bh[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const float mu = pos[3 * i + r];
            const float ext = fabsf(m[r][0]) * k[0] + fabsf(m[r][1]) * k[1] + fabsf(m[r][2]) * k[2];
            // conservative pad: the traversal slab test must never cull a peak on the face
            const float pad = 2e-6f * (fabsf(mu) + ext) + 1e-7f;
            bl[r] = mu - ext - pad;
            bh[r] = mu + ext + pad;
            lo[r] = bl[r];
            hi[r] = bh[r];
        }
        aabb[2 * i] = make_float4(bl[0], bl[1], bl[2], 0.f);
        aabb[2 * i + 1] = make_float4(bh[0], bh[1], bh[2], 0.f);
    }
    // wave reduction then one atomic per wave
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[r] = fminf(lo[r], __shfl_xor(lo[r], off));
            hi[r] = fmaxf(hi[r], __shfl_xor(hi[r], off));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            atomicMin(bounds_i + r, f2o(lo[r]));
            atomicMax(bounds_i + 3 + r, f2o(hi[r]));
        }
    }
}

__device__ __forceinline__ unsigned long long spread21(unsigned long long v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ void __launch_bounds__(256) k_morton(const float4* __restrict__ aabb, uint32_t n,
                                                const int32_t* __restrict__ bounds_i, float* __restrict__ bounds,
                                                unsigned long long* __restrict__ keys, uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float lo[3] = {o2f(bounds_i[0]), o2f(bounds_i[1]), o2f(bounds_i[2])};
    const float hi[3] = {o2f(bounds_i[3]), o2f(bounds_i[4]), o2f(bounds_i[5])};
    if (i == 0) {
        for (int r = 0; r < 3; ++r) {
            bounds[r] = lo[r];
            bounds[3 + r] = hi[r];
        }
    }
    if (i >= n) return;
    const float4 a = aabb[2 * i], b = aabb[2 * i + 1];
    const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
    unsigned long long q[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const float ext = hi[r] - lo[r];
        float u = ext > 0.f ? (c[r] - lo[r]) / ext : 0.5f;
        u = fminf(fmaxf(u, 0.f), 1.f);
        q[r] = (unsigned long long)fminf(u * 2097152.f, 2097151.f);
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    idx[i] = i;
}

__device__ __forceinline__ int delta(const unsigned long long* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const unsigned long long a = k[i], b = k[j];
    if (a == b) return 64 + __clz(uint32_t(i ^ j));
    return __clzll(a ^ b);
}

__global__ void __launch_bounds__(256) k_karras(const unsigned long long* __restrict__ k, int n,
                                                float4* __restrict__ nodes, int32_t* __restrict__ parent,
                                                uint32_t* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(k, n, i, i - d);
    int lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(k, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == gamma) ? ~gamma : gamma;
    const int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    float4* nd = nodes + 4 * i;
    nd[0].w = __int_as_float(left);
    nd[1].w = __int_as_float(right);
    parent[left >= 0 ? left : (n - 1) + ~left] = i;
    parent[right >= 0 ? right : (n - 1) + ~right] = i;
    flags[i] = 0u;
    if (i == 0) parent[0] = -1;
}

__global__ void __launch_bounds__(256) k_refit(const float4* __restrict__ aabb, const uint32_t* __restrict__ leaf_id,
                                               int n, float4* nodes, const int32_t* __restrict__ parent,
                                               uint32_t* flags) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t g = leaf_id[j];
    float4 lo = aabb[2 * g], hi = aabb[2 * g + 1];
    int child = ~j;
    int p = parent[(n - 1) + j];
    while (p >= 0) {
        float4* nd = nodes + 4 * p;
        const bool is_left = __float_as_int(nd[0].w) == child;
        float4* dlo = is_left ? &nd[0] : &nd[2];
        float4* dhi = is_left ? &nd[1] : &nd[3];
        dlo->x = lo.x; dlo->y = lo.y; dlo->z = lo.z;
        dhi->x = hi.x; dhi->y = hi.y; dhi->z = hi.z;
        __threadfence();
        if (atomicAdd(&flags[p], 1u) == 0u) return;  // first arrival: the sibling finishes
        __threadfence();
        const float4 a0 = nd[0], a1 = nd[1];
        const float4 b0 = nd[2], b1 = nd[3];
        lo = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.f);
        hi = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.f);
        child = p;
        p = parent[p];
    }
}

// ---------------------------------------------------------------- 4-wide hierarchy (wide.hpp)
__global__ void __launch_bounds__(256) k_wide_flag(const int32_t* __restrict__ parent, int m, uint32_t* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) flag[i] = wide::is_wide(parent, i);
}
__global__ void __launch_bounds__(256) k_wide_number(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pre,
                                                     int m, int32_t* __restrict__ wmap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) wide::number(flag, pre, i, wmap);
}
__global__ void __launch_bounds__(256) k_wide_fill(const float4* __restrict__ bin, const int32_t* __restrict__ wmap, int m,
                                                   float4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) wide::fill(bin, wmap, i, out);
}
// Depth of the deepest leaf of the binary tree; the walk stops at wide::kDepthCap, which the
// stack check then rejects (a malformed parent chain cannot loop).
__global__ void __launch_bounds__(256) k_depth(const int32_t* __restrict__ parent, int n, int32_t* depth) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int d = 0;
    for (int p = parent[(n - 1) + j]; p >= 0 && d < wide::kDepthCap; p = parent[p]) ++d;
    atomicMax(depth, d);
}
__global__ void k_wide_count(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pre, int m, int32_t* count) {
    *count = int32_t(pre[m - 1] + flag[m - 1]);
}
__global__ void __launch_bounds__(256) k_wide_need(const int32_t* __restrict__ parent, const int32_t* __restrict__ wmap,
                                                   const float4* __restrict__ out, int m, int32_t* need) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) atomicMax(need, wide::stack_need(parent, wmap, out, i));
}

// ---------------------------------------------------------------- PLOC builder (ploc.hpp)
__global__ void __launch_bounds__(256) k_ploc_init(const float4* __restrict__ aabb, const uint32_t* __restrict__ leaf_id,
                                                   int n, int32_t* __restrict__ c, float4* __restrict__ b,
                                                   ploc::State* __restrict__ st, int32_t* __restrict__ parent) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t g = leaf_id[j];
    c[j] = ~j;
    b[2 * j] = aabb[2 * g];
    b[2 * j + 1] = aabb[2 * g + 1];
    if (j == 0) {
        st[0] = ploc::State{n, n - 2};
        parent[0] = -1;  // node 0 is the last merge, the root
    }
}
__global__ void __launch_bounds__(256) k_ploc_nearest(ploc::Iter it) {
    ploc::nearest(it, int(blockIdx.x * blockDim.x + threadIdx.x));
}
__global__ void __launch_bounds__(256) k_ploc_flags(ploc::Iter it) {
    ploc::flags(it, int(blockIdx.x * blockDim.x + threadIdx.x));
}
__global__ void __launch_bounds__(256) k_ploc_compact(ploc::Iter it) {
    ploc::compact(it, int(blockIdx.x * blockDim.x + threadIdx.x));
}
// Structure check of a finished tree, one thread per element (n-1 nodes, then n leaves):
// every internal child of node k has an index above k (made earlier), so the child links have
// no cycle, and parent and child links agree.  Any violation sets *bad.
__global__ void __launch_bounds__(256) k_ploc_check(const float4* __restrict__ nodes,
                                                    const int32_t* __restrict__ parent, int n,
                                                    int32_t* __restrict__ bad) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= 2 * n - 1) return;
    bool ok = true;
    if (x < n - 1) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int ch = __float_as_int(nodes[4 * size_t(x) + s].w);
            if (ch >= 0) ok = ok && ch > x && ch < n - 1 && parent[ch] == x;
            else ok = ok && ~ch < n && parent[(n - 1) + ~ch] == x;
        }
    }
    if (x != 0) {
        const int p = parent[x];
        const int self = x < n - 1 ? x : ~(x - (n - 1));
        ok = ok && p >= 0 && p < n - 1 &&
             (__float_as_int(nodes[4 * size_t(p)].w) == self || __float_as_int(nodes[4 * size_t(p) + 1].w) == self);
    } else {
        ok = ok && parent[0] == -1;
    }
    if (!ok) atomicOr(bad, 1);
}

__global__ void k_clear_flags(uint32_t* flags, int m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) flags[i] = 0u;
}

// Per-trace records in leaf order from the trace's own particle tensors
// (processHit reads particleDensity/particleMaterial/shadingNormal, gaussianParticles.cuh:343-440).
__global__ void __launch_bounds__(256) k_records(const float* __restrict__ density, const float* __restrict__ material,
                                                 const float* __restrict__ snormal,
                                                 const uint32_t* __restrict__ leaf_id, uint32_t n, KernelOpts o,
                                                 float4* __restrict__ grec, float4* __restrict__ srec,
                                                 float4* __restrict__ prec) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = leaf_id[j];
    const float* p = density + 12 * size_t(i);
    float m[3][3];
    quat_to_rot(make_float4(p[4], p[5], p[6], p[7]), m);
    const float s[3] = {p[8], p[9], p[10]};
    const float is[3] = {1.0f / s[0], 1.0f / s[1], 1.0f / s[2]};
    // rows of A = diag(1/s) R^T: row r is column r of R scaled by 1/s_r
    grec[4 * j + 0] = make_float4(m[0][0] * is[0], m[1][0] * is[0], m[2][0] * is[0], p[0]);
    grec[4 * j + 1] = make_float4(m[0][1] * is[1], m[1][1] * is[1], m[2][1] * is[1], p[1]);
    grec[4 * j + 2] = make_float4(m[0][2] * is[2], m[1][2] * is[2], m[2][2] * is[2], p[2]);
    grec[4 * j + 3] = make_float4(p[3], kernel_scale(p[3], o), __int_as_float(int(i)), 0.f);
    // material may be null (the primary-only tracer has none)
    const float* mt = material ? material + 5 * size_t(i) : nullptr;
    const float sn[3] = {snormal ? snormal[3 * i] : 0.f, snormal ? snormal[3 * i + 1] : 0.f,
                         snormal ? snormal[3 * i + 2] : 0.f};
    srec[2 * j + 0] = mt ? make_float4(mt[0], mt[1], mt[2], mt[3]) : make_float4(0.f, 0.f, 0.f, 0.f);
    srec[2 * j + 1] = make_float4(mt ? mt[4] : 0.f, sn[0], sn[1], sn[2]);
    prec[j] = make_float4(m[0][0] * s[0] + m[0][1] * s[1] + m[0][2] * s[2],
                          m[1][0] * s[0] + m[1][1] * s[1] + m[1][2] * s[2],
                          m[2][0] * s[0] + m[2][1] * s[1] + m[2][2] * s[2], snormal ? 1.f : 0.f);
}

__global__ void __launch_bounds__(256) k_scatter_mat(const float* __restrict__ leaf, const uint32_t* __restrict__ leaf_id,
                                                     uint32_t n, float* __restrict__ out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = leaf_id[j];
#pragma unroll
    for (int c = 0; c < 5; ++c) out[5 * size_t(i) + c] = leaf[5 * size_t(j) + c];
}

__global__ void __launch_bounds__(256) k_fill_zero(uint32_t* __restrict__ p, size_t words) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += size_t(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

hipError_t fill_zero(void* p, size_t bytes, hipStream_t s) {
    if (!bytes) return hipSuccess;
    const size_t words = bytes / 4;  // every buffer zeroed here is a float/int array
    const unsigned grid = unsigned(std::min<size_t>((words + 255) / 256, 8192));
    k_fill_zero<<<grid, 256, 0, s>>>(static_cast<uint32_t*>(p), words);
    return hipGetLastError();
}

template <class T>
static hipError_t grow(T*& p, size_t count) {
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        if (e != hipSuccess) return e;
    }
    return hipMalloc((void**)&p, count * sizeof(T) + 256);
}

hipError_t bvh_reserve(BvhBuffers& b, uint32_t n) {
    if (n <= b.cap && b.aabb) return hipSuccess;
    hipError_t e;
    const size_t N = n, M = n > 1 ? n - 1 : 1;
#define G(ptr, cnt) if ((e = grow(ptr, cnt)) != hipSuccess) return e;
    G(b.aabb, 2 * N);
    G(b.keys, N);
    G(b.keys_sorted, N);
    G(b.idx, N);
    G(b.leaf_id, N);
    G(b.nodes, 4 * M);
    G(b.wide, size_t(wide::kWords) * M);
    G(b.wflag, M);
    G(b.wpre, M);
    G(b.wmap, M);
    G(b.parent, 2 * N);
    G(b.flags, M);
    G(b.grec, 4 * N);
    G(b.srec, 2 * N);
    G(b.prec, N);
    G(b.dmat_leaf, 5 * N);
#undef G
    if (!b.bounds_i) {
        if ((e = hipMalloc((void**)&b.bounds_i, 6 * sizeof(int32_t))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&b.bounds, 6 * sizeof(float))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&b.wneed, 3 * sizeof(int32_t))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&b.counters, size_t(kCounterCopies) * kCounterStride * sizeof(unsigned long long))) !=
            hipSuccess)
            return e;
        if ((e = hipMalloc((void**)&b.error_flag, sizeof(int32_t))) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&b.tape_state, sizeof(int32_t))) != hipSuccess) return e;
    }
    size_t tmp = 0;
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, b.keys, b.keys_sorted, b.idx, b.leaf_id, int(n), 0, 63);
    if (e != hipSuccess) return e;
    if (tmp > b.sort_tmp_bytes) {
        if (b.sort_tmp) (void)hipFree(b.sort_tmp);
        b.sort_tmp = nullptr;
        if ((e = hipMalloc(&b.sort_tmp, tmp)) != hipSuccess) return e;
        b.sort_tmp_bytes = tmp;
    }
    tmp = 0;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, b.wflag, b.wpre, int(M));
    if (e != hipSuccess) return e;
    if (tmp > b.wscan_tmp_bytes) {
        if (b.wscan_tmp) (void)hipFree(b.wscan_tmp);
        b.wscan_tmp = nullptr;
        if ((e = hipMalloc(&b.wscan_tmp, tmp + 256)) != hipSuccess) return e;
        b.wscan_tmp_bytes = tmp + 256;
    }
    b.cap = n;
    return hipSuccess;
}

void bvh_free(BvhBuffers& b) {
    void* ptrs[] = {b.aabb, b.keys, b.keys_sorted, b.idx, b.leaf_id, b.nodes, b.parent, b.flags, b.bounds_i,
                    b.bounds, b.sort_tmp, b.grec, b.srec, b.prec, b.dmat_leaf, b.counters, b.error_flag,
                    b.tape_state, b.pc[0], b.pc[1], b.pb[0], b.pb[1], b.pnn, b.pval, b.pscan, b.pstate, b.pbad,
                    b.scan_tmp, b.wide, b.wflag, b.wpre, b.wmap, b.wneed, b.wscan_tmp};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    b = BvhBuffers();
}

static inline unsigned blocks(uint32_t n, unsigned bs) { return (n + bs - 1) / bs; }

// PLOC over the sorted leaves: nodes, boxes and parent links of the whole tree.  The host
// reads the cluster count back every kPlocBatch iterations (a rebuild synchronises anyway, for
// the depth check).  Returns hipSuccess with *converged = false when kPlocMaxIter iterations
// did not finish or the finished tree fails k_ploc_check (the caller then builds the LBVH).
constexpr int kPlocBatch = 8;
constexpr int kPlocMaxIter = 4096;

static hipError_t ploc_reserve(BvhBuffers& b, uint32_t n) {
    if (n <= b.pcap && b.pc[0]) return hipSuccess;
    hipError_t e;
    const size_t N = n;
#define G(ptr, cnt) if ((e = grow(ptr, cnt)) != hipSuccess) return e;
    G(b.pc[0], N);
    G(b.pc[1], N);
    G(b.pb[0], 2 * N);
    G(b.pb[1], 2 * N);
    G(b.pnn, N);
    G(b.pval, N);
    G(b.pscan, N);
    G(b.pstate, 2);
    G(b.pbad, 1);
#undef G
    size_t tmp = 0;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, b.pval, b.pscan, int(n))) != hipSuccess) return e;
    if (b.scan_tmp) (void)hipFree(b.scan_tmp);
    b.scan_tmp = nullptr;
    if ((e = hipMalloc(&b.scan_tmp, tmp + 256)) != hipSuccess) return e;
    b.scan_tmp_bytes = tmp + 256;
    b.pcap = n;
    return hipSuccess;
}

static hipError_t ploc_build(BvhBuffers& b, uint32_t n, hipStream_t s, bool* converged) {
    *converged = false;
    hipError_t e = ploc_reserve(b, n);
    if (e != hipSuccess) return e;
    k_ploc_init<<<blocks(n, 256), 256, 0, s>>>(b.aabb, b.leaf_id, int(n), b.pc[0], b.pb[0], b.pstate, b.parent);
    int known = int(n);  // an upper bound of the cluster count (it only decreases)
    for (int it = 0; it < kPlocMaxIter;) {
        for (int k = 0; k < kPlocBatch; ++k, ++it) {
            const int a = it & 1, z = a ^ 1;
            const ploc::Iter I{b.pc[a], b.pb[a], b.pc[z], b.pb[z], b.pnn, b.pval, b.pscan, b.pstate + a, b.pstate + z,
                               b.nodes, b.parent, int(n)};
            const unsigned grid = blocks(uint32_t(known), 256);
            k_ploc_nearest<<<grid, 256, 0, s>>>(I);
            k_ploc_flags<<<grid, 256, 0, s>>>(I);
            size_t tmp = b.scan_tmp_bytes;
            if ((e = hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, tmp, b.pval, b.pscan, known, s)) != hipSuccess)
                return e;
            k_ploc_compact<<<grid, 256, 0, s>>>(I);
        }
        ploc::State h{0, 0};
        if ((e = hipMemcpyAsync(&h, b.pstate + (it & 1), sizeof(h), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        if (h.cnt <= 1) {
            if (h.cnt != 1 || h.base != -1) return hipSuccess;
            int32_t bad = 1;
            if ((e = fill_zero(b.pbad, sizeof(int32_t), s)) != hipSuccess) return e;
            k_ploc_check<<<blocks(2 * n - 1, 256), 256, 0, s>>>(b.nodes, b.parent, int(n), b.pbad);
            if ((e = hipMemcpyAsync(&bad, b.pbad, sizeof(bad), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
            if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
            *converged = bad == 0;
            return hipSuccess;
        }
        known = h.cnt;
    }
    return hipSuccess;
}

hipError_t bvh_build(BvhBuffers& b, const KernelOpts& o, const float* pos, const float* rot, const float* scl,
                     const float* dns, uint32_t n, bool rebuild, int builder, hipStream_t s) {
    hipError_t e = bvh_reserve(b, n);
    if (e != hipSuccess) return e;
    b.n = n;
    k_init_bounds<<<1, 64, 0, s>>>(b.bounds_i);
    k_instances<<<blocks(n, 256), 256, 0, s>>>(pos, rot, scl, dns, n, o, b.aabb, b.bounds_i);
    k_morton<<<blocks(n, 256), 256, 0, s>>>(b.aabb, n, b.bounds_i, b.bounds, b.keys, b.idx);
    if (rebuild) {
        e = hipcub::DeviceRadixSort::SortPairs(b.sort_tmp, b.sort_tmp_bytes, b.keys, b.keys_sorted, b.idx, b.leaf_id,
                                               int(n), 0, 63, s);
        if (e != hipSuccess) return e;
        bool done = false;
        if (n > 1 && builder == 1 && (e = ploc_build(b, n, s, &done)) != hipSuccess) return e;
        // PLOC writes nodes, boxes and parents; otherwise the radix tree, then its boxes
        if (n > 1 && !done) {
            k_karras<<<blocks(n - 1, 256), 256, 0, s>>>(b.keys_sorted, int(n), b.nodes, b.parent, b.flags);
            k_refit<<<blocks(n, 256), 256, 0, s>>>(b.aabb, b.leaf_id, int(n), b.nodes, b.parent, b.flags);
        }
    } else if (n > 1) {
        k_clear_flags<<<blocks(n - 1, 256), 256, 0, s>>>(b.flags, int(n - 1));
        k_refit<<<blocks(n, 256), 256, 0, s>>>(b.aabb, b.leaf_id, int(n), b.nodes, b.parent, b.flags);
    }
    if (n > 1) {
        const int m = int(n - 1);
        if (rebuild) {
            k_wide_flag<<<blocks(m, 256), 256, 0, s>>>(b.parent, m, b.wflag);
            size_t tmp = b.wscan_tmp_bytes;
            if ((e = hipcub::DeviceScan::ExclusiveSum(b.wscan_tmp, tmp, b.wflag, b.wpre, m, s)) != hipSuccess) return e;
            k_wide_number<<<blocks(m, 256), 256, 0, s>>>(b.wflag, b.wpre, m, b.wmap);
        }
        k_wide_fill<<<blocks(m, 256), 256, 0, s>>>(b.nodes, b.wmap, m, b.wide);
    }
    return hipGetLastError();
}

hipError_t bvh_stack_need(BvhBuffers& b, hipStream_t s, int32_t* host_need, int32_t* host_count, int32_t* host_depth) {
    hipError_t e = fill_zero(b.wneed, 3 * sizeof(int32_t), s);
    if (e != hipSuccess) return e;
    if (b.n > 1) {
        k_wide_need<<<blocks(b.n - 1, 256), 256, 0, s>>>(b.parent, b.wmap, b.wide, int(b.n - 1), b.wneed);
        k_wide_count<<<1, 1, 0, s>>>(b.wflag, b.wpre, int(b.n - 1), b.wneed + 1);
        k_depth<<<blocks(b.n, 256), 256, 0, s>>>(b.parent, int(b.n), b.wneed + 2);
    }
    int32_t h[3] = {0, 0, 0};
    e = hipMemcpyAsync(h, b.wneed, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(s);
    *host_need = h[0];
    *host_count = h[1];
    *host_depth = h[2];
    return e;
}

hipError_t prep_records(BvhBuffers& b, const KernelOpts& o, const float* density, const float* material,
                        const float* snormal, hipStream_t s) {
    if (b.n == 0) return hipSuccess;
    k_records<<<blocks(b.n, 256), 256, 0, s>>>(density, material, snormal, b.leaf_id, b.n, o, b.grec, b.srec, b.prec);
    return hipGetLastError();
}

hipError_t scatter_material_grad(const BvhBuffers& b, float* dmat, hipStream_t s) {
    if (b.n == 0) return hipSuccess;
    k_scatter_mat<<<blocks(b.n, 256), 256, 0, s>>>(b.dmat_leaf, b.leaf_id, b.n, dmat);
    return hipGetLastError();
}

}  // namespace ptir
