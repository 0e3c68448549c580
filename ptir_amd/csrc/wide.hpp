// Four-wide hierarchy collapsed from the binary one (LBVH or PLOC), and its traversal node
// test.  Every binary internal node at even depth becomes a 4-wide node whose children are
// its grandchildren (or its children where those are leaves), so a walk makes half the
// dependent node fetches of the binary walk and tests four boxes per fetch.
//
// Node layout (8 float4 = 128 B, structure of arrays over the four children):
//   [0] lo.x  [1] lo.y  [2] lo.z  [3] hi.x  [4] hi.y  [5] hi.z   (child c in component c)
//   [6] child links (int bits): >= 0 a 4-wide node, < 0 leaf slot ~link, kEmpty unused
//   [7] (number of 4-wide children, 0, 0, 0)
// An unused child has the box of the point (kFar, kFar, kFar), which no ray inside the scene
// box enters before its end, and the link kEmpty.
//
// The build phases are functions of one binary node index each, run by the kernels of
// bvh.hip and, on the host, by tests/host_emu (the same code).
#pragma once

#include "device.hpp"

namespace ptir {
namespace wide {

constexpr int kWords = 8;
constexpr int32_t kEmpty = 0x7fffffff;
constexpr float kFar = 1e30f;
constexpr int kDepthCap = 4096;  // a malformed parent chain cannot loop

// binary internal node i: 1 when it is a 4-wide node (even depth; the root has depth 0)
__host__ __device__ inline uint32_t is_wide(const int32_t* parent, int i) {
    int d = 0;
    for (int p = parent[i]; p >= 0 && d < kDepthCap; p = parent[p]) ++d;
    return (d & 1) == 0 ? 1u : 0u;
}

// the 4-wide node of binary node i from its flag and the exclusive prefix sum of the flags
__host__ __device__ inline void number(const uint32_t* flag, const uint32_t* pre, int i, int32_t* wmap) {
    wmap[i] = flag[i] ? int32_t(pre[i]) : -1;
}

// Writes the 4-wide node of binary node i (nothing when i is not one).  Binary node layout:
// {lo_l.xyz, link_l} {hi_l.xyz, link_r} {lo_r.xyz, -} {hi_r.xyz, -}.
__host__ __device__ inline void fill(const float4* bin, const int32_t* wmap, int i, float4* out) {
    const int32_t w = wmap[i];
    if (w < 0) return;
    float lo[3][4], hi[3][4];
    int32_t link[4];
    int c = 0, inner = 0;
    auto add = [&](float4 a, float4 b, int32_t l) {
        lo[0][c] = a.x;
        lo[1][c] = a.y;
        lo[2][c] = a.z;
        hi[0][c] = b.x;
        hi[1][c] = b.y;
        hi[2][c] = b.z;
        link[c] = l;
        inner += l >= 0 ? 1 : 0;
        ++c;
    };
    const float4* nd = bin + 4 * size_t(i);
    for (int s = 0; s < 2; ++s) {
        const int32_t ch = __float_as_int(s == 0 ? nd[0].w : nd[1].w);
        if (ch < 0) {
            add(nd[2 * s], nd[2 * s + 1], ch);
        } else {
            const float4* cd = bin + 4 * size_t(ch);
            for (int t = 0; t < 2; ++t) {
                const int32_t g = __float_as_int(t == 0 ? cd[0].w : cd[1].w);
                add(cd[2 * t], cd[2 * t + 1], g < 0 ? g : wmap[g]);
            }
        }
    }
    for (; c < 4; ++c) {
        for (int r = 0; r < 3; ++r) lo[r][c] = hi[r][c] = kFar;
        link[c] = kEmpty;
    }
    float4* o = out + size_t(kWords) * w;
    for (int r = 0; r < 3; ++r) {
        o[r] = make_float4(lo[r][0], lo[r][1], lo[r][2], lo[r][3]);
        o[3 + r] = make_float4(hi[r][0], hi[r][1], hi[r][2], hi[r][3]);
    }
    o[6] = make_float4(__int_as_float(link[0]), __int_as_float(link[1]), __int_as_float(link[2]),
                       __int_as_float(link[3]));
    o[7] = make_float4(__int_as_float(inner), 0.f, 0.f, 0.f);
}

// Traversal stack entries a walk can hold while at the 4-wide node of binary node i: at every
// 4-wide ancestor it pushed at most (its 4-wide children - 1) entries, and at the node itself
// it pushes at most that many again.  The largest value over all nodes bounds the stack.
__host__ __device__ inline int stack_need(const int32_t* parent, const int32_t* wmap, const float4* out, int i) {
    const int32_t w = wmap[i];
    if (w < 0) return 0;
    auto inner = [&](int32_t v) { return __float_as_int(out[size_t(kWords) * v + 7].x); };
    int need = inner(w) > 1 ? inner(w) - 1 : 0;
    int d = 0;
    for (int p = parent[i]; p >= 0 && d < kDepthCap; p = parent[p], ++d)
        if (wmap[p] >= 0) need += inner(wmap[p]) - 1;
    return need;
}

// A node's boxes and links, loaded together (one wait for the whole record).
struct Node4 {
    float4 lx, ly, lz, hx, hy, hz, links;
};
__device__ __forceinline__ Node4 load4(const float4* nd) {
    return Node4{nd[0], nd[1], nd[2], nd[3], nd[4], nd[5], nd[6]};
}

// Slab test of the four children of a node against a ray (inverse direction inv, finite: see
// slab_inv) in
// (tmin, cut): bit c of the result is set when the ray enters child c's box before cut;
// tn[c] is the entry distance.
__device__ __forceinline__ uint32_t test4(const Node4& nd, V3 o, V3 inv, float tmin, float cut, float tn[4]) {
    const float4 lx = nd.lx, ly = nd.ly, lz = nd.lz, hx = nd.hx, hy = nd.hy, hz = nd.hz;
    const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, lya[4] = {ly.x, ly.y, ly.z, ly.w}, lza[4] = {lz.x, lz.y, lz.z, lz.w};
    const float hxa[4] = {hx.x, hx.y, hx.z, hx.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
    const float oix = o.x * inv.x, oiy = o.y * inv.y, oiz = o.z * inv.z;
    uint32_t m = 0u;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float ax = fmaf(lxa[c], inv.x, -oix), bx = fmaf(hxa[c], inv.x, -oix);
        const float ay = fmaf(lya[c], inv.y, -oiy), by = fmaf(hya[c], inv.y, -oiy);
        const float az = fmaf(lza[c], inv.z, -oiz), bz = fmaf(hza[c], inv.z, -oiz);
        const float n = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
        const float f = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
        tn[c] = n;
        m |= (n <= f && f >= tmin && n <= cut) ? (1u << c) : 0u;
    }
    return m;
}

__device__ __forceinline__ void links4(const Node4& nd, int32_t l[4]) {
    const float4 q = nd.links;
    l[0] = __float_as_int(q.x);
    l[1] = __float_as_int(q.y);
    l[2] = __float_as_int(q.z);
    l[3] = __float_as_int(q.w);
}

// Sorts (key, link) pairs ascending by key (equal keys in no particular order): a five-comparator
// network over four entries.
__device__ __forceinline__ void sort4(float k[4], int32_t l[4]) {
    auto cs = [&](int a, int b) {
        const bool sw = k[b] < k[a];
        const float ka = sw ? k[b] : k[a], kb = sw ? k[a] : k[b];
        const int32_t la = sw ? l[b] : l[a], lb = sw ? l[a] : l[b];
        k[a] = ka;
        k[b] = kb;
        l[a] = la;
        l[b] = lb;
    };
    cs(0, 1);
    cs(2, 3);
    cs(0, 2);
    cs(1, 3);
    cs(1, 2);
}

}  // namespace wide
}  // namespace ptir
