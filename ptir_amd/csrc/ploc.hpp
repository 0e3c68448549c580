// PLOC hierarchy builder: Parallel Locally-Ordered Clustering (Meister and Bittner, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction", IEEE TVCG 2018).
// No counterpart in the reference, whose hierarchy is OptiX's own (optixTracer.cpp:633-907).
//
// The leaves are the Gaussians in Morton order (the leaf slots of the LBVH, so the per-trace
// records and every trace result are the same for both builders).  Each iteration over the
// current clusters, in their Morton order:
//   1. nearest: every cluster i finds the neighbour j within kRadius positions whose union
//      box has the smallest surface area.  Pairs {i,j} are ordered by (area, |i-j|, parity of
//      min(i,j), min(i,j)), a strict order, so the closest pair of the iteration is mutual and
//      the count strictly decreases; on equal areas (duplicate boxes) the order pairs
//      neighbours (2k, 2k+1), so runs of equal boxes form balanced subtrees, not chains;
//   2. flags:   i is kept unless it is the upper end of a mutual pair; the lower end of a
//      mutual pair also makes a node: val[i] = kept | (makes node) << 32;
//   3. (an exclusive sum of val over the clusters, done by the caller)
//   4. compact: kept clusters move to their scanned position; a merged pair becomes a new
//      internal node, numbered downward from N-2, so the last merge is node 0, the root.
// The node records are those of bvh.hip (two child boxes, two child links; a leaf link is
// ~slot) and the parent links those of k_refit, so refits and the depth check work on either
// tree.  Each phase is a function of one cluster index, run by one GPU thread per index
// (bvh.hip) or by a host loop (the host test of this file).
#pragma once

namespace ptir {
namespace ploc {

constexpr int kRadius = 16;

struct State {
    int32_t cnt;   // clusters left
    int32_t base;  // index of the next node to allocate (counts down; -1 when done)
};

struct Iter {
    const int32_t* c_in;         // [cnt] cluster links (>= 0 internal node, < 0 ~leaf slot)
    const float4* b_in;          // [2 cnt] cluster boxes (lo, hi)
    int32_t* c_out;
    float4* b_out;
    int32_t* nn;                 // [n] nearest neighbour
    unsigned long long* val;     // [n] flags (phase 2)
    const unsigned long long* scan;  // [n] exclusive sum of val (phase 4)
    const State* st_in;
    State* st_out;
    float4* nodes;               // [(n-1) * 4]
    int32_t* parent;             // [2n-1]: internal nodes, then leaves
    int32_t n;
};

__host__ __device__ __forceinline__ float union_area(float4 alo, float4 ahi, float4 blo, float4 bhi) {
    const float dx = fmaxf(ahi.x, bhi.x) - fminf(alo.x, blo.x);
    const float dy = fmaxf(ahi.y, bhi.y) - fminf(alo.y, blo.y);
    const float dz = fmaxf(ahi.z, bhi.z) - fminf(alo.z, blo.z);
    return dx * dy + dy * dz + dz * dx;  // half the surface area
}

// phase 1
__host__ __device__ __forceinline__ void nearest(const Iter& it, int i) {
    const int cnt = it.st_in->cnt;
    if (cnt <= 1 || i >= cnt) return;
    const float4 lo = it.b_in[2 * i], hi = it.b_in[2 * i + 1];
    const int j0 = i - kRadius < 0 ? 0 : i - kRadius;
    const int j1 = i + kRadius > cnt - 1 ? cnt - 1 : i + kRadius;
    float best = 0.f;
    int bj = -1, bdist = 0, bmin = 0;
    for (int j = j0; j <= j1; ++j) {
        if (j == i) continue;
        const float d = union_area(lo, hi, it.b_in[2 * j], it.b_in[2 * j + 1]);
        const int dist = j > i ? j - i : i - j;
        const int mn = j < i ? j : i;
        const bool better =
            bj < 0 || d < best ||
            (d == best && (dist < bdist || (dist == bdist && ((mn & 1) < (bmin & 1) ||
                                                              ((mn & 1) == (bmin & 1) && mn < bmin)))));
        if (better) {
            best = d;
            bj = j;
            bdist = dist;
            bmin = mn;
        }
    }
    it.nn[i] = bj;
}

// phase 2
__host__ __device__ __forceinline__ void flags(const Iter& it, int i) {
    if (i >= it.n) return;
    const int cnt = it.st_in->cnt;
    unsigned long long v = 0ull;
    if (cnt > 1 && i < cnt) {
        const int j = it.nn[i];
        const bool mutual = it.nn[j] == i;
        v = (!mutual || i < j) ? 1ull : 0ull;
        if (mutual && i < j) v |= 1ull << 32;
    }
    it.val[i] = v;
}

// phase 4
__host__ __device__ __forceinline__ void compact(const Iter& it, int i) {
    const int cnt = it.st_in->cnt;
    if (cnt <= 1) {  // done: carry the root cluster and the state
        if (i == 0) {
            it.c_out[0] = it.c_in[0];
            it.b_out[0] = it.b_in[0];
            it.b_out[1] = it.b_in[1];
            *it.st_out = *it.st_in;
        }
        return;
    }
    if (i >= cnt) return;
    const int j = it.nn[i];
    const bool mutual = it.nn[j] == i;
    const unsigned long long s = it.scan[i];
    const int pos = int(uint32_t(s));
    const int made = int(s >> 32);
    if (i == cnt - 1) {
        const int kept = (!mutual || i < j) ? 1 : 0;
        const int makes = (mutual && i < j) ? 1 : 0;
        *it.st_out = State{pos + kept, it.st_in->base - made - makes};
    }
    if (mutual && i > j) return;  // merged into j's node
    const float4 lo = it.b_in[2 * i], hi = it.b_in[2 * i + 1];
    if (!mutual) {
        it.c_out[pos] = it.c_in[i];
        it.b_out[2 * pos] = lo;
        it.b_out[2 * pos + 1] = hi;
        return;
    }
    const int node = it.st_in->base - made;
    const int32_t cl = it.c_in[i], cr = it.c_in[j];
    const float4 rlo = it.b_in[2 * j], rhi = it.b_in[2 * j + 1];
    float4* nd = it.nodes + 4 * size_t(node);
    nd[0] = make_float4(lo.x, lo.y, lo.z, __int_as_float(cl));
    nd[1] = make_float4(hi.x, hi.y, hi.z, __int_as_float(cr));
    nd[2] = make_float4(rlo.x, rlo.y, rlo.z, 0.f);
    nd[3] = make_float4(rhi.x, rhi.y, rhi.z, 0.f);
    it.parent[cl >= 0 ? cl : (it.n - 1) + ~cl] = node;
    it.parent[cr >= 0 ? cr : (it.n - 1) + ~cr] = node;
    it.c_out[pos] = node;
    it.b_out[2 * pos] = make_float4(fminf(lo.x, rlo.x), fminf(lo.y, rlo.y), fminf(lo.z, rlo.z), 0.f);
    it.b_out[2 * pos + 1] = make_float4(fmaxf(hi.x, rhi.x), fmaxf(hi.y, rhi.y), fmaxf(hi.z, rhi.z), 0.f);
}

}  // namespace ploc
}  // namespace ptir
