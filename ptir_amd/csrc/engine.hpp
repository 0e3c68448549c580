// Traversal engine for incoherent rays: persistent waves whose lanes each walk one ray and, when
// the walk is over, take the next ray from a queue (dynamic fetch, after Aila and Laine,
// "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009).  In a wave that walks a
// fixed set of 64 rays, the wave runs until its longest walk ends; with shadow and bounce rays
// that kept about one lane in ten busy.  Here a lane whose walk ended gets a new ray at the next
// phase boundary.
//
// A walk follows collect()'s rules (trace_path.hpp) exactly -- the K nearest candidates of the
// ray in (tmin, tmax), sorted by (t, leaf slot), node steps and leaf tests in separate phases
// with a queue of four leaves per lane -- so every ray gets the candidates collect() returns.
#pragma once

#include "trace_path.hpp"

namespace ptir {
namespace eng {

template <int K, int SD>
struct Walk {
    uint64_t kb[K];  // K-buffer keys (kb_key)
    V3 inv;
    float tmin, tmax;
    int link, sp, q0, q1, q2, q3;
    bool walk;

    __device__ __forceinline__ void idle() {
        walk = false;
        q0 = q1 = q2 = q3 = -1;
    }

    // A new walk of ray r in (lo, hi).
    __device__ __forceinline__ void start(const DeviceScene& S, const Ray& r, float lo, float hi) {
#pragma unroll
        for (int j = 0; j < K; ++j) kb[j] = kKbEmpty;
        count(S, C_TRAVERSALS);
        inv = slab_inv(r.d);
        tmin = lo;
        tmax = hi;
        q0 = q1 = q2 = q3 = -1;
        sp = 0;
        link = S.root;
        walk = S.n > 0 && S.root >= 0;
        if (S.n > 0 && S.root < 0) q0 = ~S.root;  // one Gaussian: its leaf only
    }

    __device__ __forceinline__ bool done() const { return !walk && q0 < 0; }

    __device__ __forceinline__ void push(int s) {
        q3 = q2;
        q2 = q1;
        q1 = q0;
        q0 = s;
    }


    // One node phase and one leaf phase; every active lane of the wave calls together (a lane
    // with no walk stands by).
    __device__ __forceinline__ void burst(const DeviceScene& S, const Ray& r, const Lane& L) {
        const bool wide_lane = (S.wide_walks & 2) != 0;
        for (;;) {
            if (__ballot(walk && q0 < 0) == 0) break;  // every walking lane has a leaf queued
            if (wide_lane && walk && q0 < 0) {
                const wide::Node4 nd = wide::load4(S.wnodes + wide::kWords * link);
                count(S, C_NODES);
                float tn[4];
                int32_t cl[4];
                const uint32_t m = wide::test4(nd, r.o, inv, tmin, fminf(tmax, kb_t(kb[K - 1])), tn);
                wide::links4(nd, cl);
                int ni = 0;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const bool hit = ((m >> c) & 1u) != 0u && cl[c] != wide::kEmpty;
                    if (hit && cl[c] < 0) push(~cl[c]);
                    const bool in = hit && cl[c] >= 0;
                    ni += in ? 1 : 0;
                    tn[c] = in ? tn[c] : __int_as_float(0x7f800000);
                }
                if (ni == 0) {
                    if (sp == 0) walk = false;
                    else link = int(L.at(--sp));
                } else {
                    wide::sort4(tn, cl);
                    if (sp + ni - 1 > SD) {  // cannot happen after the build-time stack check
                        atomicOr(S.error_flag, 1);
                        walk = false;
                    } else {
#pragma unroll
                        for (int j = 3; j >= 1; --j)
                            if (j < ni) L.at(sp++) = uint32_t(cl[j]);
                        link = cl[0];
                    }
                }
            } else if (!wide_lane && walk && q2 < 0) {    // binary nodes: room for both children
                const float4* nd = S.nodes + 4 * link;
                const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
                count(S, C_NODES);
                const float cut = fminf(tmax, kb_t(kb[K - 1]));
                float tl, tr;
                bool hl = slab(n0, n1, r, inv, tmin, cut, tl);
                bool hr = slab(n2, n3, r, inv, tmin, cut, tr);
                const int cl = __float_as_int(n0.w), cr = __float_as_int(n1.w);
                lane_step2<SD>(S, L, hl, hr, cl, cr, tl <= tr, link, sp, walk, q0, q1, q2, q3);
            }
        }
        // the queue fills from its front: q0 < 0 means empty, and the shift leaves it so
        while (__ballot(q0 >= 0) != 0) {
            const int slot = q0;
            q0 = q1;
            q1 = q2;
            q2 = q3;
            q3 = -1;
            bool ins = false;
            uint64_t k = 0;
            if (slot >= 0) {
                float t;
                count(S, C_LEAF_TESTS);
                if (cand_test(S, slot, r, tmin, tmax, t)) {
                    k = kb_key(t, uint32_t(slot));
                    ins = k < kb[K - 1];
                }
            }
            if (__ballot(ins) != 0) kb_insert<K>(kb, ins, k);
        }
    }
};

// The lanes of a wave that want a job take consecutive indices from *head (one atomic per
// wave); returns this lane's index (meaningful when want).
__device__ __forceinline__ uint32_t fetch(uint32_t* head, bool want) {
    const uint64_t m = __ballot(want);
    if (m == 0) return 0u;
    const int leader = __ffsll((long long)m) - 1;
    const int lane = int(__lane_id());
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(head, uint32_t(__popcll(m)));
    base = __shfl(base, leader);
    return base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
}

constexpr int kRefillLanes = 16;

// A shadow ray job: the NEE sample of a path and the ray to test (16 words).  The engine writes
// back the random stream after the test and whether the ray is occluded.
struct ShadowJob {
    float4 a;  // origin, direction.x
    float4 b;  // direction.y, direction.z, random stream (low, high)
    float4 c;  // NEE contribution (before MIS), environment pdf
    float4 d;  // scatter pdf, the path's slot in the NEE stage, occluded (0/1), unused
};

// pathTracer.cuh:161-214 (the stochastic shadow ray of shadow_occluded) over a queue of jobs:
// every job gets the candidates, draws and result shadow_occluded would give it.  Every lane of
// the wave runs it; returns when the queue is exhausted and every lane's job is done.
template <int K, int SD>
__device__ __forceinline__ void occlusion(const DeviceScene& S, ShadowJob* jobs, uint32_t njobs, uint32_t* head,
                                          const Lane& L) {
    Walk<K, SD> w;
    w.idle();
    bool has = false;
    uint32_t j = 0u;
    Ray ray{v3s(0.f), v3s(0.f)};
    Rng rng;
    rng.state = 0ull;
    rng.inc = 1ull;
    float startT = 0.f, endT = 0.f;
    auto finish = [&](bool occ) {
        ShadowJob* q = jobs + j;
        q->b.z = __uint_as_float(uint32_t(rng.state));
        q->b.w = __uint_as_float(uint32_t(rng.state >> 32));
        q->d.z = occ ? 1.f : 0.f;
        has = false;
    };
    const uint64_t lanes = __ballot(1);
    bool dry = false;  // the queue is exhausted (wave-uniform, sticky)
    for (;;) {
        // refill once a quarter of the wave is idle (or all of it): one atomic on the shared
        // queue head per refill, not per finished ray
        const uint64_t idle = __ballot(!has);
        const bool refill = !dry && (__popcll(idle) >= kRefillLanes || idle == lanes);
        const bool want = refill && !has;
        const uint32_t idx = fetch(head, want);
        if (refill) dry = __ballot(want && idx >= njobs) != 0;
        if (want && idx < njobs) {
            j = idx;
            const ShadowJob q = jobs[j];
            ray.o = v3(q.a.x, q.a.y, q.a.z);
            ray.d = v3(q.a.w, q.b.x, q.b.y);
            rng.state = uint64_t(__float_as_uint(q.b.z)) | (uint64_t(__float_as_uint(q.b.w)) << 32);
            float t0, t1;
            scene_interval(S, ray, t0, t1);
            startT = fmaxf(0.f, t0 - kEpsT);
            endT = fminf(kInfDist, t1) + kEpsT;
            has = true;
            if (startT < endT) w.start(S, ray, startT + kEpsT, endT);
            else finish(false);
        }
        if (__ballot(has) == 0) {
            if (dry) break;
            continue;  // every new job ended at once: take more
        }
        w.burst(S, ray, L);
        if (has && w.done()) {
            // the chunk, sorted, through the lane's LDS words (its stack is free now), then
            // shadow_occluded's loop over it
            int cnt = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                L.at(k) = uint32_t(w.kb[k] >> 32);
                L.at(K + k) = kb_s(w.kb[k]);
                cnt += kb_s(w.kb[k]) != kInvalid ? 1 : 0;
            }
            float batchEnd = startT;
            bool occ = false;
            for (int k = 0; k < cnt; ++k) {
                const float t = __uint_as_float(L.at(k));
                const int slot = int(L.at(K + k));
                if (t > endT) break;
                batchEnd = fmaxf(batchEnd, t);
                count(S, C_SHADOW_CANDIDATES);
                if (self_occluded(S, slot, t, ray.d)) continue;
                HitGeom h;
                const float a = hit_alpha(S, slot, ray, h);
                if (a > 0.f && rng.next_1d() < a) {
                    occ = true;
                    break;
                }
            }
            if (occ || cnt == 0 || batchEnd <= startT) {
                finish(occ);
            } else {
                startT = batchEnd;
                if (startT < endT) w.start(S, ray, startT + kEpsT, endT);
                else finish(false);
            }
        }
    }
}

}  // namespace eng
}  // namespace ptir
