// Wavefront schedule of the forward path tracer (device code; kernels and launch in wavefront.hip) (no counterpart in the reference, whose raygen
// program follows each path from start to end, referenceOptix.cu:27-71).
//
// In the one-lane-per-path kernel (k_trace_fwd) a wave keeps all 64 lanes until its longest
// path ends: once most paths have escaped, the shadow rays and the later bounces run on a few
// lanes per wave.  Here the path is split at its traversals into stages:
//   k_wf_primary  the primary segment of every path, on the pixel tiles of k_trace_fwd (the
//                 rays of a wave walk the hierarchy together); paths whose first vertex is an
//                 interaction go on to a queue;
//   k_wf_nee_setup, k_wf_occlusion, k_wf_nee_finish
//                 next-event estimation of the queued paths (with MIS on): the environment
//                 sample and the shadow ray to test (a job per path that has one), the
//                 stochastic shadow rays on the traversal engine of engine.hpp (lanes take a
//                 new ray when theirs is done), then the NEE record of each path;
//   k_wf_bounce   one launch per further path vertex: BRDF sample, Russian roulette, the next
//                 segment and its contribution; paths that go on are queued for the next launch.
// The NEE and bounce kernels are persistent: each wave takes the next 64 queue entries until
// the queue is exhausted, so every lane of a wave holds a live path.  Queues are compacted with
// one atomic per wave; a queue entry is the path's index and, in a structure-of-arrays buffer
// of its stage at the same slot, its state (39 words), so that the waves that push and take
// consecutive slots read and write consecutive words.  Each stage calls the device functions of the one-lane kernel in the same order
// with the same random stream, so every path takes the same decisions and writes the same
// outputs, tape and cache records.
#pragma once

#include "engine.hpp"

namespace ptir {
namespace wf {

// State word w of path pix at state[w * P + pix].
enum : int {
    kRng = 0,                 // 2 words: PCG state, low then high
    kThr = 2,
    kL = 5,
    kLnb = 8,
    kDirect = 11,
    kIndirect = 14,
    kELight = 17,
    kEScatter = 18,
    kEContrib = 19,
    kDir = 22,                // direction of the segment that ended at the current vertex
    kPos = 25,
    kSn = 28,
    kAlb = 31,
    kRough = 34,
    kMetal = 35,
    kUsed = 36,               // tape: hit entries used so far, pool chunk, overflow flags
    kChunk = 37,
    kFlags = 38,
    kWords = 39
};

constexpr int kBlocks = 2048;  // persistent grid of the queue kernels (8 blocks per CU)

struct IO {
    float* sa;        // states of the stages with even index, word w of slot i at [w * P + i]
    float* sb;        // and with odd index
    uint32_t* qa;     // path indices of the stages with even index
    uint32_t* qb;     // and with odd index
    uint32_t* ctrl;   // per stage s: ctrl[2s] = entries pushed, ctrl[2s + 1] = entries taken
    uint32_t* jctl;   // shadow jobs: [0] pushed, [1] taken by the engine, [2] taken by the finish
    eng::ShadowJob* jobs;  // [P]
    size_t P;
    __device__ __forceinline__ uint32_t* queue(int s) const { return (s & 1) ? qb : qa; }
    __device__ __forceinline__ float* state(int s) const { return (s & 1) ? sb : sa; }
};

// One slot of a stage's state buffer.
struct Slot {
    float* base;  // &state[slot]
    size_t P;
    __device__ __forceinline__ float& at(int w) const { return base[size_t(w) * P]; }
};
__device__ __forceinline__ Slot slot_of(const IO& io, int s, uint32_t i) { return Slot{io.state(s) + i, io.P}; }

__device__ __forceinline__ void put3(const Slot& q, int w, V3 v) {
    q.at(w) = v.x;
    q.at(w + 1) = v.y;
    q.at(w + 2) = v.z;
}
__device__ __forceinline__ V3 get3(const Slot& q, int w) { return v3(q.at(w), q.at(w + 1), q.at(w + 2)); }
__device__ __forceinline__ void putu(const Slot& q, int w, uint32_t v) { q.at(w) = __uint_as_float(v); }
__device__ __forceinline__ uint32_t getu(const Slot& q, int w) { return __float_as_uint(q.at(w)); }

// What a path carries from one stage to the next.
struct Carry {
    uint32_t used, chunk, flags;
};

__device__ __forceinline__ void save(const Slot& q, const Path& p, const Rng& rng, const Carry& c) {
    putu(q, kRng, uint32_t(rng.state));
    putu(q, kRng + 1, uint32_t(rng.state >> 32));
    put3(q, kThr, p.thr);
    put3(q, kL, p.L);
    put3(q, kLnb, p.Lnb);
    put3(q, kDirect, p.direct);
    put3(q, kIndirect, p.indirect);
    q.at(kELight) = p.elightPdf;
    q.at(kEScatter) = p.escatterPdf;
    put3(q, kEContrib, p.econtrib);
    put3(q, kDir, p.cur.ray.d);
    put3(q, kPos, p.cur.pos);
    put3(q, kSn, p.cur.sn);
    put3(q, kAlb, p.cur.alb);
    q.at(kRough) = p.cur.rough;
    q.at(kMetal) = p.cur.metal;
    putu(q, kUsed, c.used);
    putu(q, kChunk, c.chunk);
    putu(q, kFlags, c.flags);
}

// The path as the one-lane kernel holds it at the vertex where it was saved: an interaction
// (ivalid) with the fields the next stages read.
__device__ __forceinline__ void load(const Slot& q, Path& p, Rng& rng, Carry& c) {
    path_init(p);
    rng.state = uint64_t(getu(q, kRng)) | (uint64_t(getu(q, kRng + 1)) << 32);
    rng.inc = 1ull;
    p.thr = get3(q, kThr);
    p.L = get3(q, kL);
    p.Lnb = get3(q, kLnb);
    p.direct = get3(q, kDirect);
    p.indirect = get3(q, kIndirect);
    p.elightPdf = q.at(kELight);
    p.escatterPdf = q.at(kEScatter);
    p.econtrib = get3(q, kEContrib);
    p.cur.ray.d = get3(q, kDir);
    p.cur.pos = get3(q, kPos);
    p.cur.sn = get3(q, kSn);
    p.cur.alb = get3(q, kAlb);
    p.cur.rough = q.at(kRough);
    p.cur.metal = q.at(kMetal);
    p.cur.ivalid = true;
    c.used = getu(q, kUsed);
    c.chunk = getu(q, kChunk);
    c.flags = getu(q, kFlags);
}

// The end of k_trace_fwd for a path: tape flags, radiance-cache target, pbr/light/components.
__device__ __forceinline__ void finish(const DeviceScene& S, const FwdOut& out, size_t pix, const Path& p, const Carry& c) {
    if (out.tape.base) {
        uint32_t* prec = out.tape.path(pix);
        prec[0] = c.flags;
        prec[3] = c.chunk;
    }
    float* crec = out.cache ? out.cache + size_t(nrc::kRecFloats) * pix : nullptr;
    if (crec && S.cache_mode == 2u && crec[nrc::kRecFlag] != 0.f) {
        float* tg = crec + nrc::kRecTarget;
        const float* th = crec + nrc::kRecThr;
        tg[0] = th[0] > 0.f ? (p.L.x - tg[0]) / th[0] : 0.f;
        tg[1] = th[1] > 0.f ? (p.L.y - tg[1]) / th[1] : 0.f;
        tg[2] = th[2] > 0.f ? (p.L.z - tg[2]) / th[2] : 0.f;
    }
    float* o;
    o = out.pbr + 3 * pix; o[0] = p.L.x; o[1] = p.L.y; o[2] = p.L.z;
    o = out.light + 3 * pix; o[0] = p.Lnb.x; o[1] = p.Lnb.y; o[2] = p.Lnb.z;
    o = out.comp + 6 * pix;
    o[0] = p.direct.x; o[1] = p.direct.y; o[2] = p.direct.z;
    o[3] = p.indirect.x; o[4] = p.indirect.y; o[5] = p.indirect.z;
}

// Append the lanes with `want` to the queue of stage s (one atomic per wave) and return each
// such lane's slot, where the lane then saves the path's state.  Every active lane of the wave
// calls it.
__device__ __forceinline__ uint32_t push(const IO& io, int s, bool want, uint32_t pix) {
    const uint64_t m = __ballot(want);
    if (m == 0) return 0u;
    const int leader = __ffsll((long long)m) - 1;
    const int lane = int(__lane_id());
    uint32_t base = 0u;
    if (lane == leader) base = atomicAdd(io.ctrl + 2 * s, uint32_t(__popcll(m)));
    base = __shfl(base, leader);
    const uint32_t slot = base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
    if (want) io.queue(s)[slot] = pix;
    return slot;
}

// The next 64 entries of a queue for this wave (first index; >= count when exhausted).
__device__ __forceinline__ uint32_t take_at(uint32_t* taken) {
    uint32_t base = 0u;
    if (__lane_id() == 0u) base = atomicAdd(taken, 64u);
    return __shfl(base, 0);
}
__device__ __forceinline__ uint32_t take(const IO& io, int s) { return take_at(io.ctrl + 2 * s + 1); }

__device__ __forceinline__ void tape_resume(SegTape& tp, const FwdOut& out, size_t pix, size_t paths, const Carry& c) {
    path_tape_begin(tp, out.tape, pix, paths);
    tp.chunk = c.chunk;
}

// k_trace_fwd up to the primary interaction (and the NEE record without MIS).  Returns whether the
// path goes on to the queue of stage enable_mis ? 0 : 1, with its state in p, rng and c; a path
// that ends here has its outputs written.
template <int K, int SD>
__device__ __forceinline__ bool primary_path(const DeviceScene& S, const FwdOut& out, const Lane& L, uint32_t* wstack,
                                             const PixelId& px, Path& p, Rng& rng, Carry& c) {
    const size_t pix = px.pix;
    const size_t paths = size_t(S.B) * S.H * S.W;
    const bool rec = out.tape.base != nullptr;
    uint32_t* prec = rec ? out.tape.path(pix) : nullptr;
    if (rec && pix == 0) {
        uint32_t* h = out.tape.base;
        h[0] = kTapeMagic; h[1] = kTapeVersion; h[2] = uint32_t(S.B); h[3] = uint32_t(S.H); h[4] = uint32_t(S.W);
        h[5] = out.tape.segs; h[6] = S.n; h[7] = S.frame; h[8] = out.tape.cap; h[9] = tape_pool_chunks(paths);
    }
    c = Carry{0u, kNoChunk, 0u};
    SegTape tp;
    tp.hits = nullptr;
    float* crec = out.cache ? out.cache + size_t(nrc::kRecFloats) * pix : nullptr;
    if (crec) crec[nrc::kRecFlag] = 0.f;
    rng.seed(Rng::launch_seed(uint32_t(px.x), uint32_t(px.y), uint32_t(px.z), S.frame));
    path_init(p);
    count(S, C_PATHS);
    if (rec) {
        path_tape_begin(tp, out.tape, pix, paths);
        seg_tape_begin(tp, c.used);
    }
    ray_intersect<K, SD>(S, primary_ray(S, pix), p.cur, rng, false, true, true, out.vis, L, tp,
                         S.wave_traversal ? wstack : nullptr);
    if (rec) {
        store_seg(prec, 0, p.cur, tp);
        c.flags |= tp.overflow ? kTapeOverflow : 0u;
        c.used += tp.n;
        c.chunk = tp.chunk;
    }
    {
        const Seg& s = p.cur;
        float* o;
        o = out.rad + 3 * pix; o[0] = s.radiance.x; o[1] = s.radiance.y; o[2] = s.radiance.z;
        out.dns[pix] = 1.f - s.T;
        o = out.hit + 2 * pix; o[0] = s.hitDist; o[1] = s.lastHit;
        out.hit2[pix] = s.hitDist2;
        out.dist[pix] = s.distortion;
        o = out.nrm + 3 * pix; o[0] = s.normal.x; o[1] = s.normal.y; o[2] = s.normal.z;
        o = out.snrm + 3 * pix; o[0] = s.sn.x; o[1] = s.sn.y; o[2] = s.sn.z;
        o = out.mat + 5 * pix; o[0] = s.alb.x; o[1] = s.alb.y; o[2] = s.alb.z; o[3] = s.rough; o[4] = s.metal;
        out.hits[pix] = s.hits;
    }
    bool go = false;
    if (p.cur.ivalid) {
        if (S.enable_mis) {
            go = true;  // NEE next
        } else {
            if (rec) {
                prec[1] = uint32_t(rng.state);
                prec[2] = uint32_t(rng.state >> 32);
                float4* e = reinterpret_cast<float4*>(prec + 4);
                e[0] = make_float4(p.eray.d.x, p.eray.d.y, p.eray.d.z, p.elightPdf);
                e[1] = make_float4(p.escatterPdf, p.econtrib.x, p.econtrib.y, p.econtrib.z);
            }
            go = 1u < S.max_bounces;  // the first loop iteration (its accumulate adds nothing)
        }
    }
    if (!go) finish(S, out, pix, p, c);
    return go;
}

// sample_nee up to its shadow ray, for a queued path with state in p, rng: returns whether the
// path has a shadow ray to test, and then fills the job (the ray, the random stream, the
// contribution and pdfs the NEE record takes if the ray is not occluded).
__device__ __forceinline__ bool nee_setup(const DeviceScene& S, Path& p, Rng& rng, uint32_t slot, eng::ShadowJob& job) {
    emit_clear(p);
    if (!has_env_light(S)) return false;
    float envPdf;
    const V3 wi = sample_env(S, rng, envPdf);
    const V3 Li = background(S, wi);
    if (envPdf <= 0.f || is_zero(Li)) return false;
    if (dot(p.cur.sn, wi) <= 0.f) return false;
    V3 n, wo, alb;
    float met;
    shading_frame(S, p.cur, n, wo, alb, met);
    float sPdf;
    const BG bc = eval_ls<false>(wo, n, alb, met, clamp_rough(p.cur.rough), wi, sPdf, S.enable_metallic);
    const V3 o = p.cur.pos + wi * kSelfOccOffset;
    const V3 contrib = p.thr * bc.value * Li / fmaxf(envPdf, 1e-6f);
    job.a = make_float4(o.x, o.y, o.z, wi.x);
    job.b = make_float4(wi.y, wi.z, __uint_as_float(uint32_t(rng.state)), __uint_as_float(uint32_t(rng.state >> 32)));
    job.c = make_float4(contrib.x, contrib.y, contrib.z, envPdf);
    job.d = make_float4(sPdf, __uint_as_float(slot), 0.f, 0.f);
    return true;
}

// The rest of sample_nee once the shadow ray is tested: the emission record of an unoccluded
// ray, and the random stream after the test.
__device__ __forceinline__ void nee_apply(Path& p, Rng& rng, const eng::ShadowJob& job) {
    emit_clear(p);
    rng.state = uint64_t(__float_as_uint(job.b.z)) | (uint64_t(__float_as_uint(job.b.w)) << 32);
    if (job.d.z != 0.f) return;  // occluded
    p.eray = Ray{v3(job.a.x, job.a.y, job.a.z), v3(job.a.w, job.b.x, job.b.y)};
    p.elightPdf = job.c.w;
    p.escatterPdf = job.d.x;
    p.econtrib = v3(job.c.x, job.c.y, job.c.z);
}

// The NEE record of the tape after sample_nee; returns whether the path goes on to stage 1 (a
// path that ends here has its outputs written).
__device__ __forceinline__ bool nee_done(const DeviceScene& S, const FwdOut& out, uint32_t pix, const Path& p,
                                         const Rng& rng, const Carry& c) {
    if (out.tape.base) {
        uint32_t* prec = out.tape.path(pix);
        prec[1] = uint32_t(rng.state);
        prec[2] = uint32_t(rng.state >> 32);
        float4* e = reinterpret_cast<float4*>(prec + 4);
        e[0] = make_float4(p.eray.d.x, p.eray.d.y, p.eray.d.z, p.elightPdf);
        e[1] = make_float4(p.escatterPdf, p.econtrib.x, p.econtrib.y, p.econtrib.z);
    }
    const bool go = 1u < S.max_bounces;
    if (!go) finish(S, out, pix, p, c);
    return go;
}

// One iteration of k_trace_fwd's bounce loop for a path queued at vertex b - 1: BRDF sample (and
// the cache record at the first secondary vertex), roulette, segment b, its contribution.
// State in p, rng, c.  Returns whether the path goes on to stage b + 1.
template <int K, int SD>
__device__ __forceinline__ bool bounce_path(const DeviceScene& S, const FwdOut& out, const Lane& L, uint32_t pix,
                                            uint32_t b, Path& p, Rng& rng, Carry& c) {
    const size_t paths = size_t(S.B) * S.H * S.W;
    p.numBounces = b;
    bool done = false, go = false;
    float* crec = out.cache ? out.cache + size_t(nrc::kRecFloats) * pix : nullptr;
    if (crec && b == 2u) {
        cache_record(S, p, crec);
        if (S.cache_mode == 1u) done = true;  // the cache stands in for the rest of the path
    }
    if (!done) {
        sample_next(S, p, rng);
        if (vmax(p.thr) < 1e-4f || !russian_roulette(S, p, rng)) done = true;
    }
    if (!done) {
        const bool rec = out.tape.base != nullptr;
        SegTape tp;
        tp.hits = nullptr;
        if (rec) {
            tape_resume(tp, out, pix, paths, c);
            seg_tape_begin(tp, c.used);
        }
        ray_intersect<K, SD>(S, p.cur.ray, p.cur, rng, true, false, true, out.vis, L, tp);
        if (rec) {
            store_seg(out.tape.path(pix), b, p.cur, tp);
            c.flags |= tp.overflow ? kTapeOverflow : 0u;
            c.used += tp.n;
            c.chunk = tp.chunk;
        }
        accumulate(S, p);
        go = b + 1u < S.max_bounces && p.cur.ivalid;
        done = !go;
    }
    if (done) finish(S, out, pix, p, c);
    return go;
}

}  // namespace wf

}  // namespace ptir
