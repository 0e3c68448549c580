// Device code shared by the path-tracing kernels (forward, replay backward, exact adjoint):
// k-buffer traversal, segment tracing, NEE, BRDF, environment and the path logic of
// referenceOptix.cu:27-71 / referenceBwdOptix.cu:27-74, plus the launch mapping.
//
// One lane owns one path sample (launch index x, y, z of the reference,
// src/kernels/cuda/referenceOptix.cu:27-71 and referenceBwdOptix.cu:27-74).  A 128-lane
// workgroup covers a 16x8 pixel tile (each wave an 8x8 tile) so that the primary rays
// of a wave traverse the same part of the hierarchy.  Workgroup ids are remapped so
// that neighbouring tiles land on the same XCD (and its L2).
//
// Hit collection replaces OptiX's any-hit k-buffer (referenceOptix.cu:95-133): a
// stack-based walk of the 4-wide hierarchy (wide.hpp) keeps the K closest candidates with
// t* in (tmin, tmax) sorted by (t*, leaf slot) in registers and prunes boxes that start
// beyond the K-th distance.
// The traversal stack lives in LDS; once the walk ends the same LDS words hold the
// sorted K-buffer, which the compositing loop then consumes in order.
#pragma once

#include "internal.hpp"

namespace ptir {

struct Ray {
    V3 o, d;
};

// ---------------------------------------------------------------- per-lane LDS view
// Waves per SIMD the trace kernels are compiled for: four when the per-lane LDS area is
// the short stack (the LDS then holds four waves per SIMD), else what registers allow.
constexpr int waves_per_eu(int K, int SD) { return (SD <= kStackShort && 2 * K <= kStackShort) ? 4 : 1; }

template <int K, int SD>
struct LdsLayout {
    static constexpr int kWords = (2 * K > SD) ? 2 * K : SD;  // per-lane stack of SD entries
};

struct Lane {
    uint32_t* base;  // &lds[threadIdx.x], stride kTraceBlock
    __device__ __forceinline__ uint32_t& at(int i) const { return base[i * kTraceBlock]; }
};

// ---------------------------------------------------------------- candidates and traversal
// Candidate rule (the deterministic restatement of the reference's IS program,
// gaussianParticles.cuh:484-500): t* = -gro.grdu/|grdu|^2 in (tmin, tmax), the peak
// point lies in the instance cube |q|_inf <= kscale, and |q|^2 < 9 |grdu|^2.
__device__ __forceinline__ bool cand_test(const DeviceScene& S, int slot, const Ray& r, float tmin, float tmax,
                                          float& t) {
    const float4* g = S.grec + 4 * slot;
    const float4 a0 = g[0], a1 = g[1], a2 = g[2], p = g[3];
    const V3 dp = r.o - v3(a0.w, a1.w, a2.w);
    const V3 gro = v3(dot(xyz(a0), dp), dot(xyz(a1), dp), dot(xyz(a2), dp));
    const V3 grdu = v3(dot(xyz(a0), r.d), dot(xyz(a1), r.d), dot(xyz(a2), r.d));
    const float dd = dot(grdu, grdu);
    // the hardware reciprocal (1 ulp) instead of an IEEE division: every path, schedule and the
    // backward's re-trace compute t* here, so they agree bit for bit
    t = -dot(gro, grdu) * __builtin_amdgcn_rcpf(dd);
    if (!(t > tmin && t < tmax)) return false;
    const V3 q = gro + grdu * t;
    if (fmaxf(fabsf(q.x), fmaxf(fabsf(q.y), fabsf(q.z))) > p.y) return false;
    return dot(q, q) < 9.0f * dd;
}

// Reciprocal of a ray direction for the slab tests: a zero (or denormal) component becomes
// +-1e20, which leaves that axis unconstrained for an origin between the planes and empty
// outside, and keeps lo * inv - o * inv finite, so each plane distance is one fused multiply-add.
__device__ __forceinline__ V3 slab_inv(V3 d) {
    auto one = [](float c) { return fabsf(c) > 1e-20f ? 1.0f / c : copysignf(1e20f, c); };
    return v3(one(d.x), one(d.y), one(d.z));
}

__device__ __forceinline__ bool slab(float4 lo, float4 hi, const Ray& r, V3 inv, float tmin, float tmax, float& tn) {
    const V3 oi = r.o * inv;
    const float ax = fmaf(lo.x, inv.x, -oi.x), bx = fmaf(hi.x, inv.x, -oi.x);
    const float ay = fmaf(lo.y, inv.y, -oi.y), by = fmaf(hi.y, inv.y, -oi.y);
    const float az = fmaf(lo.z, inv.z, -oi.z), bz = fmaf(hi.z, inv.z, -oi.z);
    tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    return tn <= tf && tf >= tmin && tn <= tmax;
}

// K-buffer entries as 64-bit keys, the distance's bits above the leaf slot: candidate
// distances are positive (t > tmin > 0), where the bits order like the floats, so one
// unsigned compare orders entries by (t, slot).  An empty entry is (+inf, kInvalid).
__device__ __forceinline__ uint64_t kb_key(float t, uint32_t s) { return (uint64_t(__float_as_uint(t)) << 32) | s; }
__device__ __forceinline__ float kb_t(uint64_t k) { return __uint_as_float(uint32_t(k >> 32)); }
__device__ __forceinline__ uint32_t kb_s(uint64_t k) { return uint32_t(k); }
constexpr uint64_t kKbEmpty = (uint64_t(0x7f800000u) << 32) | uint64_t(kInvalid);

// Sorted insertion of k into the lanes' K-buffers, for the lanes with ok set (on which
// k < kb[K-1]); the other lanes keep theirs.  From the end down, every entry above k's place
// moves up by one and k lands in the gap: one 64-bit compare and two selects per entry, with
// no branch, run by the whole wave together.  (Stopping once no lane still moves, or skipping
// entries known to be empty, costs more in tests and branches than it saves.)
template <int K>
__device__ __forceinline__ void kb_insert(uint64_t (&kb)[K], bool ok, uint64_t k) {
    bool m = ok;  // k's place is at or below j
#pragma unroll
    for (int j = K - 1; j >= 1; --j) {
        const bool bp = m && k < kb[j - 1];
        kb[j] = bp ? kb[j - 1] : (m ? k : kb[j]);
        m = bp;
    }
    kb[0] = m ? k : kb[0];
}

// The rest of a per-lane step over a binary node once its two child boxes are tested (hl, hr
// entered; cl, cr the child links; lf: the left child is entered first): the leaves entered
// join the leaf queue (left, then right, at its front), the farther of two inner children goes
// on the stack and the walk moves to the nearer, a single inner child is taken, or the walk
// pops.  Written with selects so that the lanes of a wave do not split over the cases.
template <int SD>
__device__ __forceinline__ void lane_step2(const DeviceScene& S, const Lane& L, bool hl, bool hr, int cl, int cr, bool lf,
                                           int& link, int& sp, bool& walk, int& q0, int& q1, int& q2, int& q3) {
    const bool ll = hl && cl < 0, lr = hr && cr < 0;    // leaves entered
    const bool gl = hl && cl >= 0, gr = hr && cr >= 0;  // inner children entered
    const int n = int(ll) + int(lr);
    const int first = ll ? ~cl : ~cr;
    const int n0 = n == 0 ? q0 : (n == 1 ? first : ~cr);
    const int n1 = n == 0 ? q1 : (n == 1 ? q0 : ~cl);
    const int n2 = n == 0 ? q2 : (n == 1 ? q1 : q0);
    const int n3 = n == 0 ? q3 : (n == 1 ? q2 : q1);
    q0 = n0;
    q1 = n1;
    q2 = n2;
    q3 = n3;
    const bool both = gl && gr, none = !gl && !gr;
    if (both && sp >= SD) {  // cannot happen after the build-time depth check
        atomicOr(S.error_flag, 1);
        walk = false;
        return;
    }
    if (both) L.at(sp) = uint32_t(lf ? cr : cl);
    if (none && sp > 0) link = int(L.at(sp - 1));
    if (both || !none) link = both ? (lf ? cl : cr) : (gl ? cl : cr);
    if (none && sp == 0) walk = false;
    sp += both ? 1 : (none && sp > 0 ? -1 : 0);
}

// Collect the K closest candidates in (tmin, tmax).  Returns the count; the sorted
// entries are left in LDS: t at words [0,K), leaf slot at words [K,2K).
template <int K, int SD>
__device__ __forceinline__ int collect(const DeviceScene& S, const Ray& r, float tmin, float tmax, const Lane& L) {
    uint64_t kb[K];
#pragma unroll
    for (int j = 0; j < K; ++j) kb[j] = kKbEmpty;
    // The leaf tests of the lanes that have one (has), then the wave's insertion outside the
    // branches on has and on the test (measured: the forward is 10% faster so than with the
    // insertion and the queue shift inside them).
    auto leaf = [&](bool has, int slot) {
        bool ins = false;
        uint64_t k = 0;
        if (has) {
            float t;
            count(S, C_LEAF_TESTS);
            if (cand_test(S, slot, r, tmin, tmax, t)) {
                k = kb_key(t, uint32_t(slot));
                ins = k < kb[K - 1];
            }
        }
        if (__ballot(ins) != 0) kb_insert<K>(kb, ins, k);
    };
    count(S, C_TRAVERSALS);
    const V3 inv = slab_inv(r.d);
    if (S.n > 0 && S.root < 0) {
        leaf(true, ~S.root);
    } else if (S.n > 0 && (S.wide_walks & 2)) {
        // Node steps and leaf tests run in separate phases so that each phase keeps most
        // lanes of the wave busy.  A lane with an empty leaf queue steps through one 4-wide
        // node: the leaves it enters join the queue (at most four), the inner children it
        // enters are taken nearest first (the others go on the stack, farthest deepest).  The
        // node phase ends once every lane still walking has a leaf queued, then all queues
        // drain together.  The K nearest candidates do not depend on the order of the tests;
        // culling against the K-th distance only gets later, never wrong.
        int link = S.root;
        int sp = 0;
        bool walk = true;             // nodes left to visit
        int q0 = -1, q1 = -1, q2 = -1, q3 = -1, q4 = -1, q5 = -1, q6 = -1, q7 = -1;  // queued leaf slots
        for (;;) {
            for (;;) {
                if (__ballot(walk && q0 < 0) == 0) break;  // every walking lane has a leaf queued
                if (walk && q4 < 0) {  // room for four leaves
                    const wide::Node4 nd = wide::load4(S.wnodes + wide::kWords * link);
                    count(S, C_NODES);
                    float tn[4];
                    int32_t cl[4];
                    const uint32_t m = wide::test4(nd, r.o, inv, tmin, fminf(tmax, kb_t(kb[K - 1])), tn);
                    wide::links4(nd, cl);
                    // leaves entered go to the queue's front (child order), then the nearest
                    // inner child is taken and the others pushed (farthest deepest), or the
                    // walk pops: written with selects, so the lanes do not split over the cases
                    int ni = 0;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const bool hit = ((m >> c) & 1u) != 0u && cl[c] != wide::kEmpty;
                        const bool lf = hit && cl[c] < 0;
                        q7 = lf ? q6 : q7;
                        q6 = lf ? q5 : q6;
                        q5 = lf ? q4 : q5;
                        q4 = lf ? q3 : q4;
                        q3 = lf ? q2 : q3;
                        q2 = lf ? q1 : q2;
                        q1 = lf ? q0 : q1;
                        q0 = lf ? ~cl[c] : q0;
                        const bool in = hit && cl[c] >= 0;
                        ni += in ? 1 : 0;
                        tn[c] = in ? tn[c] : __int_as_float(0x7f800000);
                    }
                    wide::sort4(tn, cl);
                    const bool pop = ni == 0;
                    const bool over = !pop && sp + ni - 1 > SD;  // cannot happen after the build-time stack check
                    if (over) atomicOr(S.error_flag, 1);
                    const int top = int(L.at(sp > 0 ? sp - 1 : 0));
                    if (!pop && !over) {
                        if (ni > 3) L.at(sp + ni - 4) = uint32_t(cl[3]);
                        if (ni > 2) L.at(sp + ni - 3) = uint32_t(cl[2]);
                        if (ni > 1) L.at(sp + ni - 2) = uint32_t(cl[1]);
                    }
                    walk = walk && !over && !(pop && sp == 0);
                    link = pop ? (sp > 0 ? top : link) : cl[0];
                    sp = pop ? (sp > 0 ? sp - 1 : 0) : (over ? sp : sp + ni - 1);
                }
            }
            // a queue is filled from its front, so q0 < 0 means it is empty and the shift
            // leaves it so
            while (__ballot(q0 >= 0) != 0) {
                const int slot = q0;
                q0 = q1;
                q1 = q2;
                q2 = q3;
                q3 = q4;
                q4 = q5;
                q5 = q6;
                q6 = q7;
                q7 = -1;
                leaf(slot >= 0, slot);
            }
            if (__ballot(walk) == 0) break;
        }
    } else if (S.n > 0) {
        // The same phases over the binary nodes: a lane queues the leaves it reaches (up to
        // four) and keeps walking while it has room for both children.
        int link = S.root;
        int sp = 0;
        bool walk = true;             // nodes left to visit
        int q0 = -1, q1 = -1, q2 = -1, q3 = -1;  // queued leaf slots, q0 first
        for (;;) {
            for (;;) {
                if (__ballot(walk && q0 < 0) == 0) break;  // every walking lane has a leaf queued
                if (walk && q2 < 0) {                         // room for both children
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
            while (__ballot(q0 >= 0) != 0) {
                const int slot = q0;
                q0 = q1;
                q1 = q2;
                q2 = q3;
                q3 = -1;
                leaf(slot >= 0, slot);
            }
            if (__ballot(walk) == 0) break;
        }
    }
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        L.at(j) = uint32_t(kb[j] >> 32);
        L.at(K + j) = kb_s(kb[j]);
        cnt += kb_s(kb[j]) != kInvalid ? 1 : 0;
    }
    return cnt;
}

// pathTracer.cuh:46-52 ray vs scene box
__device__ __forceinline__ void scene_interval(const DeviceScene& S, const Ray& r, float& t0, float& t1) {
    const float* b = S.bounds;
    const float ax = (b[0] - r.o.x) / r.d.x, bx = (b[3] - r.o.x) / r.d.x;
    const float ay = (b[1] - r.o.y) / r.d.y, by = (b[4] - r.o.y) / r.d.y;
    const float az = (b[2] - r.o.z) / r.d.z, bz = (b[5] - r.o.z) / r.d.z;
    t0 = fmaxf(0.f, fmaxf(fminf(ax, bx), fmaxf(fminf(ay, by), fminf(az, bz))));
    t1 = fminf(fmaxf(ax, bx), fminf(fmaxf(ay, by), fmaxf(az, bz)));
}

// ---------------------------------------------------------------- path state
struct Seg {  // payLoad.h:38-100 rayPayload
    Ray ray;
    bool ivalid;
    V3 pos, sn, alb;
    float rough, metal;
    V3 galb;
    float grough, gmetal;
    V3 radiance, light, contribution, normal;
    float T, hitDist, hitDist2, lastHit, distortion, hits;
    float scatterPdf, lightPdf;
    V3 dirGrad;
    bool hit, valid;
    __device__ __forceinline__ void reset(const Ray& r, float last) {
        ray = r;
        ivalid = false;
        pos = sn = alb = galb = v3s(0.f);
        rough = metal = grough = gmetal = 0.f;
        radiance = light = contribution = normal = dirGrad = v3s(0.f);
        T = 1.f;
        hitDist = hitDist2 = distortion = hits = 0.f;
        lastHit = last;
        scatterPdf = lightPdf = 0.f;
        hit = valid = false;
    }
};

struct Path {  // payLoad.h:142-168
    unsigned active, numBounces;
    Seg cur;
    // emitter payload: only these fields are read
    Ray eray;
    float elightPdf, escatterPdf;
    V3 econtrib;
    V3 L, dL, direct, indirect, Lnb, dLnb, thr;
    // PendingRayDirectionGrad
    bool pvalid;
    Ray pray;
    float popacity, pmaxHit;
    V3 pdir;
    unsigned pbounces;
    // tape hit-list offset and hit count of the current and of the pending segment
    uint32_t cseg, ccnt, pseg, pcnt;
};

__device__ __forceinline__ void emit_clear(Path& p) {
    p.eray = Ray{v3s(0.f), v3s(0.f)};
    p.elightPdf = p.escatterPdf = 0.f;
    p.econtrib = v3s(0.f);
}

// ---------------------------------------------------------------- environment (environment.cuh)
__device__ __forceinline__ bool has_env(const DeviceScene& S) { return S.env && S.env_w > 0 && S.env_h > 0; }
__device__ __forceinline__ bool has_env_light(const DeviceScene& S) {
    return has_env(S) && S.tab && S.tab_w > 0 && S.tab_h > 0;
}

// rotation with offset (0,0): rotZ = pi/2, rotX = 0 (environment.cuh:123-158)
__device__ __forceinline__ V3 env_rotate(V3 d) {
    float sz, cz;
    sincosf(0.5f * kPiF, &sz, &cz);
    return v3(d.x * cz - d.y * sz, d.x * sz + d.y * cz, d.z);
}
__device__ __forceinline__ V3 env_inverse_rotate(V3 e) {
    float sz, cz;
    sincosf(0.5f * kPiF, &sz, &cz);
    return v3(e.x * cz + e.y * sz, -e.x * sz + e.y * cz, e.z);
}

struct Foot {
    int x0, x1, y0, y1;
    float w00, w10, w01, w11;
};

__device__ __forceinline__ Foot footprint(const DeviceScene& S, float u, float v) {
    u = u - floorf(u);
    v = fminf(fmaxf(v, 0.f), 1.f);
    const float x = u * float(S.env_w) - 0.5f, y = v * float(S.env_h) - 0.5f;
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const float ax = x - float(x0), ay = y - float(y0);
    auto wrap = [](int v, int s) { int w = v % s; return w < 0 ? w + s : w; };
    Foot f;
    f.x0 = wrap(x0, S.env_w);
    f.x1 = wrap(x0 + 1, S.env_w);
    f.y0 = max(0, min(y0, S.env_h - 1));
    f.y1 = max(0, min(y0 + 1, S.env_h - 1));
    f.w00 = (1.f - ax) * (1.f - ay);
    f.w10 = ax * (1.f - ay);
    f.w01 = (1.f - ax) * ay;
    f.w11 = ax * ay;
    return f;
}

__device__ __forceinline__ V3 bilinear(const DeviceScene& S, const Foot& f) {
    const float4 a = S.env[f.y0 * S.env_w + f.x0], b = S.env[f.y0 * S.env_w + f.x1];
    const float4 c = S.env[f.y1 * S.env_w + f.x0], d = S.env[f.y1 * S.env_w + f.x1];
    return xyz(a) * f.w00 + xyz(b) * f.w10 + xyz(c) * f.w01 + xyz(d) * f.w11;
}

__device__ __forceinline__ void equirect_uv(V3 dir, float& u, float& v) {
    const float theta = atan2f(dir.x, dir.y);
    const float zcl = fminf(1.f, fmaxf(-1.f, -dir.z));
    const float phi = asinf(zcl);
    u = (theta + kPiF) * (0.5f * (1.0f / kPiF));
    v = 0.5f + phi * (1.0f / kPiF);
}

__device__ __forceinline__ V3 env_rotated(const DeviceScene& S, V3 dir) {
    float u, v;
    equirect_uv(dir, u, v);
    count(S, C_ENV_LOOKUPS);
    return bilinear(S, footprint(S, u, v));
}

__device__ __forceinline__ V3 background(const DeviceScene& S, V3 rd) {
    if (!has_env(S)) return v3s(0.f);
    return env_rotated(S, env_rotate(rd));
}

// getBackgroundColorBwd + getBackgroundRayDirectionGradFast (environment.cuh:421-473)
__device__ __forceinline__ V3 background_bwd(const DeviceScene& S, V3 rd, V3 cg, V3& dirGrad, float* denv) {
    if (!has_env(S)) return v3s(0.f);
    const V3 dir = env_rotate(rd);
    float u, v;
    equirect_uv(dir, u, v);
    const Foot f = footprint(S, u, v);
    count(S, C_ENV_LOOKUPS);
    count(S, C_ENV_ATOMICS);
    const V3 bg = bilinear(S, f);
    auto add = [&](int x, int y, float w) {
        float* p = denv + 4 * (size_t(y) * S.env_w + x);
        atomicAdd(p + 0, w * cg.x);
        atomicAdd(p + 1, w * cg.y);
        atomicAdd(p + 2, w * cg.z);
    };
    add(f.x0, f.y0, f.w00);
    add(f.x1, f.y0, f.w10);
    add(f.x0, f.y1, f.w01);
    add(f.x1, f.y1, f.w11);
    const V3 un = safe_normalize(rd);
    V3 t, b;
    onb(un, t, b);
    const float scl = rsqrtf(1.0f + kEnvDirEps * kEnvDirEps);
    const V3 tb = env_rotated(S, env_rotate((un + kEnvDirEps * t) * scl));
    const V3 bb = env_rotated(S, env_rotate((un + kEnvDirEps * b) * scl));
    const float inv = 1.0f / kEnvDirEps;
    const float tg = dot((tb - bg) * inv, cg);
    const float bgd = dot((bb - bg) * inv, cg);
    dirGrad += t * tg + b * bgd;
    return bg;
}

// first index i in [0, n) with c[i * stride] > u (the CDFs end at exactly 1 > u)
__device__ __forceinline__ int cdf_search(const float* c, int n, int stride, float u) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (c[size_t(mid) * stride] > u) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// sampleEnvironmentAliasDirection (environment.cuh:227-255); draws u1..u4 in order.  With a
// CDF table (ptir_env_table_build) the first two draws pick the row and the column by
// inverting the marginal and the row's conditional CDF instead.
__device__ __forceinline__ V3 sample_env(const DeviceScene& S, Rng& rng, float& pdf) {
    const int nc = S.tab_h * S.tab_w;
    const float* pdfs = S.tab + 2 * nc;
    int cell;
    if (S.tab_kind == 1) {
        const float ur = rng.next_1d();
        const int yr = cdf_search(S.tab + nc, S.tab_h, S.tab_w, ur);
        const float uc = rng.next_1d();
        cell = yr * S.tab_w + cdf_search(S.tab + size_t(yr) * S.tab_w, S.tab_w, 1, uc);
    } else {
        const float* prob = S.tab;
        const float* alias = S.tab + nc;
        const int init = min(int(rng.next_1d() * float(nc)), nc - 1);
        const int al = max(0, min(int(alias[init] + 0.5f), nc - 1));
        cell = rng.next_1d() < prob[init] ? init : al;
    }
    const int x = cell % S.tab_w, y = cell / S.tab_w;
    const float u1 = rng.next_1d();
    const float u2 = rng.next_1d();
    float u = (float(x) + u1) / float(S.tab_w);
    float v = (float(y) + u2) / float(S.tab_h);
    u = u - floorf(u);
    v = fminf(fmaxf(v, 0.f), 1.f);
    const float theta = u * 2.0f * kPiF - kPiF;
    const float phi = (v - 0.5f) * kPiF;
    float st, ct, sp, cp;
    sincosf(theta, &st, &ct);
    sincosf(phi, &sp, &cp);
    pdf = pdfs[cell];
    return env_inverse_rotate(v3(st * cp, ct * cp, -sp));
}

__device__ __forceinline__ float env_pdf(const DeviceScene& S, V3 rd) {
    if (!has_env_light(S)) return 0.f;
    const V3 dir = env_rotate(rd);
    float u, v;
    equirect_uv(dir, u, v);
    int x = int(u * float(S.tab_w)) % S.tab_w;
    if (x < 0) x += S.tab_w;
    const int y = max(0, min(int(v * float(S.tab_h)), S.tab_h - 1));
    return S.tab[2 * S.tab_h * S.tab_w + y * S.tab_w + x];
}

// ---------------------------------------------------------------- BRDF (material.cuh)
__device__ __forceinline__ float sat(float v) { return fminf(fmaxf(v, 0.f), 1.f); }
__device__ __forceinline__ V3 sat3(V3 v) { return v3(sat(v.x), sat(v.y), sat(v.z)); }
__device__ __forceinline__ float clamp_rough(float r) { return fminf(fmaxf(r, kMinRough), 1.f); }
__device__ __forceinline__ V3 nonneg(V3 v) { return v3(fmaxf(v.x, 0.f), fmaxf(v.y, 0.f), fmaxf(v.z, 0.f)); }
__device__ __forceinline__ V3 nonneg_mask(V3 v) { return v3(v.x > 0.f ? 1.f : 0.f, v.y > 0.f ? 1.f : 0.f, v.z > 0.f ? 1.f : 0.f); }
__device__ __forceinline__ V3 fsn(V3 v, V3 fb) {
    const float l2 = dot(v, v);
    return l2 > kBrdfEps ? v * rsqrtf(l2) : fb;
}
__device__ __forceinline__ V3 to_world(V3 n, V3 l) {
    V3 t, b;
    onb(n, t, b);
    return t * l.x + b * l.y + n * l.z;
}
__device__ __forceinline__ V3 sample_diffuse(V3 n, float u1, float u2) {
    const float xi = sat(u1), r = sqrtf(xi), phi = 2.0f * kPiF * sat(u2);
    float s, c;
    sincosf(phi, &s, &c);
    const float z = sqrtf(fmaxf(0.f, 1.f - xi));
    return to_world(n, v3(r * c, r * s, z));
}
__device__ __forceinline__ V3 sample_ggx_h(V3 n, float rough, float u1, float u2) {
    const float a = rough * rough, a2 = a * a, xi = sat(u1);
    const float den = fmaxf(1.f + (a2 - 1.f) * xi, kBrdfEps);
    const float ct = sqrtf(fmaxf(0.f, (1.f - xi) / den));
    const float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
    float s, c;
    sincosf(2.0f * kPiF * sat(u2), &s, &c);
    return to_world(n, v3(st * c, st * s, ct));
}

struct BG {
    V3 value, dAlb, dMet, dRough, dDir;
};

template <bool GRADS>
__device__ __forceinline__ BG eval_ls(V3 wo, V3 n, V3 albedo, float metal, float roughness, V3 ldir, float& pdf, bool metallic) {
    BG res;
    res.value = res.dAlb = res.dMet = res.dRough = res.dDir = v3s(0.f);
    const V3 L = fsn(ldir, n);
    const float NV = sat(dot(n, wo)), NL = sat(dot(n, L));
    pdf = 0.f;
    if (NV <= 0.f || NL <= 0.f) return res;
    const float rough = clamp_rough(roughness);
    const V3 f0 = v3s(0.04f) + (albedo - v3s(0.04f)) * metal;
    const V3 H = fsn(wo + L, n);
    const float NH = sat(dot(n, H)), VH = sat(dot(wo, H));
    const float x = 1.f - sat(VH), x2 = x * x, q = x2 * x2 * x;
    const V3 F = f0 + (v3s(1.f) - f0) * q;
    const float invPi = 1.0f / kPiF;
    const V3 diffCol = albedo * (1.f - metal);
    const float diffScale = NL * invPi;
    const V3 diffuse = diffCol * (v3s(1.f) - F) * diffScale;
    const float al = rough * rough, a2 = al * al;
    const float dDen = fmaxf(NH * NH * (a2 - 1.f) + 1.f, kBrdfEps);
    const float D = a2 * invPi / fmaxf(dDen * dDen, kBrdfEps);
    const float k = 0.5f * rough * rough;
    const float Dv = fmaxf(NV * (1.f - k) + k, kBrdfEps);
    const float Dl = fmaxf(NL * (1.f - k) + k, kBrdfEps);
    const float Gv = NV / Dv, Gl = NL / Dl, G = Gv * Gl;
    const float specScale = D * G / fmaxf(4.f * NV, kBrdfEps);
    const V3 spec = F * specScale;
    const float specPdf = VH > 0.f ? D * NH / fmaxf(4.f * VH, kBrdfEps) : 0.f;
    pdf = 0.5f * (diffScale + specPdf);
    const V3 value = diffuse + spec;
    res.value = nonneg(value);
    if (!GRADS) return res;
    const V3 mask = nonneg_mask(value);
    const float dF_dAlb = metal * (1.f - q);
    if (metallic) {
        const V3 dF_dMet = (albedo - v3s(0.04f)) * (1.f - q);
        const V3 dDiff_dMet = (-albedo * (v3s(1.f) - F) - diffCol * dF_dMet) * diffScale;
        res.dMet = (dDiff_dMet + dF_dMet * specScale) * mask;
    }
    const V3 dDiff_dAlb = ((v3s(1.f) - F) - albedo * dF_dAlb) * ((1.f - metal) * diffScale);
    res.dAlb = (dDiff_dAlb + v3s(dF_dAlb * specScale)) * mask;
    const float dD_da2 = invPi * (1.f / fmaxf(dDen * dDen, kBrdfEps) - (2.f * a2 * NH * NH) / fmaxf(dDen * dDen * dDen, kBrdfEps));
    const float dD_dr = dD_da2 * (4.f * rough * rough * rough);
    const float dGv_dk = -NV * (1.f - NV) / (Dv * Dv);
    const float dGl_dk = -NL * (1.f - NL) / (Dl * Dl);
    const float dG_dr = (Gl * dGv_dk + Gv * dGl_dk) * rough;
    const float dSpec_dr = (dD_dr * G + D * dG_dr) / fmaxf(4.f * NV, kBrdfEps);
    const float mr = (roughness > kMinRough && roughness < 1.f) ? 1.f : 0.f;
    res.dRough = F * (dSpec_dr * mr) * mask;
    return res;
}

__device__ __forceinline__ float pdf_rough_grad(V3 wo, V3 n, float roughness, V3 ldir) {
    const V3 L = fsn(ldir, n);
    const float NV = sat(dot(n, wo)), NL = sat(dot(n, L));
    if (NV <= 0.f || NL <= 0.f) return 0.f;
    if (!(roughness > kMinRough && roughness < 1.f)) return 0.f;
    const float rough = clamp_rough(roughness);
    const V3 H = fsn(wo + L, n);
    const float NH = sat(dot(n, H)), VH = sat(dot(wo, H));
    if (VH <= 0.f) return 0.f;
    const float a2 = rough * rough * rough * rough;
    const float dDen = fmaxf(NH * NH * (a2 - 1.f) + 1.f, kBrdfEps);
    const float dD_da2 = (1.0f / kPiF) * (1.f / fmaxf(dDen * dDen, kBrdfEps) - (2.f * a2 * NH * NH) / fmaxf(dDen * dDen * dDen, kBrdfEps));
    const float dD_dr = dD_da2 * (4.f * rough * rough * rough);
    return 0.5f * dD_dr * NH / fmaxf(4.f * VH, kBrdfEps);
}

template <bool GRADS>
__device__ __forceinline__ BG sample_throughput(V3 wo, V3 n, V3 albedo, float metal, float roughness, V3 rnd, V3& L, float& pdf,
                                bool metallic) {
    const float rough = clamp_rough(roughness);
    BG res;
    res.value = res.dAlb = res.dMet = res.dRough = res.dDir = v3s(0.f);
    L = n;
    pdf = 0.f;
    V3 dDir = v3s(0.f);
    if (rnd.z < 0.5f) {
        L = sample_diffuse(n, rnd.x, rnd.y);
        if (!GRADS && !(dot(n, L) > 0.f)) return res;
    } else {
        const V3 H = sample_ggx_h(n, rough, rnd.x, rnd.y);
        L = H * (2.f * dot(wo, H)) - wo;
        const bool ok = dot(n, L) > 0.f;
        if (!GRADS && !ok) return res;
        if (GRADS && (roughness > kMinRough && roughness < 1.f) && ok) {
            const float rs = clamp_rough(roughness + kRoughGradEps);
            const float step = rs - rough;
            if (step > kBrdfEps) {
                const V3 Hs = sample_ggx_h(n, rs, rnd.x, rnd.y);
                const V3 Ls = Hs * (2.f * dot(wo, Hs)) - wo;
                if (dot(n, Ls) > 0.f) dDir = (Ls - L) / step;
            }
        }
    }
    if (dot(n, L) <= 0.f) return res;
    float p;
    const BG bc = eval_ls<GRADS>(wo, n, albedo, metal, roughness, L, p, metallic);
    pdf = p;
    if (p <= kBrdfEps) {
        pdf = 0.f;
        return res;
    }
    const float invPdf = 1.f / fmaxf(p, kBrdfEps);
    const V3 value = bc.value * invPdf;
    res.value = nonneg(value);
    res.dDir = dDir;
    if (GRADS) {
        const V3 mask = nonneg_mask(value);
        const float dPdf = pdf_rough_grad(wo, n, roughness, L);
        res.dAlb = bc.dAlb * invPdf * mask;
        res.dRough = (bc.dRough * invPdf - bc.value * (dPdf * invPdf * invPdf)) * mask;
        if (metallic) res.dMet = bc.dMet * invPdf * mask;
    }
    return res;
}

__device__ __forceinline__ void shading_frame(const DeviceScene& S, const Seg& s, V3& n, V3& wo, V3& albedo, float& metal) {
    n = fsn(s.sn, v3(0.f, 0.f, 1.f));
    wo = fsn(-s.ray.d, n);
    albedo = sat3(s.alb);
    metal = S.enable_metallic ? sat(s.metal) : 0.f;
}

__device__ __forceinline__ float mis(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 / fmaxf(a2 + b2, 1e-6f);
}

// ---------------------------------------------------------------- segment tracing
__device__ __forceinline__ bool self_occluded(const DeviceScene& S, int slot, float t, V3 d) {
    if (t > kMaxSelfOcc || !S.has_snormal) return false;
    const float4 m1 = S.srec[2 * slot + 1];
    return dot(v3(m1.y, m1.z, m1.w), d) > kSelfOccDot;
}

struct HitGeom {
    V3 gro, grd;
    float gd, alpha, sig;
};

__device__ __forceinline__ float hit_alpha(const DeviceScene& S, int slot, const Ray& r, HitGeom& h) {
    const float4* g = S.grec + 4 * slot;
    const float4 a0 = g[0], a1 = g[1], a2 = g[2], p = g[3];
    const V3 dp = r.o - v3(a0.w, a1.w, a2.w);
    h.gro = v3(dot(xyz(a0), dp), dot(xyz(a1), dp), dot(xyz(a2), dp));
    h.grd = safe_normalize(v3(dot(xyz(a0), r.d), dot(xyz(a1), r.d), dot(xyz(a2), r.d)));
    const V3 gc = cross(h.grd, h.gro);
    h.gd = dot(gc, gc);
    const float gres = response(S.kernel_degree, h.gd);
    const float a = fminf(0.99f, gres * p.x);
    h.sig = p.x;
    return (gres > S.min_response && a > kAlphaMin) ? a : 0.f;
}

__device__ __forceinline__ uint32_t orig_id(const DeviceScene& S, int slot) {
    return uint32_t(__float_as_int(S.grec[4 * slot + 3].z));
}

__device__ __forceinline__ V3 sh_radiance(const DeviceScene& S, uint32_t id, V3 d) {
    const float* c = S.sh + size_t(S.sh_stride) * id;
    auto C = [&](int k) { return v3(c[3 * k], c[3 * k + 1], c[3 * k + 2]); };
    V3 r = 0.28209479177387814f * C(0);
    const int deg = S.sph_degree;
    if (deg > 0) {
        const float x = d.x, y = d.y, z = d.z;
        const float C1 = 0.4886025119029199f;
        r = r - (C1 * y) * C(1) + (C1 * z) * C(2) - (C1 * x) * C(3);
        if (deg > 1) {
            const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
            r = r + (1.0925484305920792f * xy) * C(4) + (-1.0925484305920792f * yz) * C(5) +
                (0.31539156525252005f * (2.f * zz - xx - yy)) * C(6) + (-1.0925484305920792f * xz) * C(7) +
                (0.5462742152960396f * (xx - yy)) * C(8);
            if (deg > 2) {
                r = r + (-0.5900435899266435f * y * (3.f * xx - yy)) * C(9) + (2.890611442640554f * xy * z) * C(10) +
                    (-0.4570457994644658f * y * (4.f * zz - xx - yy)) * C(11) +
                    (0.3731763325901154f * z * (2.f * zz - 3.f * xx - 3.f * yy)) * C(12) +
                    (-0.4570457994644658f * x * (4.f * zz - xx - yy)) * C(13) +
                    (1.445305721320277f * z * (xx - yy)) * C(14) + (-0.5900435899266435f * x * (xx - 3.f * yy)) * C(15);
            }
        }
    }
    r = r + v3s(0.5f);
    return v3(fmaxf(r.x, 0.f), fmaxf(r.y, 0.f), fmaxf(r.z, 0.f));
}

// pathTracer.cuh:216-311 rayIntersect with gaussianParticles.cuh:343-440 processHit
// Tape writer of one segment: accepted hits in processing order, appended to the path's
// hit list (own entries, then a pool chunk taken on first need).
struct SegTape {
    uint2* hits;         // the path's own entries; null: do not record
    uint2* pool;
    uint32_t* pool_next;
    uint32_t pool_chunks;
    uint32_t cap;        // own entries per path
    uint32_t chunk;      // the path's pool chunk (kNoChunk: none yet)
    uint32_t off;        // index of this segment's first hit in the path's hit list
    uint32_t n;
    bool overflow;
    float tlast;         // distance of the last recorded hit
    uint32_t ties;       // recorded hits at exactly tlast (the tail the flush bound may cut)
    float tbefore;       // distance of the last recorded hit before that tail
    float endT;          // flush bound of the backward re-trace
    __device__ __forceinline__ void record(uint32_t slot, float w, float t) {
        const uint32_t i = off + n;
        uint2* dst = nullptr;
        if (i < cap) {
            dst = hits + i;
        } else if (i - cap < kTapeChunk) {
            if (chunk == kNoChunk) chunk = atomicAdd(pool_next, 1u);
            if (chunk < pool_chunks) dst = pool + size_t(chunk) * kTapeChunk + (i - cap);
        }
        if (!dst) {
            overflow = true;
            return;
        }
        *dst = make_uint2(slot, __float_as_uint(w));
        ++n;
        if (t == tlast) {
            ++ties;
        } else {
            tbefore = tlast;
            tlast = t;
            ties = 1u;
        }
    }
};

// Wave-coherent variant of collect() for the rays of one wave (an 8x8 pixel tile of
// primary rays): the wave walks the hierarchy once, visiting a node when any lane's ray
// enters it before that lane's K-th nearest candidate, with a wave-uniform stack, so node
// and leaf records are read with uniform (scalar) loads and every lane runs every box
// and leaf test in lockstep.  Each lane keeps its own K-buffer and applies exactly the
// per-lane rules of collect() (a leaf counts for a lane only when the lane's ray enters
// its box before the lane's cut), so every lane gets the K candidates collect() returns.
// Every active lane of the wave must call it together; `need` = false opts a lane out.
template <int K, int SD>
__device__ __forceinline__ int wave_collect(const DeviceScene& S, const Ray& r, float tmin, float tmax, bool need,
                                            const Lane& L, uint32_t* wstack) {
    uint64_t kb[K];
#pragma unroll
    for (int j = 0; j < K; ++j) kb[j] = kKbEmpty;
    auto leaf = [&](bool hit, int slot) {
        if (hit) count(S, C_LEAF_TESTS);
        float t;
        const bool ok = cand_test(S, slot, r, tmin, tmax, t);
        // only a candidate that beats the lane's K-th entry changes its buffer; skip the
        // insertion when no lane has one
        const uint64_t k = kb_key(t, uint32_t(slot));
        const bool ins = hit & ok & (k < kb[K - 1]);
        if (__ballot(ins) != 0) kb_insert<K>(kb, ins, k);
    };
    if (need) count(S, C_TRAVERSALS);
    const V3 inv = slab_inv(r.d);
    if (S.n == 0) need = false;
    if (S.n > 0 && S.root < 0) {
        leaf(need, ~S.root);
    } else if (__ballot(need) != 0 && (S.wide_walks & 1)) {
        int link = S.root;  // wave-uniform
        int sp = 0;         // wave-uniform
        for (;;) {
            link = __builtin_amdgcn_readfirstlane(link);
            const wide::Node4 nd = wide::load4(S.wnodes + wide::kWords * link);
            if (need) count(S, C_NODES);
            float tn[4];
            int32_t cl[4];
            const uint32_t m = need ? wide::test4(nd, r.o, inv, tmin, fminf(tmax, kb_t(kb[K - 1])), tn) : 0u;
            wide::links4(nd, cl);
#pragma unroll
            for (int c = 0; c < 4; ++c) cl[c] = __builtin_amdgcn_readfirstlane(cl[c]);
            // leaves that some lane enters: every lane tests them (a lane counts a leaf only
            // when its own ray enters the box)
            uint32_t go = 0u;  // wave-uniform: inner children that some lane enters
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool hit = ((m >> c) & 1u) != 0u;
                const bool any = __ballot(hit) != 0 && cl[c] != wide::kEmpty;
                if (any && cl[c] < 0) leaf(hit, ~cl[c]);
                if (any && cl[c] >= 0) go |= 1u << c;
            }
            if (go == 0u) {
                if (sp == 0) break;
                --sp;
                link = int(wstack[sp]);
                continue;
            }
            // order of the inner children: entry distance as seen by the first lane that enters
            // one of them (children it does not enter come after, in child order)
            const int leader = __ffsll((long long)__ballot((m & go) != 0u)) - 1;
            float key[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float mine = ((m >> c) & 1u) != 0u ? tn[c] : 3.0e38f;
                const float lead = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine), leader));
                key[c] = ((go >> c) & 1u) != 0u ? lead : __int_as_float(0x7f800000);
            }
            wide::sort4(key, cl);
            const int ni = __popc(go);
            if (sp + ni - 1 > kStackDeep) {  // cannot happen after the build-time stack check
                atomicOr(S.error_flag, 1);
                break;
            }
#pragma unroll
            for (int j = 3; j >= 1; --j)
                if (j < ni) wstack[sp++] = uint32_t(cl[j]);  // every lane stores the same uniform value
            link = cl[0];
        }
    } else if (__ballot(need) != 0) {
        int link = S.root;  // wave-uniform
        int sp = 0;         // wave-uniform
        for (;;) {
            link = __builtin_amdgcn_readfirstlane(link);
            const float4* nd = S.nodes + 4 * link;
            const float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3];
            if (need) count(S, C_NODES);
            const float cut = fminf(tmax, kb_t(kb[K - 1]));
            float tl = 0.f, tr = 0.f;
            const bool hl = need && slab(n0, n1, r, inv, tmin, cut, tl);
            const bool hr = need && slab(n2, n3, r, inv, tmin, cut, tr);
            const int cl = __builtin_amdgcn_readfirstlane(__float_as_int(n0.w));
            const int cr = __builtin_amdgcn_readfirstlane(__float_as_int(n1.w));
            const uint64_t bl = __ballot(hl), br = __ballot(hr);
            if (cl < 0 && bl != 0) leaf(hl, ~cl);
            if (cr < 0 && br != 0) leaf(hr, ~cr);
            const bool gl = cl >= 0 && bl != 0, gr = cr >= 0 && br != 0;
            if (gl && gr) {
                // near child first, as seen by the first lane that enters both
                uint64_t both = __ballot(hl && hr);
                if (both == 0) both = bl;
                const int leader = __ffsll((long long)both) - 1;
                const bool lf = __shfl(tl <= tr ? 1 : 0, leader) != 0;
                if (sp >= kStackDeep) {  // cannot happen after the build-time depth check
                    atomicOr(S.error_flag, 1);
                    break;
                }
                wstack[sp] = uint32_t(lf ? cr : cl);  // every lane stores the same uniform value
                ++sp;
                link = lf ? cl : cr;
            } else if (gl) {
                link = cl;
            } else if (gr) {
                link = cr;
            } else {
                if (sp == 0) break;
                --sp;
                link = int(wstack[sp]);
            }
        }
    }
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        L.at(j) = uint32_t(kb[j] >> 32);
        L.at(K + j) = kb_s(kb[j]);
        cnt += kb_s(kb[j]) != kInvalid ? 1 : 0;
    }
    return cnt;
}

template <int K, int SD>
__device__ __forceinline__ void ray_intersect(const DeviceScene& S, const Ray ray, Seg& pl, Rng& rng, bool selfOcc, bool primary,
                              bool gbuffer, uint32_t* vis, const Lane& L, SegTape& tp, uint32_t* wstack = nullptr) {
    float t0, t1;
    scene_interval(S, ray, t0, t1);
    const float sPdf = pl.scatterPdf, lPdf = pl.lightPdf;
    pl.reset(ray, fmaxf(0.f, t0 - kEpsT));
    pl.scatterPdf = sPdf;
    pl.lightPdf = lPdf;
    float depth = 0.f;
    V3 alb = v3s(0.f), snAcc = v3s(0.f);
    float rough = 0.f, metal = 0.f;
    const float minT = S.min_transmittance;
    // radiance and normals feed only the G-buffer outputs (and, with INDIRECT, the
    // secondary radiance); the replay in the backward needs neither
    const bool need_rad = (primary && gbuffer) || S.render_opts == 1;
    const bool need_nrm = primary && gbuffer && S.enable_normals;
    const float dlen = sqrtf(dot(ray.d, ray.d));
    count(S, C_SEGMENTS);
    // gaussianParticles.cuh:343-440 processHit and pathTracer.cuh:245-292 for one candidate
    auto proc = [&](float t, int slot) {
        count(S, C_CANDIDATES);
        if (selfOcc && self_occluded(S, slot, t, ray.d)) {
            pl.lastHit = fmaxf(pl.lastHit, t);
            return;
        }
        HitGeom h;
        const float a = hit_alpha(S, slot, ray, h);
        const bool acc = a > 0.f;
        if (acc) {
            const float w = a * pl.T;
            const float hitT = t * dlen;
            const uint32_t id = orig_id(S, slot);
            if (need_rad) pl.radiance += sh_radiance(S, id, ray.d) * w;
            const float4 m0 = S.srec[2 * slot], m1 = S.srec[2 * slot + 1];
            alb += xyz(m0) * w;
            rough += m0.w * w;
            metal += m1.x * w;
            pl.T *= (1.f - a);
            depth += hitT * w;
            pl.hitDist2 += hitT * hitT * w;
            if (S.enable_normals) {
                snAcc += w * v3(m1.y, m1.z, m1.w);
                if (need_nrm) {
                    const float4 rs = S.prec[slot];
                    const float proj = dot(h.grd, -h.gro);
                    pl.normal += w * safe_normalize((h.gro + h.grd * (proj - sqrtf(9.f - h.gd))) * xyz(rs));
                }
            }
            if (tp.hits) tp.record(uint32_t(slot), w, t);
            pl.hit = true;
            const float op = 1.f - pl.T;
            pl.distortion = fmaxf(op * pl.hitDist2 - depth * depth, 0.f);
            // the flag is idempotent: skip the store when a previous path already set it
            if (vis && vis[id] == 0u) vis[id] = 1u;
            count(S, C_ACCEPTED);
            if (primary) count(S, C_ACCEPTED_PRIMARY);
        }
        pl.lastHit = fmaxf(pl.lastHit, t);
        if (S.enable_hitcounts) pl.hits += acc ? 1.f : 0.f;
    };
    if (wstack) {
        // all lanes of the wave arrive here together (the primary segment)
        bool done = false;
        for (;;) {
            const bool need = !done && pl.lastHit <= t1 && pl.T > minT;
            if (__ballot(need) == 0) break;
            const int cnt = wave_collect<K, SD>(S, ray, pl.lastHit + kEpsT, t1 + kEpsT, need, L, wstack);
            if (need) {
                if (cnt == 0) done = true;
                for (int k = 0; k < cnt; ++k) {
                    if (!(pl.T > minT)) break;
                    proc(__uint_as_float(L.at(k)), int(L.at(K + k)));
                }
            }
        }
    } else {
        while (pl.lastHit <= t1 && pl.T > minT) {
            const int cnt = collect<K, SD>(S, ray, pl.lastHit + kEpsT, t1 + kEpsT, L);
            if (cnt == 0) break;
            for (int k = 0; k < cnt; ++k) {
                if (!(pl.T > minT)) break;
                proc(__uint_as_float(L.at(k)), int(L.at(K + k)));
            }
        }
    }
    if (tp.hits) {
        // the backward flush re-traces (tmin, min(lastHit, t1) + 1e-9f) (pathTracer.cuh:326-329):
        // in float32 the bound usually equals lastHit, which drops the last processed hit
        tp.endT = fminf(pl.lastHit, t1) + kEpsT;
        if (tp.n > 0 && tp.tlast >= tp.endT) {
            tp.n -= tp.ties;  // every hit at the last distance falls outside the bound
            if (tp.n > 0 && tp.tbefore >= tp.endT) tp.overflow = true;  // cannot happen; the replay would handle it
        }
    }
    const float op = 1.f - pl.T;
    pl.hitDist = depth;
    pl.valid = false;
    if (pl.hit && op > 0.5f) {
        const float inv = 1.f / fmaxf(op, 1e-12f);
        pl.ivalid = true;
        pl.pos = ray.o + ray.d * (depth * inv);
        pl.sn = snAcc * inv;
        pl.alb = alb * inv;
        pl.rough = rough * inv;
        pl.metal = metal * inv;
    } else {
        const float u = rng.next_1d();
        pl.valid = u < pl.T;
    }
}

// pathTracer.cuh:161-214 stochastic shadow ray
template <int K, int SD>
__device__ __forceinline__ bool shadow_occluded(const DeviceScene& S, const Ray ray, Rng& rng, float tmin, float tmax, const Lane& L) {
    float t0, t1;
    scene_interval(S, ray, t0, t1);
    float startT = fmaxf(tmin, t0 - kEpsT);
    const float endT = fminf(tmax, t1) + kEpsT;
    if (startT >= endT) return false;
    while (startT < endT) {
        const int cnt = collect<K, SD>(S, ray, startT + kEpsT, endT, L);
        if (cnt == 0) break;
        float batchEnd = startT;
        for (int k = 0; k < cnt; ++k) {
            const float t = __uint_as_float(L.at(k));
            const int slot = int(L.at(K + k));
            if (t > endT) break;
            batchEnd = fmaxf(batchEnd, t);
            count(S, C_SHADOW_CANDIDATES);
            if (self_occluded(S, slot, t, ray.d)) continue;
            HitGeom h;
            const float a = hit_alpha(S, slot, ray, h);
            if (a > 0.f && rng.next_1d() < a) return true;
        }
        if (batchEnd <= startT) break;
        startT = batchEnd;
    }
    return false;
}

// pathTracer.cuh:313-390 material-gradient flush over a replayed segment
template <int K, int SD>
__device__ __forceinline__ void intersect_bwd(const DeviceScene& S, const Ray ray, float opacity, float maxHit, V3 ga, float gr,
                              float gm, bool selfOcc, float* dmat_leaf, const Lane& L) {
    const float inv = 1.f / fmaxf(opacity, 1e-12f);
    ga = ga * inv;
    gr *= inv;
    gm *= inv;
    float t0, t1;
    scene_interval(S, ray, t0, t1);
    float startT = fmaxf(0.f, t0 - kEpsT);
    const float endT = fminf(maxHit, t1) + kEpsT;
    float T = 1.f;
    const float minT = S.min_transmittance;
    while (startT < endT && T > minT) {
        const int cnt = collect<K, SD>(S, ray, startT + kEpsT, endT, L);
        if (cnt == 0) break;
        for (int k = 0; k < cnt; ++k) {
            if (!(T > minT)) break;
            const float t = __uint_as_float(L.at(k));
            const int slot = int(L.at(K + k));
            if (selfOcc && self_occluded(S, slot, t, ray.d)) {
                startT = fmaxf(startT, t);
                continue;
            }
            HitGeom h;
            const float a = hit_alpha(S, slot, ray, h);
            if (a > 0.f) {
                const float w = a * T;
                float* q = dmat_leaf + 5 * size_t(slot);
                atomicAdd(q + 0, w * ga.x);
                atomicAdd(q + 1, w * ga.y);
                atomicAdd(q + 2, w * ga.z);
                atomicAdd(q + 3, w * gr);
                if (S.enable_metallic) atomicAdd(q + 4, w * gm);
                count(S, C_MAT_ATOMICS);
                T *= (1.f - a);
            }
            startT = fmaxf(startT, t);
        }
    }
}

// ---------------------------------------------------------------- path logic
template <int K, int SD>
__device__ __forceinline__ void sample_nee(const DeviceScene& S, Path& p, Rng& rng, const Lane& L) {
    emit_clear(p);
    if (!has_env_light(S)) return;
    float envPdf;
    const V3 wi = sample_env(S, rng, envPdf);
    const V3 Li = background(S, wi);
    if (envPdf <= 0.f || is_zero(Li)) return;
    if (dot(p.cur.sn, wi) <= 0.f) return;
    V3 n, wo, alb;
    float met;
    shading_frame(S, p.cur, n, wo, alb, met);
    float sPdf;
    const BG bc = eval_ls<false>(wo, n, alb, met, clamp_rough(p.cur.rough), wi, sPdf, S.enable_metallic);
    const Ray lr{p.cur.pos + wi * kSelfOccOffset, wi};
    if (shadow_occluded<K, SD>(S, lr, rng, 0.f, kInfDist, L)) return;
    p.eray = lr;
    p.elightPdf = envPdf;
    p.escatterPdf = sPdf;
    p.econtrib = p.thr * bc.value * Li / fmaxf(envPdf, 1e-6f);
}

__device__ __forceinline__ void accumulate(const DeviceScene& S, Path& p) {
    p.cur.contribution = v3s(0.f);
    const bool first = p.numBounces == 1u;
    const bool mis_on = S.enable_mis != 0;
    const float bMis = (mis_on && first) ? mis(p.cur.scatterPdf, p.cur.lightPdf) : 1.f;
    const float lMis = (mis_on && first) ? mis(p.elightPdf, p.escatterPdf) : 1.f;
    const V3 nee = p.econtrib * lMis;
    if (S.render_opts == 1) {
        if (mis_on && first) p.L += nee;
        if (p.numBounces > 0u && p.cur.ivalid) {
            p.cur.contribution = p.thr * p.cur.radiance;
            if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
            p.L += p.cur.contribution;
            p.indirect += p.cur.contribution;
        } else if (p.cur.valid) {
            if (!has_env(S)) return;
            const V3 bg = background(S, p.cur.ray.d);
            p.thr = p.thr * p.cur.T;
            p.cur.light += bg;
            p.Lnb += bg;
            p.cur.contribution = p.thr * p.cur.light;
            if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
            p.L += p.cur.contribution;
            p.direct += p.cur.contribution;
        }
        return;
    }
    if (mis_on && first) {
        p.L += nee;
        p.direct += nee;
    }
    if (p.cur.valid) {
        if (!has_env(S)) return;
        const V3 bg = background(S, p.cur.ray.d);
        p.thr = p.thr * p.cur.T;
        p.cur.light += bg;
        p.Lnb += bg;
        p.cur.contribution = p.thr * p.cur.light;
        if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
        p.cur.radiance += p.cur.contribution;
        p.L += p.cur.contribution;
        // per-component value selects: a branch between the two members becomes a
        // pointer select, which keeps the whole path state out of registers
        const bool isd = p.numBounces < 2u;
        const V3 c = p.cur.contribution;
        p.direct.x = isd ? p.direct.x + c.x : p.direct.x;
        p.direct.y = isd ? p.direct.y + c.y : p.direct.y;
        p.direct.z = isd ? p.direct.z + c.z : p.direct.z;
        p.indirect.x = isd ? p.indirect.x : p.indirect.x + c.x;
        p.indirect.y = isd ? p.indirect.y : p.indirect.y + c.y;
        p.indirect.z = isd ? p.indirect.z : p.indirect.z + c.z;
    }
}

__device__ __forceinline__ void accumulate_bwd(const DeviceScene& S, Path& p, float* denv) {
    p.cur.contribution = v3s(0.f);
    const bool first = p.numBounces == 1u;
    const bool mis_on = S.enable_mis != 0;
    const float bMis = (mis_on && first) ? mis(p.cur.scatterPdf, p.cur.lightPdf) : 1.f;
    const float lMis = (mis_on && first) ? mis(p.elightPdf, p.escatterPdf) : 1.f;
    if (S.render_opts == 1) {
        if (mis_on && first) p.L -= p.econtrib * lMis;
        if (p.numBounces > 0u && p.cur.ivalid) {
            p.cur.contribution = p.thr * p.cur.radiance;
            if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
            p.L -= p.cur.contribution;
        } else if (p.cur.valid) {
            if (!has_env(S)) return;
            p.thr = p.thr * p.cur.T;
            V3 eg = p.dL * p.thr;
            if (mis_on && first) eg = eg * bMis;
            eg += p.dLnb;
            const V3 vis = background_bwd(S, p.cur.ray.d, eg, p.cur.dirGrad, denv);
            p.cur.contribution = p.thr * vis;
            if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
            p.L -= p.cur.contribution;
        }
        return;
    }
    if (mis_on && first) p.L -= p.econtrib * lMis;
    if (p.cur.valid) {
        if (!has_env(S)) return;
        p.thr = p.thr * p.cur.T;
        V3 eg = p.dL * p.thr;
        if (mis_on && first) eg = eg * bMis;
        eg += p.dLnb;
        const V3 vis = background_bwd(S, p.cur.ray.d, eg, p.cur.dirGrad, denv);
        p.cur.contribution = p.thr * vis;
        if (mis_on && first) p.cur.contribution = p.cur.contribution * bMis;
        p.cur.radiance += p.cur.contribution;
        p.L -= p.cur.contribution;
    }
}

__device__ __forceinline__ void sample_next(const DeviceScene& S, Path& p, Rng& rng) {
    V3 n, wo, alb;
    float met;
    shading_frame(S, p.cur, n, wo, alb, met);
    const float r0 = rng.next_1d();
    const float r1 = rng.next_1d();
    const float r2 = rng.next_1d();
    V3 L = p.cur.ray.d;
    float sPdf = 0.f;
    const BG bt = sample_throughput<false>(wo, n, alb, met, clamp_rough(p.cur.rough), v3(r0, r1, r2), L, sPdf,
                                           S.enable_metallic);
    const V3 nt = p.thr * bt.value;
    if (is_zero(nt)) {
        p.thr = v3s(0.f);
        p.active = 0;
        return;
    }
    p.thr = nt;
    const Ray nr{p.cur.pos + safe_normalize(L) * kSelfOccOffset, L};
    const float lpdf = S.enable_mis ? env_pdf(S, L) : 0.f;
    p.cur.reset(nr, 0.f);
    if (S.enable_mis) {
        p.cur.scatterPdf = sPdf;
        p.cur.lightPdf = lpdf;
    }
}

// Where the backward's material-gradient flushes come from: a re-trace of the segment
// (pathTracer.cuh:313-390) or, with a tape, the segment's recorded hits.  Both give the
// same per-hit products w * (g / opacity), with w = alpha * T computed identically.
// Material gradients of one workgroup gathered in LDS before they reach memory: an open-
// addressing table from leaf slot to the five gradient sums.  The primary rays of a 16 x 8 pixel
// tile hit far fewer distinct Gaussians than hits (a Gaussian covers many neighbouring pixels),
// so one global atomic per distinct Gaussian and float replaces one per hit.  A hit whose slot
// finds no room within kMatProbe probes goes to global memory directly.
constexpr int kMatTable = 1024;  // entries (power of two)
constexpr int kMatProbe = 16;
constexpr uint32_t kMatEmpty = 0xffffffffu;
struct MatTable {
    uint32_t key[kMatTable];
    float val[5][kMatTable];
};

// Every thread of the workgroup calls it (it synchronises the workgroup).
__device__ __forceinline__ void mat_table_clear(MatTable& t) {
    for (int i = threadIdx.x; i < kMatTable; i += blockDim.x) {
        t.key[i] = kMatEmpty;
#pragma unroll
        for (int c = 0; c < 5; ++c) t.val[c][i] = 0.f;
    }
    __syncthreads();
}

__device__ __forceinline__ void mat_add(MatTable* t, float* dmat_leaf, uint32_t slot, const float v[5], int nv) {
    if (t) {
        uint32_t h = (slot * 2654435761u) >> (32 - 10);
        static_assert(kMatTable == 1 << 10, "hash width");
        for (int k = 0; k < kMatProbe; ++k, h = (h + 1u) & uint32_t(kMatTable - 1)) {
            uint32_t cur = t->key[h];
            if (cur == kMatEmpty) cur = atomicCAS(&t->key[h], kMatEmpty, slot);
            if (cur == kMatEmpty || cur == slot) {
                for (int c = 0; c < nv; ++c) atomicAdd(&t->val[c][h], v[c]);
                return;
            }
        }
    }
    float* q = dmat_leaf + 5 * size_t(slot);
    for (int c = 0; c < nv; ++c) atomicAdd(q + c, v[c]);
}

// Every thread of the workgroup calls it after its last mat_add (it synchronises the workgroup).
__device__ __forceinline__ void mat_table_flush(MatTable& t, float* dmat_leaf) {
    __syncthreads();
    for (int i = threadIdx.x; i < kMatTable; i += blockDim.x) {
        const uint32_t slot = t.key[i];
        if (slot == kMatEmpty) continue;
        float* q = dmat_leaf + 5 * size_t(slot);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const float v = t.val[c][i];
            if (v != 0.f) atomicAdd(q + c, v);
        }
    }
}

struct Flush {
    const BwdIO* io;
    size_t pix, paths;
    const Lane* L;
    uint32_t chunk;  // the path's tape pool chunk (kNoChunk: none)
    MatTable* tab;   // the workgroup's gradient table, or null (atomics straight to memory)
};

__device__ __forceinline__ void tape_flush(const DeviceScene& S, const Flush& f, uint32_t off, uint32_t cnt, float opacity,
                                           V3 ga, float gr, float gm) {
    const float inv = 1.f / fmaxf(opacity, 1e-12f);
    ga = ga * inv;
    gr *= inv;
    gm *= inv;
    const Tape& tape = f.io->tape;
    uint2* own = tape.hits(f.pix, f.paths);
    uint2* pl = tape.pool(f.paths);
    const int nv = S.enable_metallic ? 5 : 4;
    for (uint32_t i = off; i < off + cnt; ++i) {
        const uint2 e = *tape.entry(own, pl, f.chunk, i);
        const float w = __uint_as_float(e.y);
        const float v[5] = {w * ga.x, w * ga.y, w * ga.z, w * gr, w * gm};
        mat_add(f.tab, f.io->dmat_leaf, e.x, v, nv);
        count(S, C_MAT_ATOMICS);
    }
}

template <int K, int SD, bool TAPE>
__device__ __forceinline__ void flush_seg(const DeviceScene& S, const Flush& f, const Ray& ray, float opacity, float maxHit,
                                          uint32_t seg, uint32_t cnt, V3 ga, float gr, float gm, bool selfOcc) {
    if constexpr (TAPE) {
        tape_flush(S, f, seg, cnt, opacity, ga, gr, gm);
    } else {
        intersect_bwd<K, SD>(S, ray, opacity, maxHit, ga, gr, gm, selfOcc, f.io->dmat_leaf, *f.L);
    }
}

template <int K, int SD, bool TAPE>
__device__ __forceinline__ void pending_bwd(const DeviceScene& S, Path& p, const Flush& f) {
    if (!p.pvalid) return;
    const float dR = dot(p.cur.dirGrad, p.pdir);
    p.pvalid = false;
    if (dR == 0.f) return;
    flush_seg<K, SD, TAPE>(S, f, p.pray, p.popacity, p.pmaxHit, p.pseg, p.pcnt, v3s(0.f), dR, 0.f, p.pbounces > 1u);
}

__device__ __forceinline__ void nee_grad_bwd(const DeviceScene& S, Path& p, const Seg& cs) {
    if (p.numBounces != 1u) return;
    if (p.elightPdf <= 0.f) return;
    V3 n, wo, alb;
    float met;
    shading_frame(S, cs, n, wo, alb, met);
    float sPdf;
    const BG bc = eval_ls<true>(wo, n, alb, met, cs.rough, p.eray.d, sPdf, S.enable_metallic);
    const float lMis = mis(p.elightPdf, p.escatterPdf);
    const float scale = lMis / fmaxf(p.elightPdf, 1e-6f);
    const V3 bg = background(S, p.eray.d);
    const V3 dLdB = p.dL * p.thr * bg * scale;
    p.cur.galb += dLdB * bc.dAlb;
}

template <int K, int SD, bool TAPE>
__device__ __forceinline__ void sample_next_bwd(const DeviceScene& S, Path& p, Rng& rng, const Flush& f) {
    V3 n, wo, alb;
    float met;
    shading_frame(S, p.cur, n, wo, alb, met);
    const float r0 = rng.next_1d();
    const float r1 = rng.next_1d();
    const float r2 = rng.next_1d();
    V3 Ld = p.cur.ray.d;
    float sPdf = 0.f;
    const BG bt = sample_throughput<true>(wo, n, alb, met, p.cur.rough, v3(r0, r1, r2), Ld, sPdf, S.enable_metallic);
    const V3 nt = p.thr * bt.value;
    if (is_zero(nt)) {
        p.thr = v3s(0.f);
        p.active = 0;
        return;
    }
    const bool mis_on = S.enable_mis != 0;
    float lPdf = 0.f;
    bool firstNee = false;
    if (mis_on) {
        lPdf = env_pdf(S, Ld);
        firstNee = p.numBounces == 1u && vmax(nt) >= 1e-4f;
        if (firstNee) nee_grad_bwd(S, p, p.cur);
    }
    V3 dep = p.L;
    if (mis_on && firstNee) dep -= p.econtrib * mis(p.elightPdf, p.escatterPdf);
    const V3 num = p.dL * dep;
    const V3 dLdB = v3(bt.value.x > kBrdfEps ? num.x / bt.value.x : 0.f, bt.value.y > kBrdfEps ? num.y / bt.value.y : 0.f,
                       bt.value.z > kBrdfEps ? num.z / bt.value.z : 0.f);
    p.cur.galb += dLdB * bt.dAlb;
    p.cur.grough += dot(dLdB, bt.dRough);
    if (S.enable_metallic) p.cur.gmetal += dot(dLdB, bt.dMet);
    p.pvalid = dot(bt.dDir, bt.dDir) > 0.f;
    p.pray = p.cur.ray;
    p.popacity = 1.f - p.cur.T;
    p.pmaxHit = p.cur.lastHit;
    p.pdir = bt.dDir;
    p.pbounces = p.numBounces;
    p.pseg = p.cseg;
    p.pcnt = p.ccnt;
    flush_seg<K, SD, TAPE>(S, f, p.cur.ray, 1.f - p.cur.T, p.cur.lastHit, p.cseg, p.ccnt, p.cur.galb, p.cur.grough,
                       p.cur.gmetal, p.numBounces > 1u);
    p.thr = nt;
    const Ray nr{p.cur.pos + safe_normalize(Ld) * kSelfOccOffset, Ld};
    p.cur.reset(nr, 0.f);
    if (mis_on) {
        p.cur.scatterPdf = sPdf;
        p.cur.lightPdf = lPdf;
    }
}

// Russian roulette after the BRDF sample of path vertex numBounces (no counterpart in the
// reference, which only cuts throughputs below 1e-4, referenceOptix.cu:62-65): from vertex
// rr_start on the path survives with probability q = min(1, max throughput) and divides its
// throughput by q.  One draw, taken identically by the forward and the replay backward
// (which treats 1/q as a constant of the path).
__device__ __forceinline__ bool russian_roulette(const DeviceScene& S, Path& p, Rng& rng) {
    if (S.rr_start == 0u || p.numBounces < S.rr_start) return true;
    const float q = fminf(1.0f, vmax(p.thr));
    if (!(rng.next_1d() < q)) return false;
    p.thr = p.thr * (1.0f / q);
    return true;
}

// ---------------------------------------------------------------- tape records
__device__ __forceinline__ void store_seg(uint32_t* rec, uint32_t s, const Seg& pl, const SegTape& tp) {
    float4* q = reinterpret_cast<float4*>(rec + kTapePathWords + kTapeSegWords * s);
    const uint32_t packed = (tp.n & 0xFFFFu) | (pl.ivalid ? kSegIValid : 0u) | (pl.valid ? kSegValid : 0u);
    q[0] = make_float4(pl.T, tp.endT, pl.lastHit, __uint_as_float(packed));
    q[1] = make_float4(pl.pos.x, pl.pos.y, pl.pos.z, pl.sn.x);
    q[2] = make_float4(pl.sn.y, pl.sn.z, pl.alb.x, pl.alb.y);
    q[3] = make_float4(pl.alb.z, pl.rough, pl.metal, __uint_as_float(tp.off));
}

// Segment s from the tape, traced along `ray`; keeps the MIS pdfs already set on pl.
// Returns the hit count; `off` receives the offset of its hits in the path's hit list.
__device__ __forceinline__ uint32_t load_seg(const uint32_t* rec, uint32_t s, const Ray& ray, Seg& pl, uint32_t& off) {
    const float4* q = reinterpret_cast<const float4*>(rec + kTapePathWords + kTapeSegWords * s);
    const float4 a = q[0], b = q[1], c = q[2], d = q[3];
    const float sPdf = pl.scatterPdf, lPdf = pl.lightPdf;
    pl.reset(ray, a.z);
    pl.scatterPdf = sPdf;
    pl.lightPdf = lPdf;
    const uint32_t packed = __float_as_uint(a.w);
    pl.T = a.x;
    pl.ivalid = (packed & kSegIValid) != 0u;
    pl.valid = (packed & kSegValid) != 0u;
    pl.pos = v3(b.x, b.y, b.z);
    pl.sn = v3(b.w, c.x, c.y);
    pl.alb = v3(c.z, c.w, d.x);
    pl.rough = d.y;
    pl.metal = d.z;
    off = __float_as_uint(d.w);
    return packed & 0xFFFFu;
}

// Start recording a path: its own hit entries, no pool chunk yet.
__device__ __forceinline__ void path_tape_begin(SegTape& tp, const Tape& t, size_t pix, size_t paths) {
    tp.hits = t.hits(pix, paths);
    tp.pool = t.pool(paths);
    tp.pool_next = t.base + kTapePoolNext;
    tp.pool_chunks = tape_pool_chunks(paths);
    tp.cap = t.cap;
    tp.chunk = kNoChunk;
}

// Start recording a segment after `used` entries of the path's hit list.
__device__ __forceinline__ void seg_tape_begin(SegTape& tp, uint32_t used) {
    tp.off = used;
    tp.n = 0;
    tp.overflow = false;
    tp.tlast = tp.tbefore = -__int_as_float(0x7f800000);
    tp.ties = 0u;
    tp.endT = 0.f;
}

// ---------------------------------------------------------------- radiance-cache records
// At the first secondary vertex: the query (position, outgoing direction, shading normal,
// roughness, albedo as the BRDF sees them) and the path throughput that scales the cache's
// answer.  For training, the target slot holds L so far; the end of the path turns it into
// the suffix radiance per unit throughput.
__device__ __forceinline__ void cache_record(const DeviceScene& S, const Path& p, float* r) {
    V3 n, wo, alb;
    float met;
    shading_frame(S, p.cur, n, wo, alb, met);
    r[nrc::kRecFlag] = 1.f;
    r[nrc::kRecPos] = p.cur.pos.x; r[nrc::kRecPos + 1] = p.cur.pos.y; r[nrc::kRecPos + 2] = p.cur.pos.z;
    r[nrc::kRecDir] = wo.x; r[nrc::kRecDir + 1] = wo.y; r[nrc::kRecDir + 2] = wo.z;
    r[nrc::kRecNrm] = n.x; r[nrc::kRecNrm + 1] = n.y; r[nrc::kRecNrm + 2] = n.z;
    r[nrc::kRecRough] = clamp_rough(p.cur.rough);
    r[nrc::kRecAlb] = alb.x; r[nrc::kRecAlb + 1] = alb.y; r[nrc::kRecAlb + 2] = alb.z;
    r[nrc::kRecThr] = p.thr.x; r[nrc::kRecThr + 1] = p.thr.y; r[nrc::kRecThr + 2] = p.thr.z;
    r[nrc::kRecTarget] = p.L.x; r[nrc::kRecTarget + 1] = p.L.y; r[nrc::kRecTarget + 2] = p.L.z;
}

// ---------------------------------------------------------------- launch mapping
struct PixelId {
    int x, y, z;  // launch index: pixel and image (the path's random stream)
    bool ok;
    size_t pix;   // path index into the ray and output arrays
};

// Block b of the launch runs on XCD (b mod 8); give each XCD a contiguous range of tiles.
__device__ __forceinline__ PixelId pixel_of(const DeviceScene& S) {
    const int nb = S.num_blocks;
    const int b = int(blockIdx.x);
    const int per = nb / 8, rem = nb % 8;
    const int xcd = b % 8, k = b / 8;
    const int tile = (xcd < rem) ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    PixelId p;
    if (S.num_images > 0) {
        // ragged batch: the image whose tile range holds this tile (last first-tile <= tile)
        int lo = 0, hi = S.num_images - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.images[mid].z <= tile) lo = mid;
            else hi = mid - 1;
        }
        const int4 im = S.images[lo];
        const int tiles_x = (im.y + 15) / 16;
        const int r = tile - im.z;
        const int ty = r / tiles_x, tx = r - ty * tiles_x;
        p.x = tx * 16 + wave * 8 + (lane & 7);
        p.y = ty * 8 + (lane >> 3);
        p.z = lo;
        p.ok = p.x < im.y && p.y < im.x;
        p.pix = size_t(im.w) + size_t(p.y) * im.y + p.x;
        return p;
    }
    const int tpi = S.tiles_x * S.tiles_y;
    const int z = tile / tpi;
    const int r = tile - z * tpi;
    const int ty = r / S.tiles_x, tx = r - ty * S.tiles_x;
    p.x = tx * 16 + wave * 8 + (lane & 7);
    p.y = ty * 8 + (lane >> 3);
    p.z = z;
    p.ok = p.x < S.W && p.y < S.H && z < S.B;
    p.pix = (size_t(z) * S.H + p.y) * S.W + p.x;
    return p;
}

__device__ __forceinline__ Ray primary_ray(const DeviceScene& S, size_t pix) {
    const float* M = S.ray_to_world;
    const float* o = S.ray_ori + 3 * pix;
    const float* d = S.ray_dir + 3 * pix;
    const float ox = o[0], oy = o[1], oz = o[2], dx = d[0], dy = d[1], dz = d[2];
    Ray r;
    r.o = v3(M[0] * ox + M[1] * oy + M[2] * oz + M[3], M[4] * ox + M[5] * oy + M[6] * oz + M[7],
             M[8] * ox + M[9] * oy + M[10] * oz + M[11]);
    r.d = v3(M[0] * dx + M[1] * dy + M[2] * dz, M[4] * dx + M[5] * dy + M[6] * dz, M[8] * dx + M[9] * dy + M[10] * dz);
    return r;
}

__device__ __forceinline__ void path_init(Path& p) {
    p.active = 1u;
    p.numBounces = 0u;
    emit_clear(p);
    p.L = p.dL = p.direct = p.indirect = p.Lnb = p.dLnb = v3s(0.f);
    p.thr = v3s(1.f);
    p.pvalid = false;
    p.pbounces = 0u;
    p.popacity = p.pmaxHit = 0.f;
    p.pdir = v3s(0.f);
    p.cseg = p.ccnt = p.pseg = p.pcnt = 0u;
    p.cur.reset(Ray{v3s(0.f), v3s(0.f)}, 0.f);
}


// referenceOptix.cu:27-71: one path from its primary ray to its end (the body of k_trace_fwd).
// wstack: the wave's node stack for the wave-coherent primary traversal, or null.
template <int K, int SD>
__device__ __forceinline__ void fwd_path(const DeviceScene& S, const FwdOut& out, const Lane& L, uint32_t* wstack,
                                         const PixelId& px) {
    const size_t pix = px.pix;
    const size_t paths = size_t(S.B) * S.H * S.W;
    const bool rec = out.tape.base != nullptr;
    uint32_t* prec = rec ? out.tape.path(pix) : nullptr;
    if (rec && pix == 0) {
        uint32_t* h = out.tape.base;
        h[0] = kTapeMagic; h[1] = kTapeVersion; h[2] = uint32_t(S.B); h[3] = uint32_t(S.H); h[4] = uint32_t(S.W);
        h[5] = out.tape.segs; h[6] = S.n; h[7] = S.frame; h[8] = out.tape.cap; h[9] = tape_pool_chunks(paths);
    }
    uint32_t flags = 0u;
    SegTape tp;
    tp.hits = nullptr;
    float* crec = out.cache ? out.cache + size_t(nrc::kRecFloats) * pix : nullptr;
    if (crec) crec[nrc::kRecFlag] = 0.f;
    Rng rng;
    rng.seed(Rng::launch_seed(uint32_t(px.x), uint32_t(px.y), uint32_t(px.z), S.frame));
    Path p;
    path_init(p);
    count(S, C_PATHS);
    uint32_t used = 0u;  // tape hit entries used by this path
    if (rec) {
        path_tape_begin(tp, out.tape, pix, paths);
        seg_tape_begin(tp, used);
    }
    ray_intersect<K, SD>(S, primary_ray(S, pix), p.cur, rng, false, true, true, out.vis, L, tp,
                     S.wave_traversal ? wstack : nullptr);
    if (rec) {
        store_seg(prec, 0, p.cur, tp);
        flags |= tp.overflow ? kTapeOverflow : 0u;
        used += tp.n;
    }
    {
        const Seg& c = p.cur;
        float* o;
        o = out.rad + 3 * pix; o[0] = c.radiance.x; o[1] = c.radiance.y; o[2] = c.radiance.z;
        out.dns[pix] = 1.f - c.T;
        o = out.hit + 2 * pix; o[0] = c.hitDist; o[1] = c.lastHit;
        out.hit2[pix] = c.hitDist2;
        out.dist[pix] = c.distortion;
        o = out.nrm + 3 * pix; o[0] = c.normal.x; o[1] = c.normal.y; o[2] = c.normal.z;
        o = out.snrm + 3 * pix; o[0] = c.sn.x; o[1] = c.sn.y; o[2] = c.sn.z;
        o = out.mat + 5 * pix; o[0] = c.alb.x; o[1] = c.alb.y; o[2] = c.alb.z; o[3] = c.rough; o[4] = c.metal;
        out.hits[pix] = c.hits;
    }
    if (p.cur.ivalid) {
        if (S.enable_mis) sample_nee<K, SD>(S, p, rng, L);
        if (rec) {
            prec[1] = uint32_t(rng.state);
            prec[2] = uint32_t(rng.state >> 32);
            float4* e = reinterpret_cast<float4*>(prec + 4);
            e[0] = make_float4(p.eray.d.x, p.eray.d.y, p.eray.d.z, p.elightPdf);
            e[1] = make_float4(p.escatterPdf, p.econtrib.x, p.econtrib.y, p.econtrib.z);
        }
        for (unsigned depth = 0; depth < S.max_bounces && p.active; ++depth) {
            accumulate(S, p);
            p.active &= (depth + 1u < S.max_bounces) && p.cur.ivalid;
            if (!p.active) break;
            p.numBounces = depth + 1u;
            if (crec && p.numBounces == 2u) {
                cache_record(S, p, crec);
                if (S.cache_mode == 1u) break;  // the cache stands in for the rest of the path
            }
            sample_next(S, p, rng);
            if (vmax(p.thr) < 1e-4f) break;
            if (!russian_roulette(S, p, rng)) break;
            if (rec) seg_tape_begin(tp, used);
            ray_intersect<K, SD>(S, p.cur.ray, p.cur, rng, true, false, true, out.vis, L, tp);
            if (rec) {
                store_seg(prec, p.numBounces, p.cur, tp);
                flags |= tp.overflow ? kTapeOverflow : 0u;
                used += tp.n;
            }
        }
    }
    if (rec) {
        prec[0] = flags;
        prec[3] = tp.chunk;
    }
    if (crec && S.cache_mode == 2u && crec[nrc::kRecFlag] != 0.f) {
        // training target: the radiance the rest of the path carried, per unit throughput at the vertex
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

static inline void tiling(DeviceScene& S) {
    if (S.num_images > 0) return;  // ragged: num_blocks is the table's tile count
    S.tiles_x = (S.W + 15) / 16;
    S.tiles_y = (S.H + 7) / 8;
    S.num_blocks = S.tiles_x * S.tiles_y * S.B;
}

// Kernel instances: K = 8 and 16 with the short or the deep stack, K = 32 (whose LDS
// area is 64 words anyway) with the deep one.
template <class F>
static inline hipError_t dispatch(int kbuf, bool short_stack, F&& f) {
    switch (kbuf) {
    case 8: short_stack ? f.template run<8, kStackShort>() : f.template run<8, kStackDeep>(); break;
    case 32: f.template run<32, kStackDeep>(); break;
    default: short_stack ? f.template run<16, kStackShort>() : f.template run<16, kStackDeep>(); break;
    }
    return hipGetLastError();
}

}  // namespace ptir
