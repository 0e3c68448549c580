// Forward and path-replay backward kernels of the differentiable Gaussian path tracer
// (referenceOptix.cu:27-71, referenceBwdOptix.cu:27-74).  The device code they share with
// the exact-adjoint kernel (traversal, segments, NEE, BRDF, environment, launch mapping)
// is in trace_path.hpp.
#include "trace_path.hpp"

namespace ptir {

// referenceOptix.cu:27-71
template <int K, int SD>
// nodes and grec come again as restrict-qualified arguments: the kernel never writes the
// hierarchy or the records, and saying so lets wave-uniform reads of them become scalar loads
// With the short stack the LDS allows four waves per SIMD; asking the compiler for four
// (at most 128 VGPRs, a few spilled) is faster than two waves without spills.
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(waves_per_eu(K, SD))))
k_trace_fwd(DeviceScene S, FwdOut out,
                                                          const float4* __restrict__ nodes,
                                                          const float4* __restrict__ wnodes,
                                                          const float4* __restrict__ grec) {
    S.nodes = nodes;
    S.wnodes = wnodes;
    S.grec = grec;
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    __shared__ uint32_t wst[(kTraceBlock / 64) * kStackDeep];  // wave-uniform node stacks
    const Lane L{lds + threadIdx.x};
    uint32_t* wstack = wst + (threadIdx.x / 64) * kStackDeep;
    const PixelId px = pixel_of(S);
    if (!px.ok) return;
    fwd_path<K, SD>(S, out, L, wstack, px);
}

// One path of the backward (the kernel below).
template <int K, int SD, bool TAPE>
__device__ __forceinline__ void bwd_path(const DeviceScene& S, const BwdIO& io, const Lane& L, MatTable* tab) {
    const PixelId px = pixel_of(S);
    if (!px.ok) return;
    const size_t pix = px.pix;
    const size_t paths = size_t(S.B) * S.H * S.W;
    const uint32_t* prec = io.tape.base ? io.tape.path(pix) : nullptr;
    if constexpr (TAPE) {
        const uint32_t* h = io.tape.base;
        if (h[0] != kTapeMagic || h[1] != kTapeVersion || h[2] != uint32_t(S.B) || h[3] != uint32_t(S.H) ||
            h[4] != uint32_t(S.W) || h[5] != io.tape.segs || h[6] != S.n || h[7] != S.frame || h[8] != io.tape.cap ||
            h[9] != tape_pool_chunks(paths)) {
            atomicOr(io.tape_state, 1);  // the replay launch that follows handles every path
            return;
        }
        if (prec[0] & kTapeOverflow) return;
    } else {
        if (io.replay_mode == 1 && *io.tape_state == 0 && !(prec[0] & kTapeOverflow)) return;
    }
    const Flush f{&io, pix, paths, &L, prec ? prec[3] : kNoChunk, tab};
    Rng rng;
    rng.seed(Rng::launch_seed(uint32_t(px.x), uint32_t(px.y), uint32_t(px.z), S.frame));
    Path p;
    path_init(p);
    count(S, C_PATHS);
    p.L = v3(io.pbr[3 * pix], io.pbr[3 * pix + 1], io.pbr[3 * pix + 2]);
    p.dL = v3(io.pbr_grad[3 * pix], io.pbr_grad[3 * pix + 1], io.pbr_grad[3 * pix + 2]);
    p.Lnb = v3(io.light[3 * pix], io.light[3 * pix + 1], io.light[3 * pix + 2]);
    if (io.light_grad) p.dLnb = v3(io.light_grad[3 * pix], io.light_grad[3 * pix + 1], io.light_grad[3 * pix + 2]);
    if constexpr (TAPE) {
        p.ccnt = load_seg(prec, 0, primary_ray(S, pix), p.cur, p.cseg);
        count(S, C_SEGMENTS);
    } else {
        SegTape none;
        none.hits = nullptr;
        ray_intersect<K, SD>(S, primary_ray(S, pix), p.cur, rng, false, true, false, nullptr, L, none);
    }
    if (!p.cur.ivalid) return;
    if (io.mat_grad) {
        const float* mg = io.mat_grad + 5 * pix;
        p.cur.galb += v3(mg[0], mg[1], mg[2]);
        p.cur.grough += mg[3];
        if (S.enable_metallic) p.cur.gmetal += mg[4];
    }
    if constexpr (TAPE) {
        // the NEE record and the RNG state after it (shadow-ray draws included)
        rng.state = uint64_t(prec[1]) | (uint64_t(prec[2]) << 32);
        if (S.enable_mis) {
            const float4* e = reinterpret_cast<const float4*>(prec + 4);
            const float4 e0 = e[0], e1 = e[1];
            p.eray.d = v3(e0.x, e0.y, e0.z);
            p.elightPdf = e0.w;
            p.escatterPdf = e1.x;
            p.econtrib = v3(e1.y, e1.z, e1.w);
        }
    } else {
        if (S.enable_mis) sample_nee<K, SD>(S, p, rng, L);
    }
    for (unsigned depth = 0; depth < S.max_bounces && p.active; ++depth) {
        accumulate_bwd(S, p, io.denv);
        pending_bwd<K, SD, TAPE>(S, p, f);
        p.active &= (depth + 1u < S.max_bounces) && p.cur.ivalid;
        if (!p.active) break;
        p.numBounces = depth + 1u;
        if (S.cache_mode == 1u && p.numBounces == 2u) break;  // the forward ended here with a cache query
        sample_next_bwd<K, SD, TAPE>(S, p, rng, f);
        if (vmax(p.thr) < 1e-4f) break;
        if (!russian_roulette(S, p, rng)) break;
        if constexpr (TAPE) {
            p.ccnt = load_seg(prec, p.numBounces, p.cur.ray, p.cur, p.cseg);
            count(S, C_SEGMENTS);
            if (!p.cur.ivalid) (void)rng.next_1d();  // the forward's escape draw
        } else {
            SegTape none;
            none.hits = nullptr;
            ray_intersect<K, SD>(S, p.cur.ray, p.cur, rng, true, false, false, nullptr, L, none);
        }
    }
}

// referenceBwdOptix.cu:27-74.  TAPE = false replays the path by re-tracing every
// segment; TAPE = true reads the segments and their hits from the forward's tape and gathers
// the workgroup's material gradients in LDS (MatTable) before they reach memory.
template <int K, int SD, bool TAPE>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(waves_per_eu(K, TAPE ? 0 : SD))))
k_trace_bwd(DeviceScene S, BwdIO io) {
    __shared__ uint32_t lds[(TAPE ? 1 : LdsLayout<K, SD>::kWords) * kTraceBlock];
    const Lane L{lds + threadIdx.x};
    if constexpr (TAPE) {
        __shared__ MatTable tab;
        mat_table_clear(tab);
        bwd_path<K, SD, TAPE>(S, io, L, &tab);
        mat_table_flush(tab, io.dmat_leaf);
    } else {
        bwd_path<K, SD, TAPE>(S, io, L, nullptr);
    }
}

struct FwdLaunch {
    const DeviceScene& S;
    const FwdOut& out;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_trace_fwd<K, SD><<<S.num_blocks, kTraceBlock, 0, s>>>(S, out, S.nodes, S.wnodes, S.grec);
    }
};

struct BwdLaunch {
    const DeviceScene& S;
    const BwdIO& io;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_trace_bwd<K, SD, false><<<S.num_blocks, kTraceBlock, 0, s>>>(S, io);
    }
};

hipError_t launch_trace_fwd(const DeviceScene& S0, const FwdOut& out, int kbuf, bool short_stack, hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    return dispatch(kbuf, short_stack, FwdLaunch{S, out, s});
}

hipError_t launch_trace_bwd(const DeviceScene& S0, const BwdIO& io, int kbuf, bool short_stack, hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    BwdIO r = io;
    if (io.tape.base) {
        // tape pass for every path it covers, then a re-trace pass for the rest
        k_trace_bwd<16, kStackDeep, true><<<S.num_blocks, kTraceBlock, 0, s>>>(S, io);
        r.replay_mode = 1;
    } else {
        r.replay_mode = 0;
    }
    return dispatch(kbuf, short_stack, BwdLaunch{S, r, s});
}

}  // namespace ptir
