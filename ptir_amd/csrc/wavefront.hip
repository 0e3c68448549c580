// Wavefront schedule of the forward: kernels and launch.  The per-path stage code, and what
// the schedule is, are in wavefront_path.hpp.
#include "wavefront_path.hpp"

namespace ptir {

template <int K, int SD>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(waves_per_eu(K, SD))))
k_wf_primary(DeviceScene S, FwdOut out, wf::IO io, const float4* __restrict__ nodes, const float4* __restrict__ wnodes,
          const float4* __restrict__ grec) {
    S.nodes = nodes;
    S.wnodes = wnodes;
    S.grec = grec;
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    __shared__ uint32_t wst[(kTraceBlock / 64) * kStackDeep];
    const Lane L{lds + threadIdx.x};
    uint32_t* wstack = wst + (threadIdx.x / 64) * kStackDeep;
    const PixelId px = pixel_of(S);
    if (!px.ok) return;
    Path p;
    Rng rng;
    wf::Carry c;
    const bool go = wf::primary_path<K, SD>(S, out, L, wstack, px, p, rng, c);
    const int st = S.enable_mis ? 0 : 1;
    const uint32_t slot = wf::push(io, st, go, uint32_t(px.pix));
    if (go) wf::save(wf::slot_of(io, st, slot), p, rng, c);
}

// NEE, part 1 (persistent: every wave takes 64 queue entries at a time until the queue is
// exhausted): the environment sample of each queued path.  A path with a shadow ray to test
// becomes a job of the occlusion engine (its state stays in its stage-0 slot); the others get
// their NEE record and go on at once.
__global__ void __launch_bounds__(kTraceBlock) k_wf_nee_setup(DeviceScene S, FwdOut out, wf::IO io) {
    const uint32_t* q = io.queue(0);
    const uint32_t n = io.ctrl[0];
    for (;;) {
        const uint32_t base = wf::take(io, 0);
        if (base >= n) break;
        const uint32_t i = base + __lane_id();
        uint32_t pix = 0u;
        bool shadow = false, go = false;
        Path p;
        Rng rng;
        wf::Carry c;
        eng::ShadowJob job;
        if (i < n) {
            pix = q[i];
            wf::load(wf::slot_of(io, 0, i), p, rng, c);
            shadow = wf::nee_setup(S, p, rng, i, job);
            if (!shadow) go = wf::nee_done(S, out, pix, p, rng, c);
        }
        const uint32_t js = eng::fetch(io.jctl, shadow);  // append the job
        if (shadow) io.jobs[js] = job;
        const uint32_t slot = wf::push(io, 1, go, pix);
        if (go) wf::save(wf::slot_of(io, 1, slot), p, rng, c);
    }
}

// NEE, part 2: the stochastic shadow rays of the jobs on the traversal engine (persistent).  The
// engine keeps a walk, a ray and a job per lane live across its loop (about 170 VGPRs at K = 16):
// two waves per SIMD without spills.
template <int K, int SD>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_wf_occlusion(DeviceScene S, wf::IO io, const float4* __restrict__ nodes, const float4* __restrict__ wnodes,
          const float4* __restrict__ grec) {
    S.nodes = nodes;
    S.wnodes = wnodes;
    S.grec = grec;
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    const Lane L{lds + threadIdx.x};
    eng::occlusion<K, SD>(S, io.jobs, io.jctl[0], io.jctl + 1, L);
}

// NEE, part 3 (persistent): each job's result into its path, the NEE record, and the path on to
// stage 1.
__global__ void __launch_bounds__(kTraceBlock) k_wf_nee_finish(DeviceScene S, FwdOut out, wf::IO io) {
    const uint32_t* q = io.queue(0);
    const uint32_t n = io.jctl[0];
    for (;;) {
        const uint32_t base = wf::take_at(io.jctl + 2);
        if (base >= n) break;
        const uint32_t j = base + __lane_id();
        uint32_t pix = 0u;
        bool go = false;
        Path p;
        Rng rng;
        wf::Carry c;
        if (j < n) {
            const eng::ShadowJob job = io.jobs[j];
            const uint32_t i = __float_as_uint(job.d.y);
            pix = q[i];
            wf::load(wf::slot_of(io, 0, i), p, rng, c);
            wf::nee_apply(p, rng, job);
            go = wf::nee_done(S, out, pix, p, rng, c);
        }
        const uint32_t slot = wf::push(io, 1, go, pix);
        if (go) wf::save(wf::slot_of(io, 1, slot), p, rng, c);
    }
}

template <int K, int SD>
__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(waves_per_eu(K, SD))))
k_wf_bounce(DeviceScene S, FwdOut out, wf::IO io, uint32_t b) {
    __shared__ uint32_t lds[LdsLayout<K, SD>::kWords * kTraceBlock];
    const Lane L{lds + threadIdx.x};
    const int st = int(b);
    const uint32_t* q = io.queue(st);
    const uint32_t n = io.ctrl[2 * st];
    for (;;) {
        const uint32_t base = wf::take(io, st);
        if (base >= n) break;
        const uint32_t i = base + __lane_id();
        uint32_t pix = 0u;
        bool go = false;
        Path p;
        Rng rng;
        wf::Carry c;
        if (i < n) {
            pix = q[i];
            wf::load(wf::slot_of(io, st, i), p, rng, c);
            go = wf::bounce_path<K, SD>(S, out, L, pix, b, p, rng, c);
        }
        const uint32_t slot = wf::push(io, st + 1, go, pix);
        if (go) wf::save(wf::slot_of(io, st + 1, slot), p, rng, c);
    }
}

// ---------------------------------------------------------------- launch
uint64_t wavefront_bytes(uint64_t paths, uint32_t max_bounces) {
    const uint64_t stages = uint64_t(max_bounces > 1 ? max_bounces : 1) + 1;
    // states, queues, control words, then the shadow jobs (16-byte aligned)
    return 4ull * (2ull * uint64_t(wf::kWords) * paths + 2ull * paths + 2ull * stages + 64ull) + 16ull +
           sizeof(eng::ShadowJob) * paths;
}

struct WfPrimary {
    const DeviceScene& S;
    const FwdOut& out;
    const wf::IO& io;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_wf_primary<K, SD><<<S.num_blocks, kTraceBlock, 0, s>>>(S, out, io, S.nodes, S.wnodes, S.grec);
    }
};
struct WfOcclusion {
    const DeviceScene& S;
    const wf::IO& io;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_wf_occlusion<K, SD><<<wf::kBlocks, kTraceBlock, 0, s>>>(S, io, S.nodes, S.wnodes, S.grec);
    }
};
struct WfBounce {
    const DeviceScene& S;
    const FwdOut& out;
    const wf::IO& io;
    uint32_t b;
    hipStream_t s;
    template <int K, int SD>
    void run() const {
        k_wf_bounce<K, SD><<<wf::kBlocks, kTraceBlock, 0, s>>>(S, out, io, b);
    }
};

hipError_t launch_trace_fwd_wavefront(const DeviceScene& S0, const FwdOut& out, void* work, int kbuf, bool short_stack,
                                      hipStream_t s) {
    DeviceScene S = S0;
    tiling(S);
    if (S.num_blocks == 0) return hipSuccess;
    const size_t P = size_t(S.B) * S.H * S.W;
    const uint32_t stages = (S.max_bounces > 1 ? S.max_bounces : 1) + 1;
    wf::IO io;
    io.P = P;
    io.sa = static_cast<float*>(work);
    io.sb = io.sa + size_t(wf::kWords) * P;
    io.qa = reinterpret_cast<uint32_t*>(io.sb + size_t(wf::kWords) * P);
    io.qb = io.qa + P;
    io.ctrl = io.qb + P;
    io.jctl = io.ctrl + 2ull * stages;
    {
        const uintptr_t j = reinterpret_cast<uintptr_t>(io.ctrl + 2ull * stages + 64ull);
        io.jobs = reinterpret_cast<eng::ShadowJob*>((j + 15u) & ~uintptr_t(15u));
    }
    hipError_t e = fill_zero(io.ctrl, (2ull * stages + 4ull) * sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    e = dispatch(kbuf, short_stack, WfPrimary{S, out, io, s});
    if (e == hipSuccess && S.enable_mis) {
        k_wf_nee_setup<<<wf::kBlocks, kTraceBlock, 0, s>>>(S, out, io);
        e = hipGetLastError();
        if (e == hipSuccess) e = dispatch(kbuf, short_stack, WfOcclusion{S, io, s});
        if (e == hipSuccess) {
            k_wf_nee_finish<<<wf::kBlocks, kTraceBlock, 0, s>>>(S, out, io);
            e = hipGetLastError();
        }
    }
    for (uint32_t b = 1; e == hipSuccess && b < S.max_bounces; ++b) e = dispatch(kbuf, short_stack, WfBounce{S, out, io, b, s});
    return e;
}

}  // namespace ptir
