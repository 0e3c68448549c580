// C ABI of the tracer (declared in include/ptir_amd.h).
#include "../../include/ptir_amd.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <algorithm>
#include <string>
#include <vector>

#include "internal.hpp"

struct ptir_tracer {
    ptir_options opts;
    ptir::BvhBuffers bvh;
    bool built = false;
    bool short_stack = false;  // the built hierarchy fits the short traversal stack
    int32_t stack_need = 0;    // traversal stack entries the built hierarchy needs
    int32_t wide_nodes = 0;    // 4-wide nodes of the built hierarchy
    void* wf = nullptr;        // wavefront-schedule workspace
    uint64_t wf_bytes = 0;
    ptir::GrtRecBuffers grt_rec;  // hit records of the 3DGRT backward
    std::string err;
};

namespace {

thread_local std::string g_create_error;

// Device launch of the exact backward (see ptir_trace_bwd_exact).
constexpr bool kExactAdjointOnDevice = false;

ptir_status fail(ptir_tracer* t, ptir_status st, const std::string& msg) {
    if (t) t->err = msg;
    else g_create_error = msg;
    return st;
}

ptir_status hip_fail(ptir_tracer* t, hipError_t e, const char* where) {
    return fail(t, PTIR_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

ptir::KernelOpts kopts(const ptir_options& o) {
    return ptir::KernelOpts{int(o.kernel_degree), o.min_response, o.density_clamping};
}

bool valid_degree(float d) {
    const int i = int(d);
    return float(i) == d && (i == 0 || i == 1 || i == 2 || i == 3 || i == 4 || i == 5 || i == 8);
}

ptir_status check_scene(ptir_tracer* t, const ptir_scene* sc, const ptir_launch* la) {
    if (!sc || !la) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null scene or launch");
    if (!t->built) return fail(t, PTIR_ERR_NOT_BUILT, "trace called before build_bvh");
    if (sc->num_gaussians != t->bvh.n)
        return fail(t, PTIR_ERR_NOT_BUILT, "number of Gaussians differs from the last build_bvh (" +
                                               std::to_string(sc->num_gaussians) + " vs " + std::to_string(t->bvh.n) + ")");
    if (sc->num_lights || sc->num_light_entries || sc->num_mesh_lights)
        return fail(t, PTIR_ERR_UNSUPPORTED, "point, sphere and mesh lights are not supported; pass empty light tensors");
    if (sc->num_gaussians && (!sc->particle_density || !sc->particle_material || !sc->particle_radiance))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null particle tensor");
    if (la->num_images < 0) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "negative num_images");
    if (la->num_images > 0 && (!la->images || la->num_tiles <= 0 || la->num_paths <= 0))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "a ragged batch needs its image table and positive tile and path counts");
    if (la->num_images == 0 && (la->B < 0 || la->H < 0 || la->W < 0))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "negative ray shape");
    const long long paths = la->num_images > 0 ? (long long)la->num_paths : (long long)la->B * la->H * la->W;
    if (paths > 0 && (!la->ray_ori || !la->ray_dir || !la->ray_to_world))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null ray tensor");
    if (paths > (1ll << 31) - 1) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "more than 2^31-1 paths in one launch");
    if (la->render_opts > 1u) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "render_opts must be 0 or 1");
    if (la->cache_mode > 2u) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "cache_mode must be 0, 1 or 2");
    if (la->cache_mode && (la->max_bounces < 3u || la->render_opts != 0u))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "cache_mode 1 and 2 need max_bounces >= 3 and render_opts 0");
    if (la->sph_degree < 0 || la->sph_degree > 3) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "sph_degree must be in [0,3]");
    const int ncoef = (la->sph_degree + 1) * (la->sph_degree + 1);
    if (sc->sh_stride < 3 * ncoef)
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "particle_radiance has fewer coefficients than sph_degree needs");
    if (sc->environment && (sc->env_h <= 0 || sc->env_w <= 0))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "environment with a non-positive size");
    if (sc->environment && sc->env_h == 6 * sc->env_w)
        return fail(t, PTIR_ERR_UNSUPPORTED, "cube-map environments ([6N,N,4]) are not supported");
    if (sc->env_table && (sc->tab_h <= 0 || sc->tab_w <= 0))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "environment table with a non-positive size");
    if (sc->tab_kind != PTIR_ENV_TABLE_ALIAS && sc->tab_kind != PTIR_ENV_TABLE_CDF)
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "tab_kind must be PTIR_ENV_TABLE_ALIAS or PTIR_ENV_TABLE_CDF");
    return PTIR_OK;
}

ptir::DeviceScene make_scene(ptir_tracer* t, const ptir_scene* sc, const ptir_launch* la) {
    ptir::DeviceScene S{};
    const auto& b = t->bvh;
    S.grec = b.grec;
    S.srec = b.srec;
    S.prec = b.prec;
    S.nodes = b.nodes;
    S.wnodes = b.wide;
    S.bounds = b.bounds;
    S.sh = sc->particle_radiance;
    S.env = reinterpret_cast<const float4*>(sc->environment);
    S.env_w = sc->environment ? sc->env_w : 0;
    S.env_h = sc->environment ? sc->env_h : 0;
    S.tab = sc->env_table;
    S.tab_w = sc->env_table ? sc->tab_w : 0;
    S.tab_h = sc->env_table ? sc->tab_h : 0;
    S.tab_kind = sc->tab_kind;
    S.ray_to_world = la->ray_to_world;
    S.ray_ori = la->ray_ori;
    S.ray_dir = la->ray_dir;
    S.n = b.n;
    S.root = b.n > 1 ? 0 : ~0;
    S.sh_stride = sc->sh_stride;
    S.sph_degree = la->sph_degree;
    S.kernel_degree = int(t->opts.kernel_degree);
    S.min_response = t->opts.min_response;
    S.min_transmittance = la->min_transmittance;
    // OptixTracer::trace: maxBounces = renderOpts ? min(maxBounces, 2) : maxBounces (optixTracer.cpp:959)
    S.max_bounces = la->render_opts ? (la->max_bounces < 2u ? la->max_bounces : 2u) : la->max_bounces;
    S.render_opts = la->render_opts;
    S.cache_mode = la->cache_mode;
    S.frame = la->frame;
    if (la->num_images > 0) {
        S.B = 1;
        S.H = 1;
        S.W = la->num_paths;
        S.images = reinterpret_cast<const int4*>(la->images);
        S.num_images = la->num_images;
        S.num_blocks = la->num_tiles;
    } else {
        S.B = la->B;
        S.H = la->H;
        S.W = la->W;
        S.images = nullptr;
        S.num_images = 0;
    }
    S.rr_start = la->rr_start;
    S.enable_normals = t->opts.enable_normals;
    S.enable_hitcounts = t->opts.enable_hitcounts;
    S.enable_mis = t->opts.enable_mis;
    S.enable_metallic = t->opts.enable_metallic;
    S.has_snormal = sc->shading_normal != nullptr;
    S.wide_walks = t->opts.wide_walks;
    S.wave_traversal = t->opts.wave_traversal;
    S.counters = nullptr;
    S.error_flag = b.error_flag;
    return S;
}

// The traversal-stack error flag is sticky: a kernel sets it, and it is cleared only by a
// rebuild of the hierarchy and by the call that reports it (a counting launch or ptir_check).
ptir_status begin_counters(ptir_tracer* t, ptir::DeviceScene& S, ptir_counters* c, hipStream_t s) {
    if (!c) return PTIR_OK;
    hipError_t e = ptir::fill_zero(t->bvh.counters, size_t(ptir::kCounterCopies) * ptir::kCounterStride * sizeof(unsigned long long), s);
    if (e != hipSuccess) return hip_fail(t, e, "zeroing counters");
    S.counters = t->bvh.counters;
    return PTIR_OK;
}

ptir_status end_counters(ptir_tracer* t, ptir_counters* c, hipStream_t s) {
    if (!c) return PTIR_OK;
    std::vector<unsigned long long> all(size_t(ptir::kCounterCopies) * ptir::kCounterStride);
    int32_t flag = 0;
    hipError_t e = hipMemcpyAsync(all.data(), t->bvh.counters, all.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, t->bvh.error_flag, sizeof(flag), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(t, e, "reading counters");
    unsigned long long h[ptir::C_COUNT] = {};
    for (int k = 0; k < ptir::kCounterCopies; ++k)
        for (int i = 0; i < ptir::C_COUNT; ++i) h[i] += all[size_t(k) * ptir::kCounterStride + i];
    c->paths = h[ptir::C_PATHS];
    c->segments = h[ptir::C_SEGMENTS];
    c->candidates = h[ptir::C_CANDIDATES];
    c->accepted = h[ptir::C_ACCEPTED];
    c->accepted_primary = h[ptir::C_ACCEPTED_PRIMARY];
    c->shadow_candidates = h[ptir::C_SHADOW_CANDIDATES];
    c->env_lookups = h[ptir::C_ENV_LOOKUPS];
    c->mat_atomics = h[ptir::C_MAT_ATOMICS];
    c->env_atomics = h[ptir::C_ENV_ATOMICS];
    c->traversals = h[ptir::C_TRAVERSALS];
    c->nodes = h[ptir::C_NODES];
    c->leaf_tests = h[ptir::C_LEAF_TESTS];
    if (flag) {
        e = ptir::fill_zero(t->bvh.error_flag, sizeof(int32_t), s);
        if (e != hipSuccess) return hip_fail(t, e, "clearing the error flag");
        return fail(t, PTIR_ERR_BVH_DEPTH, "traversal stack overflow");
    }
    return PTIR_OK;
}

}  // namespace

namespace ptir {
// Error text of a failed call that has no handle (ptir_create, the environment table
// builders); ptir_last_error(NULL) returns it.
void set_thread_error(const std::string& msg) { g_create_error = msg; }
}  // namespace ptir

extern "C" {

ptir_status ptir_create(const ptir_options* options, ptir_tracer** out) {
    if (!options || !out) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null options or output handle");
    *out = nullptr;
    if (!valid_degree(options->kernel_degree))
        return fail(nullptr, PTIR_ERR_UNSUPPORTED, "kernel_degree must be one of 0,1,2,3,4,5,8");
    if (!(options->min_response > 0.f && options->min_response < 1.f))
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "min_response must be in (0,1)");
    if (options->sph_max_degree < 0 || options->sph_max_degree > 3)
        return fail(nullptr, PTIR_ERR_UNSUPPORTED, "particle_radiance_sph_degree must be in [0,3]");
    if (options->visualize_lights) return fail(nullptr, PTIR_ERR_UNSUPPORTED, "visualize_lights is not supported");
    if (options->kbuffer != 8 && options->kbuffer != 16 && options->kbuffer != 32)
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "kbuffer must be 8, 16 or 32");
    if (options->wave_traversal != 0 && options->wave_traversal != 1)
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "wave_traversal must be 0 or 1");
    if (options->wide_walks < 0 || options->wide_walks > 3)
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "wide_walks must be in [0,3]");
    if (options->schedule != 0 && options->schedule != 1)
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "schedule must be 0 (one lane per path) or 1 (wavefront)");
    if (options->builder != 0 && options->builder != 1)
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "builder must be 0 (LBVH) or 1 (PLOC)");
    ptir_tracer* t = new (std::nothrow) ptir_tracer();
    if (!t) return fail(nullptr, PTIR_ERR_OUT_OF_MEMORY, "host allocation failed");
    t->opts = *options;
    *out = t;
    return PTIR_OK;
}

ptir_status ptir_destroy(ptir_tracer* t) {
    if (!t) return PTIR_OK;
    ptir::bvh_free(t->bvh);
    if (t->wf) (void)hipFree(t->wf);
    ptir::grt_rec_free(t->grt_rec);
    delete t;
    return PTIR_OK;
}

ptir_status ptir_build_bvh(ptir_tracer* t, const float* pos, const float* rot, const float* scl, const float* dns,
                           uint32_t n, uint32_t rebuild, int32_t allow_update, void* stream) {
    (void)allow_update;
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    if (n > 0 && (!pos || !rot || !scl || !dns)) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null Gaussian tensor");
    if (n >= (1u << 30)) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "too many Gaussians");
    if (!rebuild && (!t->built || t->bvh.n != n))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "refit requested but the hierarchy was built for a different N");
    hipStream_t s = static_cast<hipStream_t>(stream);
    t->built = false;
    if (n == 0) {
        t->bvh.n = 0;
        t->built = true;
        return PTIR_OK;
    }
    hipError_t e = ptir::bvh_build(t->bvh, kopts(t->opts), pos, rot, scl, dns, n, rebuild != 0, t->opts.builder, s);
    if (e != hipSuccess) return hip_fail(t, e, "bvh_build");
    if (rebuild) {
        int32_t need = 0, wide_count = 0, depth = 0;
        e = ptir::fill_zero(t->bvh.error_flag, sizeof(int32_t), s);
        if (e == hipSuccess) e = ptir::bvh_stack_need(t->bvh, s, &need, &wide_count, &depth);
        if (e != hipSuccess) return hip_fail(t, e, "bvh_stack_need");
        // a binary walk pushes at most one entry per level
        need = std::max(need, depth + 1);
        if (need > ptir::kStackDeep)
            return fail(t, PTIR_ERR_BVH_DEPTH, "the hierarchy's walks need " + std::to_string(need) +
                                                   " stack entries, more than the traversal stack holds");
        t->short_stack = need <= ptir::kStackShort;
        t->stack_need = need;
        t->wide_nodes = n > 1 ? wide_count : 0;
    }
    t->built = true;
    return PTIR_OK;
}

ptir_status ptir_check(ptir_tracer* t, void* stream) {
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    if (!t->bvh.error_flag) return PTIR_OK;  // nothing built, nothing traced
    hipStream_t s = static_cast<hipStream_t>(stream);
    int32_t flag = 0;
    hipError_t e = hipMemcpyAsync(&flag, t->bvh.error_flag, sizeof(flag), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(t, e, "reading the error flag");
    if (!flag) return PTIR_OK;
    e = ptir::fill_zero(t->bvh.error_flag, sizeof(int32_t), s);
    if (e != hipSuccess) return hip_fail(t, e, "clearing the error flag");
    return fail(t, PTIR_ERR_BVH_DEPTH, "a traversal stack overflowed in a launch since the last check");
}

ptir_status ptir_hierarchy_info(ptir_tracer* t, int32_t* wide_nodes, int32_t* stack_need) {
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    if (!t->built) return fail(t, PTIR_ERR_NOT_BUILT, "no hierarchy built");
    if (wide_nodes) *wide_nodes = t->wide_nodes;
    if (stack_need) *stack_need = t->stack_need;
    return PTIR_OK;
}

ptir_status ptir_trace(ptir_tracer* t, const ptir_scene* sc, const ptir_launch* la, const ptir_fwd_outputs* out,
                       ptir_counters* counters, void* stream) {
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    ptir_status st = check_scene(t, sc, la);
    if (st != PTIR_OK) return st;
    if (!out || !out->radiance || !out->density || !out->hit_distance || !out->hit_distance2 || !out->distortion ||
        !out->normal || !out->shading_normal || !out->material || !out->hits_count || !out->pbr || !out->light ||
        !out->pbr_components || (sc->num_gaussians && !out->visibility))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null output tensor");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = ptir::prep_records(t->bvh, kopts(t->opts), sc->particle_density, sc->particle_material,
                                      sc->shading_normal, s);
    if (e != hipSuccess) return hip_fail(t, e, "prep_records");
    ptir::DeviceScene S = make_scene(t, sc, la);
    if ((st = begin_counters(t, S, counters, s)) != PTIR_OK) return st;
    ptir::FwdOut o{out->radiance, out->density, out->hit_distance, out->hit_distance2, out->distortion, out->normal,
                   out->shading_normal, out->material, out->hits_count, out->pbr, out->light, out->pbr_components,
                   reinterpret_cast<uint32_t*>(out->visibility), ptir::Tape{nullptr, 0u, 0u}, nullptr};
    if (la->cache_mode && !out->cache_records)
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "cache_mode needs cache_records");
    if (la->cache_mode) o.cache = out->cache_records;
    if (out->tape) {
        const uint64_t need = ptir::tape_bytes(uint64_t(S.B) * S.H * S.W, S.max_bounces, out->tape_hits_per_path);
        if (out->tape_hits_per_path == 0 || out->tape_hits_per_path > 0xFFFFu - ptir::kTapeChunk ||
            out->tape_bytes < need)
            return fail(t, PTIR_ERR_INVALID_ARGUMENT, "tape too small for this launch (" + std::to_string(out->tape_bytes) +
                                                         " < " + std::to_string(need) + " bytes)");
        o.tape = ptir::Tape{static_cast<uint32_t*>(out->tape), out->tape_hits_per_path, S.max_bounces};
        // the pool's chunk counter starts at zero (a kernel, so the call stays capturable)
        e = ptir::fill_zero(o.tape.base + ptir::kTapePoolNext, sizeof(uint32_t), s);
        if (e != hipSuccess) return hip_fail(t, e, "zeroing the tape pool counter");
    }
    if (t->opts.schedule == 1) {
        const uint64_t need = ptir::wavefront_bytes(uint64_t(S.B) * S.H * S.W, S.max_bounces);
        if (t->wf_bytes < need) {
            // grow the workspace; the old one may still be in use by queued work
            if (t->wf) {
                e = hipDeviceSynchronize();
                if (e == hipSuccess) e = hipFree(t->wf);
                t->wf = nullptr;
                t->wf_bytes = 0;
                if (e != hipSuccess) return hip_fail(t, e, "releasing the wavefront workspace");
            }
            e = hipMalloc(&t->wf, need);
            if (e != hipSuccess) {
                t->wf = nullptr;
                return fail(t, PTIR_ERR_OUT_OF_MEMORY, std::string("wavefront workspace: ") + hipGetErrorString(e));
            }
            t->wf_bytes = need;
        }
        e = ptir::launch_trace_fwd_wavefront(S, o, t->wf, t->opts.kbuffer, t->short_stack, s);
    } else {
        e = ptir::launch_trace_fwd(S, o, t->opts.kbuffer, t->short_stack, s);
    }
    if (e != hipSuccess) return hip_fail(t, e, "trace kernel launch");
    return end_counters(t, counters, s);
}

ptir_status ptir_trace_bwd(ptir_tracer* t, const ptir_scene* sc, const ptir_launch* la, const ptir_bwd_inputs* in,
                           const ptir_bwd_outputs* out, ptir_counters* counters, void* stream) {
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    ptir_status st = check_scene(t, sc, la);
    if (st != PTIR_OK) return st;
    if (!in || !out) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null backward inputs or outputs");
    const long long paths = la->num_images > 0 ? (long long)la->num_paths : (long long)la->B * la->H * la->W;
    if (paths > 0 && (!in->pbr || !in->light || !in->pbr_grad))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null pbr, light or pbr_grad");
    const size_t N = sc->num_gaussians;
    if (N && (!out->density_grad || !out->material_grad || !out->radiance_grad || !out->shading_normal_grad))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null gradient output");
    if (sc->environment && !out->environment_grad)
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null environment gradient");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    // the reference allocates every gradient with torch::zeros (optixTracer.cpp:1058-1063)
    if (N) {
        if (e == hipSuccess) e = ptir::fill_zero(out->density_grad, N * 12 * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(out->radiance_grad, N * size_t(sc->sh_stride) * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(out->shading_normal_grad, N * 3 * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(t->bvh.dmat_leaf, N * 5 * sizeof(float), s);
    }
    if (e == hipSuccess && sc->environment)
        e = ptir::fill_zero(out->environment_grad, size_t(sc->env_h) * sc->env_w * 4 * sizeof(float), s);
    if (e != hipSuccess) return hip_fail(t, e, "zeroing gradients");
    e = ptir::prep_records(t->bvh, kopts(t->opts), sc->particle_density, sc->particle_material, sc->shading_normal, s);
    if (e != hipSuccess) return hip_fail(t, e, "prep_records");
    ptir::DeviceScene S = make_scene(t, sc, la);
    if ((st = begin_counters(t, S, counters, s)) != PTIR_OK) return st;
    ptir::BwdIO io{in->pbr, in->light, in->pbr_grad, in->light_grad, in->material_grad, t->bvh.dmat_leaf,
                   sc->environment ? out->environment_grad : nullptr, ptir::Tape{nullptr, 0u, 0u}, t->bvh.tape_state, 0};
    if (in->tape) {
        const uint64_t need = ptir::tape_bytes(uint64_t(S.B) * S.H * S.W, S.max_bounces, in->tape_hits_per_path);
        if (in->tape_hits_per_path == 0 || in->tape_hits_per_path > 0xFFFFu - ptir::kTapeChunk || in->tape_bytes < need)
            return fail(t, PTIR_ERR_INVALID_ARGUMENT, "tape too small for this launch");
        io.tape = ptir::Tape{static_cast<uint32_t*>(const_cast<void*>(in->tape)), in->tape_hits_per_path, S.max_bounces};
        e = ptir::fill_zero(t->bvh.tape_state, sizeof(int32_t), s);
        if (e != hipSuccess) return hip_fail(t, e, "zeroing the tape state");
    }
    e = ptir::launch_trace_bwd(S, io, t->opts.kbuffer, t->short_stack, s);
    if (e != hipSuccess) return hip_fail(t, e, "trace_bwd kernel launch");
    if (N) {
        e = ptir::scatter_material_grad(t->bvh, out->material_grad, s);
        if (e != hipSuccess) return hip_fail(t, e, "scatter_material_grad");
    }
    return end_counters(t, counters, s);
}

ptir_status ptir_trace_bwd_exact(ptir_tracer* t, const ptir_scene* sc, const ptir_launch* la, const ptir_bwd_inputs* in,
                                 const ptir_exact_grads* ex, const ptir_bwd_outputs* out, void* stream) {
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    ptir_status st = check_scene(t, sc, la);
    if (st != PTIR_OK) return st;
    if (!in || !out) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null backward inputs or outputs");
    if (la->cache_mode != 0u) return fail(t, PTIR_ERR_UNSUPPORTED, "the exact backward does not support cache modes");
    // The kernel of this entry point faulted the device on its first run and the cause is not
    // found (its device code runs clean under address, undefined-behaviour and memory
    // sanitizers on the host, tests/host_emu, and matches the oracle's dual numbers there), so
    // the device launch stays switched off until it is.
    if (!kExactAdjointOnDevice)
        return fail(t, PTIR_ERR_UNSUPPORTED, "the exact backward's device kernel is switched off in this build (it "
                                             "faulted the device; its device code is verified on the host only)");
    const long long paths = la->num_images > 0 ? (long long)la->num_paths : (long long)la->B * la->H * la->W;
    if (paths > 0 && !in->pbr_grad) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null pbr_grad");
    const size_t N = sc->num_gaussians;
    if (N && (!out->density_grad || !out->material_grad || !out->radiance_grad || !out->shading_normal_grad))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null gradient output");
    if (sc->environment && !out->environment_grad)
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null environment gradient");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    if (N) {
        if (e == hipSuccess) e = ptir::fill_zero(out->density_grad, N * 12 * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(out->material_grad, N * 5 * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(out->radiance_grad, N * size_t(sc->sh_stride) * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(out->shading_normal_grad, N * 3 * sizeof(float), s);
    }
    if (e == hipSuccess && sc->environment)
        e = ptir::fill_zero(out->environment_grad, size_t(sc->env_h) * sc->env_w * 4 * sizeof(float), s);
    if (e != hipSuccess) return hip_fail(t, e, "zeroing gradients");
    if (paths == 0) return PTIR_OK;
    e = ptir::prep_records(t->bvh, kopts(t->opts), sc->particle_density, sc->particle_material, sc->shading_normal, s);
    if (e != hipSuccess) return hip_fail(t, e, "prep_records");
    ptir::DeviceScene S = make_scene(t, sc, la);
    const ptir_exact_grads none{};
    const ptir_exact_grads& x = ex ? *ex : none;
    ptir::AdjIO io{in->pbr_grad, in->light_grad, in->material_grad, x.radiance_grad, x.density_grad,
                   x.hit_distance_grad, x.shading_normal_grad, x.pbr_components_grad, sc->particle_density,
                   out->density_grad, out->material_grad, out->radiance_grad, out->shading_normal_grad,
                   sc->environment ? out->environment_grad : nullptr, nullptr};
    const uint64_t bytes = ptir::adjoint_scratch_bytes(uint64_t(paths), S.max_bounces);
    void* scratch = nullptr;
    e = hipMallocAsync(&scratch, bytes, s);
    if (e != hipSuccess) return fail(t, PTIR_ERR_OUT_OF_MEMORY, std::string("exact backward scratch: ") + hipGetErrorString(e));
    io.scratch = static_cast<float4*>(scratch);
    e = ptir::launch_trace_adjoint(S, io, t->opts.kbuffer, t->short_stack, s);
    const hipError_t ef = hipFreeAsync(scratch, s);
    if (e != hipSuccess) return hip_fail(t, e, "trace_adjoint kernel launch");
    if (ef != hipSuccess) return hip_fail(t, ef, "exact backward scratch free");
    return PTIR_OK;
}

ptir_status ptir_ragged_layout(const int32_t* hw, int32_t n, int32_t* table, int32_t* num_tiles, int32_t* num_paths) {
    if (n <= 0 || !hw || !table || !num_tiles || !num_paths)
        return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "ragged layout: bad arguments");
    long long tiles = 0, paths = 0;
    for (int i = 0; i < n; ++i) {
        const int32_t h = hw[2 * i], w = hw[2 * i + 1];
        if (h <= 0 || w <= 0) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "ragged layout: image sizes must be positive");
        table[4 * i] = h;
        table[4 * i + 1] = w;
        table[4 * i + 2] = int32_t(tiles);
        table[4 * i + 3] = int32_t(paths);
        tiles += (long long)((w + 15) / 16) * ((h + 7) / 8);
        paths += (long long)h * w;
        if (tiles > (1ll << 31) - 1 || paths > (1ll << 31) - 1)
            return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "ragged layout: more than 2^31-1 paths or tiles");
    }
    *num_tiles = int32_t(tiles);
    *num_paths = int32_t(paths);
    return PTIR_OK;
}

// ---------------------------------------------------------------- 3DGRT primary-only tracer
namespace {
// The 3DGRT scene as the path tracer's: no material, no environment, no lights.
ptir_scene grt_as_scene(const ptir_grt_scene* g) {
    ptir_scene sc{};
    sc.num_gaussians = g->num_gaussians;
    sc.particle_density = g->particle_density;
    sc.particle_material = nullptr;
    sc.particle_radiance = g->particle_radiance;
    sc.sh_stride = g->sh_stride;
    sc.shading_normal = g->shading_normal;
    sc.tab_kind = PTIR_ENV_TABLE_ALIAS;
    return sc;
}

ptir_status check_grt(ptir_tracer* t, const ptir_grt_scene* g, const ptir_launch* la) {
    if (!t) return fail(nullptr, PTIR_ERR_INVALID_ARGUMENT, "null tracer");
    if (!g || !la) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null scene or launch");
    if (!t->built) return fail(t, PTIR_ERR_NOT_BUILT, "trace called before build_bvh");
    if (g->num_gaussians != t->bvh.n)
        return fail(t, PTIR_ERR_NOT_BUILT, "number of Gaussians differs from the last build_bvh (" +
                                               std::to_string(g->num_gaussians) + " vs " + std::to_string(t->bvh.n) + ")");
    if (g->num_gaussians && (!g->particle_density || !g->particle_radiance))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null particle tensor");
    if (la->num_images != 0) return fail(t, PTIR_ERR_UNSUPPORTED, "the primary-only tracer takes [B,H,W] launches");
    if (la->B < 0 || la->H < 0 || la->W < 0) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "negative ray shape");
    const long long paths = (long long)la->B * la->H * la->W;
    if (paths > 0 && (!la->ray_ori || !la->ray_dir || !la->ray_to_world))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null ray tensor");
    if (paths > (1ll << 31) - 1) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "more than 2^31-1 paths in one launch");
    if (la->render_opts != 0u) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "render_opts must be 0");
    if (la->sph_degree < 0 || la->sph_degree > 3) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "sph_degree must be in [0,3]");
    if (g->sh_stride < 3 * (la->sph_degree + 1) * (la->sph_degree + 1))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "particle_radiance has fewer coefficients than sph_degree needs");
    return PTIR_OK;
}
}  // namespace

ptir_status ptir_grt_trace(ptir_tracer* t, const ptir_grt_scene* g, const ptir_launch* la, const ptir_grt_outputs* out,
                           void* stream) {
    ptir_status st = check_grt(t, g, la);
    if (st != PTIR_OK) return st;
    if (!out || !out->radiance || !out->density || !out->hit_distance || !out->hit_distance2 || !out->distortion ||
        !out->normal || !out->shading_normal || !out->hits_count || (g->num_gaussians && !out->visibility))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null output tensor");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = ptir::prep_records(t->bvh, kopts(t->opts), g->particle_density, nullptr, g->shading_normal, s);
    if (e != hipSuccess) return hip_fail(t, e, "prep_records");
    const ptir_scene sc = grt_as_scene(g);
    ptir::DeviceScene S = make_scene(t, &sc, la);
    S.max_bounces = 1u;
    if (out->hit_tape && (out->hit_tape_cap == 0 || out->hit_tape_cap > 0xFFFFu))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "hit_tape_cap must be in [1, 65535]");
    const ptir::grt::FwdOut o{out->radiance, out->density, out->hit_distance, out->hit_distance2, out->distortion,
                              out->normal, out->shading_normal, out->hits_count,
                              reinterpret_cast<uint32_t*>(out->visibility), out->hit_tape,
                              out->hit_tape ? out->hit_tape_cap : 0u};
    e = ptir::launch_grt_fwd(S, o, t->opts.kbuffer, t->short_stack, s);
    if (e != hipSuccess) return hip_fail(t, e, "primary-only trace launch");
    return PTIR_OK;
}

ptir_status ptir_grt_trace_bwd(ptir_tracer* t, const ptir_grt_scene* g, const ptir_launch* la,
                               const ptir_grt_bwd_inputs* in, const ptir_grt_bwd_outputs* out, void* stream) {
    ptir_status st = check_grt(t, g, la);
    if (st != PTIR_OK) return st;
    if (!in || !out) return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null backward inputs or outputs");
    const long long paths = (long long)la->B * la->H * la->W;
    if (paths > 0 && (!in->radiance || !in->density || !in->hit_distance || !in->hit_distance2 || !in->distortion ||
                      !in->shading_normal))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null forward output");
    const size_t N = g->num_gaussians;
    if (N && (!out->density_grad || !out->radiance_grad || (g->shading_normal && !out->shading_normal_grad)))
        return fail(t, PTIR_ERR_INVALID_ARGUMENT, "null gradient output");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = hipSuccess;
    // the reference allocates every gradient with torch::zeros (threedgrt optixTracer.cpp:953-955)
    if (N) {
        e = ptir::fill_zero(out->density_grad, N * 12 * sizeof(float), s);
        if (e == hipSuccess) e = ptir::fill_zero(out->radiance_grad, N * size_t(g->sh_stride) * sizeof(float), s);
        if (e == hipSuccess && out->shading_normal_grad)
            e = ptir::fill_zero(out->shading_normal_grad, N * 3 * sizeof(float), s);
        if (e != hipSuccess) return hip_fail(t, e, "zeroing gradients");
    }
    e = ptir::prep_records(t->bvh, kopts(t->opts), g->particle_density, nullptr, g->shading_normal, s);
    if (e != hipSuccess) return hip_fail(t, e, "prep_records");
    const ptir_scene sc = grt_as_scene(g);
    ptir::DeviceScene S = make_scene(t, &sc, la);
    S.max_bounces = 1u;
    ptir::grt::BwdIO io{in->radiance, in->density, in->hit_distance, in->hit_distance2, in->distortion,
                        in->shading_normal, in->radiance_grad, in->density_grad, in->hit_distance_grad,
                        in->hit_distance2_grad, in->distortion_grad, in->shading_normal_grad,
                        g->particle_density, g->shading_normal, out->density_grad, out->radiance_grad,
                        g->shading_normal ? out->shading_normal_grad : nullptr};
    io.tape = in->hit_tape;
    io.tape_cap = in->hit_tape ? in->hit_tape_cap : 0u;
    // Hit records (64 B per slot, 32 slots per pixel) replace most of the per-hit gradient
    // atomics; where they cannot be had (slot indices past 32 bits, or no memory for them),
    // every hit adds its gradients by atomics.
    const ptir::GrtRecBuffers* rec = nullptr;
    if (paths > 0 && N > 0 && uint64_t(paths) * ptir::grt::kRecPerPixel < (uint64_t(1) << 32)) {
        if (ptir::grt_rec_reserve(t->grt_rec, size_t(paths), N) == hipSuccess) rec = &t->grt_rec;
        else (void)hipGetLastError();
    }
    e = ptir::launch_grt_bwd(S, io, t->opts.kbuffer, t->short_stack, rec, N, s);
    if (e != hipSuccess) return hip_fail(t, e, "primary-only backward launch");
    return PTIR_OK;
}

uint64_t ptir_grt_tape_bytes(int32_t B, int32_t H, int32_t W, uint32_t hits_per_pixel) {
    if (B < 0 || H < 0 || W < 0 || hits_per_pixel == 0 || hits_per_pixel > 0xFFFFu) return 0;
    return ptir::grt::tape_bytes((uint64_t)B * H * W, hits_per_pixel);
}

uint64_t ptir_tape_bytes(int32_t B, int32_t H, int32_t W, uint32_t max_bounces, uint32_t hits_per_path) {
    if (B < 0 || H < 0 || W < 0 || hits_per_path == 0 || hits_per_path > 0xFFFFu || max_bounces > 64u) return 0;
    return ptir::tape_bytes((uint64_t)B * H * W, max_bounces, hits_per_path);
}

const char* ptir_last_error(const ptir_tracer* t) { return t ? t->err.c_str() : g_create_error.c_str(); }

const char* ptir_version(void) { return "ptir_amd 0.1.0 (gfx950)"; }

}  // extern "C"
