/*
 * ptir_amd.h -- C ABI of the MI355X (gfx950) differentiable Gaussian path tracer.
 *
 * Drop-in boundary for the native surface of the reference's `threedgptir_tracer`
 * plugin.  Every entry point takes plain device pointers, sizes and a HIP stream
 * (passed as `void*`, NULL = default stream) and returns a ptir_status.  Nothing in
 * this header depends on torch.  All device buffers are float32, contiguous, row-major,
 * with the reference's shapes.  The library never synchronises the stream; it only
 * enqueues work on it.  A tracer handle is not re-entrant: calls on one handle must be
 * serialised by the caller (as in the reference, whose params buffer is shared).
 */
#ifndef PTIR_AMD_H
#define PTIR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ptir_status {
    PTIR_OK = 0,
    PTIR_ERR_INVALID_ARGUMENT = 1,   /* bad shape, size, null pointer or option */
    PTIR_ERR_UNSUPPORTED = 2,        /* option the build does not implement */
    PTIR_ERR_NOT_BUILT = 3,          /* trace before build_bvh, or N changed since */
    PTIR_ERR_HIP = 4,                /* a HIP runtime call failed (see ptir_last_error) */
    PTIR_ERR_BVH_DEPTH = 5,          /* hierarchy deeper than the traversal stack */
    PTIR_ERR_OUT_OF_MEMORY = 6
} ptir_status;

typedef struct ptir_tracer ptir_tracer;

/* Constructor flags.  Replaces OptixTracer::OptixTracer
 * (threedgptir_tracer/src/optixTracer.cpp:274-343, bindings.cpp:35): the path, cuda_path,
 * pipeline and primitive strings of the reference select OptiX programs and have no
 * counterpart; only "reference"/"referenceBwd" pipelines on "instances" primitives exist. */
typedef struct ptir_options {
    float kernel_degree;          /* generalized Gaussian degree: 0,1,2,3,4,5,8 (PTIR: 4) */
    float min_response;           /* particle_kernel_min_response (PTIR: 0.0113) */
    int32_t density_clamping;     /* particle_kernel_density_clamping (PTIR: 1) */
    int32_t sph_max_degree;       /* particle_radiance_sph_degree (SH storage degree, <= 3) */
    int32_t enable_normals;
    int32_t enable_hitcounts;
    int32_t enable_mis;
    int32_t enable_metallic;
    int32_t visualize_lights;     /* must be 0 (light visualisation is not built) */
    int32_t kbuffer;              /* hits collected per traversal chunk: 8, 16 (reference) or 32 */
    int32_t wave_traversal;       /* 0: every ray walks the hierarchy on its own; 1: the primary rays
                                     of an 8x8 pixel tile walk it together (the same hits in the
                                     same order, faster; secondary rays always walk on their own) */
    int32_t schedule;             /* forward schedule (no counterpart in the reference): 0 = one lane
                                     follows its path from start to end (the reference's raygen);
                                     1 = wavefront: the primary segments, the NEE of the paths that
                                     reached an interaction and each further bounce run as separate
                                     launches over compacted queues of live paths, with the same
                                     per-path decisions and outputs.  The tracer then keeps a device
                                     workspace of 164 bytes per path, grown on demand (by a host-
                                     synchronous allocation: make the first call outside graph capture) */
    int32_t builder;              /* hierarchy builder (no counterpart in the reference, whose hierarchy
                                     is OptiX's): 0 = LBVH (radix splits of the Morton order); 1 = PLOC
                                     (bottom-up clustering of the same Morton-ordered leaves by smallest
                                     union surface area; fewer nodes visited per ray, a longer build with
                                     a host synchronisation every 8 clustering rounds).  The trace results
                                     do not depend on it; a refit keeps the tree either builder made */
    int32_t wide_walks;           /* which walks use the 4-wide nodes collapsed from the binary tree
                                     (no counterpart in the reference): bit 0 the wave-coherent primary
                                     walk, bit 1 the per-lane walks; the other walks use the binary
                                     nodes.  Half the dependent node fetches, four boxes per fetch.
                                     The trace results do not depend on it */
} ptir_options;

/* Scene inputs of one trace call (OptixTracer::trace argument list,
 * threedgptir_tracer/src/optixTracer.cpp:909-929).  Light tensors other than the
 * environment must be empty (num_lights == 0, num_light_entries == 0, num_mesh_lights == 0). */
typedef struct ptir_scene {
    uint32_t num_gaussians;
    const float* particle_density;    /* [N,12] pos3, density, quat(wxyz)4, scale3, pad */
    const float* particle_material;   /* [N,5]  albedo3, roughness, metallic */
    const float* particle_radiance;   /* [N, sh_stride] SH, coefficient-major, RGB inner */
    int32_t sh_stride;                /* floats per Gaussian in particle_radiance */
    const float* shading_normal;      /* [N,3] */
    const float* environment;         /* [env_h, env_w, 4] or NULL */
    int32_t env_h, env_w;
    const float* env_table;           /* [3, tab_h, tab_w] in the tab_kind layout, or NULL */
    int32_t tab_h, tab_w;
    uint32_t num_lights;              /* lights [L,9]; must be 0 */
    uint32_t num_light_entries;       /* lightAliasTable [5,E]; must be 0 */
    uint32_t num_mesh_lights;         /* meshLights [M,8]; must be 0 */
    int32_t tab_kind;                 /* layout of env_table: PTIR_ENV_TABLE_ALIAS (the reference's
                                         (prob, alias, pdf), built by the caller) or
                                         PTIR_ENV_TABLE_CDF (ptir_env_table_build) */
} ptir_scene;

enum { PTIR_ENV_TABLE_ALIAS = 0, PTIR_ENV_TABLE_CDF = 1 };

/* Launch description.  rays are [B,H,W,3]; ray_to_world is a DEVICE pointer to a
 * [*,4,4] float32 matrix of which only the first 12 floats are used (the reference reads
 * them from rayToWorld[0] on the host, optixTracer.cpp:962; here the kernel reads them on
 * the device, so no host synchronisation happens and the call can be graph-captured). */
typedef struct ptir_launch {
    uint32_t frame;
    const float* ray_to_world;
    const float* ray_ori;
    const float* ray_dir;
    int32_t B, H, W;
    uint32_t render_opts;             /* 0 = NONE, 1 = INDIRECT */
    int32_t sph_degree;               /* evaluated SH degree (n_active_features) */
    float min_transmittance;
    uint32_t max_bounces;             /* path vertices incl. the primary hit */
    /* Neural radiance cache (no counterpart in the reference, which has none): 0 = none;
     * 1 = end every path at its first secondary vertex and write a cache query there (the
     * cache's answer is added by ptir_cache_query); 2 = trace the whole path and write the
     * query with its training target (the radiance the rest of the path carried, per unit
     * throughput).  1 and 2 need max_bounces >= 3 and ptir_fwd_outputs.cache_records.
     * ptir_trace_bwd must be given the same value.  In mode 1 it expects the pbr after
     * ptir_cache_query added the cache term: that term scales with the throughput of the
     * first bounce, so it belongs to the radiance the first vertex's BRDF gradient
     * multiplies; the cache's answer itself is held constant (no gradient flows into the
     * cache or into the material seen at the cache vertex). */
    uint32_t cache_mode;
    /* Ragged batch (no counterpart in the reference, which traces one [B,H,W] batch under one
     * camera transform, optixTracer.cpp:962): num_images > 0 traces images of different
     * sizes, e.g. several views at their own resolutions and samples per pixel, in one
     * launch.  images is a DEVICE int32 [num_images, 4] table of rows (H, W, first tile,
     * first path) filled by ptir_ragged_layout; rays and outputs are flat, [num_paths, C] in
     * image order (pixel (y, x) of image i is path first_path_i + y W_i + x, and its
     * launch index, i.e. its random stream, is (x, y, i)).  B, H, W are ignored; use world-
     * space rays with an identity ray_to_world.  ptir_trace_bwd takes the same fields. */
    int32_t num_images;
    const int32_t* images;
    int32_t num_tiles;
    int32_t num_paths;
    /* Russian roulette (no counterpart in the reference): 0 = off; else, after sampling the
     * direction that leaves path vertex rr_start or later, the path continues with
     * probability q = min(1, max throughput) and divides its throughput by q (one random
     * draw, replayed by the backward, which takes 1/q as a constant). */
    uint32_t rr_start;
} ptir_launch;

/* Host-side layout of a ragged batch of n images of sizes hw[2i] x hw[2i+1] (H, W):
 * writes the [n, 4] table rows (H, W, first tile, first path) into table (host memory; the
 * caller copies it to the device) and the totals.  The tracer covers each image with 16x8
 * pixel tiles. */
ptir_status ptir_ragged_layout(const int32_t* hw, int32_t n, int32_t* table, int32_t* num_tiles, int32_t* num_paths);

/* Forward outputs, all [B,H,W,C] float32 unless noted (optixTracer.cpp:931-944).
 * The library writes every element: G-buffer outputs always, pbr/light/components zero
 * where the primary interaction is invalid.  visibility is [N] float32 whose bits are
 * set to the int32 value 1 for every Gaussian accepted on a traced path, and must be
 * zeroed by the caller before the call (the reference allocates it with torch::zeros). */
typedef struct ptir_fwd_outputs {
    float* radiance;       /* [..,3] */
    float* density;        /* [..,1] */
    float* hit_distance;   /* [..,2] integrated depth, last processed hit distance */
    float* hit_distance2;  /* [..,1] */
    float* distortion;     /* [..,1] */
    float* normal;         /* [..,3] */
    float* shading_normal; /* [..,3] */
    float* material;       /* [..,5] */
    float* hits_count;     /* [..,1] */
    float* visibility;     /* [N] */
    float* pbr;            /* [..,3] */
    float* light;          /* [..,3] */
    float* pbr_components; /* [..,2,3] direct, indirect */
    /* Optional path tape (NULL = none): device memory of tape_bytes >= ptir_tape_bytes(...)
     * into which the forward records, per path, what the backward would otherwise re-trace.
     * No counterpart in the reference; pass the same buffer to ptir_trace_bwd. */
    void* tape;
    uint64_t tape_bytes;
    uint32_t tape_hits_per_path;     /* accepted-hit slots per path, shared by its segments */
    /* [B,H,W,20] radiance-cache records (cache_mode 1 or 2, else NULL).  Per path:
     * flag (1 = the path reached its first secondary vertex), position[3], outgoing
     * direction[3], shading normal[3], roughness, albedo[3], throughput[3], target[3]. */
    float* cache_records;
} ptir_fwd_outputs;

/* Backward inputs: forward rayPbr/rayLight and the incoming gradients.  The reference's
 * backward reads only these among the 10 forward outputs / 10 output gradients it is
 * passed (referenceBwdOptix.cu:38-51).  light_grad and material_grad may be NULL (zero). */
typedef struct ptir_bwd_inputs {
    const float* pbr;            /* [..,3] */
    const float* light;          /* [..,3] */
    const float* pbr_grad;       /* [..,3] */
    const float* light_grad;     /* [..,3] or NULL */
    const float* material_grad;  /* [..,5] or NULL */
    /* Optional tape written by the forward of the same launch (NULL = re-trace every
     * segment, as the reference does).  A tape from a different launch is detected and
     * ignored; paths whose hits overflowed the tape are re-traced. */
    const void* tape;
    uint64_t tape_bytes;
    uint32_t tape_hits_per_path;
} ptir_bwd_inputs;

/* Backward outputs (OptixTracer::traceBwd, optixTracer.cpp:1058-1063).  The library
 * zeroes and writes every element.  In the reference semantics only material and env
 * gradients are non-zero; density/SH/shading-normal gradients are written as zeros. */
typedef struct ptir_bwd_outputs {
    float* density_grad;          /* [N,12] */
    float* material_grad;         /* [N,5] */
    float* radiance_grad;         /* [N, sh_stride] */
    float* shading_normal_grad;   /* [N,3] */
    float* environment_grad;      /* [env_h, env_w, 4] or NULL when there is no env */
} ptir_bwd_outputs;

/* Work counters of one call (optional): units of the algorithmic byte model. */
typedef struct ptir_counters {
    uint64_t paths, segments, candidates, accepted, accepted_primary, shadow_candidates, env_lookups;
    uint64_t mat_atomics, env_atomics;
    uint64_t traversals, nodes, leaf_tests;  /* hierarchy work (no counterpart in the byte model) */
} ptir_counters;

/* OptixTracer::OptixTracer (optixTracer.cpp:274-343).  Allocates no device memory. */
ptir_status ptir_create(const ptir_options* options, ptir_tracer** out);

/* ~OptixTracer (optixTracer.cpp:345-371).  Frees the handle's device buffers. */
ptir_status ptir_destroy(ptir_tracer* tracer);

/* OptixTracer::buildBVH (optixTracer.cpp:633-907): pos [N,3], rot [N,4] (unit wxyz),
 * scl [N,3] (activated), dns [N] (activated), all device pointers.  rebuild = 0 refits
 * the existing hierarchy (topology kept, boxes recomputed); refitting with a different N
 * is an error here (the reference prints and rebuilds).  allow_update is accepted for
 * parity and has no effect: every hierarchy can be refitted. */
ptir_status ptir_build_bvh(ptir_tracer* tracer, const float* pos, const float* rot, const float* scl, const float* dns,
                           uint32_t num_gaussians, uint32_t rebuild, int32_t allow_update, void* stream);

/* Reports a traversal-stack overflow in any launch since the last report.  No counterpart in
 * the reference, whose traversal is OptiX's own (its pipelines are compiled with exceptions
 * off, optixTracer.cpp:399).  The kernels set a sticky device flag; this call copies it back
 * (synchronising the stream), clears it and returns PTIR_ERR_BVH_DEPTH when it was set.
 * A counting launch (counters != NULL) reports and clears it too; a rebuild clears it.
 * The launch that overflowed dropped the subtrees it could not push. */
ptir_status ptir_check(ptir_tracer* tracer, void* stream);

/* Shape of the last built hierarchy (no counterpart in the reference, whose hierarchy is
 * OptiX's own): the number of 4-wide nodes the traversal walks and the traversal stack
 * entries its walks can need (at most 64; up to 36 runs the kernels' short LDS stack).  Either
 * output may be NULL.  Synchronises nothing: the values were read back by the last rebuild. */
ptir_status ptir_hierarchy_info(ptir_tracer* tracer, int32_t* wide_nodes, int32_t* stack_need);

/* OptixTracer::trace (optixTracer.cpp:909-1013).  counters may be NULL; when not NULL
 * the kernels count work units into device memory and the call copies them back
 * synchronously (use only for measurement). */
ptir_status ptir_trace(ptir_tracer* tracer, const ptir_scene* scene, const ptir_launch* launch,
                       const ptir_fwd_outputs* out, ptir_counters* counters, void* stream);

/* OptixTracer::traceBwd (optixTracer.cpp:1015-1145), path-replay backward. */
ptir_status ptir_trace_bwd(ptir_tracer* tracer, const ptir_scene* scene, const ptir_launch* launch,
                           const ptir_bwd_inputs* in, const ptir_bwd_outputs* out, ptir_counters* counters,
                           void* stream);

/* Gradients of the forward outputs that ptir_trace_bwd ignores, for ptir_trace_bwd_exact.
 * Each may be NULL (zero).  The normal, hit_distance2, distortion and hits_count outputs
 * take no gradient. */
typedef struct ptir_exact_grads {
    const float* radiance_grad;        /* [..,3] primary radiance */
    const float* density_grad;         /* [..,1] primary opacity */
    const float* hit_distance_grad;    /* [..,2] channel 0 (integrated depth); channel 1 is ignored */
    const float* shading_normal_grad;  /* [..,3] */
    const float* pbr_components_grad;  /* [..,2,3] direct, indirect */
} ptir_exact_grads;

/* Exact-adjoint backward (no counterpart in the reference, whose backward leaves the density,
 * SH and shading-normal gradients at zero and approximates the rest, referenceBwdOptix.cu:27-74):
 * the gradient of sum_outputs <output_grad, output> with respect to every parameter of the
 * trace -- particle_density [N,12] (position, density, quaternion, scale; the quaternion is
 * differentiated as given, without normalisation), particle_material, particle_radiance,
 * shading_normal and environment -- with the paths' discrete decisions (candidate sets,
 * accepted hits, escape, shadow and roulette draws, sampled cells) held fixed.  Each path is
 * replayed from its random stream, so the launch must match the forward's.  in->pbr_grad is
 * required; light_grad and material_grad may be NULL; in->pbr, in->light and the tape are not
 * read.  extra may be NULL.  cache_mode must be 0.  Allocates stream-ordered scratch of
 * 112 * max(max_bounces, 1) bytes per path for the duration of the call.
 * In this build the device launch is switched off: after argument checks the call returns
 * PTIR_ERR_UNSUPPORTED (its kernel faulted the device and the cause is not found; the device
 * code itself is verified on the host against the FP64 oracle's derivatives). */
ptir_status ptir_trace_bwd_exact(ptir_tracer* tracer, const ptir_scene* scene, const ptir_launch* launch,
                                 const ptir_bwd_inputs* in, const ptir_exact_grads* extra,
                                 const ptir_bwd_outputs* out, void* stream);

/* ------------------------------------------------------------------------------------
 * 3DGRT primary-only tracer (threedgrt_tracer/, the stage-1 tracer of the same repository)
 * on the same handle, hierarchy and k-buffer: ptir_create/ptir_build_bvh as above
 * (ptir_options' MIS, metallic and schedule fields are ignored), then these two calls.
 * The launch uses B, H, W, the ray tensors, frame, sph_degree and min_transmittance;
 * render_opts must be 0 and num_images 0. */
typedef struct ptir_grt_scene {
    uint32_t num_gaussians;
    const float* particle_density;   /* [N,12] pos3, density, quat(wxyz)4, scale3, pad */
    const float* particle_radiance;  /* [N, sh_stride] SH coefficients */
    int32_t sh_stride;
    const float* shading_normal;     /* [N,3] (NULL = none) */
} ptir_grt_scene;

/* threedgrt OptixTracer::trace outputs (threedgrt_tracer/src/optixTracer.cpp:855-925), all
 * [B,H,W,C] float32: the primary segment's radiance, opacity, integrated depth and last hit
 * distance, depth second moment, distortion, normal, shading normal (the weighted sum, not
 * normalised by the opacity, unlike ptir_trace's), hit count.  visibility as in
 * ptir_fwd_outputs (zeroed by the caller). */
typedef struct ptir_grt_outputs {
    float* radiance;        /* [..,3] */
    float* density;         /* [..,1] */
    float* hit_distance;    /* [..,2] */
    float* hit_distance2;   /* [..,1] */
    float* distortion;      /* [..,1] */
    float* normal;          /* [..,3] */
    float* shading_normal;  /* [..,3] */
    float* hits_count;      /* [..,1] */
    float* visibility;      /* [N] */
    uint32_t* hit_tape;     /* optional (NULL = none): ptir_grt_tape_bytes(B,H,W,hit_tape_cap) bytes
                               of device memory that receive each pixel's accepted hits, for
                               ptir_grt_trace_bwd to read instead of re-tracing */
    uint32_t hit_tape_cap;  /* hits per pixel the tape holds, [1, 65535] */
} ptir_grt_outputs;

/* threedgrt OptixTracer::trace (threedgrt_tracer/src/optixTracer.cpp:855-925;
 * kernels referenceOptix.cu:105-194). */
ptir_status ptir_grt_trace(ptir_tracer* tracer, const ptir_grt_scene* scene, const ptir_launch* launch,
                           const ptir_grt_outputs* out, void* stream);

/* Forward outputs of ptir_grt_trace and their gradients (each gradient may be NULL = zero);
 * hit_distance_grad is [..,1]: the integrated depth (channel 0) only. */
typedef struct ptir_grt_bwd_inputs {
    const float *radiance, *density, *hit_distance, *hit_distance2, *distortion, *shading_normal;
    const float *radiance_grad, *density_grad, *hit_distance_grad, *hit_distance2_grad, *distortion_grad,
        *shading_normal_grad;
    const uint32_t* hit_tape;  /* the hit tape the forward wrote, or NULL: every pixel re-traces.  A
                                  tape of another launch shape, frame or Gaussian count is ignored;
                                  pixels whose hits overflowed it re-trace. */
    uint32_t hit_tape_cap;     /* as given to the forward */
} ptir_grt_bwd_inputs;

/* Gradients, zeroed by the call (the reference allocates them with torch::zeros):
 * density_grad [N,12] (position, density, quaternion as given, scale), radiance_grad
 * [N, sh_stride], shading_normal_grad [N,3] (may be NULL when the scene has no shading normal). */
typedef struct ptir_grt_bwd_outputs {
    float* density_grad;
    float* radiance_grad;
    float* shading_normal_grad;
} ptir_grt_bwd_outputs;

/* threedgrt OptixTracer::traceBwd (threedgrt_tracer/src/optixTracer.cpp:927-1012; kernels
 * referenceBwdOptix.cu:103-184, gaussianParticles.cuh:473-772): per pixel, the forward's
 * candidates re-collected up to its last hit distance (float32 end bound, which drops the
 * hits at exactly that distance, as the reference does) and each hit's adjoint through the
 * forward's totals.  The normal and hit-count outputs take no gradient. */
ptir_status ptir_grt_trace_bwd(ptir_tracer* tracer, const ptir_grt_scene* scene, const ptir_launch* launch,
                               const ptir_grt_bwd_inputs* in, const ptir_grt_bwd_outputs* out, void* stream);

/* Bytes of a path tape for a [B,H,W] launch with max_bounces path vertices and room for
 * hits_per_path accepted hits per path (0 on invalid sizes).  No counterpart in the
 * reference: it lets ptir_trace_bwd skip re-tracing (a 64 B record per path and per
 * bounce, plus 8 B per hit slot). */
uint64_t ptir_tape_bytes(int32_t B, int32_t H, int32_t W, uint32_t max_bounces, uint32_t hits_per_path);

/* Bytes of a 3DGRT hit tape for a [B,H,W] launch holding hits_per_pixel accepted hits per
 * pixel (0 on invalid sizes): 64 B, 4 B per pixel, 8 B per hit slot.  No counterpart in the
 * reference: it lets ptir_grt_trace_bwd skip re-tracing. */
uint64_t ptir_grt_tape_bytes(int32_t B, int32_t H, int32_t W, uint32_t hits_per_pixel);

/* ------------------------------------------------------------------------------------
 * Neural radiance cache.  No counterpart in the reference (its tiny-cuda-nn submodule is
 * empty and unused).  A multiresolution hash grid (16 levels x 2 features, 2^19 entries
 * per level, resolutions 16..2048 over the scene box) and the SH (bands 0..3) of the
 * outgoing direction, roughness, albedo and normal feed an MLP of four hidden ReLU layers
 * of width 64 and a 3-channel exp output, in bf16 on MFMA with FP32 accumulation.  It is
 * trained online on the records of a cache_mode 2 trace and queried on those of a
 * cache_mode 1 trace.  The handle owns FP32 master weights, the table, Adam state and
 * the scratch of one training batch of up to max_batch records. */
typedef struct ptir_cache ptir_cache;

typedef struct ptir_cache_options {
    uint64_t max_batch;      /* records per training call, at most 2^20 */
    uint32_t seed;           /* initial weights and table */
    float learning_rate;     /* Adam */
    float beta1, beta2, epsilon;
} ptir_cache_options;

ptir_status ptir_cache_create(const ptir_cache_options* options, ptir_cache** out, void* stream);
ptir_status ptir_cache_destroy(ptir_cache* cache);
/* Sizes of the FP32 parameters: weights W0..W3 [64,64] (out, in) then W4 [16,64] (rows 3..15
 * zero); table [16, 2^19, 2]. */
uint64_t ptir_cache_num_weights(void);
uint64_t ptir_cache_table_floats(void);
/* Device-to-device copies of the parameters out of / into the cache (either pointer may be
 * NULL).  Writing repacks the bf16 operands. */
ptir_status ptir_cache_read_params(ptir_cache* cache, float* weights, float* table, void* stream);
ptir_status ptir_cache_write_params(ptir_cache* cache, const float* weights, const float* table, void* stream);
/* Gradients left by the last ptir_cache_backprop (cleared by ptir_cache_step) and its loss
 * (device [1]); any pointer may be NULL. */
ptir_status ptir_cache_read_grads(ptir_cache* cache, float* weight_grad, float* table_grad, float* loss,
                                  void* stream);
/* Cache answers for n records (scene_box: device [6] lo.xyz hi.xyz).  For every record
 * with flag 1, adds throughput * answer to pbr [n,3] and to the indirect half of
 * pbr_components [n,2,3] (either may be NULL); prediction [n,3] (or NULL) receives the
 * answer of every record with flag 1 (other rows are left as they are). */
ptir_status ptir_cache_query(ptir_cache* cache, const float* records, uint64_t n, const float* scene_box, float* pbr,
                             float* pbr_components, float* prediction, void* stream);
/* Loss and gradients of one training batch of n <= max_batch records (flag 0 records are
 * ignored; the loss is the mean over all n): overwrites weight_grad and loss, adds into
 * table_grad. */
ptir_status ptir_cache_backprop(ptir_cache* cache, const float* records, uint64_t n, const float* scene_box,
                                void* stream);
/* One Adam step on weights and table; clears both gradients. */
ptir_status ptir_cache_step(ptir_cache* cache, void* stream);
/* ptir_cache_backprop followed by ptir_cache_step. */
ptir_status ptir_cache_train(ptir_cache* cache, const float* records, uint64_t n, const float* scene_box,
                             void* stream);
const char* ptir_cache_last_error(const ptir_cache* cache);

/* ------------------------------------------------------------------------------------
 * Environment importance table built on the device (layout PTIR_ENV_TABLE_CDF).
 * Replaces the caller-side alias table build (threedgrut/model/aliastable.py:244-297,
 * packed by threedgptir_tracer/tracer.py:819-847): the same cell weights (area resampling
 * of the [env_h, env_w, 4] environment to tab_h x tab_w, Rec.709 luminance of the clamped
 * RGB, times the equirect cell solid angle; solid angles alone when the environment carries
 * no finite positive light) and the same per-steradian pdf w / (sum(w) * solid_angle), but
 * sampled by CDF inversion instead of a Vose alias table, with the prefix sums in FP64.
 * table is device [3, tab_h, tab_w] float32: plane 0 the inclusive conditional CDF of each
 * row, plane 1 column 0 the inclusive marginal CDF of the rows (other columns zero), plane 2
 * the pdf.  Every CDF ends at exactly 1.  The sampler draws four numbers like the alias
 * sampler (row, column, position in the cell), so the rest of a path's random stream is the
 * same under either layout; the chosen cells differ, so parity with the alias sampler is
 * distributional.  Scratch of ptir_env_table_scratch_bytes() is allocated stream-ordered
 * (hipMallocAsync / hipFreeAsync) on the stream; nothing synchronises.  Errors are
 * reported by ptir_last_error(NULL). */
uint64_t ptir_env_table_scratch_bytes(int32_t tab_h, int32_t tab_w);
ptir_status ptir_env_table_build(const float* environment, int32_t env_h, int32_t env_w, int32_t tab_h, int32_t tab_w,
                                 float* table, void* stream);

/* ptir_env_table_build without any allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  Its scratch is a static device array of the library (1.08 MB per
 * GPU, held until the process ends), so tables are limited to 131072 cells (e.g. 256 x 512)
 * and 4096 rows; larger ones return PTIR_ERR_INVALID_ARGUMENT.  Same kernels and results as
 * ptir_env_table_build.  Every call on one GPU shares that array: calls that overlap on one
 * GPU (different streams or threads) race and must be serialised by the caller. */
ptir_status ptir_env_table_build_noalloc(const float* environment, int32_t env_h, int32_t env_w, int32_t tab_h,
                                         int32_t tab_w, float* table, void* stream);

/* ------------------------------------------------------------------------------------
 * Render-side image work on the device (the trainer-side torch code of the reference's
 * Tracer.render).  Errors are reported by ptir_last_error(NULL).
 *
 * Camera rays for S samples per pixel (threedgptir_tracer/tracer.py:646-706): for sample s,
 * view b and pixel (y, x), origin = ray_ori[b,y,x] and direction =
 * normalize(((px - 0.5 + jx - cx) / fx, (py - 0.5 + jy - cy) / fy, 1)) with (px, py) =
 * pixel_coords[b,y,x] ([B,H,W,2]), intrinsics = device [4] (fx, fy, cx, cy) and the jitter
 * (jx, jy) = jitter[s] ([S,2]) or jitter[s,y,x] ([S,H,W,2], jitter_per_pixel = 1).  Output
 * [S*B,H,W,3], sample-major. */
ptir_status ptir_camera_rays(const float* pixel_coords, const float* intrinsics, const float* jitter,
                             int32_t jitter_per_pixel, const float* ray_ori, int32_t S, int32_t B, int32_t H,
                             int32_t W, float* out_ori, float* out_dir, void* stream);

/* Resolve of the spp-averaged tracer outputs into render outputs (tracer.py:505-526,
 * threedgrut/model/filters.py:6-69): out_dist = dist / opacity (non-finite -> 0),
 * out_normals = normals / max(|normals|, 1e-12), and image [B,H,W,12] =
 * (pbr * opacity, direct * opacity, indirect * opacity, light) filtered by `iterations`
 * passes of the 3x3 B-spline a-trous kernel with replicate borders (0 = no filter).  work is
 * a second [B,H,W,12] buffer.  The backward takes the packed gradient of image in `image`
 * (overwritten) and writes every input gradient. */
typedef struct ptir_resolve {
    int32_t B, H, W, iterations;
    const float *opacity, *dist, *pbr, *direct, *indirect, *light, *normals;
    float *out_dist, *out_normals;
    float *image, *work;
} ptir_resolve;
ptir_status ptir_resolve_forward(const ptir_resolve* r, void* stream);
ptir_status ptir_resolve_backward(const ptir_resolve* r, const float* grad_dist, const float* grad_normals,
                                  float* d_opacity, float* d_dist, float* d_pbr, float* d_direct, float* d_indirect,
                                  float* d_light, float* d_normals, void* stream);

/* ------------------------------------------------------------------------------------
 * Adam step of the optimised tensors (the optimiser of the reference's PTIR inversion:
 * torch.optim.Adam, or SelectiveAdam of threedgrut/optimizers/optimizers.cu:49-80 with
 * visibility-masked rows and no bias correction).  Each group is a [rows, cols] float32
 * tensor with its gradient and moment buffers (device pointers) and learning rate;
 * visibility, if not NULL, is a [rows] array whose nonzero words mark the rows to update
 * (the tracer's visibility output).  bias_correction = 1: torch.optim.Adam,
 * p -= (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps) with t = *step + 1;
 * 0: p -= lr m / (sqrt(v) + eps).  step is a device int32 incremented by the call, so the
 * step can be captured in a graph and replayed.  Errors: ptir_last_error(NULL). */
typedef struct ptir_adam_group {
    float* param;
    const float* grad;
    float* exp_avg;
    float* exp_avg_sq;
    const uint32_t* visibility;
    uint64_t rows;
    uint32_t cols;
    float lr;
} ptir_adam_group;
ptir_status ptir_adam_step(const ptir_adam_group* groups, int32_t num_groups, float beta1, float beta2, float eps,
                           int32_t bias_correction, int32_t* step, void* stream);

/* Human-readable text of the last error on this handle (or, when tracer is NULL, of the
 * last failed call without a handle on this thread: ptir_create, the table builders).  Valid until the next call on the handle. */
const char* ptir_last_error(const ptir_tracer* tracer);

/* Library version string. */
const char* ptir_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PTIR_AMD_H */
