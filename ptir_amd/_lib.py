"""ctypes binding of libptir_amd.so (the C ABI declared in include/ptir_amd.h).

The library is built in-tree by ``__graft_entry__.build()`` (ptir_amd/csrc/Makefile).
Loading fails loudly when it is missing: there is no CPU or eager fallback.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptir_amd.so")

PTIR_STATUS = {
    0: "PTIR_OK",
    1: "PTIR_ERR_INVALID_ARGUMENT",
    2: "PTIR_ERR_UNSUPPORTED",
    3: "PTIR_ERR_NOT_BUILT",
    4: "PTIR_ERR_HIP",
    5: "PTIR_ERR_BVH_DEPTH",
    6: "PTIR_ERR_OUT_OF_MEMORY",
}

EXPORTED = ["ptir_create", "ptir_destroy", "ptir_build_bvh", "ptir_check", "ptir_hierarchy_info", "ptir_grt_trace", "ptir_grt_trace_bwd",
            "ptir_trace", "ptir_trace_bwd", "ptir_tape_bytes", "ptir_grt_tape_bytes",
            "ptir_last_error", "ptir_version", "ptir_cache_create", "ptir_cache_destroy", "ptir_cache_num_weights",
            "ptir_cache_table_floats", "ptir_cache_read_params", "ptir_cache_write_params", "ptir_cache_read_grads",
            "ptir_cache_query", "ptir_cache_backprop", "ptir_cache_step", "ptir_cache_train", "ptir_cache_last_error",
            "ptir_env_table_scratch_bytes", "ptir_env_table_build", "ptir_env_table_build_noalloc",
            "ptir_camera_rays", "ptir_resolve_forward", "ptir_resolve_backward", "ptir_adam_step",
            "ptir_ragged_layout", "ptir_trace_bwd_exact"]


class Options(C.Structure):
    _fields_ = [("kernel_degree", C.c_float), ("min_response", C.c_float), ("density_clamping", C.c_int32),
                ("sph_max_degree", C.c_int32), ("enable_normals", C.c_int32), ("enable_hitcounts", C.c_int32),
                ("enable_mis", C.c_int32), ("enable_metallic", C.c_int32), ("visualize_lights", C.c_int32),
                ("kbuffer", C.c_int32), ("wave_traversal", C.c_int32), ("schedule", C.c_int32),
                ("builder", C.c_int32), ("wide_walks", C.c_int32)]


class Scene(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint32), ("particle_density", C.c_void_p), ("particle_material", C.c_void_p),
                ("particle_radiance", C.c_void_p), ("sh_stride", C.c_int32), ("shading_normal", C.c_void_p),
                ("environment", C.c_void_p), ("env_h", C.c_int32), ("env_w", C.c_int32),
                ("env_table", C.c_void_p), ("tab_h", C.c_int32), ("tab_w", C.c_int32),
                ("num_lights", C.c_uint32), ("num_light_entries", C.c_uint32), ("num_mesh_lights", C.c_uint32),
                ("tab_kind", C.c_int32)]

ENV_TABLE_ALIAS, ENV_TABLE_CDF = 0, 1


class AdamGroup(C.Structure):
    _fields_ = [("param", C.c_void_p), ("grad", C.c_void_p), ("exp_avg", C.c_void_p), ("exp_avg_sq", C.c_void_p),
                ("visibility", C.c_void_p), ("rows", C.c_uint64), ("cols", C.c_uint32), ("lr", C.c_float)]


class Resolve(C.Structure):
    _fields_ = [("B", C.c_int32), ("H", C.c_int32), ("W", C.c_int32), ("iterations", C.c_int32)] + [
        (k, C.c_void_p) for k in ("opacity", "dist", "pbr", "direct", "indirect", "light", "normals", "out_dist",
                                  "out_normals", "image", "work")]


class Launch(C.Structure):
    _fields_ = [("frame", C.c_uint32), ("ray_to_world", C.c_void_p), ("ray_ori", C.c_void_p), ("ray_dir", C.c_void_p),
                ("B", C.c_int32), ("H", C.c_int32), ("W", C.c_int32), ("render_opts", C.c_uint32),
                ("sph_degree", C.c_int32), ("min_transmittance", C.c_float), ("max_bounces", C.c_uint32),
                ("cache_mode", C.c_uint32), ("num_images", C.c_int32), ("images", C.c_void_p),
                ("num_tiles", C.c_int32), ("num_paths", C.c_int32), ("rr_start", C.c_uint32)]


class FwdOutputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("radiance", "density", "hit_distance", "hit_distance2", "distortion",
                                          "normal", "shading_normal", "material", "hits_count", "visibility", "pbr",
                                          "light", "pbr_components", "tape")] + [
        ("tape_bytes", C.c_uint64), ("tape_hits_per_path", C.c_uint32), ("cache_records", C.c_void_p)]


class BwdInputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("pbr", "light", "pbr_grad", "light_grad", "material_grad", "tape")] + [
        ("tape_bytes", C.c_uint64), ("tape_hits_per_path", C.c_uint32)]


class BwdOutputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("density_grad", "material_grad", "radiance_grad", "shading_normal_grad",
                                          "environment_grad")]


class ExactGrads(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("radiance_grad", "density_grad", "hit_distance_grad",
                                          "shading_normal_grad", "pbr_components_grad")]


class Counters(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("paths", "segments", "candidates", "accepted", "accepted_primary",
                                          "shadow_candidates", "env_lookups", "mat_atomics", "env_atomics",
                                          "traversals", "nodes", "leaf_tests")]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class GrtScene(C.Structure):
    _fields_ = [("num_gaussians", C.c_uint32), ("particle_density", C.c_void_p), ("particle_radiance", C.c_void_p),
                ("sh_stride", C.c_int32), ("shading_normal", C.c_void_p)]


class GrtOutputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("radiance", "density", "hit_distance", "hit_distance2", "distortion",
                                           "normal", "shading_normal", "hits_count", "visibility", "hit_tape")] + \
        [("hit_tape_cap", C.c_uint32)]


class GrtBwdInputs(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("radiance", "density", "hit_distance", "hit_distance2", "distortion",
                                           "shading_normal", "radiance_grad", "density_grad", "hit_distance_grad",
                                           "hit_distance2_grad", "distortion_grad", "shading_normal_grad",
                                           "hit_tape")] + [("hit_tape_cap", C.c_uint32)]


class GrtBwdOutputs(C.Structure):
    _fields_ = [("density_grad", C.c_void_p), ("radiance_grad", C.c_void_p), ("shading_normal_grad", C.c_void_p)]


class CacheOptions(C.Structure):
    _fields_ = [("max_batch", C.c_uint64), ("seed", C.c_uint32), ("learning_rate", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("epsilon", C.c_float)]


_lib = None


def lib() -> C.CDLL:
    """Load libptir_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(hipcc --offload-arch=gfx950); there is no fallback path")
        L = C.CDLL(LIB_PATH)
        L.ptir_create.argtypes = [C.POINTER(Options), C.POINTER(C.c_void_p)]
        L.ptir_destroy.argtypes = [C.c_void_p]
        L.ptir_build_bvh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                     C.c_uint32, C.c_int32, C.c_void_p]
        L.ptir_trace.argtypes = [C.c_void_p, C.POINTER(Scene), C.POINTER(Launch), C.POINTER(FwdOutputs),
                                 C.POINTER(Counters), C.c_void_p]
        L.ptir_trace_bwd.argtypes = [C.c_void_p, C.POINTER(Scene), C.POINTER(Launch), C.POINTER(BwdInputs),
                                     C.POINTER(BwdOutputs), C.POINTER(Counters), C.c_void_p]
        L.ptir_trace_bwd_exact.argtypes = [C.c_void_p, C.POINTER(Scene), C.POINTER(Launch), C.POINTER(BwdInputs),
                                           C.POINTER(ExactGrads), C.POINTER(BwdOutputs), C.c_void_p]
        L.ptir_check.argtypes = [C.c_void_p, C.c_void_p]
        L.ptir_hierarchy_info.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ptir_grt_trace.argtypes = [C.c_void_p, C.POINTER(GrtScene), C.POINTER(Launch), C.POINTER(GrtOutputs),
                                     C.c_void_p]
        L.ptir_grt_trace_bwd.argtypes = [C.c_void_p, C.POINTER(GrtScene), C.POINTER(Launch), C.POINTER(GrtBwdInputs),
                                         C.POINTER(GrtBwdOutputs), C.c_void_p]
        L.ptir_tape_bytes.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32]
        L.ptir_tape_bytes.restype = C.c_uint64
        L.ptir_grt_tape_bytes.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]
        L.ptir_grt_tape_bytes.restype = C.c_uint64
        L.ptir_last_error.argtypes = [C.c_void_p]
        L.ptir_last_error.restype = C.c_char_p
        L.ptir_version.restype = C.c_char_p
        vp, u64 = C.c_void_p, C.c_uint64
        L.ptir_cache_create.argtypes = [C.POINTER(CacheOptions), C.POINTER(C.c_void_p), vp]
        L.ptir_cache_destroy.argtypes = [vp]
        L.ptir_cache_num_weights.restype = u64
        L.ptir_cache_table_floats.restype = u64
        L.ptir_cache_read_params.argtypes = [vp, vp, vp, vp]
        L.ptir_cache_write_params.argtypes = [vp, vp, vp, vp]
        L.ptir_cache_read_grads.argtypes = [vp, vp, vp, vp, vp]
        L.ptir_cache_query.argtypes = [vp, vp, u64, vp, vp, vp, vp, vp]
        L.ptir_cache_backprop.argtypes = [vp, vp, u64, vp, vp]
        L.ptir_cache_step.argtypes = [vp, vp]
        L.ptir_cache_train.argtypes = [vp, vp, u64, vp, vp]
        L.ptir_cache_last_error.argtypes = [vp]
        L.ptir_cache_last_error.restype = C.c_char_p
        L.ptir_env_table_scratch_bytes.argtypes = [C.c_int32, C.c_int32]
        L.ptir_env_table_scratch_bytes.restype = u64
        L.ptir_env_table_build.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, vp]
        L.ptir_env_table_build_noalloc.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, vp]
        i32 = C.c_int32
        L.ptir_camera_rays.argtypes = [vp, vp, vp, i32, vp, i32, i32, i32, i32, vp, vp, vp]
        L.ptir_resolve_forward.argtypes = [C.POINTER(Resolve), vp]
        L.ptir_resolve_backward.argtypes = [C.POINTER(Resolve)] + [vp] * 10
        L.ptir_adam_step.argtypes = [C.POINTER(AdamGroup), i32, C.c_float, C.c_float, C.c_float, i32, vp, vp]
        L.ptir_ragged_layout.argtypes = [vp, i32, vp, vp, vp]
        for name in ("ptir_create", "ptir_destroy", "ptir_build_bvh", "ptir_check", "ptir_hierarchy_info", "ptir_grt_trace",
                     "ptir_grt_trace_bwd", "ptir_trace", "ptir_trace_bwd", "ptir_trace_bwd_exact",
                     "ptir_cache_create", "ptir_cache_destroy", "ptir_cache_read_params", "ptir_cache_write_params",
                     "ptir_cache_read_grads", "ptir_cache_query", "ptir_cache_backprop", "ptir_cache_step", "ptir_cache_train",
                     "ptir_env_table_build", "ptir_env_table_build_noalloc", "ptir_camera_rays", "ptir_resolve_forward",
                     "ptir_resolve_backward", "ptir_adam_step", "ptir_ragged_layout"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


class PtirError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{PTIR_STATUS.get(status, status)}: {message}")
        self.status = status


def check(status: int, handle) -> None:
    if status != 0:
        msg = lib().ptir_last_error(handle)
        raise PtirError(status, msg.decode() if msg else "")


def check_cache(status: int, handle) -> None:
    if status != 0:
        msg = lib().ptir_cache_last_error(handle)
        raise PtirError(status, msg.decode() if msg else "")
