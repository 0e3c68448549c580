"""Python surface of the MI355X tracer, mirroring threedgptir_tracer/tracer.py.

* `HipTracer` has the method set of the reference's pybind class `OptixTracer`
  (threedgptir_tracer/bindings.cpp:32-38, include/3dgptir/optixTracer.h:130-217):
  ``trace`` (13 outputs), ``trace_bwd`` (5 outputs) and ``build_bvh``, same argument
  order, same output shapes.  It calls the C ABI of libptir_amd.so through ctypes with
  the current torch stream; outputs are allocated here with torch.
* `Tracer` keeps the reference's `Tracer(conf)` / `build_acc` / `render` surface and its
  `_Autograd` function (tracer.py:136-384,428-975), so a PTIR model can swap it in.

Errors raise `PtirError` (the reference swallows CUDA/OptiX errors,
include/3dgptir/cuoptixMacros.h:22-45).
"""

from __future__ import annotations

import ctypes as C
import logging
from dataclasses import dataclass
from enum import IntEnum

import torch

from . import _lib
from . import envmap, render

logger = logging.getLogger(__name__)


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _need(t: torch.Tensor, name: str, shape_tail=None, dtype=torch.float32) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if shape_tail is not None and tuple(t.shape[-len(shape_tail):]) != tuple(shape_tail):
        raise ValueError(f"{name} must have trailing shape {tuple(shape_tail)}, got {tuple(t.shape)}")
    return t.contiguous()


@dataclass
class RaggedBatch:
    """Layout of a ragged launch: images of their own sizes traced together (several views
    at their own resolutions and samples per pixel).  Rays and outputs of such a launch are
    flat, [num_paths, C] in image order; `split` cuts a flat output back into images."""

    hw: list
    table: torch.Tensor  # device int32 [n, 4]: H, W, first tile, first path
    num_tiles: int
    num_paths: int

    @classmethod
    def build(cls, hw, device) -> "RaggedBatch":
        hw = [(int(h), int(w)) for h, w in hw]
        flat = (C.c_int32 * (2 * len(hw)))(*[v for p in hw for v in p])
        table = (C.c_int32 * (4 * len(hw)))()
        tiles, paths = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().ptir_ragged_layout(flat, len(hw), table, C.byref(tiles), C.byref(paths)), None)
        t = torch.tensor(list(table), dtype=torch.int32).view(len(hw), 4).to(device)
        return cls(hw, t, int(tiles.value), int(paths.value))

    def offsets(self):
        return [int(v) for v in self.table[:, 3].cpu()]

    def split(self, flat: torch.Tensor):
        """[num_paths, C...] (or [1,1,num_paths,C...]) -> list of [H, W, C...] views."""
        if flat.dim() >= 3 and tuple(flat.shape[:2]) == (1, 1):
            flat = flat.reshape(self.num_paths, *flat.shape[3:])
        out, start = [], 0
        for h, w in self.hw:
            out.append(flat[start:start + h * w].reshape(h, w, *flat.shape[1:]))
            start += h * w
        return out


# which hierarchy walks use the 4-wide nodes (ptir_options.wide_walks): none, the wave-coherent
# primary walk, the per-lane walks, or all of them
WIDE_WALKS = {"none": 0, "wave": 1, "lane": 2, "all": 3}
# forward schedules (ptir_options.schedule): one lane per path, or wavefront stages
SCHEDULES = {"megakernel": 0, "wavefront": 1}


class HipTracer:
    """Drop-in for the reference's `OptixTracer` pybind class."""

    def __init__(self, path: str = "", cuda_path: str = "", pipeline: str = "reference",
                 backward_pipeline: str = "referenceBwd", primitive: str = "instances",
                 particle_kernel_degree: float = 4.0, particle_kernel_min_response: float = 0.0113,
                 particle_kernel_density_clamping: bool = True, particle_radiance_sph_degree: int = 3,
                 enable_normals: bool = True, enable_hitcounts: bool = True, enable_mis: bool = True,
                 enable_metallic: bool = False, visualize_lights: bool = False, kbuffer: int = 16,
                 use_tape: bool = True, tape_hits_per_path: int = 64, traversal: str = "wave",
                 env_sampling: str = "alias", gradients: str = "ptir", schedule: str = "megakernel",
                 builder: str = "lbvh", wide_walks: str = "none"):
        if pipeline != "reference" or backward_pipeline not in ("referenceBwd", "reference" + "Bwd"):
            raise ValueError(f"unsupported pipeline pair ({pipeline}, {backward_pipeline}); "
                             "only the PTIR 'reference' path tracer is built")
        if primitive != "instances":
            raise ValueError(f"unsupported primitive type {primitive!r}; only 'instances' is built")
        if traversal not in ("lane", "wave"):
            raise ValueError(f"traversal must be 'lane' or 'wave', not {traversal!r}")
        if env_sampling not in ("alias", "cdf"):
            raise ValueError(f"env_sampling must be 'alias' (caller-built alias table) or 'cdf' "
                             f"(envmap.build_cdf_table), not {env_sampling!r}")
        if schedule not in SCHEDULES:
            raise ValueError(f"schedule must be 'megakernel' (one lane per path) or 'wavefront', not {schedule!r}")
        if builder not in ("lbvh", "ploc"):
            raise ValueError(f"builder must be 'lbvh' or 'ploc', not {builder!r}")
        if wide_walks not in WIDE_WALKS:
            raise ValueError(f"wide_walks must be one of {tuple(WIDE_WALKS)}, not {wide_walks!r}")
        if gradients not in ("ptir", "exact"):
            raise ValueError(f"gradients must be 'ptir' (the reference's backward) or 'exact', not {gradients!r}")
        # which backward the autograd function runs: the reference's path-replay backward
        # (material and environment gradients) or the exact adjoint (every parameter)
        self.gradients = gradients
        # layout of the environment table passed to trace / trace_bwd
        self.tab_kind = _lib.ENV_TABLE_ALIAS if env_sampling == "alias" else _lib.ENV_TABLE_CDF
        o = _lib.Options(float(particle_kernel_degree), float(particle_kernel_min_response),
                         int(bool(particle_kernel_density_clamping)), int(particle_radiance_sph_degree),
                         int(bool(enable_normals)), int(bool(enable_hitcounts)), int(bool(enable_mis)),
                         int(bool(enable_metallic)), int(bool(visualize_lights)), int(kbuffer),
                         {"lane": 0, "wave": 1}[traversal], SCHEDULES[schedule],
                         {"lbvh": 0, "ploc": 1}[builder], WIDE_WALKS[wide_walks])
        h = C.c_void_p()
        L = _lib.lib()
        _lib.check(L.ptir_create(C.byref(o), C.byref(h)), None)
        self._h = h
        self.options = o
        self.sph_max_degree = int(particle_radiance_sph_degree)
        self.last_counters: dict | None = None
        # the autograd function records a path tape in the forward so that the backward
        # replays paths without re-tracing them (same gradients, no BVH traversal)
        self.use_tape = bool(use_tape)
        self.tape_hits_per_path = int(tape_hits_per_path)

    def tape_bytes(self, B: int, H: int, W: int, max_bounces: int) -> int:
        return int(_lib.lib().ptir_tape_bytes(B, H, W, max_bounces, self.tape_hits_per_path))

    def new_tape(self, ray_dir: torch.Tensor, max_bounces: int) -> torch.Tensor:
        """Device buffer for the path tape of one trace call on rays of ray_dir's shape
        ([B,H,W,3], or flat [P,3] for a ragged launch)."""
        B, H, W = ray_dir.shape[:3] if ray_dir.dim() == 4 else (1, 1, ray_dir.numel() // 3)
        n = self.tape_bytes(B, H, W, int(max_bounces))
        if n == 0:
            raise ValueError("invalid launch size for a tape")
        return torch.empty(n, dtype=torch.uint8, device=ray_dir.device)

    def close(self) -> None:
        """Free the device buffers of this tracer (also done on garbage collection)."""
        h = getattr(self, "_h", None)
        self._h = None
        if h is not None and h.value:
            try:
                _lib.lib().ptir_destroy(h)
            except Exception:  # interpreter shutdown: the library may already be unloaded
                pass

    def __del__(self):
        self.close()

    def check(self) -> None:
        """Raise PtirError (PTIR_ERR_BVH_DEPTH) when a traversal stack overflowed in any launch
        since the last check; synchronises the current stream.  ptir_check."""
        _lib.check(_lib.lib().ptir_check(self._h, _stream()), self._h)

    def hierarchy_info(self) -> dict:
        """The last built hierarchy: its 4-wide node count and the traversal stack entries its
        walks can need (no counterpart in the reference).  ptir_hierarchy_info."""
        nodes, need = C.c_int32(0), C.c_int32(0)
        _lib.check(_lib.lib().ptir_hierarchy_info(self._h, C.byref(nodes), C.byref(need)), self._h)
        return {"wide_nodes": int(nodes.value), "stack_need": int(need.value)}

    # -------------------------------------------------------------- build_bvh
    def build_bvh(self, mog_pos, mog_rot, mog_scl, mog_dns, rebuild=True, allow_update=False):
        """OptixTracer::buildBVH (optixTracer.cpp:633-907)."""
        pos = _need(mog_pos, "mog_pos", (3,))
        n = pos.shape[0]
        rot = _need(mog_rot, "mog_rot", (4,))
        scl = _need(mog_scl, "mog_scl", (3,))
        dns = _need(mog_dns, "mog_dns").reshape(-1)
        if rot.shape[0] != n or scl.shape[0] != n or dns.shape[0] != n:
            raise ValueError("position, rotation, scale and density must have the same number of Gaussians")
        _lib.check(_lib.lib().ptir_build_bvh(self._h, _ptr(pos), _ptr(rot), _ptr(scl), _ptr(dns), n, int(bool(rebuild)),
                                             int(bool(allow_update)), _stream()), self._h)

    # -------------------------------------------------------------- shared
    def _scene(self, particle_density, particle_material, particle_radiance, particle_shading_normal, environment,
               environment_alias_table, lights, light_alias_table, mesh_lights):
        pd = _need(particle_density, "particle_density", (12,))
        n = pd.shape[0]
        pm = _need(particle_material, "particle_material", (5,))
        pr = _need(particle_radiance, "particle_radiance")
        sn = _need(particle_shading_normal, "particle_shading_normal", (3,))
        if pm.shape[0] != n or pr.shape[0] != n or sn.shape[0] != n:
            raise ValueError("particle tensors disagree on the number of Gaussians")
        keep = [pd, pm, pr, sn]
        sc = _lib.Scene()
        sc.num_gaussians = n
        sc.particle_density, sc.particle_material, sc.particle_radiance = _ptr(pd), _ptr(pm), _ptr(pr)
        sc.sh_stride = pr.shape[1] if pr.dim() > 1 else 0
        sc.shading_normal = _ptr(sn)
        env = environment
        if env is not None and env.numel() > 0 and env.dim() >= 3 and env.shape[0] > 0 and env.shape[1] > 0:
            env = _need(env, "environment", (4,))
            keep.append(env)
            sc.environment, sc.env_h, sc.env_w = _ptr(env), env.shape[0], env.shape[1]
        else:
            env = None
        tab = environment_alias_table
        if tab is not None and tab.dim() == 3 and tab.shape[0] == 3 and tab.numel() > 0:
            tab = _need(tab, "environment_alias_table")
            keep.append(tab)
            sc.env_table, sc.tab_h, sc.tab_w = _ptr(tab), tab.shape[1], tab.shape[2]
        sc.tab_kind = self.tab_kind
        sc.num_lights = 0 if lights is None else lights.shape[0]
        sc.num_light_entries = 0 if light_alias_table is None or light_alias_table.dim() < 2 else light_alias_table.shape[1]
        sc.num_mesh_lights = 0 if mesh_lights is None else mesh_lights.shape[0]
        return sc, keep, env

    @staticmethod
    def _launch(frame, ray_to_world, ray_ori, ray_dir, render_opts, sph_degree, min_transmittance, max_bounces,
                cache_mode=0, ragged: RaggedBatch | None = None, rr_start: int = 0):
        ro = _need(ray_ori, "ray_ori", (3,))
        rd = _need(ray_dir, "ray_dir", (3,))
        if ragged is not None:
            if ro.numel() != 3 * ragged.num_paths or rd.numel() != 3 * ragged.num_paths:
                raise ValueError(f"a ragged launch takes {ragged.num_paths} rays, got {ro.numel() // 3}")
            ro, rd = ro.reshape(1, 1, -1, 3), rd.reshape(1, 1, -1, 3)
        if ro.dim() != 4 or ro.shape != rd.shape:
            raise ValueError(f"rays must be [B,H,W,3] with equal shapes, got {tuple(ro.shape)} and {tuple(rd.shape)}")
        T = ray_to_world
        if not T.is_cuda:
            T = T.to(ro.device)
        T = _need(T.float(), "ray_to_world")
        if T.numel() < 12:
            raise ValueError("ray_to_world must hold at least 12 floats ([B,4,4])")
        la = _lib.Launch()
        la.frame, la.ray_to_world, la.ray_ori, la.ray_dir = int(frame) & 0xFFFFFFFF, _ptr(T), _ptr(ro), _ptr(rd)
        la.B, la.H, la.W = ro.shape[0], ro.shape[1], ro.shape[2]
        la.render_opts, la.sph_degree = int(render_opts), int(sph_degree)
        la.min_transmittance, la.max_bounces = float(min_transmittance), int(max_bounces)
        la.cache_mode = int(cache_mode)
        la.rr_start = int(rr_start)
        if ragged is not None:
            la.num_images, la.images = len(ragged.hw), _ptr(ragged.table)
            la.num_tiles, la.num_paths = ragged.num_tiles, ragged.num_paths
        return la, [ro, rd, T]

    # -------------------------------------------------------------- trace
    def trace(self, frame_id, ray_to_world, ray_ori, ray_dir, particle_density, particle_material,
              particle_radiance, particle_shading_normal, environment, environment_alias_table, lights=None,
              light_alias_table=None, mesh_light_vertices=None, mesh_light_triangles=None, mesh_lights=None,
              mesh_light_triangle_alias_table=None, render_opts=0, sph_degree=3, min_transmittance=1e-3,
              max_bounces=3, counters=False, tape: torch.Tensor | None = None, cache_mode: int = 0,
              cache_records: torch.Tensor | None = None, ragged: RaggedBatch | None = None, rr_start: int = 0):
        """OptixTracer::trace (optixTracer.cpp:909-1013): returns the 13 outputs
        (rad, dns, hit[2], hit2, distortion, nrm, snrm, material, hits, visibility, pbr,
        light, pbr_components).

        cache_mode / cache_records (no reference counterpart): 1 ends every path at its first
        secondary vertex and writes a radiance-cache query there, 2 traces on and writes the
        query with its training target; cache_records is [B,H,W,20] float32.

        ragged (no reference counterpart): trace the images of a RaggedBatch in one launch;
        rays are flat [num_paths, 3] (world space) and the outputs [1, 1, num_paths, C].
        rr_start (no reference counterpart): Russian roulette from that path vertex on."""
        sc, keep, _ = self._scene(particle_density, particle_material, particle_radiance, particle_shading_normal,
                                  environment, environment_alias_table, lights, light_alias_table, mesh_lights)
        la, keep2 = self._launch(frame_id, ray_to_world, ray_ori, ray_dir, render_opts, sph_degree,
                                 min_transmittance, max_bounces, cache_mode, ragged, rr_start)
        if ragged is not None:
            keep2.append(ragged.table)
        B, H, W = la.B, la.H, la.W
        dev = ray_dir.device
        f = dict(dtype=torch.float32, device=dev)
        rad = torch.empty(B, H, W, 3, **f)
        dns = torch.empty(B, H, W, 1, **f)
        hit = torch.empty(B, H, W, 2, **f)
        hit2 = torch.empty(B, H, W, 1, **f)
        dist = torch.empty(B, H, W, 1, **f)
        nrm = torch.empty(B, H, W, 3, **f)
        snrm = torch.empty(B, H, W, 3, **f)
        mat = torch.empty(B, H, W, 5, **f)
        hits = torch.empty(B, H, W, 1, **f)
        vis = torch.zeros(sc.num_gaussians, 1, **f)
        pbr = torch.empty(B, H, W, 3, **f)
        light = torch.empty(B, H, W, 3, **f)
        comp = torch.empty(B, H, W, 2, 3, **f)
        out = _lib.FwdOutputs(*[_ptr(t) for t in (rad, dns, hit, hit2, dist, nrm, snrm, mat, hits, vis, pbr, light, comp)])
        if tape is not None:
            tape = _need(tape, "tape", dtype=torch.uint8)
            out.tape, out.tape_bytes, out.tape_hits_per_path = _ptr(tape), tape.numel(), self.tape_hits_per_path
        if cache_records is not None:
            if tuple(cache_records.shape) != (B, H, W, 20):
                raise ValueError(f"cache_records must be [{B},{H},{W},20], got {tuple(cache_records.shape)}")
            if not cache_records.is_contiguous():
                raise ValueError("cache_records must be contiguous (the tracer writes into it)")
            out.cache_records = _ptr(_need(cache_records, "cache_records"))
        cnt = _lib.Counters() if counters else None
        st = _lib.lib().ptir_trace(self._h, C.byref(sc), C.byref(la), C.byref(out),
                                   C.byref(cnt) if cnt is not None else None, _stream())
        _lib.check(st, self._h)
        del keep, keep2
        if cnt is not None:
            self.last_counters = cnt.as_dict()
        return rad, dns, hit, hit2, dist, nrm, snrm, mat, hits, vis, pbr, light, comp

    # -------------------------------------------------------------- trace_bwd
    def trace_bwd(self, frame_id, ray_to_world, ray_ori, ray_dir, ray_rad, ray_dns, ray_hit, ray_hit2, ray_dist,
                  ray_nrm, ray_snrm, ray_mat, ray_pbr, ray_light, particle_density, particle_material,
                  particle_radiance, particle_shading_normal, environment, environment_alias_table, lights,
                  light_alias_table, mesh_light_vertices, mesh_light_triangles, mesh_lights,
                  mesh_light_triangle_alias_table, ray_rad_grd, ray_dns_grd, ray_hit_grd, ray_hit2_grd, ray_dist_grd,
                  ray_nrm_grd, ray_snrm_grd, ray_mat_grd, ray_pbr_grd, ray_light_grd, render_opts=0, sph_degree=3,
                  min_transmittance=1e-3, max_bounces=3, counters=False, tape: torch.Tensor | None = None,
                  cache_mode: int = 0, ragged: RaggedBatch | None = None, rr_start: int = 0):
        """OptixTracer::traceBwd (optixTracer.cpp:1015-1145): returns (dDensity [N,12],
        dMaterial [N,5], dSH, dShadingNormal [N,3], dEnv [H,W,4]).  cache_mode must be the
        forward's (with 1 the replay ends at the first secondary vertex, where the forward
        handed the path to the radiance cache)."""
        sc, keep, env = self._scene(particle_density, particle_material, particle_radiance, particle_shading_normal,
                                    environment, environment_alias_table, lights, light_alias_table, mesh_lights)
        la, keep2 = self._launch(frame_id, ray_to_world, ray_ori, ray_dir, render_opts, sph_degree,
                                 min_transmittance, max_bounces, cache_mode, ragged, rr_start)
        if ragged is not None:
            keep2.append(ragged.table)
        B, H, W = la.B, la.H, la.W

        def pix(t, name, c, allow_none=False):
            if t is None:
                if allow_none:
                    return None
                raise ValueError(f"{name} is required")
            t = _need(t, name)
            if tuple(t.shape) != (B, H, W, c):
                raise ValueError(f"{name} must be [{B},{H},{W},{c}], got {tuple(t.shape)}")
            return t

        pbr = pix(ray_pbr, "ray_pbr", 3)
        light = pix(ray_light, "ray_light", 3)
        pbr_g = pix(ray_pbr_grd, "ray_pbr_grd", 3)
        light_g = pix(ray_light_grd, "ray_light_grd", 3, True)
        mat_g = pix(ray_mat_grd, "ray_mat_grd", 5, True)
        n = sc.num_gaussians
        f = dict(dtype=torch.float32, device=ray_dir.device)
        d_dns = torch.empty(n, particle_density.shape[1], **f)
        d_mat = torch.empty(n, particle_material.shape[1], **f)
        d_sh = torch.empty(n, particle_radiance.shape[1], **f)
        d_sn = torch.empty(n, particle_shading_normal.shape[1], **f)
        d_env = torch.empty_like(environment) if environment is not None else torch.zeros(0, 0, 4, **f)
        bi = _lib.BwdInputs(_ptr(pbr), _ptr(light), _ptr(pbr_g), _ptr(light_g), _ptr(mat_g))
        if tape is not None:
            tape = _need(tape, "tape", dtype=torch.uint8)
            bi.tape, bi.tape_bytes, bi.tape_hits_per_path = _ptr(tape), tape.numel(), self.tape_hits_per_path
        bo = _lib.BwdOutputs(_ptr(d_dns), _ptr(d_mat), _ptr(d_sh), _ptr(d_sn), _ptr(d_env) if env is not None else None)
        if env is None and d_env.numel():
            d_env.zero_()
        cnt = _lib.Counters() if counters else None
        st = _lib.lib().ptir_trace_bwd(self._h, C.byref(sc), C.byref(la), C.byref(bi), C.byref(bo),
                                       C.byref(cnt) if cnt is not None else None, _stream())
        _lib.check(st, self._h)
        del keep, keep2
        if cnt is not None:
            self.last_counters = cnt.as_dict()
        return d_dns, d_mat, d_sh, d_sn, d_env


    # -------------------------------------------------------------- exact adjoint
    def trace_bwd_exact(self, frame_id, ray_to_world, ray_ori, ray_dir, particle_density, particle_material,
                        particle_radiance, particle_shading_normal, environment, environment_alias_table,
                        ray_pbr_grd, ray_light_grd=None, ray_mat_grd=None, ray_rad_grd=None, ray_dns_grd=None,
                        ray_hit_grd=None, ray_snrm_grd=None, ray_comp_grd=None, render_opts=0, sph_degree=3,
                        min_transmittance=1e-3, max_bounces=3, ragged: RaggedBatch | None = None,
                        rr_start: int = 0):
        """Exact-adjoint backward (no counterpart in the reference; ptir_trace_bwd_exact): the
        gradients (dDensity [N,12], dMaterial [N,5], dSH, dShadingNormal [N,3], dEnv [H,W,4])
        of sum <output_grad, output> over the given output gradients, every one of them
        propagated to every parameter with the paths' discrete decisions held fixed.
        ray_hit_grd may hold one channel (integrated depth) or two (the second is ignored);
        ray_comp_grd is [B,H,W,2,3] (direct, indirect).  The launch arguments must be the
        forward's."""
        sc, keep, env = self._scene(particle_density, particle_material, particle_radiance, particle_shading_normal,
                                    environment, environment_alias_table, None, None, None)
        la, keep2 = self._launch(frame_id, ray_to_world, ray_ori, ray_dir, render_opts, sph_degree,
                                 min_transmittance, max_bounces, 0, ragged, rr_start)
        if ragged is not None:
            keep2.append(ragged.table)
        B, H, W = la.B, la.H, la.W

        def pix(t, name, c):
            if t is None:
                return None
            t = _need(t, name)
            if t.numel() != B * H * W * c:
                raise ValueError(f"{name} must hold [{B},{H},{W},{c}] values, got {tuple(t.shape)}")
            return t

        if ray_hit_grd is not None and ray_hit_grd.numel() == B * H * W:
            ray_hit_grd = torch.cat([ray_hit_grd.reshape(B, H, W, 1), torch.zeros_like(ray_hit_grd.reshape(B, H, W, 1))],
                                    dim=-1)
        if ray_pbr_grd is None:
            raise ValueError("ray_pbr_grd is required")
        grads = [pix(ray_pbr_grd, "ray_pbr_grd", 3), pix(ray_light_grd, "ray_light_grd", 3),
                 pix(ray_mat_grd, "ray_mat_grd", 5), pix(ray_rad_grd, "ray_rad_grd", 3),
                 pix(ray_dns_grd, "ray_dns_grd", 1), pix(ray_hit_grd, "ray_hit_grd", 2),
                 pix(ray_snrm_grd, "ray_snrm_grd", 3), pix(ray_comp_grd, "ray_comp_grd", 6)]
        n = sc.num_gaussians
        f = dict(dtype=torch.float32, device=ray_dir.device)
        d_dns = torch.empty(n, particle_density.shape[1], **f)
        d_mat = torch.empty(n, particle_material.shape[1], **f)
        d_sh = torch.empty(n, particle_radiance.shape[1], **f)
        d_sn = torch.empty(n, particle_shading_normal.shape[1], **f)
        d_env = torch.empty_like(environment) if env is not None else torch.zeros(0, 0, 4, **f)
        bi = _lib.BwdInputs(None, None, _ptr(grads[0]), _ptr(grads[1]), _ptr(grads[2]))
        ex = _lib.ExactGrads(*[_ptr(g) for g in grads[3:]])
        bo = _lib.BwdOutputs(_ptr(d_dns), _ptr(d_mat), _ptr(d_sh), _ptr(d_sn), _ptr(d_env) if env is not None else None)
        st = _lib.lib().ptir_trace_bwd_exact(self._h, C.byref(sc), C.byref(la), C.byref(bi), C.byref(ex),
                                             C.byref(bo), _stream())
        _lib.check(st, self._h)
        del keep, keep2, grads
        return d_dns, d_mat, d_sh, d_sn, d_env


# ----------------------------------------------------------------------------- light buffers
@dataclass(frozen=True)
class LightBuffers:
    """Packed scene lights (tracer.py:40-119).  Only empty buffers are accepted by the
    tracer; point/sphere/mesh lights raise PTIR_ERR_UNSUPPORTED."""

    lights: torch.Tensor
    alias_table: torch.Tensor
    mesh_vertices: torch.Tensor
    mesh_triangles: torch.Tensor
    mesh_lights: torch.Tensor
    mesh_triangle_alias_table: torch.Tensor

    @classmethod
    def empty(cls, device) -> "LightBuffers":
        f = dict(dtype=torch.float32, device=device)
        return cls(torch.empty((0, 9), **f), torch.empty((5, 0), **f), torch.empty((0, 3), **f),
                   torch.empty((0, 3), dtype=torch.int32, device=device), torch.empty((0, 8), **f),
                   torch.empty((3, 0), **f))

    @classmethod
    def from_batch(cls, batch, device) -> "LightBuffers":
        e = cls.empty(device)

        def get(name, fb):
            v = getattr(batch, name, None)
            return fb if v is None else v

        return cls(get("lights", e.lights), get("light_alias_table", e.alias_table),
                   get("mesh_light_vertices", e.mesh_vertices), get("mesh_light_triangles", e.mesh_triangles),
                   get("mesh_lights", e.mesh_lights),
                   get("mesh_light_triangle_alias_table", e.mesh_triangle_alias_table)).to(device)

    def to(self, device) -> "LightBuffers":
        return LightBuffers(self.lights.to(device=device, dtype=torch.float32).contiguous(),
                            self.alias_table.to(device=device, dtype=torch.float32).contiguous(),
                            self.mesh_vertices.to(device=device, dtype=torch.float32).contiguous(),
                            self.mesh_triangles.to(device=device, dtype=torch.int32).contiguous(),
                            self.mesh_lights.to(device=device, dtype=torch.float32).contiguous(),
                            self.mesh_triangle_alias_table.to(device=device, dtype=torch.float32).contiguous())

    def tensors(self):
        return (self.lights, self.alias_table, self.mesh_vertices, self.mesh_triangles, self.mesh_lights,
                self.mesh_triangle_alias_table)


def _conf_get(node, key, default=None):
    if node is None:
        return default
    if isinstance(node, dict):
        return node.get(key, default)
    getter = getattr(node, "get", None)
    if callable(getter):
        try:
            return getter(key, default)
        except TypeError:
            pass
    return getattr(node, key, default)


# ----------------------------------------------------------------------------- Tracer
class Tracer:
    class _Autograd(torch.autograd.Function):
        """tracer.py:137-384: packs particle tensors, calls trace / trace_bwd and splits
        the gradients back onto the parameters."""

        @staticmethod
        def forward(ctx, tracer_wrapper, frame_id, ray_to_world, ray_ori, ray_dir, mog_pos, mog_rot, mog_scl,
                    mog_dns, mog_sph, mog_snrm, mog_malb, mog_mrgh, mog_mmet, environment, environment_alias_table,
                    light_buffers, render_opts, sph_degree, min_transmittance, max_bounces):
            particle_density = torch.concat([mog_pos, mog_dns, mog_rot, mog_scl, torch.zeros_like(mog_dns)], dim=1)
            particle_material = torch.concat([mog_malb, mog_mrgh, mog_mmet], dim=1)
            light_buffers = light_buffers.to(ray_dir.device)
            lights, lat, mlv, mlt, ml, mlta = light_buffers.tensors()
            tape = None
            exact = getattr(tracer_wrapper, "gradients", "ptir") == "exact"
            if not exact and getattr(tracer_wrapper, "use_tape", False) and any(ctx.needs_input_grad):
                tape = tracer_wrapper.new_tape(ray_dir, max_bounces)
            (ray_radiance, ray_density, ray_hit_distance, ray_hit_distance_second_moment, ray_depth_distortion,
             ray_normals, ray_shadingnormal, ray_material, hits_count, mog_visibility, ray_pbr, ray_light,
             pbr_components) = tracer_wrapper.trace(frame_id, ray_to_world, ray_ori, ray_dir, particle_density,
                                                    particle_material, mog_sph, mog_snrm, environment,
                                                    environment_alias_table, lights, lat, mlv, mlt, ml, mlta,
                                                    render_opts, sph_degree, min_transmittance, max_bounces,
                                                    **({"tape": tape} if tape is not None else {}))
            ctx.tape = tape
            ctx.save_for_backward(ray_to_world, ray_ori, ray_dir, ray_radiance, ray_density, ray_hit_distance,
                                  ray_hit_distance_second_moment, ray_depth_distortion, ray_normals,
                                  ray_shadingnormal, ray_material, ray_pbr, ray_light, particle_density,
                                  particle_material, mog_sph, mog_snrm, environment, environment_alias_table, lights,
                                  lat, mlv, mlt, ml, mlta)
            ctx.frame_id = frame_id
            ctx.render_opts = render_opts
            ctx.sph_degree = sph_degree
            ctx.min_transmittance = min_transmittance
            ctx.max_bounces = max_bounces
            ctx.tracer_wrapper = tracer_wrapper
            return (ray_radiance, ray_density, ray_hit_distance[:, :, :, 0:1], ray_hit_distance_second_moment,
                    ray_depth_distortion, ray_normals, ray_shadingnormal, ray_material, hits_count, mog_visibility,
                    ray_pbr, ray_light, pbr_components)

        @staticmethod
        def backward(ctx, ray_radiance_grd, ray_density_grd, ray_hit_distance_grd, ray_hit_distance_second_moment_grd,
                     ray_depth_distortion_grd, ray_normals_grd, ray_shadingnormal_grd, ray_material_grd,
                     ray_hits_count_grd_UNUSED, mog_visibility_grd_UNUSED, ray_pbr_grd, ray_light_grd,
                     pbr_components_grd):
            (ray_to_world, ray_ori, ray_dir, ray_radiance, ray_density, ray_hit_distance,
             ray_hit_distance_second_moment, ray_depth_distortion, ray_normals, ray_shadingnormal, ray_material,
             ray_pbr, ray_light, particle_density, particle_material, mog_sph, mog_snrm, environment,
             environment_alias_table, lights, lat, mlv, mlt, ml, mlta) = ctx.saved_tensors
            if ray_light_grd is None:
                ray_light_grd = torch.zeros_like(ray_light)
            if ray_pbr_grd is None:
                ray_pbr_grd = torch.zeros_like(ray_pbr)
            if ray_material_grd is not None:
                ray_material_grd = ray_material_grd.contiguous()
            if getattr(ctx.tracer_wrapper, "gradients", "ptir") == "exact":
                c = lambda t: None if t is None else t.contiguous()  # noqa: E731
                (particle_density_grd, particle_material_grd, mog_sph_grd, mog_sn_grd,
                 environment_grd) = ctx.tracer_wrapper.trace_bwd_exact(
                    ctx.frame_id, ray_to_world, ray_ori, ray_dir, particle_density, particle_material, mog_sph,
                    mog_snrm, environment, environment_alias_table, ray_pbr_grd.contiguous(),
                    ray_light_grd.contiguous(), ray_material_grd, c(ray_radiance_grd), c(ray_density_grd),
                    c(ray_hit_distance_grd), c(ray_shadingnormal_grd), c(pbr_components_grd),
                    ctx.render_opts, ctx.sph_degree, ctx.min_transmittance, ctx.max_bounces)
                return Tracer._Autograd._split(particle_density_grd, particle_material_grd, mog_sph_grd, mog_sn_grd,
                                               environment_grd)
            (particle_density_grd, particle_material_grd, mog_sph_grd, mog_sn_grd,
             environment_grd) = ctx.tracer_wrapper.trace_bwd(
                ctx.frame_id, ray_to_world, ray_ori, ray_dir, ray_radiance, ray_density, ray_hit_distance,
                ray_hit_distance_second_moment, ray_depth_distortion, ray_normals, ray_shadingnormal, ray_material,
                ray_pbr, ray_light, particle_density, particle_material, mog_sph, mog_snrm, environment,
                environment_alias_table, lights, lat, mlv, mlt, ml, mlta, ray_radiance_grd, ray_density_grd,
                ray_hit_distance_grd, ray_hit_distance_second_moment_grd, ray_depth_distortion_grd, ray_normals_grd,
                ray_shadingnormal_grd, ray_material_grd, ray_pbr_grd.contiguous(), ray_light_grd.contiguous(),
                ctx.render_opts, ctx.sph_degree, ctx.min_transmittance, ctx.max_bounces,
                **({"tape": ctx.tape} if ctx.tape is not None else {}))
            ctx.tape = None
            return Tracer._Autograd._split(particle_density_grd, particle_material_grd, mog_sph_grd, mog_sn_grd,
                                           environment_grd)

        @staticmethod
        def _split(particle_density_grd, particle_material_grd, mog_sph_grd, mog_sn_grd, environment_grd):
            mog_pos_grd, mog_dns_grd, mog_rot_grd, mog_scl_grd, _ = torch.split(particle_density_grd, [3, 1, 4, 3, 1],
                                                                                dim=1)
            mog_malb_grd, mog_mrgh_grd, mog_mmet_grd = torch.split(particle_material_grd, [3, 1, 1], dim=1)
            return (None, None, None, None, None, mog_pos_grd, mog_rot_grd, mog_scl_grd, mog_dns_grd, mog_sph_grd,
                    mog_sn_grd, mog_malb_grd, mog_mrgh_grd, mog_mmet_grd, environment_grd, None, None, None, None,
                    None, None)

    class RenderOpts(IntEnum):
        NONE = 0
        INDIRECT = 1
        DEFAULT = NONE

    def __init__(self, conf):
        """Tracer(conf) of the reference (tracer.py:428-469): conf.render.* configures the
        native tracer; conf.loss.* decides which outputs a training render must produce."""
        self.device = "cuda"
        self.conf = conf
        r = _conf_get(conf, "render")
        self.render_conf = r
        self.loss_conf = _conf_get(conf, "loss")
        self.filter_iterations = render.filter_iterations(_conf_get(r, "filter_type", "none"))
        self.visualize_lights = bool(_conf_get(r, "visualize_lights", False))
        self.density_clamping = bool(_conf_get(r, "particle_kernel_density_clamping", True))
        self.max_bvh_updates = int(_conf_get(r, "max_consecutive_bvh_update", 15))
        metallic = _conf_get(r, "enable_metallic", None)
        if metallic is None:
            metallic = _conf_get(_conf_get(conf, "model"), "optimize_material_metallic", False)
        self.tracer_wrapper = HipTracer(
            "", "", _conf_get(r, "pipeline_type", "reference"), _conf_get(r, "backward_pipeline_type", "referenceBwd"),
            _conf_get(r, "primitive_type", "instances"), _conf_get(r, "particle_kernel_degree", 4),
            _conf_get(r, "particle_kernel_min_response", 0.0113), self.density_clamping,
            _conf_get(r, "particle_radiance_sph_degree", 3), _conf_get(r, "enable_normals", True),
            _conf_get(r, "enable_hitcounts", True), _conf_get(r, "enable_mis", True), bool(metallic),
            self.visualize_lights, kbuffer=int(_conf_get(r, "kbuffer_depth", 16)),
            env_sampling=str(_conf_get(r, "env_sampling", "alias")),
            traversal=str(_conf_get(r, "primary_traversal", "wave")),
            schedule=str(_conf_get(r, "schedule", "megakernel")),
            builder=str(_conf_get(r, "bvh_builder", "lbvh")),
            wide_walks=str(_conf_get(r, "wide_walks", "none")))
        # report a traversal-stack overflow after each render (one stream synchronisation;
        # skipped while a graph is being captured)
        self.check_errors = bool(_conf_get(r, "check_errors", True))
        self.updates_since_rebuild = 0
        self._built_geometry = None
        self._spp_fallback_warned = False
        self.timings = {}

    # -------------------------------------------------------------- acceleration structure
    @staticmethod
    def _geometry_key(gaussians):
        """Identity of the geometry tensors: storage, shape and in-place version counter.
        An optimiser step bumps the version of every parameter it updates."""
        key = []
        for name in ("positions", "rotation", "scale", "density"):
            t = getattr(gaussians, name)
            key.append((t.data_ptr(), tuple(t.shape), t._version))
        return tuple(key)

    def build_acc(self, gaussians, rebuild=True):
        """build_acc of the reference (tracer.py:716-738): a full rebuild when asked, when
        density clamping is on, or after max_consecutive_bvh_update refits, else a refit.
        When the geometry tensors are exactly those of the last build (PTIR's inversion stage
        freezes geometry, configs/base_ptir.yaml:13), nothing is rebuilt: the hierarchy
        already describes them."""
        key = self._geometry_key(gaussians)
        if key == self._built_geometry:
            return
        full = bool(rebuild) or self.density_clamping or self.updates_since_rebuild >= self.max_bvh_updates
        allow_update = self.max_bvh_updates > 1 and not self.density_clamping
        if self._built_geometry is None:
            full = True
        self.tracer_wrapper.build_bvh(
            gaussians.positions.view(-1, 3).contiguous(),
            gaussians.rotation_activation(gaussians.rotation).view(-1, 4).contiguous(),
            gaussians.scale_activation(gaussians.scale).view(-1, 3).contiguous(),
            gaussians.density_activation(gaussians.density).view(-1, 1).contiguous(), full, allow_update)
        self.updates_since_rebuild = 0 if full else self.updates_since_rebuild + 1
        self._built_geometry = key

    # -------------------------------------------------------------- render
    def _wanted_outputs(self, train: bool):
        """(outputs returned, outputs accumulated) of a render; None = all.  Only a training
        render with render.hide_intermediate_outputs drops outputs (tracer.py:470-503)."""
        if not train or not bool(_conf_get(self.render_conf, "hide_intermediate_outputs", False)):
            return None, None
        keep = {"pred_pbr", "pred_material"}
        lc = self.loss_conf
        if bool(_conf_get(lc, "use_light_consistency", False)) or float(_conf_get(lc, "lambda_light", 0.0)) > 0.0:
            keep.add("pred_light")
        if bool(_conf_get(lc, "use_edge_aware_smoothness", False)):
            keep.update(_conf_get(lc, "edge_aware_smoothness_outputs", []) or [])
        acc = set(keep)
        if "pred_pbr" in acc:
            acc.add("pred_opacity")
        if acc & {"pred_direct", "pred_indirect"}:
            acc.update(("pbr_components", "pred_opacity"))
        return keep, acc

    def _spp_plan(self, batch, train: bool, frame_id: int) -> render.SppPlan:
        spp = max(1, int(_conf_get(self.render_conf, "inversion_spp" if train else "render_spp", 1)))
        dev = batch.rays_dir.device
        intr = pinhole = None
        if spp > 1:
            pinhole = getattr(batch, "pixel_coords", None) is not None
            intr = render.pinhole_intrinsics(batch, dev) if pinhole else None
            if intr is None:
                if not self._spp_fallback_warned:
                    logger.warning("spp %d needs pixel_coords and distortion-free pinhole intrinsics on the batch; "
                                   "rendering 1 sample per pixel", spp)
                    self._spp_fallback_warned = True
                spp = 1
        chunk = min(spp, max(1, int(_conf_get(self.render_conf, "spp_chunk", spp))))
        pos = render.sample_positions(spp, frame_id, dev, batch.rays_dir.dtype) if spp > 1 else None
        return render.SppPlan(spp, chunk, pos, intr)

    @staticmethod
    def _environment(gaussians, tracer_wrapper, device):
        env = gaussians.get_environment()
        if env is None:
            return torch.empty((0, 0, 4), dtype=torch.float32, device=device), torch.empty(0, device=device)
        if tracer_wrapper.tab_kind == _lib.ENV_TABLE_CDF:
            return env, envmap.build_cdf_table(env.detach().contiguous())
        t = getattr(gaussians, "environment_alias_table", None)
        if t is None:
            return env, torch.empty(0, dtype=torch.float32, device=device)
        planes = [t.prob, t.alias.to(torch.float32), t.pdf]
        table = torch.stack([p.reshape(t.height, t.width).float() for p in planes]).to(device).contiguous()
        return env, table

    def render(self, gaussians, gpu_batch, train=False, frame_id=0, sh_indirect: bool = False):
        """Tracer.render of the reference (tracer.py:740-975): samples per pixel in chunks,
        each chunk one trace through the autograd function, the model's background on the
        radiance, the spp average, then the resolve of PBR outputs on the device."""
        keep, acc_keys = self._wanted_outputs(train)
        base = gpu_batch.rays_dir.shape[0]
        dev = gpu_batch.rays_dir.device
        plan = self._spp_plan(gpu_batch, train, frame_id)
        env, table = self._environment(gaussians, self.tracer_wrapper, dev)
        lights = LightBuffers.from_batch(gpu_batch, device=dev)
        to_world = gpu_batch.T_to_world.contiguous()
        opts = int(Tracer.RenderOpts.INDIRECT if sh_indirect else Tracer.RenderOpts.DEFAULT)
        names = ("pred_rgb", "pred_opacity", "pred_dist", "pred_depth_second_moment", "pred_depth_distortion",
                 "pred_normals", "pred_shadingnormal", "pred_material", "hits_count", "pred_pbr", "pred_light",
                 "pbr_components")
        sums: dict = {}
        visibility = None
        for start, count in plan.chunks():
            if plan.positions is None:
                ori, dirs = gpu_batch.rays_ori.contiguous(), gpu_batch.rays_dir.contiguous()
            else:
                ori, dirs = render.camera_rays(gpu_batch, plan.positions[start:start + count], plan.intrinsics)
            out = Tracer._Autograd.apply(
                self.tracer_wrapper, frame_id + start, to_world, ori, dirs, gaussians.positions.contiguous(),
                gaussians.get_rotation().contiguous(), gaussians.get_scale().contiguous(),
                gaussians.get_density().contiguous(), gaussians.get_features().contiguous(),
                gaussians.get_shading_normal().contiguous(), gaussians.get_material_albedo().contiguous(),
                gaussians.get_material_roughness().contiguous(), gaussians.get_material_metallic().contiguous(),
                env, table, lights, opts, gaussians.n_active_features,
                _conf_get(self.render_conf, "min_transmittance", 1e-3), _conf_get(self.render_conf, "max_bounces", 3))
            rgb, opacity = gaussians.background(to_world, dirs, out[0], out[1], train)
            values = (rgb, opacity, *out[2:9], out[10], out[11], out[12].detach())
            for name, v in zip(names, values):
                if acc_keys is not None and name not in acc_keys:
                    continue
                # samples of a chunk are the leading factor of the batch: sum them
                total = v.reshape(count, base, *v.shape[1:]).sum(dim=0)
                sums[name] = total if name not in sums else sums[name] + total
            visibility = out[9] if visibility is None else torch.maximum(visibility, out[9])
        mean = {k: v / plan.spp for k, v in sums.items()}
        res = self._resolve(mean)
        out = {k: res[k] for k in ("pred_rgb", "pred_opacity", "pred_dist", "pred_depth_second_moment",
                                   "pred_depth_distortion", "pred_normals", "pred_shadingnormal", "pred_material",
                                   "pred_pbr", "pred_light", "pred_direct", "pred_indirect", "hits_count")
               if res.get(k) is not None and (keep is None or k in keep)}
        out["frame_time_ms"] = 0.0
        out["mog_visibility"] = visibility
        if self.check_errors and not torch.cuda.is_current_stream_capturing():
            self.tracer_wrapper.check()
        return render.apply_gt_mask(out, gpu_batch)

    def _resolve(self, mean: dict) -> dict:
        """Depth / opacity, opacity-weighted and filtered PBR terms, unit normals."""
        res = dict(mean)
        comp = res.pop("pbr_components", None)
        op = mean.get("pred_opacity")
        if op is None:
            if comp is not None:
                res["pred_direct"], res["pred_indirect"] = comp[..., 0, :], comp[..., 1, :]
            return res
        z3 = torch.zeros(op.shape[:3] + (3,), dtype=torch.float32, device=op.device)
        pick = lambda k: mean[k] if mean.get(k) is not None else z3  # noqa: E731
        direct = comp[..., 0, :] if comp is not None else z3
        indirect = comp[..., 1, :] if comp is not None else z3
        dist = mean.get("pred_dist", torch.zeros_like(op))
        d, pbr, dr, ind, light, nrm = render.Resolve.apply(self.filter_iterations, op, dist, pick("pred_pbr"), direct,
                                                           indirect, pick("pred_light"), pick("pred_normals"))
        for k, v in (("pred_dist", d), ("pred_pbr", pbr), ("pred_light", light), ("pred_normals", nrm)):
            if mean.get(k) is not None:
                res[k] = v
        if comp is not None:
            res["pred_direct"], res["pred_indirect"] = dr, ind
        return res
