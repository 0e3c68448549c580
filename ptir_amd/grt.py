"""Drop-in for the 3DGRT primary-only tracer (threedgrt_tracer/ of the reference, the stage-1
tracer PTIR trains geometry with): `GRTTracer` mirrors the pybind class `OptixTracer`
(threedgrt_tracer/bindings.cpp:32-38) and `Tracer` mirrors threedgrt_tracer/tracer.py:49-334.

It runs on the same C-ABI handle, hierarchy (LBVH or PLOC) and k-buffer as the path tracer
(`ptir_grt_trace`, `ptir_grt_trace_bwd` of libptir_amd.so); the backward gives gradients of
positions, densities, quaternions, scales, SH coefficients and shading normals.
"""

from __future__ import annotations

import ctypes as C
from enum import IntEnum

import torch

from . import _lib
from .tracer import HipTracer, _conf_get, _need, _ptr, _stream


class GRTTracer:
    """OptixTracer of threedgrt_tracer (src/optixTracer.cpp:222-1012)."""

    def __init__(self, path: str = "", cuda_path: str = "", pipeline: str = "reference",
                 backward_pipeline: str = "referenceBwd", primitive: str = "instances",
                 particle_kernel_degree: float = 4.0, particle_kernel_min_response: float = 0.0113,
                 particle_kernel_density_clamping: bool = True, particle_radiance_sph_degree: int = 3,
                 enable_normals: bool = True, enable_hitcounts: bool = True, *, kbuffer: int = 16,
                 traversal: str = "wave", builder: str = "lbvh", wide_walks: str = "none"):
        # the path tracer's handle: ptir_create and ptir_build_bvh are shared
        self._t = HipTracer(path, cuda_path, pipeline, backward_pipeline, primitive, particle_kernel_degree,
                            particle_kernel_min_response, particle_kernel_density_clamping,
                            particle_radiance_sph_degree, enable_normals, enable_hitcounts, False, False, False,
                            kbuffer=kbuffer, traversal=traversal, builder=builder, wide_walks=wide_walks)

    def build_bvh(self, mog_pos, mog_rot, mog_scl, mog_dns, rebuild=True, allow_update=False):
        """OptixTracer::buildBVH (threedgrt_tracer/src/optixTracer.cpp:578-851)."""
        self._t.build_bvh(mog_pos, mog_rot, mog_scl, mog_dns, rebuild, allow_update)

    def check(self) -> None:
        self._t.check()

    @staticmethod
    def _scene(particle_density, particle_radiance, particle_shading_normal):
        pd = _need(particle_density, "particle_density", (12,))
        pr = _need(particle_radiance, "particle_radiance")
        n = pd.shape[0]
        keep = [pd, pr]
        sc = _lib.GrtScene()
        sc.num_gaussians = n
        sc.particle_density, sc.particle_radiance = _ptr(pd), _ptr(pr)
        sc.sh_stride = pr.shape[1] if pr.dim() > 1 else 0
        if pr.shape[0] != n:
            raise ValueError("particle tensors disagree on the number of Gaussians")
        if particle_shading_normal is not None and particle_shading_normal.numel() > 0:
            sn = _need(particle_shading_normal, "particle_shading_normal", (3,))
            if sn.shape[0] != n:
                raise ValueError("particle tensors disagree on the number of Gaussians")
            keep.append(sn)
            sc.shading_normal = _ptr(sn)
        return sc, keep

    def new_tape(self, ray_dir, hits_per_pixel: int = 128) -> torch.Tensor:
        """A hit tape for launches of ray_dir's [B,H,W] shape: passed as `tape=` to trace, it
        receives each pixel's accepted hits (up to hits_per_pixel); passed to trace_bwd, the
        backward reads them instead of re-tracing (pixels with more hits re-trace, and each
        such pixel costs its whole wave a hierarchy walk, so the default is generous)."""
        B, H, W = ray_dir.shape[:3]
        n = _lib.lib().ptir_grt_tape_bytes(B, H, W, hits_per_pixel)
        if n == 0:
            raise ValueError("invalid tape size")
        t = torch.empty((n + 3) // 4, dtype=torch.int32, device=ray_dir.device)
        t.hits_per_pixel = hits_per_pixel
        return t

    def trace(self, frame_id, ray_to_world, ray_ori, ray_dir, particle_density, particle_radiance,
              particle_shading_normal, render_opts=0, sph_degree=3, min_transmittance=1e-3, *, tape=None):
        """OptixTracer::trace (threedgrt_tracer/src/optixTracer.cpp:855-925): rad, dns,
        hit[2], hit2, distortion, nrm, snrm, hits, visibility.  tape: a new_tape() to record
        the hits in (no counterpart in the reference)."""
        sc, keep = self._scene(particle_density, particle_radiance, particle_shading_normal)
        la, keep2 = HipTracer._launch(frame_id, ray_to_world, ray_ori, ray_dir, render_opts, sph_degree,
                                      min_transmittance, 1)
        B, H, W = la.B, la.H, la.W
        f = dict(dtype=torch.float32, device=ray_dir.device)
        rad = torch.empty(B, H, W, 3, **f)
        dns = torch.empty(B, H, W, 1, **f)
        hit = torch.empty(B, H, W, 2, **f)
        hit2 = torch.empty(B, H, W, 1, **f)
        dist = torch.empty(B, H, W, 1, **f)
        nrm = torch.empty(B, H, W, 3, **f)
        snrm = torch.empty(B, H, W, 3, **f)
        hits = torch.empty(B, H, W, 1, **f)
        vis = torch.zeros(sc.num_gaussians, 1, **f)
        out = _lib.GrtOutputs(*[_ptr(t) for t in (rad, dns, hit, hit2, dist, nrm, snrm, hits, vis)])
        if tape is not None:
            if tape.numel() * 4 < _lib.lib().ptir_grt_tape_bytes(B, H, W, tape.hits_per_pixel):
                raise ValueError("tape is too small for this launch")
            out.hit_tape, out.hit_tape_cap = _ptr(tape), tape.hits_per_pixel
        _lib.check(_lib.lib().ptir_grt_trace(self._t._h, C.byref(sc), C.byref(la), C.byref(out), _stream()),
                   self._t._h)
        del keep, keep2
        return rad, dns, hit, hit2, dist, nrm, snrm, hits, vis

    def trace_bwd(self, frame_id, ray_to_world, ray_ori, ray_dir, ray_rad, ray_dns, ray_hit, ray_hit2, ray_dist,
                  ray_nrm, ray_snrm, particle_density, particle_radiance, particle_shading_normal, ray_rad_grd,
                  ray_dns_grd, ray_hit_grd, ray_hit2_grd, ray_dist_grd, ray_nrm_grd, ray_snrm_grd, render_opts=0,
                  sph_degree=3, min_transmittance=1e-3, *, tape=None):
        """OptixTracer::traceBwd (threedgrt_tracer/src/optixTracer.cpp:927-1012): returns
        (particle_density_grd [N,12], particle_radiance_grd, particle_shading_normal_grd).
        ray_hit_grd is the gradient of the integrated depth ([..,1], channel 0 of ray_hit);
        the normal's gradient is not used, as in the reference."""
        sc, keep = self._scene(particle_density, particle_radiance, particle_shading_normal)
        la, keep2 = HipTracer._launch(frame_id, ray_to_world, ray_ori, ray_dir, render_opts, sph_degree,
                                      min_transmittance, 1)
        B, H, W = la.B, la.H, la.W

        def pix(t, name, c, allow_none=False):
            if t is None and allow_none:
                return None
            t = _need(t, name)
            if t.numel() != B * H * W * c:
                raise ValueError(f"{name} must have {B * H * W * c} elements, got {t.numel()}")
            keep.append(t)
            return t

        ins = [pix(ray_rad, "ray_rad", 3), pix(ray_dns, "ray_dns", 1), pix(ray_hit, "ray_hit", 2),
               pix(ray_hit2, "ray_hit2", 1), pix(ray_dist, "ray_dist", 1), pix(ray_snrm, "ray_snrm", 3),
               pix(ray_rad_grd, "ray_rad_grd", 3, True), pix(ray_dns_grd, "ray_dns_grd", 1, True),
               pix(ray_hit_grd[..., :1] if ray_hit_grd is not None and ray_hit_grd.shape[-1] == 2 else ray_hit_grd,
                   "ray_hit_grd", 1, True),
               pix(ray_hit2_grd, "ray_hit2_grd", 1, True), pix(ray_dist_grd, "ray_dist_grd", 1, True),
               pix(ray_snrm_grd, "ray_snrm_grd", 3, True)]
        bi = _lib.GrtBwdInputs(*[_ptr(t) for t in ins])
        if tape is not None:
            if tape.numel() * 4 < _lib.lib().ptir_grt_tape_bytes(B, H, W, tape.hits_per_pixel):
                raise ValueError("tape is too small for this launch")
            bi.hit_tape, bi.hit_tape_cap = _ptr(tape), tape.hits_per_pixel
        n = sc.num_gaussians
        pd = keep[0]
        f = dict(dtype=torch.float32, device=pd.device)
        d_dns = torch.empty(n, 12, **f)
        d_sh = torch.empty(n, max(sc.sh_stride, 1), **f)
        has_sn = particle_shading_normal is not None and particle_shading_normal.numel() > 0
        d_sn = torch.empty(n, 3, **f) if has_sn else torch.zeros(n, 3, **f)
        bo = _lib.GrtBwdOutputs(_ptr(d_dns), _ptr(d_sh), _ptr(d_sn) if has_sn else None)
        _lib.check(_lib.lib().ptir_grt_trace_bwd(self._t._h, C.byref(sc), C.byref(la), C.byref(bi), C.byref(bo),
                                                  _stream()), self._t._h)
        del keep, keep2
        return d_dns, d_sh, d_sn


class Tracer:
    """threedgrt_tracer/tracer.py:49-334 Tracer: primary-only render and its autograd function."""

    class _Autograd(torch.autograd.Function):
        @staticmethod
        def forward(ctx, tracer_wrapper, frame_id, ray_to_world, ray_ori, ray_dir, mog_pos, mog_rot, mog_scl, mog_dns,
                    mog_sph, mog_snrm, render_opts, sph_degree, min_transmittance):
            particle_density = torch.concat([mog_pos, mog_dns, mog_rot, mog_scl, torch.zeros_like(mog_dns)], dim=1)
            # the backward reads this forward's hits from a tape instead of re-tracing them
            tape = tracer_wrapper.new_tape(ray_dir) if any(ctx.needs_input_grad[5:11]) else None
            outs = tracer_wrapper.trace(frame_id, ray_to_world, ray_ori, ray_dir, particle_density, mog_sph, mog_snrm,
                                        render_opts, sph_degree, min_transmittance, tape=tape)
            rad, dns, hit, hit2, dist, nrm, snrm, hits, vis = outs
            ctx.save_for_backward(ray_to_world, ray_ori, ray_dir, rad, dns, hit, hit2, dist, nrm, snrm,
                                  particle_density, mog_sph, mog_snrm)
            ctx.frame_id, ctx.render_opts = frame_id, render_opts
            ctx.sph_degree, ctx.min_transmittance = sph_degree, min_transmittance
            ctx.tracer_wrapper = tracer_wrapper
            ctx.tape = tape
            return rad, dns, hit[..., 0:1], hit2, dist, nrm, snrm, hits, vis

        @staticmethod
        def backward(ctx, rad_grd, dns_grd, hit_grd, hit2_grd, dist_grd, nrm_grd, snrm_grd, _hits_grd, _vis_grd):
            (ray_to_world, ray_ori, ray_dir, rad, dns, hit, hit2, dist, nrm, snrm, particle_density, mog_sph,
             mog_snrm) = ctx.saved_tensors
            d_dns, d_sph, d_sn = ctx.tracer_wrapper.trace_bwd(
                ctx.frame_id, ray_to_world, ray_ori, ray_dir, rad, dns, hit, hit2, dist, nrm, snrm, particle_density,
                mog_sph, mog_snrm, rad_grd.contiguous(), dns_grd.contiguous(), hit_grd.contiguous(),
                hit2_grd.contiguous(), dist_grd.contiguous(), nrm_grd.contiguous(), snrm_grd.contiguous(),
                ctx.render_opts, ctx.sph_degree, ctx.min_transmittance, tape=ctx.tape)
            pos_g, dns_g, rot_g, scl_g, _ = torch.split(d_dns, [3, 1, 4, 3, 1], dim=1)
            return (None, None, None, None, None, pos_g, rot_g, scl_g, dns_g, d_sph, d_sn, None, None, None)

    class RenderOpts(IntEnum):
        NONE = 0
        DEFAULT = NONE

    def __init__(self, conf):
        """Tracer(conf) of threedgrt_tracer/tracer.py:207-238 (conf.render.*)."""
        self.device = "cuda"
        self.conf = conf
        r = _conf_get(conf, "render")
        self.render_conf = r
        self.num_update_bvh = 0
        self.tracer_wrapper = GRTTracer(
            "", "", _conf_get(r, "pipeline_type", "reference"), _conf_get(r, "backward_pipeline_type", "referenceBwd"),
            _conf_get(r, "primitive_type", "instances"), _conf_get(r, "particle_kernel_degree", 4),
            _conf_get(r, "particle_kernel_min_response", 0.0113),
            bool(_conf_get(r, "particle_kernel_density_clamping", True)),
            _conf_get(r, "particle_radiance_sph_degree", 3), _conf_get(r, "enable_normals", True),
            _conf_get(r, "enable_hitcounts", True), kbuffer=int(_conf_get(r, "kbuffer_depth", 16)),
            traversal=str(_conf_get(r, "primary_traversal", "wave")), builder=str(_conf_get(r, "bvh_builder", "lbvh")))
        self.timings = {}

    def build_acc(self, gaussians, rebuild=True):
        """threedgrt_tracer/tracer.py:240-262: a refit after at most max_consecutive_bvh_update
        updates, when density clamping is off; a full build otherwise."""
        r = self.render_conf
        clamping = bool(_conf_get(r, "particle_kernel_density_clamping", True))
        max_updates = int(_conf_get(r, "max_consecutive_bvh_update", 15))
        allow_update = max_updates > 1 and not clamping
        rebuild_bvh = rebuild or clamping or self.num_update_bvh >= max_updates
        self.tracer_wrapper.build_bvh(
            gaussians.positions.view(-1, 3).contiguous(), gaussians.get_rotation().view(-1, 4).contiguous(),
            gaussians.get_scale().view(-1, 3).contiguous(), gaussians.get_density().view(-1, 1).contiguous(),
            rebuild_bvh, allow_update)
        self.num_update_bvh = 0 if rebuild_bvh else self.num_update_bvh + 1

    def render(self, gaussians, gpu_batch, train=False, frame_id=0, sh_indirect: bool = False):
        """threedgrt_tracer/tracer.py:264-334."""
        rgb, opacity, dist, dist2, distortion, normals, snormal, hits, vis = Tracer._Autograd.apply(
            self.tracer_wrapper, frame_id, gpu_batch.T_to_world.contiguous(), gpu_batch.rays_ori.contiguous(),
            gpu_batch.rays_dir.contiguous(), gaussians.positions.contiguous(), gaussians.get_rotation().contiguous(),
            gaussians.get_scale().contiguous(), gaussians.get_density().contiguous(),
            gaussians.get_features().contiguous(), gaussians.get_shading_normal().contiguous(),
            Tracer.RenderOpts.DEFAULT, gaussians.n_active_features, _conf_get(self.render_conf, "min_transmittance",
                                                                               1e-3))
        rgb, opacity = gaussians.background(gpu_batch.T_to_world.contiguous(), gpu_batch.rays_dir.contiguous(), rgb,
                                            opacity, train)
        dist = torch.nan_to_num(dist / opacity, 0.0, 0.0)
        return {"pred_rgb": rgb, "pred_opacity": opacity, "pred_dist": dist, "pred_depth_second_moment": dist2,
                "pred_depth_distortion": distortion, "pred_normals": torch.nn.functional.normalize(normals, dim=3),
                "pred_shadingnormal": snormal, "hits_count": hits, "frame_time_ms": 0.0, "mog_visibility": vis}
