"""Shared helpers of the GPU tests: run the HIP path and the oracle on the same inputs."""

from __future__ import annotations

import numpy as np
import torch

from oracle import oracle as O
from ptir_amd import envmap
from ptir_amd.tracer import HipTracer

FWD_NAMES = ["rad", "dns", "hit", "hit2", "distortion", "nrm", "snrm", "mat", "hits", "visibility", "pbr", "light",
             "comp"]


def env_table(scene):
    return envmap.pack_alias_table(envmap.build_environment_alias_table(scene.env))


def gpu_forward(tracer: HipTracer, scene, table, device, *, frame=None, render_opts=0, counters=False, build=True):
    g = scene.to(device)
    if build:
        tracer.build_bvh(g.positions, g.rotation, g.scale, g.density)
    tab = table.to(device) if table is not None else torch.empty(0, device=device)
    args = (scene.frame if frame is None else frame, g.ray_to_world, g.rays_ori, g.rays_dir, g.particle_density(),
            g.particle_material(), g.features, g.shading_normal, g.env, tab)
    outs = tracer.trace(*args, render_opts=render_opts, sph_degree=scene.sph_degree,
                        min_transmittance=scene.min_transmittance, max_bounces=scene.max_bounces, counters=counters)
    return outs, args


def gpu_backward(tracer: HipTracer, scene, outs, args, pbr_grad, light_grad=None, mat_grad=None, render_opts=0,
                 counters=False):
    frame, T, ro, rd = args[:4]
    return tracer.trace_bwd(frame, T, ro, rd, *outs[:8], outs[10], outs[11], *args[4:], None, None, None, None, None,
                            None, None, None, None, None, None, None, None, mat_grad, pbr_grad, light_grad,
                            render_opts=render_opts, sph_degree=scene.sph_degree,
                            min_transmittance=scene.min_transmittance, max_bounces=scene.max_bounces,
                            counters=counters)


def oracle_forward(scene, table, precision=64, render_opts=0, frame=None, **kw):
    inp = O.Inputs.from_scene(scene, table, render_opts=render_opts,
                              **({} if frame is None else {"frame": frame}), **kw)
    rays = O.Rays(scene.rays_ori, scene.rays_dir, scene.ray_to_world)
    return inp, rays, O.trace(inp, rays, precision=precision, counters=True)


def to_np(outs):
    res = {}
    for k, v in zip(FWD_NAMES, outs):
        if k == "visibility":
            res[k] = (v.view(torch.int32).reshape(-1) != 0).cpu().numpy().astype(np.float64)
        else:
            res[k] = v.detach().double().cpu().numpy()
    res["comp"] = res["comp"].reshape(*res["comp"].shape[:3], 2, 3)
    return res


def pixel_agreement(a, b, rtol, atol):
    """Fraction of pixels whose every channel agrees within atol + rtol*|b|."""
    a = a.reshape(-1, int(np.prod(a.shape[3:])) if a.ndim > 3 else 1)
    b = b.reshape(a.shape)
    ok = np.abs(a - b) <= atol + rtol * np.abs(b)
    return float(ok.all(axis=1).mean())


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
