"""The launch-configuration autotuner (ptir_amd/autotune.py) on the GPU: it times every
(k-buffer, traversal, schedule, builder) candidate, returns the fastest, and the tracer built from its
choice writes the same outputs as the default configuration, bit for bit (the candidates
differ in speed only: k-buffer depth and traversal change how the hits are collected, not
which hits or their order, and the two schedules run the same device functions)."""

import pytest
import torch

from gpu_helpers import env_table
from ptir_amd import scenes
from ptir_amd.autotune import autotune
from ptir_amd.tracer import HipTracer

pytestmark = pytest.mark.gpu


def test_autotune_returns_the_fastest_and_the_same_image(device):
    sc = scenes.small_scene(11, n=4096, res=64, spp=1, max_bounces=3)
    tab = env_table(sc).to(device)
    g = sc.to(device)
    args = (g.particle_density(), g.particle_material(), g.features, g.shading_normal, g.env, tab)
    geom = (g.positions, g.rotation, g.scale, g.density)
    kw = dict(max_bounces=3, sph_degree=sc.sph_degree, min_transmittance=sc.min_transmittance)
    print("autotune small scene", flush=True)
    res = autotune(geom, args, sc.frame, g.ray_to_world, g.rays_ori, g.rays_dir,
                   tracer_kwargs=dict(enable_metallic=True), trace_kwargs=kw, backward=True, reps=1)
    assert len(res.timings) == 48
    assert all(t > 0 for t in res.timings.values())
    assert res.ms == min(res.timings.values())
    assert res.timings[(res.kbuffer, res.traversal, res.schedule, res.builder, res.wide_walks)] == res.ms
    d = res.as_dict()
    assert set(d) == {"kbuffer", "traversal", "schedule", "builder", "wide_walks", "ms", "timings_ms"}
    assert len(d["timings_ms"]) == 48

    outs = []
    for tkw in (dict(), res.tracer_kwargs()):
        tr = HipTracer(enable_metallic=True, **tkw)
        tr.build_bvh(*geom)
        outs.append(tr.trace(sc.frame, g.ray_to_world, g.rays_ori, g.rays_dir, *args, **kw))
    # rad, dns, hit, hit2, dist, nrm, snrm, mat, hits, vis, pbr, light, comp: bit for bit
    for i, (a, b) in enumerate(zip(*outs)):
        assert torch.equal(a, b), f"output {i}"
