"""The PLOC hierarchy builder (ptir_options.builder = 1) on the GPU against the LBVH: the same
Morton-ordered leaves under another tree, so every forward output must be equal bit for bit
(the candidate rule is exact and hits are ordered by distance), the backward must give the
same gradients, and refitting a PLOC tree must trace like rebuilding it.  On the object-scale
workload the PLOC tree must visit fewer nodes per traversal.  The builder's phases are
checked on the host as well (tests/test_ploc_host.py)."""

import time

import pytest
import torch

from gpu_helpers import env_table
from ptir_amd import scenes
from ptir_amd.tracer import HipTracer

pytestmark = pytest.mark.gpu


def _setup(sc, device):
    tab = env_table(sc).to(device)
    g = sc.to(device)
    args = (g.particle_density(), g.particle_material(), g.features, g.shading_normal, g.env, tab)
    return g, args


def _trace(tr, g, args, sc, **kw):
    return tr.trace(sc.frame, g.ray_to_world, g.rays_ori, g.rays_dir, *args, sph_degree=sc.sph_degree,
                    min_transmittance=sc.min_transmittance, max_bounces=sc.max_bounces, **kw)


@pytest.mark.parametrize("which", ["small", "object"])
def test_ploc_traces_what_the_lbvh_traces(device, which):
    sc = scenes.small_scene(5, n=4096, res=64, spp=1, max_bounces=3) if which == "small" else scenes.object_scene(0)
    g, args = _setup(sc, device)
    res, cnt = {}, {}
    for builder in ("lbvh", "ploc"):
        tr = HipTracer(enable_metallic=True, builder=builder)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
        torch.cuda.synchronize()
        print(f"{which} {builder}: build {1e3 * (time.perf_counter() - t0):.2f} ms", flush=True)
        info = tr.hierarchy_info()
        print(f"{which} {builder}: {info}", flush=True)
        # the 4-wide nodes are the even-depth binary nodes: at least a third of the n - 1
        assert (sc.num_gaussians - 1) / 3 <= info["wide_nodes"] < sc.num_gaussians
        assert 0 <= info["stack_need"] <= 64
        res[builder] = _trace(tr, g, args, sc, counters=True)
        cnt[builder] = tr.last_counters
        tr.check()
        print(f"{which} {builder}: {cnt[builder]}", flush=True)
    for i, (a, b) in enumerate(zip(res["lbvh"], res["ploc"])):
        assert torch.equal(a, b), f"output {i}"
    for k in ("paths", "segments", "candidates", "accepted", "accepted_primary", "shadow_candidates", "traversals"):
        assert cnt["lbvh"][k] == cnt["ploc"][k], k
    if which == "object":
        assert cnt["ploc"]["nodes"] < 0.9 * cnt["lbvh"]["nodes"]


def test_ploc_backward_equals_lbvh_backward(device):
    sc = scenes.small_scene(6, n=4096, res=64, spp=1, max_bounces=3)
    g, args = _setup(sc, device)
    grads = []
    for builder in ("lbvh", "ploc"):
        tr = HipTracer(enable_metallic=True, builder=builder)
        tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
        for tape in (tr.new_tape(g.rays_dir, 3), None):
            outs = _trace(tr, g, args, sc, tape=tape)
            gp = torch.full_like(outs[10], 0.5)
            grads.append(tr.trace_bwd(sc.frame, g.ray_to_world, g.rays_ori, g.rays_dir, *outs[:8], outs[10], outs[11],
                                      *args, None, None, None, None, None, None, None, None, None, None, None, None,
                                      None, None, gp, None, sph_degree=sc.sph_degree,
                                      min_transmittance=sc.min_transmittance, max_bounces=3, tape=tape))
    for other in grads[1:]:
        for a, b in zip(grads[0], other):
            torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-6)  # atomics: order of the sums


def test_ploc_refit_equals_rebuild(device):
    sc = scenes.small_scene(2, n=2048, res=32, spp=1)
    g, _ = _setup(sc, device)
    gen = torch.Generator().manual_seed(9)
    moved = (g.positions + 0.05 * torch.randn(g.positions.shape, generator=gen).to(device)).contiguous()
    dens = torch.cat([moved, g.density, g.rotation, g.scale, torch.zeros_like(g.density)], 1).contiguous()
    args = (dens, g.particle_material(), g.features, g.shading_normal, g.env, env_table(sc).to(device))
    refit = HipTracer(enable_metallic=True, builder="ploc")
    refit.build_bvh(g.positions, g.rotation, g.scale, g.density)
    refit.build_bvh(moved, g.rotation, g.scale, g.density, rebuild=False)
    rebuilt = HipTracer(enable_metallic=True, builder="ploc")
    rebuilt.build_bvh(moved, g.rotation, g.scale, g.density)
    a = refit.trace(0, g.ray_to_world, g.rays_ori, g.rays_dir, *args, max_bounces=2)
    b = rebuilt.trace(0, g.ray_to_world, g.rays_ori, g.rays_dir, *args, max_bounces=2)
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.parametrize("n", [1, 2, 3, 1000])
def test_ploc_tiny_and_duplicate_scenes(device, n):
    """One, two and three Gaussians, and 1,000 copies of one Gaussian (every union area ties)."""
    sc = scenes.small_scene(3, n=max(n, 2), res=16, spp=1)
    g, _ = _setup(sc, device)
    pos, rot, scl, dns = (t[:1].repeat(n, 1) if n == 1000 else t[:n] for t in
                          (g.positions, g.rotation, g.scale, g.density))
    dens = torch.cat([pos, dns, rot, scl, torch.zeros_like(dns)], 1).contiguous()
    mat = g.particle_material()[:1].repeat(n, 1) if n == 1000 else g.particle_material()[:n]
    feat = g.features[:1].repeat(n, 1) if n == 1000 else g.features[:n]
    sn = g.shading_normal[:1].repeat(n, 1) if n == 1000 else g.shading_normal[:n]
    args = (dens.contiguous(), mat.contiguous(), feat.contiguous(), sn.contiguous(), g.env,
            env_table(sc).to(device))
    outs = []
    for builder in ("lbvh", "ploc"):
        tr = HipTracer(enable_metallic=True, builder=builder)
        tr.build_bvh(pos.contiguous(), rot.contiguous(), scl.contiguous(), dns.contiguous())
        outs.append(tr.trace(0, g.ray_to_world, g.rays_ori, g.rays_dir, *args, max_bounces=2))
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(*outs)):
        assert torch.equal(a, b), f"output {i}"


@pytest.mark.parametrize("which", ["small", "object"])
def test_wide_walks_trace_what_the_binary_walks_trace(device, which):
    """The 4-wide nodes (collapsed from the same binary tree) under the wave-coherent walk, the
    per-lane walks, or both: every output equal bit for bit, fewer node visits per traversal,
    both schedules."""
    sc = scenes.small_scene(5, n=4096, res=64, spp=1, max_bounces=3) if which == "small" else scenes.object_scene(0)
    g, args = _setup(sc, device)
    for schedule in ("megakernel", "wavefront"):
        res, cnt = {}, {}
        for ww in ("none", "wave", "lane", "all"):
            tr = HipTracer(enable_metallic=True, builder="ploc", schedule=schedule, wide_walks=ww)
            tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
            res[ww] = _trace(tr, g, args, sc, counters=True)
            cnt[ww] = tr.last_counters
            tr.check()
            print(f"{which} {schedule} wide={ww}: {cnt[ww]}", flush=True)
        for ww in ("wave", "lane", "all"):
            for i, (a, b) in enumerate(zip(res["none"], res[ww])):
                assert torch.equal(a, b), f"{schedule} wide={ww} output {i}"
            for k in ("paths", "segments", "candidates", "accepted", "accepted_primary", "shadow_candidates",
                      "traversals"):
                assert cnt["none"][k] == cnt[ww][k], (ww, k)
        assert cnt["all"]["nodes"] < 0.7 * cnt["none"]["nodes"]
