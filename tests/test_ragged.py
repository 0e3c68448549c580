"""Ragged launches (several images of their own sizes in one launch) and Russian roulette,
on the GPU against the regular launch and the FP64 oracle.

A ragged image i traced alone by the oracle with launch-index offset z = i sees the same
random streams as in the ragged launch, so every image is checked path by path."""

import numpy as np
import pytest
import torch

from gpu_helpers import env_table, pixel_agreement, rel_l2
from oracle import oracle as O
from ptir_amd import scenes
from ptir_amd.tracer import HipTracer, RaggedBatch

pytestmark = pytest.mark.gpu

KEYS = ("dns", "hit", "mat", "pbr", "light")
IDX = {"rad": 0, "dns": 1, "hit": 2, "hit2": 3, "distortion": 4, "nrm": 5, "snrm": 6, "mat": 7, "hits": 8,
       "pbr": 10, "light": 11, "comp": 12}


def _trace_args(g, tab, device, ori, dirs, frame=0):
    return (frame, torch.eye(4, device=device)[None], ori, dirs, g.particle_density(), g.particle_material(),
            g.features, g.shading_normal, g.env, tab.to(device))


def test_ragged_launch_of_equal_images_equals_the_batched_launch(device):
    sc = scenes.small_scene(0, n=2048, res=32, spp=4)
    tab = env_table(sc)
    g = sc.to(device)
    T = g.ray_to_world[0]
    ori = (T[:3, :3] @ g.rays_ori.reshape(-1, 3).T).T + T[:3, 3]
    dirs = (T[:3, :3] @ g.rays_dir.reshape(-1, 3).T).T
    ori, dirs = ori.reshape(g.rays_ori.shape).contiguous(), dirs.reshape(g.rays_dir.shape).contiguous()
    tr = HipTracer(enable_metallic=True)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    regular = tr.trace(*_trace_args(g, tab, device, ori, dirs), max_bounces=2)
    rb = RaggedBatch.build([(32, 32)] * 4, device)
    ragged = tr.trace(*_trace_args(g, tab, device, ori.reshape(-1, 3), dirs.reshape(-1, 3)), max_bounces=2,
                      ragged=rb)
    torch.cuda.synchronize()
    for k, i in IDX.items():
        assert torch.equal(regular[i].reshape(-1), ragged[i].reshape(-1)), k
    gp = torch.rand(regular[10].shape, generator=torch.Generator().manual_seed(1)).to(device) * 1e-3
    common = dict(max_bounces=2)

    def bwd(outs, o, d, **kw):
        return tr.trace_bwd(0, torch.eye(4, device=device)[None], o, d, *outs[:8], outs[10], outs[11],
                            *_trace_args(g, tab, device, o, d)[4:], None, None, None, None, None, None, None, None,
                            None, None, None, None, None, None, kw.pop("gp"), None, **common, **kw)

    a = bwd(regular, ori, dirs, gp=gp)
    b = bwd(ragged, ori.reshape(-1, 3), dirs.reshape(-1, 3), gp=gp.reshape(ragged[10].shape), ragged=rb)
    torch.cuda.synchronize()
    for i in (1, 4):
        assert rel_l2(b[i].cpu().numpy(), a[i].cpu().numpy()) < 1e-5


@pytest.mark.parametrize("rr_start", [0, 2])
def test_ragged_images_of_different_sizes_match_the_oracle(device, rr_start):
    """A shrunken copy of the batched workload: 16 views of their own resolutions and
    samples per pixel (60 images), maxBounces 4, with and without Russian roulette."""
    sc = scenes.batched_scene(0, n=20_000, res_scale=1 / 32, env_hw=(64, 128))
    tab = env_table(sc)
    g = sc.to(device)
    rb = RaggedBatch.build(sc.extra["images"], device)
    tr = HipTracer(enable_metallic=True)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    print("ragged images", len(rb.hw), "paths", rb.num_paths, "rr", rr_start, flush=True)
    outs = tr.trace(*_trace_args(g, tab, device, g.rays_ori, g.rays_dir), max_bounces=sc.max_bounces, ragged=rb,
                    rr_start=rr_start)
    torch.cuda.synchronize()
    inp = O.Inputs.from_scene(sc, tab, rr_start=rr_start)
    ori, dirs = sc.rays_ori.reshape(-1, 3), sc.rays_dir.reshape(-1, 3)
    worst = {k: 1.0 for k in KEYS}
    for i, ((h, w), start) in enumerate(zip(rb.hw, rb.offsets())):
        sl = slice(start, start + h * w)
        rays = O.Rays(ori[sl].reshape(1, h, w, 3), dirs[sl].reshape(1, h, w, 3), torch.eye(4), z_offset=i)
        ref = O.trace(inp, rays, nthreads=8)
        for k in KEYS:
            got = outs[IDX[k]].reshape(rb.num_paths, -1)[sl].double().cpu().numpy().reshape(1, h, w, -1)
            worst[k] = min(worst[k], pixel_agreement(got, ref[k], 1e-3, 1e-5))
    print("ragged worst-image agreement", worst, flush=True)
    for k in ("dns", "hit"):
        assert worst[k] >= 0.999, k
    for k in ("mat", "pbr", "light"):
        assert worst[k] >= 0.98, k


def test_russian_roulette_backward_matches_oracle_and_tape(device):
    sc = scenes.small_scene(3, n=2048, res=32, spp=2, max_bounces=4)
    tab = env_table(sc)
    g = sc.to(device)
    tr = HipTracer(enable_metallic=True)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    args = (0, g.ray_to_world, g.rays_ori, g.rays_dir, g.particle_density(), g.particle_material(), g.features,
            g.shading_normal, g.env, tab.to(device))
    kw = dict(max_bounces=4, rr_start=2)
    tape = tr.new_tape(g.rays_dir, 4)
    outs = tr.trace(*args, tape=tape, **kw)
    gp = (torch.rand(outs[10].shape, generator=torch.Generator().manual_seed(2)) * 1e-3).to(device)
    common = (0, g.ray_to_world, g.rays_ori, g.rays_dir, *outs[:8], outs[10], outs[11], *args[4:], None, None, None,
              None, None, None, None, None, None, None, None, None, None, None, gp, None)
    with_tape = tr.trace_bwd(*common, tape=tape, **kw)
    replay = tr.trace_bwd(*common, **kw)
    torch.cuda.synchronize()
    for i in (1, 4):
        assert rel_l2(with_tape[i].cpu().numpy(), replay[i].cpu().numpy()) < 1e-5
    inp = O.Inputs.from_scene(sc, tab, rr_start=2)
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    ref = O.trace(inp, rays)
    pbr = outs[10].double().cpu().numpy()
    print("rr pbr agreement", pixel_agreement(pbr, ref["pbr"], 1e-3, 1e-5), flush=True)
    assert pixel_agreement(pbr, ref["pbr"], 1e-3, 1e-5) >= 0.995
    # gradients on the paths whose forward agrees (the backward is linear in the output gradient)
    ok = (np.abs(pbr - ref["pbr"]) <= 1e-5 + 1e-3 * np.abs(ref["pbr"])).all(-1, keepdims=True)
    gpm = (gp * torch.from_numpy(ok.astype(np.float32)).to(device)).contiguous()
    common = common[:-2] + (gpm, None)
    got = tr.trace_bwd(*common, tape=tape, **kw)
    dmat, denv = O.trace_bwd(inp, rays, pbr, outs[11].double().cpu().numpy(), gpm.cpu().numpy(),
                             np.zeros(outs[11].shape, np.float32), np.zeros(outs[7].shape, np.float32))
    torch.cuda.synchronize()
    assert rel_l2(got[1].cpu().numpy(), dmat) < 1e-3
    assert rel_l2(got[4].cpu().numpy(), denv) < 1e-3
