"""The wavefront schedule (ptir_options.schedule = 1) against the one-lane-per-path kernel on the
GPU: the same launch must write the same 13 outputs, bit for bit, in every mode (MIS on and off,
INDIRECT, Russian roulette, both radiance-cache modes with their records, a ragged batch), and
the path tape it records must drive the backward to the same gradients.  Both schedules run the
same device functions with the same random streams; contraction of float expressions is fixed
per source expression (-ffp-contract=on), so the two compilations round alike.  The stage code
is also checked path by path on the host (tests/test_wavefront_host.py)."""

import pytest
import torch

from gpu_helpers import env_table
from ptir_amd import scenes
from ptir_amd.tracer import HipTracer, RaggedBatch

pytestmark = pytest.mark.gpu

NAMES = ["rad", "dns", "hit", "hit2", "dist", "nrm", "snrm", "mat", "hits", "vis", "pbr", "light", "comp"]


def _pair(sc, device, mis=True, **kw):
    tab = env_table(sc).to(device)
    g = sc.to(device)
    args = (g.particle_density(), g.particle_material(), g.features, g.shading_normal, g.env, tab)
    res = []
    for sched in ("megakernel", "wavefront"):
        tr = HipTracer(enable_metallic=True, enable_mis=mis, schedule=sched)
        tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
        res.append((tr, g, args))
    return res


def _trace(tr, g, args, sc, rays=None, **kw):
    o, d = rays if rays is not None else (g.rays_ori, g.rays_dir)
    return tr.trace(sc.frame, g.ray_to_world, o, d, *args, sph_degree=sc.sph_degree,
                    min_transmittance=sc.min_transmittance, **kw)


def _same(a, b, label):
    for n, x, y in zip(NAMES, a, b):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32)), f"{label}: {n} differs"


@pytest.mark.parametrize("cfg", [
    dict(mis=True, max_bounces=3),
    dict(mis=False, max_bounces=4),
    dict(mis=True, max_bounces=2, render_opts=1),
    dict(mis=True, max_bounces=5, rr_start=2),
    dict(mis=True, max_bounces=1),
], ids=["mis_mb3", "nomis_mb4", "indirect", "roulette_mb5", "mb1"])
def test_wavefront_writes_what_the_one_lane_kernel_writes(device, cfg):
    cfg = dict(cfg)
    mis = cfg.pop("mis")
    sc = scenes.small_scene(6, n=4096, res=64, spp=1, max_bounces=cfg["max_bounces"])
    (ta, g, args), (tb, _, _) = _pair(sc, device, mis=mis)
    print("schedules", cfg, flush=True)
    _same(_trace(ta, g, args, sc, **cfg), _trace(tb, g, args, sc, **cfg), str(cfg))


@pytest.mark.parametrize("mode", [1, 2])
def test_wavefront_cache_modes_write_the_same_records(device, mode):
    sc = scenes.small_scene(7, n=4096, res=64, spp=1, max_bounces=4)
    (ta, g, args), (tb, _, _) = _pair(sc, device)
    recs = []
    outs = []
    for tr in (ta, tb):
        rec = torch.zeros(*sc.rays_dir.shape[:3], 20, device=device)
        outs.append(_trace(tr, g, args, sc, max_bounces=4, cache_mode=mode, cache_records=rec))
        recs.append(rec)
    _same(outs[0], outs[1], f"cache mode {mode}")
    assert torch.equal(recs[0].view(torch.int32), recs[1].view(torch.int32))


def test_wavefront_ragged_batch(device):
    sc = scenes.batched_scene(1, n=20_000, n_views=3, res_scale=0.125, max_bounces=4, rr_start=2)
    (ta, g, args), (tb, _, _) = _pair(sc, device)
    rb = RaggedBatch.build(sc.extra["images"], device)
    o = g.rays_ori.reshape(-1, 3).contiguous()
    d = g.rays_dir.reshape(-1, 3).contiguous()
    kw = dict(max_bounces=sc.max_bounces, rr_start=2, ragged=rb)
    _same(_trace(ta, g, args, sc, rays=(o, d), **kw), _trace(tb, g, args, sc, rays=(o, d), **kw), "ragged")


def test_wavefront_tape_drives_the_same_backward(device):
    sc = scenes.small_scene(8, n=4096, res=64, spp=1, max_bounces=3)
    (ta, g, args), (tb, _, _) = _pair(sc, device)
    grads = []
    for tr in (ta, tb):
        tape = tr.new_tape(g.rays_dir, 3)
        outs = _trace(tr, g, args, sc, max_bounces=3, tape=tape)
        gp = torch.ones_like(outs[10]) * 0.5
        gm = torch.full_like(outs[7], 0.25)
        grads.append(tr.trace_bwd(sc.frame, g.ray_to_world, g.rays_ori, g.rays_dir, *outs[:8], outs[10], outs[11],
                                  *args, None, None, None, None, None, None, None, None, None, None, None, None, None,
                                  gm, gp, None, sph_degree=sc.sph_degree, min_transmittance=sc.min_transmittance,
                                  max_bounces=3, tape=tape))
    for a, b in zip(grads[0], grads[1]):
        torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-6)  # atomics: order of the sums
