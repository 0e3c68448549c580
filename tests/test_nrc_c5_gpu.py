"""The radiance cache on the object-scale workload, against the 1,024-spp FP64 oracle image
of the plain tracer (maxBounces 4) on a 96x96 crop (tests/golden/c5_reference.npz, written
by tests/golden/make_c5_reference.py).

The cache form ends every path at its first secondary vertex and adds the cache's answer
there; training uses full paths of every 4th pixel row of the 4 x 800 x 800 batch.  The crop
carries about 1.2% of its light beyond that vertex, so the test checks what the cache must
do on it: remove the bias of the cut (at 64 spp the cut tracer's MSE against the FP64 image
exceeds the cache form's) without adding variance (the cache form's MSE stays within 5% of
the plain tracer's at 1 and 64 spp).  Frames 10,000 on are disjoint from the reference's
frames 0..1023.
"""

import os

import numpy as np
import pytest
import torch

from ptir_amd import envmap, scenes
from ptir_amd.cache import CachedPathTracer, RadianceCache
from ptir_amd.tracer import HipTracer

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_cache_removes_the_cut_bias_on_the_object_scale_crop(device):
    ref = np.load(os.path.join(HERE, "golden", "c5_reference.npz"))
    _, y0, x0, ch, cw = (int(v) for v in ref["crop"])
    assert int(ref["spp"]) == 1024 and int(ref["max_bounces"]) == 4
    sc = scenes.object_scene(0)
    tab = envmap.pack_alias_table(envmap.build_environment_alias_table(sc.env)).to(device)
    g = sc.to(device)
    tr = HipTracer(enable_metallic=True)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    args = (g.particle_density(), g.particle_material(), g.features, g.shading_normal, g.env, tab)
    box = torch.cat([g.positions.min(0).values - 0.05, g.positions.max(0).values + 0.05]).contiguous()
    cached = CachedPathTracer(tr, RadianceCache(max_batch=65536, seed=0), box, max_bounces=4)
    T = g.ray_to_world
    print("training 150 steps", flush=True)
    for s in range(150):
        cached.train(100_000 + s, T, g.rays_ori, g.rays_dir, args, offset=s % 4)
    torch.cuda.synchronize()
    assert cached.train_samples > 0
    o = g.rays_ori[:1, y0:y0 + ch, x0:x0 + cw].contiguous()
    d = g.rays_dir[:1, y0:y0 + ch, x0:x0 + cw].contiguous()
    refimg = torch.from_numpy(ref["image"]).double().to(device)

    def mse_at(render, spps):
        acc = torch.zeros(ch, cw, 3, dtype=torch.float64, device=device)
        out, n = {}, 0
        for S in spps:
            while n < S:
                acc += render(10_000 + n)[0].double()
                n += 1
            out[S] = float(((acc / n - refimg) ** 2).mean())
        return out

    plain = mse_at(lambda f: tr.trace(f, T, o, d, *args, max_bounces=4)[10], (1, 64))
    cut = mse_at(lambda f: tr.trace(f, T, o, d, *args, max_bounces=2)[10], (1, 64))
    nrc = mse_at(lambda f: cached.render(f, T, o, d, args)[0][10], (1, 64))
    print(f"MSE vs FP64: plain {plain}, cut {cut}, cache {nrc}", flush=True)
    assert nrc[64] < cut[64]
    for S in (1, 64):
        assert nrc[S] < 1.05 * plain[S]
