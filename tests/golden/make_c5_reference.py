"""Writes tests/golden/c5_reference.npz: the FP64 reference image of the radiance-cache
comparison on the object-scale scene (300k Gaussians on the sphere-torus shell, env
512x256 with a 500x sun, view 0 at 800x800).

The FP64 CPU oracle renders the plain path tracer with maxBounces 4 (three bounces) at
1,024 samples per pixel (frames 0..1023, pixel-centre rays) on a 96x96 crop.  The crop is
the one, among a grid of candidates, whose 4-spp oracle image carries the largest share of
indirect light (the part a radiance cache stands in for).

Run:  python tests/golden/make_c5_reference.py   (about 15 minutes on 8 cores)
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from ptir_amd import envmap, scenes  # noqa: E402

CROP = 96
SPP = 1024
MAX_BOUNCES = 4


def main():
    sc = scenes.object_scene(0)
    tab = envmap.pack_alias_table(envmap.build_environment_alias_table(sc.env))
    inp4 = O.Inputs.from_scene(sc, tab, max_bounces=MAX_BOUNCES, frame=0)
    best = None
    for y0 in range(160, 640 - CROP + 1, 96):
        for x0 in range(160, 640 - CROP + 1, 96):
            o = sc.rays_ori[:1, y0:y0 + CROP, x0:x0 + CROP].contiguous()
            d = sc.rays_dir[:1, y0:y0 + CROP, x0:x0 + CROP].contiguous()
            acc_i = acc_t = 0.0
            for f in range(4):
                inp = O.Inputs.from_scene(sc, tab, max_bounces=MAX_BOUNCES, frame=f)
                out = O.trace(inp, O.Rays(o, d, sc.ray_to_world), nthreads=os.cpu_count())
                acc_i += out["comp"][..., 1, :].sum()
                acc_t += out["pbr"].sum()
            share = acc_i / max(acc_t, 1e-12)
            if best is None or share > best[0]:
                best = (share, y0, x0)
    share, y0, x0 = best
    print(f"crop at ({y0}, {x0}): indirect share {share:.3f}", flush=True)
    o = sc.rays_ori[:1, y0:y0 + CROP, x0:x0 + CROP].contiguous()
    d = sc.rays_dir[:1, y0:y0 + CROP, x0:x0 + CROP].contiguous()
    rays = O.Rays(o, d, sc.ray_to_world)
    acc = np.zeros((CROP, CROP, 3))
    t0 = time.time()
    for f in range(SPP):
        inp = O.Inputs.from_scene(sc, tab, max_bounces=MAX_BOUNCES, frame=f)
        acc += O.trace(inp, rays, nthreads=os.cpu_count())["pbr"][0]
        if f % 128 == 127:
            print(f"  {f + 1} spp, {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c5_reference.npz"), image=(acc / SPP).astype(np.float32),
                        crop=np.array([0, y0, x0, CROP, CROP]), spp=SPP, max_bounces=MAX_BOUNCES,
                        indirect_share=share)


if __name__ == "__main__":
    main()
