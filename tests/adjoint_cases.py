"""Shared cases of the exact-adjoint tests (GPU kernel and its host build) against the FP64
dual-number oracle.

For output gradients g (random, on every forward output that takes one) the exact backward
returns grad_theta of sum <g, output>.  The oracle's dual-number forward gives, along a
parameter direction t, the derivative of every output; so for every t

    sum_theta grad_theta . t  ==  sum_outputs <g, d output / d t>.

The oracle's derivatives are pinned by central differences (tests/test_oracle_dual.py).
The check runs per parameter block (positions, densities, rotations, scales, materials, SH,
shading normals, environment texels) with random dense directions.  Output gradients are zero
on pixels whose oracle decision margin is below 3e-3: their paths lie close to a discrete
decision (a hit accepted or not, a draw against a probability, a BRDF pdf against its 1e-6
cut-off), where float32 and FP64 can take different branches or, next to the pdf cut-off,
differ by far more than float32 rounding (gradients of 1e6 and more).

Tolerance: |lhs - rhs| <= 1e-3 * sum |g * d output / d t| (the sum of the magnitudes of the
terms).
"""

import zlib

import numpy as np

from gpu_helpers import env_table
from oracle import oracle as O
from ptir_amd import scenes

MARGIN = 3e-3
TOL = 1e-3

BLOCKS = {"position": ("density", slice(0, 3)), "sigma": ("density", slice(3, 4)),
          "rotation": ("density", slice(4, 8)), "scale": ("density", slice(8, 11)),
          "material": ("material", slice(0, 5)), "sh": ("sh", None), "snormal": ("snormal", slice(0, 3)),
          "env": ("env", slice(0, 3))}
# output name -> (gradient keyword of trace_bwd_exact, channels)
OUTS = {"pbr": ("ray_pbr_grd", 3), "light": ("ray_light_grd", 3), "mat": ("ray_mat_grd", 5),
        "rad": ("ray_rad_grd", 3), "dns": ("ray_dns_grd", 1), "hit": ("ray_hit_grd", 2),
        "snrm": ("ray_snrm_grd", 3), "comp": ("ray_comp_grd", 6)}

CASES = {
    "mis_metallic_mb3": dict(max_bounces=3, enable_mis=True, render_opts=0, rr_start=0),
    "no_mis_mb3": dict(max_bounces=3, enable_mis=False, render_opts=0, rr_start=0),
    "indirect_mb2": dict(max_bounces=2, enable_mis=True, render_opts=1, rr_start=0),
    "roulette_mb4": dict(max_bounces=4, enable_mis=True, render_opts=0, rr_start=1),
}


def setup(case):
    """(scene, table, oracle inputs, oracle rays, FP64 forward outputs, rng) of a case."""
    cfg = CASES[case]
    sc = scenes.small_scene(3, n=1024, res=24, spp=1, max_bounces=cfg["max_bounces"])
    tab = env_table(sc)
    inp = O.Inputs.from_scene(sc, tab, enable_mis=cfg["enable_mis"], render_opts=cfg["render_opts"],
                              rr_start=cfg["rr_start"])
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    return sc, tab, inp, rays, O.trace(inp, rays), np.random.default_rng(zlib.crc32(case.encode()))


def output_grads(keep_mask, rng):
    """Random output gradients, zero outside keep_mask [B,H,W]."""
    B, H, W = keep_mask.shape
    keep = keep_mask.astype(np.float32)[..., None]
    g = {}
    for k, (_, c) in OUTS.items():
        a = rng.standard_normal((B, H, W, c)).astype(np.float32) * keep
        if k == "hit":
            a[..., 1] = 0.0  # the last processed hit distance takes no gradient
        g[k] = a
    return g


def direction(shape, sl, rng):
    t = np.zeros(shape, np.float32)
    if len(shape) == 3:  # environment texels: rgb
        t[..., sl] = rng.standard_normal(t[..., sl].shape)
    elif sl is None:
        t[:] = rng.standard_normal(t.shape)
    else:
        t[:, sl] = rng.standard_normal(t[:, sl].shape)
    return t


def check_blocks(sc, inp, rays, g, grads, rng, label, reps=2):
    """Compare grad . t with the oracle's sum <g, d output / d t> for every block."""
    B, H, W = sc.rays_dir.shape[:3]
    shapes = {k: v.shape for k, v in grads.items()}
    for block, (name, sl) in BLOCKS.items():
        for rep in range(reps):
            t = direction(shapes[name], sl, rng)
            _, der = O.trace_dual(inp, rays, {name: t})
            terms = np.concatenate([(g[k].astype(np.float64) * der[k].reshape(B, H, W, c)).reshape(-1)
                                    for k, (_, c) in OUTS.items()])
            lhs = float((grads[name] * t).sum())
            rhs = float(terms.sum())
            scale = float(np.abs(terms).sum())
            print(f"[{label}] {block}/{rep}: exact {lhs:+.6e} dual {rhs:+.6e} scale {scale:.3e}", flush=True)
            assert abs(lhs - rhs) <= TOL * scale + 1e-9, (block, rep, lhs, rhs, scale)
            assert scale > 0.0, block  # the direction moves some output
