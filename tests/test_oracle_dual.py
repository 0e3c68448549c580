"""Pins of the dual-number oracle: the exact directional derivative of every output along a
direction in parameter space (oracle.trace_dual) equals central differences of the FP64
forward, for every parameter block (positions, rotations, scales, densities, SH, shading
normals, materials, environment texels).

Perturbations are +-2^-14 along directions with entries in {-1, 0, 1}: exact in the
float32 inputs.  A pixel whose path takes a different discrete decision at +h and -h (a
hit accepted or not, an escape or shadow draw) has no derivative; those pixels are
recognised by a finite difference far from both the dual derivative and from the smooth
scale (as are paths that cross a clamp, max(0, .) of an SH colour or a BRDF value), and
must be rare (at most 2% of the pixels).  Everything else agrees to the O(h^2) error of the difference.
"""

import zlib

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ptir_amd import envmap, scenes

H = 2.0 ** -14
OUTS = ("rad", "dns", "hit", "hit2", "distortion", "snrm", "mat", "pbr", "light")


def _scene(max_bounces, mis=True):
    sc = scenes.small_scene(4, n=512, res=12, spp=1, max_bounces=max_bounces)
    tab = envmap.pack_alias_table(envmap.build_environment_alias_table(sc.env))
    return sc, tab, dict(enable_mis=mis, enable_metallic=True)


def _inputs(sc, tab, opts, **over):
    t = {"density": sc.particle_density(), "material": sc.particle_material(), "sh": sc.features,
         "snormal": sc.shading_normal, "env": sc.env}
    t.update(over)
    return O.Inputs(t["density"], t["material"], t["sh"], t["snormal"], t["env"], tab, sph_degree=sc.sph_degree,
                    min_transmittance=sc.min_transmittance, max_bounces=sc.max_bounces, frame=sc.frame, **opts)


def _direction(sc, block, rng):
    """Random {-1,0,1} direction touching one parameter block of ~40 Gaussians (or texels)."""
    n = sc.num_gaussians
    pick = rng.choice(n, size=40, replace=False)
    shapes = {"density": (n, 12), "material": (n, 5), "sh": tuple(sc.features.shape), "snormal": (n, 3),
              "env": tuple(sc.env.shape)}
    cols = {"position": ("density", slice(0, 3)), "sigma": ("density", slice(3, 4)),
            "rotation": ("density", slice(4, 8)), "scale": ("density", slice(8, 11)),
            "material": ("material", slice(0, 5)), "sh": ("sh", slice(0, 48)), "snormal": ("snormal", slice(0, 3)),
            "env": ("env", None)}
    name, sl = cols[block]
    t = np.zeros(shapes[name], np.float32)
    if name == "env":
        ys = rng.integers(0, shapes[name][0], 60)
        xs = rng.integers(0, shapes[name][1], 60)
        t[ys, xs, :3] = rng.choice([-1.0, 1.0], size=(60, 3))
    else:
        t[pick, sl] = rng.choice([-1.0, 1.0], size=(40, sl.stop - sl.start))
    return name, t


@pytest.mark.parametrize("max_bounces", [2, 3])
@pytest.mark.parametrize("block", ["position", "sigma", "rotation", "scale", "material", "sh", "snormal", "env"])
def test_dual_derivative_matches_central_differences(max_bounces, block):
    sc, tab, opts = _scene(max_bounces)
    rng = np.random.default_rng(zlib.crc32(f"{max_bounces}/{block}".encode()))
    name, t = _direction(sc, block, rng)
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    base = {"density": sc.particle_density(), "material": sc.particle_material(), "sh": sc.features,
            "snormal": sc.shading_normal, "env": sc.env}
    val, der = O.trace_dual(_inputs(sc, tab, opts), rays, {name: t}, nthreads=8)
    plus = O.trace(_inputs(sc, tab, opts, **{name: base[name] + H * torch.from_numpy(t)}), rays, nthreads=8)
    minus = O.trace(_inputs(sc, tab, opts, **{name: base[name] - H * torch.from_numpy(t)}), rays, nthreads=8)
    ref = O.trace(_inputs(sc, tab, opts), rays, nthreads=8)
    touched = 0
    for k in OUTS:
        np.testing.assert_allclose(val[k], ref[k], rtol=1e-12, atol=1e-14)  # the dual run's values
        fd = (plus[k] - minus[k]) / (2 * H)
        d = der[k]
        if k == "hit":  # channel 1, the last processed hit distance, is not differentiated
            fd, d = fd[..., :1], d[..., :1]
        scale = np.abs(d).max() + 1e-12
        ok = np.abs(fd - d) <= 2e-4 * scale + 1e-3 * np.abs(d)
        bad = ~ok.reshape(ok.shape[0], ok.shape[1], ok.shape[2], -1).all(-1)
        # a path whose discrete decisions differ between +h and -h, or that crosses a clamp
        # (max(0, .) of SH colours and BRDF values) between them, jumps: O(1/h) differences
        assert bad.mean() <= 0.02, (k, bad.sum())
        touched += int((np.abs(d) > 1e-9 * scale).any(-1).sum()) if d.ndim > 3 else 0
    assert np.abs(der["pbr"]).max() > 0 or block in ("sh",)  # the direction moves the image
