"""Pins of the oracle's 3DGRT primary-only tracer (oracle_grt_trace / oracle_grt_trace_bwd,
restating threedgrt_tracer/src/kernels/cuda/referenceOptix.cu:105-194, referenceBwdOptix.cu:
103-184 and gaussianParticles.cuh:473-772).

The reference backward is the adjoint of the G-buffer through the forward's totals.  Where it
is exact -- the gradients of radiance, opacity, integrated depth and the shading-normal sum --
it must equal the FP64 dual-number derivative of the forward (itself pinned by central
differences in tests/test_oracle_dual.py): for output gradients g and any parameter direction
t, sum grad . t == sum <g, d output / d t>.  The reference's float32 end bound drops the hits at
the last forward distance, so this comparison runs with that bound widened (inclusive_end);
the exclusive bound is pinned on its own.  The second-moment and distortion gradients follow
the hit distances only (the reference omits their path through alpha), so they are exact
along directions that leave every alpha unchanged: moving Gaussians along parallel rays.
Gaussian densities are kept below 0.99 (the alpha clamp, whose gradient the reference does not
zero) and pixels near a discrete decision are masked, as in the exact-adjoint tests.
"""

import zlib

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ptir_amd import scenes

MARGIN = 3e-3
TOL = 1e-6  # FP64 against FP64: rounding only

BLOCKS = {"position": ("density", slice(0, 3)), "sigma": ("density", slice(3, 4)),
          "rotation": ("density", slice(4, 8)), "scale": ("density", slice(8, 11)),
          "sh": ("sh", None), "snormal": ("snormal", slice(0, 3))}


def _scene(seed=3, n=1024, res=24):
    sc = scenes.small_scene(seed, n=n, res=res, spp=1, max_bounces=1)
    sc.density = torch.clamp(sc.density, max=0.9)
    return sc


def _inputs(sc, **kw):
    return O.grt_inputs(sc.particle_density(), sc.features, sc.shading_normal, sph_degree=sc.sph_degree,
                        min_transmittance=sc.min_transmittance, **kw)


def _check(inp, rays, fwd, g, grads, direction_of, outs, rng, label, reps=2):
    for block, (name, sl) in BLOCKS.items():
        for rep in range(reps):
            t = direction_of(name, sl)
            if t is None:
                continue
            _, der = O.grt_trace_dual(inp, rays, {name: t})
            terms = np.concatenate([(g[k].astype(np.float64) * der[k][..., :g[k].shape[-1]]).reshape(-1) for k in outs])
            lhs = float((grads[name] * t).sum())
            rhs = float(terms.sum())
            scale = float(np.abs(terms).sum())
            print(f"[{label}] {block}/{rep}: restated {lhs:+.9e} dual {rhs:+.9e} scale {scale:.3e}", flush=True)
            assert scale > 0.0, block
            assert abs(lhs - rhs) <= TOL * scale + 1e-12, (block, rep, lhs, rhs, scale)


@pytest.mark.parametrize("degree", [4, 2, 8])
def test_grt_backward_is_the_derivative_of_the_gbuffer(degree):
    sc = _scene()
    inp = _inputs(sc, kernel_degree=float(degree))
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    fwd = O.grt_trace(inp, rays)
    keep = (fwd["margin"] >= MARGIN).astype(np.float32)[..., None]
    assert keep.mean() > 0.7 and (fwd["dns"] > 0.1).mean() > 0.2
    rng = np.random.default_rng(zlib.crc32(f"grt{degree}".encode()))
    B, H, W = sc.rays_dir.shape[:3]
    g = {"rad": rng.standard_normal((B, H, W, 3)) * keep, "dns": rng.standard_normal((B, H, W, 1)) * keep,
         "hit": rng.standard_normal((B, H, W, 1)) * keep, "snrm": rng.standard_normal((B, H, W, 3)) * keep}
    g = {k: v.astype(np.float32) for k, v in g.items()}
    dd, dsh, dsn = O.grt_trace_bwd(inp, rays, fwd, g, inclusive_end=True)
    grads = {"density": dd, "sh": dsh, "snormal": dsn}
    shapes = {"density": dd.shape, "sh": dsh.shape, "snormal": dsn.shape}

    def direction(name, sl):
        t = np.zeros(shapes[name], np.float32)
        if sl is None:
            t[:, :3 * (sc.sph_degree + 1) ** 2] = rng.standard_normal((t.shape[0], 3 * (sc.sph_degree + 1) ** 2))
        else:
            t[:, sl] = rng.standard_normal(t[:, sl].shape)
        return t

    _check(inp, rays, fwd, g, grads, direction, ["rad", "dns", "hit", "snrm"], rng, f"deg{degree}")


def test_grt_second_moment_and_distortion_gradients_follow_the_hit_distances():
    """Parallel rays along +z through the Gaussians; each Gaussian moves along +z (alphas fixed)."""
    sc = _scene(seed=5, n=768)
    res = 32
    xs = np.linspace(-0.9, 0.9, res, dtype=np.float32)
    ori = np.zeros((1, res, res, 3), np.float32)
    ori[0, :, :, 0] = xs[None, :]
    ori[0, :, :, 1] = xs[:, None]
    ori[0, :, :, 2] = -3.0
    dirs = np.zeros_like(ori)
    dirs[..., 2] = 1.0
    inp = _inputs(sc)
    rays = O.Rays(ori, dirs, np.eye(4, dtype=np.float32)[None])
    fwd = O.grt_trace(inp, rays)
    keep = (fwd["margin"] >= MARGIN).astype(np.float32)[..., None]
    assert (fwd["distortion"][..., 0] > 0).mean() > 0.1
    rng = np.random.default_rng(11)
    g = {"hit2": (rng.standard_normal((1, res, res, 1)) * keep).astype(np.float32),
         "distortion": (rng.standard_normal((1, res, res, 1)) * keep).astype(np.float32),
         "hit": (rng.standard_normal((1, res, res, 1)) * keep).astype(np.float32)}
    dd, _, _ = O.grt_trace_bwd(inp, rays, fwd, g, inclusive_end=True)
    for rep in range(3):
        t = np.zeros_like(dd, dtype=np.float32)
        t[:, 2] = rng.standard_normal(t.shape[0])  # z of every position
        _, der = O.grt_trace_dual(inp, rays, {"density": t})
        assert np.abs(der["dns"]).max() < 1e-9  # alphas do not move
        terms = np.concatenate([(g[k].astype(np.float64) * der[k][..., :1]).reshape(-1)
                                for k in ("hit2", "distortion", "hit")])
        lhs = float((dd * t).sum())
        rhs = float(terms.sum())
        scale = float(np.abs(terms).sum())
        print(f"moment/distortion {rep}: restated {lhs:+.9e} dual {rhs:+.9e} scale {scale:.3e}")
        assert scale > 0 and abs(lhs - rhs) <= 1e-6 * scale


def test_grt_backward_excludes_the_last_forward_hit():
    """One Gaussian on one ray: its hit is the last forward hit, at distance t*; the float32 end
    bound t* + 1e-9 rounds back to t*, so the reference backward gives it no gradient."""
    dens = np.array([[0.0, 0.0, 0.0, 0.8, 1.0, 0.0, 0.0, 0.0, 0.2, 0.2, 0.2, 0.0]], np.float32)
    sh = np.zeros((1, 48), np.float32)
    sh[0, 0:3] = 0.5
    inp = O.grt_inputs(dens, sh, np.array([[0, 0, 1]], np.float32), sph_degree=0)
    ori = np.array([[[[0.05, 0.02, -2.0]]]], np.float32)
    dirs = np.array([[[[0.0, 0.0, 1.0]]]], np.float32)
    rays = O.Rays(ori, dirs, np.eye(4, dtype=np.float32)[None])
    fwd = O.grt_trace(inp, rays)
    assert fwd["dns"][0, 0, 0, 0] > 0.5 and fwd["hit"][0, 0, 0, 1] > 1.0
    g = {"rad": np.ones((1, 1, 1, 3), np.float32), "dns": np.ones((1, 1, 1, 1), np.float32)}
    excl = O.grt_trace_bwd(inp, rays, fwd, g)
    incl = O.grt_trace_bwd(inp, rays, fwd, g, inclusive_end=True)
    assert all(np.abs(a).max() == 0.0 for a in excl)
    assert np.abs(incl[0]).max() > 0.0 and np.abs(incl[1]).max() > 0.0
