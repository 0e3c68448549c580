"""Pins of the oracle's ptir-mode backward against central differences of its own FP64 forward.

The reference backward (threedgptir_tracer/src/kernels/cuda/referenceBwdOptix.cu:27-74,
include/3dgptir/kernels/cuda/pathTracer.cuh:313-634) is a path-replay estimator.  With the
random numbers fixed it is the exact derivative of the forward for the quantities that
enter the path only through continuous factors:

* albedo: the BRDF value at every vertex (value = f cos / pdf, the pdf does not depend
  on albedo) and the NEE term at the first vertex (accumulateNeeGradBwd, pathTracer.cuh:506-548);
* metallic: the BRDF values, but NOT the NEE term, whose metallic gradient the reference
  leaves disabled (pathTracer.cuh:544-546), so it is exact only with MIS (and NEE) off;
* environment texels: the escape contributions and the `light` output through the
  bilinear weights (environment.cuh:96-114), but NOT the NEE light, whose env gradient is
  disabled (pathTracer.cuh:533-536), so again exact only with MIS off.

None of them changes a discrete decision of the path (hit acceptance, escape draws,
shadow-ray draws depend on geometry only), and the path value is a polynomial of low
degree in each of them, so central differences with a power-of-two step (exact in the
float32 inputs) agree with the analytic gradient to rounding.  A dropped term, a wrong
index or a transposed channel in the backward fails these checks.

Roughness is not checked this way: the reference differentiates the GGX sample direction
by a forward difference (material.cuh:367-378) and drops the NEE roughness term, so its
roughness gradient is not the derivative of the forward by construction.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ptir_amd import envmap, scenes


def _setup(max_bounces, mis, seed=3, n=640, res=14):
    sc = scenes.small_scene(seed, n=n, res=res, spp=1, max_bounces=max_bounces)
    sc.enable_metallic = True
    tab = envmap.pack_alias_table(envmap.build_environment_alias_table(sc.env))
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    rng = np.random.default_rng(seed + 100)
    shape = tuple(sc.rays_dir.shape[:3]) + (3,)
    gp = rng.normal(size=shape).astype(np.float32)
    gl = rng.normal(size=shape).astype(np.float32)
    return sc, tab, rays, gp, gl, dict(enable_mis=mis, enable_metallic=True)


def _loss(sc, tab, rays, gp, gl, opts, material=None, env=None):
    mat = sc.particle_material() if material is None else material
    inp = O.Inputs(sc.particle_density(), mat, sc.features, sc.shading_normal, sc.env if env is None else env, tab,
                   sph_degree=sc.sph_degree, min_transmittance=sc.min_transmittance, max_bounces=sc.max_bounces,
                   frame=sc.frame, **opts)
    out = O.trace(inp, rays, nthreads=4)
    return float((out["pbr"] * gp).sum() + (out["light"] * gl).sum()), out, inp


def _grads(sc, tab, rays, gp, gl, opts):
    _, out, inp = _loss(sc, tab, rays, gp, gl, opts)
    dmat, denv = O.trace_bwd(inp, rays, out["pbr"], out["light"], gp, gl, np.zeros(out["mat"].shape, np.float32),
                             nthreads=4)
    return dmat, denv


def _pick(g, k, rng):
    """k entries with the largest gradients plus k random non-zero ones."""
    flat = np.abs(g).reshape(-1)
    nz = np.nonzero(flat)[0]
    assert nz.size > 2 * k, "scene produced too few non-zero gradients"
    top = nz[np.argsort(flat[nz])[-k:]]
    rest = rng.choice(np.setdiff1d(nz, top), size=k, replace=False)
    return np.concatenate([top, rest])


@pytest.mark.parametrize("max_bounces,mis", [(2, True), (3, True), (2, False), (3, False)])
def test_albedo_gradient_is_the_derivative_of_the_forward(max_bounces, mis):
    sc, tab, rays, gp, gl, opts = _setup(max_bounces, mis)
    dmat, _ = _grads(sc, tab, rays, gp, gl, opts)
    rng = np.random.default_rng(7)
    h = 2.0 ** -9
    base = sc.particle_material()
    for flat in _pick(dmat[:, :3], 6, rng):
        i, c = divmod(int(flat), 3)
        hi, lo = base.clone(), base.clone()
        hi[i, c] += h
        lo[i, c] -= h
        fd = (_loss(sc, tab, rays, gp, gl, opts, material=hi)[0] - _loss(sc, tab, rays, gp, gl, opts, material=lo)[0]) / (2 * h)
        assert fd == pytest.approx(dmat[i, c], rel=1e-6, abs=1e-9), (i, c)


@pytest.mark.parametrize("max_bounces", [2, 3])
def test_metallic_gradient_is_the_derivative_of_the_forward_without_nee(max_bounces):
    sc, tab, rays, gp, gl, opts = _setup(max_bounces, mis=False)
    dmat, _ = _grads(sc, tab, rays, gp, gl, opts)
    rng = np.random.default_rng(8)
    h = 2.0 ** -9
    base = sc.particle_material()
    for i in _pick(dmat[:, 4], 6, rng):
        hi, lo = base.clone(), base.clone()
        hi[i, 4] += h
        lo[i, 4] -= h
        fd = (_loss(sc, tab, rays, gp, gl, opts, material=hi)[0] - _loss(sc, tab, rays, gp, gl, opts, material=lo)[0]) / (2 * h)
        assert fd == pytest.approx(dmat[i, 4], rel=1e-6, abs=1e-9), int(i)


def test_metallic_gradient_with_nee_omits_the_nee_term():
    """With MIS on, the forward's metallic derivative includes the NEE term the reference
    does not differentiate: the two must differ somewhere (else the checks above would not
    be telling apart the two estimators)."""
    sc, tab, rays, gp, gl, opts = _setup(2, mis=True)
    dmat, _ = _grads(sc, tab, rays, gp, gl, opts)
    rng = np.random.default_rng(9)
    h = 2.0 ** -9
    base = sc.particle_material()
    worst = 0.0
    for i in _pick(dmat[:, 4], 6, rng):
        hi, lo = base.clone(), base.clone()
        hi[i, 4] += h
        lo[i, 4] -= h
        fd = (_loss(sc, tab, rays, gp, gl, opts, material=hi)[0] - _loss(sc, tab, rays, gp, gl, opts, material=lo)[0]) / (2 * h)
        worst = max(worst, abs(fd - dmat[i, 4]) / max(abs(fd), 1e-12))
    assert worst > 1e-3


@pytest.mark.parametrize("max_bounces", [2, 3])
def test_environment_gradient_is_the_derivative_of_the_forward_without_nee(max_bounces):
    sc, tab, rays, gp, gl, opts = _setup(max_bounces, mis=False)
    _, denv = _grads(sc, tab, rays, gp, gl, opts)
    assert np.all(denv[..., 3] == 0)
    rng = np.random.default_rng(10)
    h = 2.0 ** -8
    for flat in _pick(denv[..., :3], 8, rng):
        y, rem = divmod(int(flat), denv.shape[1] * 3)
        x, c = divmod(rem, 3)
        hi, lo = sc.env.clone(), sc.env.clone()
        hi[y, x, c] += h
        lo[y, x, c] -= h
        fd = (_loss(sc, tab, rays, gp, gl, opts, env=hi)[0] - _loss(sc, tab, rays, gp, gl, opts, env=lo)[0]) / (2 * h)
        assert fd == pytest.approx(denv[y, x, c], rel=1e-6, abs=1e-9), (y, x, c)


def test_backward_is_linear_in_the_output_gradients():
    sc, tab, rays, gp, gl, opts = _setup(3, mis=True)
    _, out, inp = _loss(sc, tab, rays, gp, gl, opts)
    z = np.zeros(out["mat"].shape, np.float32)
    a = O.trace_bwd(inp, rays, out["pbr"], out["light"], gp, gl, z, nthreads=4)
    b = O.trace_bwd(inp, rays, out["pbr"], out["light"], 2 * gp, gl, z, nthreads=4)
    c = O.trace_bwd(inp, rays, out["pbr"], out["light"], gp, 0 * gl, z, nthreads=4)
    d = O.trace_bwd(inp, rays, out["pbr"], out["light"], 2 * gp, 0 * gl, z, nthreads=4)
    # g(2 gp, gl) - g(gp, gl) = g(2 gp, 0) - g(gp, 0)
    for k in range(2):
        np.testing.assert_allclose(b[k] - a[k], d[k] - c[k], rtol=1e-9, atol=1e-12)
