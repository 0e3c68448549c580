"""The render pipeline around the tracer: sample positions, camera rays, resolve, masks, and
Tracer(conf).build_acc(...).render(...) end to end.

The HIP kernels (camera rays, resolve forward and backward) are checked against plain
float32 torch restatements of the reference's trainer-side code (threedgptir_tracer/
tracer.py:505-526,646-706; threedgrut/model/filters.py:6-69); the render against HipTracer
plus the oracle.
"""

import math
import types

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from gpu_helpers import env_table, oracle_forward, pixel_agreement, to_np
from ptir_amd import render, scenes
from ptir_amd.tracer import HipTracer, Tracer


# ----------------------------------------------------------------------------- CPU
def test_sample_positions_tables_and_stratified_grid():
    for spp, pat in render.SPP_PATTERNS.items():
        got = render.sample_positions(spp, 0, "cpu")
        assert got.shape == (spp, 2)
        torch.testing.assert_close(got, torch.tensor(pat, dtype=torch.float32))
    # 4 spp: the rotated-grid pattern, one sample per row and column of a 4x4 grid
    p4 = render.sample_positions(4, 0, "cpu") * 4
    assert sorted((p4[:, 0] - 0.5).round().tolist()) == [0, 1, 2, 3]
    assert sorted((p4[:, 1] - 0.5).round().tolist()) == [0, 1, 2, 3]
    # 6 spp: a 3x2 stratified grid under one shift mod 1 -> one sample per stratum
    a = render.sample_positions(6, 5, "cpu")
    cells = {(int(x * 3), int(y * 2)) for x, y in a.tolist()}
    assert len(cells) == 6
    # the shift depends on the frame, deterministically
    torch.testing.assert_close(a, render.sample_positions(6, 5, "cpu"))
    assert not torch.equal(a, render.sample_positions(6, 6, "cpu"))
    g = torch.Generator().manual_seed(6 * 1315423911 + 6 * 2654435761)
    shift = torch.rand((1, 2), generator=g)
    k = torch.arange(6.0)
    grid = torch.stack(((k % 3 + 0.5) / 3, (torch.floor(k / 3) + 0.5) / 2), 1)
    torch.testing.assert_close(a, torch.remainder(grid + shift, 1.0))


def test_pinhole_intrinsics_forms():
    b = types.SimpleNamespace(intrinsics=[10.0, 11.0, 3.0, 4.0])
    torch.testing.assert_close(render.pinhole_intrinsics(b, "cpu"), torch.tensor([10.0, 11.0, 3.0, 4.0]))
    b = types.SimpleNamespace(intrinsics=[[10.0, 0, 3.0], [0, 11.0, 4.0], [0, 0, 1]])
    torch.testing.assert_close(render.pinhole_intrinsics(b, "cpu"), torch.tensor([10.0, 11.0, 3.0, 4.0]))
    assert render.pinhole_intrinsics(types.SimpleNamespace(intrinsics=[1.0, 2.0]), "cpu") is None
    params = {"focal_length": [9.0], "principal_point": [2.0, 3.0], "radial_coeffs": [0.0, 0.0]}
    b = types.SimpleNamespace(intrinsics_OpenCVPinholeCameraModelParameters=params)
    torch.testing.assert_close(render.pinhole_intrinsics(b, "cpu"), torch.tensor([9.0, 9.0, 2.0, 3.0]))
    params["radial_coeffs"] = [0.1]
    assert render.pinhole_intrinsics(b, "cpu") is None
    assert render.pinhole_intrinsics(types.SimpleNamespace(), "cpu") is None


def test_gt_mask_touches_only_material_light_and_shading_normal():
    outs = {k: torch.ones(2, 4, 5, 3) for k in ("pred_pbr", "pred_light", "pred_material", "pred_shadingnormal")}
    mask = torch.zeros(4, 5)
    mask[1:3, 2:4] = 1.0
    res = render.apply_gt_mask(outs, types.SimpleNamespace(mask=mask))
    assert torch.equal(res["pred_pbr"], outs["pred_pbr"])
    for k in render.GT_MASKED_OUTPUTS:
        assert float(res[k].sum()) == 2 * 4 * 3
    assert res[render.GT_MASK_APPLIED]
    again = render.apply_gt_mask(res, types.SimpleNamespace(mask=torch.zeros(4, 5)))
    assert torch.equal(again["pred_light"], res["pred_light"])  # applied once
    for m in (mask[None, ..., None], mask[..., None], mask[None].expand(2, -1, -1)):
        r = render.apply_gt_mask(outs, types.SimpleNamespace(mask=m))
        torch.testing.assert_close(r["pred_light"], res["pred_light"])
    with pytest.raises(ValueError):
        render.apply_gt_mask(outs, types.SimpleNamespace(mask=torch.zeros(3, 3)))
    assert render.apply_gt_mask(outs, types.SimpleNamespace()) is outs


def test_filter_option():
    assert render.filter_iterations(None) == 0
    assert render.filter_iterations("none") == 0
    assert render.filter_iterations("ATROUS") == 1
    with pytest.raises(ValueError):
        render.filter_iterations("gaussian")


# ----------------------------------------------------------------------------- torch references
def _atrous_reference(x, iterations):
    """3x3 B-spline a-trous filter with replicate padding on [B,H,W,C] (filters.py:46-69)."""
    k = torch.tensor([[1, 2, 1], [2, 4, 2], [1, 2, 1]], dtype=x.dtype, device=x.device) / 16
    y = x.permute(0, 3, 1, 2)
    for i in range(iterations):
        d = 2 ** i
        y = F.conv2d(F.pad(y, (d, d, d, d), mode="replicate"), k.view(1, 1, 3, 3).repeat(y.shape[1], 1, 1, 1),
                     dilation=d, groups=y.shape[1])
    return y.permute(0, 2, 3, 1)


def _resolve_reference(it, op, dist, pbr, direct, indirect, light, nrm):
    return (torch.nan_to_num(dist / op, 0.0, 0.0), _atrous_reference(pbr * op, it), _atrous_reference(direct * op, it),
            _atrous_reference(indirect * op, it), _atrous_reference(light, it), F.normalize(nrm, dim=3))


# ----------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("iterations,shape", [(1, (2, 9, 7)), (0, (1, 5, 6)), (3, (1, 17, 13)), (2, (1, 1, 6))])
def test_resolve_forward_and_backward_match_torch(device, iterations, shape):
    gen = torch.Generator().manual_seed(iterations)
    B, H, W = shape
    op = torch.rand(B, H, W, 1, generator=gen)
    op[0, 0, 0] = 0.0  # depth / 0 -> 0
    ins = [op, torch.rand(B, H, W, 1, generator=gen)] + [torch.rand(B, H, W, 3, generator=gen) for _ in range(5)]
    ins[6][0, -1, -1] = 0.0  # zero normal
    dev_in = [t.to(device).requires_grad_(True) for t in ins]
    ref_in = [t.to(device).requires_grad_(True) for t in ins]
    got = render.Resolve.apply(iterations, *dev_in)
    ref = _resolve_reference(iterations, *ref_in)
    for a, b in zip(got, ref):
        torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-6)
    gouts = [torch.randn(r.shape, generator=gen).to(device) for r in ref]
    ga = torch.autograd.grad(got, dev_in, gouts)
    gb = torch.autograd.grad(ref, ref_in, gouts)
    for k, (a, b) in enumerate(zip(ga, gb)):
        if k == 6:  # zero normal: torch's normalize has gradient 1/eps there too
            torch.testing.assert_close(a[0, :-1], b[0, :-1], rtol=1e-4, atol=1e-5)
        elif k in (0, 1):
            # at zero opacity the depth quotient is non-finite and sent to 0; torch's chain
            # rule gives 0 * inf = nan for opacity and depth there, the kernel drops that term
            assert bool(torch.isfinite(a).all()) and not bool(torch.isfinite(b[0, 0, 0, 0]))
            a, b = a.clone(), b.clone()
            a[0, 0, 0, 0] = b[0, 0, 0, 0] = 0.0
            torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5)
        else:
            torch.testing.assert_close(a, b, rtol=1e-4, atol=1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("spp,per_pixel", [(4, False), (3, True)])
def test_camera_rays_match_torch(device, spp, per_pixel):
    B, H, W = 2, 5, 7
    gen = torch.Generator().manual_seed(spp)
    ys, xs = torch.meshgrid(torch.arange(H) + 0.5, torch.arange(W) + 0.5, indexing="ij")
    batch = types.SimpleNamespace(rays_ori=torch.randn(B, H, W, 3, generator=gen).to(device),
                                  pixel_coords=torch.stack([xs, ys], -1).expand(B, H, W, 2).contiguous().to(device))
    K = torch.tensor([12.0, 13.0, W / 2, H / 2], device=device)
    pos = torch.rand((spp, H, W, 2) if per_pixel else (spp, 2), generator=gen).to(device)
    o, d = render.camera_rays(batch, pos, K)
    j = pos.view(spp, 1, H, W, 2) if per_pixel else pos.view(spp, 1, 1, 1, 2)
    px = batch.pixel_coords.unsqueeze(0) - 0.5
    ref = torch.stack(((px[..., 0] + j[..., 0] - K[2]) / K[0], (px[..., 1] + j[..., 1] - K[3]) / K[1],
                       torch.ones(spp, B, H, W, device=device)), -1)
    ref = F.normalize(ref, dim=-1).reshape(spp * B, H, W, 3)
    torch.testing.assert_close(d, ref, rtol=1e-6, atol=1e-6)
    torch.testing.assert_close(o, batch.rays_ori.unsqueeze(0).expand(spp, -1, -1, -1, -1).reshape(spp * B, H, W, 3))


class _Model:
    """Duck-typed stand-in of the PTIR model that Tracer.render reads (model.py:1049-1063):
    raw parameters plus their activations, materials, environment and background."""

    def __init__(self, sc, device, table=None):
        g = sc.to(device)
        # the parameters are stored activated (identity activations), so the render traces
        # bit-for-bit the tensors the direct traces of the tests use
        self.positions = g.positions
        self.rotation = g.rotation
        self.scale = g.scale
        self.density = g.density
        self.rotation_activation = self.scale_activation = self.density_activation = lambda t: t
        self._g = g
        self.n_active_features = sc.sph_degree
        self.environment_alias_table = table

    def get_rotation(self):
        return self.rotation_activation(self.rotation)

    def get_scale(self):
        return self.scale_activation(self.scale)

    def get_density(self):
        return self.density_activation(self.density)

    def get_features(self):
        return self._g.features

    def get_shading_normal(self):
        return self._g.shading_normal

    def get_material_albedo(self):
        return self._g.albedo

    def get_material_roughness(self):
        return self._g.roughness

    def get_material_metallic(self):
        return self._g.metallic

    def get_environment(self):
        return self._g.env

    def background(self, T, rays, rgb, opacity, train):
        return rgb, opacity


def _conf(**render_opts):
    r = dict(max_bounces=2, min_transmittance=1e-3, enable_metallic=True, render_spp=1, inversion_spp=1,
             filter_type="atrous")
    r.update(render_opts)
    return {"render": r, "loss": {}}


@pytest.mark.gpu
def test_tracer_render_matches_direct_trace_and_oracle(device):
    """Tracer(conf).build_acc(model).render(batch) at 1 spp equals HipTracer.trace on the
    batch's rays followed by the torch restatement of the resolve, and its pbr before the
    resolve is the oracle's."""
    sc = scenes.small_scene(0, spp=1)
    tab = env_table(sc)
    from ptir_amd import envmap

    alias = envmap.build_environment_alias_table(sc.env)
    model = _Model(sc, device, alias)
    tracer = Tracer(_conf())
    tracer.build_acc(model)
    g = sc.to(device)
    batch = types.SimpleNamespace(rays_ori=g.rays_ori, rays_dir=g.rays_dir, T_to_world=g.ray_to_world)
    out = tracer.render(model, batch, train=False, frame_id=0)
    ht = HipTracer(enable_metallic=True)
    ht.build_bvh(g.positions, g.rotation, g.scale, g.density)
    raw = ht.trace(0, g.ray_to_world, g.rays_ori, g.rays_dir, g.particle_density(), g.particle_material(), g.features,
                   g.shading_normal, g.env, tab.to(device), max_bounces=2)
    ref = _resolve_reference(1, raw[1], raw[2][..., 0:1], raw[10], raw[12][..., 0, :], raw[12][..., 1, :], raw[11],
                             raw[5])
    for key, r in zip(("pred_dist", "pred_pbr", "pred_direct", "pred_indirect", "pred_light", "pred_normals"), ref):
        torch.testing.assert_close(out[key], r, rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(out["pred_opacity"], raw[1])
    torch.testing.assert_close(out["pred_material"], raw[7])
    _, _, oref = oracle_forward(sc, tab)
    assert pixel_agreement(to_np(raw)["pbr"], oref["pbr"], 1e-3, 1e-5) >= 0.999
    # build_acc with unchanged geometry does not rebuild; an optimiser-style in-place update does
    key = tracer._built_geometry
    tracer.build_acc(model)
    assert tracer._built_geometry == key
    with torch.no_grad():
        model.positions.add_(0.0)
    tracer.build_acc(model, rebuild=False)
    assert tracer._built_geometry != key


@pytest.mark.gpu
@pytest.mark.parametrize("sampling", ["alias", "cdf"])
def test_tracer_render_spp_and_gradients(device, sampling):
    """4 spp through pinhole intrinsics, traced in chunks of 2 samples: equals the resolve of
    the sample means of the direct traces (chunk c traced with frame frame_id + first sample
    of c); a loss on pred_pbr backpropagates to the materials and the environment."""
    from ptir_amd import envmap

    sc = scenes.small_scene(1, n=2048, res=24, spp=1)
    model = _Model(sc, device, envmap.build_environment_alias_table(sc.env))
    g = model._g
    g.albedo.requires_grad_(True)
    g.env.requires_grad_(True)
    tracer = Tracer(_conf(render_spp=4, inversion_spp=4, spp_chunk=2, env_sampling=sampling))
    tracer.build_acc(model)
    H = W = 24
    f = 0.5 * W / math.tan(math.radians(20.0))
    ys, xs = torch.meshgrid(torch.arange(H) + 0.5, torch.arange(W) + 0.5, indexing="ij")
    batch = types.SimpleNamespace(rays_ori=g.rays_ori, rays_dir=g.rays_dir, T_to_world=g.ray_to_world,
                                  pixel_coords=torch.stack([xs, ys], -1).unsqueeze(0).to(device),
                                  intrinsics=[f, f, W / 2, H / 2])
    out = tracer.render(model, batch, train=True, frame_id=3)
    K = torch.tensor([f, f, W / 2, H / 2], device=device)
    pos = render.sample_positions(4, 3, device)
    tab = (envmap.build_cdf_table(g.env.detach()) if sampling == "cdf"
           else envmap.pack_alias_table(model.environment_alias_table, device))
    mat = torch.cat([g.albedo, g.roughness, g.metallic], 1).detach()
    sums = [0, 0, 0]
    with torch.no_grad():
        for start in (0, 2):
            o, d = render.camera_rays(batch, pos[start:start + 2], K)
            raw = tracer.tracer_wrapper.trace(3 + start, g.ray_to_world, o, d, g.particle_density(), mat, g.features,
                                              g.shading_normal, g.env.detach(), tab, max_bounces=2)
            for k, idx in enumerate((1, 10, 11)):
                sums[k] = sums[k] + raw[idx].sum(0, keepdim=True)
    op, pbr, light = (t / 4 for t in sums)
    torch.testing.assert_close(out["pred_opacity"], op, rtol=1e-6, atol=1e-7)
    torch.testing.assert_close(out["pred_pbr"], _atrous_reference(pbr * op, 1), rtol=1e-5, atol=1e-6)
    torch.testing.assert_close(out["pred_light"], _atrous_reference(light, 1), rtol=1e-5, atol=1e-6)
    loss = out["pred_pbr"].square().mean()
    loss.backward()
    assert float(g.albedo.grad.abs().sum()) > 0 and bool(torch.isfinite(g.albedo.grad).all())
    assert float(g.env.grad.abs().sum()) > 0 and bool(torch.isfinite(g.env.grad).all())
