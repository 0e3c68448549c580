"""The 3DGRT primary-only tracer on the GPU (ptir_grt_trace / ptir_grt_trace_bwd through
ptir_amd.grt) against the FP64 oracle's restatement (pinned in tests/test_grt_oracle.py):
the nine forward outputs per pixel, the gradients of every Gaussian parameter block, the
autograd function, and invariance to the hierarchy builder and the primary traversal.
Output gradients are zero on pixels within float32 reach of a discrete decision (oracle
decision margin < 3e-3)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ptir_amd import scenes
from ptir_amd.grt import GRTTracer, Tracer

pytestmark = pytest.mark.gpu

MARGIN = 3e-3
NAMES = ["rad", "dns", "hit", "hit2", "distortion", "nrm", "snrm", "hits"]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _scene(seed, n=2048, res=48):
    sc = scenes.small_scene(seed, n=n, res=res, spp=1, max_bounces=1)
    sc.density = torch.clamp(sc.density, max=0.9)
    return sc


def _gpu(tr, sc, device):
    g = sc.to(device)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    args = (sc.frame, g.ray_to_world, g.rays_ori, g.rays_dir, g.particle_density(), g.features, g.shading_normal)
    return tr.trace(*args, 0, sc.sph_degree, sc.min_transmittance), args


@pytest.mark.parametrize("seed,dense", [(3, False), (9, False), (5, True)])
def test_grt_forward_and_backward_match_the_oracle(device, seed, dense):
    """dense: four times the Gaussians at 0.3 times the density, so that rays cross many of
    them before the transmittance ends: a third of the pixels then hold more hits than the
    backward's 32 records per pixel, and those past the records add their gradients by atomics."""
    sc = _scene(seed, n=8192 if dense else 2048)
    if dense:
        sc.density = sc.density * 0.3
    inp = O.grt_inputs(sc.particle_density(), sc.features, sc.shading_normal, sph_degree=sc.sph_degree,
                       min_transmittance=sc.min_transmittance)
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    ref = O.grt_trace(inp, rays)
    keep = ref["margin"] >= MARGIN
    print(f"grt seed {seed}: {keep.mean():.4f} of pixels kept", flush=True)
    tr = GRTTracer()
    outs, args = _gpu(tr, sc, device)
    torch.cuda.synchronize()
    for name, v in zip(NAMES, outs[:8]):
        got = v.cpu().double().numpy()
        close = np.isclose(got, ref[name], rtol=1e-3, atol=1e-4).all(axis=-1)
        assert close[keep].mean() >= 0.999, (name, close[keep].mean())
    vis = outs[8].view(torch.int32).cpu().numpy().reshape(-1)
    assert np.array_equal(vis != 0, ref["visibility"] > 0)
    tr.check()
    most = float(outs[7].max())
    print(f"grt seed {seed}: at most {most:.0f} hits per pixel", flush=True)
    if dense:
        assert most > 40, most  # past the 32 records per pixel

    rng = np.random.default_rng(seed)
    B, H, W = sc.rays_dir.shape[:3]
    k = keep.astype(np.float32)[..., None]
    g = {"rad": rng.standard_normal((B, H, W, 3)) * k, "dns": rng.standard_normal((B, H, W, 1)) * k,
         "hit": rng.standard_normal((B, H, W, 1)) * k, "hit2": rng.standard_normal((B, H, W, 1)) * k,
         "distortion": rng.standard_normal((B, H, W, 1)) * k, "snrm": rng.standard_normal((B, H, W, 3)) * k}
    g = {key: v.astype(np.float32) for key, v in g.items()}
    dd, dsh, dsn = O.grt_trace_bwd(inp, rays, ref, g)
    t = {key: torch.from_numpy(v).to(device) for key, v in g.items()}
    hd, hsh, hsn = tr.trace_bwd(*args[:4], *outs[:7], *args[4:], t["rad"], t["dns"], t["hit"], t["hit2"],
                                t["distortion"], torch.zeros_like(outs[5]), t["snrm"], 0, sc.sph_degree,
                                sc.min_transmittance)
    torch.cuda.synchronize()
    hd, hsh, hsn = (x.cpu().double().numpy() for x in (hd, hsh, hsn))
    for name, a, b in (("position", hd[:, 0:3], dd[:, 0:3]), ("density", hd[:, 3], dd[:, 3]),
                       ("rotation", hd[:, 4:8], dd[:, 4:8]), ("scale", hd[:, 8:11], dd[:, 8:11]),
                       ("sh", hsh, dsh), ("snormal", hsn, dsn)):
        err = rel_l2(a, b)
        print(f"grt seed {seed} {name}: relative L2 {err:.2e}", flush=True)
        assert np.isfinite(a).all() and err < 2e-3, name

    # the backward that reads the forward's hit tape (16 hits per pixel: pixels with more
    # re-trace) gives the re-tracing backward's gradients (to the float rounding of the
    # atomics' order) and the same forward outputs
    tape = tr.new_tape(args[3], hits_per_pixel=16)
    outs_t = tr.trace(*args, 0, sc.sph_degree, sc.min_transmittance, tape=tape)
    for a, b in zip(outs, outs_t):
        assert torch.equal(a, b)
    td, tsh, tsn = tr.trace_bwd(*args[:4], *outs_t[:7], *args[4:], t["rad"], t["dns"], t["hit"], t["hit2"],
                                t["distortion"], torch.zeros_like(outs[5]), t["snrm"], 0, sc.sph_degree,
                                sc.min_transmittance, tape=tape)
    torch.cuda.synchronize()
    redo = float((tape[16:16 + B * H * W] < 0).float().mean())  # count words with bit 31 set
    print(f"grt seed {seed}: {redo:.3f} of pixels re-traced with the tape", flush=True)
    for name, a, b in (("density", td, hd), ("sh", tsh, hsh), ("snormal", tsn, hsn)):
        a = a.cpu().double().numpy()
        err = rel_l2(a, b)
        print(f"grt seed {seed} {name}: tape vs re-trace relative L2 {err:.2e}", flush=True)
        assert err < 1e-5, name


def test_grt_outputs_do_not_depend_on_builder_or_traversal(device):
    sc = _scene(4)
    ref = None
    for kw in (dict(), dict(traversal="lane"), dict(builder="ploc"), dict(kbuffer=8), dict(kbuffer=32)):
        outs, _ = _gpu(GRTTracer(**kw), sc, device)
        cur = [o.clone() for o in outs]
        if ref is None:
            ref = cur
        else:
            for i, (a, b) in enumerate(zip(ref, cur)):
                assert torch.equal(a, b), (kw, i)


class _Gaussians:
    """The duck-typed model the 3DGRT Tracer reads (positions, get_*, background)."""

    def __init__(self, sc, device):
        g = sc.to(device)
        self.positions = g.positions.clone().requires_grad_(True)
        self.rotation = g.rotation.clone().requires_grad_(True)
        self.scale = g.scale.clone().requires_grad_(True)
        self.density = g.density.clone().requires_grad_(True)
        self.features = g.features.clone().requires_grad_(True)
        self.snormal = g.shading_normal.clone().requires_grad_(True)
        self.n_active_features = sc.sph_degree

    def get_rotation(self):
        return self.rotation

    def get_scale(self):
        return self.scale

    def get_density(self):
        return self.density

    def get_features(self):
        return self.features

    def get_shading_normal(self):
        return self.snormal

    def background(self, T, dirs, rgb, opacity, train):
        return rgb + (1.0 - opacity) * 0.25, opacity


class _Batch:
    def __init__(self, sc, device):
        g = sc.to(device)
        self.T_to_world = g.ray_to_world
        self.rays_ori = g.rays_ori
        self.rays_dir = g.rays_dir


def test_grt_tracer_render_backpropagates_to_every_parameter(device):
    sc = _scene(6, n=2048, res=32)
    conf = {"render": {"min_transmittance": sc.min_transmittance, "particle_kernel_density_clamping": True}}
    tracer = Tracer(conf)
    gs = _Gaussians(sc, device)
    tracer.build_acc(gs)
    out = tracer.render(gs, _Batch(sc, device))
    assert set(out) >= {"pred_rgb", "pred_opacity", "pred_dist", "pred_normals", "mog_visibility"}
    loss = out["pred_rgb"].square().mean() + out["pred_opacity"].mean() + out["pred_dist"].mean() + \
        out["pred_shadingnormal"].sum() * 1e-3
    loss.backward()
    torch.cuda.synchronize()
    vis = out["mog_visibility"].view(torch.int32).reshape(-1) != 0
    for name in ("positions", "rotation", "scale", "density", "features", "snormal"):
        gr = getattr(gs, name).grad
        assert gr is not None and bool(torch.isfinite(gr).all()), name
        nz = gr.abs().reshape(gr.shape[0], -1).sum(1) > 0
        assert bool(nz.any()), name
        assert not bool((nz & ~vis).any()), name  # only Gaussians that were hit
