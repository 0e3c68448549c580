"""Render-side work around the tracer: samples per pixel, camera rays, resolve, masks.

The reference does this with torch ops in `Tracer.render` (threedgptir_tracer/tracer.py:
391-426, 505-706, 740-975), threedgrut/model/filters.py:6-69 and
threedgrut/model/ptir_helper.py:100-122.  Here the per-pixel work runs in HIP kernels of
libptir_amd.so:

* ``ptir_camera_rays`` expands a batch of pinhole views into S jittered samples per pixel;
* ``ptir_resolve_forward`` / ``ptir_resolve_backward`` turn the spp-averaged tracer
  outputs into render outputs (depth / opacity, opacity-weighted PBR terms, the a-trous
  B-spline filter, unit normals) and differentiate that map.

What stays in Python is bookkeeping: which sample positions a given spp uses, how the
pinhole intrinsics are found on a batch, and the ground-truth mask of the PBR outputs.
"""

from __future__ import annotations

import ctypes as C
import logging
import math
from dataclasses import dataclass

import torch

from . import _lib

logger = logging.getLogger(__name__)

# Sub-pixel sample positions for the spp counts the reference tabulates (tracer.py:391-426);
# parity of the rendered image needs the same positions.
SPP_PATTERNS = {
    1: ((0.5, 0.5),),
    2: ((0.25, 0.25), (0.75, 0.75)),
    4: ((0.375, 0.125), (0.875, 0.375), (0.625, 0.875), (0.125, 0.625)),
    8: ((0.5625, 0.6875), (0.4375, 0.3125), (0.8125, 0.4375), (0.3125, 0.8125), (0.1875, 0.1875),
        (0.0625, 0.5625), (0.6875, 0.0625), (0.9375, 0.9375)),
    16: ((0.5625, 0.4375), (0.4375, 0.6875), (0.3125, 0.375), (0.75, 0.5625), (0.1875, 0.625), (0.625, 0.1875),
         (0.1875, 0.3125), (0.6875, 0.8125), (0.375, 0.125), (0.5, 0.9375), (0.25, 0.875), (0.125, 0.25),
         (0.0, 0.5), (0.9375, 0.75), (0.875, 0.0625), (0.0625, 0.0)),
    32: tuple(((i % 8 + 0.5) / 8.0, (i // 8 + 0.5) / 4.0) for i in range(32)),
}

# Outputs whose pixels a ground-truth foreground mask zeroes (ptir_helper.py:30-44):
# material-like images; the PBR images themselves stay unmasked.
GT_MASKED_OUTPUTS = ("pred_light", "pred_shadingnormal", "pred_material")
GT_MASK_APPLIED = "_pbr_gt_mask_applied"


def sample_positions(spp: int, frame_id: int, device, dtype=torch.float32) -> torch.Tensor:
    """[spp, 2] sub-pixel positions in [0,1)^2 of the samples of one pixel.

    Tabulated counts use the fixed patterns; any other count uses a ceil(sqrt(spp))-wide
    stratified grid shifted (mod 1) by one uniform offset drawn from a torch generator
    seeded with ((frame_id + 1) * 1315423911 + spp * 2654435761), as the reference does."""
    if spp in SPP_PATTERNS:
        return torch.tensor(SPP_PATTERNS[spp], dtype=dtype, device=device)
    cols = int(math.ceil(math.sqrt(spp)))
    rows = int(math.ceil(spp / cols))
    k = torch.arange(spp, dtype=dtype, device=device)
    grid = torch.stack(((k % cols + 0.5) / cols, (torch.floor(k / cols) + 0.5) / rows), dim=1)
    gen = torch.Generator(device=device)
    gen.manual_seed((int(frame_id) + 1) * 1315423911 + spp * 2654435761)
    offset = torch.rand((1, 2), dtype=dtype, device=device, generator=gen)
    return torch.remainder(grid + offset, 1.0)


def _flat(v, device):
    return None if v is None else torch.as_tensor(v, dtype=torch.float32, device=device).reshape(-1)


def pinhole_intrinsics(batch, device) -> torch.Tensor | None:
    """Device [4] (fx, fy, cx, cy) of a batch, or None when its camera is not a distortion-free
    pinhole.  Accepted forms: `intrinsics` as 4 values or a 3x3 K, or
    `intrinsics_OpenCVPinholeCameraModelParameters` with zero distortion coefficients."""
    direct = _flat(getattr(batch, "intrinsics", None), device)
    if direct is not None:
        if direct.numel() == 4:
            return direct.contiguous()
        if direct.numel() == 9:
            return direct[[0, 4, 2, 5]].contiguous()
        return None
    params = getattr(batch, "intrinsics_OpenCVPinholeCameraModelParameters", None)
    if params is None:
        return None
    field = params.get if isinstance(params, dict) else (lambda k: getattr(params, k, None))
    for name in ("radial_coeffs", "tangential_coeffs", "thin_prism_coeffs"):
        c = _flat(field(name), device)
        if c is not None and c.numel() and bool((c.abs() > 1e-8).any()):
            return None
    f, p = _flat(field("focal_length"), device), _flat(field("principal_point"), device)
    if f is None or p is None or f.numel() < 1 or p.numel() < 2:
        return None
    return torch.stack((f[0], f[1] if f.numel() > 1 else f[0], p[0], p[1])).contiguous()


def camera_rays(batch, positions: torch.Tensor, intrinsics: torch.Tensor):
    """Rays of every sample of every pixel: ([S*B,H,W,3] origins, [S*B,H,W,3] unit directions).
    positions is [S,2] (shared by all pixels) or [S,H,W,2]."""
    ori = batch.rays_ori.contiguous().float()
    B, H, W, _ = ori.shape
    S = positions.shape[0]
    pix = batch.pixel_coords.to(device=ori.device, dtype=torch.float32).reshape(B, H, W, 2).contiguous()
    per_pixel = positions.dim() == 4
    jit = positions.to(device=ori.device, dtype=torch.float32).contiguous()
    out_o = torch.empty((S * B, H, W, 3), dtype=torch.float32, device=ori.device)
    out_d = torch.empty_like(out_o)
    stream = torch.cuda.current_stream(ori.device).cuda_stream
    _lib.check(_lib.lib().ptir_camera_rays(pix.data_ptr(), intrinsics.data_ptr(), jit.data_ptr(), int(per_pixel),
                                           ori.data_ptr(), S, B, H, W, out_o.data_ptr(), out_d.data_ptr(),
                                           C.c_void_p(stream)), None)
    return out_o, out_d


def _resolve_struct(iterations, tensors, outs, image, work):
    r = _lib.Resolve()
    op = tensors[0]
    r.B, r.H, r.W = op.shape[0], op.shape[1], op.shape[2]
    r.iterations = iterations
    (r.opacity, r.dist, r.pbr, r.direct, r.indirect, r.light, r.normals) = (t.data_ptr() for t in tensors)
    r.out_dist, r.out_normals = (None, None) if outs is None else (outs[0].data_ptr(), outs[1].data_ptr())
    r.image, r.work = image.data_ptr(), work.data_ptr()
    return r


class Resolve(torch.autograd.Function):
    """Averaged tracer outputs -> render outputs, forward and backward on the device.

    inputs  : opacity [B,H,W,1], dist [B,H,W,1], pbr, direct, indirect, light, normals [B,H,W,3]
    outputs : dist / opacity, filter(pbr * opacity), filter(direct * opacity),
              filter(indirect * opacity), filter(light), normals / |normals|
    `iterations` = 0 disables the filter."""

    @staticmethod
    def forward(ctx, iterations, opacity, dist, pbr, direct, indirect, light, normals):
        ins = [t.detach().contiguous().float() for t in (opacity, dist, pbr, direct, indirect, light, normals)]
        B, H, W, _ = ins[0].shape
        image = torch.empty((B, H, W, 12), dtype=torch.float32, device=ins[0].device)
        work = torch.empty_like(image)
        out_dist = torch.empty_like(ins[1])
        out_nrm = torch.empty_like(ins[6])
        stream = torch.cuda.current_stream(ins[0].device).cuda_stream
        r = _resolve_struct(int(iterations), ins, (out_dist, out_nrm), image, work)
        _lib.check(_lib.lib().ptir_resolve_forward(C.byref(r), C.c_void_p(stream)), None)
        ctx.iterations = int(iterations)
        ctx.save_for_backward(*ins)
        return (out_dist, image[..., 0:3].contiguous(), image[..., 3:6].contiguous(), image[..., 6:9].contiguous(),
                image[..., 9:12].contiguous(), out_nrm)

    @staticmethod
    def backward(ctx, g_dist, g_pbr, g_dir, g_ind, g_light, g_nrm):
        ins = ctx.saved_tensors
        z3 = lambda g: torch.zeros_like(ins[2]) if g is None else g.float()  # noqa: E731
        image = torch.cat([z3(g_pbr), z3(g_dir), z3(g_ind), z3(g_light)], dim=-1).contiguous()
        work = torch.empty_like(image)
        gd = (torch.zeros_like(ins[1]) if g_dist is None else g_dist.float()).contiguous()
        gn = (torch.zeros_like(ins[6]) if g_nrm is None else g_nrm.float()).contiguous()
        grads = [torch.empty_like(t) for t in ins]  # opacity, dist, pbr, direct, indirect, light, normals
        stream = torch.cuda.current_stream(ins[0].device).cuda_stream
        r = _resolve_struct(ctx.iterations, ins, None, image, work)
        _lib.check(_lib.lib().ptir_resolve_backward(C.byref(r), gd.data_ptr(), gn.data_ptr(),
                                                    *(g.data_ptr() for g in grads), C.c_void_p(stream)), None)
        return (None, *grads)


def filter_iterations(filter_type) -> int:
    """Iterations of the a-trous filter for the render.filter_type option ('none' or 'atrous')."""
    kind = "none" if filter_type is None else str(filter_type).lower()
    if kind == "none":
        return 0
    if kind == "atrous":
        return 1
    raise ValueError(f"render.filter_type {filter_type!r} is not one of 'none', 'atrous'")


def apply_gt_mask(outputs: dict, batch) -> dict:
    """Zero the material-like outputs outside the batch's foreground mask (a [H,W],
    [H,W,1], [B,H,W] or [B,H,W,1] tensor in [0,1]); no-op without a mask or when done."""
    mask = getattr(batch, "mask", None)
    if mask is None or outputs.get(GT_MASK_APPLIED, False):
        return outputs
    out = dict(outputs)
    changed = False
    for key in GT_MASKED_OUTPUTS:
        img = outputs.get(key)
        if img is None or img.dim() != 4:
            continue
        m = mask.detach().to(device=img.device, dtype=img.dtype).clamp(0.0, 1.0)
        if m.dim() == 2:
            m = m[None, ..., None]
        elif m.dim() == 3:
            m = m[None] if (m.shape[-1] == 1 and tuple(m.shape[:2]) == tuple(img.shape[1:3])) else m[..., None]
        elif m.dim() != 4:
            raise ValueError(f"a ground-truth mask has 2 to 4 dimensions, this one has {m.dim()}")
        m = m[..., :1]
        if m.shape[0] == 1 and img.shape[0] != 1:
            m = m.expand(img.shape[0], -1, -1, -1)
        if tuple(m.shape[:3]) != tuple(img.shape[:3]):
            raise ValueError(f"mask {tuple(m.shape)} does not cover output {tuple(img.shape)}")
        out[key] = img * m
        changed = True
    if changed:
        out[GT_MASK_APPLIED] = True
    return out


@dataclass
class SppPlan:
    """How one render call spends its samples: `spp` samples per pixel traced in chunks of
    at most `chunk` samples; `positions` is None when the batch's own rays are used."""

    spp: int
    chunk: int
    positions: torch.Tensor | None
    intrinsics: torch.Tensor | None

    def chunks(self):
        for start in range(0, self.spp, self.chunk):
            yield start, min(self.chunk, self.spp - start)
