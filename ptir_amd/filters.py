"""Output post-processing of the tracer's caller (threedgrut/model/filters.py:6-69 and
threedgrut/model/ptir_helper.py:100-122): an optional 3x3 B-spline a-trous filter on the
PBR outputs and the ground-truth foreground mask."""

from __future__ import annotations

import torch
import torch.nn as nn
import torch.nn.functional as F

PBR_GT_MASK_APPLIED_KEY = "_pbr_gt_mask_applied"
# pred_pbr/direct/indirect stay unmasked (ptir_helper.py:49-62)
PBR_GT_MASK_OUTPUT_KEYS = ("pred_light", "pred_shadingnormal", "pred_material")


class Filter(nn.Module):
    """'none' or 'atrous' (B-spline [1,2,1]^T[1,2,1]/16, dilation 2^i, replicate padding)."""

    def __init__(self, filter_type: str | None = "none", iterations: int = 1, padding_mode: str = "replicate"):
        super().__init__()
        if iterations < 0:
            raise ValueError(f"iterations must be non-negative, got {iterations}")
        self.filter_type = "none" if filter_type is None else str(filter_type).lower()
        if self.filter_type not in ("none", "atrous"):
            raise ValueError(f"Unknown render.filter_type: {self.filter_type}")
        self.iterations = int(iterations)
        self.padding_mode = padding_mode
        k = torch.tensor([[1, 2, 1], [2, 4, 2], [1, 2, 1]], dtype=torch.float32)
        self.register_buffer("kernel", (k / k.sum()).view(1, 1, 3, 3), persistent=False)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if self.filter_type == "none" or self.iterations == 0:
            return x
        if x.ndim != 4:
            raise ValueError(f"Filter expects [B, H, W, C], got {tuple(x.shape)}")
        y = x.permute(0, 3, 1, 2).contiguous()
        kern = self.kernel.to(device=y.device, dtype=y.dtype).repeat(y.shape[1], 1, 1, 1)
        for i in range(self.iterations):
            d = 2 ** i
            y = F.pad(y, (d, d, d, d), mode=self.padding_mode)
            y = F.conv2d(y, kern, dilation=d, groups=y.shape[1])
        return y.permute(0, 2, 3, 1).contiguous()


def apply_gt_mask_to_tensor(tensor, mask):
    """Multiply an NHWC image by a 2D/3D/4D foreground mask clamped to [0,1]."""
    if tensor is None or mask is None or tensor.ndim != 4:
        return tensor
    m = mask.detach().to(device=tensor.device, dtype=tensor.dtype).clamp(0.0, 1.0)
    if m.ndim == 2:
        m = m[None, :, :, None]
    elif m.ndim == 3:
        m = m.unsqueeze(0) if (m.shape[-1] == 1 and m.shape[:2] == tensor.shape[1:3]) else m.unsqueeze(-1)
    elif m.ndim != 4:
        raise ValueError(f"GT mask must be 2D, 3D, or 4D, got {m.ndim}D")
    if m.shape[-1] != 1:
        m = m[..., :1]
    if m.shape[0] == 1 and tensor.shape[0] != 1:
        m = m.expand(tensor.shape[0], -1, -1, -1)
    if m.shape[:3] != tensor.shape[:3]:
        raise ValueError(f"GT mask shape {tuple(m.shape)} is not compatible with {tuple(tensor.shape)}")
    return tensor * m


def post_processing(outputs: dict, gpu_batch, visualize_lights: bool = False) -> dict:
    """Multiply PBR image outputs by the batch foreground mask when the batch has one."""
    if outputs.get(PBR_GT_MASK_APPLIED_KEY):
        return outputs
    mask = getattr(gpu_batch, "mask", None)
    if mask is None:
        return outputs
    out = dict(outputs)
    applied = False
    for key in PBR_GT_MASK_OUTPUT_KEYS:
        if visualize_lights and key in ("pred_pbr", "pred_direct"):
            continue
        v = outputs.get(key)
        if v is None:
            continue
        out[key] = apply_gt_mask_to_tensor(v, mask)
        applied = True
    if applied:
        out[PBR_GT_MASK_APPLIED_KEY] = True
    return out
