"""Environment importance tables consumed by the tracer.

`build_environment_alias_table` restates the reference's caller-side table build
(threedgrut/model/aliastable.py:49-116,244-297): the env is area-resampled to 64x128,
weighted by Rec.709 luminance times the equirect texel solid angle, and turned into a
Vose alias table (FP64 weights, FP32 storage) plus the per-steradian pdf
``w / (sum(w) * solid_angle)``.  `pack_alias_table` packs it as the ``[3, h, w]``
(prob, alias-as-float, pdf) tensor the tracer reads (threedgptir_tracer/tracer.py:819-847).

The table is a caller-side input of the tracer, built every 32 training steps in the
reference; it is not on the per-path hot loop.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

DEFAULT_ALIAS_TABLE_SIZE = (64, 128)
LUMINANCE = (0.2126, 0.7152, 0.0722)


@dataclass(frozen=True)
class EnvAliasTable:
    width: int
    height: int
    numCells: int
    prob: torch.Tensor
    alias: torch.Tensor
    pdf: torch.Tensor


def build_alias_table(weights) -> tuple[torch.Tensor, torch.Tensor]:
    """Vose alias table from non-negative weights (aliastable.py:49-116).

    Small/large work lists are popped from the back, exactly as the reference does, so
    the (prob, alias) pairing is the reference's."""
    w = torch.as_tensor(weights).detach().reshape(-1).to("cpu", torch.float64)
    w = torch.where(torch.isfinite(w) & (w > 0), w, torch.zeros_like(w))
    n = w.numel()
    if n == 0:
        raise ValueError("weights must contain at least one element.")
    total = float(w.sum())
    if total <= 0.0:
        return torch.ones(n, dtype=torch.float32), torch.arange(n, dtype=torch.int32)
    scaled = w.numpy() * (float(n) / total)
    prob = np.empty(n, dtype=np.float32)
    alias = np.arange(n, dtype=np.int64)
    small = [int(i) for i in np.nonzero(scaled < 1.0)[0]]
    large = [int(i) for i in np.nonzero(scaled >= 1.0)[0]]
    while small and large:
        s = small.pop()
        l = large.pop()
        prob[s] = np.float32(np.clip(scaled[s], 0.0, 1.0))
        alias[s] = l
        scaled[l] -= 1.0 - scaled[s]
        (small if scaled[l] < 1.0 else large).append(l)
    for i in small + large:
        prob[i] = 1.0
        alias[i] = i
    return torch.from_numpy(prob), torch.from_numpy(alias.astype(np.int32))


def equirect_solid_angles(h: int, w: int, dtype=torch.float32) -> torch.Tensor:
    """Solid angle of each equirect texel (aliastable.py:163-176)."""
    edges = (torch.arange(h + 1, dtype=dtype) / float(h) - 0.5) * torch.pi
    rows = (2.0 * torch.pi / float(w)) * (torch.sin(edges[1:]) - torch.sin(edges[:-1]))
    return torch.clamp(rows, min=0.0).reshape(h, 1).expand(h, w).contiguous()


def build_environment_alias_table(environment: torch.Tensor, target_size=DEFAULT_ALIAS_TABLE_SIZE) -> EnvAliasTable:
    env = torch.as_tensor(environment).detach().cpu()
    if env.ndim != 3 or env.shape[-1] < 3:
        raise ValueError(f"Environment must have shape [H, W, C>=3], got {tuple(env.shape)}")
    th, tw = int(target_size[0]), int(target_size[1])
    if tuple(env.shape[:2]) != (th, tw):
        env = F.interpolate(env.float().permute(2, 0, 1).unsqueeze(0), size=(th, tw), mode="area")
        env = env.squeeze(0).permute(1, 2, 0).contiguous()
    rgb = env[..., :3].float()
    rgb = torch.where(torch.isfinite(rgb) & (rgb > 0), rgb, torch.zeros_like(rgb))
    lum = torch.sum(rgb * rgb.new_tensor(LUMINANCE), dim=-1)
    h, w = lum.shape
    sa = equirect_solid_angles(h, w, lum.dtype)
    sw = (lum * sa).contiguous()
    total = sw.sum()
    if not bool(torch.isfinite(total)) or float(total) <= 0.0:
        sw = sa
        total = sw.sum()
    pdf = torch.where(sa > 0, sw / torch.clamp(total * sa, min=torch.finfo(sw.dtype).tiny), torch.zeros_like(sw))
    prob, alias = build_alias_table(sw)
    return EnvAliasTable(width=w, height=h, numCells=h * w, prob=prob, alias=alias,
                         pdf=pdf.reshape(-1).float().contiguous())


def pack_alias_table(table: EnvAliasTable | None, device="cpu") -> torch.Tensor:
    """[3, h, w] float32 = (prob, alias, pdf), or an empty tensor (tracer.py:824-847)."""
    if table is None:
        return torch.empty(0, dtype=torch.float32, device=device)
    return torch.cat([
        table.prob.reshape(1, table.height, table.width),
        table.alias.reshape(1, table.height, table.width).to(torch.float32),
        table.pdf.reshape(1, table.height, table.width),
    ], dim=0).to(device).contiguous()
