"""Environment importance tables consumed by the tracer.

Two layouts of the ``[3, h, w]`` float32 table the tracer reads (``ptir_scene.tab_kind``):

* **alias** -- the reference's: the caller area-resamples the environment to 64x128,
  weights each cell by Rec.709 luminance times its equirect solid angle and builds a
  Vose alias table on the host (threedgrut/model/aliastable.py:49-116,244-297; packed by
  threedgptir_tracer/tracer.py:819-847 as (prob, alias-as-float, pdf)).
  :func:`build_environment_alias_table` restates that build so the tracer can be fed
  exactly the cells the reference would sample.
* **cdf** -- built on the GPU by :func:`build_cdf_table` (``ptir_env_table_build``):
  the same weights and per-steradian pdf, inverted through a marginal CDF over rows and a
  conditional CDF per row, with FP64 prefix sums.  No host round trip, so it can be
  rebuilt every training step (the reference rebuilds the alias table every 32 steps).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

TABLE_GRID = (64, 128)  # rows (elevation) x columns (azimuth), aliastable.py:24
REC709 = (0.2126, 0.7152, 0.0722)


@dataclass(frozen=True)
class AliasTable:
    """Per-cell arrays of an alias table over an h x w equirect grid (row-major)."""

    prob: torch.Tensor   # [h*w] float32 acceptance probability of the cell itself
    alias: torch.Tensor  # [h*w] int32 cell taken on rejection
    pdf: torch.Tensor    # [h*w] float32 density per steradian
    height: int
    width: int

    @property
    def num_cells(self) -> int:
        return self.height * self.width


def vose_alias(weights) -> tuple[torch.Tensor, torch.Tensor]:
    """(prob, alias) of Vose's method over non-negative weights, in float64 arithmetic.

    The under-full and over-full work lists are consumed from their ends, which fixes the
    (prob, alias) pairing to the one aliastable.py:49-116 produces; the sampled cells of the
    tracer depend on that pairing."""
    w = torch.as_tensor(weights).detach().reshape(-1).to("cpu", torch.float64).numpy()
    if w.size == 0:
        raise ValueError("an alias table needs at least one weight")
    w = np.where(np.isfinite(w) & (w > 0), w, 0.0)
    n = w.size
    total = w.sum()
    if total <= 0.0:
        return torch.ones(n, dtype=torch.float32), torch.arange(n, dtype=torch.int32)
    level = w * (n / total)  # mean 1
    prob = np.ones(n, dtype=np.float32)
    alias = np.arange(n, dtype=np.int64)
    under = list(np.flatnonzero(level < 1.0))
    over = list(np.flatnonzero(level >= 1.0))
    while under and over:
        lo, hi = under.pop(), over.pop()
        prob[lo] = np.float32(min(max(level[lo], 0.0), 1.0))
        alias[lo] = hi
        level[hi] -= 1.0 - level[lo]
        if level[hi] < 1.0:
            under.append(hi)
        else:
            over.append(hi)
    # whatever is left holds (up to rounding) exactly one unit of mass: keep it
    return torch.from_numpy(prob), torch.from_numpy(alias.astype(np.int32))


def cell_solid_angles(h: int, w: int, dtype=torch.float32) -> torch.Tensor:
    """[h, w] solid angle of each equirect cell: (2 pi / w)(sin(lat_top) - sin(lat_bottom))."""
    lat = (torch.arange(h + 1, dtype=dtype) / h - 0.5) * torch.pi
    band = (2.0 * torch.pi / w) * torch.diff(torch.sin(lat))
    return band.clamp_min(0.0)[:, None].expand(h, w).contiguous()


def _cell_weights(environment: torch.Tensor, grid) -> tuple[torch.Tensor, torch.Tensor]:
    env = torch.as_tensor(environment).detach().cpu()
    if env.dim() != 3 or env.shape[-1] < 3:
        raise ValueError(f"environment must be [H, W, C] with C >= 3, not {tuple(env.shape)}")
    h, w = int(grid[0]), int(grid[1])
    if tuple(env.shape[:2]) != (h, w):
        env = F.interpolate(env.float().permute(2, 0, 1)[None], size=(h, w), mode="area")[0].permute(1, 2, 0)
    rgb = env[..., :3].float()
    rgb = torch.where(torch.isfinite(rgb) & (rgb > 0), rgb, 0.0)
    lum = (rgb * rgb.new_tensor(REC709)).sum(-1)
    sa = cell_solid_angles(h, w, lum.dtype)
    return (lum * sa).contiguous(), sa


def build_environment_alias_table(environment: torch.Tensor, target_size=TABLE_GRID) -> AliasTable:
    """Alias table of the reference's caller-side build (aliastable.py:244-297)."""
    wts, sa = _cell_weights(environment, target_size)
    total = wts.sum()
    if not bool(torch.isfinite(total)) or float(total) <= 0.0:
        wts, total = sa, sa.sum()  # no usable light: sample by solid angle
    pdf = torch.where(sa > 0, wts / torch.clamp(total * sa, min=torch.finfo(wts.dtype).tiny), 0.0)
    prob, alias = vose_alias(wts)
    h, w = wts.shape
    return AliasTable(prob=prob, alias=alias, pdf=pdf.reshape(-1).float().contiguous(), height=h, width=w)


def pack_alias_table(table: AliasTable | None, device="cpu") -> torch.Tensor:
    """The tracer's [3, h, w] alias layout (prob, alias as float, pdf), or an empty tensor."""
    if table is None:
        return torch.empty(0, dtype=torch.float32, device=device)
    planes = [table.prob, table.alias.to(torch.float32), table.pdf]
    return torch.stack([p.reshape(table.height, table.width).float() for p in planes]).to(device).contiguous()


def build_cdf_table(environment: torch.Tensor, grid=TABLE_GRID, out: torch.Tensor | None = None,
                    allocate: bool = True) -> torch.Tensor:
    """[3, h, w] CDF-layout table of a device environment [H, W, 4], built on the GPU on the
    current stream (ptir_env_table_build).  allocate=False uses the allocation-free entry
    point (tables up to 131072 cells), for code under graph capture."""
    from . import _lib

    env = environment
    if not (isinstance(env, torch.Tensor) and env.is_cuda and env.dtype == torch.float32):
        raise TypeError("environment must be a float32 device tensor")
    if env.dim() != 3 or env.shape[-1] != 4:
        raise ValueError(f"environment must be [H, W, 4], not {tuple(env.shape)}")
    env = env.contiguous()
    h, w = int(grid[0]), int(grid[1])
    if out is None:
        out = torch.empty(3, h, w, dtype=torch.float32, device=env.device)
    elif tuple(out.shape) != (3, h, w) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 [3, h, w] tensor")
    stream = torch.cuda.current_stream(env.device).cuda_stream
    fn = _lib.lib().ptir_env_table_build if allocate else _lib.lib().ptir_env_table_build_noalloc
    _lib.check(fn(env.data_ptr(), env.shape[0], env.shape[1], h, w, out.data_ptr(), C.c_void_p(stream)), None)
    return out
