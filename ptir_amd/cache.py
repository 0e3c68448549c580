"""Neural radiance cache over the C ABI (no counterpart in the reference, which has none).

`RadianceCache` owns a cache handle of libptir_amd.so: a multiresolution hash grid and a
64-wide MLP that run on bf16 MFMA kernels.  A trace with ``cache_mode=2`` writes training
records (query at the first secondary vertex plus the radiance the rest of the path carried);
`train` fits the cache to them.  A trace with ``cache_mode=1`` ends every path at its first
secondary vertex and writes a query; `query` adds throughput x answer to the traced pbr
and to the indirect component, which is then the radiance-cache estimate of the pixel.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib

REC_FLOATS = 20


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dev(t: torch.Tensor, name: str, tail=None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise TypeError(f"{name} must be a float32 device tensor")
    if tail is not None and tuple(t.shape[-len(tail):]) != tuple(tail):
        raise ValueError(f"{name} must have trailing shape {tuple(tail)}, got {tuple(t.shape)}")
    return t.contiguous()


class RadianceCache:
    """Hash grid (16 levels x 2 features, 2^19 entries per level) + MLP (4 x 64 ReLU, exp out)."""

    def __init__(self, max_batch: int = 65536, seed: int = 0, learning_rate: float = 1e-2, beta1: float = 0.9,
                 beta2: float = 0.99, epsilon: float = 1e-10):
        L = _lib.lib()
        o = _lib.CacheOptions(int(max_batch), int(seed) & 0xFFFFFFFF, float(learning_rate), float(beta1), float(beta2),
                              float(epsilon))
        h = C.c_void_p()
        _lib.check_cache(L.ptir_cache_create(C.byref(o), C.byref(h), _stream()), None)
        self._h = h
        self.max_batch = int(max_batch)
        self.num_weights = int(L.ptir_cache_num_weights())
        self.table_floats = int(L.ptir_cache_table_floats())

    def close(self) -> None:
        h = getattr(self, "_h", None)
        self._h = None
        if h is not None and h.value:
            try:
                _lib.lib().ptir_cache_destroy(h)
            except Exception:  # interpreter shutdown
                pass

    def __del__(self):
        self.close()

    # -------------------------------------------------------------- parameters
    def read_params(self, device="cuda") -> tuple[torch.Tensor, torch.Tensor]:
        w = torch.empty(self.num_weights, dtype=torch.float32, device=device)
        t = torch.empty(self.table_floats, dtype=torch.float32, device=device)
        _lib.check_cache(_lib.lib().ptir_cache_read_params(self._h, w.data_ptr(), t.data_ptr(), _stream()), self._h)
        return w, t.view(16, -1, 2)

    def write_params(self, weights: torch.Tensor | None = None, table: torch.Tensor | None = None) -> None:
        w = None if weights is None else _dev(weights, "weights").reshape(-1)
        t = None if table is None else _dev(table, "table").reshape(-1)
        if w is not None and w.numel() != self.num_weights:
            raise ValueError(f"weights must hold {self.num_weights} floats")
        if t is not None and t.numel() != self.table_floats:
            raise ValueError(f"table must hold {self.table_floats} floats")
        _lib.check_cache(_lib.lib().ptir_cache_write_params(self._h, None if w is None else w.data_ptr(),
                                                            None if t is None else t.data_ptr(), _stream()), self._h)

    def read_grads(self, device="cuda") -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        w = torch.empty(self.num_weights, dtype=torch.float32, device=device)
        t = torch.empty(self.table_floats, dtype=torch.float32, device=device)
        loss = torch.empty(1, dtype=torch.float32, device=device)
        _lib.check_cache(_lib.lib().ptir_cache_read_grads(self._h, w.data_ptr(), t.data_ptr(), loss.data_ptr(),
                                                          _stream()), self._h)
        return w, t.view(16, -1, 2), loss

    # -------------------------------------------------------------- use
    def query(self, records: torch.Tensor, scene_box: torch.Tensor, pbr: torch.Tensor | None = None,
              pbr_components: torch.Tensor | None = None, prediction: bool = False) -> torch.Tensor | None:
        """Add the cache term of every record with flag 1 into pbr / pbr_components (in place);
        return the [n,3] answers when `prediction` is set."""
        rec = _dev(records, "records", (REC_FLOATS,))
        n = rec.numel() // REC_FLOATS
        box = _dev(scene_box, "scene_box").reshape(-1)
        if box.numel() != 6:
            raise ValueError("scene_box must hold 6 floats (lo.xyz, hi.xyz)")
        for t, name, c in ((pbr, "pbr", 3), (pbr_components, "pbr_components", 6)):
            if t is not None and (not t.is_contiguous() or t.numel() != n * c or t.dtype != torch.float32):
                raise ValueError(f"{name} must be a contiguous float32 tensor of {n}x{c} values")
        pred = torch.zeros(n, 3, dtype=torch.float32, device=rec.device) if prediction else None
        _lib.check_cache(_lib.lib().ptir_cache_query(
            self._h, rec.data_ptr(), n, box.data_ptr(), None if pbr is None else pbr.data_ptr(),
            None if pbr_components is None else pbr_components.data_ptr(), None if pred is None else pred.data_ptr(),
            _stream()), self._h)
        return pred

    def backprop(self, records: torch.Tensor, scene_box: torch.Tensor) -> None:
        rec = _dev(records, "records", (REC_FLOATS,))
        box = _dev(scene_box, "scene_box").reshape(-1)
        _lib.check_cache(_lib.lib().ptir_cache_backprop(self._h, rec.data_ptr(), rec.numel() // REC_FLOATS,
                                                        box.data_ptr(), _stream()), self._h)

    def step(self) -> None:
        _lib.check_cache(_lib.lib().ptir_cache_step(self._h, _stream()), self._h)

    def train(self, records: torch.Tensor, scene_box: torch.Tensor) -> None:
        """One online training step (loss, gradients, Adam) on a batch of training records."""
        rec = _dev(records, "records", (REC_FLOATS,))
        box = _dev(scene_box, "scene_box").reshape(-1)
        _lib.check_cache(_lib.lib().ptir_cache_train(self._h, rec.data_ptr(), rec.numel() // REC_FLOATS,
                                                     box.data_ptr(), _stream()), self._h)


class CachedPathTracer:
    """The radiance-cache form of the path tracer (no counterpart in the reference).

    Rendering traces every path to its first secondary vertex (cache_mode 1) and adds the
    cache's answer there, weighted by the path throughput.  Training traces a subset of the
    pixels as full paths (cache_mode 2, up to max_bounces vertices), keeps the records of
    the paths that reached the cache vertex (each holds the query and the radiance the rest
    of its path carried) and takes one Adam step on up to `train_batch` of them.  The plain
    multi-bounce tracer remains the reference the cache is judged against."""

    def __init__(self, tracer, cache: RadianceCache, scene_box: torch.Tensor, max_bounces: int = 4,
                 train_batch: int = 65536, train_stride: int = 4):
        if max_bounces < 3:
            raise ValueError("the cache stands in for bounces beyond the first: max_bounces must be >= 3")
        if train_batch > cache.max_batch:
            raise ValueError("train_batch exceeds the cache's max_batch")
        self.tracer = tracer
        self.cache = cache
        self.box = scene_box.contiguous()
        self.max_bounces = int(max_bounces)
        self.train_batch = int(train_batch)
        self.train_stride = int(train_stride)
        self.batch = None
        self.train_paths = 0
        self.train_samples = 0

    def _records(self, ray_dir):
        B, H, W = ray_dir.shape[:3]
        return torch.zeros((B, H, W, REC_FLOATS), dtype=torch.float32, device=ray_dir.device)

    def render(self, frame, ray_to_world, ray_ori, ray_dir, scene_args, **kw):
        """Forward outputs of the cache form (the 13 tracer outputs; pbr and the indirect part of
        pbr_components include the cache term) and the cache records of the launch."""
        rec = self._records(ray_dir)
        outs = self.tracer.trace(frame, ray_to_world, ray_ori, ray_dir, *scene_args, max_bounces=self.max_bounces,
                                 cache_mode=1, cache_records=rec, **kw)
        self.cache.query(rec.view(-1, REC_FLOATS), self.box, pbr=outs[10], pbr_components=outs[12])
        return outs, rec

    def train(self, frame, ray_to_world, ray_ori, ray_dir, scene_args, offset: int = 0, **kw) -> int:
        """One online training step from full paths of every train_stride-th pixel row
        (starting at offset).  Returns the number of training samples used."""
        o = ray_ori[:, offset::self.train_stride].contiguous()
        d = ray_dir[:, offset::self.train_stride].contiguous()
        rec = self._records(d)
        self.tracer.trace(frame, ray_to_world, o, d, *scene_args, max_bounces=self.max_bounces, cache_mode=2,
                          cache_records=rec, **kw)
        flat = rec.view(-1, REC_FLOATS)
        live = flat[flat[:, 0] != 0.0][: self.train_batch].contiguous()
        self.train_paths += flat.shape[0]
        n = live.shape[0]
        self.train_samples += n
        if n:
            self.cache.train(live, self.box)
        return n
