"""Adam on the device for the tensors a PTIR inversion optimises.

`DeviceAdam` applies one ``ptir_adam_step`` per call: torch.optim.Adam semantics
(bias-corrected), or the reference's SelectiveAdam (``selective=True``: rows of invisible
Gaussians untouched, no bias correction; threedgrut/optimizers/optimizers.cu:49-80).  Its
step count is a device integer, so a training step that ends with it can be captured in a
CUDA/HIP graph and replayed (torch.optim.Adam keeps the count on the host).
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib


class DeviceAdam:
    def __init__(self, params, betas=(0.9, 0.999), eps=1e-15, selective=False):
        """params: iterable of (tensor, lr).  Tensors are updated in place from `.grad`;
        per-Gaussian tensors are [N, C] (visibility selects rows), others any shape."""
        self.groups = [(p, float(lr)) for p, lr in params]
        for p, _ in self.groups:
            if not (p.is_cuda and p.dtype == torch.float32 and p.is_contiguous()):
                raise ValueError("DeviceAdam optimises contiguous float32 device tensors")
        self.betas = (float(betas[0]), float(betas[1]))
        self.eps = float(eps)
        self.selective = bool(selective)
        dev = self.groups[0][0].device if self.groups else torch.device("cuda")
        self.state = [(torch.zeros_like(p), torch.zeros_like(p)) for p, _ in self.groups]
        self.step_count = torch.zeros(1, dtype=torch.int32, device=dev)

    @torch.no_grad()
    def step(self, visibility: torch.Tensor | None = None) -> None:
        """One update.  visibility: the tracer's [N] visibility (nonzero words = visible),
        applied to tensors whose first dimension is N when the optimiser is selective."""
        arr = (_lib.AdamGroup * max(1, len(self.groups)))()
        keep = []
        n = 0
        for (p, lr), (m, v) in zip(self.groups, self.state):
            if p.grad is None:
                continue
            g = p.grad.contiguous()
            keep.append(g)
            rows = p.shape[0] if p.dim() > 0 else 1
            cols = p.numel() // max(rows, 1)
            vis = None
            if self.selective and visibility is not None and p.dim() > 0 and visibility.numel() == rows:
                vis = visibility.contiguous().view(torch.int32)
                keep.append(vis)
            arr[n] = _lib.AdamGroup(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(),
                                    None if vis is None else vis.data_ptr(), rows, cols, lr)
            n += 1
        stream = torch.cuda.current_stream(self.step_count.device).cuda_stream
        _lib.check(_lib.lib().ptir_adam_step(arr, n, self.betas[0], self.betas[1], self.eps, int(not self.selective),
                                             self.step_count.data_ptr(), C.c_void_p(stream)), None)

    def zero_grad(self) -> None:
        for p, _ in self.groups:
            if p.grad is not None:
                p.grad.zero_()
