"""DeviceAdam (ptir_adam_step) against torch.optim.Adam and a torch restatement of the
reference's SelectiveAdam (threedgrut/optimizers/optimizers.cu:49-80)."""

import pytest
import torch

from ptir_amd import _lib
from ptir_amd.optim import DeviceAdam




@pytest.mark.gpu
def test_matches_torch_adam_over_several_steps(device):
    gen = torch.Generator().manual_seed(0)
    shapes = [(100, 3), (100, 1), (16, 32, 4)]
    lrs = [0.035, 0.025, 0.01]
    ours = [torch.rand(s, generator=gen).to(device) for s in shapes]
    ref = [t.clone().requires_grad_(True) for t in ours]
    opt = DeviceAdam(list(zip(ours, lrs)), eps=1e-15)
    topt = torch.optim.Adam([{"params": [r], "lr": lr} for r, lr in zip(ref, lrs)], eps=1e-15)
    for _ in range(5):
        grads = [torch.randn(s, generator=gen).to(device) for s in shapes]
        for o, r, g in zip(ours, ref, grads):
            o.grad = g.clone()
            r.grad = g.clone()
        opt.step()
        topt.step()
    torch.cuda.synchronize()
    assert int(opt.step_count.item()) == 5
    for o, r in zip(ours, ref):
        torch.testing.assert_close(o, r.detach(), rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_selective_adam_updates_only_visible_rows(device):
    gen = torch.Generator().manual_seed(1)
    p = torch.rand(64, 3, generator=gen).to(device)
    p0 = p.clone()
    vis = torch.zeros(64, dtype=torch.int32, device=device)
    vis[::3] = 1
    opt = DeviceAdam([(p, 0.1)], betas=(0.9, 0.999), eps=1e-15, selective=True)
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    ref = p0.clone()
    for _ in range(3):
        g = torch.randn(64, 3, generator=gen).to(device)
        p.grad = g
        opt.step(visibility=vis.view(torch.float32))
        sel = vis.bool()[:, None]
        # the kernel's coefficients are float32 numbers: b, and 1 - b rounded to float32
        b1, b2 = torch.tensor(0.9), torch.tensor(0.999)
        m = torch.where(sel, b1 * m + (1 - b1) * g, m)
        v = torch.where(sel, b2 * v + (1 - b2) * g * g, v)
        ref = torch.where(sel, ref - 0.1 * m / (v.sqrt() + 1e-15), ref)
    torch.cuda.synchronize()
    torch.testing.assert_close(p, ref, rtol=1e-5, atol=1e-6)
    assert torch.equal(p[1], p0[1])


def test_rejects_bad_arguments():
    L = _lib.lib()
    assert L.ptir_adam_step(None, 1, 0.9, 0.999, 1e-8, 1, None, None) == 1
    assert L.ptir_adam_step(None, 0, 1.5, 0.999, 1e-8, 1, None, None) == 1
