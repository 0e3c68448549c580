"""GPU exact adjoint (ptir_trace_bwd_exact) against the FP64 dual-number oracle (the identity
and tolerance of tests/adjoint_cases.py), on pixels whose paths take the same decisions on the
GPU and in the oracle.

Skipped: the kernel faulted the device on its first run and the cause is not found, so the
library switches its launch off (PTIR_ERR_UNSUPPORTED).  Its device code is checked on the host
by tests/test_exact_adjoint_host.py.  test_exact_backward_reports_switched_off checks that the
entry point refuses cleanly instead of launching.
"""

import numpy as np
import pytest
import torch

import adjoint_cases as A
from gpu_helpers import gpu_forward, to_np
from ptir_amd._lib import PtirError
from ptir_amd.tracer import HipTracer

pytestmark = pytest.mark.gpu


def _agree(gpu, ref):
    """Pixels whose outputs agree between the GPU and the oracle (the same path)."""
    ok = None
    for k in ("pbr", "light", "rad", "dns", "mat", "comp"):
        a = gpu[k].reshape(*gpu[k].shape[:3], -1)
        b = ref[k].reshape(a.shape)
        m = (np.abs(a - b) <= 2e-4 + 2e-3 * np.abs(b)).all(-1)
        ok = m if ok is None else ok & m
    return ok


def _exact(tr, sc, args, cfg, g, device):
    B, H, W = sc.rays_dir.shape[:3]
    kw = {A.OUTS[k][0]: torch.from_numpy(v).to(device) for k, v in g.items()}
    kw["ray_comp_grd"] = kw["ray_comp_grd"].reshape(B, H, W, 2, 3)
    frame, T, ro, rd = args[:4]
    return tr.trace_bwd_exact(frame, T, ro, rd, *args[4:10], **kw, render_opts=cfg["render_opts"],
                              sph_degree=sc.sph_degree, min_transmittance=sc.min_transmittance,
                              max_bounces=sc.max_bounces, rr_start=cfg["rr_start"])


def test_exact_backward_reports_switched_off(device):
    sc, tab, inp, rays, ref, rng = A.setup("no_mis_mb3")
    tr = HipTracer(enable_metallic=True, enable_mis=False, gradients="exact")
    outs, args = gpu_forward(tr, sc, tab, device)
    g = A.output_grads(np.ones(sc.rays_dir.shape[:3], bool), rng)
    with pytest.raises(PtirError, match="switched off"):
        _exact(tr, sc, args, A.CASES["no_mis_mb3"], g, device)


@pytest.mark.skip(reason="the exact-adjoint kernel faulted the device on its first run and the cause is not "
                         "found; its launch is switched off and its device code is verified on the host "
                         "(tests/test_exact_adjoint_host.py)")
@pytest.mark.parametrize("case", list(A.CASES))
def test_exact_adjoint_matches_dual_derivatives(device, case):
    cfg = A.CASES[case]
    sc, tab, inp, rays, ref, rng = A.setup(case)
    tr = HipTracer(enable_metallic=True, enable_mis=cfg["enable_mis"], gradients="exact")
    print(f"[{case}] forward", flush=True)
    outs, args = gpu_forward(tr, sc, tab, device, render_opts=cfg["render_opts"])
    if cfg["rr_start"]:
        outs = tr.trace(*args, render_opts=cfg["render_opts"], sph_degree=sc.sph_degree,
                        min_transmittance=sc.min_transmittance, max_bounces=sc.max_bounces, rr_start=cfg["rr_start"])
    gpu = to_np(outs)
    robust = ref["margin"] >= A.MARGIN
    agree = _agree(gpu, ref)
    print(f"[{case}] margin >= {A.MARGIN}: {robust.mean():.3f}; outputs agree there: {agree[robust].mean():.4f}",
          flush=True)
    assert robust.mean() > 0.7 and agree[robust].mean() >= 0.995
    g = A.output_grads(robust & agree, rng)
    d_dns, d_mat, d_sh, d_sn, d_env = _exact(tr, sc, args, cfg, g, device)
    torch.cuda.synchronize()
    grads = {"density": d_dns.double().cpu().numpy(), "material": d_mat.double().cpu().numpy(),
             "sh": d_sh.double().cpu().numpy(), "snormal": d_sn.double().cpu().numpy(),
             "env": d_env.double().cpu().numpy()}
    for k, v in grads.items():
        assert np.isfinite(v).all(), k
    A.check_blocks(sc, inp, rays, g, grads, rng, case)
