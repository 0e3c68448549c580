"""Environment importance table in the CDF layout (ptir_env_table_build).

CPU: the oracle's float64 restatement (oracle.env_cdf_table) is pinned against
torch's area resampling, the alias build's pdf (same weights, same formula), the
normalisation of the pdf over the sphere, and a chi-square test of CDF-inversion sampling.
GPU: the device build equals the restatement, and a trace in CDF mode equals the oracle's.
"""

import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from gpu_helpers import gpu_forward, oracle_forward, pixel_agreement, to_np
from oracle import oracle as O
from ptir_amd import envmap, scenes
from ptir_amd.tracer import HipTracer


def _env(h, w, seed, sun=False):
    g = torch.Generator().manual_seed(seed)
    if sun:
        return scenes.sun_sky_env(h, w, g, sun_gain=200.0)[0]
    return scenes.lognormal_env(h, w, g)


# ----------------------------------------------------------------------------- CPU pins
@pytest.mark.parametrize("hw,grid", [((32, 64), (64, 128)), ((256, 512), (64, 128)), ((37, 91), (16, 24)),
                                     ((64, 128), (64, 128))])
def test_restated_resampling_equals_torch_area_interpolation(hw, grid):
    env = _env(*hw, seed=1)
    tab = O.env_cdf_table(env, *grid)
    ref = F.interpolate(env.double().permute(2, 0, 1)[None], size=grid, mode="area")[0].permute(1, 2, 0)
    rgb = ref[..., :3].clamp_min(0)
    lum = (rgb * torch.tensor([0.2126, 0.7152, 0.0722], dtype=torch.float64)).sum(-1)
    sa = envmap.cell_solid_angles(*grid, torch.float64)
    pdf = (lum * sa / (lum * sa).sum() / sa).numpy()
    np.testing.assert_allclose(tab[2], pdf, rtol=2e-6)


@pytest.mark.parametrize("seed,sun", [(0, False), (3, True)])
def test_cdf_table_has_the_alias_tables_pdf_and_normalised_cdfs(seed, sun):
    env = _env(128, 256, seed, sun)
    tab = O.env_cdf_table(env)
    alias = envmap.build_environment_alias_table(env)
    np.testing.assert_allclose(tab[2].reshape(-1), alias.pdf.numpy(), rtol=2e-5)
    sa = envmap.cell_solid_angles(64, 128, torch.float64).numpy()
    assert float((tab[2].astype(np.float64) * sa).sum()) == pytest.approx(1.0, rel=1e-6)
    assert np.all(np.diff(tab[0], axis=1) >= 0) and np.all(tab[0][:, -1] == 1.0)
    m = tab[1][:, 0]
    assert np.all(np.diff(m) >= 0) and m[-1] == 1.0
    assert np.all(tab[1][:, 1:] == 0)
    # row masses of the marginal CDF are the rows' shares of the pdf mass
    row_mass = (tab[2].astype(np.float64) * sa).sum(1)
    np.testing.assert_allclose(np.diff(np.concatenate([[0.0], m])), row_mass, rtol=1e-4, atol=1e-7)


def test_cdf_table_falls_back_to_solid_angle_without_light():
    tab = O.env_cdf_table(torch.zeros(16, 32, 4))
    np.testing.assert_allclose(tab[2], 1.0 / (4 * math.pi), rtol=1e-6)


def test_cdf_sampling_histogram_matches_pdf():
    env = _env(64, 128, 5, sun=True)
    tab = torch.from_numpy(O.env_cdf_table(env))
    d = torch.zeros(1, 12)
    d[0, 3], d[0, 4], d[0, 8:11] = 0.5, 1.0, 0.1
    inp = O.Inputs(d, torch.zeros(1, 5), torch.zeros(1, 48), torch.tensor([[0.0, 0.0, 1.0]]), env, tab, tab_kind=1)
    n = 200_000
    dirs, pdfs = O.env_sample(inp, 13, n)
    np.testing.assert_allclose([O.env_pdf(inp, v) for v in dirs[:2000]], pdfs[:2000], rtol=1e-9)
    e = np.stack([-dirs[:, 1], dirs[:, 0], dirs[:, 2]], -1)  # the env frame is rotated by pi/2 about z
    u = (np.arctan2(e[:, 0], e[:, 1]) + np.pi) / (2 * np.pi)
    v = 0.5 + np.arcsin(np.clip(-e[:, 2], -1, 1)) / np.pi
    cell = np.clip((v * 64).astype(int), 0, 63) * 128 + np.clip((u * 128).astype(int), 0, 127)
    hist = np.bincount(cell, minlength=64 * 128) / n
    p_cell = tab[2].double().numpy().reshape(-1) * envmap.cell_solid_angles(64, 128, torch.float64).reshape(-1).numpy()
    big = p_cell > 20.0 / n
    chi = ((hist[big] - p_cell[big]) ** 2 / p_cell[big]).sum() * n
    dof = big.sum() - 1
    assert chi < dof + 6 * math.sqrt(2 * dof)


def test_env_table_build_rejects_bad_arguments():
    from ptir_amd import _lib

    L = _lib.lib()
    assert L.ptir_env_table_build(None, 4, 8, 4, 8, None, None) == 1
    assert "null" in L.ptir_last_error(None).decode()
    assert L.ptir_env_table_scratch_bytes(64, 128) == 8 * (64 * 128 + 64 + 2)
    assert L.ptir_env_table_scratch_bytes(0, 128) == 0


# ----------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("hw,grid,sun", [((32, 64), (64, 128), False), ((256, 512), (64, 128), True),
                                         ((1024, 2048), (64, 128), True), ((37, 91), (16, 24), False),
                                         ((512, 1024), (256, 512), True)])
def test_device_cdf_table_equals_restatement(device, hw, grid, sun):
    env = _env(*hw, seed=7, sun=sun)
    print("device table", hw, grid, flush=True)
    got = envmap.build_cdf_table(env.to(device), grid).cpu().numpy()
    ref = O.env_cdf_table(env, *grid)
    # the CDFs come from FP64 prefix sums rounded once to float32: at most one ulp apart
    np.testing.assert_allclose(got[0], ref[0], rtol=0, atol=1.2e-7)
    np.testing.assert_allclose(got[1], ref[1], rtol=0, atol=1.2e-7)
    np.testing.assert_allclose(got[2], ref[2], rtol=1e-6, atol=0)
    assert np.all(got[0][:, -1] == 1.0) and got[1][-1, 0] == 1.0


@pytest.mark.gpu
def test_device_cdf_table_without_light(device):
    got = envmap.build_cdf_table(torch.zeros(16, 32, 4, device=device), (8, 16)).cpu().numpy()
    np.testing.assert_allclose(got[2], 1.0 / (4 * math.pi), rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("max_bounces", [2, 3])
def test_trace_with_cdf_table_matches_oracle(device, max_bounces):
    sc = scenes.small_scene(0, max_bounces=max_bounces)
    tab = envmap.build_cdf_table(sc.env.to(device))
    tr = HipTracer(enable_metallic=True, env_sampling="cdf")
    outs, _ = gpu_forward(tr, sc, tab, device)
    torch.cuda.synchronize()
    g = to_np(outs)
    _, _, ref = oracle_forward(sc, tab.cpu(), tab_kind=1)
    agree = {k: pixel_agreement(g[k], ref[k], 1e-3, 1e-5) for k in ("dns", "pbr", "light", "comp")}
    print("cdf-mode agreement", agree, flush=True)
    # one secondary ray: as tight as the alias path; two: more decisions per path can flip
    bar = 0.999 if max_bounces == 2 else 0.995
    for k, v in agree.items():
        assert v >= bar, k
    # every differing path is within float32 reach of a discrete decision
    ok = np.ones(ref["pbr"].shape[:3], bool)
    for k in ("dns", "pbr", "light", "comp"):
        x = g[k].reshape(*g[k].shape[:3], -1)
        y = ref[k].reshape(x.shape)
        ok &= (np.abs(x - y) <= 1e-5 + 1e-3 * np.abs(y)).all(-1)
    assert float(ref["margin"][~ok].max(initial=0.0)) < 1e-2
