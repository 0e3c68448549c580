"""The 3DGRT primary-only tracer's device code (ptir_amd/csrc/grt_path.hpp) compiled for the
host with g++ and run one lane at a time (tests/host_emu/grt_host.cpp), against the FP64
oracle (oracle_grt_trace / oracle_grt_trace_bwd, pinned in tests/test_grt_oracle.py): the
G-buffer per pixel, and the gradients of every Gaussian parameter block.  Output gradients are
zero on pixels within float32 reach of a discrete decision (oracle decision margin < 3e-3).
With PTIR_HOST_EMU_SANITIZE=1 every memory access of the device code is bounds-checked."""

import os

import numpy as np
import pytest
import torch

import host_emu_util as E
from gpu_helpers import env_table
from oracle import oracle as O
from ptir_amd import scenes

MARGIN = 3e-3


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return E.build("grt_host.cpp", str(tmp_path_factory.mktemp("host_emu") / "grt_host"))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("seed,wave", [(3, 0), (8, 1)])
def test_grt_host_build_matches_oracle(harness, tmp_path, seed, wave):
    sc = scenes.small_scene(seed, n=1024, res=24, spp=1, max_bounces=1)
    sc.density = torch.clamp(sc.density, max=0.9)
    inp = O.grt_inputs(sc.particle_density(), sc.features, sc.shading_normal, sph_degree=sc.sph_degree,
                       min_transmittance=sc.min_transmittance)
    rays = O.Rays(sc.rays_ori, sc.rays_dir, sc.ray_to_world)
    ref = O.grt_trace(inp, rays)
    keep = ref["margin"] >= MARGIN
    assert keep.mean() > 0.7
    rng = np.random.default_rng(seed)
    B, H, W = sc.rays_dir.shape[:3]
    k = keep.astype(np.float32)[..., None]
    g = {"rad": rng.standard_normal((B, H, W, 3)) * k, "dns": rng.standard_normal((B, H, W, 1)) * k,
         "hit": rng.standard_normal((B, H, W, 1)) * k, "hit2": rng.standard_normal((B, H, W, 1)) * k,
         "distortion": rng.standard_normal((B, H, W, 1)) * k, "snrm": rng.standard_normal((B, H, W, 3)) * k}
    g = {key: v.astype(np.float32) for key, v in g.items()}
    wd = str(tmp_path)
    E.write_launch(wd, sc, env_table(sc), max_bounces=1, wave_traversal=wave)
    for key, name in (("rad", "g_rad"), ("dns", "g_dns"), ("hit", "g_hit"), ("hit2", "g_hit2"),
                      ("distortion", "g_dist"), ("snrm", "g_snrm")):
        E.write_f32(wd, name, g[key])
    E.run(harness, wd)

    def rd(name, shape):
        return np.fromfile(os.path.join(wd, name + ".f32"), np.float32).reshape(shape).astype(np.float64)

    got = {"rad": rd("f_rad", (B, H, W, 3)), "dns": rd("f_dns", (B, H, W, 1)), "hit": rd("f_hit", (B, H, W, 2)),
           "hit2": rd("f_hit2", (B, H, W, 1)), "distortion": rd("f_dist", (B, H, W, 1)),
           "nrm": rd("f_nrm", (B, H, W, 3)), "snrm": rd("f_snrm", (B, H, W, 3)), "hits": rd("f_hits", (B, H, W, 1))}
    for key, v in got.items():
        r = ref[key]
        close = np.isclose(v, r, rtol=1e-3, atol=1e-4).all(axis=-1)
        assert close[keep].mean() >= 0.999, (key, close[keep].mean())
    vis = rd("f_vis", (sc.num_gaussians,))
    assert np.array_equal(vis > 0, ref["visibility"] > 0)

    N = sc.num_gaussians
    dd, dsh, dsn = O.grt_trace_bwd(inp, rays, ref, g)
    hd, hsh, hsn = rd("d_density", (N, 12)), rd("d_sh", tuple(sc.features.shape)), rd("d_snormal", (N, 3))
    for name, a, b in (("position", hd[:, 0:3], dd[:, 0:3]), ("density", hd[:, 3], dd[:, 3]),
                       ("rotation", hd[:, 4:8], dd[:, 4:8]), ("scale", hd[:, 8:11], dd[:, 8:11]),
                       ("sh", hsh, dsh), ("snormal", hsn, dsn)):
        err = rel_l2(a, b)
        print(f"seed {seed} {name}: relative L2 {err:.2e}", flush=True)
        assert np.isfinite(a).all() and err < 2e-3, name
    # the backward that reads the forward's hit tape (8 hits per pixel: the pixels with more
    # re-trace) adds the same gradients in the same order as the one that re-traces
    redo = float(rd("tape_redo", (1,))[0])
    print(f"seed {seed}: {redo:.3f} of pixels re-traced", flush=True)
    assert 0.0 < redo < 1.0
    assert np.array_equal(rd("t_density", (N, 12)), hd)
    assert np.array_equal(rd("t_sh", tuple(sc.features.shape)), hsh)
    assert np.array_equal(rd("t_snormal", (N, 3)), hsn)
