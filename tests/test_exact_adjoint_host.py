"""The exact-adjoint device code (ptir_amd/csrc/adjoint_path.hpp) compiled for the host with g++
and run one lane at a time (tests/host_emu/adjoint_host.cpp), against the FP64 dual-number
oracle: see tests/adjoint_cases.py for the identity checked and its tolerance.  The same
source is the kernel's on the GPU; this runs it without one, with every memory access of the
device code bounds-checked when PTIR_HOST_EMU_SANITIZE=1 (AddressSanitizer and
UndefinedBehaviorSanitizer)."""

import os
import subprocess

import numpy as np
import pytest

import adjoint_cases as A

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "host_emu")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("host_emu") / "adjoint_host")
    flags = ["-O2"]
    if os.environ.get("PTIR_HOST_EMU_SANITIZE") == "1":
        flags = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    cmd = ["g++", "-std=c++17", *flags, "-I", os.path.join(EMU, "shim"), "-include",
           os.path.join(EMU, "shim", "hip", "hip_runtime.h"), os.path.join(EMU, "adjoint_host.cpp"), "-o", out]
    subprocess.run(cmd, check=True, cwd=ROOT)
    return out


def _run(harness, workdir, sc, tab, cfg, g):
    f32 = lambda name, a: np.ascontiguousarray(np.asarray(a, dtype=np.float32)).tofile(  # noqa: E731
        os.path.join(workdir, name + ".f32"))
    f32("density", sc.particle_density())
    f32("material", sc.particle_material())
    f32("sh", sc.features)
    f32("snormal", sc.shading_normal)
    f32("env", sc.env)
    f32("table", tab)
    f32("ori", sc.rays_ori)
    f32("dir", sc.rays_dir)
    f32("ray_to_world", np.asarray(sc.ray_to_world).reshape(-1)[:12])
    for k in A.OUTS:
        f32("g_" + k, g[k])
    B, H, W = sc.rays_dir.shape[:3]
    meta = dict(N=sc.num_gaussians, B=B, H=H, W=W, env_h=sc.env.shape[0], env_w=sc.env.shape[1], tab_h=tab.shape[1],
                tab_w=tab.shape[2], sh_stride=sc.features.shape[1], kernel_degree=4, min_response=0.0113,
                density_clamping=1, sph_degree=sc.sph_degree, max_bounces=sc.max_bounces,
                min_transmittance=sc.min_transmittance, frame=sc.frame, render_opts=cfg["render_opts"],
                enable_mis=int(cfg["enable_mis"]), enable_metallic=1, enable_normals=1, rr_start=cfg["rr_start"],
                tab_kind=0)
    with open(os.path.join(workdir, "meta.txt"), "w") as fh:
        for k, v in meta.items():
            fh.write(f"{k} {v}\n")
    r = subprocess.run([harness, workdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    N = sc.num_gaussians
    rd = lambda name, shape: np.fromfile(os.path.join(workdir, name + ".f32"), np.float32).reshape(shape).astype(  # noqa: E731
        np.float64)
    return {"density": rd("d_density", (N, 12)), "material": rd("d_material", (N, 5)),
            "sh": rd("d_sh", tuple(sc.features.shape)), "snormal": rd("d_snormal", (N, 3)),
            "env": rd("d_env", tuple(sc.env.shape))}


@pytest.mark.parametrize("case", list(A.CASES))
def test_exact_adjoint_host_build_matches_dual_derivatives(harness, tmp_path, case):
    sc, tab, inp, rays, ref, rng = A.setup(case)
    keep = ref["margin"] >= A.MARGIN
    assert keep.mean() > 0.7
    g = A.output_grads(keep, rng)
    grads = _run(harness, str(tmp_path), sc, tab, A.CASES[case], g)
    for k, v in grads.items():
        assert np.isfinite(v).all(), k
    A.check_blocks(sc, inp, rays, g, grads, rng, case)
