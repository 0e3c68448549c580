"""Build and feed the host runs of the tracer's device code (tests/host_emu/*.cpp): g++
compiles ptir_amd/csrc/*.hpp for the CPU with the shim of tests/host_emu/shim, and a launch
is handed over as meta.txt plus float32 files (tests/host_emu/host_scene.hpp reads them).
PTIR_HOST_EMU_SANITIZE=1 builds with AddressSanitizer and UndefinedBehaviorSanitizer."""

import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMU = os.path.join(HERE, "host_emu")


def build(source: str, out: str) -> str:
    flags = ["-O2"]
    if os.environ.get("PTIR_HOST_EMU_SANITIZE") == "1":
        flags = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    cmd = ["g++", "-std=c++17", *flags, "-I", os.path.join(EMU, "shim"), "-include",
           os.path.join(EMU, "shim", "hip", "hip_runtime.h"), os.path.join(EMU, source), "-o", out]
    subprocess.run(cmd, check=True, cwd=ROOT)
    return out


def write_f32(workdir: str, name: str, a) -> None:
    np.ascontiguousarray(np.asarray(a, dtype=np.float32)).tofile(os.path.join(workdir, name + ".f32"))


def write_launch(workdir: str, sc, tab, *, max_bounces=None, render_opts=0, enable_mis=True, rr_start=0,
                 cache_mode=0, **extra) -> None:
    """Scene inputs and launch options of a SyntheticScene (kernel degree 4, min response 0.0113)."""
    write_f32(workdir, "density", sc.particle_density())
    write_f32(workdir, "material", sc.particle_material())
    write_f32(workdir, "sh", sc.features)
    write_f32(workdir, "snormal", sc.shading_normal)
    write_f32(workdir, "env", sc.env)
    write_f32(workdir, "table", tab)
    write_f32(workdir, "ori", sc.rays_ori)
    write_f32(workdir, "dir", sc.rays_dir)
    write_f32(workdir, "ray_to_world", np.asarray(sc.ray_to_world).reshape(-1)[:12])
    B, H, W = sc.rays_dir.shape[:3]
    meta = dict(N=sc.num_gaussians, B=B, H=H, W=W, env_h=sc.env.shape[0], env_w=sc.env.shape[1],
                tab_h=tab.shape[1], tab_w=tab.shape[2], sh_stride=sc.features.shape[1], kernel_degree=4,
                min_response=0.0113, density_clamping=1, sph_degree=sc.sph_degree,
                max_bounces=sc.max_bounces if max_bounces is None else max_bounces,
                min_transmittance=sc.min_transmittance, frame=sc.frame, render_opts=render_opts,
                enable_mis=int(enable_mis), enable_metallic=1, enable_normals=1, rr_start=rr_start, tab_kind=0,
                cache_mode=cache_mode)
    meta.update(extra)
    with open(os.path.join(workdir, "meta.txt"), "w") as fh:
        for k, v in meta.items():
            fh.write(f"{k} {v}\n")


def run(binary: str, workdir: str) -> str:
    r = subprocess.run([binary, workdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout
