#!/usr/bin/env python
"""Throughput of the differentiable Gaussian path tracer on one MI355X.

Metric: forward + backward path samples per second on the object-scale workload
(300,000 Gaussians on a sphere-torus shell, env 512x256 with a sun lobe, 4 views of
800x800, 1 spp, maxBounces 3 = primary + 2 secondary rays with NEE and MIS), FP32.
One step is one training step of the tracer: BVH build, forward trace, L2 loss gradient
against a seeded target, backward trace (material + environment gradients, path-replay
semantics of the reference).  Inputs are synthetic and resident in HBM before timing.

Printed JSON fields follow the driver contract; `roofline` is for the dominant kernel
(algorithmic bytes from the kernels' own work counters / average launch time measured
with HIP events on the tracer's stream), `cpu_baseline` is the FP64 CPU oracle timed on
a strided sample of the same paths on this host's cores.
"""

from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# Algorithmic bytes per work unit (see DESIGN.md "Measurement").
B_CAND, B_ACC, B_SH, B_ENV, B_MAT_ATOM, B_ENV_ATOM, B_IO_FWD, B_IO_BWD = 44, 32, 192, 64, 40, 96, 152, 92
B_TAPE_PATH, B_TAPE_SEG, B_TAPE_HIT = 64, 64, 8  # path tape: record header, segment record, hit entry


def fwd_bytes(c: dict) -> float:
    return (c["candidates"] * B_CAND + c["accepted"] * B_ACC + c["accepted_primary"] * B_SH +
            c["shadow_candidates"] * B_CAND + c["env_lookups"] * B_ENV + c["paths"] * B_IO_FWD)


def bwd_bytes(c: dict) -> float:
    """Re-tracing backward: the replayed traversal plus the gradient atomics."""
    return (c["candidates"] * B_CAND + c["accepted"] * B_ACC + c["shadow_candidates"] * B_CAND +
            c["env_lookups"] * B_ENV + c["mat_atomics"] * B_MAT_ATOM + c["env_atomics"] * B_ENV_ATOM +
            c["paths"] * B_IO_BWD)


def bwd_tape_bytes(c: dict) -> float:
    """Tape backward: path records and hit entries instead of traversal."""
    return (c["paths"] * (B_IO_BWD + B_TAPE_PATH) + c["segments"] * B_TAPE_SEG + c["mat_atomics"] * (B_MAT_ATOM + 8) +
            c["env_atomics"] * B_ENV_ATOM + c["env_lookups"] * B_ENV)


def read_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (profiles/traffic_object.md, a line `traffic <kernel> <bytes>`), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_object.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"^traffic\s+(\S+)\s+([0-9.eE+]+)", line.strip().strip("`"))
        if m and m.group(1) == kernel:
            return float(m.group(2))
    return None


def cpu_baseline(scene, table, stride: int) -> dict:
    """FP64 oracle (test infrastructure, used here only as the reported baseline) on
    every `stride`-th path of the same workload, forward + backward."""
    import numpy as np

    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(cores, 16))
    inp = O.Inputs.from_scene(scene, table)
    rays = O.Rays(scene.rays_ori, scene.rays_dir, scene.ray_to_world, path_stride=stride)
    t0 = time.perf_counter()
    out = O.trace(inp, rays, nthreads=cores, counters=True)
    g = np.full(out["pbr"].shape, 1e-6, dtype=np.float32)
    O.trace_bwd(inp, rays, out["pbr"], out["light"], g, np.zeros_like(g), np.zeros(out["mat"].shape, np.float32),
                nthreads=cores)
    dt = time.perf_counter() - t0
    n = out["counters"]["paths"]
    return {"value": n / dt, "unit": "path samples/s", "cores": cores, "kind": "port",
            "sample": f"every {stride}th path of the 4x800x800 batch ({n} paths), FP64 forward + backward "
                      f"incl. the oracle's own BVH build, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kbuffer", type=int, default=16)
    ap.add_argument("--cpu-stride", type=int, default=8, help="oracle path sample stride for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tape", action="store_true", help="backward re-traces every segment (reference behaviour)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy a float32 buffer of BYTES bytes 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration // 4
        src = torch.ones(n, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 5}))
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the tracer shards by view or pixel tile, "
                                   "but no multi-GPU path is built; run with --gpus 1"}))
        return 2

    import torch

    from ptir_amd import envmap, scenes
    from ptir_amd.tracer import HipTracer

    dev = torch.device("cuda:0")
    scene = scenes.object_scene(0)
    table = envmap.pack_alias_table(envmap.build_environment_alias_table(scene.env))
    g = scene.to(dev)
    tab = table.to(dev)
    target = scenes.target_image(scene).to(dev)
    tracer = HipTracer(enable_metallic=True, kbuffer=args.kbuffer)
    paths = scene.num_paths
    dens, mat = g.particle_density(), g.particle_material()
    fargs = (scene.frame, g.ray_to_world, g.rays_ori, g.rays_dir, dens, mat, g.features, g.shading_normal, g.env, tab)
    kw = dict(render_opts=0, sph_degree=scene.sph_degree, min_transmittance=scene.min_transmittance,
              max_bounces=scene.max_bounces)
    ev = {"fwd": [], "bwd": []}

    tape = None if args.no_tape else tracer.new_tape(g.rays_dir, scene.max_bounces)
    tk = {} if tape is None else {"tape": tape}

    def step(timed=False, counters=False):
        tracer.build_bvh(g.positions, g.rotation, g.scale, g.density)
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        e0.record()
        outs = tracer.trace(*fargs, counters=counters, **tk, **kw)
        e1.record()
        cf = tracer.last_counters
        grad = (2.0 / outs[10].numel()) * (outs[10] - target)
        e_b = torch.cuda.Event(enable_timing=True)
        e_b.record()
        tracer.trace_bwd(scene.frame, g.ray_to_world, g.rays_ori, g.rays_dir, *outs[:8], outs[10], outs[11], dens, mat,
                         g.features, g.shading_normal, g.env, tab, None, None, None, None, None, None, None, None,
                         None, None, None, None, None, None, grad, None, counters=counters, **tk, **kw)
        e2.record()
        if timed:
            ev["fwd"].append((e0, e1))
            ev["bwd"].append((e_b, e2))
        return cf, tracer.last_counters

    cf, cb = step(counters=True)  # work units of one step (untimed)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(timed=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_fwd = sum(a.elapsed_time(b) for a, b in ev["fwd"]) / args.steps
    ms_bwd = sum(a.elapsed_time(b) for a, b in ev["bwd"]) / args.steps
    value = paths * args.steps / dt
    fb = fwd_bytes(cf) + (0 if tape is None else cf["paths"] * B_TAPE_PATH + cf["segments"] * B_TAPE_SEG +
                          cf["accepted"] * B_TAPE_HIT)
    bb = bwd_bytes(cb) if tape is None else bwd_tape_bytes(cb)
    dom, ms_dom, by = ("k_trace_bwd", ms_bwd, bb) if ms_bwd >= ms_fwd else ("k_trace_fwd", ms_fwd, fb)
    achieved = by / (ms_dom * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": read_traffic(dom),
                "kernel": dom, "algorithmic_bytes": by, "ms_per_launch": round(ms_dom, 3),
                "ms_fwd": round(ms_fwd, 3), "ms_bwd": round(ms_bwd, 3)}
    line = {
        "metric": "fwd+bwd path samples/s, 800x800 2-bounce; NRC vs plain at equal image MSE",
        "value": round(value, 1), "unit": "path samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "object-scale: 300k Gaussians, 4 views 800x800, 1 spp, maxBounces 3, env 512x256, "
                               "plain tracer (radiance cache not built), ptir-mode backward",
                   "gaussians": scene.num_gaussians, "paths_per_step": paths, "kbuffer": args.kbuffer,
                   "backward": "re-trace" if tape is None else "path tape"},
        "roofline": roofline,
        "work_units": {"fwd": cf, "bwd": cb},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, table, args.cpu_stride)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
