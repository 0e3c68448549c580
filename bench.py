#!/usr/bin/env python
"""Throughput of the differentiable Gaussian path tracer on one MI355X.

Metric: forward + backward path samples per second on the object-scale workload
(300,000 Gaussians on a sphere-torus shell, env 512x256 with a sun lobe, 4 views of
800x800, 1 spp, maxBounces 3 = primary + 2 secondary rays with NEE and MIS), FP32.
The `nrc` block compares the radiance-cache form with the plain tracer at equal image MSE
against a 1,024-spp FP64 oracle image; the `workloads` block times the whole inversion
step (captured as one graph) on the batched ragged and the scene-scale workloads.
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
import csv
import glob
import json
import os
import re
import shutil
import signal
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# Algorithmic bytes per work unit: what each event the kernels count moves (a candidate's
# geometry record, an accepted hit's material and normal, a primary hit's SH coefficients, an
# env texel fetch, a gradient read-modify-write, a path's inputs and outputs, tape records).
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


def _pmc_values(outdir: str) -> dict:
    """{kernel name: {dispatch id: counter value summed over its rows}} of one rocprofv3 --pmc run."""
    vals: dict = {}
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                d = vals.setdefault(r["Kernel_Name"], {})
                d[int(r["Dispatch_Id"])] = d.get(int(r["Dispatch_Id"]), 0.0) + float(r["Counter_Value"])
    return vals


def _pmc_run(counter: str, argv: list, timeout: int) -> dict:
    """One `rocprofv3 --pmc <counter>` child run of `argv` (from /tmp, TMPDIR=/tmp)."""
    outdir = tempfile.mkdtemp(prefix="ptir_pmc_", dir="/tmp")
    cmd = ["rocprofv3", "--pmc", counter, "-d", outdir, "-o", "pmc", "--output-format", "csv", "--"] + argv
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                         start_new_session=True)
    try:
        p.wait(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        shutil.rmtree(outdir, ignore_errors=True)
        raise RuntimeError(f"rocprofv3 --pmc {counter} did not finish in {timeout} s")
    # rocprofv3 itself may end with a signal after the program and its results are complete
    vals = _pmc_values(outdir)
    shutil.rmtree(outdir, ignore_errors=True)
    if not vals:
        raise RuntimeError(f"rocprofv3 --pmc {counter} wrote no counters (exit {p.returncode})")
    return vals


def measure_traffic(tcfg: dict, tape: bool) -> dict:
    """HBM bytes per launch of the tracer kernels, measured on this build: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) of one object-scale step in child runs under rocprofv3, each
    scaled by what the same counter reads for a 1 GiB device copy.  A kernel's figure is its
    timed dispatch (the last; the first is the counting step), summed over its instances."""
    me = [sys.executable, os.path.join(ROOT, "bench.py")]
    probe = me + ["--no-cpu-baseline", "--no-nrc", "--no-workloads", "--no-traffic", "--steps", "1", "--warmup", "0",
                  "--tracer-config", json.dumps(tcfg)] + ([] if tape else ["--no-tape"])
    copy = me + ["--copy-calibration", str(1 << 30)]
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        print(f"traffic: rocprofv3 --pmc {counter}", file=sys.stderr, flush=True)
        step = _pmc_run(counter, probe, 300)
        cal = _pmc_run(counter, copy, 120)
        copies = [v for k, v in cal.items() if "copyBuffer" in k or "copy" in k.lower()]
        ref = max((max(d.values()) for d in copies), default=0.0)
        if ref <= 0:
            raise RuntimeError(f"no copy kernel in the {counter} calibration run")
        scale = (1 << 30) / 1024 / ref  # bytes per KiB counted
        for name, d in step.items():
            last = d[max(d)] * 1024 * scale
            for short in ("k_trace_fwd", "k_trace_bwd", "k_wf_primary", "k_wf_nee", "k_wf_bounce"):
                if short in name:
                    got.setdefault(short, {}).setdefault(counter, 0.0)
                    got[short][counter] += last
        got.setdefault("_scale", {})[counter] = round(scale, 3)
    return got


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


def time_primary_only(steps: int = 10, warmup: int = 2) -> dict:
    """The 3DGRT primary-only tracer on the object-scale rays: G-buffer forward and the
    backward to every Gaussian parameter (random output gradients), timed with HIP events.
    The forward records its hits on a tape (128 per pixel) that the backward reads, as the
    autograd Tracer does; `ms_bwd_retrace` is the backward that re-traces every pixel."""
    import torch

    from ptir_amd import scenes
    from ptir_amd.grt import GRTTracer

    dev = torch.device("cuda:0")
    sc = scenes.object_scene(0)
    g = sc.to(dev)
    out = {"gaussians": sc.num_gaussians, "rays": sc.num_paths}
    for builder in ("lbvh", "ploc"):
        tr = GRTTracer(builder=builder)
        args = (0, g.ray_to_world, g.rays_ori, g.rays_dir, g.particle_density(), g.features, g.shading_normal)
        gen = torch.Generator(device=dev).manual_seed(0)
        tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
        tape = tr.new_tape(g.rays_dir, hits_per_pixel=128)
        outs = tr.trace(*args, 0, sc.sph_degree, sc.min_transmittance)
        grads = [torch.randn(o.shape, device=dev, generator=gen) for o in outs[:7]]
        grads[2] = grads[2][..., :1].contiguous()
        ev, ev_re = [], []
        for i in range(warmup + steps):
            e0, e1, e2, e3 = (torch.cuda.Event(enable_timing=True) for _ in range(4))
            tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
            e0.record()
            outs = tr.trace(*args, 0, sc.sph_degree, sc.min_transmittance, tape=tape)
            e1.record()
            tr.trace_bwd(*args[:4], *outs[:7], *args[4:], *grads, 0, sc.sph_degree, sc.min_transmittance, tape=tape)
            e2.record()
            tr.trace_bwd(*args[:4], *outs[:7], *args[4:], *grads, 0, sc.sph_degree, sc.min_transmittance)
            e3.record()
            if i >= warmup:
                ev.append((e0, e1, e2))
                ev_re.append((e2, e3))
        torch.cuda.synchronize()
        fwd = sum(a.elapsed_time(b) for a, b, _ in ev) / steps
        bwd = sum(b.elapsed_time(c) for _, b, c in ev) / steps
        bwd_re = sum(a.elapsed_time(b) for a, b in ev_re) / steps
        redo = float((tape[16:16 + sc.num_paths] < 0).float().mean())
        out[builder] = {"ms_fwd": round(fwd, 3), "ms_bwd": round(bwd, 3), "ms_bwd_retrace": round(bwd_re, 3),
                        "tape_retraced_pixels": round(redo, 4),
                        "rays_per_s_fwd_bwd": round(sc.num_paths / ((fwd + bwd) * 1e-3), 1)}
        del tr, outs, grads, tape
    torch.cuda.empty_cache()
    return out


def nrc_comparison(train_steps: int = 150, timed_steps: int = 5, spps=(1, 4, 16, 64)) -> dict:
    """Radiance cache (NRC) form against the plain tracer on the object-scale workload.

    Both render paths of up to maxBounces 4 (three bounces): the plain tracer traces them
    out; the cache form ends them at the first secondary vertex with a cache query and trains
    the cache online, one Adam step per training step, on up to 65,536 path suffixes from full
    paths of every 4th pixel row.  Reported: fwd+bwd path samples/s of each form on the full
    4 x 800 x 800 batch, the MSE of S-spp images of a 96x96 crop against a 1,024-spp FP64
    oracle image of the plain tracer (tests/golden/c5_reference.npz), and, at equal MSE, the
    samples and time each form needs."""
    import numpy as np
    import torch

    from ptir_amd import envmap, scenes
    from ptir_amd.cache import CachedPathTracer, RadianceCache
    from ptir_amd.tracer import HipTracer

    ref_path = os.path.join(ROOT, "tests", "golden", "c5_reference.npz")
    ref = np.load(ref_path)
    _, y0, x0, ch, cw = (int(v) for v in ref["crop"])
    dev = torch.device("cuda:0")
    sc = scenes.object_scene(0)
    tab = envmap.pack_alias_table(envmap.build_environment_alias_table(sc.env)).to(dev)
    g = sc.to(dev)
    tr = HipTracer(enable_metallic=True)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    scene_args = (g.particle_density(), g.particle_material(), g.features, g.shading_normal, g.env, tab)
    box = torch.cat([g.positions.min(0).values - 0.05, g.positions.max(0).values + 0.05]).contiguous()
    cached = CachedPathTracer(tr, RadianceCache(max_batch=65536, seed=0), box, max_bounces=4)
    T = g.ray_to_world
    for s in range(train_steps):
        cached.train(100_000 + s, T, g.rays_ori, g.rays_dir, scene_args, offset=s % 4)
    torch.cuda.synchronize()
    _, _, loss = cached.cache.read_grads()
    target = scenes.target_image(sc).to(dev)
    tape = tr.new_tape(g.rays_dir, 4)

    def bwd(outs, frame, cache_mode):
        grad = ((outs[10] - target) * (2.0 / outs[10].numel())).contiguous()
        tr.trace_bwd(frame, T, g.rays_ori, g.rays_dir, *outs[:8], outs[10], outs[11], *scene_args, None, None, None,
                     None, None, None, None, None, None, None, None, None, None, None, grad, None, max_bounces=4,
                     tape=tape, cache_mode=cache_mode)

    def plain_step(f):
        bwd(tr.trace(f, T, g.rays_ori, g.rays_dir, *scene_args, max_bounces=4, tape=tape), f, 0)

    def nrc_step(f):
        cached.train(200_000 + f, T, g.rays_ori, g.rays_dir, scene_args, offset=f % 4)
        outs, _ = cached.render(f, T, g.rays_ori, g.rays_dir, scene_args, tape=tape)
        bwd(outs, f, 1)

    timing = {}
    for name, fn in (("plain", plain_step), ("nrc", nrc_step)):
        fn(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(1, timed_steps + 1):
            fn(f)
        torch.cuda.synchronize()
        timing[name] = (time.perf_counter() - t0) / timed_steps
    o = g.rays_ori[:1, y0:y0 + ch, x0:x0 + cw].contiguous()
    d = g.rays_dir[:1, y0:y0 + ch, x0:x0 + cw].contiguous()
    refimg = torch.from_numpy(ref["image"]).double().to(dev)

    def mse_curve(render):
        acc = torch.zeros(ch, cw, 3, dtype=torch.float64, device=dev)
        out, n = {}, 0
        for S in spps:
            while n < S:
                # frames from 10,000 on: disjoint from the reference image's frames 0..1023
                acc += render(10_000 + n)[0].double()
                n += 1
            out[S] = float(((acc / n - refimg) ** 2).mean())
        return out

    mse_plain = mse_curve(lambda f: tr.trace(f, T, o, d, *scene_args, max_bounces=4)[10])
    mse_nrc = mse_curve(lambda f: cached.render(f, T, o, d, scene_args)[0][10])
    mse_cut = mse_curve(lambda f: tr.trace(f, T, o, d, *scene_args, max_bounces=2)[10])
    paths = sc.num_paths

    def spp_for(mse_target):
        """Samples the plain tracer needs for mse_target (log-log interpolation of its curve)."""
        xs = np.log([float(S) for S in spps])
        ys = np.log([mse_plain[S] for S in spps])
        slope = (ys[-1] - ys[0]) / (xs[-1] - xs[0])
        return float(np.exp(xs[0] + (np.log(mse_target) - ys[0]) / slope))

    equal = {}
    for S in spps:
        sp = spp_for(mse_nrc[S])
        equal[str(S)] = {"nrc_spp": S, "plain_spp_same_mse": round(sp, 2),
                         "time_ratio_plain_over_nrc": round((sp * timing["plain"]) / (S * timing["nrc"]), 3)}
    return {
        "workload": "object-scale, maxBounces 4; cache queried at the first secondary vertex",
        "plain": {"ms_per_step": round(timing["plain"] * 1e3, 3), "path_samples_per_s": round(paths / timing["plain"], 1),
                  "mse_vs_fp64": {str(k): v for k, v in mse_plain.items()}},
        "nrc": {"ms_per_step": round(timing["nrc"] * 1e3, 3), "path_samples_per_s": round(paths / timing["nrc"], 1),
                "mse_vs_fp64": {str(k): v for k, v in mse_nrc.items()},
                "train_steps_before_timing": train_steps, "train_samples_per_step": round(
                    cached.train_samples / max(1, train_steps + timed_steps + 1), 1),
                "last_train_loss": float(loss.item())},
        "cut_at_first_secondary_vertex": {"mse_vs_fp64": {str(k): v for k, v in mse_cut.items()}},
        "equal_mse": equal,
        "reference": {"spp": int(ref["spp"]), "crop": [y0, x0, ch, cw], "precision": "fp64 oracle",
                      "indirect_share": float(ref["indirect_share"])},
    }


def time_inversion_workload(name: str, steps: int, warmup: int) -> dict:
    """Steps/s of the whole inversion step (device env table, BVH refit, forward, loss
    gradient, backward, Adam) on a workload, captured as one graph and replayed (the capture
    first fits the tape's per-path hit capacity so that the backward re-traces nothing)."""
    import torch

    from ptir_amd import envmap, scenes
    from ptir_amd.autotune import autotune
    from ptir_amd.inversion import InversionStep
    from ptir_amd.tracer import HipTracer, RaggedBatch

    dev = torch.device("cuda:0")
    t_gen = time.perf_counter()
    sc = scenes.WORKLOADS[name](0)
    gen_s = time.perf_counter() - t_gen
    kw = {"rr_start": sc.extra.get("rr_start", 0)}
    if name == "batched":
        kw["ragged"] = RaggedBatch.build(sc.extra["images"], dev)
    g = sc.to(dev)
    target = scenes.target_image(sc).to(dev)
    # the autotuner over {one-lane kernel, wavefront} x {LBVH, PLOC} at k-buffer 16 (build + forward,
    # one timed call each; k-buffer 32 and per-lane primary traversal lose on every workload here)
    tab = envmap.build_cdf_table(g.env)
    scene_args = (g.particle_density(), g.particle_material(), g.features, g.shading_normal, g.env, tab)
    tune = autotune((g.positions, g.rotation, g.scale, g.density), scene_args, 0, g.ray_to_world, g.rays_ori,
                    g.rays_dir, tracer_kwargs=dict(enable_metallic=True, env_sampling="cdf"),
                    trace_kwargs=dict(max_bounces=sc.max_bounces, **kw), backward=False, reps=1, kbuffers=(16,),
                    traversals=("wave",))
    del tab, scene_args
    step = InversionStep(HipTracer(enable_metallic=True, env_sampling="cdf", **tune.tracer_kwargs()), g, target,
                         **kw)
    step.capture()
    for _ in range(warmup):
        step.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    paths = sc.num_paths
    out = {"gaussians": sc.num_gaussians, "paths_per_step": paths, "max_bounces": sc.max_bounces,
           "rr_start": kw["rr_start"], "ms_per_step": round(dt * 1e3, 3), "path_samples_per_s": round(paths / dt, 1),
           "steps": steps, "graph": True, "scene_generation_s": round(gen_s, 1), "loss_last": round(step.loss(), 6),
           "tracer": tune.tracer_kwargs(), "autotune": tune.as_dict(),
           "tape_hits_per_path": step.tracer.tape_hits_per_path, "tape_gb": round(step.tape.numel() / 1e9, 2)}
    if name == "batched":
        out["images"] = len(sc.extra["images"])
    del step, g
    torch.cuda.empty_cache()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--kbuffer", type=int, default=0, help="k-buffer depth (0: the autotuner's choice)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use k-buffer 16, per-wave primary traversal, the one-lane kernel and PLOC instead of tuning")
    ap.add_argument("--cpu-stride", type=int, default=8, help="oracle path sample stride for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tape", action="store_true", help="backward re-traces every segment (reference behaviour)")
    ap.add_argument("--no-nrc", action="store_true", help="skip the radiance-cache comparison")
    ap.add_argument("--primary-only", action="store_true",
                    help="time only the 3DGRT primary-only tracer on the object-scale rays (for profiling)")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the batched (ragged, 500k) and scene-scale (3M) inversion-step timings")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy a float32 buffer of BYTES bytes 5 times (HBM counter calibration)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="take roofline.traffic from profiles/traffic_object.md instead of measuring it "
                         "with rocprofv3 --pmc child runs of this build")
    ap.add_argument("--tracer-config", default="", metavar="JSON",
                    help="tracer configuration to use instead of tuning (the traffic runs pass the tuned one)")
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
    if args.primary_only:
        print(json.dumps({"primary_only": time_primary_only(steps=args.steps, warmup=args.warmup)}))
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the tracer shards by view or pixel tile, "
                                   "but no multi-GPU path is built; run with --gpus 1"}))
        return 2

    import torch

    from ptir_amd import envmap, scenes
    from ptir_amd.autotune import autotune
    from ptir_amd.tracer import HipTracer

    dev = torch.device("cuda:0")
    scene = scenes.object_scene(0)
    table = envmap.pack_alias_table(envmap.build_environment_alias_table(scene.env))
    g = scene.to(dev)
    tab = table.to(dev)
    target = scenes.target_image(scene).to(dev)
    paths = scene.num_paths
    dens, mat = g.particle_density(), g.particle_material()
    fargs = (scene.frame, g.ray_to_world, g.rays_ori, g.rays_dir, dens, mat, g.features, g.shading_normal, g.env, tab)
    kw = dict(render_opts=0, sph_degree=scene.sph_degree, min_transmittance=scene.min_transmittance,
              max_bounces=scene.max_bounces)
    tuned = None
    tcfg = dict(kbuffer=16, traversal="wave", schedule="megakernel", builder="ploc", wide_walks="all")
    if args.tracer_config:
        tcfg = json.loads(args.tracer_config)
    elif not args.no_autotune:
        print("autotuning the object-scale step", file=sys.stderr, flush=True)
        tuned = autotune((g.positions, g.rotation, g.scale, g.density), fargs[4:], scene.frame, g.ray_to_world,
                         g.rays_ori, g.rays_dir, tracer_kwargs=dict(enable_metallic=True), trace_kwargs=kw,
                         backward=True, reps=2)
        tcfg = tuned.tracer_kwargs()
    if args.kbuffer:
        tcfg["kbuffer"] = args.kbuffer
    tracer = HipTracer(enable_metallic=True, **tcfg)
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
    fwd_name = "k_trace_fwd" if tcfg["schedule"] == "megakernel" else "k_wf_primary+k_wf_nee+k_wf_bounce"
    dom, ms_dom, by = ("k_trace_bwd", ms_bwd, bb) if ms_bwd >= ms_fwd else (fwd_name, ms_fwd, fb)
    achieved = by / (ms_dom * 1e-3) / 1e9
    traffic, source = None, ""
    if not args.no_traffic:
        try:
            meas = measure_traffic(tcfg, tape is not None)
            parts = [meas.get(k) for k in dom.split("+")]
            if all(p and "FETCH_SIZE" in p and "WRITE_SIZE" in p for p in parts):
                traffic = sum(p["FETCH_SIZE"] + p["WRITE_SIZE"] for p in parts)
                sc = meas["_scale"]
                source = (f"measured on this build: rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, child runs of one "
                          f"step, the timed dispatch, scaled by a 1 GiB copy (x{sc['FETCH_SIZE']} / "
                          f"x{sc['WRITE_SIZE']})")
            else:
                source = "measurement found no dispatch of the kernel"
        except (OSError, RuntimeError, ValueError, KeyError) as exc:
            source = f"measurement failed ({exc})"
    if traffic is None:
        traffic = read_traffic(dom)
        source = ("profiles/traffic_object.md (committed rocprofv3 PMC summary)" +
                  (f"; {source}" if source else ""))
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_source": source,
                "kernel": dom, "algorithmic_bytes": by, "ms_per_launch": round(ms_dom, 3),
                "ms_fwd": round(ms_fwd, 3), "ms_bwd": round(ms_bwd, 3)}
    line = {
        "metric": "fwd+bwd path samples/s, 800x800 2-bounce; NRC vs plain at equal image MSE",
        "value": round(value, 1), "unit": "path samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "object-scale: 300k Gaussians, 4 views 800x800, 1 spp, maxBounces 3, env 512x256, "
                               "plain tracer fwd+bwd (ptir-mode backward); radiance-cache comparison in nrc",
                   "gaussians": scene.num_gaussians, "paths_per_step": paths, **tcfg,
                   "backward": "re-trace" if tape is None else "path tape",
                   "autotune": None if tuned is None else tuned.as_dict()},
        "roofline": roofline,
        "work_units": {"fwd": cf, "bwd": cb},
    }
    if not args.no_workloads:
        print("primary-only tracer", file=sys.stderr, flush=True)
        line["primary_only"] = time_primary_only()
    if not args.no_nrc:
        del tape, tk
        tape, tk = None, {}
        torch.cuda.empty_cache()
        print("radiance-cache comparison", file=sys.stderr, flush=True)
        line["nrc"] = nrc_comparison()
    if not args.no_workloads:
        del tape, tk
        torch.cuda.empty_cache()
        line["workloads"] = {}
        for name in ("batched", "scene"):
            print(f"timing workload {name}", file=sys.stderr, flush=True)
            line["workloads"][name] = time_inversion_workload(name, steps=3, warmup=1)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, table, args.cpu_stride)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
