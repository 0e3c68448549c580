"""Launch-configuration autotuner (no counterpart in the reference, whose OptiX programs have one
configuration).

For one workload -- the scene tensors and launch arguments of a trace -- it times every
candidate configuration of the tracer and returns the fastest:

* ``kbuffer``    k-buffer depth {8, 16, 32}: hits collected per traversal chunk.  Fewer means a
                 tighter cull but more restarts of the walk; more means fewer restarts but a
                 larger sorted buffer in registers;
* ``traversal``  {"wave", "lane"}: primary rays of an 8x8 tile walk the hierarchy together or
                 each on its own;
* ``schedule``   {"megakernel", "wavefront"}: one lane per path, or stages over compacted
                 queues of live paths;
* ``builder``    {"lbvh", "ploc"}: the hierarchy over the same Morton-ordered leaves, by radix
                 splits of the codes or by bottom-up clustering of smallest union area (fewer
                 nodes per ray, a longer build; the timed step includes the build);
* ``wide_walks`` {"none", "all"}: every walk over the binary nodes, or over the 4-wide nodes
                 collapsed from them (half the dependent node fetches, four boxes per fetch).

All candidates compute the same paths (the same hits in the same order, the same decisions), so
the choice changes speed only.  The block size is not searched: a block is two 8x8 wave tiles
(128 lanes), and residency is set per wave by the LDS traversal stack and the registers, so the
block size only changes the granularity of the tile-to-CU mapping.  Leaves hold one Gaussian
each (the candidate test is the leaf test), which is not a search axis either.

The objective is the hierarchy build plus the forward (with the path tape the autograd
function records) plus, when ``backward=True``, the tape backward, timed with HIP events on the
current stream after one untimed call per configuration.
"""

from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import torch

from .tracer import HipTracer

KBUFFERS = (8, 16, 32)
TRAVERSALS = ("wave", "lane")
SCHEDULES = ("megakernel", "wavefront")
BUILDERS = ("lbvh", "ploc")
WIDE = ("none", "all")


@dataclass
class TuneResult:
    kbuffer: int
    traversal: str
    schedule: str
    builder: str
    wide_walks: str
    ms: float
    timings: dict = field(default_factory=dict)  # (kbuffer, traversal, schedule, builder, wide_walks) -> ms

    def tracer_kwargs(self) -> dict:
        return dict(kbuffer=self.kbuffer, traversal=self.traversal, schedule=self.schedule, builder=self.builder,
                    wide_walks=self.wide_walks)

    def as_dict(self) -> dict:
        return {"kbuffer": self.kbuffer, "traversal": self.traversal, "schedule": self.schedule,
                "builder": self.builder, "wide_walks": self.wide_walks, "ms": round(self.ms, 3),
                "timings_ms": {f"k{k}/{t}/{s}/{b}/{w}": round(v, 3)
                               for (k, t, s, b, w), v in sorted(self.timings.items())}}


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def autotune(geometry, scene_args, frame, ray_to_world, ray_ori, ray_dir, *, tracer_kwargs=None, trace_kwargs=None,
             backward: bool = True, reps: int = 3, kbuffers=KBUFFERS, traversals=TRAVERSALS,
             schedules=SCHEDULES, builders=BUILDERS, wide=WIDE) -> TuneResult:
    """Time every (kbuffer, traversal, schedule, builder, wide_walks) on one workload and return the
    fastest.

    geometry: (positions, rotations, scales, densities) for build_bvh; scene_args: the six scene
    tensors of HipTracer.trace (particle density, material, radiance, shading normal,
    environment, environment table); tracer_kwargs: the other HipTracer options (metallic, MIS,
    env sampling, ...); trace_kwargs: the launch options of HipTracer.trace (max_bounces,
    render_opts, ragged, rr_start, ...)."""
    tracer_kwargs = dict(tracer_kwargs or {})
    trace_kwargs = dict(trace_kwargs or {})
    max_bounces = int(trace_kwargs.get("max_bounces", 3))
    timings = {}
    for k, trav, sched, bld, ww in itertools.product(kbuffers, traversals, schedules, builders, wide):
        tr = HipTracer(kbuffer=k, traversal=trav, schedule=sched, builder=bld, wide_walks=ww, **tracer_kwargs)
        tape = tr.new_tape(ray_dir, max_bounces)

        def step():
            tr.build_bvh(*geometry)
            outs = tr.trace(frame, ray_to_world, ray_ori, ray_dir, *scene_args, tape=tape, **trace_kwargs)
            if backward:
                grad = torch.ones_like(outs[10])
                bwd_kw = {k2: v for k2, v in trace_kwargs.items() if k2 != "counters"}
                tr.trace_bwd(frame, ray_to_world, ray_ori, ray_dir, *outs[:8], outs[10], outs[11], *scene_args,
                             None, None, None, None, None, None, None, None, None, None, None, None, None, None,
                             grad, None, tape=tape, **bwd_kw)

        timings[(k, trav, sched, bld, ww)] = _time(step, reps)
        del tr, tape
    best = min(timings, key=timings.get)
    return TuneResult(best[0], best[1], best[2], best[3], best[4], timings[best], timings)
