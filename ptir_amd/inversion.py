"""One PTIR inversion training step on the device, optionally captured as one graph.

PTIR's second stage freezes the Gaussians' geometry and optimises their materials and the
environment map by differentiable path tracing (configs/base_ptir.yaml:13,30-38; the
trainer loop threedgrut/trainer.py:1620-1836 with the tracer of threedgptir_tracer/).  One
step here is:

  1. the environment's importance table (CDF layout, built on the GPU);
  2. a refit of the hierarchy (topology kept; no host synchronisation);
  3. the forward trace, recording the path tape;
  4. the gradient of the L2 loss of the PBR image against a target;
  5. the backward trace (material and environment gradients);
  6. Adam on albedo, roughness, metallic and the environment.

Nothing in it synchronises with the host or allocates outside torch's allocator, so the
whole step is captured once with torch.cuda.graph and replayed; `replay()` then costs one
graph launch per training step.

Before capture, the tape's per-path hit capacity is fitted to the scene (`fit_tape`): a path
whose accepted hits do not fit its tape entries and a pool chunk is re-traced by the
backward, which on dense scenes costs more than the forward.  The capacity doubles until an
eager step's backward re-traces nothing (or a memory cap is reached).
"""

from __future__ import annotations

import torch

from . import _lib, envmap
from .optim import DeviceAdam
from .tracer import HipTracer, RaggedBatch


class InversionStep:
    def __init__(self, tracer: HipTracer, scene, target: torch.Tensor, *, lr_albedo=0.035, lr_roughness=0.025,
                 lr_metallic=0.0, lr_environment=0.035, selective=False, ragged: RaggedBatch | None = None,
                 rr_start: int = 0, refit: bool = True, table_grid=envmap.TABLE_GRID, adam_eps: float = 1e-15,
                 fit_tape: bool = True, tape_max_bytes: int = 64 << 30):
        """scene: a SyntheticScene on the device (its geometry is frozen; albedo, roughness,
        metallic and env are optimised in place).  target: the PBR image the loss compares
        with, shaped like the tracer's pbr output."""
        if tracer.tab_kind != _lib.ENV_TABLE_CDF:
            raise ValueError("the inversion step rebuilds the environment table on the device: "
                             "use HipTracer(env_sampling='cdf')")
        self.tracer = tracer
        self.sc = scene
        self.target = target.contiguous()
        self.ragged = ragged
        self.rr_start = int(rr_start)
        self.refit = bool(refit)
        self.table = torch.empty((3, table_grid[0], table_grid[1]), dtype=torch.float32, device=target.device)
        sc = scene
        self.density = sc.particle_density()
        self.params = {"albedo": sc.albedo, "roughness": sc.roughness, "metallic": sc.metallic, "env": sc.env}
        for t in self.params.values():
            t.grad = torch.zeros_like(t)
        self.adam = DeviceAdam([(sc.albedo, lr_albedo), (sc.roughness, lr_roughness), (sc.metallic, lr_metallic),
                                (sc.env, lr_environment)], eps=adam_eps, selective=selective)
        self.fit = bool(fit_tape)
        self.tape_max_bytes = int(tape_max_bytes)
        self.tape = self._new_tape()
        self.tape_retraced = None  # paths re-traced by the last fitted step's backward
        self.pbr = None  # the PBR image of the last step (a graph-owned buffer after capture)
        tracer.build_bvh(sc.positions, sc.rotation, sc.scale, sc.density)  # the initial full build
        self.graph = None

    def _new_tape(self):
        sc = self.sc
        return self.tracer.new_tape(sc.rays_dir.reshape(-1, 3) if self.ragged is not None else sc.rays_dir,
                                    sc.max_bounces)

    def fit_tape(self, max_hits: int = 4096) -> int:
        """Run eager steps, doubling the tracer's tape_hits_per_path until the backward
        replays every path from the tape; returns the capacity chosen.  The steps train the
        parameters like any other step."""
        tr = self.tracer
        while True:
            self.step(counters=True)
            torch.cuda.synchronize()
            self.tape_retraced = int(tr.last_counters["traversals"])
            if self.tape_retraced == 0 or 2 * tr.tape_hits_per_path > max_hits:
                return tr.tape_hits_per_path
            old = tr.tape_hits_per_path
            tr.tape_hits_per_path = 2 * old
            sc = self.sc
            paths = sc.rays_dir.numel() // 3
            if tr.tape_bytes(1, 1, paths, sc.max_bounces) > self.tape_max_bytes:
                tr.tape_hits_per_path = old
                return old
            self.tape = None
            self.tape = self._new_tape()

    def _kw(self):
        sc = self.sc
        return dict(render_opts=0, sph_degree=sc.sph_degree, min_transmittance=sc.min_transmittance,
                    max_bounces=sc.max_bounces, tape=self.tape, ragged=self.ragged, rr_start=self.rr_start)

    def step(self, counters: bool = False) -> None:
        """Enqueue one training step on the current stream (counters: count the backward's
        work units into tracer.last_counters, synchronising)."""
        sc, tr = self.sc, self.tracer
        envmap.build_cdf_table(sc.env.detach(), (self.table.shape[1], self.table.shape[2]), out=self.table,
                               allocate=False)
        if self.refit:
            tr.build_bvh(sc.positions, sc.rotation, sc.scale, sc.density, rebuild=False)
        material = torch.cat([sc.albedo, sc.roughness, sc.metallic], dim=1)
        args = (sc.frame, sc.ray_to_world, sc.rays_ori, sc.rays_dir, self.density, material, sc.features,
                sc.shading_normal, sc.env, self.table)
        outs = tr.trace(*args, **self._kw())
        pbr = outs[10]
        self.pbr = pbr
        # elementwise only: reductions can record memsets, which graphs replay wrongly here
        grad = ((pbr - self.target.reshape(pbr.shape)) * (2.0 / pbr.numel())).contiguous()
        d_dns, d_mat, d_sh, d_sn, d_env = tr.trace_bwd(
            sc.frame, sc.ray_to_world, sc.rays_ori, sc.rays_dir, *outs[:8], outs[10], outs[11], *args[4:], None, None,
            None, None, None, None, None, None, None, None, None, None, None, None, grad, None, counters=counters,
            **self._kw())
        self.params["albedo"].grad.copy_(d_mat[:, 0:3])
        self.params["roughness"].grad.copy_(d_mat[:, 3:4])
        self.params["metallic"].grad.copy_(d_mat[:, 4:5])
        self.params["env"].grad.copy_(d_env)
        self.adam.step(visibility=outs[9] if self.adam.selective else None)

    def capture(self) -> None:
        """Fit the tape (see fit_tape), warm up on a side stream, then capture one step as a graph."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            # warm-up (one step, more while the tape grows): the allocator sees every buffer once
            if self.fit:
                self.fit_tape()
            else:
                self.step()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.step()
        torch.cuda.synchronize()

    def loss(self) -> float:
        """L2 loss of the last step's PBR image (evaluated eagerly, outside any graph)."""
        return float((self.pbr - self.target.reshape(self.pbr.shape)).square().mean())

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("capture() the step before replaying it")
        self.graph.replay()
