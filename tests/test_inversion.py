"""The whole inversion training step (env table, BVH refit, forward, loss gradient,
backward, Adam) captured as one graph: every replay equals the eager step, and the loss
goes down."""

import pytest
import torch

from gpu_helpers import rel_l2
from ptir_amd import scenes
from ptir_amd.inversion import InversionStep
from ptir_amd.tracer import HipTracer, RaggedBatch

pytestmark = pytest.mark.gpu


def _state(step):
    return [t.detach().clone() for t in (step.sc.albedo, step.sc.roughness, step.sc.env)]


def _make(device, ragged=False, tracer_kw=None, **kw_step):
    if ragged:
        sc = scenes.batched_scene(0, n=20_000, res_scale=1 / 32, env_hw=(64, 128))
        rb = RaggedBatch.build(sc.extra["images"], device)
        kw = dict(ragged=rb, rr_start=sc.extra["rr_start"])
    else:
        sc = scenes.small_scene(0, n=2048, res=32, spp=2, max_bounces=3)
        kw = {}
    g = sc.to(device)
    target = scenes.target_image(sc).to(device)
    tracer = HipTracer(enable_metallic=True, env_sampling="cdf", **(tracer_kw or {}))
    return InversionStep(tracer, g, target, **kw, **kw_step)


@pytest.mark.parametrize("ragged", [False, True])
def test_graph_replays_equal_eager_steps(device, ragged):
    """Gradient atomics add in a launch-dependent order, and Adam with a tiny eps turns the
    rounding noise of near-zero gradients into full +-lr steps; a large eps keeps the update
    proportional to the gradient, so graph and eager steps must agree to float rounding."""
    opt = dict(adam_eps=1.0, lr_albedo=0.5, lr_roughness=0.5, lr_environment=0.5, fit_tape=False)
    eager = _make(device, ragged, **opt)
    graph = _make(device, ragged, **opt)
    ref = []
    for _ in range(3):
        eager.step()
        torch.cuda.synchronize()
        ref.append(_state(eager))
    graph.capture()  # the warm-up inside performs step 1
    got = [_state(graph)]
    for _ in range(2):
        graph.replay()
        torch.cuda.synchronize()
        got.append(_state(graph))
        _ = (torch.randn(4096, device=device) * 3).sum().item()  # eager work between replays
    for k, (a, b) in enumerate(zip(got, ref)):
        for x, y in zip(a, b):
            # gradient atomics add in a different order: equal up to float rounding
            assert rel_l2(x.cpu().numpy(), y.cpu().numpy()) < 1e-5, k
    assert int(graph.adam.step_count.item()) == 3


def test_inversion_reduces_the_loss(device):
    step = _make(device)
    step.capture()
    first = step.loss()
    for _ in range(20):
        step.replay()
    torch.cuda.synchronize()
    last = step.loss()
    print(f"inversion loss {first:.5f} -> {last:.5f}", flush=True)
    assert last < 0.9 * first


@pytest.mark.parametrize("ragged", [False, True])
def test_fit_tape_grows_until_nothing_is_retraced(device, ragged):
    """A tape with room for 2 hits per path (and one pool chunk per 64 paths) leaves many
    paths to the backward's re-trace; fitting doubles the capacity until none is left, and
    the captured step then replays from the tape alone."""
    step = _make(device, ragged, tracer_kw=dict(tape_hits_per_path=2))
    tr = step.tracer
    step.step(counters=True)
    torch.cuda.synchronize()
    before = int(tr.last_counters["traversals"])
    print(f"re-traced segments at 2 hits per path: {before}", flush=True)
    assert before > 0
    cap = step.fit_tape()
    assert cap == tr.tape_hits_per_path and cap > 2 and cap & (cap - 1) == 0
    assert step.tape_retraced == 0
    step.fit = False
    step.capture()
    step.replay()
    step.step(counters=True)  # eager step after the replays: the tape still covers every path
    torch.cuda.synchronize()
    assert int(tr.last_counters["traversals"]) == 0
