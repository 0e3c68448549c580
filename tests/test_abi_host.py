"""CPU tests of the C ABI library (load, exports, argument checking) and the host-side
Python surface.  Nothing here launches a kernel."""

import ctypes as C
import os
import re

import pytest
import torch

from ptir_amd import _lib, envmap, scenes
from ptir_amd.tracer import HipTracer, LightBuffers, Tracer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "ptir_amd.h")).read()
    return sorted(set(re.findall(r"^(?:ptir_status|const char\*|uint64_t)\s+(ptir_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    assert _declared_symbols() == sorted(_lib.EXPORTED)


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    for name in _declared_symbols():
        assert hasattr(L, name), name
    assert b"gfx950" in L.ptir_version()


def _opts(**kw):
    o = dict(kernel_degree=4.0, min_response=0.0113, density_clamping=1, sph_max_degree=3, enable_normals=1,
             enable_hitcounts=1, enable_mis=1, enable_metallic=0, visualize_lights=0, kbuffer=16,
             wave_traversal=1, schedule=1)
    o.update(kw)
    return _lib.Options(**o)


def test_tape_size_formula():
    L = _lib.lib()
    # header 64 words + per path (16 + 16 * segments) words + 8 bytes per own hit slot, then
    # a pool of one 128-entry chunk per 64 paths (rounded up)
    assert L.ptir_tape_bytes(2, 3, 4, 3, 32) == 4 * (64 + 24 * (16 + 48)) + 8 * 24 * 32 + 8 * 128 * 1
    assert L.ptir_tape_bytes(1, 100, 100, 2, 64) == 4 * (64 + 10000 * (16 + 32)) + 8 * 10000 * 64 + 8 * 128 * 157
    assert L.ptir_tape_bytes(-1, 3, 4, 3, 32) == 0
    assert L.ptir_tape_bytes(1, 1, 1, 3, 0) == 0


def test_create_destroy_and_option_errors():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.ptir_create(C.byref(_opts()), C.byref(h)) == 0
    assert h.value
    assert L.ptir_destroy(h) == 0
    for bad, status in [(dict(kernel_degree=6.0), 2), (dict(kbuffer=12), 1), (dict(visualize_lights=1), 2),
                        (dict(min_response=0.0), 1), (dict(sph_max_degree=4), 2), (dict(wave_traversal=2), 1),
                        (dict(schedule=2), 1)]:
        h = C.c_void_p()
        assert L.ptir_create(C.byref(_opts(**bad)), C.byref(h)) == status
        assert not h.value
        assert L.ptir_last_error(None)


def test_trace_before_build_and_refit_without_build_are_errors():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.ptir_create(C.byref(_opts()), C.byref(h)) == 0
    sc = _lib.Scene()
    la = _lib.Launch()
    out = _lib.FwdOutputs()
    assert L.ptir_trace(h, C.byref(sc), C.byref(la), C.byref(out), None, None) == 3  # NOT_BUILT
    assert b"before build_bvh" in L.ptir_last_error(h)
    # refit (rebuild = 0) of a hierarchy that was never built is rejected before any GPU work
    assert L.ptir_build_bvh(h, None, None, None, None, 10, 0, 0, None) == 1
    assert L.ptir_build_bvh(h, None, None, None, None, 10, 1, 0, None) == 1  # null tensors
    # an empty scene builds without touching the GPU
    assert L.ptir_build_bvh(h, None, None, None, None, 0, 1, 0, None) == 0
    sc.num_lights = 1
    assert L.ptir_trace(h, C.byref(sc), C.byref(la), C.byref(out), None, None) == 2  # lights unsupported
    L.ptir_destroy(h)


def test_hip_tracer_rejects_unsupported_pipelines_and_host_tensors():
    with pytest.raises(ValueError):
        HipTracer(pipeline="referenceSlang")
    with pytest.raises(ValueError):
        HipTracer(primitive="icosahedron")
    with pytest.raises(_lib.PtirError):
        HipTracer(particle_kernel_degree=7)
    t = HipTracer()
    with pytest.raises(ValueError):
        t.build_bvh(torch.zeros(4, 3), torch.zeros(4, 4), torch.zeros(4, 3), torch.zeros(4, 1))


def test_light_buffers_empty_shapes():
    lb = LightBuffers.empty("cpu")
    shapes = [tuple(x.shape) for x in lb.tensors()]
    assert shapes == [(0, 9), (5, 0), (0, 3), (0, 3), (0, 8), (3, 0)]
    assert lb.mesh_triangles.dtype == torch.int32


def test_env_table_packing_layout():
    env = scenes.lognormal_env(16, 32, torch.Generator().manual_seed(1))
    t = envmap.build_environment_alias_table(env, target_size=(8, 16))
    packed = envmap.pack_alias_table(t)
    assert packed.shape == (3, 8, 16)
    torch.testing.assert_close(packed[0].reshape(-1), t.prob)
    torch.testing.assert_close(packed[1].reshape(-1), t.alias.float())
    torch.testing.assert_close(packed[2].reshape(-1), t.pdf)
    assert envmap.pack_alias_table(None).numel() == 0
