"""ptir_env_table_build_noalloc: the allocation-free variant of ptir_env_table_build.

Called through the C ABI; its output must equal the allocating entry point's bit for bit
(same kernels), on two calls in a row with the second smaller than the first (the static
scratch keeps the first call's contents), and it must reject tables beyond its limit.
"""

import ctypes as C

import pytest
import torch

from ptir_amd import _lib, scenes


def _env(h, w, seed):
    g = torch.Generator().manual_seed(seed)
    return scenes.sun_sky_env(h, w, g, sun_gain=300.0)[0]


def test_noalloc_rejects_tables_beyond_its_scratch():
    L = _lib.lib()
    fake = C.c_void_p(0x1000)  # never dereferenced: the size check comes first
    assert L.ptir_env_table_build_noalloc(fake, 64, 128, 512, 512, fake, None) == 1
    assert "at most 131072 cells" in L.ptir_last_error(None).decode()
    assert L.ptir_env_table_build_noalloc(fake, 64, 128, 8192, 2, fake, None) == 1


@pytest.mark.gpu
def test_noalloc_table_equals_allocating_build_on_consecutive_calls(device):
    L = _lib.lib()
    stream = torch.cuda.current_stream(device).cuda_stream
    for (H, W, h, w, seed) in [(512, 1024, 256, 512, 1), (256, 512, 64, 128, 2)]:
        print("noalloc table", (H, W, h, w), flush=True)
        env = _env(H, W, seed).to(device)
        ref = torch.empty(3, h, w, device=device)
        got = torch.full((3, h, w), float("nan"), device=device)
        _lib.check(L.ptir_env_table_build(env.data_ptr(), H, W, h, w, ref.data_ptr(), C.c_void_p(stream)), None)
        _lib.check(L.ptir_env_table_build_noalloc(env.data_ptr(), H, W, h, w, got.data_ptr(), C.c_void_p(stream)),
                   None)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        assert bool(torch.isfinite(got).all())
