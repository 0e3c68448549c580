"""The wavefront schedule's stage code (ptir_amd/csrc/wavefront_path.hpp) against the one-lane-
per-path body (fwd_path, trace_path.hpp), both compiled for the host and run one lane at a time
(tests/host_emu/wavefront_host.cpp): per path, the primary stage, NEE and every bounce stage in
turn must write exactly what the one-lane body writes -- all 13 outputs, visibility, radiance-
cache records and the path tape (header, records, hit lists) -- bit for bit, under MIS on and
off, INDIRECT, Russian roulette, both cache modes, tapes whose hits spill into pool chunks, and
per-wave and per-lane primary traversal, over the binary and the 4-wide nodes."""

import pytest

import host_emu_util as E
from gpu_helpers import env_table
from ptir_amd import scenes

CASES = {
    "mis_tape": dict(max_bounces=3, enable_mis=True, tape=1, tape_cap=16, wave_traversal=1, wide_walks=3),
    "no_mis": dict(max_bounces=3, enable_mis=False, tape=1, tape_cap=64, wave_traversal=0, wide_walks=0),
    "indirect": dict(max_bounces=2, enable_mis=True, render_opts=1, tape=1, tape_cap=8, wave_traversal=1,
                     wide_walks=1),
    "roulette_mb5": dict(max_bounces=5, enable_mis=True, rr_start=1, tape=1, tape_cap=32, wave_traversal=1,
                         wide_walks=2),
    "cache_query": dict(max_bounces=4, enable_mis=True, cache_mode=1, tape=1, tape_cap=64, wave_traversal=1),
    "cache_train": dict(max_bounces=4, enable_mis=True, cache_mode=2, tape=0, wave_traversal=0),
    "one_vertex": dict(max_bounces=1, enable_mis=True, tape=1, tape_cap=64, wave_traversal=1),
}


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return E.build("wavefront_host.cpp", str(tmp_path_factory.mktemp("host_emu") / "wavefront_host"))


@pytest.mark.parametrize("case", list(CASES))
def test_wavefront_stages_write_what_the_one_lane_body_writes(harness, tmp_path, case):
    sc = scenes.small_scene(5, n=1024, res=24, spp=1, max_bounces=2)
    E.write_launch(str(tmp_path), sc, env_table(sc), **CASES[case])
    out = E.run(harness, str(tmp_path))
    print(out)
    assert "mismatches 0" in out
