"""The exact-adjoint device code (ptir_amd/csrc/adjoint_path.hpp) compiled for the host with g++
and run one lane at a time (tests/host_emu/adjoint_host.cpp), against the FP64 dual-number
oracle: see tests/adjoint_cases.py for the identity checked and its tolerance.  The same
source is the kernel's on the GPU; this runs it without one (with every memory access of the
device code bounds-checked when PTIR_HOST_EMU_SANITIZE=1)."""

import os

import numpy as np
import pytest

import adjoint_cases as A
import host_emu_util as E


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return E.build("adjoint_host.cpp", str(tmp_path_factory.mktemp("host_emu") / "adjoint_host"))


@pytest.mark.parametrize("case", list(A.CASES))
def test_exact_adjoint_host_build_matches_dual_derivatives(harness, tmp_path, case):
    sc, tab, inp, rays, ref, rng = A.setup(case)
    cfg = A.CASES[case]
    keep = ref["margin"] >= A.MARGIN
    assert keep.mean() > 0.7
    g = A.output_grads(keep, rng)
    wd = str(tmp_path)
    E.write_launch(wd, sc, tab, render_opts=cfg["render_opts"], enable_mis=cfg["enable_mis"], rr_start=cfg["rr_start"])
    for k in A.OUTS:
        E.write_f32(wd, "g_" + k, g[k])
    E.run(harness, wd)
    N = sc.num_gaussians

    def rd(name, shape):
        return np.fromfile(os.path.join(wd, name + ".f32"), np.float32).reshape(shape).astype(np.float64)

    grads = {"density": rd("d_density", (N, 12)), "material": rd("d_material", (N, 5)),
             "sh": rd("d_sh", tuple(sc.features.shape)), "snormal": rd("d_snormal", (N, 3)),
             "env": rd("d_env", tuple(sc.env.shape))}
    for k, v in grads.items():
        assert np.isfinite(v).all(), k
    A.check_blocks(sc, inp, rays, g, grads, rng, case)
