"""The PLOC hierarchy builder's phases (ptir_amd/csrc/ploc.hpp) run on the host, one loop per
kernel launch (tests/host_emu/ploc_host.cpp): the tree must be well formed (every leaf and node
reached once from root 0, consistent parent links, exact child boxes), equal to a sequential
restatement of the clustering, and, on Gaussian boxes, of lower surface-area cost than a
Morton-split (LBVH) tree of the same leaves."""

import os

import numpy as np
import pytest

import host_emu_util as E


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return E.build("ploc_host.cpp", str(tmp_path_factory.mktemp("host_emu") / "ploc_host"))


def _run(harness, tmp_path, boxes):
    wd = str(tmp_path)
    with open(os.path.join(wd, "meta.txt"), "w") as fh:
        fh.write(f"n {len(boxes)}\n")
    E.write_f32(wd, "boxes", boxes)
    out = E.run(harness, wd)
    assert out.strip().endswith("ok"), out
    first = out.split()
    return {"depth": int(first[1]), "sah": float(first[3]), "iterations": int(first[5])}


def _morton_sorted(centres, half):
    lo, hi = centres.min(0), centres.max(0)
    q = np.clip(((centres - lo) / np.maximum(hi - lo, 1e-30) * 1024).astype(np.int64), 0, 1023)
    code = np.zeros(len(q), dtype=np.int64)
    for bit in range(10):
        for ax in range(3):
            code |= ((q[:, ax] >> bit) & 1) << (3 * bit + (2 - ax))
    order = np.argsort(code, kind="stable")
    return np.concatenate([centres - half, centres + half], 1)[order].astype(np.float32)


def _shell(n, seed):
    g = np.random.default_rng(seed)
    v = g.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    half = np.exp(-4.0 + 0.4 * g.normal(size=(n, 3)))
    return _morton_sorted(v, half)


def _sah_of_median_split(boxes):
    """Half-area cost of the tree that splits the Morton order in halves (an LBVH-like tree)."""
    root = np.concatenate([boxes[:, :3].min(0), boxes[:, 3:].max(0)])

    def ha(b):
        d = b[3:] - b[:3]
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0]

    total = 0.0
    stack = [(0, len(boxes))]
    while stack:
        a, b = stack.pop()
        if b - a < 2:
            continue
        sub = boxes[a:b]
        total += ha(np.concatenate([sub[:, :3].min(0), sub[:, 3:].max(0)])) / ha(root)
        m = (a + b) // 2
        stack += [(a, m), (m, b)]
    return total


@pytest.mark.parametrize("n", [2, 3, 17, 5000])
def test_ploc_random_boxes(harness, tmp_path, n):
    g = np.random.default_rng(n)
    c = g.uniform(-1, 1, size=(n, 3))
    h = g.uniform(0.001, 0.2, size=(n, 3))
    r = _run(harness, tmp_path, np.concatenate([c - h, c + h], 1).astype(np.float32))
    assert r["depth"] <= n - 1


def test_ploc_identical_boxes_tie_everywhere(harness, tmp_path):
    boxes = np.tile(np.array([[0, 0, 0, 1, 1, 1]], np.float32), (1000, 1))
    r = _run(harness, tmp_path, boxes)
    assert r["depth"] <= 20  # ties pair neighbours: a balanced tree


def test_ploc_beats_morton_halving_on_gaussian_shell(harness, tmp_path):
    boxes = _shell(20000, 3)
    r = _run(harness, tmp_path, boxes)
    base = _sah_of_median_split(boxes)
    print(f"PLOC half-area cost {r['sah']:.1f} vs Morton halving {base:.1f}, depth {r['depth']}")
    assert r["sah"] < 0.95 * base
