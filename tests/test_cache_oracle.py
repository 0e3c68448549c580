"""Pins of the FP64 radiance-cache oracle (oracle/nrc_oracle.py) on the CPU.

The reference has no radiance cache, so the oracle restates the build's own
specification; these checks tie each of its parts to something independent of it:
torch's bfloat16 conversion, SH orthonormality by quadrature, hand-computed hash
indices, exact trilinear identities, torch float64 autograd, central differences and
torch.optim.Adam.
"""

import math

import numpy as np
import pytest
import torch

from oracle import nrc_oracle as N


def test_bf16_rounding_matches_torch_and_ties_to_even():
    assert N.bf16(1.0 + 2.0 ** -8) == 1.0                      # halfway, mantissa 0 is even
    assert N.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6       # halfway, rounds up to even
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(N.bf16(x), ref)


def test_level_resolutions_are_geometric_from_16_to_2048():
    res = N.level_resolutions()
    assert len(res) == 16 and res[0] == 16 and res[-1] == 2048
    ratio = (2048 / 16) ** (1 / 15)
    for l, r in enumerate(res):
        assert abs(r - 16 * ratio ** l) <= 0.5
    dense = [N.level_is_dense(r) for r in res]
    assert dense == [True] * 5 + [False] * 11  # (58+1)^3 fits 2^19 entries, (81)^3 does not


def test_hash_and_dense_indices_by_hand():
    # hashed level: (x * 1) xor (y * 2654435761 mod 2^32) xor (z * 805459861 mod 2^32), low 19 bits
    x, y, z = 5, 7, 11
    h = (x ^ ((y * 2654435761) % 2**32) ^ ((z * 805459861) % 2**32)) % 2**19
    assert int(N.grid_index(2048, x, y, z)) == h
    # dense level of resolution 16: 17 vertices per axis
    assert int(N.grid_index(16, 3, 4, 5)) == 3 + 4 * 17 + 5 * 17 * 17


def test_sh_bands_are_orthonormal():
    # Gauss-Legendre in cos(theta) x uniform phi integrates products of degree <= 6 exactly
    xs, ws = np.polynomial.legendre.leggauss(16)
    nphi = 32
    phi = (np.arange(nphi) + 0.5) * 2 * math.pi / nphi
    ct, ph = np.meshgrid(xs, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    d = np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3)
    w = (ws[:, None] * np.full(nphi, 2 * math.pi / nphi)[None, :]).reshape(-1)
    Y = N.sh_bands4(d)
    gram = (Y * w[:, None]).T @ Y
    assert np.allclose(gram, np.eye(16), atol=1e-12)
    # band 0 and band 1 in closed form
    assert np.allclose(Y[:, 0], 0.5 / math.sqrt(math.pi))
    assert np.allclose(Y[:, 2], math.sqrt(3 / (4 * math.pi)) * d[:, 2])


def test_trilinear_weights_partition_unity_and_hit_vertices_exactly():
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, (500, 3)).astype(np.float32)
    for res in (16, 58, 2048):
        _, w = N.grid_corners(p, res)
        assert np.allclose(w.sum(1), 1.0, atol=1e-12)
    # at a vertex of a dense level the lookup returns the table entry itself
    res = 16
    table = rng.standard_normal((16, N.T, 2))
    v = np.array([[3, 5, 7]])
    p = (v / res).astype(np.float32)
    idx, w = N.grid_corners(p, res)
    f = (w[:, :, None] * table[0][idx]).sum(1)
    assert np.allclose(f[0], table[0][N.grid_index(res, 3, 5, 7)], atol=1e-12)


def test_dense_level_reproduces_linear_fields():
    # a table linear in the vertex coordinates is interpolated exactly
    res = 30
    s = res + 1
    ii = np.arange(s ** 3)
    X, Y, Z = ii % s, (ii // s) % s, ii // (s * s)
    tab = np.zeros((N.T, 2))
    tab[:s ** 3, 0] = 0.3 * X - 0.2 * Y + 0.7 * Z
    tab[:s ** 3, 1] = 1.0
    rng = np.random.default_rng(2)
    p = rng.uniform(0, 1, (300, 3)).astype(np.float32)
    idx, w = N.grid_corners(p, res)
    f = (w[:, :, None] * tab[idx]).sum(1)
    u = p.astype(np.float64) * res
    assert np.allclose(f[:, 0], 0.3 * u[:, 0] - 0.2 * u[:, 1] + 0.7 * u[:, 2], atol=1e-5)
    assert np.allclose(f[:, 1], 1.0)


def _random_weights(rng, scale=0.3):
    ws = [rng.standard_normal((64, 64)) * scale for _ in range(4)]
    ws.append(rng.standard_normal((3, 64)) * scale)
    return ws


def test_mlp_backward_matches_torch_autograd():
    rng = np.random.default_rng(3)
    n = 257
    x = rng.standard_normal((n, 64))
    ws = _random_weights(rng)
    t = np.abs(rng.standard_normal((n, 3)))
    z, hs = N.mlp_forward(x, ws, round_bf16=False)
    loss, dz = N.loss_and_grad(z, t)
    dws, dx = N.mlp_backward(dz, hs, ws, round_bf16=False)

    tw = [torch.tensor(w, requires_grad=True) for w in ws]
    tx = torch.tensor(x, requires_grad=True)
    h = tx
    for l in range(4):
        h = torch.relu(h @ tw[l].T)
    tz = h @ tw[4].T
    p = torch.exp(torch.clamp(tz, max=N.Z_MAX))
    tl = (((p - torch.tensor(t)) ** 2) / (p.detach() ** 2 + N.LOSS_EPS)).sum() / n
    tl.backward()
    assert math.isclose(loss, float(tl.detach()), rel_tol=1e-12)
    for a, b in zip(dws, tw):
        assert np.allclose(a, b.grad.numpy(), rtol=1e-10, atol=1e-14)
    assert np.allclose(dx, tx.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_mlp_gradient_central_differences():
    # the stop-gradient denominator is part of the loss definition; FD on a loss whose
    # denominator is frozen at the base point checks the numerator's derivative
    rng = np.random.default_rng(4)
    n = 33
    x = rng.standard_normal((n, 64))
    ws = _random_weights(rng)
    t = np.abs(rng.standard_normal((n, 3)))
    z0, hs = N.mlp_forward(x, ws, round_bf16=False)
    p0 = N.predict(z0)
    den = p0 * p0 + N.LOSS_EPS

    def frozen_loss(wl):
        z, _ = N.mlp_forward(x, wl, round_bf16=False)
        return float(np.sum((N.predict(z) - t) ** 2 / den)) / n

    _, dz = N.loss_and_grad(z0, t)
    dws, _ = N.mlp_backward(dz, hs, ws, round_bf16=False)
    eps = 1e-6
    for l, (i, j) in [(0, (3, 5)), (2, (10, 60)), (4, (1, 7)), (3, (63, 0))]:
        wp = [w.copy() for w in ws]
        wm = [w.copy() for w in ws]
        wp[l][i, j] += eps
        wm[l][i, j] -= eps
        fd = (frozen_loss(wp) - frozen_loss(wm)) / (2 * eps)
        assert math.isclose(fd, dws[l][i, j], rel_tol=1e-5, abs_tol=1e-9)


def test_grid_backward_is_the_adjoint_of_the_lookup():
    # <dL/dtable, delta> = <dx, d(features)> for any table perturbation (the lookup is linear)
    rng = np.random.default_rng(5)
    n = 64
    rec = np.zeros((n, N.REC_FLOATS), dtype=np.float32)
    rec[:, N.R_POS:N.R_POS + 3] = rng.uniform(-1, 1, (n, 3))
    lo, hi = np.array([-1, -1, -1], np.float32), np.array([1, 1, 1], np.float32)
    dx = np.zeros((n, 64))
    dx[:, :32] = rng.standard_normal((n, 32))
    g = N.grid_backward(rec, dx, lo, hi)
    delta = rng.standard_normal((16, N.T, 2))
    f = N.encode(rec, delta, lo, hi, round_bf16=False)
    assert math.isclose(float(np.sum(g * delta)), float(np.sum(dx[:, :32] * f[:, :32])), rel_tol=1e-9)


def test_adam_matches_torch_optim():
    rng = np.random.default_rng(6)
    p0 = rng.standard_normal(100)
    grads = [rng.standard_normal(100) for _ in range(5)]
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.99), eps=1e-10)
    p, m, v = p0.copy(), np.zeros(100), np.zeros(100)
    for k, g in enumerate(grads, start=1):
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = N.adam(p, g, m, v, k, 1e-2)
    assert np.allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_encode_layout():
    rng = np.random.default_rng(7)
    n = 10
    rec = np.zeros((n, N.REC_FLOATS), dtype=np.float32)
    rec[:, N.R_POS:N.R_POS + 3] = rng.uniform(-1, 1, (n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, N.R_DIR:N.R_DIR + 3] = d
    rec[:, N.R_NRM:N.R_NRM + 3] = d[:, ::-1]
    rec[:, N.R_ROUGH] = 0.3
    rec[:, N.R_ALB:N.R_ALB + 3] = [0.1, 0.2, 0.9]
    x = N.encode(rec, np.zeros((16, N.T, 2)), np.full(3, -1, np.float32), np.full(3, 1, np.float32),
                 round_bf16=False)
    assert np.all(x[:, :32] == 0)
    assert np.allclose(x[:, 32:48], N.sh_bands4(rec[:, N.R_DIR:N.R_DIR + 3].astype(np.float64)))
    assert np.allclose(x[:, 48], 0.3, atol=1e-7) and np.allclose(x[:, 49:52], [0.1, 0.2, 0.9], atol=1e-7)
    assert np.allclose(x[:, 52:55], rec[:, N.R_NRM:N.R_NRM + 3] * 0.5 + 0.5)
    assert np.all(x[:, 55] == 1.0) and np.all(x[:, 56:] == 0.0)
