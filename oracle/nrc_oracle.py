"""FP64 restatement of the neural radiance cache (hash-grid encoding + 64-wide MLP).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's comparison leg as the checker
of the HIP kernels in ptir_amd/csrc/nrc.hip.  The product package never imports it.

The reference has no radiance cache (its tiny-cuda-nn submodule is an empty directory and
nothing imports it), so this restates the build's own specification, which the HIP kernels
follow:

* Encoding of a cache query (position, outgoing direction, shading normal, roughness,
  albedo) into 64 features:
  - [0, 32)  multiresolution hash grid: 16 levels x 2 features, 2^19 entries per level,
             resolution round(16 * 128^(l/15)) (16 ... 2048).  The position is normalised to
             the scene box; a level whose (res+1)^3 vertices fit the table is indexed densely
             (x + y (res+1) + z (res+1)^2), the others by the spatial hash
             (x * 1) ^ (y * 2654435761) ^ (z * 805459861) mod 2^19 (Teschner et al. 2003, the
             primes of Mueller et al., Instant NGP 2022).  Trilinear interpolation.
  - [32, 48) real spherical harmonics of bands 0..3 of the outgoing direction (degree 4).
  - 48 roughness, 49..51 albedo, 52..54 shading normal * 0.5 + 0.5, 55 constant 1, rest 0.
  Features are rounded to bfloat16 (round to nearest even) before the network.
* MLP: four hidden layers of width 64 with ReLU, then a 3-channel output through exp
  (exponent clamped at 15).  Weights and activations bfloat16, accumulation FP32 on the GPU
  (FP64 here).  Rounding points: inputs, each hidden activation, each back-propagated
  activation gradient.
* Loss: relative L2, mean over samples of sum over channels of (p - t)^2 / (sg(p)^2 + 0.01)
  (Mueller et al., "Real-time Neural Radiance Caching for Path Tracing", 2021).
* Adam (Kingma & Ba 2015) with bias correction on the FP32 master weights and the table.
"""

from __future__ import annotations

import math

import numpy as np

LEVELS = 16
FEATS = 2
LOG2_T = 19
T = 1 << LOG2_T
BASE_RES = 16
MAX_RES = 2048
WIDTH = 64
N_IN = 64
N_HIDDEN = 4
N_OUT = 3
OUT_PAD = 16
Z_MAX = 15.0
LOSS_EPS = 0.01
PRIMES = (1, 2654435761, 805459861)


def level_resolutions() -> list[int]:
    g = math.log(MAX_RES / BASE_RES) / (LEVELS - 1)
    return [int(math.floor(BASE_RES * math.exp(l * g) + 0.5)) for l in range(LEVELS)]


def level_is_dense(res: int) -> bool:
    return (res + 1) ** 3 <= T


def bf16(x) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    out = np.where(np.isnan(f), f, out)
    return out.astype(np.float64)


def grid_index(level_res: int, X, Y, Z) -> np.ndarray:
    X = np.asarray(X, dtype=np.uint64)
    Y = np.asarray(Y, dtype=np.uint64)
    Z = np.asarray(Z, dtype=np.uint64)
    if level_is_dense(level_res):
        s = np.uint64(level_res + 1)
        return (X + Y * s + Z * s * s).astype(np.int64)
    m32 = np.uint64(0xFFFFFFFF)
    h = ((X * np.uint64(PRIMES[0])) & m32) ^ ((Y * np.uint64(PRIMES[1])) & m32) ^ ((Z * np.uint64(PRIMES[2])) & m32)
    return (h & np.uint64(T - 1)).astype(np.int64)


def grid_corners(p01: np.ndarray, level_res: int):
    """Corner indices [n,8] and trilinear weights [n,8] of one level for positions in [0,1]^3.

    Positions are scaled in float32 as the kernel does (u = p * res)."""
    u = (p01.astype(np.float32) * np.float32(level_res)).astype(np.float32)
    i0 = np.minimum(np.floor(u), level_res - 1).astype(np.int64)
    f = (u - i0.astype(np.float32)).astype(np.float64)
    idx = np.zeros((p01.shape[0], 8), dtype=np.int64)
    w = np.zeros((p01.shape[0], 8), dtype=np.float64)
    for c in range(8):
        cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
        idx[:, c] = grid_index(level_res, i0[:, 0] + cx, i0[:, 1] + cy, i0[:, 2] + cz)
        wx = f[:, 0] if cx else 1.0 - f[:, 0]
        wy = f[:, 1] if cy else 1.0 - f[:, 1]
        wz = f[:, 2] if cz else 1.0 - f[:, 2]
        w[:, c] = wx * wy * wz
    return idx, w


def normalize_positions(pos: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    ext = np.maximum(hi.astype(np.float32) - lo.astype(np.float32), np.float32(1e-12))
    p = ((pos.astype(np.float32) - lo.astype(np.float32)) / ext).astype(np.float32)
    return np.clip(p, 0.0, 1.0).astype(np.float32)


def sh_bands4(d: np.ndarray) -> np.ndarray:
    """Real spherical harmonics of bands 0..3 (16 values) of unit directions [n,3]."""
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    xx, yy, zz = x * x, y * y, z * z
    out = np.empty((d.shape[0], 16), dtype=np.float64)
    out[:, 0] = 0.28209479177387814
    out[:, 1] = -0.48860251190291987 * y
    out[:, 2] = 0.48860251190291987 * z
    out[:, 3] = -0.48860251190291987 * x
    out[:, 4] = 1.0925484305920792 * x * y
    out[:, 5] = -1.0925484305920792 * y * z
    out[:, 6] = 0.94617469575755997 * zz - 0.31539156525251999
    out[:, 7] = -1.0925484305920792 * x * z
    out[:, 8] = 0.54627421529603959 * (xx - yy)
    out[:, 9] = 0.59004358992664352 * y * (3.0 * xx - yy)
    out[:, 10] = 2.8906114426405538 * x * y * z
    out[:, 11] = 0.45704579946446572 * y * (5.0 * zz - 1.0)
    out[:, 12] = 0.3731763325901154 * z * (5.0 * zz - 3.0)
    out[:, 13] = 0.45704579946446572 * x * (5.0 * zz - 1.0)
    out[:, 14] = 1.4453057213202769 * z * (xx - yy)
    out[:, 15] = 0.59004358992664352 * x * (xx - 3.0 * yy)
    return out


# Query record layout (float32, 20 per query), written by the tracer's cache mode.
REC_FLOATS = 20
R_FLAG, R_POS, R_DIR, R_NRM, R_ROUGH, R_ALB, R_THR, R_TARGET = 0, 1, 4, 7, 10, 11, 14, 17


def encode(records: np.ndarray, table: np.ndarray, lo: np.ndarray, hi: np.ndarray, round_bf16: bool = True):
    """Features [n,64] of query records [n,20]; table [LEVELS, T, 2]."""
    rec = np.asarray(records, dtype=np.float32)
    n = rec.shape[0]
    x = np.zeros((n, N_IN), dtype=np.float64)
    p01 = normalize_positions(rec[:, R_POS:R_POS + 3], lo, hi)
    for l, res in enumerate(level_resolutions()):
        idx, w = grid_corners(p01, res)
        for f in range(FEATS):
            x[:, 2 * l + f] = np.sum(w * table[l][idx, f].astype(np.float64), axis=1)
    x[:, 32:48] = sh_bands4(rec[:, R_DIR:R_DIR + 3].astype(np.float64))
    x[:, 48] = rec[:, R_ROUGH]
    x[:, 49:52] = rec[:, R_ALB:R_ALB + 3]
    x[:, 52:55] = rec[:, R_NRM:R_NRM + 3].astype(np.float64) * 0.5 + 0.5
    x[:, 55] = 1.0
    return bf16(x) if round_bf16 else x


def mlp_forward(x: np.ndarray, weights: list[np.ndarray], round_bf16: bool = True):
    """weights: [W0..W3] of [64,64] (out, in) and W4 of [3,64].  Returns (z [n,3], hs) where
    hs[l] is the input of layer l (hs[0] = x)."""
    rnd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    ws = [rnd(w) for w in weights]
    hs = [rnd(x)]
    h = hs[0]
    for l in range(N_HIDDEN):
        h = rnd(np.maximum(h @ ws[l].T, 0.0))
        hs.append(h)
    z = h @ ws[N_HIDDEN].T
    return z, hs


def predict(z: np.ndarray) -> np.ndarray:
    return np.exp(np.minimum(z, Z_MAX))


def loss_and_grad(z: np.ndarray, target: np.ndarray):
    """Relative L2 loss (mean over samples) and its gradient w.r.t. z."""
    n = z.shape[0]
    p = predict(z)
    t = np.asarray(target, dtype=np.float64)
    den = p * p + LOSS_EPS
    loss = float(np.sum((p - t) ** 2 / den)) / max(n, 1)
    dz = 2.0 * (p - t) * p / den / max(n, 1)
    dz = np.where(z > Z_MAX, 0.0, dz)
    return loss, dz


def mlp_backward(dz: np.ndarray, hs: list[np.ndarray], weights: list[np.ndarray], round_bf16: bool = True):
    """Gradients of the weights (list like `weights`) and of the input features [n,64]."""
    rnd = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    ws = [rnd(w) for w in weights]
    g = rnd(dz)
    dws = [None] * (N_HIDDEN + 1)
    for l in range(N_HIDDEN, -1, -1):
        dws[l] = g.T @ hs[l]
        dh = g @ ws[l]
        if l == 0:
            return dws, dh
        g = rnd(dh * (hs[l] > 0.0))
    raise AssertionError("unreachable")


def grid_backward(records: np.ndarray, dx: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Gradient of the table [LEVELS, T, 2] from feature gradients dx [n,64]."""
    rec = np.asarray(records, dtype=np.float32)
    p01 = normalize_positions(rec[:, R_POS:R_POS + 3], lo, hi)
    g = np.zeros((LEVELS, T, FEATS), dtype=np.float64)
    for l, res in enumerate(level_resolutions()):
        idx, w = grid_corners(p01, res)
        for f in range(FEATS):
            np.add.at(g[l, :, f], idx.reshape(-1), (w * dx[:, 2 * l + f:2 * l + f + 1]).reshape(-1))
    return g


def adam(param: np.ndarray, grad: np.ndarray, m: np.ndarray, v: np.ndarray, step: int, lr: float,
         beta1: float = 0.9, beta2: float = 0.99, eps: float = 1e-10):
    """One Adam step (step counts from 1); returns new (param, m, v)."""
    m = beta1 * m + (1.0 - beta1) * grad
    v = beta2 * v + (1.0 - beta2) * grad * grad
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return param - lr * mh / (np.sqrt(vh) + eps), m, v
