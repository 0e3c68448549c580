"""GPU tests of the neural radiance cache: the HIP kernels against the FP64 oracle of
oracle/nrc_oracle.py (same bf16-rounded weights and features), the tracer's cache modes
against the plain tracer, and what a trained cache gains over the tracer without it.

Tolerances.  The GPU accumulates in FP32 and evaluates the encoding in FP32; the oracle
rounds to bf16 at the same points but computes in FP64.  A feature that lands on a bf16
rounding boundary can differ by one bf16 ulp (2^-8 relative), so outputs are held to
5e-3 relative for 99% of samples and 5e-2 for all, and gradients to 1e-2 relative L2.
"""

import numpy as np
import pytest
import torch

from gpu_helpers import env_table, rel_l2
from oracle import nrc_oracle as NO
from ptir_amd import scenes
from ptir_amd.cache import RadianceCache
from ptir_amd.tracer import HipTracer

pytestmark = pytest.mark.gpu

BOX = np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0], dtype=np.float32)


def _records(n, seed, flag_frac=1.0):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, NO.REC_FLOATS), dtype=np.float32)
    rec[:, NO.R_FLAG] = (rng.uniform(size=n) < flag_frac).astype(np.float32)
    rec[:, NO.R_POS:NO.R_POS + 3] = rng.uniform(-1.05, 1.05, (n, 3))  # a few outside the box (clamped)
    for k in (NO.R_DIR, NO.R_NRM):
        d = rng.standard_normal((n, 3))
        rec[:, k:k + 3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rec[:, NO.R_ROUGH] = rng.uniform(0.05, 1.0, n)
    rec[:, NO.R_ALB:NO.R_ALB + 3] = rng.uniform(0, 1, (n, 3))
    rec[:, NO.R_THR:NO.R_THR + 3] = rng.uniform(0, 2, (n, 3))
    rec[:, NO.R_TARGET:NO.R_TARGET + 3] = np.exp(rng.standard_normal((n, 3)))
    return rec


def _params(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(NO.N_HIDDEN * 4096 + 16 * 64, dtype=np.float32)
    w[:NO.N_HIDDEN * 4096] = rng.uniform(-1, 1, NO.N_HIDDEN * 4096) * np.sqrt(6 / 64)
    w[NO.N_HIDDEN * 4096:NO.N_HIDDEN * 4096 + 3 * 64] = rng.uniform(-1, 1, 3 * 64) * 0.1
    table = rng.uniform(-1, 1, (NO.LEVELS, NO.T, 2)).astype(np.float32)
    return w, table


def _weights_list(w):
    ws = [w[l * 4096:(l + 1) * 4096].reshape(64, 64).astype(np.float64) for l in range(NO.N_HIDDEN)]
    ws.append(w[NO.N_HIDDEN * 4096:NO.N_HIDDEN * 4096 + 3 * 64].reshape(3, 64).astype(np.float64))
    return ws


def _oracle(rec, w, table):
    x = NO.encode(rec, table, BOX[:3], BOX[3:])
    z, hs = NO.mlp_forward(x, _weights_list(w))
    return z, hs


@pytest.fixture(scope="module")
def loaded(device):
    w, table = _params(11)
    c = RadianceCache(max_batch=4096, seed=5)
    c.write_params(torch.from_numpy(w).to(device), torch.from_numpy(table).to(device))
    return c, w, table


def test_query_matches_oracle(loaded, device):
    c, w, table = loaded
    rec = _records(3001, 1, flag_frac=0.8)   # not a multiple of the 64-sample tile
    rt = torch.from_numpy(rec).to(device)
    box = torch.from_numpy(BOX).to(device)
    pbr = torch.rand(3001, 3, device=device)
    comp = torch.rand(3001, 2, 3, device=device)
    pbr0, comp0 = pbr.clone(), comp.clone()
    pred = c.query(rt, box, pbr=pbr, pbr_components=comp, prediction=True)
    torch.cuda.synchronize()
    z, _ = _oracle(rec, w, table)
    ref = NO.predict(z)
    on = rec[:, NO.R_FLAG] != 0
    g = pred.double().cpu().numpy()
    rel = np.abs(g[on] - ref[on]) / np.abs(ref[on])
    print(f"query: median rel {np.median(rel):.2e}, p99 {np.quantile(rel, 0.99):.2e}, max {rel.max():.2e}", flush=True)
    assert np.quantile(rel, 0.99) < 5e-3 and rel.max() < 5e-2
    assert np.all(g[~on] == 0)  # rows with flag 0 are left as they were
    thr = rec[:, NO.R_THR:NO.R_THR + 3]
    add = np.where(on[:, None], thr * g, 0.0)
    assert np.allclose(pbr.cpu().numpy(), pbr0.cpu().numpy() + add, rtol=1e-6, atol=1e-6)
    cc = comp.cpu().numpy()
    assert np.allclose(cc[:, 1], comp0[:, 1].cpu().numpy() + add, rtol=1e-6, atol=1e-6)
    assert np.array_equal(cc[:, 0], comp0[:, 0].cpu().numpy())


def test_backprop_matches_oracle(loaded, device):
    c, w, table = loaded
    n = 2050
    rec = _records(n, 2, flag_frac=0.9)
    c.backprop(torch.from_numpy(rec).to(device), torch.from_numpy(BOX).to(device))
    dw, dtab, loss = c.read_grads()
    torch.cuda.synchronize()
    z, hs = _oracle(rec, w, table)
    on = rec[:, NO.R_FLAG] != 0
    t = rec[:, NO.R_TARGET:NO.R_TARGET + 3]
    l_on, dz = NO.loss_and_grad(z[on], t[on])
    dz_all = np.zeros_like(z)
    dz_all[on] = dz * on.sum() / n  # the loss is a mean over all n records
    ref_loss = l_on * on.sum() / n
    dws, dx = NO.mlp_backward(dz_all, hs, _weights_list(w))
    g = dw.double().cpu().numpy()
    errs = []
    for l in range(NO.N_HIDDEN):
        errs.append(rel_l2(g[l * 4096:(l + 1) * 4096].reshape(64, 64), dws[l]))
    g4 = g[NO.N_HIDDEN * 4096:].reshape(16, 64)
    errs.append(rel_l2(g4[:3], dws[4]))
    assert np.all(g4[3:] == 0)
    gt = NO.grid_backward(rec, dx, BOX[:3], BOX[3:])
    etab = rel_l2(dtab.double().cpu().numpy(), gt)
    print(f"backprop: layer rel L2 {['%.1e' % e for e in errs]}, table {etab:.1e}, loss {float(loss):.5f} vs "
          f"{ref_loss:.5f}", flush=True)
    assert max(errs) < 1e-2 and etab < 1e-2
    assert abs(float(loss) - ref_loss) <= 1e-3 * ref_loss


def test_adam_step_matches_oracle_and_clears_gradients(device):
    w, table = _params(12)
    c = RadianceCache(max_batch=1024, seed=1, learning_rate=1e-3)
    wt, tt = torch.from_numpy(w).to(device), torch.from_numpy(table).to(device)
    c.write_params(wt, tt)
    rec = torch.from_numpy(_records(1024, 3)).to(device)
    c.backprop(rec, torch.from_numpy(BOX).to(device))
    dw, dtab, _ = c.read_grads()
    c.step()
    w1, t1 = c.read_params()
    dw2, dtab2, _ = c.read_grads()
    torch.cuda.synchronize()
    ref_w, _, _ = NO.adam(w.astype(np.float64), dw.double().cpu().numpy(), 0.0, 0.0, 1, 1e-3)
    ref_t, _, _ = NO.adam(table.astype(np.float64), dtab.double().cpu().numpy(), 0.0, 0.0, 1, 1e-3)
    assert np.allclose(w1.double().cpu().numpy(), ref_w, rtol=1e-5, atol=1e-7)
    assert np.allclose(t1.double().cpu().numpy(), ref_t, rtol=1e-5, atol=1e-7)
    assert float(dw2.abs().max()) == 0.0 and float(dtab2.abs().max()) == 0.0


def test_flag0_records_do_not_poison_training(device):
    """Records of paths that never reached the cache vertex hold whatever their buffer held;
    training must ignore them: NaN/inf-filled flag-0 rows give the gradients of zero-filled
    ones, and they are finite."""
    w, table = _params(13)
    rec = _records(2048, 4)
    rec[::3, NO.R_FLAG] = 0.0
    clean = rec.copy()
    clean[::3, 1:] = 0.0
    dirty = rec.copy()
    dirty[::3, 1:] = np.nan
    dirty[1::6, 1:] = np.where(dirty[1::6, NO.R_FLAG:NO.R_FLAG + 1] == 0.0, np.inf, dirty[1::6, 1:])
    grads = []
    for r in (clean, dirty):
        c = RadianceCache(max_batch=2048, seed=3)
        c.write_params(torch.from_numpy(w).to(device), torch.from_numpy(table).to(device))
        c.backprop(torch.from_numpy(r).to(device), torch.from_numpy(BOX).to(device))
        dw, dtab, loss = c.read_grads()
        torch.cuda.synchronize()
        grads.append((dw.cpu(), dtab.cpu(), loss.cpu()))
    for k, (a, b) in enumerate(zip(grads[0], grads[1])):
        assert bool(torch.isfinite(b).all())
        if k == 0:  # weight gradients: per-tile products reduced in a fixed order
            assert torch.equal(a, b)
        else:  # table gradient and loss: float atomics, equal up to their order
            torch.testing.assert_close(a, b, rtol=1e-5, atol=1e-9)


def test_training_fits_a_smooth_radiance_field(device):
    c = RadianceCache(max_batch=8192, seed=2)
    box = torch.from_numpy(BOX).to(device)

    def batch(seed):
        rec = _records(8192, seed)
        p, d = rec[:, NO.R_POS:NO.R_POS + 3], rec[:, NO.R_DIR:NO.R_DIR + 3]
        f = 0.6 + 0.4 * np.sin(3 * p[:, 0]) * np.cos(2 * p[:, 1]) + 0.3 * d[:, 2] * rec[:, NO.R_ALB:NO.R_ALB + 1].T[0]
        rec[:, NO.R_TARGET:NO.R_TARGET + 3] = (f[:, None] * np.array([1.0, 0.7, 0.4])).astype(np.float32)
        return rec

    losses = []
    for s in range(300):
        c.train(torch.from_numpy(batch(100 + s)).to(device), box)
        if s in (0, 299):
            _, _, loss = c.read_grads()
            losses.append(float(loss))
    held = batch(999)
    pred = c.query(torch.from_numpy(held).to(device), box, prediction=True).cpu().numpy()
    t = held[:, NO.R_TARGET:NO.R_TARGET + 3]
    rel = float(np.median(np.abs(pred - t) / t))
    print(f"training: loss {losses[0]:.4f} -> {losses[1]:.5f}, held-out median rel error {rel:.3f}", flush=True)
    assert losses[1] < 0.05 * losses[0]
    assert rel < 0.08


# ----------------------------------------------------------------------------- tracer cache modes
@pytest.fixture(scope="module")
def small_cache_scene():
    sc = scenes.small_scene(0, res=48, spp=2, max_bounces=3)
    return sc, env_table(sc)


def _trace(tr, sc, tab, device, mb, cache_mode=0, frame=0):
    g = sc.to(device)
    args = (frame, g.ray_to_world, g.rays_ori, g.rays_dir, g.particle_density(), g.particle_material(), g.features,
            g.shading_normal, g.env, tab.to(device))
    rec = torch.zeros(*g.rays_dir.shape[:3], 20, device=device) if cache_mode else None
    outs = tr.trace(*args, sph_degree=3, max_bounces=mb, cache_mode=cache_mode, cache_records=rec)
    return outs, args, rec


def test_trace_cache_modes_are_consistent_with_the_plain_tracer(small_cache_scene, device):
    sc, tab = small_cache_scene
    tr = HipTracer(enable_metallic=True)
    g = sc.to(device)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    plain3, _, _ = _trace(tr, sc, tab, device, 3)
    plain2, _, _ = _trace(tr, sc, tab, device, 2)
    rec2, _, r2 = _trace(tr, sc, tab, device, 3, cache_mode=2)
    rec1, _, r1 = _trace(tr, sc, tab, device, 3, cache_mode=1)
    torch.cuda.synchronize()
    # recording does not change the path; ending at the first secondary vertex is maxBounces 2
    assert torch.equal(rec2[10], plain3[10]) and torch.equal(rec2[12], plain3[12])
    assert torch.equal(rec1[10], plain2[10]) and torch.equal(rec1[12], plain2[12])
    # the same vertex is recorded in both modes, and prefix + throughput * target = full path
    f1, f2 = r1[..., 0], r2[..., 0]
    assert torch.equal(f1, f2) and float(f1.sum()) > 100
    assert torch.equal(r1[..., 1:17], r2[..., 1:17]) or torch.allclose(r1[..., 1:17], r2[..., 1:17])
    on = f2 != 0
    thr, tgt = r2[..., 14:17], r2[..., 17:20]
    recon = rec1[10] + torch.where(on[..., None], thr * tgt, torch.zeros_like(tgt))
    assert torch.allclose(recon, plain3[10], rtol=1e-4, atol=1e-5)
    # records hold unit directions and normals and clamped roughness
    d = r2[on][:, 4:7].norm(dim=1)
    assert torch.allclose(d, torch.ones_like(d), atol=1e-4)
    assert float(r2[on][:, 10].min()) >= 0.05


def test_cache_mode_backward_ends_at_the_cache_vertex(small_cache_scene, device):
    sc, tab = small_cache_scene
    tr = HipTracer(enable_metallic=True)
    g = sc.to(device)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    outs, args, _ = _trace(tr, sc, tab, device, 3, cache_mode=1)
    grad = torch.rand_like(outs[10])

    def bwd(mb, mode):
        frame, T, ro, rd = args[:4]
        return tr.trace_bwd(frame, T, ro, rd, *outs[:8], outs[10], outs[11], *args[4:], None, None, None, None, None,
                            None, None, None, None, None, None, None, None, None, grad, None, sph_degree=3,
                            max_bounces=mb, cache_mode=mode)

    a = bwd(3, 1)
    b = bwd(2, 0)
    c = bwd(3, 0)
    torch.cuda.synchronize()
    # gradients are summed by float atomics, so equal up to the order of the additions
    for k in (1, 4):
        assert rel_l2(a[k].cpu().numpy(), b[k].cpu().numpy()) < 1e-5, k
    assert float(a[1].abs().sum()) > 0
    # without the cache mode the replay would go on past the cache vertex
    assert rel_l2(c[1].cpu().numpy(), a[1].cpu().numpy()) > 1e-3


def test_trained_cache_recovers_the_light_beyond_the_first_bounce(device):
    """What the cache gains.  With paths cut at the first secondary vertex, the trained cache
    supplies the radiance of the remaining bounces.  The small scene (random Gaussians in a
    cube under a log-normal sky) carries about 10% of its light beyond that vertex; against a
    512-spp image of the plain tracer with maxBounces 4, the cache's 256-spp mean image is
    much closer than the cut tracer's, and at 1 spp its error is below both the cut and the
    full plain tracer's (whose later bounces add variance)."""
    sc = scenes.small_scene(0, res=96, spp=1, max_bounces=4)
    tab = env_table(sc)
    tr = HipTracer(enable_metallic=True)
    g = sc.to(device)
    tr.build_bvh(g.positions, g.rotation, g.scale, g.density)
    box = torch.cat([g.positions.min(0).values - 0.05, g.positions.max(0).values + 0.05]).contiguous()
    cache = RadianceCache(max_batch=96 * 96, seed=0)
    for s in range(400):
        _, _, rec = _trace(tr, sc, tab, device, 4, cache_mode=2, frame=100_000 + s)
        cache.train(rec.view(-1, 20), box)

    def mean_image(fn, frames):
        acc = torch.zeros(96, 96, 3, dtype=torch.float64, device=device)
        for f in frames:
            acc += fn(f).double()
        return acc / len(frames)

    def plain(mb):
        return lambda f: _trace(tr, sc, tab, device, mb, frame=f)[0][10][0]

    def cached(f):
        outs, _, rec = _trace(tr, sc, tab, device, 4, cache_mode=1, frame=f)
        pbr = outs[10]
        cache.query(rec.view(-1, 20), box, pbr=pbr)
        return pbr[0]

    ref = mean_image(plain(4), range(1000, 1512))
    mse = lambda img: float(((img - ref) ** 2).mean())  # noqa: E731
    bias_cut = mse(mean_image(plain(2), range(256)))
    bias_cache = mse(mean_image(cached, range(256)))
    frames = range(2000, 2016)
    one_plain = np.mean([mse(plain(4)(f).double()) for f in frames])
    one_cut = np.mean([mse(plain(2)(f).double()) for f in frames])
    one_cache = np.mean([mse(cached(f).double()) for f in frames])
    print(f"cache gain: 256-spp MSE cut {bias_cut:.3e} vs cache {bias_cache:.3e}; 1-spp MSE plain {one_plain:.3e}, "
          f"cut {one_cut:.3e}, cache {one_cache:.3e}", flush=True)
    assert bias_cache < 0.6 * bias_cut
    assert one_cache < one_plain and one_cache < one_cut
