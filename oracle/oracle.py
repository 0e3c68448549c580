"""ctypes front end of the CPU oracle (oracle/ptir_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product package ptir_amd never imports it.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> str:
    subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)
    return LIB_PATH


class OracleScene(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("density", C.c_void_p), ("material", C.c_void_p), ("sh", C.c_void_p),
        ("sh_stride", C.c_int32), ("snormal", C.c_void_p),
        ("env_h", C.c_int32), ("env_w", C.c_int32), ("env", C.c_void_p),
        ("tab_h", C.c_int32), ("tab_w", C.c_int32), ("tab", C.c_void_p),
        ("kernel_degree", C.c_float), ("min_response", C.c_float), ("density_clamping", C.c_int32),
        ("sph_degree", C.c_int32), ("enable_normals", C.c_int32), ("enable_hitcounts", C.c_int32),
        ("enable_mis", C.c_int32), ("enable_metallic", C.c_int32), ("render_opts", C.c_int32),
        ("min_transmittance", C.c_float), ("max_bounces", C.c_uint32), ("frame", C.c_uint32), ("kbuf", C.c_int32),
        ("tab_kind", C.c_int32), ("rr_start", C.c_uint32),
        ("t_density", C.c_void_p), ("t_material", C.c_void_p), ("t_sh", C.c_void_p), ("t_snormal", C.c_void_p),
        ("t_env", C.c_void_p),
    ]


class OracleRays(C.Structure):
    _fields_ = [("B", C.c_int32), ("H", C.c_int32), ("W", C.c_int32), ("ray_to_world", C.c_void_p),
                ("ori", C.c_void_p), ("dir", C.c_void_p), ("path_stride", C.c_int64), ("path_offset", C.c_int64),
                ("z_offset", C.c_int64)]


_FWD_KEYS = ["rad", "dns", "hit", "hit2", "distortion", "nrm", "snrm", "mat", "hits", "pbr", "light", "comp",
             "visibility"]
_FWD_CH = {"rad": 3, "dns": 1, "hit": 2, "hit2": 1, "distortion": 1, "nrm": 3, "snrm": 3, "mat": 5, "hits": 1,
           "pbr": 3, "light": 3, "comp": 6}


class OracleFwdOut(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in _FWD_KEYS] + [("margin", C.c_void_p)]


class OracleBwdIn(C.Structure):
    _fields_ = [("pbr", C.c_void_p), ("light", C.c_void_p), ("pbr_grad", C.c_void_p), ("light_grad", C.c_void_p),
                ("mat_grad", C.c_void_p)]


class OracleBwdOut(C.Structure):
    _fields_ = [("dmat", C.c_void_p), ("denv", C.c_void_p)]


class OracleGrtIn(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ["rad", "dns", "hit", "hit2", "distortion", "snrm", "rad_grad", "dns_grad",
                                           "hit_grad", "hit2_grad", "distortion_grad", "snrm_grad"]] + \
               [("inclusive_end", C.c_int32)]


class OracleGrtOut(C.Structure):
    _fields_ = [("ddensity", C.c_void_p), ("dsh", C.c_void_p), ("dsnormal", C.c_void_p)]


class OracleCounters(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ["paths", "segments", "candidates", "accepted", "accepted_primary",
                                          "shadow_candidates", "env_lookups", "mat_atomics", "env_atomics"]]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.oracle_trace.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleRays), C.POINTER(OracleFwdOut), C.c_int,
                                   C.c_int, C.POINTER(OracleCounters)]
        L.oracle_trace_bwd.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleRays), C.POINTER(OracleBwdIn),
                                       C.POINTER(OracleBwdOut), C.c_int, C.c_int, C.POINTER(OracleCounters)]
        L.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]
        L.oracle_launch_seed.argtypes = [C.c_uint32] * 4
        L.oracle_launch_seed.restype = C.c_uint64
        L.oracle_response.argtypes = [C.c_int, C.c_double]
        L.oracle_response.restype = C.c_double
        L.oracle_kernel_scale.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double]
        L.oracle_kernel_scale.restype = C.c_double
        L.oracle_brdf_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_int,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_brdf_sample.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p,
                                         C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_env_lookup.argtypes = [C.POINTER(OracleScene), C.c_void_p, C.c_void_p]
        L.oracle_env_pdf.argtypes = [C.POINTER(OracleScene), C.c_void_p]
        L.oracle_env_pdf.restype = C.c_double
        L.oracle_env_sample.argtypes = [C.POINTER(OracleScene), C.c_uint64, C.c_int, C.c_void_p, C.c_void_p]
        L.oracle_candidates.argtypes = [C.POINTER(OracleScene), C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                        C.c_void_p]
        L.oracle_candidates.restype = C.c_int
        L.oracle_scene_aabb.argtypes = [C.POINTER(OracleScene), C.c_void_p]
        L.oracle_grt_trace.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleRays), C.POINTER(OracleFwdOut), C.c_int]
        L.oracle_grt_trace_dual.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleRays), C.POINTER(OracleFwdOut),
                                            C.POINTER(OracleFwdOut), C.c_int]
        L.oracle_grt_trace_bwd.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleRays), C.POINTER(OracleGrtIn),
                                           C.POINTER(OracleGrtOut), C.c_int]
        L.oracle_trace_dual.argtypes = [C.POINTER(OracleScene), C.POINTER(OracleRays), C.POINTER(OracleFwdOut),
                                        C.POINTER(OracleFwdOut), C.c_int]
        _lib = L
    return _lib


def _f32(x):
    if x is None:
        return None
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def _ptr(a):
    return None if a is None else a.ctypes.data


class Inputs:
    """Holds the float32 arrays alive for the duration of oracle calls."""

    def __init__(self, density, material, sh, snormal, env, table, *, sph_degree=3, kernel_degree=4.0,
                 min_response=0.0113, density_clamping=True, enable_normals=True, enable_hitcounts=True,
                 enable_mis=True, enable_metallic=False, render_opts=0, min_transmittance=1e-3, max_bounces=3,
                 frame=0, kbuf=16, tab_kind=0, rr_start=0):
        self.density = _f32(density).reshape(-1, 12)
        self.material = _f32(material).reshape(-1, 5)
        self.sh = _f32(sh).reshape(self.density.shape[0], -1)
        self.snormal = None if snormal is None else _f32(snormal).reshape(-1, 3)
        self.env = None if env is None or np.asarray(env.shape).prod() == 0 else _f32(env)
        self.table = None if table is None or np.asarray(table.shape).prod() == 0 else _f32(table)
        s = OracleScene()
        s.n = self.density.shape[0]
        s.density, s.material, s.sh = _ptr(self.density), _ptr(self.material), _ptr(self.sh)
        s.sh_stride = self.sh.shape[1]
        s.snormal = _ptr(self.snormal)
        if self.env is not None:
            s.env_h, s.env_w, s.env = self.env.shape[0], self.env.shape[1], _ptr(self.env)
        if self.table is not None:
            s.tab_h, s.tab_w, s.tab = self.table.shape[1], self.table.shape[2], _ptr(self.table)
        s.kernel_degree, s.min_response, s.density_clamping = kernel_degree, min_response, int(density_clamping)
        s.sph_degree = sph_degree
        s.enable_normals, s.enable_hitcounts = int(enable_normals), int(enable_hitcounts)
        s.enable_mis, s.enable_metallic, s.render_opts = int(enable_mis), int(enable_metallic), int(render_opts)
        s.min_transmittance, s.max_bounces, s.frame, s.kbuf = min_transmittance, max_bounces, frame, kbuf
        s.tab_kind = int(tab_kind)
        s.rr_start = int(rr_start)
        self.s = s

    @classmethod
    def from_scene(cls, scene, table, **kw):
        opts = dict(sph_degree=scene.sph_degree, enable_metallic=scene.enable_metallic,
                    min_transmittance=scene.min_transmittance, max_bounces=scene.max_bounces, frame=scene.frame)
        opts.update(kw)
        return cls(scene.particle_density(), scene.particle_material(), scene.features, scene.shading_normal,
                   scene.env, table, **opts)


class Rays:
    def __init__(self, ori, dir, ray_to_world, path_stride=1, path_offset=0, z_offset=0):
        self.ori = _f32(ori)
        self.dir = _f32(dir)
        self.T = _f32(ray_to_world).reshape(-1)[:12].copy()
        B, H, W, _ = self.dir.shape
        r = OracleRays()
        r.B, r.H, r.W = B, H, W
        r.ray_to_world, r.ori, r.dir = _ptr(self.T), _ptr(self.ori), _ptr(self.dir)
        r.path_stride, r.path_offset = path_stride, path_offset
        r.z_offset = z_offset
        self.r = r
        self.shape = (B, H, W)


def trace(inp: Inputs, rays: Rays, precision=64, nthreads=None, counters=False):
    """Forward pass; returns dict of float64 arrays with the reference's output shapes."""
    B, H, W = rays.shape
    out = {k: np.zeros((B, H, W, c), dtype=np.float64) for k, c in _FWD_CH.items()}
    out["visibility"] = np.zeros(inp.s.n, dtype=np.float64)
    out["margin"] = np.ones((B, H, W), dtype=np.float64)
    o = OracleFwdOut(**{k: _ptr(out[k]) for k in _FWD_KEYS + ["margin"]})
    cnt = OracleCounters()
    rc = lib().oracle_trace(C.byref(inp.s), C.byref(rays.r), C.byref(o), precision, nthreads or os.cpu_count(),
                            C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"oracle_trace failed: {rc}")
    out["comp"] = out["comp"].reshape(B, H, W, 2, 3)
    if counters:
        out["counters"] = cnt.as_dict()
    return out


TANGENT_FIELDS = {"density": "t_density", "material": "t_material", "sh": "t_sh", "snormal": "t_snormal",
                  "env": "t_env"}


def trace_dual(inp: Inputs, rays: Rays, tangents: dict, nthreads=None):
    """FP64 forward on dual numbers: (outputs, derivatives of the outputs along the direction
    `tangents` = {input name: array shaped like that input}, names of TANGENT_FIELDS)."""
    keep = {}
    for name, field in TANGENT_FIELDS.items():
        t = tangents.get(name)
        arr = None if t is None else _f32(t)
        if arr is not None:
            ref = getattr(inp, name if name != "env" else "env")
            if ref is None or arr.size != ref.size:
                raise ValueError(f"tangent {name} must match the input's shape")
        keep[name] = arr
        setattr(inp.s, field, _ptr(arr))
    B, H, W = rays.shape
    out = {k: np.zeros((B, H, W, c), dtype=np.float64) for k, c in _FWD_CH.items()}
    der = {k: np.zeros((B, H, W, c), dtype=np.float64) for k, c in _FWD_CH.items()}
    out["visibility"] = np.zeros(inp.s.n, dtype=np.float64)
    out["margin"] = np.ones((B, H, W), dtype=np.float64)
    o = OracleFwdOut(**{k: _ptr(out[k]) for k in _FWD_KEYS + ["margin"]})
    d = OracleFwdOut(**{k: _ptr(der[k]) for k in _FWD_CH})
    try:
        rc = lib().oracle_trace_dual(C.byref(inp.s), C.byref(rays.r), C.byref(o), C.byref(d),
                                     nthreads or os.cpu_count())
    finally:
        for field in TANGENT_FIELDS.values():
            setattr(inp.s, field, None)
    if rc != 0:
        raise RuntimeError(f"oracle_trace_dual failed: {rc}")
    for r in (out, der):
        r["comp"] = r["comp"].reshape(B, H, W, 2, 3)
    return out, der


def trace_bwd(inp: Inputs, rays: Rays, pbr, light, pbr_grad, light_grad, mat_grad, precision=64, nthreads=None,
              counters=False):
    """ptir-mode backward; returns (dmat [N,5], denv [H,W,4]) as float64."""
    pbr = np.ascontiguousarray(np.asarray(pbr, dtype=np.float64))
    light = np.ascontiguousarray(np.asarray(light, dtype=np.float64))
    pg, lg, mg = _f32(pbr_grad), _f32(light_grad), _f32(mat_grad)
    bi = OracleBwdIn(_ptr(pbr), _ptr(light), _ptr(pg), _ptr(lg), _ptr(mg))
    dmat = np.zeros((inp.s.n, 5), dtype=np.float64)
    denv = np.zeros(inp.env.shape if inp.env is not None else (0, 0, 4), dtype=np.float64)
    bo = OracleBwdOut(_ptr(dmat), _ptr(denv) if denv.size else None)
    cnt = OracleCounters()
    rc = lib().oracle_trace_bwd(C.byref(inp.s), C.byref(rays.r), C.byref(bi), C.byref(bo), precision,
                                nthreads or os.cpu_count(), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"oracle_trace_bwd failed: {rc}")
    if counters:
        return dmat, denv, cnt.as_dict()
    return dmat, denv


def pcg32(seed, seq, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_pcg32(seed, seq, n, out.ctypes.data)
    return out


def launch_seed(x, y, z, f):
    return int(lib().oracle_launch_seed(x, y, z, f))


def response(degree, d2):
    return float(lib().oracle_response(degree, d2))


def kernel_scale(density, min_response=0.0113, clamping=True, degree=4.0):
    return float(lib().oracle_kernel_scale(density, min_response, int(clamping), degree))


def brdf_eval(wo, n, albedo, metal, rough, l, enable_metallic=True):
    a = [np.ascontiguousarray(np.asarray(x, dtype=np.float64)) for x in (wo, n, albedo, l)]
    value, dalb, drough, dmet = (np.zeros(3) for _ in range(4))
    pdf = np.zeros(1)
    lib().oracle_brdf_eval(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, metal, rough, a[3].ctypes.data,
                           int(enable_metallic), value.ctypes.data, pdf.ctypes.data, dalb.ctypes.data,
                           drough.ctypes.data, dmet.ctypes.data)
    return value, float(pdf[0]), dalb, drough, dmet


def brdf_sample(wo, n, albedo, metal, rough, rnd, enable_metallic=True):
    a = [np.ascontiguousarray(np.asarray(x, dtype=np.float64)) for x in (wo, n, albedo, rnd)]
    value, d = np.zeros(3), np.zeros(3)
    pdf = np.zeros(1)
    lib().oracle_brdf_sample(a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data, metal, rough, a[3].ctypes.data,
                             int(enable_metallic), value.ctypes.data, d.ctypes.data, pdf.ctypes.data)
    return value, d, float(pdf[0])


def env_lookup(inp: Inputs, d):
    d = np.ascontiguousarray(np.asarray(d, dtype=np.float64))
    rgb = np.zeros(3)
    lib().oracle_env_lookup(C.byref(inp.s), d.ctypes.data, rgb.ctypes.data)
    return rgb


def env_pdf(inp: Inputs, d):
    d = np.ascontiguousarray(np.asarray(d, dtype=np.float64))
    return float(lib().oracle_env_pdf(C.byref(inp.s), d.ctypes.data))


def env_sample(inp: Inputs, seed, n):
    dirs, pdfs = np.zeros((n, 3)), np.zeros(n)
    lib().oracle_env_sample(C.byref(inp.s), seed, n, dirs.ctypes.data, pdfs.ctypes.data)
    return dirs, pdfs


def candidates(inp: Inputs, o, d, use_bvh=True, cap=100000):
    o = np.ascontiguousarray(np.asarray(o, dtype=np.float64))
    d = np.ascontiguousarray(np.asarray(d, dtype=np.float64))
    ts = np.zeros(cap)
    ids = np.zeros(cap, dtype=np.uint32)
    m = lib().oracle_candidates(C.byref(inp.s), o.ctypes.data, d.ctypes.data, int(use_bvh), cap, ts.ctypes.data,
                                ids.ctypes.data)
    m = min(m, cap)
    return ts[:m], ids[:m]


def scene_aabb(inp: Inputs):
    out = np.zeros(6)
    lib().oracle_scene_aabb(C.byref(inp.s), out.ctypes.data)
    return out


def env_cdf_table(env, h=64, w=128):
    """CPU restatement, in float64, of the device-built environment CDF table
    (ptir_env_table_build): weights of threedgrut/model/aliastable.py:244-297 (area
    resampling, Rec.709 luminance of the clamped RGB, times the equirect cell solid angle),
    per-steradian pdf w / (sum(w) sa), inclusive conditional CDF per row (plane 0), inclusive
    marginal CDF of the rows in column 0 of plane 1, pdf in plane 2; float32 [3, h, w]."""
    e = np.asarray(env.detach().cpu().numpy() if hasattr(env, "detach") else env, dtype=np.float64)[..., :3]
    H, W = e.shape[:2]
    rgb = np.zeros((h, w, 3))
    for y in range(h):
        y0, y1 = (y * H) // h, -((-(y + 1) * H) // h)
        for x in range(w):
            x0, x1 = (x * W) // w, -((-(x + 1) * W) // w)
            rgb[y, x] = e[y0:y1, x0:x1].reshape(-1, 3).mean(0)
    rgb = np.where(np.isfinite(rgb) & (rgb > 0), rgb, 0.0)
    lum = rgb @ np.array([0.2126, 0.7152, 0.0722])
    edges = (np.arange(h + 1) / h - 0.5) * np.pi
    sa_row = np.clip((2 * np.pi / w) * (np.sin(edges[1:]) - np.sin(edges[:-1])), 0, None)
    sa = np.repeat(sa_row[:, None], w, axis=1)
    wts = lum * sa
    total = wts.sum()
    if not (np.isfinite(total) and total > 0):
        wts = sa.copy()
        total = wts.sum()
    rows = wts.sum(1)
    out = np.zeros((3, h, w))
    with np.errstate(invalid="ignore", divide="ignore"):
        cond = np.where(rows[:, None] > 0, np.cumsum(wts, 1) / rows[:, None], (np.arange(w) + 1.0) / w)
    cond[:, -1] = 1.0
    marg = np.cumsum(rows) / total
    marg[-1] = 1.0
    out[0] = cond
    out[1, :, 0] = marg
    out[2] = np.where(sa > 0, wts / (total * np.where(sa > 0, sa, 1.0)), 0.0)
    return out.astype(np.float32)


# ---------------------------------------------------------------- 3DGRT primary-only tracer
GRT_KEYS = ["rad", "dns", "hit", "hit2", "distortion", "nrm", "snrm", "hits"]


def grt_inputs(density, sh, snormal, **kw) -> Inputs:
    """Inputs of the 3DGRT tracer: no material, no environment light."""
    inp = Inputs(density, np.zeros((np.asarray(density).reshape(-1, 12).shape[0], 5), np.float32), sh, snormal, None,
                 None, max_bounces=1, enable_mis=False, **kw)
    inp.s.material = None
    return inp


def grt_trace(inp: Inputs, rays: Rays, nthreads=None):
    """3DGRT forward: rad, dns, hit (2 channels), hit2, distortion, nrm, snrm (raw weighted sum),
    hits, visibility, margin; float64."""
    B, H, W = rays.shape
    out = {k: np.zeros((B, H, W, c), dtype=np.float64) for k, c in _FWD_CH.items()}
    out["visibility"] = np.zeros(inp.s.n, dtype=np.float64)
    out["margin"] = np.ones((B, H, W), dtype=np.float64)
    o = OracleFwdOut(**{k: _ptr(out[k]) for k in _FWD_KEYS + ["margin"]})
    rc = lib().oracle_grt_trace(C.byref(inp.s), C.byref(rays.r), C.byref(o), nthreads or os.cpu_count())
    if rc != 0:
        raise RuntimeError(f"oracle_grt_trace failed: {rc}")
    return {k: out[k] for k in GRT_KEYS + ["visibility", "margin"]}


def grt_trace_dual(inp: Inputs, rays: Rays, tangents: dict, nthreads=None):
    """3DGRT forward on dual numbers: (outputs, derivatives along `tangents`: density, sh, snormal)."""
    keep = {}
    for name in ("density", "sh", "snormal"):
        t = tangents.get(name)
        arr = None if t is None else _f32(t)
        keep[name] = arr
        setattr(inp.s, TANGENT_FIELDS[name], _ptr(arr))
    B, H, W = rays.shape
    out = {k: np.zeros((B, H, W, c), dtype=np.float64) for k, c in _FWD_CH.items()}
    der = {k: np.zeros((B, H, W, c), dtype=np.float64) for k, c in _FWD_CH.items()}
    out["visibility"] = np.zeros(inp.s.n, dtype=np.float64)
    out["margin"] = np.ones((B, H, W), dtype=np.float64)
    o = OracleFwdOut(**{k: _ptr(out[k]) for k in _FWD_KEYS + ["margin"]})
    d = OracleFwdOut(**{k: _ptr(der[k]) for k in _FWD_CH})
    try:
        rc = lib().oracle_grt_trace_dual(C.byref(inp.s), C.byref(rays.r), C.byref(o), C.byref(d),
                                         nthreads or os.cpu_count())
    finally:
        for field in TANGENT_FIELDS.values():
            setattr(inp.s, field, None)
    if rc != 0:
        raise RuntimeError(f"oracle_grt_trace_dual failed: {rc}")
    return out, der


def grt_trace_bwd(inp: Inputs, rays: Rays, fwd: dict, grads: dict, inclusive_end=False, nthreads=None):
    """3DGRT backward (the reference's processHitBwd): fwd = grt_trace outputs; grads = output
    gradients keyed rad [..,3], dns, hit [..,1], hit2, distortion, snrm [..,3] (missing = 0).
    Returns (d_density [N,12], d_sh [N, sh_stride], d_snormal [N,3])."""
    f = {k: np.ascontiguousarray(np.asarray(fwd[k], dtype=np.float64)) for k in
         ("rad", "dns", "hit", "hit2", "distortion", "snrm")}
    g = {k: _f32(grads.get(k)) for k in ("rad", "dns", "hit", "hit2", "distortion", "snrm")}
    gi = OracleGrtIn(_ptr(f["rad"]), _ptr(f["dns"]), _ptr(f["hit"]), _ptr(f["hit2"]), _ptr(f["distortion"]),
                     _ptr(f["snrm"]), _ptr(g["rad"]), _ptr(g["dns"]), _ptr(g["hit"]), _ptr(g["hit2"]),
                     _ptr(g["distortion"]), _ptr(g["snrm"]), int(bool(inclusive_end)))
    n = inp.s.n
    dd = np.zeros((n, 12), np.float64)
    dsh = np.zeros((n, inp.s.sh_stride), np.float64)
    dsn = np.zeros((n, 3), np.float64)
    go = OracleGrtOut(_ptr(dd), _ptr(dsh), _ptr(dsn))
    rc = lib().oracle_grt_trace_bwd(C.byref(inp.s), C.byref(rays.r), C.byref(gi), C.byref(go),
                                    nthreads or os.cpu_count())
    if rc != 0:
        raise RuntimeError(f"oracle_grt_trace_bwd failed: {rc}")
    return dd, dsh, dsn
