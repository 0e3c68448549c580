"""Seeded synthetic scenes for the differentiable Gaussian path tracer.

Each generator returns a :class:`SyntheticScene` holding activated Gaussian parameters
(the tensors `Tracer._Autograd.apply` receives in the reference,
threedgptir_tracer/tracer.py:862-889), an environment map ``[H, W, 4]`` (alpha = 1),
the ray batch ``[B, H, W, 3]`` already expanded over samples per pixel, and the
launch scalars.  Everything is generated on the CPU with ``torch.Generator`` seeds, so
the GPU path and the CPU oracle see bit-identical float32 inputs.

Workloads:

* ``small``  : 4,096 Gaussians uniform in [-1,1]^3, env 64x32, one 64x64 view, 4 spp,
  maxBounces 2 (primary + one secondary ray with NEE).
* ``object`` : 300,000 Gaussians in a +-0.01 shell around the union of a sphere
  (r = 0.6) and a torus (R = 0.5, r = 0.15), log-scales N(-5, 0.4) with the shortest
  axis on the surface normal, 8 material clusters, env 512x256 with a 500x sun lobe,
  4 views 800x800 at radius 3, 1 spp, maxBounces 3.
* ``batched``: 500,000 Gaussians on the same shell, 16 views with ragged resolutions
  and spp in one launch, env 1024x512, maxBounces 4 with Russian roulette from bounce 2.
* ``scene``  : 3,000,000 Gaussians (ground clusters, box clusters, background shell),
  env 2048x1024, 8 views 1280x720, maxBounces 5 with Russian roulette from bounce 2.

The `maxBounces` convention is the reference's: path vertices including the primary hit
(src/kernels/cuda/referenceOptix.cu:52-67).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import torch

# Sample-position patterns of the reference's spp expansion (tracer.py:391-426).
MULTISPP_PATTERNS = {
    1: ((0.5, 0.5),),
    2: ((0.25, 0.25), (0.75, 0.75)),
    4: ((0.375, 0.125), (0.875, 0.375), (0.625, 0.875), (0.125, 0.625)),
    8: (
        (0.5625, 0.6875), (0.4375, 0.3125), (0.8125, 0.4375), (0.3125, 0.8125),
        (0.1875, 0.1875), (0.0625, 0.5625), (0.6875, 0.0625), (0.9375, 0.9375),
    ),
}


@dataclass
class SyntheticScene:
    name: str
    positions: torch.Tensor      # [N,3]
    rotation: torch.Tensor       # [N,4] unit quaternion (w,x,y,z)
    scale: torch.Tensor          # [N,3] activated (exp)
    density: torch.Tensor        # [N,1] activated (sigmoid)
    features: torch.Tensor       # [N,3*(deg+1)^2] SH, coefficient-major, RGB inner
    shading_normal: torch.Tensor  # [N,3]
    albedo: torch.Tensor         # [N,3]
    roughness: torch.Tensor      # [N,1]
    metallic: torch.Tensor       # [N,1]
    env: torch.Tensor            # [H,W,4]
    rays_ori: torch.Tensor       # [B,H,W,3]
    rays_dir: torch.Tensor       # [B,H,W,3]
    ray_to_world: torch.Tensor   # [1,4,4]
    sph_degree: int = 3
    max_bounces: int = 3
    spp: int = 1
    frame: int = 0
    min_transmittance: float = 1e-3
    enable_metallic: bool = True
    extra: dict = field(default_factory=dict)

    @property
    def num_gaussians(self) -> int:
        return int(self.positions.shape[0])

    @property
    def num_paths(self) -> int:
        return int(self.rays_dir.shape[0] * self.rays_dir.shape[1] * self.rays_dir.shape[2])

    def particle_density(self) -> torch.Tensor:
        """[N,12] = [pos3, sigma, quat4, scale3, 0] (tracer.py:163-166)."""
        return torch.cat(
            [self.positions, self.density, self.rotation, self.scale, torch.zeros_like(self.density)], dim=1
        ).contiguous()

    def particle_material(self) -> torch.Tensor:
        """[N,5] = [albedo3, roughness, metallic] (tracer.py:167)."""
        return torch.cat([self.albedo, self.roughness, self.metallic], dim=1).contiguous()

    def to(self, device) -> "SyntheticScene":
        out = SyntheticScene(**{k: getattr(self, k) for k in self.__dataclass_fields__})
        for k in self.__dataclass_fields__:
            v = getattr(out, k)
            if isinstance(v, torch.Tensor):
                setattr(out, k, v.to(device))
        return out


# ----------------------------------------------------------------------------- helpers
def _gen(seed: int) -> torch.Generator:
    g = torch.Generator()
    g.manual_seed(int(seed))
    return g


def random_unit_quaternions(n: int, g: torch.Generator) -> torch.Tensor:
    q = torch.randn(n, 4, generator=g, dtype=torch.float64)
    q = q / q.norm(dim=1, keepdim=True)
    return q.float()


def random_unit_vectors(n: int, g: torch.Generator) -> torch.Tensor:
    v = torch.randn(n, 3, generator=g, dtype=torch.float64)
    return (v / v.norm(dim=1, keepdim=True)).float()


def quaternion_from_matrix(m: torch.Tensor) -> torch.Tensor:
    """Unit quaternions (w,x,y,z) of rotation matrices [N,3,3] (column convention)."""
    m = m.double()
    n = m.shape[0]
    q = torch.empty(n, 4, dtype=torch.float64)
    tr = m[:, 0, 0] + m[:, 1, 1] + m[:, 2, 2]
    # Shepperd's method: pick the largest of 4 candidate divisors for stability.
    cands = torch.stack([tr, m[:, 0, 0], m[:, 1, 1], m[:, 2, 2]], dim=1)
    k = cands.argmax(dim=1)
    for case in range(4):
        sel = k == case
        if not sel.any():
            continue
        a = m[sel]
        if case == 0:
            s = torch.sqrt(1.0 + a[:, 0, 0] + a[:, 1, 1] + a[:, 2, 2]) * 2
            q[sel] = torch.stack([0.25 * s, (a[:, 2, 1] - a[:, 1, 2]) / s, (a[:, 0, 2] - a[:, 2, 0]) / s,
                                  (a[:, 1, 0] - a[:, 0, 1]) / s], dim=1)
        elif case == 1:
            s = torch.sqrt(1.0 + a[:, 0, 0] - a[:, 1, 1] - a[:, 2, 2]) * 2
            q[sel] = torch.stack([(a[:, 2, 1] - a[:, 1, 2]) / s, 0.25 * s, (a[:, 0, 1] + a[:, 1, 0]) / s,
                                  (a[:, 0, 2] + a[:, 2, 0]) / s], dim=1)
        elif case == 2:
            s = torch.sqrt(1.0 + a[:, 1, 1] - a[:, 0, 0] - a[:, 2, 2]) * 2
            q[sel] = torch.stack([(a[:, 0, 2] - a[:, 2, 0]) / s, (a[:, 0, 1] + a[:, 1, 0]) / s, 0.25 * s,
                                  (a[:, 1, 2] + a[:, 2, 1]) / s], dim=1)
        else:
            s = torch.sqrt(1.0 + a[:, 2, 2] - a[:, 0, 0] - a[:, 1, 1]) * 2
            q[sel] = torch.stack([(a[:, 1, 0] - a[:, 0, 1]) / s, (a[:, 0, 2] + a[:, 2, 0]) / s,
                                  (a[:, 1, 2] + a[:, 2, 1]) / s, 0.25 * s], dim=1)
    q = q / q.norm(dim=1, keepdim=True)
    return q.float()


def frames_from_normals(nrm: torch.Tensor, g: torch.Generator) -> torch.Tensor:
    """Rotation matrices whose third column is the given unit normal, random spin."""
    n = nrm.double()
    a = torch.where(n[:, 0:1].abs() < 0.9, torch.tensor([[1.0, 0.0, 0.0]], dtype=torch.float64),
                    torch.tensor([[0.0, 1.0, 0.0]], dtype=torch.float64))
    t = torch.linalg.cross(a, n)
    t = t / t.norm(dim=1, keepdim=True)
    b = torch.linalg.cross(n, t)
    phi = torch.rand(n.shape[0], 1, generator=g, dtype=torch.float64) * 2 * math.pi
    t2 = t * torch.cos(phi) + b * torch.sin(phi)
    b2 = torch.linalg.cross(n, t2)
    return torch.stack([t2, b2, n], dim=2)


def lognormal_env(h: int, w: int, g: torch.Generator, sigma: float = 1.0) -> torch.Tensor:
    rgb = torch.exp(torch.randn(h, w, 3, generator=g) * sigma)
    return torch.cat([rgb, torch.ones(h, w, 1)], dim=2).contiguous()


def env_direction_grid(h: int, w: int) -> torch.Tensor:
    """World direction of every texel centre of an equirect env in the tracer's mapping
    (environment.cuh:123-158,203-219 with zero offset): u -> theta, v -> phi."""
    v = (torch.arange(h, dtype=torch.float64) + 0.5) / h
    u = (torch.arange(w, dtype=torch.float64) + 0.5) / w
    theta = u * 2 * math.pi - math.pi
    phi = (v - 0.5) * math.pi
    th, ph = theta.view(1, w).expand(h, w), phi.view(h, 1).expand(h, w)
    e = torch.stack([torch.sin(th) * torch.cos(ph), torch.cos(th) * torch.cos(ph), -torch.sin(ph)], dim=-1)
    # inverse of the pi/2 rotation about z: world = (e.y, -e.x, e.z)
    return torch.stack([e[..., 1], -e[..., 0], e[..., 2]], dim=-1)


def sun_sky_env(h: int, w: int, g: torch.Generator, sun_gain: float = 500.0, sun_radius_deg: float = 3.0):
    sky = lognormal_env(h, w, g, sigma=0.5)
    sun = random_unit_vectors(1, g).double()[0]
    sun[2] = abs(sun[2]) + 0.3
    sun = sun / sun.norm()
    dirs = env_direction_grid(h, w)
    ang = torch.arccos(torch.clamp((dirs * sun).sum(-1), -1, 1))
    lobe = 1.0 + (sun_gain - 1.0) * torch.exp(-((ang / math.radians(sun_radius_deg)) ** 2))
    sky[..., :3] = (sky[..., :3].double() * lobe.unsqueeze(-1)).float()
    return sky.contiguous(), sun.float()


def look_at(cam: torch.Tensor) -> torch.Tensor:
    """Camera-to-world rotation (OpenCV: x right, y down, z forward) looking at origin."""
    c = cam.double()
    f = -c / c.norm()
    up = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)
    if abs(float((f * up).sum())) > 0.999:
        up = torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64)
    x = torch.linalg.cross(f, up)
    x = x / x.norm()
    y = torch.linalg.cross(f, x)
    return torch.stack([x, y, f], dim=1)


def pinhole_rays(h: int, w: int, fov_deg: float, spp: int, pattern=None):
    """Camera-space unit ray directions [spp, H, W, 3] like tracer.py:646-706."""
    f = 0.5 * w / math.tan(math.radians(fov_deg) / 2)
    cx, cy = w / 2.0, h / 2.0
    pat = torch.tensor(pattern if pattern is not None else MULTISPP_PATTERNS[spp], dtype=torch.float64)
    ys, xs = torch.meshgrid(torch.arange(h, dtype=torch.float64), torch.arange(w, dtype=torch.float64), indexing="ij")
    dirs = torch.stack([
        (xs.unsqueeze(0) + pat[:, 0].view(-1, 1, 1) - cx) / f,
        (ys.unsqueeze(0) + pat[:, 1].view(-1, 1, 1) - cy) / f,
        torch.ones(pat.shape[0], h, w, dtype=torch.float64),
    ], dim=-1)
    return torch.nn.functional.normalize(dirs, dim=-1)


def world_rays(cams, h, w, fov_deg, spp):
    """World-space rays for several cameras looking at the origin: [V*spp, H, W, 3]."""
    oris, dirs = [], []
    local = pinhole_rays(h, w, fov_deg, spp)
    for s in range(spp):
        for c in cams:
            R = look_at(c)
            d = torch.einsum("ij,hwj->hwi", R, local[s])
            dirs.append(d.float())
            oris.append(c.float().view(1, 1, 3).expand(h, w, 3))
    return torch.stack(oris).contiguous(), torch.stack(dirs).contiguous()


def _materials_uniform(n, g):
    albedo = torch.rand(n, 3, generator=g)
    rough = 0.1 + 0.8 * torch.rand(n, 1, generator=g)
    metal = torch.rand(n, 1, generator=g)
    return albedo, rough, metal


# ----------------------------------------------------------------------------- workloads
def small_scene(seed: int = 0, n: int = 4096, res: int = 64, spp: int = 4, env_hw=(32, 64),
                max_bounces: int = 2) -> SyntheticScene:
    """4,096 Gaussians, one 64x64 pinhole view at radius 4, 40 deg FOV, 4 spp, 1 bounce."""
    g = _gen(seed)
    pos = torch.rand(n, 3, generator=g) * 2 - 1
    scale = torch.exp(-3.5 + 0.3 * torch.randn(n, 3, generator=g))
    rot = random_unit_quaternions(n, g)
    dns = torch.sigmoid(torch.randn(n, 1, generator=g))
    albedo, rough, metal = _materials_uniform(n, g)
    snrm = random_unit_vectors(n, g)
    sh = 0.1 * torch.randn(n, 48, generator=g)
    env = lognormal_env(env_hw[0], env_hw[1], g)
    cam = torch.tensor([2.6, -2.2, 2.0], dtype=torch.float64)
    cam = 4.0 * cam / cam.norm()
    R = look_at(cam)
    T = torch.eye(4, dtype=torch.float64)
    T[:3, :3] = R
    T[:3, 3] = cam
    dirs = pinhole_rays(res, res, 40.0, spp).float().contiguous()
    oris = torch.zeros_like(dirs)
    return SyntheticScene("small", pos, rot, scale, dns, sh, snrm, albedo, rough, metal, env, oris, dirs,
                          T.float().unsqueeze(0), sph_degree=3, max_bounces=max_bounces, spp=spp)


def _union_shell(n: int, g: torch.Generator, shell: float = 0.01):
    """Points on the boundary of (sphere r=0.6) U (torus R=0.5, r=0.15 about z), offset
    by U(-shell, shell) along the SDF normal.  Returns (points, normals)."""
    a_s = 4 * math.pi * 0.6 ** 2
    a_t = 4 * math.pi ** 2 * 0.5 * 0.15
    pts, nrms = [], []
    have = 0
    while have < n:
        m = 2 * (n - have) + 1024
        pick_sphere = torch.rand(m, generator=g, dtype=torch.float64) < a_s / (a_s + a_t)
        # sphere samples
        sv = torch.randn(m, 3, generator=g, dtype=torch.float64)
        sv = sv / sv.norm(dim=1, keepdim=True)
        sp = 0.6 * sv
        # torus samples: area element proportional to (R + r cos v)
        u = torch.rand(m, generator=g, dtype=torch.float64) * 2 * math.pi
        v = torch.rand(m, generator=g, dtype=torch.float64) * 2 * math.pi
        keep_t = torch.rand(m, generator=g, dtype=torch.float64) < (0.5 + 0.15 * torch.cos(v)) / 0.65
        tn = torch.stack([torch.cos(v) * torch.cos(u), torch.cos(v) * torch.sin(u), torch.sin(v)], dim=1)
        tp = torch.stack([(0.5 + 0.15 * torch.cos(v)) * torch.cos(u), (0.5 + 0.15 * torch.cos(v)) * torch.sin(u),
                          0.15 * torch.sin(v)], dim=1)
        p = torch.where(pick_sphere.unsqueeze(1), sp, tp)
        nn_ = torch.where(pick_sphere.unsqueeze(1), sv, tn)
        ok = torch.where(pick_sphere, torch.ones_like(keep_t), keep_t)
        # keep only the union boundary: sphere points outside the torus, torus points outside the sphere
        q = torch.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2) - 0.5
        d_torus = torch.sqrt(q ** 2 + p[:, 2] ** 2) - 0.15
        d_sphere = p.norm(dim=1) - 0.6
        ok = ok & torch.where(pick_sphere, d_torus > -1e-9, d_sphere > -1e-9)
        p, nn_ = p[ok], nn_[ok]
        pts.append(p)
        nrms.append(nn_)
        have += p.shape[0]
    p = torch.cat(pts)[:n]
    nrm = torch.cat(nrms)[:n]
    off = (torch.rand(n, 1, generator=g, dtype=torch.float64) * 2 - 1) * shell
    return (p + off * nrm).float(), nrm.float()


def _shell_gaussians(n, g, log_mean=-5.0, log_std=0.4):
    pos, nrm = _union_shell(n, g)
    ls = log_mean + log_std * torch.randn(n, 3, generator=g)
    ls, _ = torch.sort(ls, dim=1, descending=True)  # smallest scale on the local z (normal) axis
    scale = torch.exp(ls)
    rot = quaternion_from_matrix(frames_from_normals(nrm, g))
    dns = torch.sigmoid(torch.randn(n, 1, generator=g))
    # 8 material clusters by octant, small jitter
    centers_alb = torch.rand(8, 3, generator=g)
    centers_rgh = 0.1 + 0.8 * torch.rand(8, 1, generator=g)
    centers_met = torch.rand(8, 1, generator=g)
    cl = (pos[:, 0] > 0).long() + 2 * (pos[:, 1] > 0).long() + 4 * (pos[:, 2] > 0).long()
    albedo = (centers_alb[cl] + 0.02 * torch.randn(n, 3, generator=g)).clamp(0, 1)
    rough = (centers_rgh[cl] + 0.02 * torch.randn(n, 1, generator=g)).clamp(0.05, 1)
    metal = (centers_met[cl] + 0.02 * torch.randn(n, 1, generator=g)).clamp(0, 1)
    sh = 0.1 * torch.randn(n, 48, generator=g)
    return pos, rot, scale, dns, sh, nrm, albedo, rough, metal


def orbit_cameras(n_views: int, radius: float, g: torch.Generator):
    cams = []
    for i in range(n_views):
        az = 2 * math.pi * (i + float(torch.rand(1, generator=g))) / n_views
        el = math.radians(-20 + 50 * float(torch.rand(1, generator=g)))
        cams.append(torch.tensor([radius * math.cos(el) * math.cos(az), radius * math.cos(el) * math.sin(az),
                                  radius * math.sin(el)], dtype=torch.float64))
    return cams


def object_scene(seed: int = 0, n: int = 300_000, res=(800, 800), n_views: int = 4, spp: int = 1,
                 env_hw=(256, 512), max_bounces: int = 3) -> SyntheticScene:
    """300k Gaussians on the sphere-torus shell, 4 world-space views, 1 spp, 2 bounces."""
    g = _gen(seed)
    pos, rot, scale, dns, sh, nrm, albedo, rough, metal = _shell_gaussians(n, g)
    env, sun = sun_sky_env(env_hw[0], env_hw[1], g)
    cams = orbit_cameras(n_views, 3.0, g)
    oris, dirs = world_rays(cams, res[0], res[1], 40.0, spp)
    scene = SyntheticScene("object", pos, rot, scale, dns, sh, nrm, albedo, rough, metal, env, oris, dirs,
                           torch.eye(4).unsqueeze(0), sph_degree=3, max_bounces=max_bounces, spp=spp)
    scene.extra["sun"] = sun
    return scene


def shrink(scene: SyntheticScene, h: int, w: int, views: int | None = None) -> SyntheticScene:
    """Same scene, a centred h x w crop of the first `views` ray batches (parity-test size)."""
    out = scene.to("cpu")
    b = out.rays_dir.shape[0] if views is None else views
    H, W = out.rays_dir.shape[1:3]
    y0, x0 = (H - h) // 2, (W - w) // 2
    out.rays_dir = out.rays_dir[:b, y0:y0 + h, x0:x0 + w].contiguous()
    out.rays_ori = out.rays_ori[:b, y0:y0 + h, x0:x0 + w].contiguous()
    return out


def target_image(scene: SyntheticScene, seed: int = 1) -> torch.Tensor:
    """Seeded random L2 target for the pbr output, [B,H,W,3]."""
    g = _gen(seed)
    return torch.rand(*scene.rays_dir.shape[:3], 3, generator=g)


def batched_scene(seed: int = 0, n: int = 500_000, n_views: int = 16, res_choices=(256, 512, 768, 1024),
                  spp_choices=(1, 2, 4, 8), env_hw=(512, 1024), max_bounces: int = 4, rr_start: int = 2,
                  res_scale: float = 1.0) -> SyntheticScene:
    """Ragged batch: 500k Gaussians on the sphere-torus shell, 16 views on a sphere of radius
    3, each with its own square resolution and samples per pixel; up to 3 bounces with
    Russian roulette from bounce 2.  Every (view, sample) is one image of a ragged launch:
    rays are flat world-space [1, 1, P, 3] in image order, extra["images"] lists the (H, W)
    of each image and extra["view_of_image"] / extra["spp"] the view it belongs to and that
    view's samples per pixel.  res_scale shrinks every resolution (parity-test sizes)."""
    g = _gen(seed)
    pos, rot, scale, dns, sh, nrm, albedo, rough, metal = _shell_gaussians(n, g)
    env, sun = sun_sky_env(env_hw[0], env_hw[1], g)
    cams = orbit_cameras(n_views, 3.0, g)
    res = [res_choices[int(i)] for i in torch.randint(len(res_choices), (n_views,), generator=g)]
    spps = [spp_choices[int(i)] for i in torch.randint(len(spp_choices), (n_views,), generator=g)]
    oris, dirs, images, view_of = [], [], [], []
    for v, (c, r, sp) in enumerate(zip(cams, res, spps)):
        r = max(8, int(round(r * res_scale)))
        o, d = world_rays([c], r, r, 40.0, sp)  # [sp, r, r, 3]
        for k in range(sp):
            oris.append(o[k].reshape(-1, 3))
            dirs.append(d[k].reshape(-1, 3))
            images.append((r, r))
            view_of.append(v)
    ori = torch.cat(oris).reshape(1, 1, -1, 3).contiguous()
    dirv = torch.cat(dirs).reshape(1, 1, -1, 3).contiguous()
    scene = SyntheticScene("batched", pos, rot, scale, dns, sh, nrm, albedo, rough, metal, env, ori, dirv,
                           torch.eye(4).unsqueeze(0), sph_degree=3, max_bounces=max_bounces, spp=1)
    scene.extra.update(sun=sun, images=images, view_of_image=view_of, spp=spps, rr_start=rr_start)
    return scene


def _box_surface(m: int, lo: torch.Tensor, hi: torch.Tensor, g: torch.Generator):
    """m points on the surface of the box [lo, hi] (area-weighted faces) and their normals."""
    ext = hi - lo
    areas = torch.stack([ext[1] * ext[2], ext[1] * ext[2], ext[0] * ext[2], ext[0] * ext[2], ext[0] * ext[1],
                         ext[0] * ext[1]])
    face = torch.multinomial(areas / areas.sum(), m, replacement=True, generator=g)
    u = torch.rand(m, 3, generator=g, dtype=torch.float64)
    p = lo + u * ext
    axis = face // 2
    side = face % 2
    p[torch.arange(m), axis] = torch.where(side == 0, lo[axis], hi[axis])
    nrm = torch.zeros(m, 3, dtype=torch.float64)
    nrm[torch.arange(m), axis] = torch.where(side == 0, -1.0, 1.0).double()
    return p, nrm


def scene_scale(seed: int = 0, n: int = 3_000_000, res=(720, 1280), n_views: int = 8, env_hw=(1024, 2048),
                max_bounces: int = 5, rr_start: int = 2) -> SyntheticScene:
    """Scene-scale unbounded workload: 60% of the Gaussians in 200 Gaussian-mixture clusters
    over a 40 x 40 ground area (z in [0, 2]), 25% on the faces of 50 boxes (walls and
    furniture), 15% on a background shell of radius 50 with log-scales N(0, 0.5); env
    2048x1024; a camera path of 8 views at 1280x720 moving through the scene at eye height,
    1 spp, 4 bounces with Russian roulette from bounce 2."""
    g = _gen(seed)
    n_g, n_b = int(0.60 * n), int(0.25 * n)
    n_s = n - n_g - n_b
    # ground clusters
    centers = torch.cat([torch.rand(200, 2, generator=g, dtype=torch.float64) * 40 - 20,
                         torch.rand(200, 1, generator=g, dtype=torch.float64) * 2], 1)
    spread = torch.stack([0.3 + 1.2 * torch.rand(200, generator=g, dtype=torch.float64)] * 2 +
                         [0.1 + 0.4 * torch.rand(200, generator=g, dtype=torch.float64)], 1)
    cl = torch.randint(200, (n_g,), generator=g)
    pg = centers[cl] + spread[cl] * torch.randn(n_g, 3, generator=g, dtype=torch.float64)
    pg[:, 2] = pg[:, 2].clamp(0.0, 2.0)
    ng = random_unit_vectors(n_g, g).double()
    # box clusters
    bl = torch.cat([torch.rand(50, 2, generator=g, dtype=torch.float64) * 36 - 18,
                    torch.zeros(50, 1, dtype=torch.float64)], 1)
    be = torch.cat([0.5 + 4.5 * torch.rand(50, 2, generator=g, dtype=torch.float64),
                    0.5 + 2.5 * torch.rand(50, 1, generator=g, dtype=torch.float64)], 1)
    per = torch.full((50,), n_b // 50, dtype=torch.long)
    per[: n_b - int(per.sum())] += 1
    pb, nb = [], []
    for k in range(50):
        p_, n_ = _box_surface(int(per[k]), bl[k], bl[k] + be[k], g)
        pb.append(p_)
        nb.append(n_)
    pb, nb = torch.cat(pb), torch.cat(nb)
    # background shell
    vs = random_unit_vectors(n_s, g).double()
    ps = 50.0 * vs
    pos = torch.cat([pg, pb, ps]).float()
    nrm = torch.cat([ng, nb, -vs]).float()
    ls = torch.cat([-3.5 + 0.4 * torch.randn(n_g, 3, generator=g), -3.8 + 0.3 * torch.randn(n_b, 3, generator=g),
                    0.5 * torch.randn(n_s, 3, generator=g)])
    ls, _ = torch.sort(ls, dim=1, descending=True)  # flattest axis on the surface normal
    scale = torch.exp(ls)
    rot = quaternion_from_matrix(frames_from_normals(nrm, g))
    dns = torch.sigmoid(torch.randn(n, 1, generator=g))
    albedo, rough, metal = _materials_uniform(n, g)
    sh = 0.1 * torch.randn(n, 48, generator=g)
    env, sun = sun_sky_env(env_hw[0], env_hw[1], g, sun_gain=100.0)
    oris, dirs = [], []
    local = pinhole_rays(res[0], res[1], 60.0, 1)[0]
    for v in range(n_views):
        t = v / max(1, n_views - 1)
        cam = torch.tensor([-12.0 + 24.0 * t, -6.0 + 4.0 * math.sin(3.0 * t), 1.6], dtype=torch.float64)
        target = cam + torch.tensor([math.cos(0.8 * v), math.sin(0.8 * v), -0.15], dtype=torch.float64)
        f = target - cam
        R = look_at(-f / f.norm())  # look_at looks from its argument towards the origin
        oris.append(cam.float().view(1, 1, 3).expand(res[0], res[1], 3))
        dirs.append(torch.einsum("ij,hwj->hwi", R, local).float())
    scene = SyntheticScene("scene", pos, rot, scale, dns, sh, nrm, albedo, rough, metal, env,
                           torch.stack(oris).contiguous(), torch.stack(dirs).contiguous(), torch.eye(4).unsqueeze(0),
                           sph_degree=3, max_bounces=max_bounces, spp=1)
    scene.extra.update(sun=sun, rr_start=rr_start)
    return scene


WORKLOADS = {
    "small": small_scene,
    "object": object_scene,
    "batched": batched_scene,
    "scene": scene_scale,
}
