"""Synthetic scenes and the training dataset (the input side of the hot path).

The reference trains on DiLiGenT-MV normal maps loaded from disk
(models/dataset_loader.py:39-150); none ship with it and there is no network
here, so scenes are analytic SDFs rendered to noisy world-space normal maps
plus masks by sphere tracing, with cameras in the reference's convention
(OpenCV: camera x right, y down, z forward; pose = camera-to-world).

``Dataset`` keeps the reference's interface: normals [n,H,W,3], masks,
intrinsics_all(_inv), pose_all, V_inverse_all [n,H,W,3,3] (dataset_loader.py:
114-137), gen_random_patches (223-277) and near_far_from_sphere (279-297).
The patch sampler draws the view indices from the device RNG (the reference
uses numpy's host RNG, dataset_loader.py:252, then copies to the GPU).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch


# --------------------------------------------------------------------------- #
# analytic SDF primitives
# --------------------------------------------------------------------------- #
def sd_sphere(p, c, r):
    return (p - c).norm(dim=-1) - r


def sd_box(p, c, half):
    q = (p - c).abs() - half
    return q.clamp(min=0).norm(dim=-1) + q.max(dim=-1).values.clamp(max=0)


def sd_torus(p, c, R, r):
    q = p - c
    xz = torch.sqrt(q[..., 0] ** 2 + q[..., 2] ** 2) - R
    return torch.sqrt(xz ** 2 + q[..., 1] ** 2) - r


def sd_capsule(p, a, b, r):
    pa, ba = p - a, b - a
    h = ((pa * ba).sum(-1) / (ba * ba).sum(-1)).clamp(0, 1)
    return (pa - ba * h[..., None]).norm(dim=-1) - r


def smin(a, b, k=0.05):
    h = (0.5 + 0.5 * (b - a) / k).clamp(0, 1)
    return b * (1 - h) + a * h - k * h * (1 - h)


@dataclass
class Scene:
    """An analytic SDF f: [N,3] -> [N] (torch, any device)."""
    name: str
    fn: Callable = field(repr=False)

    def __call__(self, p):
        return self.fn(p)


def sphere_scene(radius=0.5):
    """Sphere of radius 0.5 at the origin (the CPU-sized workload)."""
    return Scene("sphere", lambda p: sd_sphere(p, torch.zeros(3, device=p.device, dtype=p.dtype),
                                               radius))


def torus_box_spheres_scene(seed=0):
    """Smooth union of a torus (R 0.4, r 0.12), a box of side 0.3 and 3 seeded
    random spheres (the full-size single-object workload)."""
    rng = np.random.default_rng(seed)
    sph_c = rng.uniform(-0.45, 0.45, (3, 3))
    sph_r = rng.uniform(0.08, 0.16, 3)
    box_c = np.array([0.0, 0.15, 0.0])

    def fn(p):
        t = lambda a: torch.as_tensor(a, device=p.device, dtype=p.dtype)
        d = sd_torus(p, t([0.0, -0.05, 0.0]), 0.4, 0.12)
        d = smin(d, sd_box(p, t(box_c), 0.15))
        for c, r in zip(sph_c, sph_r):
            d = smin(d, sd_sphere(p, t(c), float(r)))
        return d

    return Scene(f"torus_box_spheres_{seed}", fn)


def random_primitives_scene(seed=0, n=8, ridge_amp=0.005, ridge_freq=40.0):
    """n seeded random primitives (spheres, boxes, tori, capsules), smooth
    union, plus a sinusoidal displacement of amplitude ridge_amp."""
    rng = np.random.default_rng(seed)
    prims = []
    for _ in range(n):
        kind = rng.integers(0, 4)
        c = rng.uniform(-0.4, 0.4, 3)
        if kind == 0:
            prims.append(("s", c, rng.uniform(0.08, 0.2)))
        elif kind == 1:
            prims.append(("b", c, rng.uniform(0.05, 0.15, 3)))
        elif kind == 2:
            prims.append(("t", c, (rng.uniform(0.12, 0.25), rng.uniform(0.03, 0.07))))
        else:
            prims.append(("c", c, (c + rng.uniform(-0.25, 0.25, 3), rng.uniform(0.03, 0.08))))

    def fn(p):
        t = lambda a: torch.as_tensor(np.asarray(a), device=p.device, dtype=p.dtype)
        d = None
        for kind, c, prm in prims:
            if kind == "s":
                di = sd_sphere(p, t(c), float(prm))
            elif kind == "b":
                di = sd_box(p, t(c), t(prm))
            elif kind == "t":
                di = sd_torus(p, t(c), float(prm[0]), float(prm[1]))
            else:
                di = sd_capsule(p, t(c), t(prm[0]), float(prm[1]))
            d = di if d is None else smin(d, di)
        if ridge_amp:
            w = ridge_freq
            d = d + ridge_amp * torch.sin(w * p[..., 0]) * torch.sin(w * p[..., 1]) * \
                torch.sin(w * p[..., 2])
        return d

    return Scene(f"random_primitives_{seed}", fn)


# --------------------------------------------------------------------------- #
# cameras (OpenCV convention, camera-to-world pose)
# --------------------------------------------------------------------------- #
def look_at_pose(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)):
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, np.float64))
    if np.linalg.norm(right) < 1e-6:
        right = np.cross(fwd, np.array([1.0, 0.0, 0.0]))
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    pose = np.eye(4)
    pose[:3, 0], pose[:3, 1], pose[:3, 2], pose[:3, 3] = right, down, fwd, eye
    return pose


def ring_cameras(n, radius=2.5, elevations=(0.0,), seed=0):
    """n cameras on rings around the y axis; elevations (degrees) cycle."""
    poses = []
    for i in range(n):
        el = math.radians(elevations[i % len(elevations)])
        az = 2 * math.pi * i / n
        eye = radius * np.array([math.cos(el) * math.sin(az), -math.sin(el),
                                 -math.cos(el) * math.cos(az)])
        poses.append(look_at_pose(eye))
    return np.stack(poses)


def intrinsics(W, H, fov_deg=35.0):
    f = 0.5 * W / math.tan(math.radians(fov_deg) / 2)
    K = np.eye(4)
    K[0, 0], K[1, 1], K[0, 2], K[1, 2] = f, f, (W - 1) / 2.0, (H - 1) / 2.0
    return K


# --------------------------------------------------------------------------- #
# rendering of normal maps by sphere tracing
# --------------------------------------------------------------------------- #
def scene_normal(scene, p, eps=1e-4):
    e = torch.eye(3, device=p.device, dtype=p.dtype) * eps
    g = torch.stack([scene(p + e[i]) - scene(p - e[i]) for i in range(3)], -1)
    return g / g.norm(dim=-1, keepdim=True).clamp_min(1e-12)


def _pixel_rays(K, pose, W, H, device, chunk):
    """Per chunk of pixels (row-major): origin [3], unit directions, and the
    entry / exit distances of the unit sphere with a flag of the rays that
    meet it."""
    Kinv = torch.inverse(torch.as_tensor(K[:3, :3], dtype=torch.float64, device=device))
    R = torch.as_tensor(pose[:3, :3], dtype=torch.float64, device=device)
    o = torch.as_tensor(pose[:3, 3], dtype=torch.float64, device=device)
    ys, xs = torch.meshgrid(torch.arange(H, device=device, dtype=torch.float64),
                            torch.arange(W, device=device, dtype=torch.float64), indexing="ij")
    pix = torch.stack([xs, ys, torch.ones_like(xs)], -1).reshape(-1, 3)
    out = []
    for s in range(0, H * W, chunk):
        d = pix[s:s + chunk] @ Kinv.T
        d = d / d.norm(dim=-1, keepdim=True)
        d = d @ R.T
        # start at the unit sphere
        b = 2 * (o * d).sum(-1)
        c = (o * o).sum() - 1.0
        disc = b * b - 4 * c
        ok = disc > 0
        t = torch.where(ok, (-b - disc.clamp_min(0).sqrt()) / 2, torch.zeros_like(b)).clamp_min(0)
        tfar = torch.where(ok, (-b + disc.clamp_min(0).sqrt()) / 2, torch.zeros_like(b))
        out.append((o, d, t, tfar, ok))
    return out


def _sphere_trace(scene, o, d, t, tfar, ok, n_iter):
    """First hits along the rays (o [N,3] or [3], d [N,3]): t updated in place;
    returns the hit flags.  Every 8 steps the rays that hit, missed the unit
    sphere or left it (t > tfar: the scene is positive outside it, so t only
    grows) are dropped from the set; each ray's own sequence of steps does
    not depend on which other rays are traced with it."""
    hit = torch.zeros_like(ok)
    act = torch.nonzero(ok).squeeze(1)
    per_ray_o = o.dim() == 2
    for it in range(n_iter):
        if it % 8 == 0 and it:
            act = act[~hit[act] & (t[act] <= tfar[act])]
            if act.numel() == 0:
                break
        ta = t[act]
        f = scene((o[act] if per_ray_o else o) + d[act] * ta[:, None])
        h = hit[act] | ((f.abs() < 1e-5) & (ta <= tfar[act]))
        hit[act] = h
        t[act] = torch.where(h, ta, ta + f * 0.9)
    return hit


@torch.no_grad()
def render_normal_maps(scene, K, poses, W, H, device, n_iter=160, chunk=1 << 20,
                       trace_rays=1 << 23):
    """World-space unit normals [H,W,3] and masks [H,W] of the first hits of
    several views: the rays of the views are sphere-traced together, up to
    trace_rays at a time (a step of the trace is ~a few hundred small
    launches whatever the ray count, so fewer, larger batches take less
    time); per ray the result is the one render_normal_map gives."""
    views = [_pixel_rays(K, pose, W, H, device, chunk) for pose in poses]
    flat = [(v, k) for v, pieces in enumerate(views) for k in range(len(pieces))]
    hits = {}
    i = 0
    while i < len(flat):
        group, n = [], 0
        while i < len(flat) and (not group or n + len(views[flat[i][0]][flat[i][1]][1]) <= trace_rays):
            group.append(flat[i])
            n += len(views[flat[i][0]][flat[i][1]][1])
            i += 1
        pieces = [views[v][k] for v, k in group]
        O = torch.cat([pc[0].expand(len(pc[1]), 3) for pc in pieces])
        D = torch.cat([pc[1] for pc in pieces])
        T = torch.cat([pc[2] for pc in pieces])
        TF = torch.cat([pc[3] for pc in pieces])
        OK = torch.cat([pc[4] for pc in pieces])
        hit = _sphere_trace(scene, O, D, T, TF, OK, n_iter)
        s0 = 0
        for (v, k), pc in zip(group, pieces):
            m = len(pc[1])
            hits[(v, k)] = (T[s0:s0 + m], hit[s0:s0 + m])
            s0 += m
    out = []
    for v, pieces in enumerate(views):
        normals = torch.zeros(H * W, 3, dtype=torch.float64, device=device)
        mask = torch.zeros(H * W, dtype=torch.bool, device=device)
        s = 0
        for k, (o, d, _, tfar, _) in enumerate(pieces):
            t, hit = hits[(v, k)]
            p = o + d * t[:, None]
            hit = hit & (t <= tfar)
            nrm = scene_normal(scene, p)
            m = len(d)
            normals[s:s + m] = torch.where(hit[:, None], nrm, torch.zeros_like(nrm))
            mask[s:s + m] = hit
            s += m
        out.append((normals.reshape(H, W, 3), mask.reshape(H, W)))
    return out


@torch.no_grad()
def render_normal_map(scene, K, pose, W, H, device, n_iter=160, chunk=1 << 20):
    """World-space unit normals [H,W,3] and mask [H,W] of the first hit."""
    return render_normal_maps(scene, K, [pose], W, H, device, n_iter, chunk)[0]


# --------------------------------------------------------------------------- #
# dataset
# --------------------------------------------------------------------------- #
class Dataset:
    """Multi-view normal maps + masks of an analytic scene, in the reference's
    Dataset interface (models/dataset_loader.py)."""

    def __init__(self, scene: Optional[Scene], n_views: int, W: int, H: int, noise_std: float,
                 device, seed: int = 0, elevations=(0.0,), radius: float = 2.5,
                 fov_deg: float = 35.0, exclude_views=()):
        self.device = torch.device(device)
        self.scene = scene
        if scene is None:  # filled by from_arrays
            return
        K = intrinsics(W, H, fov_deg)
        poses = ring_cameras(n_views, radius, elevations, seed)
        g = torch.Generator(device="cpu").manual_seed(seed)
        normals, masks = [], []
        maps = render_normal_maps(scene, K, poses, W, H, self.device)
        for i in range(n_views):
            n, m = maps[i]
            if noise_std > 0:
                noise = torch.randn(H, W, 3, generator=g, dtype=torch.float64).to(self.device)
                n = n + noise_std * noise
                n = n / n.norm(dim=-1, keepdim=True).clamp_min(1e-12)
                n = torch.where(m[..., None], n, torch.zeros_like(n))
            normals.append(n.float())
            masks.append(m.float())
        self._setup(torch.stack(normals), torch.stack(masks), np.stack([K] * n_views),
                    np.asarray(poses), exclude_views)

    @classmethod
    def from_arrays(cls, normals, masks, intrinsics_all, pose_all, device, exclude_views=()):
        """A dataset of given views: normals [n,H,W,3], masks [n,H,W] (0 / 1),
        intrinsics_all / pose_all [n,4,4] (camera-to-world pose)."""
        ds = cls(None, 0, 0, 0, 0.0, device)
        ds._setup(torch.as_tensor(normals, dtype=torch.float32, device=ds.device),
                  torch.as_tensor(masks, dtype=torch.float32, device=ds.device),
                  np.asarray(intrinsics_all), np.asarray(pose_all), exclude_views)
        return ds

    def _setup(self, normals, masks, K_all, poses, exclude_views):
        n_views, H, W = normals.shape[:3]
        self.W, self.H = int(W), int(H)
        self.n_images = n_views
        self.normals = normals.to(self.device)                     # [n,H,W,3]
        self.masks = masks.to(self.device)                         # [n,H,W]
        self.intrinsics_all = torch.as_tensor(K_all, dtype=torch.float32, device=self.device)
        self.intrinsics_all_inv = torch.inverse(self.intrinsics_all)
        self.pose_all = torch.as_tensor(poses, dtype=torch.float32, device=self.device)
        self.exclude_view_list = list(exclude_views)
        self.train_images = [i for i in range(n_views) if i not in self.exclude_view_list]
        self._train_idx = torch.as_tensor(self.train_images, device=self.device)
        # per-pixel V^-1 (dataset_loader.py:114-137): rows (ray dir, camera right, camera down)
        Vinv = []
        ys, xs = torch.meshgrid(torch.arange(H, dtype=torch.float32),
                                torch.arange(W, dtype=torch.float32), indexing="ij")
        pix = torch.stack([xs, ys, torch.ones_like(xs)], -1).to(self.device)
        for i in range(n_views):
            p = pix @ self.intrinsics_all_inv[i, :3, :3].T
            rays_v = p / p.norm(dim=-1, keepdim=True)
            rays_v = rays_v @ self.pose_all[i, :3, :3].T
            right = self.pose_all[i, :3, 0].expand(rays_v.shape)
            down = self.pose_all[i, :3, 1].expand(rays_v.shape)
            V = torch.stack([rays_v, right, down], -2)
            Vinv.append(torch.inverse(V))
        self.V_inverse_all = torch.stack(Vinv)                     # [n,H,W,3,3]
        self.object_bbox_min = np.array([-1.0, -1.0, -1.0])
        self.object_bbox_max = np.array([1.0, 1.0, 1.0])

    @torch.no_grad()
    def gen_random_patches(self, num_patch, patch_H=3, patch_W=3, generator=None):
        """dataset_loader.py:223-277: rays_o/rays_d [B,Ph,Pw,3], marching plane
        normal [B,3], V^-1 [B,Ph,Pw,3,3], normals [B,Ph,Pw,3], mask [B,Ph,Pw,1]."""
        dev = self.device
        cx = torch.randint(patch_W // 2, self.W - 1 - patch_W // 2, (num_patch,), device=dev,
                           generator=generator)
        cy = torch.randint(patch_H // 2, self.H - 1 - patch_H // 2, (num_patch,), device=dev,
                           generator=generator)
        xs = cx[:, None, None] + torch.arange(-patch_W // 2 + 1, patch_W // 2 + 1,
                                              device=dev).repeat(patch_H, 1)
        ys = cy[:, None, None] + torch.arange(-patch_H // 2 + 1, patch_H // 2 + 1,
                                              device=dev).reshape(-1, 1).repeat(1, patch_W)
        sel = torch.randint(0, len(self.train_images), (num_patch,), device=dev,
                            generator=generator)
        img = self._train_idx[sel]
        imgx = img.view(-1, 1, 1).expand_as(xs)
        normal = self.normals[imgx, ys, xs]
        V_inv = self.V_inverse_all[imgx, ys, xs]
        mask = self.masks[imgx, ys, xs].unsqueeze(-1)
        p = torch.stack([xs, ys, torch.ones_like(ys)], -1).float()
        # 3x3 products as broadcast multiply-adds: batched GEMMs of this shape
        # cost far more in launch and tiling than the arithmetic
        p = (self.intrinsics_all_inv[imgx, :3, :3] * p[..., None, :]).sum(-1)
        d = p / p.norm(dim=-1, keepdim=True)
        d = (self.pose_all[img, None, None, :3, :3] * d[..., None, :]).sum(-1)
        o = self.pose_all[img, None, None, :3, 3].expand(d.shape)
        plane_n = self.pose_all[img, :3, 2].expand((num_patch, 3))
        return o, d, plane_n, V_inv, normal, mask

    @staticmethod
    def near_far_from_sphere(rays_o, rays_d):
        """Intersections with the unit sphere (dataset_loader.py:279-297)."""
        a = torch.sum(rays_d ** 2, dim=-1, keepdim=True)
        b = 2.0 * torch.sum(rays_o * rays_d, dim=-1, keepdim=True)
        c = torch.sum(rays_o ** 2, dim=-1, keepdim=True) - 1.0
        mid = 0.5 * (-b) / a
        near = mid - torch.sqrt(b ** 2 - 4 * a * c) / (2 * a)
        far = mid + torch.sqrt(b ** 2 - 4 * a * c) / (2 * a)
        return near[..., 0], far[..., 0]

    def surface_points(self, n=100000, seed=0):
        """Points on the scene's zero level set seen by the training views (the
        reference's visible GT points, exp_runner.py:236-244): the first hits of
        random training pixels."""
        g = torch.Generator(device="cpu").manual_seed(seed)
        pts = []
        per = max(1, n // max(1, len(self.train_images)))
        for i in self.train_images:
            m = self.masks[i].reshape(-1).nonzero()[:, 0].cpu()
            if m.numel() == 0:
                continue
            pick = m[torch.randint(0, m.numel(), (per,), generator=g)]
            ys, xs = pick // self.W, pick % self.W
            x, ok = _first_hits(self.scene, self.intrinsics_all[i].cpu().numpy(),
                                self.pose_all[i].cpu().numpy(), xs, ys, self.device)
            pts.append(x[ok])
        return torch.cat(pts).cpu().numpy()


@torch.no_grad()
def _first_hits(scene, K, pose, xs, ys, device, n_iter=200, tol=1e-5):
    """Sphere tracing of the pixel rays inside the unit sphere; returns the
    points and a mask of the rays that converged onto the surface (a ray that
    grazes an edge the mask covers can leave the sphere without a hit)."""
    Kinv = torch.inverse(torch.as_tensor(K[:3, :3], dtype=torch.float64, device=device))
    R = torch.as_tensor(pose[:3, :3], dtype=torch.float64, device=device)
    o = torch.as_tensor(pose[:3, 3], dtype=torch.float64, device=device)
    pix = torch.stack([xs.double(), ys.double(), torch.ones(len(xs), dtype=torch.float64)], -1)
    d = pix.to(device) @ Kinv.T
    d = d / d.norm(dim=-1, keepdim=True)
    d = d @ R.T
    b = 2 * (o * d).sum(-1)
    c = (o * o).sum() - 1.0
    disc = (b * b - 4 * c).clamp_min(0).sqrt()
    t = ((-b - disc) / 2).clamp_min(0)
    t_far = (-b + disc) / 2
    for _ in range(n_iter):
        t = torch.minimum(t + scene(o + d * t[:, None]) * 0.9, t_far)
    x = o + d * t[:, None]
    return x, scene(x).abs() < tol


def config_dataset(name: str, device, seed: int = 0) -> Dataset:
    """The synthetic workloads: 'sphere128' (8 ring views 128x128, noise 0.01),
    'tbs612' (torus+box+spheres, 20 ring+elevation views 612x512, noise 0.02),
    'prims1024' (8 primitives with ridges, 50 views 1024x1024, noise 0.02)."""
    if name == "sphere128":
        return Dataset(sphere_scene(0.5), 8, 128, 128, 0.01, device, seed)
    if name == "tbs612":
        return Dataset(torus_box_spheres_scene(seed), 20, 612, 512, 0.02, device, seed,
                       elevations=(0.0, 25.0, -25.0, 45.0))
    if name == "prims1024":
        return Dataset(random_primitives_scene(seed), 50, 1024, 1024, 0.02, device, seed,
                       elevations=(0.0, 20.0, -20.0, 40.0, -40.0))
    raise ValueError(f"unknown dataset {name!r}")
