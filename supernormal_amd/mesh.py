"""Evaluation of a trained field: visible surface points and Chamfer / F-score.

The reference scores a run by extracting a marching-cubes mesh
(renderer.py:9-34), intersecting it with the rays of the training pixels
inside the masks (exp_runner.py:618-630, first hits only) and comparing the
hits with the same points on the ground-truth mesh by Chamfer distance and
F-score (models/cd_and_fscore.py:5-29).

Here the first hits are taken on the learned SDF directly: each pixel ray is
marched from the unit sphere with a fixed step until the SDF changes sign,
and the crossing is refined by bisection.  That is the point the reference's
ray-mesh intersection converges to as the marching-cubes resolution grows,
without building a mesh.  Chamfer / F-score use exact nearest neighbours
computed on the GPU in fp64 (brute force in tiles), i.e. the same numbers
scipy's KDTree gives in cd_and_fscore.py.
"""
from __future__ import annotations

import numpy as np
import torch


@torch.no_grad()
def first_hits(sdf_fn, rays_o, rays_d, near, far, step=2e-3, n_bisect=10, chunk=1 << 18,
               check_every=16):
    """First + to - crossing of sdf_fn along each ray in [near, far]; returns
    (points [M,3], hit mask [N]).  Every step evaluates all rays of the chunk
    (finished rays keep their state), so a step is a fixed set of launches;
    whether all rays are finished is read back every `check_every` steps
    (the extra steps change nothing)."""
    dev = rays_o.device
    N = rays_o.shape[0]
    pts = torch.zeros(N, 3, device=dev)
    hit = torch.zeros(N, dtype=torch.bool, device=dev)
    for s in range(0, N, chunk):
        o, d = rays_o[s:s + chunk], rays_d[s:s + chunk]
        t = near[s:s + chunk].clone()
        t_hi_lim = far[s:s + chunk]
        f_lo = sdf_fn(o + d * t[:, None])
        done = torch.zeros(o.shape[0], dtype=torch.bool, device=dev)
        found = torch.zeros_like(done)
        t_hit_lo = torch.zeros_like(t)
        t_hit_hi = torch.zeros_like(t)
        one = torch.ones_like(f_lo)
        k = 0
        while k % check_every != 0 or not bool(done.all()):
            act = ~done
            t_new = torch.where(act, t + step, t)
            f_new = torch.where(act, sdf_fn(o + d * t_new[:, None]), one)
            cross = act & (f_lo > 0) & (f_new <= 0)
            t_hit_lo = torch.where(cross, t, t_hit_lo)
            t_hit_hi = torch.where(cross, t_new, t_hit_hi)
            found |= cross
            done |= cross | (t_new >= t_hi_lim)
            t = t_new
            f_lo = torch.where(act, f_new, f_lo)
            k += 1
        a, b = t_hit_lo, t_hit_hi
        for _ in range(n_bisect):
            m = 0.5 * (a + b)
            fm = sdf_fn(o + d * m[:, None])
            a = torch.where(fm > 0, m, a)
            b = torch.where(fm > 0, b, m)
        tm = 0.5 * (a + b)
        pts[s:s + chunk] = o + d * tm[:, None]
        hit[s:s + chunk] = found
    return pts[hit], hit


@torch.no_grad()
def visible_surface_points(sdf_fn, dataset, rays_per_view=20000, seed=0, step=2e-3):
    """First hits of the learned SDF along masked training-pixel rays (the
    rays of all views marched together), without building a mesh."""
    o, d = _pixel_rays(dataset, rays_per_view, seed)
    if o is None:
        return np.zeros((0, 3), np.float32)
    near, far = dataset.near_far_from_sphere(o, d)
    pts, _ = first_hits(sdf_fn, o, d, near, far, step=step)
    return pts.cpu().numpy()


@torch.no_grad()
def nn_distances(a, b, tile=4096):
    """For each point of a, the fp64 Euclidean distance to its nearest point of b."""
    a = torch.as_tensor(a, dtype=torch.float64)
    b = torch.as_tensor(b, dtype=torch.float64)
    dev = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
    a, b = a.to(dev), b.to(dev)
    out = torch.empty(a.shape[0], dtype=torch.float64, device=dev)
    bb = (b * b).sum(-1)
    for s in range(0, a.shape[0], tile):
        x = a[s:s + tile]
        best = torch.full((x.shape[0],), float("inf"), dtype=torch.float64, device=dev)
        xx = (x * x).sum(-1, keepdim=True)
        for t in range(0, b.shape[0], 1 << 15):
            y = b[t:t + (1 << 15)]
            d2 = xx + bb[t:t + (1 << 15)][None, :] - 2.0 * (x @ y.T)
            best = torch.minimum(best, d2.min(dim=1).values)
        out[s:s + tile] = best.clamp_min(0).sqrt()
    return out.cpu().numpy()


def chamfer_distance_and_f1_score(ref_points, eval_points, f_threshold=0.5):
    """models/cd_and_fscore.py:5-29 with exact fp64 nearest neighbours:
    CD = (mean d(eval->ref) + mean d(ref->eval)) / 2, F = 2PR/(P+R) at f_threshold."""
    d_e2r = nn_distances(eval_points, ref_points)
    d_r2e = nn_distances(ref_points, eval_points)
    cd = (np.mean(d_e2r) + np.mean(d_r2e)) / 2
    precision = np.mean(d_e2r < f_threshold)
    recall = np.mean(d_r2e < f_threshold)
    f = 2 * precision * recall / (precision + recall) if precision + recall > 0 else 0.0
    return float(cd), float(f)


def extract_fields(bound_min, bound_max, resolution, query_func, chunk=64, device=None):
    """SDF values on a resolution^3 grid (renderer.py:9-24)."""
    device = device or (torch.device("cuda") if torch.cuda.is_available() else "cpu")
    axes = [torch.linspace(float(bound_min[i]), float(bound_max[i]), resolution) for i in range(3)]
    u = np.zeros([resolution] * 3, dtype=np.float32)
    with torch.no_grad():
        for xi, xs in enumerate(axes[0].split(chunk)):
            for yi, ys in enumerate(axes[1].split(chunk)):
                for zi, zs in enumerate(axes[2].split(chunk)):
                    xx, yy, zz = torch.meshgrid(xs, ys, zs, indexing="ij")
                    pts = torch.stack([xx.reshape(-1), yy.reshape(-1), zz.reshape(-1)], -1)
                    val = query_func(pts.to(device)).reshape(len(xs), len(ys), len(zs))
                    u[xi * chunk: xi * chunk + len(xs), yi * chunk: yi * chunk + len(ys),
                      zi * chunk: zi * chunk + len(zs)] = val.detach().cpu().numpy()
    return u


def extract_fields_device(bound_min, bound_max, resolution, query_func, device, slab=32):
    """Field values on a resolution^3 grid, kept on the device (renderer.py:9-24:
    torch.linspace axes, x-major grid)."""
    r = int(resolution)
    axes = [torch.linspace(float(bound_min[i]), float(bound_max[i]), r, device=device)
            for i in range(3)]
    u = torch.empty((r, r, r), dtype=torch.float32, device=device)
    yy, zz = torch.meshgrid(axes[1], axes[2], indexing="ij")
    with torch.no_grad():
        for x0 in range(0, r, slab):
            xs = axes[0][x0:x0 + slab]
            pts = torch.stack([xs[:, None, None].expand(-1, r, r), yy[None].expand(len(xs), -1, -1),
                               zz[None].expand(len(xs), -1, -1)], -1).reshape(-1, 3)
            u[x0:x0 + len(xs)] = query_func(pts.contiguous()).reshape(len(xs), r, r)
    return u


def marching_tetrahedra_device(w, iso, bound_min, bound_max):
    """Triangle soup of {w = iso} on the device (sn_mt_count / scan / sn_mt_emit):
    float32 [T, 3, 3] world-space vertices, oriented from w < iso to w > iso."""
    from . import _lib
    from ._lib import call, ptr, stream_of, workspace
    w = w.contiguous().to(torch.float32)
    rx, ry, rz = (int(v) for v in w.shape)
    dev = w.device
    s = stream_of(dev)
    n = (rx - 1) * (ry - 1) * (rz - 1)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    call("sn_mt_count", ptr(w), rx, ry, rz, float(iso), ptr(counts), s)
    packed = torch.empty((n, 2), dtype=torch.int32, device=dev)
    total = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = workspace(_lib.lib().sn_scan_workspace_bytes(n), dev)
    call("sn_counts_to_packed_info", ptr(counts), n, ptr(packed), ptr(total), ptr(ws),
         ws.numel(), s)
    t = int(total.item())
    tris = torch.empty((max(t, 1), 3, 3), dtype=torch.float32, device=dev)
    bmin = (ctypes_floats(bound_min))
    bmax = (ctypes_floats(bound_max))
    if t:
        call("sn_mt_emit", ptr(w), rx, ry, rz, float(iso), bmin, bmax, ptr(packed), ptr(tris), s)
    return tris[:t]


class DeviceMesh:
    """A marching-cubes mesh kept on the device with the grid it came from:
    triangles [T,3,3] (world space), packed_info [(rx-1)(ry-1)(rz-1), 2] of
    each cube's triangles (base, count), vertex edge ids [T,3] (or None), the
    grid's shape and bounds."""

    def __init__(self, tris, packed, shape, bound_min, bound_max, vertex_ids=None):
        self.tris, self.packed, self.shape = tris, packed, tuple(int(v) for v in shape)
        self.bound_min = [float(v) for v in bound_min]
        self.bound_max = [float(v) for v in bound_max]
        self.vertex_ids = vertex_ids

    def indexed(self):
        """(vertices [V,3] float64, triangles [T,3] int64) with one vertex per
        grid edge, as PyMCubes returns them (numpy, host)."""
        if self.vertex_ids is None:
            raise ValueError("mesh built without vertex ids")
        ids = self.vertex_ids.reshape(-1)
        uniq, inv = torch.unique(ids, return_inverse=True)
        first = torch.full((uniq.numel(),), ids.numel(), dtype=torch.int64, device=ids.device)
        first.scatter_reduce_(0, inv, torch.arange(ids.numel(), device=ids.device), "amin")
        v = self.tris.reshape(-1, 3)[first]
        return v.cpu().numpy().astype(np.float64), inv.reshape(-1, 3).cpu().numpy()


def marching_cubes_device(w, iso, bound_min, bound_max, vertex_ids=True):
    """Marching cubes of {w = iso} on the device (sn_mc_count / scan /
    sn_mc_emit; PyMCubes' mesher in renderer.py:27-34): a DeviceMesh whose
    triangles are wound towards w > iso."""
    from . import _lib
    from ._lib import call, ptr, stream_of, workspace
    w = w.contiguous().to(torch.float32)
    rx, ry, rz = (int(v) for v in w.shape)
    dev = w.device
    s = stream_of(dev)
    n = (rx - 1) * (ry - 1) * (rz - 1)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    call("sn_mc_count", ptr(w), rx, ry, rz, float(iso), ptr(counts), s)
    packed = torch.empty((n, 2), dtype=torch.int32, device=dev)
    total = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = workspace(_lib.lib().sn_scan_workspace_bytes(n), dev)
    call("sn_counts_to_packed_info", ptr(counts), n, ptr(packed), ptr(total), ptr(ws),
         ws.numel(), s)
    t = int(total.item())
    tris = torch.empty((max(t, 1), 3, 3), dtype=torch.float32, device=dev)
    ids = torch.empty((max(t, 1), 3), dtype=torch.int64, device=dev) if vertex_ids else None
    if t:
        call("sn_mc_emit", ptr(w), rx, ry, rz, float(iso), ctypes_floats(bound_min),
             ctypes_floats(bound_max), ptr(packed), ptr(tris), ptr(ids), s)
    return DeviceMesh(tris[:t], packed, (rx, ry, rz), bound_min, bound_max,
                      ids[:t] if ids is not None else None)


@torch.no_grad()
def mesh_first_hits(mesh, rays_o, rays_d, near=None, far=None):
    """(t [N], triangle [N]) of each ray's first hit with a DeviceMesh
    (sn_mesh_first_hit; trimesh's intersects_location with multiple_hits=False
    in exp_runner.py:618-630); t = inf and triangle = -1 for a miss."""
    from ._lib import call, ptr, stream_of
    o = rays_o.contiguous().to(torch.float32)
    d = rays_d.contiguous().to(torch.float32)
    n = o.shape[0]
    dev = o.device
    t = torch.empty(n, dtype=torch.float32, device=dev)
    tri = torch.empty(n, dtype=torch.int32, device=dev)
    nr = near.contiguous().to(torch.float32) if near is not None else None
    fr = far.contiguous().to(torch.float32) if far is not None else None
    rx, ry, rz = mesh.shape
    call("sn_mesh_first_hit", ptr(o), ptr(d), ptr(nr), ptr(fr), n, rx, ry, rz,
         ctypes_floats(mesh.bound_min), ctypes_floats(mesh.bound_max), ptr(mesh.packed),
         ptr(mesh.tris), ptr(t), ptr(tri), stream_of(dev))
    return t, tri


def _pixel_rays(dataset, rays_per_view, seed):
    """Unit-direction rays of masked training pixels, all views together
    (gen_rays_at(within_mask=True), exp_runner.py:621-623); rays_per_view
    None = every masked pixel."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    rays = []
    for i in dataset.train_images:
        m = dataset.masks[i].reshape(-1).nonzero()[:, 0].cpu()
        if m.numel() == 0:
            continue
        if rays_per_view is not None:
            m = m[torch.randint(0, m.numel(), (min(rays_per_view, m.numel()),), generator=g)]
        ys, xs = (m // dataset.W).to(dataset.device), (m % dataset.W).to(dataset.device)
        p = torch.stack([xs.float(), ys.float(), torch.ones_like(xs, dtype=torch.float32)], -1)
        p = p @ dataset.intrinsics_all_inv[i, :3, :3].T
        d = p / p.norm(dim=-1, keepdim=True)
        d = d @ dataset.pose_all[i, :3, :3].T
        o = dataset.pose_all[i, :3, 3].expand(d.shape).contiguous()
        rays.append((o, d))
    if not rays:
        return None, None
    return torch.cat([r[0] for r in rays]), torch.cat([r[1] for r in rays])


@torch.no_grad()
def visible_mesh_points(mesh, dataset, rays_per_view=None, seed=0):
    """find_visible_points (exp_runner.py:618-630): first hits of masked
    training-pixel rays with the extracted mesh; numpy [M,3]."""
    o, d = _pixel_rays(dataset, rays_per_view, seed)
    if o is None:
        return np.zeros((0, 3), np.float32)
    d = d / d.norm(dim=-1, keepdim=True)
    t, tri = mesh_first_hits(mesh, o, d)
    hit = tri >= 0
    return (o[hit] + d[hit] * t[hit, None]).cpu().numpy()


def ctypes_floats(v):
    import ctypes
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def extract_geometry(bound_min, bound_max, resolution, threshold, query_func, device=None,
                     mesher="mc"):
    """(vertices [V, 3] in world units, triangles [T, 3]) of {query_func =
    threshold} (renderer.py:27-34, where query_func is -sdf and the mesher is
    PyMCubes' marching cubes).  mesher "mc": marching cubes on the GPU, one
    vertex per grid edge crossing the level (PyMCubes' vertex set);
    "mt": marching tetrahedra (a triangle soup, 3 vertices per triangle).
    Triangles are oriented towards query_func < threshold (outside for
    query_func = -sdf)."""
    device = device or torch.device("cuda")
    u = extract_fields_device(bound_min, bound_max, resolution, query_func, device)
    if mesher == "mc":
        return marching_cubes_device(-u, -float(threshold), bound_min, bound_max).indexed()
    if mesher != "mt":
        raise ValueError(f"unknown mesher {mesher!r}")
    tris = marching_tetrahedra_device(-u, -float(threshold), bound_min, bound_max)
    v = tris.reshape(-1, 3).cpu().numpy().astype(np.float64)
    f = np.arange(v.shape[0], dtype=np.int64).reshape(-1, 3)
    return v, f
