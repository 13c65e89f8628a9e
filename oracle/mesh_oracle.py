"""CPU restatement of the surface extraction (marching tetrahedra) used to
check the GPU kernels (supernormal_amd/csrc/mesh.hip).  Test infrastructure:
only tests/ import it.

Pinned by: a sphere's analytic SDF sampled on a grid gives vertices on the
sphere up to the linear interpolation error; the triangle count of a closed
surface is even and every edge is shared by two triangles (closed, consistently
oriented: the signed volume equals the sphere's up to discretisation); every
case of a tetrahedron (1-3 corners inside) appears in the random-field test.
"""
from __future__ import annotations

import numpy as np

# Marching tetrahedra: each grid cube is split into 6 tetrahedra along the
# main diagonal; each tetrahedron yields 0, 1 or 2 triangles.  (The
# reference uses PyMCubes' marching cubes, which is not installed; the
# surfaces agree up to the triangulation inside each cube.)
_CUBE_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                          [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
_TETS = np.array([[0, 5, 1, 6], [0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6],
                  [0, 4, 5, 6]])


def marching_tetrahedra(u, threshold=0.0):
    """Triangle mesh of {u = threshold}: (vertices [V,3] in grid units, faces [F,3])."""
    R = np.array(u.shape)
    idx = np.stack(np.meshgrid(*[np.arange(r - 1) for r in R], indexing="ij"), -1).reshape(-1, 3)
    vals = np.stack([u[idx[:, 0] + c[0], idx[:, 1] + c[1], idx[:, 2] + c[2]] for c in _CUBE_CORNERS],
                    -1) - threshold
    inside = vals < 0
    any_cross = inside.any(1) & (~inside).any(1)
    idx, vals = idx[any_cross], vals[any_cross]
    tris = []
    for tet in _TETS:
        tv = vals[:, tet]
        tp = (idx[:, None, :] + _CUBE_CORNERS[tet][None]).astype(np.float64)
        neg = tv < 0
        code = neg.astype(np.int64) @ np.array([1, 2, 4, 8])
        # direction from the inside corners to the outside corners (orientation)
        n_in = neg.sum(1, keepdims=True)
        out_dir = (tp * (~neg)[..., None]).sum(1) / np.maximum(4 - n_in, 1) - \
            (tp * neg[..., None]).sum(1) / np.maximum(n_in, 1)

        def edge(i, j, sel):
            a, b = tv[sel, i], tv[sel, j]
            t = a / (a - b)
            return tp[sel, i] + (tp[sel, j] - tp[sel, i]) * t[:, None]

        for lone in range(4):  # one corner on its own side: one triangle
            others = [k for k in range(4) if k != lone]
            for c in (1 << lone, 15 - (1 << lone)):
                sel = code == c
                if sel.any():
                    tri = np.stack([edge(lone, k, sel) for k in others], 1)
                    tris.append((tri, out_dir[sel]))
        for a_, b_ in ((0, 1), (0, 2), (0, 3)):  # two and two: a quad
            c_ = [k for k in range(4) if k not in (a_, b_)]
            for c in ((1 << a_) | (1 << b_), 15 - ((1 << a_) | (1 << b_))):
                sel = code == c
                if sel.any():
                    p0, p1 = edge(a_, c_[0], sel), edge(a_, c_[1], sel)
                    p2, p3 = edge(b_, c_[1], sel), edge(b_, c_[0], sel)
                    tris.append((np.stack([p0, p1, p2], 1), out_dir[sel]))
                    tris.append((np.stack([p0, p2, p3], 1), out_dir[sel]))
    verts, faces = [], []
    nv = 0
    for tri, od in tris:
        nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        flip = (nrm * od).sum(1) < 0
        tri = np.where(flip[:, None, None], tri[:, ::-1], tri)
        n = tri.shape[0]
        verts.append(tri.reshape(-1, 3))
        faces.append(nv + np.arange(3 * n).reshape(n, 3))
        nv += 3 * n
    if not verts:
        return np.zeros((0, 3)), np.zeros((0, 3), dtype=np.int64)
    return np.concatenate(verts), np.concatenate(faces)




def triangles(u, threshold, bound_min, bound_max):
    """[T, 3, 3] world-space triangles of {u = threshold}, inside = u < threshold."""
    v, f = marching_tetrahedra(u, threshold)
    bmin = np.asarray(bound_min, dtype=np.float64)
    bmax = np.asarray(bound_max, dtype=np.float64)
    r = np.array(u.shape, dtype=np.float64)
    v = v / (r - 1.0)[None, :] * (bmax - bmin)[None, :] + bmin[None, :]
    return v[f]


def canonical(tris, decimals=5):
    """Triangles as a sorted array, each rotated (orientation kept) to start at
    its lexicographically smallest vertex, coordinates rounded."""
    t = np.round(np.asarray(tris, dtype=np.float64), decimals)
    out = []
    for tri in t:
        keys = [tuple(p) for p in tri]
        i = min(range(3), key=lambda k: keys[k])
        out.append(np.concatenate([tri[i], tri[(i + 1) % 3], tri[(i + 2) % 3]]))
    out = np.array(out).reshape(-1, 9)
    return out[np.lexsort(out.T[::-1])]


# ---------------------------------------------------------------------------
# Marching cubes (the mesher of renderer.py:27-34, PyMCubes) with the case
# table of supernormal_amd/csrc/mc_tables.h read as data, and the first hit of
# rays with a triangle mesh (exp_runner.py:618-630, trimesh's ray query) by
# brute force in fp64.  Pinned by: every vertex lies on a grid edge whose ends
# straddle the iso level, at the linear interpolation of the values (closed
# form); the welded mesh of a field that crosses no grid boundary is closed and
# consistently oriented, and for a sphere its signed volume is the sphere's;
# ray hits of a sphere's mesh agree with the analytic ray-sphere intersection
# within the grid's interpolation error.
# ---------------------------------------------------------------------------
_MC_EDGES = np.array([[0, 1], [1, 2], [3, 2], [0, 3], [4, 5], [5, 6], [7, 6], [4, 7],
                      [0, 4], [1, 5], [2, 6], [3, 7]])   # (lower corner, higher corner)


def mc_table(path=None):
    """(tri_count [256], tri_table [256,16]) parsed from mc_tables.h."""
    import re
    from pathlib import Path
    p = Path(path) if path else \
        Path(__file__).resolve().parents[1] / "supernormal_amd" / "csrc" / "mc_tables.h"
    text = p.read_text()

    def block(name):
        body = text[text.index(name):]
        body = body[body.index("{") + 1:body.index("};")]
        return [int(v, 0) for v in re.findall(r"-?0x[0-9a-fA-F]+|-?\d+", body)]

    cnt = np.array(block("kMcTriCount"), dtype=np.int64)
    tab = np.array(block("kMcTriTable"), dtype=np.int64).reshape(256, 16)
    return cnt, tab


def marching_cubes(u, threshold=0.0, table=None):
    """Triangles [T,3,3] in grid units and vertex edge ids [T,3] of
    {u = threshold}, cubes in x-major order, each cube's triangles in table
    order (the GPU kernels' order)."""
    cnt, tab = mc_table() if table is None else table
    u = np.asarray(u, dtype=np.float32)
    R = np.array(u.shape)
    idx = np.stack(np.meshgrid(*[np.arange(r - 1) for r in R], indexing="ij"), -1).reshape(-1, 3)
    vals = np.stack([u[idx[:, 0] + c[0], idx[:, 1] + c[1], idx[:, 2] + c[2]] for c in _CUBE_CORNERS],
                    -1) - np.float32(threshold)
    case = ((vals < 0).astype(np.int64) << np.arange(8)).sum(1)
    keep = cnt[case] > 0
    idx, vals, case = idx[keep], vals[keep], case[keep]
    tris, ids = [], []
    for q in range(5):
        sel = cnt[case] > q
        if not sel.any():
            continue
        for c in range(3):
            e = tab[case[sel], 3 * q + c]
            lo, hi = _MC_EDGES[e, 0], _MC_EDGES[e, 1]
            ci = np.arange(sel.sum())
            va, vb = vals[sel][ci, lo], vals[sel][ci, hi]
            s = va / (va - vb)
            p = (idx[sel] + _CUBE_CORNERS[lo]).astype(np.float64)
            ax = np.argmax(_CUBE_CORNERS[hi] - _CUBE_CORNERS[lo], axis=1)
            p[ci, ax] += s
            gid = ((idx[sel] + _CUBE_CORNERS[lo]) * np.array([R[1] * R[2], R[2], 1])).sum(1)
            tris.append((np.flatnonzero(sel), q, c, p, 3 * gid + ax))
    n = len(idx)
    T = np.zeros((n, 5, 3, 3))
    I = np.full((n, 5, 3), -1, dtype=np.int64)
    for rows, q, c, p, g in tris:
        T[rows, q, c] = p
        I[rows, q, c] = g
    valid = np.arange(5)[None, :] < cnt[case][:, None]
    return T[valid], I[valid]


def ray_mesh_first_hit(o, d, tris, t_near=None, t_far=None, chunk=256):
    """Nearest hit of each ray with triangles [T,3,3] (fp64 Moller-Trumbore,
    brute force): (t [N] (inf if none), triangle [N] (-1))."""
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    tr = np.asarray(tris, np.float64)
    N = o.shape[0]
    t_near = np.zeros(N) if t_near is None else np.asarray(t_near, np.float64)
    t_far = np.full(N, np.inf) if t_far is None else np.asarray(t_far, np.float64)
    e1, e2 = tr[:, 1] - tr[:, 0], tr[:, 2] - tr[:, 0]
    best = np.full(N, np.inf)
    best_i = np.full(N, -1, dtype=np.int64)
    for s in range(0, N, chunk):
        oo, dd = o[s:s + chunk, None, :], d[s:s + chunk, None, :]
        pv = np.cross(dd, e2[None])
        det = (e1[None] * pv).sum(-1)
        ok = np.abs(det) > 1e-30
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        tv = oo - tr[None, :, 0]
        u = (tv * pv).sum(-1) * inv
        qv = np.cross(tv, e1[None])
        v = (dd * qv).sum(-1) * inv
        t = (e2[None] * qv).sum(-1) * inv
        hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= t_near[s:s + chunk, None]) & \
            (t <= t_far[s:s + chunk, None])
        t = np.where(hit, t, np.inf)
        j = t.argmin(1)
        bt = t[np.arange(t.shape[0]), j]
        best[s:s + chunk] = bt
        best_i[s:s + chunk] = np.where(np.isfinite(bt), j, -1)
    return best, best_i
