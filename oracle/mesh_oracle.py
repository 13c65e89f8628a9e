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
