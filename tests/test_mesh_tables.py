"""Marching-cubes case tables (supernormal_amd/csrc/mc_tables.h, derived by
gen_mc_tables.py) and the CPU marching-cubes / ray-mesh restatements of
oracle/mesh_oracle.py, pinned by what the geometry fixes:
  * the committed header is what the generator writes;
  * every case's triangles use exactly the edges whose corners straddle the
    level, and their boundary is one closed loop per run of inside corners
    on each face;
  * vertices sit at the linear interpolation of the values along their edge
    (closed form), so the vertex set is the one any marching-cubes table
    gives (PyMCubes' included, models/renderer.py:27-34);
  * meshes of random fields that stay inside the grid are closed and
    consistently oriented (every directed edge's reverse appears once), and a
    sphere's mesh has the sphere's signed volume;
  * brute-force ray hits on a sphere's mesh agree with the analytic
    ray-sphere intersection within the grid's interpolation error.
"""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

from oracle import mesh_oracle

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "supernormal_amd" / "csrc"


def _gen():
    spec = importlib.util.spec_from_file_location("gen_mc_tables", CSRC / "gen_mc_tables.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_header_is_generated():
    assert (CSRC / "mc_tables.h").read_text() == _gen().header_text()


def test_cases_use_exactly_the_crossing_edges():
    g = _gen()
    cnt, tab = mesh_oracle.mc_table()
    for case in range(256):
        inside = [(case >> c) & 1 for c in range(8)]
        cross = {i for i, (a, b) in enumerate(g.EDGES) if inside[a] != inside[b]}
        used = {e for e in tab[case] if e >= 0}
        assert used == cross, case
        assert (tab[case, 3 * cnt[case]:] == -1).all()
        assert cnt[case] == 0 or (tab[case, :3 * cnt[case]] >= 0).all()
        # boundary of the triangle set: directed edges not cancelled by a
        # reverse; it must be the face segments, each crossing edge starting
        # exactly one and ending exactly one
        tri = tab[case, :3 * cnt[case]].reshape(-1, 3)
        dir_edges = {}
        for t in tri:
            for i in range(3):
                e = (t[i], t[(i + 1) % 3])
                dir_edges[e] = dir_edges.get(e, 0) + 1
        bnd = [e for e, c in dir_edges.items() for _ in range(c)
               if dir_edges.get((e[1], e[0]), 0) == 0]
        starts = sorted(a for a, _ in bnd)
        ends = sorted(b for _, b in bnd)
        assert starts == sorted(cross) and ends == sorted(cross), case


def test_case_counts_match_known_totals():
    # per case, a fan over each loop: triangles = crossing edges - 2 x loops;
    # with inside corners kept apart the classic case list has 0..5 triangles
    cnt, _ = mesh_oracle.mc_table()
    assert cnt[0] == 0 and cnt[255] == 0
    for c in range(8):
        assert cnt[1 << c] == 1 and cnt[255 ^ (1 << c)] == 1
    assert cnt.max() == 5


def _field_inside(rng, shape):
    u = rng.standard_normal(shape).astype(np.float32)
    u[0, :, :] = u[-1, :, :] = 1.0
    u[:, 0, :] = u[:, -1, :] = 1.0
    u[:, :, 0] = u[:, :, -1] = 1.0
    return u


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_fields_give_closed_oriented_meshes(seed):
    rng = np.random.default_rng(seed)
    u = _field_inside(rng, (10, 9, 12))
    tris, ids = mesh_oracle.marching_cubes(u, 0.1)
    assert len(tris) > 100
    edges = {}
    for t in ids:
        for i in range(3):
            e = (t[i], t[(i + 1) % 3])
            edges[e] = edges.get(e, 0) + 1
    assert all(c == 1 for c in edges.values())
    assert all(edges.get((b, a), 0) == 1 for (a, b) in edges)


def test_vertices_at_linear_interpolation():
    rng = np.random.default_rng(3)
    u = rng.standard_normal((7, 8, 6)).astype(np.float32)
    iso = np.float32(-0.2)
    tris, ids = mesh_oracle.marching_cubes(u, iso)
    R = np.array(u.shape)
    for p, gid in zip(tris.reshape(-1, 3), ids.reshape(-1)):
        ax = gid % 3
        lin = gid // 3
        lo = np.array([lin // (R[1] * R[2]), (lin // R[2]) % R[1], lin % R[2]])
        hi = lo.copy()
        hi[ax] += 1
        a, b = u[tuple(lo)] - iso, u[tuple(hi)] - iso
        assert (a < 0) != (b < 0)
        s = float(np.float32(a) / np.float32(a - b))
        want = lo.astype(np.float64)
        want[ax] += s
        np.testing.assert_allclose(p, want, atol=1e-6)


def _sphere(r, rad):
    ax = np.linspace(-1, 1, r)
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    return (np.sqrt(x * x + y * y + z * z) - rad).astype(np.float32)


def test_sphere_volume_and_orientation():
    r, rad = 41, 0.6
    tris, ids = mesh_oracle.marching_cubes(_sphere(r, rad), 0.0)
    w = tris * (2.0 / (r - 1)) - 1.0
    rn = np.linalg.norm(w.reshape(-1, 3), axis=1)
    assert np.abs(rn - rad).max() < 2.0 / (r - 1)
    vol = np.einsum("ij,ij->i", w[:, 0], np.cross(w[:, 1], w[:, 2])).sum() / 6.0
    assert abs(vol - 4.0 / 3.0 * np.pi * rad ** 3) < 0.02


def test_ray_hits_on_sphere_mesh_match_analytic():
    r, rad = 41, 0.6
    tris, _ = mesh_oracle.marching_cubes(_sphere(r, rad), 0.0)
    w = tris * (2.0 / (r - 1)) - 1.0
    rng = np.random.default_rng(4)
    n = 300
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * 2.5
    tgt = rng.uniform(-0.7, 0.7, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t, tri = mesh_oracle.ray_mesh_first_hit(o, d, w)
    b = (o * d).sum(1)
    disc = b * b - ((o * o).sum(1) - rad * rad)
    hit_a = disc > 0
    t_a = -b - np.sqrt(np.maximum(disc, 0))
    clear = np.abs(disc) > 0.02   # rays not grazing the sphere
    assert ((tri >= 0) == hit_a)[clear].all()
    both = hit_a & (tri >= 0)
    assert np.abs(t[both] - t_a[both]).max() < 0.05
