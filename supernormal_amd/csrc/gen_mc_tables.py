"""Generates mc_tables.h: the case tables of marching cubes.

Corner numbering and edge numbering follow the usual marching-cubes
convention (Lorensen & Cline 1987, as tabulated by P. Bourke and used by
PyMCubes, which the reference calls in models/renderer.py:27-34):

  corners  0 (0,0,0)  1 (1,0,0)  2 (1,1,0)  3 (0,1,0)
           4 (0,0,1)  5 (1,0,1)  6 (1,1,1)  7 (0,1,1)
  edges    0: 0-1   1: 1-2   2: 2-3   3: 3-0
           4: 4-5   5: 5-6   6: 6-7   7: 7-4
           8: 0-4   9: 1-5  10: 2-6  11: 3-7

case = sum over corners of (value < iso) << corner.  A vertex sits on every
edge whose two corners are on different sides, so the vertex set of a case
is fixed by the case; only the triangulation is a choice.  The triangles
here are derived, not transcribed: on each face of the cube, every maximal
run of consecutive inside corners (value < iso) is cut off by one segment
from the edge that enters the run to the edge that leaves it (so on a face
whose inside corners are diagonal the two are kept apart; adjacent cubes see
the same face the same way, so the surface has no cracks).  The segments of
the six faces close into loops around the inside corners, and each loop is
cut into a fan of triangles from the first vertex whose fan has no diagonal
lying in a face of the cube (every loop has one).  Triangles are wound so
that their normal (v1 - v0) x (v2 - v0) points from the inside corners to the
outside ones.

Usage:  python gen_mc_tables.py [out.h]   (default: mc_tables.h beside it)
"""
from __future__ import annotations

import sys
from pathlib import Path

CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
           (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
         (0, 4), (1, 5), (2, 6), (3, 7)]
# faces as corner cycles, counter-clockwise seen from outside the cube
FACES = [(0, 3, 2, 1),   # z = 0
         (4, 5, 6, 7),   # z = 1
         (0, 1, 5, 4),   # y = 0
         (3, 7, 6, 2),   # y = 1
         (0, 4, 7, 3),   # x = 0
         (1, 2, 6, 5)]   # x = 1


def edge_of(a, b):
    for i, (p, q) in enumerate(EDGES):
        if (p, q) == (a, b) or (p, q) == (b, a):
            return i
    raise KeyError((a, b))


def case_loops(case):
    """Loops of edge indices for one case (empty for 0 and 255)."""
    inside = [(case >> c) & 1 == 1 for c in range(8)]
    nxt = {}
    for f in FACES:
        for i in range(4):
            a, b = f[i], f[(i + 1) % 4]
            if inside[a] or not inside[b]:
                continue
            # b starts a run of inside corners; find where it ends
            j = (i + 1) % 4
            while inside[f[(j + 1) % 4]]:
                j = (j + 1) % 4
            e_in = edge_of(a, b)
            e_out = edge_of(f[j], f[(j + 1) % 4])
            if e_in in nxt:
                raise AssertionError(f"case {case}: edge {e_in} starts two segments")
            nxt[e_in] = e_out
    loops, seen = [], set()
    for e0 in sorted(nxt):
        if e0 in seen:
            continue
        loop, e = [], e0
        while e not in seen:
            seen.add(e)
            loop.append(e)
            e = nxt[e]
        if e != e0:
            raise AssertionError(f"case {case}: open chain")
        loops.append(loop)
    return loops


def _midpoint(e):
    a, b = EDGES[e]
    return [(CORNERS[a][k] + CORNERS[b][k]) * 0.5 for k in range(3)]


def _sub(p, q):
    return [p[k] - q[k] for k in range(3)]


def _cross(p, q):
    return [p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]]


def _dot(p, q):
    return sum(p[k] * q[k] for k in range(3))


def case_triangles(case):
    """Triangles (edge triples) of one case, wound outward (inside -> outside)."""
    tris = []
    inside = [(case >> c) & 1 == 1 for c in range(8)]
    for loop in case_loops(case):
        # winding of the loop: area vector against the direction from the
        # loop's inside corners (ends of its edges below iso) to its outside ones
        area = [0.0, 0.0, 0.0]
        pts = [_midpoint(e) for e in loop]
        for i in range(len(pts)):
            c = _cross(pts[i], pts[(i + 1) % len(pts)])
            area = [area[k] + c[k] for k in range(3)]
        dirv = [0.0, 0.0, 0.0]
        for e in loop:
            a, b = EDGES[e]
            ci, co = (a, b) if inside[a] else (b, a)
            dirv = [dirv[k] + CORNERS[co][k] - CORNERS[ci][k] for k in range(3)]
        flip = _dot(area, dirv) < 0
        seq = loop[::-1] if flip else loop
        seq = _fan_start(seq)
        for i in range(1, len(seq) - 1):
            tris.append((seq[0], seq[i], seq[i + 1]))
    return tris


FACE_EDGES = [{edge_of(f[i], f[(i + 1) % 4]) for i in range(4)} for f in FACES]


def _same_face(a, b):
    return any(a in fe and b in fe for fe in FACE_EDGES)


def _fan_start(loop):
    """The loop rotated so that no diagonal of the fan from its first vertex
    joins two vertices on one face of the cube (such a diagonal lies in the
    face, where the neighbouring cube's triangles meet it too)."""
    n = len(loop)
    for r in range(n):
        seq = loop[r:] + loop[:r]
        if not any(_same_face(seq[0], seq[i]) for i in range(2, n - 1)):
            return seq
    raise AssertionError(f"loop {loop}: every fan has an in-face diagonal")


def tables():
    edge_mask, tri_table, tri_count = [], [], []
    for case in range(256):
        m = 0
        for i, (a, b) in enumerate(EDGES):
            if ((case >> a) & 1) != ((case >> b) & 1):
                m |= 1 << i
        tris = case_triangles(case)
        used = 0
        for t in tris:
            for e in t:
                used |= 1 << e
        if used != m:
            raise AssertionError(f"case {case}: triangles use edges {used:#x}, crossings {m:#x}")
        if len(tris) > 5:
            raise AssertionError(f"case {case}: {len(tris)} triangles")
        row = [e for t in tris for e in t]
        tri_table.append(row + [-1] * (16 - len(row)))
        tri_count.append(len(tris))
        edge_mask.append(m)
    return edge_mask, tri_count, tri_table


def header_text():
    edge_mask, tri_count, tri_table = tables()
    lines = ["// Generated by gen_mc_tables.py; do not edit.",
             "// Marching-cubes case tables: crossing-edge mask, triangle count and",
             "// triangles (edge triples, -1 padded) of each of the 256 corner cases.",
             "#pragma once", "", "namespace sn {", "",
             "__constant__ unsigned short kMcEdgeMask[256] = {"]
    for i in range(0, 256, 8):
        lines.append("    " + ", ".join(f"0x{v:03x}" for v in edge_mask[i:i + 8]) + ",")
    lines += ["};", "", "__constant__ unsigned char kMcTriCount[256] = {"]
    for i in range(0, 256, 16):
        lines.append("    " + ", ".join(str(v) for v in tri_count[i:i + 16]) + ",")
    lines += ["};", "", "__constant__ signed char kMcTriTable[256][16] = {"]
    for row in tri_table:
        lines.append("    {" + ", ".join(str(v) for v in row) + "},")
    lines += ["};", "", "}  // namespace sn", ""]
    return "\n".join(lines)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    out = Path(argv[0]) if argv else Path(__file__).with_name("mc_tables.h")
    out.write_text(header_text())
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
