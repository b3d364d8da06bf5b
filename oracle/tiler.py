"""CPU restatement of the tiler's triangle assignment (small cases).

TEST INFRASTRUCTURE ONLY: the oracle the device assignment is checked
against; only tests/ import it.

TriangleOctTreeGrid::assignTriangles / getVertGridTreeXY /
computeMinMaxGridCell (main.cu 3986-4107), in fp32: each vertex's cell is
uint(dist / size); a vertex within `border` of the cell's far edge also
reaches the next cell, otherwise one within `border` of its near edge (not in
cell 0) reaches the previous one; both ends clamp to dim - 1; a vertex below
the origin reaches cell 0.  A triangle goes to every tile in the box spanned
by its vertices' ranges, tile index y * dimX + x.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def cell_range(v, origin, size, border, dim):
    v, origin, size, border = f32(v), f32(origin), f32(size), f32(border)
    if v < origin:
        return 0, 0
    dist = f32(v - origin)
    q = f32(dist / size)
    mn = mx = int(q)
    max_border = f32(f32(mn + 1) * size)
    d = f32(max_border - dist)
    if d <= border:
        mx = mn + 1
    elif mn != 0:
        min_border = f32(max_border - size)
        if f32(dist - min_border) <= border:
            mx = mn
            mn = mn - 1
    return min(mn, dim - 1), min(mx, dim - 1)


def assign(verts, origin, size, border, dim):
    """{tile: [triangle indices in input order]}"""
    v = np.asarray(verts, np.float32).reshape(-1, 3, 3)
    out = {}
    for t in range(v.shape[0]):
        x0 = y0 = 1 << 40
        x1 = y1 = -1
        for k in range(3):
            a, b = cell_range(v[t, k, 0], origin[0], size[0], border[0], dim[0])
            c, d = cell_range(v[t, k, 1], origin[1], size[1], border[1], dim[1])
            x0, x1, y0, y1 = min(x0, a), max(x1, b), min(y0, c), max(y1, d)
        for y in range(y0, y1 + 1):
            for x in range(x0, x1 + 1):
                out.setdefault(y * dim[0] + x, []).append(t)
    return out
