"""Pins the tiler oracle (oracle/tiler.py) and the host-side tile logic of
voxvizthesis_amd/tiler.py on hand cases: the border rule of
computeMinMaxGridCell at both edges of a cell, clamping, vertices below the
origin; the z stacking of a tile (including the flat tile, which yields no
tree in the reference); the root.gvp text."""
import numpy as np

from oracle import tiler as ot
from voxvizthesis_amd.scenes import Mesh
from voxvizthesis_amd.tiler import PagedOctTreeGenerator


def test_cell_range_border_rule():
    r = lambda v: ot.cell_range(v, 0.0, 1.0, 0.1, 3)
    assert r(0.5) == (0, 0)
    assert r(0.95) == (0, 1)        # within the border of the far edge
    assert r(1.05) == (0, 1)        # within the border of the near edge
    assert r(0.05) == (0, 0)        # near edge of cell 0: no previous cell
    assert r(1.5) == (1, 1)
    assert r(2.95) == (2, 2)        # next cell clamped to the last
    assert r(7.0) == (2, 2)
    assert r(-0.5) == (0, 0)        # below the origin


def test_assign_spans_vertex_ranges():
    tris = np.array([[[0.2, 0.2, 0], [0.3, 0.2, 0], [0.2, 0.3, 0]],       # tile 0 only
                     [[0.95, 0.5, 0], [0.9, 0.4, 0], [0.8, 0.5, 0]],       # x border: 0, 1
                     [[0.5, 1.5, 0], [2.5, 1.5, 0], [0.5, 2.5, 0]]],       # box x 0..2, y 1..2
                    np.float32)
    a = ot.assign(tris, (0, 0), (1, 1), (0.1, 0.1), (3, 3))
    assert a[0] == [0, 1] and a[1] == [1]
    assert all(2 in a[y * 3 + x] for y in (1, 2) for x in (0, 1, 2))
    assert 2 not in a.get(0, [])


def test_z_stack_and_root_text(tmp_path):
    m = Mesh(verts=np.array([[[0.1, 0.1, 0.3], [0.9, 0.1, 0.7], [0.1, 0.9, 0.5]]], np.float32))
    g = PagedOctTreeGenerator([m], voxel_size=1.0 / 64, N=64, B=8)
    g.grid()
    assert g.dim == (1, 1)
    # height 0.4 in a 1.0 tile: one level, centred, origin floored
    (oz, lo, hi), = g.z_stack(np.float32(0.3), np.float32(0.7))
    assert oz == 0.0 and np.isclose(lo, 0.0, atol=1e-7) and hi == np.float32(lo + 1.0)
    # 2.5 tall: three levels, each starting where the last ended
    st = g.z_stack(np.float32(0.0), np.float32(2.5))
    assert len(st) == 3 and [s[0] for s in st] == [-1.0, 0.0, 1.0]
    assert st[1][1] == st[0][2] and st[2][1] == st[1][2]
    # a flat tile gets no level (ceil(0 / size) = 0)
    assert g.z_stack(np.float32(0.5), np.float32(0.5)) == []
    d = tmp_path / "grid_0_0_0"
    d.mkdir()
    (d / "tree.gvx").write_text("<GigaVoxelsOctTree />")
    g.trees = [(0, 0, 0, np.float32(0.0), np.float32(1.0), 1)]
    txt = g.root_text(str(tmp_path))
    lines = txt.splitlines()
    assert lines[0] == "<OctTrees Compressed='NO' BrickXSize='8' BrickYSize='8' BrickZSize='8' >"
    assert lines[-1] == "</OctTrees>"
    assert 'TreeFile="grid_0_0_0/tree.gvx"' in lines[1] and 'SizeXMeters="1"' in lines[1]
    assert 'CenterZ="0.5"' in lines[1] and 'Radius="0.8' in lines[1]
