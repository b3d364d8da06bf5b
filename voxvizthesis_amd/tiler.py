"""Tiled octree generation: PagedGigaVoxelOctTreeGenerator (main.cu 3811-4507).

The input's x/y extent is split into tiles of N voxels (one octree each); the
device assigns every triangle to the tiles its vertices reach, with the voxel
border duplicated into neighbouring tiles (include/voxmi_tiler.h); each tile
is stacked in z by the z extent of its triangles and voxelized with its own
origin, writing grid_x_y_z/tree.gvx + voxels_<d>.gvb; root.gvp lists the
non-empty trees (writeRootFile, main.cu 4416-4502).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .voxelizer import Voxelizer

f32 = np.float32


class TileAssignment:
    """Device assignment of a triangle soup to the tile grid."""

    def __init__(self, verts: np.ndarray, grid: _lib.VoxTileGrid):
        self._L = _lib.load()
        self._h = self._L.vox_tiler_create()
        if not self._h:
            raise MemoryError("vox_tiler_create")
        try:
            v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 9)
            if not self._L.vox_tiler_assign(self._h, v.ctypes.data, v.shape[0], 0, C.byref(grid)):
                raise RuntimeError(self._L.vox_tiler_get_error(self._h).decode())
            lst = _lib.VoxTileLists()
            self._L.vox_tiler_get(self._h, C.byref(lst))
            n = lst.ntiles
            self.start = self._copy(lst.tile_start, n + 1, np.uint32)
            self.tris = self._copy(lst.tile_tris, lst.pairs, np.uint32)
            self.zmin = self._copy(lst.tile_zmin, n, np.float32)
            self.zmax = self._copy(lst.tile_zmax, n, np.float32)
        finally:
            self._L.vox_tiler_destroy(self._h)
            self._h = None

    def _copy(self, ptr, n, dtype):
        a = np.empty(int(n), dtype)
        if n and not self._L.vox_pool_copy_to_host(a.ctypes.data, ptr, a.nbytes):
            raise RuntimeError("copy failed")
        return a

    def tile(self, t: int) -> np.ndarray:
        """Triangles of tile t in input order (as the reference appends them;
        the device lists are ordered)."""
        return self.tris[self.start[t]:self.start[t + 1]]


def fmt(v) -> str:
    """Default iostream formatting of a float (6 significant digits)."""
    return "%g" % float(f32(v))


class PagedOctTreeGenerator:
    def __init__(self, meshes, voxel_size, N: int, B: int):
        self.meshes = list(meshes)
        self.vs = np.asarray(voxel_size, np.float32).reshape(-1)
        if self.vs.size == 1:
            self.vs = np.repeat(self.vs, 3)
        self.N, self.B = int(N), int(B)
        # maxVoxelMipMapSizeMeters = voxel size * dimensions; border = extra voxels
        self.size = (self.vs * f32(self.N)).astype(np.float32)
        extra = (self.N // self.B) >> 1
        self.border = (self.vs * f32(extra)).astype(np.float32)
        self.verts = np.concatenate([np.asarray(m.verts, np.float32).reshape(-1, 9)
                                     for m in self.meshes]) if self.meshes else \
            np.zeros((0, 9), np.float32)
        self.group_off = np.cumsum([0] + [m.ntri for m in self.meshes])

    def grid(self) -> _lib.VoxTileGrid:
        """TriangleOctTreeGrid (main.cu 3940-3968) over the input's bounding box."""
        flat = self.verts.reshape(-1, 3)
        self.lo = flat.min(axis=0).astype(np.float32)
        self.hi = flat.max(axis=0).astype(np.float32)
        g = _lib.VoxTileGrid()
        for a in range(2):
            g.origin[a] = float(self.lo[a])
            g.size[a] = float(self.size[a])
            g.border[a] = float(self.border[a])
            g.dim[a] = int(np.ceil(f32(self.hi[a] - self.lo[a]) / self.size[a]))
        self.dim = (g.dim[0], g.dim[1])
        # tile origins accumulate the tile size (origin.x += size.x)
        self.ox = [f32(self.lo[0])]
        for _ in range(self.dim[0]):
            self.ox.append(f32(self.ox[-1] + self.size[0]))
        self.oy = [f32(self.lo[1])]
        for _ in range(self.dim[1]):
            self.oy.append(f32(self.oy[-1] + self.size[1]))
        return g

    def z_stack(self, zmin, zmax):
        """Z levels of one tile (main.cu 4231-4243, 4330-4336): [(origin z,
        bbox z min, bbox z max)]."""
        sz = self.size[2]
        height = f32(f32(zmax) - f32(zmin))
        nz = int(np.ceil(f32(height / sz)))
        gh = f32(f32(nz) * sz)
        lo, hi = f32(zmin), f32(zmax)
        if gh > height:
            diff = f32(f32(gh - height) * f32(0.5))
            lo = f32(lo - diff)
            hi = f32(lo + sz)
        oz = f32(np.floor(lo))
        out = []
        for k in range(nz):
            out.append((oz, lo, hi))
            oz = f32(oz + sz)
            lo, hi = hi, f32(hi + sz)
        return out

    def generate(self, out_dir: str | None = None, options: int = 0):
        """Voxelize every tile; returns [(x, y, z, status, triangles)]."""
        g = self.grid()
        self.assign = TileAssignment(self.verts, g)
        results = []
        self.trees = []
        vz = Voxelizer((self.N,) * 3, (self.B,) * 3, None, options=options)
        try:
            if not (vz.initCuda() and vz.allocateVoxelMemory()):
                raise RuntimeError(vz.getErrorMessage())
            for y in range(self.dim[1]):
                for x in range(self.dim[0]):
                    t = y * self.dim[0] + x
                    idx = self.assign.tile(t)
                    if idx.size == 0:
                        continue
                    for z, (oz, blo, bhi) in enumerate(self.z_stack(self.assign.zmin[t],
                                                                    self.assign.zmax[t])):
                        d = None
                        if out_dir is not None:
                            d = os.path.join(out_dir, f"grid_{x}_{y}_{z}")
                            os.makedirs(d, exist_ok=True)
                        st = self._voxelize(vz, idx, (self.ox[x], self.oy[y], oz), d)
                        if d is not None:
                            marker = os.path.join(d, "GigaVoxelOctreeIsEmpty.txt")
                            if st == -1:
                                with open(marker, "w") as f:
                                    f.write("This octree voxelized no triangles.\n")
                            elif os.path.exists(marker):
                                os.remove(marker)
                        results.append((x, y, z, st, int(idx.size)))
                        self.trees.append((x, y, z, blo, bhi, st))
        finally:
            vz.close()
        if out_dir is not None:
            self.write_root(out_dir)
        return results

    def _voxelize(self, vz, idx, origin, out_dir):
        """Voxelize (main.cu 3225-3341): one group per input mesh with
        triangles in the tile, group bounds from those triangles."""
        vz.deallocateTriangleMemory()
        vz.setVoxelizationParams(origin, self.vs)
        for gi, m in enumerate(self.meshes):
            lo, hi = self.group_off[gi], self.group_off[gi + 1]
            sel = idx[(idx >= lo) & (idx < hi)] - lo
            if sel.size == 0:
                continue
            v = np.asarray(m.verts, np.float32).reshape(-1, 3, 3)[sel]
            vn = None if m.vnormals is None else np.asarray(m.vnormals, np.float32).reshape(
                -1, 3, 3)[sel]
            uv = None if m.uvs is None else np.asarray(m.uvs, np.float32).reshape(-1, 3, 2)[sel]
            flat = v.reshape(-1, 3)
            vz.addTriangleGroup(v, vn, uv, flat.min(axis=0), flat.max(axis=0), m.is_terrain,
                                m.image, m.addr[0], m.addr[1])
        if not (vz.allocateTriangleMemory() and vz.computeEdgesFaceNormalsAndBounds()):
            raise RuntimeError(vz.getErrorMessage())
        st = vz.generateVoxelsAndOctTree(out_dir, True, False)
        if st == 0:
            raise RuntimeError(vz.getErrorMessage())
        return st

    def root_text(self, out_dir: str, compressed: bool = False) -> str:
        """writeRootFile (main.cu 4416-4502): x outer, y inner, z chain;
        empty or missing trees skipped."""
        lines = [f"<OctTrees Compressed='{'YES' if compressed else 'NO'}' "
                 f"BrickXSize='{self.B}' BrickYSize='{self.B}' BrickZSize='{self.B}' >"]
        by_xy = {}
        for (x, y, z, blo, bhi, st) in self.trees:
            by_xy.setdefault((x, y), []).append((z, blo, bhi))
        for x in range(self.dim[0]):
            for y in range(self.dim[1]):
                for (z, blo, bhi) in sorted(by_xy.get((x, y), [])):
                    rel = f"grid_{x}_{y}_{z}"
                    if os.path.exists(os.path.join(out_dir, rel, "GigaVoxelOctreeIsEmpty.txt")):
                        continue
                    if not os.path.exists(os.path.join(out_dir, rel, "tree.gvx")):
                        continue
                    mn = np.array([self.ox[x], self.oy[y], blo], np.float32)
                    mx = np.array([f32(self.ox[x] + self.size[0]), f32(self.oy[y] + self.size[1]),
                                   bhi], np.float32)
                    c = ((mn + mx) * f32(0.5)).astype(np.float32)
                    dv = (mx - mn).astype(np.float32)
                    r = f32(np.sqrt(f32(f32(f32(dv[0] * dv[0]) + f32(dv[1] * dv[1]))
                                        + f32(dv[2] * dv[2])) * f32(0.25)))
                    lines.append(
                        f'    <OctTree CenterX="{fmt(c[0])}" CenterY="{fmt(c[1])}" '
                        f'CenterZ="{fmt(c[2])}" Radius="{fmt(r)}" '
                        f'SizeXMeters="{fmt(self.size[0])}" SizeYMeters="{fmt(self.size[1])}" '
                        f'SizeZMeters="{fmt(self.size[2])}" TreeFile="{rel}/tree.gvx" />')
        lines.append("</OctTrees>")
        return "\n".join(lines) + "\n"

    def write_root(self, out_dir: str, compressed: bool = False):
        with open(os.path.join(out_dir, "root.gvp"), "w") as f:
            f.write(self.root_text(out_dir, compressed))
