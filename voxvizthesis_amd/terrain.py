"""Procedural terrain on the device and the served paging built on it at
scale (include/voxmi_terrain.h).

A 32769^2 heightfield is 2.1 G triangles (77 GB as a soup), more than a
scene kept resident for the tiler, so `TerrainPager` builds each tile's
triangles on the device from a resident height grid when the pager asks for
the tile: the same on-demand tile voxelization, shared brick pool (a wide
pool beyond 8.2 GB), brick server, culling and load order as `TilePager`
(the reference's paged renderer, GigaVoxelsDatabasePager.cpp 220-465), with
the disk load replaced by tile generation + voxelization.  Rocks are
instanced icospheres resting on the surface, picked per tile.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib, scenes
from .pager import PagedTree, TilePager
from .voxelizer import Voxelizer

f32 = np.float32


def c5_height_fn(seed: int, ridged_seed: int, ridged_amp: float, extent: float):
    """Host restatement of the device heights (float64): 0.5 E + 0.2 E fbm
    + ridged_amp E ridged, at world (x, y)."""
    def h(x, y):
        x = np.asarray(x, np.float64) / extent
        y = np.asarray(y, np.float64) / extent
        out = 0.5 * extent + 0.2 * extent * scenes.fbm(x, y, seed)
        if ridged_amp:
            out = out + ridged_amp * extent * scenes.ridged(x, y, ridged_seed)
        return out
    return h


class DeviceTerrain:
    """An (cells+1)^2 height grid over [0, extent]^2 resident in HBM, and the
    triangles of any cell window of it (vox_terrain_*)."""

    def __init__(self, cells: int, extent: float, seed: int, ridged_seed: int = 0,
                 ridged_amp: float = 0.0, tex_width: int = 256):
        self._L = _lib.load()
        self.cells, self.n1 = int(cells), int(cells) + 1
        self.extent, self.tex_width = float(extent), int(tex_width)
        self.heights = _lib.DeviceArray.alloc(4 * self.n1 * self.n1)
        if not self._L.vox_terrain_heights(self.heights.ptr, self.n1, self.extent, int(seed),
                                           int(ridged_seed), float(ridged_amp)):
            raise RuntimeError(self._err())
        self.lo, self.hi = self.range(0, 0, self.n1, self.n1)
        self.image_host = scenes.height_ramp_texture(self.tex_width)
        self.image = _lib.DeviceArray(self.image_host)

    def _err(self) -> str:
        return self._L.vox_terrain_get_error().decode()

    def range(self, x0: int, y0: int, w: int, h: int):
        """(min, max) height over the vertex window [x0, x0+w) x [y0, y0+h)."""
        mm = (C.c_float * 2)()
        if not self._L.vox_terrain_range(self.heights.ptr, self.n1, x0, y0, w, h, mm):
            raise RuntimeError(self._err())
        return float(mm[0]), float(mm[1])

    def tile(self, x0: int, y0: int, cw: int, ch: int, out=None):
        """Device arrays (verts, vnormals, uvs) of the cells [x0, x0+cw) x
        [y0, y0+ch): 2 * cw * ch triangles; into `out` (three arrays large
        enough) when given, else new arrays."""
        nt = 2 * cw * ch
        need = (4 * 9 * nt, 4 * 9 * nt, 4 * 6 * nt)
        if out is not None:
            if any(b.nbytes < n for b, n in zip(out, need)):
                raise ValueError("tile buffers too small")
            bufs = out
        else:
            bufs = [_lib.DeviceArray.alloc(n) for n in need]
        if not self._L.vox_terrain_tile(self.heights.ptr, self.n1, self.extent, x0, y0, cw, ch,
                                        self.lo, self.hi, self.tex_width, bufs[0].ptr,
                                        bufs[1].ptr, bufs[2].ptr):
            if out is None:
                for b in bufs:
                    b.close()
            raise RuntimeError(self._err())
        return bufs

    def mesh(self, x0: int, y0: int, cw: int, ch: int) -> scenes.Mesh:
        """The same window as a host Mesh (tests)."""
        bufs = self.tile(x0, y0, cw, ch)
        nt = 2 * cw * ch
        out = []
        try:
            for b, k in zip(bufs, (9, 9, 6)):
                a = np.empty(nt * k, np.float32)
                if not self._L.vox_device_download(a.ctypes.data, b.ptr, 0, a.nbytes):
                    raise RuntimeError("download failed")
                out.append(a)
        finally:
            for b in bufs:
                b.close()
        return scenes.Mesh(verts=out[0].reshape(-1, 3, 3), vnormals=out[1].reshape(-1, 3, 3),
                           uvs=out[2].reshape(-1, 3, 2), image=self.image_host, is_terrain=True,
                           name=f"terrain_window_{x0}_{y0}_{cw}x{ch}")

    def grid(self) -> np.ndarray:
        """The whole height grid on the host (tests; (n1, n1) float32)."""
        a = np.empty(self.n1 * self.n1, np.float32)
        if not self._L.vox_device_download(a.ctypes.data, self.heights.ptr, 0, a.nbytes):
            raise RuntimeError("download failed")
        return a.reshape(self.n1, self.n1)

    def sample(self, xy: np.ndarray) -> np.ndarray:
        xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
        z = np.empty(xy.shape[0], np.float32)
        if not self._L.vox_terrain_sample(self.heights.ptr, self.n1, self.extent, xy.ctypes.data,
                                          xy.shape[0], z.ctypes.data):
            raise RuntimeError(self._err())
        return z

    def close(self):
        for a in (getattr(self, "heights", None), getattr(self, "image", None)):
            if a is not None:
                a.close()
        self.heights = self.image = None


class RockField:
    """count icosphere rocks (80 triangles, the scene generator's), radius
    log-uniform [rmin, rmax], centres uniform over the inner 98% of the
    extent, resting 0.3 r into the surface."""

    def __init__(self, terrain: DeviceTerrain, count: int, seed: int, rmin: float, rmax: float):
        rng = np.random.Generator(np.random.PCG64(seed))
        E = terrain.extent
        self.r = np.exp(rng.uniform(np.log(rmin), np.log(rmax), size=count))
        cx = rng.uniform(0.01 * E, 0.99 * E, size=count)
        cy = rng.uniform(0.01 * E, 0.99 * E, size=count)
        cz = terrain.sample(np.stack([cx, cy], axis=-1)).astype(np.float64) + 0.3 * self.r
        self.c = np.stack([cx, cy, cz], axis=-1)
        self.sphere = scenes._icosphere1()

    def select(self, lo, hi) -> np.ndarray:
        """Indices of the rocks whose bounding boxes meet the box [lo, hi]."""
        m = np.ones(self.r.size, bool)
        for a in range(3):
            m &= (self.c[:, a] + self.r >= lo[a]) & (self.c[:, a] - self.r <= hi[a])
        return np.flatnonzero(m)

    def mesh(self, idx) -> scenes.Mesh | None:
        if len(idx) == 0:
            return None
        v = self.sphere[None] * self.r[idx, None, None, None] + self.c[idx, None, None, :]
        return scenes.Mesh(verts=v.reshape(-1, 3, 3).astype(np.float32), name=f"rocks{len(idx)}")


class TerrainPager(TilePager):
    """Served paging over a procedural terrain: cells^2 metre cells (voxel
    size 1 m), tiles of N^3 voxels stacked in z over each column's height
    range, each built on the device and voxelized when first seen.

    pool_slots as for TilePager (one number: a wide pool of that many
    slots).  The frame loop, culling, load order, unloading and serving
    (f2b_stop included) are TilePager's."""

    def __init__(self, cells: int, N: int, B: int, pool_slots, seed: int = 6,
                 ridged_seed: int = 7, ridged_amp: float = 0.05, rocks: int = 100000,
                 rock_seed: int = 8, rock_radius=(2.0, 40.0), max_trees: int = 96,
                 max_uploads: int = 65536, max_loading: int = 4, stale_frames: int = 100,
                 streams: int = 2, max_width: int = 2560, max_height: int = 1440,
                 near: float = 1.0, far: float = 40000.0, options: int = 0,
                 serve_group: int = 3, f2b_stop=None, pool_normals: bool = True):
        from concurrent.futures import ThreadPoolExecutor
        self._L = _lib.load()
        self.meshes = []
        self.N, self.B = int(N), int(B)
        self.vs = np.full(3, 1.0, np.float32)
        self.max_loading, self.stale = int(max_loading), int(stale_frames)
        self.serve_group = int(serve_group)
        self.f2b_stop = f2b_stop
        self.pool_normals = bool(pool_normals)
        self.streams, self.W, self.H = int(streams), int(max_width), int(max_height)
        self.near, self.far, self.options = float(near), float(far), int(options)
        self.frame_index = 0
        self._cast = []
        self._loader = ThreadPoolExecutor(max_workers=self.max_loading)
        self.load_ms = []
        self.load_phases = []     # per load: host ms of its phases
        import threading
        self._tls = threading.local()
        self._tls_lock = threading.Lock()
        self._all_tile_bufs = []
        self.tiles_loaded = 0
        self.terrain = DeviceTerrain(cells, float(cells), seed, ridged_seed, ridged_amp)
        self.height_fn = c5_height_fn(seed, ridged_seed, ridged_amp, float(cells))
        self.rocks = RockField(self.terrain, rocks, rock_seed, *rock_radius)
        self.frames = [_lib.DeviceArray(np.zeros((self.H, self.W, 4), np.float32))
                       for _ in range(self.streams)]
        self._frame_dims = [(0, 0)] * self.streams
        # tiles of N metres; the tile's voxel space reaches extra voxels
        # before its origin and 3 * extra past its end (offsetP = p - extra
        # * delta, S = N + 4 * extra); its triangles are the cells there
        self.extra = (self.N // self.B) >> 1
        self.size = f32(self.N)
        dim = (cells + self.N - 1) // self.N
        self.trees = []
        self._win = {}
        for ty in range(dim):
            for tx in range(dim):
                win = self._window(tx, ty)
                zlo, zhi = self.terrain.range(win[0], win[1], win[2] + 1, win[3] + 1)
                sel = self.rocks.select((win[0], win[1], -1e30),
                                        (win[0] + win[2], win[1] + win[3], 1e30))
                if len(sel):
                    zhi = max(zhi, float((self.rocks.c[sel, 2] + self.rocks.r[sel]).max()))
                    zlo = min(zlo, float((self.rocks.c[sel, 2] - self.rocks.r[sel]).min()))
                t = ty * dim + tx
                self._win[t] = win
                ox, oy = f32(tx * self.N), f32(ty * self.N)
                for tz, (oz, blo, bhi) in enumerate(self._z_stack(zlo, zhi)):
                    mn = np.array([ox, oy, blo], np.float32)
                    mx = np.array([f32(ox + self.size), f32(oy + self.size), bhi], np.float32)
                    c = ((mn + mx) * f32(0.5)).astype(np.float32)
                    dv = (mx - mn).astype(np.float32)
                    r = float(np.sqrt(f32(f32(f32(dv[0] * dv[0]) + f32(dv[1] * dv[1]))
                                          + f32(dv[2] * dv[2])) * f32(0.25)))
                    org = (ox, oy, oz)
                    lo = np.asarray(org, np.float32)
                    hi = (lo + self.size).astype(np.float32)
                    self.trees.append(PagedTree(tx, ty, tz, t, org, lo, hi, c, r))
        self._setup_serving(pool_slots, max_trees, max_uploads)

    def _window(self, tx: int, ty: int):
        """Cell window (x0, y0, cw, ch) of a tile's voxel space."""
        cells = self.terrain.cells
        out = []
        for t in (tx, ty):
            a = max(0, t * self.N - self.extra)
            b = min(cells, t * self.N + self.N + 3 * self.extra)
            out.append((a, b - a))
        return out[0][0], out[1][0], out[0][1], out[1][1]

    def _z_stack(self, zmin, zmax):
        """Tile levels over a column's height range, as the tiler stacks
        them (main.cu 4231-4243, 4330-4336)."""
        sz = self.size
        height = f32(f32(zmax) - f32(zmin))
        nz = max(1, int(np.ceil(f32(height / sz))))
        gh = f32(f32(nz) * sz)
        lo, hi = f32(zmin), f32(zmax)
        if gh > height:
            diff = f32(f32(gh - height) * f32(0.5))
            lo = f32(lo - diff)
            hi = f32(lo + sz)
        oz = f32(np.floor(lo))
        out = []
        for _ in range(nz):
            out.append((oz, lo, hi))
            oz = f32(oz + sz)
            lo, hi = hi, f32(hi + sz)
        return out

    def tile_groups(self, tr: PagedTree, out=None):
        """The tile's terrain window (device arrays, in `out` when given +
        group bounds) and its rocks (host mesh or None)."""
        x0, y0, cw, ch = self._win[tr.tile]
        bufs = self.terrain.tile(x0, y0, cw, ch, out)
        zlo, zhi = self.terrain.range(x0, y0, cw + 1, ch + 1)
        step = self.terrain.extent / self.terrain.cells
        lo = (f32(x0 * step), f32(y0 * step), f32(zlo))
        hi = (f32((x0 + cw) * step), f32((y0 + ch) * step), f32(zhi))
        z0 = float(tr.origin[2]) - self.extra
        z1 = float(tr.origin[2]) + self.N + 3 * self.extra
        sel = self.rocks.select((lo[0], lo[1], z0), (hi[0], hi[1], z1))
        return bufs, 2 * cw * ch, lo, hi, self.rocks.mesh(sel)

    def _tile_buffers(self):
        """This loader thread's tile arrays, sized for the largest window and
        kept for the pager's life (no allocation or release per tile)."""
        bufs = getattr(self._tls, "bufs", None)
        if bufs is None:
            w = self.N + 4 * self.extra
            nt = 2 * w * w
            bufs = [_lib.DeviceArray.alloc(n) for n in (36 * nt, 36 * nt, 24 * nt)]
            self._tls.bufs = bufs
            with self._tls_lock:
                self._all_tile_bufs += bufs
        return bufs

    def _load(self, tr: PagedTree):
        import time
        t = [time.perf_counter()]
        vz = Voxelizer((self.N,) * 3, (self.B,) * 3, None, options=self.options)
        try:
            ok = vz.initCuda() and vz.allocateVoxelMemory() and \
                vz.setVoxelizationParams(tr.origin, self.vs)
            t.append(time.perf_counter())
            bufs, nt, lo, hi, rocks = self.tile_groups(tr, self._tile_buffers())
            t.append(time.perf_counter())
            grp = _lib.VoxGroup()
            grp.verts, grp.vnormals, grp.uvs = bufs[0].ptr, bufs[1].ptr, bufs[2].ptr
            grp.ntri = nt
            grp.bbox_min[:] = [float(v) for v in lo]
            grp.bbox_max[:] = [float(v) for v in hi]
            grp.is_terrain = 1
            grp.rgba8 = self.terrain.image.ptr
            grp.width, grp.height = self.terrain.tex_width, 1
            grp.addr0, grp.addr1 = 1, 1
            ok = ok and bool(self._L.vox_add_group_device(vz._h, C.byref(grp)))
            if rocks is not None:
                ok = ok and vz.addMesh(rocks)
            ok = ok and vz.allocateTriangleMemory() and vz.computeEdgesFaceNormalsAndBounds()
            t.append(time.perf_counter())
            st = vz.generateVoxelsAndOctTree(None) if ok else 0
            t.append(time.perf_counter())
            if st == 1:
                vz.releaseWorkMemory()
            t.append(time.perf_counter())
            self.load_phases.append(dict(zip(
                ("create", "tile", "setup", "generate", "release"),
                [(b - a) * 1e3 for a, b in zip(t, t[1:])])))
            return vz, st
        except BaseException:
            vz.close()
            raise

    def close(self):
        super().close()
        for b in getattr(self, "_all_tile_bufs", []):
            b.close()
        self._all_tile_bufs = []
        if getattr(self, "terrain", None) is not None:
            self.terrain.close()
            self.terrain = None
