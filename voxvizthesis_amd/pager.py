"""Paged serving: tiles of a large scene voxelized on demand on the device,
served from one shared brick pool.

The reference's renderer culls the paged octrees of root.gvp against the
camera frustum each frame (CullVisitor, GigaVoxelsRenderer.cpp 750-797),
queues the visible trees that are not loaded, and a pager thread loads them
from disk, most recently accessed first, then nearest (DatabasePagerThread::run
and SortByDistAndFrameIndex, GigaVoxelsDatabasePager.cpp 220-338); a loaded
tree's root brick goes on the brick upload list, loaded trees are drawn far
to near, and trees not seen for 100 frames are unloaded
(unloadStalePagedOctTreeNodes, 420-465).

Here the "load" of a tile is its voxelization on the device: the scene's
triangles stay resident in HBM, the tiler's ordered per-tile lists pick each
tile's triangles (vox_add_group_gather), and a loader thread runs the tile's
allocation, gather and generate on the handle's own stream while frames go
on (the library releases the interpreter lock in every call).  A finished
tile becomes an octree on the shared pool (its root brick admitted through
the LRU like any other) attached to one brick server; every loaded visible
tree gets a ray cast per camera stream, and a server frame after every few
of them (serve_group) admits their requests while the next trees are cast,
so a request waits for its group's casts rather than the whole frame's.
Tiles therefore enter as tree slots and memory free up, and their
voxelization overlaps with the frames' ray casts on the device.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .gvpool import BrickPool, OctTree
from .gvserve import (BrickServer, RayCaster, box_rects, camera, composite, raycast_front_to_back,
                      raycast_params)
from .tiler import PagedOctTreeGenerator
from .voxelizer import Voxelizer

f32 = np.float32


# --- frustum (vox::Frustum, VoxVizCore/Frustum.cpp 14-147) ------------------

def perspective(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """ComputePerspective / ComputeFrustum (Camera.cpp 60-89): the OpenGL
    frustum matrix, column-major as QMatrix4x4::data() holds it."""
    ymax = near * math.tan(fovy_deg * math.pi / 360.0)
    xmax = ymax * aspect
    l, r, b, t = -xmax, xmax, -ymax, ymax
    m = np.zeros(16)
    m[0] = 2 * near / (r - l)
    m[5] = 2 * near / (t - b)
    m[8] = (r + l) / (r - l)
    m[9] = (t + b) / (t - b)
    m[10] = -(far + near) / (far - near)
    m[11] = -1.0
    m[14] = -2 * far * near / (far - near)
    return m


def look_at_matrix(pos, look, up) -> np.ndarray:
    """View matrix of a camera at pos looking along look (column-major)."""
    f = np.asarray(look, np.float64)
    f = f / np.linalg.norm(f)
    u = np.asarray(up, np.float64)
    s = np.cross(f, u)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    p = np.asarray(pos, np.float64)
    m = np.zeros(16)
    m[0], m[4], m[8] = s
    m[1], m[5], m[9] = u
    m[2], m[6], m[10] = -f
    m[12] = -s.dot(p)
    m[13] = -u.dot(p)
    m[14] = f.dot(p)
    m[15] = 1.0
    return m


def frustum_planes(proj: np.ndarray, modl: np.ndarray) -> np.ndarray:
    """Frustum::setFromCamera: clip = modelview x projection, then right,
    left, bottom, top, far, near planes, each normalised by its normal."""
    clip = np.zeros(16)
    for r in range(4):
        for c in range(4):
            clip[4 * r + c] = sum(modl[4 * r + k] * proj[4 * k + c] for k in range(4))
    rows = [(3, 0, -1), (3, 0, 1), (3, 1, 1), (3, 1, -1), (3, 2, -1), (3, 2, 1)]
    planes = np.zeros((6, 4))
    for p, (a, b, sgn) in enumerate(rows):
        for k in range(4):
            planes[p, k] = clip[4 * k + a] + sgn * clip[4 * k + b]
        planes[p] /= math.sqrt(planes[p, 0] ** 2 + planes[p, 1] ** 2 + planes[p, 2] ** 2)
    return planes


def sphere_in_frustum(planes: np.ndarray, center, radius: float):
    """Frustum::sphereInFrustum: (visible, distance to the near plane)."""
    d = 0.0
    for p in range(6):
        d = float(f32(planes[p, 0]) * f32(center[0]) + f32(planes[p, 1]) * f32(center[1])
                  + f32(planes[p, 2]) * f32(center[2]) + f32(planes[p, 3]))
        if d <= -radius:
            return False, d
    return True, d


def spheres_in_frustum(planes: np.ndarray, centers, radii):
    """sphere_in_frustum for n spheres at once (centers (n, 3), radii (n,)),
    in the same float32 arithmetic: per sphere whether it is visible and its
    distance to the near plane (visible) or to the first plane it lies
    behind."""
    c = np.asarray(centers, np.float32).reshape(-1, 3)
    r = np.asarray(radii, np.float64).reshape(-1)
    pl = np.asarray(planes).astype(np.float32)
    d = ((pl[None, :, 0] * c[:, None, 0] + pl[None, :, 1] * c[:, None, 1])
         + pl[None, :, 2] * c[:, None, 2]) + pl[None, :, 3]                 # (n, 6) float32
    behind = d.astype(np.float64) <= -r[:, None]
    inside = ~behind.any(axis=1)
    first = np.where(inside, 5, behind.argmax(axis=1))
    dist = d[np.arange(c.shape[0]), first].astype(np.float64)
    return inside, dist


def front_to_back(trees, pos):
    """The trees in an order every ray from pos meets them in.  Tiles sit on
    a regular grid of columns in x and y, each column a contiguous stack of
    boxes in z: along a ray the column indices move away from the camera's
    column one step at a time (so the L1 column distance grows), and inside a
    column the boxes are met in order of their distance from the camera's
    height.  Sorting by (column distance, height distance) therefore orders
    every ray's boxes front to back."""
    def key(tr):
        sx = float(tr.box_max[0] - tr.box_min[0])
        sy = float(tr.box_max[1] - tr.box_min[1])
        dx = math.floor((float(pos[0]) - float(tr.box_min[0])) / sx)
        dy = math.floor((float(pos[1]) - float(tr.box_min[1])) / sy)
        z = float(pos[2])
        dz = max(float(tr.box_min[2]) - z, z - float(tr.box_max[2]), 0.0)
        return (abs(dx) + abs(dy), dz)
    return sorted(trees, key=key)


def load_order(requests):
    """SortByDistAndFrameIndex (GigaVoxelsDatabasePager.cpp 220-230): most
    recently accessed first, then nearest.  requests: [(tree, last access,
    distance)]."""
    return sorted(requests, key=lambda r: (-r[1], r[2]))


# --- the pager ---------------------------------------------------------------

@dataclass
class PagedTree:
    x: int
    y: int
    z: int
    tile: int                     # tiler tile index y * dimX + x
    origin: tuple                 # world position of voxel (0, 0, 0)
    box_min: np.ndarray
    box_max: np.ndarray
    center: np.ndarray
    radius: float
    state: str = "unloaded"       # unloaded | loading | loaded | empty
    slot: int = -1                # the server's tree slot while loaded
    last_access: int = -1
    requested_frame: int = -1
    loaded_frame: int = -1
    vz: Voxelizer | None = None
    octree: OctTree | None = None
    rcs: list = field(default_factory=list)


class TilePager:
    """On-demand tile voxelization + one shared-pool brick server.

    meshes: scene meshes (numpy, as voxvizthesis_amd.scenes builds them);
    voxel_size, N, B: the tiles' voxel size, grid edge and brick edge (the
    PagedGigaVoxelOctTreeGenerator parameters, main.cu 4135-4200)."""

    def __init__(self, meshes, voxel_size, N: int, B: int, pool_slots, max_trees: int = 32,
                 max_uploads: int = 65536, max_loading: int = 4, stale_frames: int = 100,
                 streams: int = 1, max_width: int = 1920, max_height: int = 1080,
                 near: float = 1e-3, far: float = 100.0, options: int = 0,
                 serve_group: int = 3, f2b_stop=None, pool_normals: bool = True):
        self._L = _lib.load()
        self.meshes = list(meshes)
        # f2b_stop: None casts and blends every tree separately, far to near
        # (the reference's renderer); a number casts each stream's trees in
        # one front-to-back pass, a pixel skipping the trees behind it once
        # its transmittance is at most f2b_stop (0: none skipped)
        self.f2b_stop = f2b_stop
        # pool_normals False: a colour-only wide pool (the ray cast samples
        # colour only), twice the slots in the same memory
        self.pool_normals = bool(pool_normals)
        self.N, self.B = int(N), int(B)
        self.gen = PagedOctTreeGenerator(self.meshes, voxel_size, self.N, self.B)
        self.max_loading, self.stale = int(max_loading), int(stale_frames)
        # trees per server frame within a frame (0: one server frame after all
        # of the frame's ray casts)
        self.serve_group = int(serve_group)
        self.streams, self.W, self.H = int(streams), int(max_width), int(max_height)
        self.near, self.far, self.options = float(near), float(far), int(options)
        self.frame_index = 0
        self._cast = []
        from concurrent.futures import ThreadPoolExecutor
        self._loader = ThreadPoolExecutor(max_workers=self.max_loading)
        self.load_ms = []
        self.tiles_loaded = 0
        # scene resident in HBM, one set of arrays per mesh
        self._dev = []
        for m in self.meshes:
            tex = m.uvs is not None and m.image is not None
            self._dev.append(dict(
                verts=_lib.DeviceArray(np.asarray(m.verts, np.float32)),
                vn=None if m.vnormals is None else _lib.DeviceArray(
                    np.asarray(m.vnormals, np.float32)),
                uv=_lib.DeviceArray(np.asarray(m.uvs, np.float32)) if tex else None,
                img=_lib.DeviceArray(np.asarray(m.image, np.uint8)) if tex else None))
        # per stream, the RGBA32F frame the trees are blended into
        self.frames = [_lib.DeviceArray(np.zeros((self.H, self.W, 4), np.float32))
                       for _ in range(self.streams)]
        self._frame_dims = [(0, 0)] * self.streams
        # per-tile triangle lists on the device, and each mesh's range in them
        g = self.gen.grid()
        self._tiler = self._L.vox_tiler_create()
        if not self._tiler or not self._L.vox_tiler_assign(self._tiler,
                                                           self.gen.verts.ctypes.data,
                                                           self.gen.verts.shape[0], 0,
                                                           C.byref(g)):
            raise RuntimeError(self._L.vox_tiler_get_error(self._tiler).decode())
        lst = _lib.VoxTileLists()
        self._L.vox_tiler_get(self._tiler, C.byref(lst))
        self._tile_tris = lst.tile_tris
        nt = int(lst.ntiles)
        first = np.asarray(self.gen.group_off, np.uint32)
        self._bounds = np.zeros(nt * first.size, np.uint32)
        if not self._L.vox_tiler_group_bounds(self._tiler, first.ctypes.data, first.size,
                                              self._bounds.ctypes.data):
            raise RuntimeError(self._L.vox_tiler_get_error(self._tiler).decode())
        self._bounds = self._bounds.reshape(nt, first.size)
        zmin = np.empty(nt, np.float32)
        zmax = np.empty(nt, np.float32)
        self._L.vox_pool_copy_to_host(zmin.ctypes.data, lst.tile_zmin, zmin.nbytes)
        self._L.vox_pool_copy_to_host(zmax.ctypes.data, lst.tile_zmax, zmax.nbytes)
        # the paged trees: every non-empty tile, stacked in z (main.cu
        # 4225-4340), with root.gvp's bounding sphere (4416-4502)
        self.trees = []
        size = self.gen.size
        for y in range(self.gen.dim[1]):
            for x in range(self.gen.dim[0]):
                t = y * self.gen.dim[0] + x
                if self._bounds[t, -1] == self._bounds[t, 0]:
                    continue
                for z, (oz, blo, bhi) in enumerate(self.gen.z_stack(zmin[t], zmax[t])):
                    mn = np.array([self.gen.ox[x], self.gen.oy[y], blo], np.float32)
                    mx = np.array([f32(self.gen.ox[x] + size[0]), f32(self.gen.oy[y] + size[1]),
                                   bhi], np.float32)
                    c = ((mn + mx) * f32(0.5)).astype(np.float32)
                    dv = (mx - mn).astype(np.float32)
                    r = float(np.sqrt(f32(f32(f32(dv[0] * dv[0]) + f32(dv[1] * dv[1]))
                                          + f32(dv[2] * dv[2])) * f32(0.25)))
                    org = (self.gen.ox[x], self.gen.oy[y], oz)
                    lo = np.asarray(org, np.float32)
                    hi = (lo + size).astype(np.float32)
                    self.trees.append(PagedTree(x, y, z, t, org, lo, hi, c, r))
        self._setup_serving(pool_slots, max_trees, max_uploads)

    def _setup_serving(self, pool_slots, max_trees: int, max_uploads: int):
        """Shared pool and server.  A tree's node texture holds at most
        8 * (1 + internal nodes) texels, plus the texture's rounding.
        pool_slots: slots per axis of the reference's pool texture, or one
        number for a wide pool of that many slots (beyond 8.2 GB of 18^3
        bricks, which 10-bit texel pointers cap)."""
        L = int(round(math.log2(self.N // self.B))) + 1
        blocks = 1 + sum(8 ** d for d in range(L - 1))
        side = int(round((8 * blocks) ** (1.0 / 3.0))) + 2
        self.cap = 8 * blocks + 2 * side * side + 64
        if isinstance(pool_slots, (int, np.integer)):
            self.pool = BrickPool((self.B,) * 3, wide_slots=int(pool_slots),
                                  normals=getattr(self, "pool_normals", True))
        else:
            self.pool = BrickPool((self.B,) * 3, pool_slots)
        self.server = BrickServer.shared(self.pool, max_trees, self.cap, max_uploads)
        self.max_trees = int(max_trees)

    # --- loading -----------------------------------------------------------
    def _load(self, tr: PagedTree):
        """A loader thread's work for one tile: voxel memory, the tile's
        groups gathered from the resident scene, setup and generate on the
        handle's own stream.  Returns (voxelizer, status)."""
        vz = Voxelizer((self.N,) * 3, (self.B,) * 3, None, options=self.options)
        ok = vz.initCuda() and vz.allocateVoxelMemory() and \
            vz.setVoxelizationParams(tr.origin, self.gen.vs)
        row = self._bounds[tr.tile]
        for gi, m in enumerate(self.meshes):
            a, b = int(row[gi]), int(row[gi + 1])
            if b == a:
                continue
            d = self._dev[gi]
            grp = _lib.VoxGroup()
            grp.verts = d["verts"].ptr
            grp.vnormals = None if d["vn"] is None else d["vn"].ptr
            grp.uvs = None if d["uv"] is None else d["uv"].ptr
            grp.ntri = m.ntri
            grp.is_terrain = 1 if m.is_terrain else 0
            grp.rgba8 = None if d["img"] is None else d["img"].ptr
            grp.width = 0 if d["img"] is None else int(m.image.shape[1])
            grp.height = 0 if d["img"] is None else int(m.image.shape[0])
            grp.addr0, grp.addr1 = int(m.addr[0]), int(m.addr[1])
            idx = C.c_void_p(self._tile_tris + 4 * a)
            ok = ok and bool(self._L.vox_add_group_gather(vz._h, C.byref(grp), idx, b - a,
                                                          int(self.gen.group_off[gi])))
        ok = ok and vz.allocateTriangleMemory() and vz.computeEdgesFaceNormalsAndBounds()
        st = vz.generateVoxelsAndOctTree(None) if ok else 0
        if st == 1:
            vz.releaseWorkMemory()
        return vz, st

    def _start_load(self, tr: PagedTree):
        import time
        tr._t0 = time.perf_counter()
        tr._job = self._loader.submit(self._load, tr)
        tr.state = "loading"

    def _finish_load(self, tr: PagedTree):
        import time
        vz, st = tr._job.result()
        tr._job = None
        tr.vz = vz
        if st == 0:
            err = vz.getErrorMessage()
            vz.close()
            tr.vz = None
            tr.state = "unloaded"
            raise RuntimeError(f"tile {tr.x},{tr.y},{tr.z}: {err}")
        self.load_ms.append((time.perf_counter() - tr._t0) * 1e3)
        if st == -1:                       # no triangle overlapped the grid
            vz.close()
            tr.vz = None
            tr.state = "empty"
            return
        tr.octree = OctTree(self.pool, tr.vz, root_brick=False)
        if sum(t.state == "loaded" for t in self.trees) >= self.max_trees:
            self._unload(min((t for t in self.trees if t.state == "loaded"),
                             key=lambda t: t.last_access))
        tr.slot = self.server.attach(tr.octree)
        tr.rcs = [RayCaster(tr.octree, self.W, self.H) for _ in range(self.streams)]
        tr.state = "loaded"
        tr.loaded_frame = self.frame_index
        self.tiles_loaded += 1

    def server_slot(self, tr: PagedTree) -> int:
        return tr.slot

    def _unload(self, tr: PagedTree):
        self.server.detach(tr.octree)
        tr.slot = -1
        for rc in tr.rcs:
            rc.close()
        tr.rcs = []
        tr.octree.close()
        tr.octree = None
        tr.vz.close()
        tr.vz = None
        tr.state = "unloaded"

    # --- one frame -----------------------------------------------------------
    def frame(self, cams, prm_extra=None, serve: bool = True):
        """cams: per camera stream a dict of vox_camera fields (position,
        look, up, fov_y_deg, width, height).  Returns per stream the list of
        (tree, ray caster) drawn far to near this frame.  serve=False leaves
        the server frame to serve() (to read the state the casts saw)."""
        f = self.frame_index
        # finished voxelizations become trees
        for tr in self.trees:
            if tr.state == "loading" and tr._job.done():
                self._finish_load(tr)
        # cull per stream (CullVisitor), every tree's sphere at once
        if getattr(self, "_spheres", None) is None or len(self._spheres[1]) != len(self.trees):
            self._spheres = (np.array([tr.center for tr in self.trees], np.float32).reshape(-1, 3),
                             np.array([tr.radius for tr in self.trees], np.float64))
        drawn, requests = [], {}
        for cm in cams:
            proj = perspective(cm["fov_y_deg"], cm["width"] / cm["height"], self.near, self.far)
            planes = frustum_planes(proj, look_at_matrix(cm["position"], cm["look"], cm["up"]))
            inside, dists = spheres_in_frustum(planes, *self._spheres)
            vis = []
            for i in np.flatnonzero(inside).tolist():
                tr = self.trees[i]
                dist = float(dists[i])
                tr.last_access = f
                if tr.state == "loaded":
                    vis.append((dist - tr.radius, i))
                elif tr.state == "unloaded":
                    requests[i] = min(requests.get(i, dist), dist)
            vis.sort(reverse=True)                       # far to near
            drawn.append([self.trees[i] for _, i in vis])
        # trees unseen for stale_frames frames are unloaded once this frame's
        # casts are queued (none of them is drawn; freeing their device
        # memory waits for the device, which then has the frame to run)
        stale = [tr for tr in self.trees
                 if tr.state == "loaded" and f - tr.last_access > self.stale]
        # start loads, most recent then nearest, as load slots free up
        loading = sum(t.state == "loading" for t in self.trees)
        order = load_order([(i, self.trees[i].last_access, d) for i, d in requests.items()])
        for i, _, _ in order:
            if loading >= self.max_loading:
                break
            tr = self.trees[i]
            tr.requested_frame = f
            self._start_load(tr)
            loading += 1
        if self.f2b_stop is not None:
            out = self._frame_f2b(cams, drawn, prm_extra, serve)
            for tr in stale:
                self._unload(tr)
            return out
        # ray casts of every stream's visible trees.  With serve, a server
        # frame follows every serve_group casts: their requests are admitted
        # while the next trees are cast (which wait for that server frame's
        # node-texture and pool writes), so a request waits for its group,
        # not for the whole frame.
        grouped = serve and self.serve_group > 0
        rcs, out, group = [], [], []
        for k, (cm, trees) in enumerate(zip(cams, drawn)):
            row, cast = [], []
            cam = camera(cm["position"], cm["look"], cm["up"], cm["fov_y_deg"], cm["width"],
                         cm["height"], f)
            rects = box_rects(cm, [tr.box_min for tr in trees], [tr.box_max for tr in trees]) \
                if trees else []
            for tr, rect in zip(trees, rects):
                if rect == ():
                    continue                              # nothing on screen
                rc = tr.rcs[k]
                prm = raycast_params(tr.box_min, tr.box_max, rect=rect, **(prm_extra or {}))
                rc.frame(cam, prm, wait=False)
                rcs.append(rc)
                cast.append(rc)
                row.append((tr, rc))
                if grouped:
                    group.append(rc)
                    if len(group) >= self.serve_group:
                        self.server.frame(group, wait=False)
                        group = []
            # far to near into the stream's frame (the renderer's blending)
            composite(cast, self.frames[k].ptr, cm["width"], cm["height"], clear=True)
            self._frame_dims[k] = (int(cm["width"]), int(cm["height"]))
            out.append(row)
        if grouped:
            if group:
                self.server.frame(group, wait=False)
            self._cast = []
        else:
            self._cast = rcs
            if serve:
                self.serve()
        for tr in stale:
            self._unload(tr)
        self.frame_index += 1
        return out

    def _frame_f2b(self, cams, drawn, prm_extra, serve: bool):
        """Per stream the visible trees in front-to-back order, cast in
        passes straight into the stream's frame, each pass resuming where the
        nearer trees left every pixel; with serve, a server frame follows
        every pass, so the nearest trees' requests are served first and a
        request waits for its pass, not the frame.  The passes hold
        serve_group trees, then serve_group again, then twice as many as the
        pass before (the far trees are small on screen and mostly hidden);
        one pass with serve_group 0."""
        f = self.frame_index
        if getattr(self, "_trans", None) is None:
            self._trans = [_lib.DeviceArray.alloc(4 * self.W * self.H)
                           for _ in range(self.streams)]
        grouped = serve and self.serve_group > 0
        rcs, out = [], []
        for k, (cm, trees) in enumerate(zip(cams, drawn)):
            row, cast, prms = [], [], []
            order = front_to_back(trees, cm["position"])
            rects = box_rects(cm, [tr.box_min for tr in order], [tr.box_max for tr in order]) \
                if order else []
            for tr, rect in zip(order, rects):
                if rect == ():
                    continue                              # nothing on screen
                prms.append(raycast_params(tr.box_min, tr.box_max, rect=rect,
                                           **(prm_extra or {})))
                cast.append(tr.rcs[k])
                row.append((tr, tr.rcs[k]))
            cam = camera(cm["position"], cm["look"], cm["up"], cm["fov_y_deg"], cm["width"],
                         cm["height"], f)
            if not cast:
                composite([], self.frames[k].ptr, cm["width"], cm["height"], clear=True)
            a, step = 0, (self.serve_group if grouped else len(cast))
            while a < len(cast):
                part = cast[a:a + step]
                raycast_front_to_back(part, cam, prms[a:a + step], self.frames[k].ptr,
                                      self.f2b_stop, self._trans[k].ptr, resume=a > 0)
                if grouped:
                    self.server.frame(part, wait=False)
                else:
                    rcs += part
                if a > 0:
                    step *= 2
                a += len(part)
            self._frame_dims[k] = (int(cm["width"]), int(cm["height"]))
            out.append(row)
        self._cast = rcs
        if serve and not grouped:
            self.serve()
        self.frame_index += 1
        return out

    def serve(self):
        """The server frame for the ray casts of the last frame."""
        if self._cast:
            self.server.frame(self._cast, wait=False)
        self._cast = []

    def image(self, k: int = 0) -> np.ndarray:
        """Stream k's composited frame, (H, W, 4) float32."""
        W, H = self._frame_dims[k]
        a = np.empty((H, W, 4), np.float32)
        if a.size and not self._L.vox_pool_copy_to_host(a.ctypes.data, self.frames[k].ptr,
                                                        a.nbytes):
            raise RuntimeError("copy failed")
        return a

    def drain(self):
        """Wait for every running voxelization and finish its load."""
        for tr in self.trees:
            if tr.state == "loading":
                tr._job.result()
                self._finish_load(tr)

    def close(self):
        for tr in getattr(self, "trees", []):
            if tr.state == "loading":
                vz, _ = tr._job.result()
                vz.close()
                tr._job = None
                tr.state = "unloaded"
            elif tr.state == "loaded":
                self._unload(tr)
        if getattr(self, "_loader", None) is not None:
            self._loader.shutdown(wait=True)
            self._loader = None
        if getattr(self, "server", None) is not None:
            self.server.close()
            self.server = None
        if getattr(self, "pool", None) is not None:
            self.pool.close()
            self.pool = None
        if getattr(self, "_tiler", None):
            self._L.vox_tiler_destroy(self._tiler)
            self._tiler = None
        for fr in getattr(self, "frames", []):
            fr.close()
        self.frames = []
        for t in getattr(self, "_trans", None) or []:
            t.close()
        self._trans = None
        for d in getattr(self, "_dev", []):
            for a in d.values():
                if a is not None:
                    a.close()
        self._dev = []
