"""Sparse-octree ray cast producing brick requests, and the brick server
that admits them into the pool (include/voxmi_serve.h).

`RayCaster` runs the front-to-back traversal of VoxViz/Shaders/GigaVoxels.frag
over an `OctTree`'s node texture and brick pool on the device and returns the
frame's per-pixel node usage and its distinct requested node ids.
`BrickServer` takes a ray caster's request list each frame and uploads the
missing bricks into least recently used slots (BrickPool::update,
GigaVoxelsBrickPool.cpp 561-682), all on the device.  `BrickServer.shared`
serves several octrees on one brick pool (one LRU for every tree, as the
reference's BrickPool singleton), with trees attached and detached as a
pager loads and unloads them.  Frames are queued without host waits; the
accessors that read results wait for the work they read.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def camera(position, look, up, fov_y_deg, width, height, frame=0) -> _lib.VoxCamera:
    c = _lib.VoxCamera()
    c.position[:] = [float(v) for v in position]
    c.look[:] = [float(v) for v in look]
    c.up[:] = [float(v) for v in up]
    c.fov_y_deg = float(fov_y_deg)
    c.width, c.height, c.frame = int(width), int(height), int(frame)
    return c


def raycast_params(vol_min, vol_max, ray_step=0.0, lod_scalar=0.0, jitter_seed=0,
                   max_traversals=0, rect=None) -> _lib.VoxRaycastParams:
    """rect: (x0, y0, x1, y1) pixels to cast (x1, y1 exclusive), or None for
    the whole viewport."""
    p = _lib.VoxRaycastParams()
    p.vol_min[:] = [float(v) for v in vol_min]
    p.vol_max[:] = [float(v) for v in vol_max]
    p.ray_step, p.lod_scalar = float(ray_step), float(lod_scalar)
    p.jitter_seed, p.max_traversals = int(jitter_seed), int(max_traversals)
    p.rect[:] = [0, 0, 0, 0] if rect is None else [int(v) for v in rect]
    return p


def box_rects(cam: dict, box_min, box_max) -> list:
    """Pixel rectangles holding the projections of m world boxes (box_min,
    box_max: (m, 3)) for a camera dict (position, look, up, fov_y_deg,
    width, height), each with a one-pixel margin; None for a box with a
    corner at or behind the camera plane (then the whole viewport is cast),
    () for one whose projection misses the viewport.  Float64 throughout,
    dot products summed x, y, z in that order."""
    lo = np.asarray(box_min, np.float64).reshape(-1, 3)
    hi = np.asarray(box_max, np.float64).reshape(-1, 3)
    pos = np.asarray(cam["position"], np.float64)
    look = np.asarray(cam["look"], np.float64)
    up = np.asarray(cam["up"], np.float64)
    left = np.cross(up, look)
    W, H = int(cam["width"]), int(cam["height"])
    tan_y = np.tan(np.radians(float(cam["fov_y_deg"])) * 0.5)
    tan_x = tan_y * W / H
    c = np.arange(8)
    sel = np.stack([(c & 1) != 0, (c & 2) != 0, (c & 4) != 0], axis=1)       # (8, 3)
    corners = np.where(sel[None], hi[:, None, :], lo[:, None, :]) - pos      # (m, 8, 3)

    def dot(v):
        return (corners[..., 0] * v[0] + corners[..., 1] * v[1]) + corners[..., 2] * v[2]

    zc = dot(look)
    behind = (zc <= 1e-9).any(axis=1)
    zs = np.where(zc <= 1e-9, 1.0, zc)
    sx = (-dot(left) / zs) / tan_x                                            # right = -left
    sy = (dot(up) / zs) / tan_y
    xs = (sx + 1.0) * 0.5 * W
    ys = (sy + 1.0) * 0.5 * H
    x0 = np.maximum(np.floor(xs.min(axis=1)).astype(np.int64) - 1, 0)
    y0 = np.maximum(np.floor(ys.min(axis=1)).astype(np.int64) - 1, 0)
    x1 = np.minimum(np.ceil(xs.max(axis=1)).astype(np.int64) + 1, W)
    y1 = np.minimum(np.ceil(ys.max(axis=1)).astype(np.int64) + 1, H)
    out = []
    for k in range(lo.shape[0]):
        if behind[k]:
            out.append(None)
        elif x1[k] <= x0[k] or y1[k] <= y0[k]:
            out.append(())
        else:
            out.append((int(x0[k]), int(y0[k]), int(x1[k]), int(y1[k])))
    return out


def box_rect(cam: dict, box_min, box_max):
    """box_rects for one box."""
    return box_rects(cam, box_min, box_max)[0]


class RayCaster:
    def __init__(self, octree, max_width, max_height, max_requests=32768):
        self._L = _lib.load()
        self.tree = octree
        self._h = self._L.vox_raycaster_create(octree._h, int(max_width), int(max_height),
                                               int(max_requests))
        if not self._h:
            raise MemoryError("vox_raycaster_create")
        err = self.getErrorMessage()
        if err:
            self.close()
            raise RuntimeError(err)

    def getErrorMessage(self) -> str:
        return self._L.vox_raycaster_get_error(self._h).decode()

    def frame(self, cam: _lib.VoxCamera, prm: _lib.VoxRaycastParams, wait: bool = True):
        """Queue one frame; with wait, return its view (counts, timings)."""
        if not self._L.vox_raycast_frame(self._h, C.byref(cam), C.byref(prm)):
            raise RuntimeError(self.getErrorMessage())
        return self.view() if wait else None

    def view(self) -> _lib.VoxFrameView:
        v = _lib.VoxFrameView()
        if not self._L.vox_raycaster_get_frame(self._h, C.byref(v)):
            raise RuntimeError(self.getErrorMessage())
        return v

    def _copy(self, ptr, n):
        a = np.empty(n, np.uint32)
        if n and not self._L.vox_pool_copy_to_host(a.ctypes.data, ptr, a.nbytes):
            raise RuntimeError("copy failed")
        return a

    def usage(self) -> np.ndarray:
        v = _lib.VoxFrameView()
        self._L.vox_raycaster_get_frame(self._h, C.byref(v))
        return self._copy(v.usage, v.width * v.height * 12).reshape(v.height, v.width, 12)

    def rgba(self) -> np.ndarray:
        v = _lib.VoxFrameView()
        self._L.vox_raycaster_get_frame(self._h, C.byref(v))
        return self._copy(v.rgba, v.width * v.height).reshape(v.height, v.width)

    def rgba_f32(self) -> np.ndarray:
        """The last frame's FragColor, (H, W, 4) float32."""
        v = self.view()
        a = np.empty(v.width * v.height * 4, np.float32)
        if a.size and not self._L.vox_pool_copy_to_host(a.ctypes.data, v.rgba_f32, a.nbytes):
            raise RuntimeError("copy failed")
        return a.reshape(v.height, v.width, 4)

    def requests(self) -> np.ndarray:
        v = _lib.VoxFrameView()
        self._L.vox_raycaster_get_frame(self._h, C.byref(v))
        return self._copy(v.requests, v.nrequests)

    def close(self):
        if getattr(self, "_h", None):
            self._L.vox_raycaster_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def composite(raycasters, dst_ptr: int, width: int, height: int, clear: bool = True):
    """vox_composite: blend the casters' last frames, in the order given, into
    the RGBA32F device buffer dst_ptr (width x height x 4 floats)."""
    L = _lib.load()
    arr = (C.c_void_p * max(len(raycasters), 1))(*[r._h for r in raycasters])
    if not L.vox_composite(arr, len(raycasters), dst_ptr, int(width), int(height),
                           1 if clear else 0):
        raise RuntimeError("vox_composite failed")


def raycast_front_to_back(raycasters, cam: _lib.VoxCamera, prms, dst_ptr: int,
                          t_stop: float = 0.0, trans_ptr=None, resume: bool = False):
    """vox_raycast_frame_multi: one front-to-back pass over the casters'
    octrees (listed in an order every ray meets them in), one parameter set
    per caster, into the RGBA32F device buffer dst_ptr (every pixel
    written); each caster gets its octree's request list.  trans_ptr
    (width*height floats) receives each pixel's transmittance; resume
    continues a frame from dst_ptr and trans_ptr."""
    L = _lib.load()
    n = len(raycasters)
    arr = (C.c_void_p * max(n, 1))(*[r._h for r in raycasters])
    parr = (_lib.VoxRaycastParams * max(n, 1))(*prms)
    if not L.vox_raycast_frame_multi(arr, n, C.byref(cam), parr, dst_ptr, trans_ptr,
                                     float(t_stop), 1 if resume else 0):
        raise RuntimeError("vox_raycast_frame_multi failed: " +
                           (raycasters[0].getErrorMessage() if n else "no octrees"))


class BrickServer:
    def __init__(self, octree, max_uploads_per_frame=0, _handle=None):
        self._L = _lib.load()
        self.tree = octree
        self._h = _handle if _handle is not None else self._L.vox_server_create(
            octree._h, int(max_uploads_per_frame))
        if not self._h:
            raise MemoryError("vox_server_create")
        err = self.getErrorMessage()
        if err:
            self.close()
            raise RuntimeError(err)

    @classmethod
    def shared(cls, pool, max_trees, max_texels_per_tree, max_uploads_per_frame=0):
        """Server over `pool` for up to max_trees attached octrees."""
        L = _lib.load()
        h = L.vox_server_create_shared(pool._h, int(max_trees), int(max_texels_per_tree),
                                       int(max_uploads_per_frame))
        return cls(None, _handle=h)

    def attach(self, octree) -> int:
        i = C.c_uint32()
        if not self._L.vox_server_attach(self._h, octree._h, C.byref(i)):
            raise RuntimeError(self.getErrorMessage())
        return i.value

    def detach(self, octree):
        if not self._L.vox_server_detach(self._h, octree._h):
            raise RuntimeError(self.getErrorMessage())

    def getErrorMessage(self) -> str:
        return self._L.vox_server_get_error(self._h).decode()

    def frame(self, raycaster, wait: bool = True):
        """Serve one ray caster's (or a list of ray casters') last frame;
        with wait, return the stats after it."""
        rcs = raycaster if isinstance(raycaster, (list, tuple)) else [raycaster]
        arr = (C.c_void_p * max(len(rcs), 1))(*[r._h for r in rcs])
        if not self._L.vox_server_frame_multi(self._h, arr, len(rcs)):
            raise RuntimeError(self.getErrorMessage())
        return self.stats() if wait else None

    def stats(self) -> _lib.VoxServerStats:
        s = _lib.VoxServerStats()
        if not self._L.vox_server_get_stats(self._h, C.byref(s)):
            raise RuntimeError(self.getErrorMessage())
        return s

    def latencies_us(self) -> np.ndarray:
        n = C.c_uint64()
        self._L.vox_server_latencies(self._h, None, 0, C.byref(n))
        out = np.empty(n.value, np.uint64)
        if n.value:
            self._L.vox_server_latencies(self._h, out.ctypes.data, n.value, C.byref(n))
        khz = self.stats().clock_khz
        return out.astype(np.float64) * (1e3 / khz)

    def state(self):
        """(slot of each node texel, node texel of each slot) as host arrays."""
        ns, sn = C.c_void_p(), C.c_void_p()
        nt, nsl = C.c_uint64(), C.c_uint32()
        self._L.vox_server_state(self._h, C.byref(ns), C.byref(sn), C.byref(nt), C.byref(nsl))
        a = np.empty(nt.value, np.uint32)
        b = np.empty(nsl.value, np.uint32)
        self._L.vox_pool_copy_to_host(a.ctypes.data, ns, a.nbytes)
        self._L.vox_pool_copy_to_host(b.ctypes.data, sn, b.nbytes)
        return a, b

    def close(self):
        if getattr(self, "_h", None):
            self._L.vox_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
