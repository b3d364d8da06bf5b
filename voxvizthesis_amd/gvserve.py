"""Sparse-octree ray cast producing brick requests, and the brick server
that admits them into the pool (include/voxmi_serve.h).

`RayCaster` runs the front-to-back traversal of VoxViz/Shaders/GigaVoxels.frag
over an `OctTree`'s node texture and brick pool on the device and returns the
frame's per-pixel node usage and its distinct requested node ids.
`BrickServer` takes a ray caster's request list each frame and uploads the
missing bricks into least recently used slots (BrickPool::update,
GigaVoxelsBrickPool.cpp 561-682), all on the device.
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
                   max_traversals=0) -> _lib.VoxRaycastParams:
    p = _lib.VoxRaycastParams()
    p.vol_min[:] = [float(v) for v in vol_min]
    p.vol_max[:] = [float(v) for v in vol_max]
    p.ray_step, p.lod_scalar = float(ray_step), float(lod_scalar)
    p.jitter_seed, p.max_traversals = int(jitter_seed), int(max_traversals)
    return p


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

    def frame(self, cam: _lib.VoxCamera, prm: _lib.VoxRaycastParams) -> _lib.VoxFrameView:
        if not self._L.vox_raycast_frame(self._h, C.byref(cam), C.byref(prm)):
            raise RuntimeError(self.getErrorMessage())
        v = _lib.VoxFrameView()
        self._L.vox_raycaster_get_frame(self._h, C.byref(v))
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


class BrickServer:
    def __init__(self, octree, max_uploads_per_frame=0):
        self._L = _lib.load()
        self.tree = octree
        self._h = self._L.vox_server_create(octree._h, int(max_uploads_per_frame))
        if not self._h:
            raise MemoryError("vox_server_create")
        err = self.getErrorMessage()
        if err:
            self.close()
            raise RuntimeError(err)

    def getErrorMessage(self) -> str:
        return self._L.vox_server_get_error(self._h).decode()

    def frame(self, raycaster: RayCaster):
        if not self._L.vox_server_frame(self._h, raycaster._h):
            raise RuntimeError(self.getErrorMessage())
        return self.stats()

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
