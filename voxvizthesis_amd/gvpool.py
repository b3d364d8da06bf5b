"""GigaVoxels node pool and brick pool on the device (include/voxmi_pool.h).

Python mirror of the runtime side the voxelizer's output is served from:
`BrickPool` follows gv::BrickPool (VoxViz/GigaVoxels/GigaVoxelsBrickPool.h)
and `OctTree` the node pool of gv::GigaVoxelsOctTree as GigaVoxelsReader::Load
and OctTreeNodePool::create3DTexture build it (GigaVoxelsReader.cpp 277-514,
GigaVoxelsOctTreeNodePool.cpp 49-223), but from the voxelizer's result in HBM
instead of from files.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class BrickPool:
    """BrickPool::initSpecial (GigaVoxelsBrickPool.cpp 153-250): a 3D texture of
    (B+2)*slots texels per axis.  slots=None takes the renderer's choice, 8^3
    for bricks wider than 14 voxels, else 16^3 (GigaVoxelsRenderer.cpp 259-272).
    """

    def __init__(self, brickDim, slots=None, wide_slots=None, normals: bool = True):
        """wide_slots: a wide pool of that many slots instead
        (vox_brick_pool_create_wide: beyond the 10-bit texel pointers, word 1
        of a resident node holds its slot); with normals=False a colour-only
        wide pool (vox_brick_pool_create_colour)."""
        self._L = _lib.load()
        u3 = C.c_uint32 * 3
        if wide_slots is not None and not normals:
            self._h = self._L.vox_brick_pool_create_colour(u3(*brickDim), int(wide_slots))
        elif wide_slots is not None:
            self._h = self._L.vox_brick_pool_create_wide(u3(*brickDim), int(wide_slots))
        else:
            self._h = self._L.vox_brick_pool_create(u3(*brickDim),
                                                    u3(*slots) if slots is not None else None)
        if not self._h:
            raise MemoryError("vox_brick_pool_create")
        err = self.getErrorMessage()
        if err:
            self.close()
            raise RuntimeError(err)

    def getErrorMessage(self) -> str:
        return self._L.vox_brick_pool_get_error(self._h).decode()

    def view(self) -> _lib.VoxBrickPoolView:
        v = _lib.VoxBrickPoolView()
        if not self._L.vox_brick_pool_get_view(self._h, C.byref(v)):
            raise RuntimeError(self.getErrorMessage())
        return v

    def dims(self):
        return tuple(self.view().dims)

    def textures(self):
        """(colors, normals) as host arrays of shape (z, y, x), uint32
        (normals None for a colour-only pool)."""
        v = self.view()
        n = v.dims[0] * v.dims[1] * v.dims[2]
        out = []
        for ptr in (v.colors, v.normals):
            if not ptr:
                out.append(None)
                continue
            a = np.empty(n, np.uint32)
            if not self._L.vox_pool_copy_to_host(a.ctypes.data, ptr, a.nbytes):
                raise RuntimeError("copy failed")
            out.append(a.reshape(v.dims[2], v.dims[1], v.dims[0]))
        return tuple(out)

    def close(self):
        if getattr(self, "_h", None):
            self._L.vox_brick_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OctTree:
    """One voxelized octree's node pool (RG32UI texels, 2 words each).

    const_mode picks word 1 of constant nodes: "text" (default) the colour the
    reader parses back from tree.gvx, "reference" 0 as the reference's reader
    leaves it (GigaVoxelsOctTree.cpp 1611-1621), "exact" the voxelizer's colour.
    The voxelizer must not generate again while the tree is in use: bricks are
    read from its result when they are uploaded.
    """

    CONST_MODES = {"text": 0, "reference": _lib.VOX_TREE_REFERENCE_CONST,
                   "exact": _lib.VOX_TREE_EXACT_CONST}

    def __init__(self, brickPool: BrickPool, voxelizer, const_mode: str = "text",
                 root_brick: bool = True):
        """root_brick=False leaves the root's brick to a shared brick server
        (VOX_TREE_NO_ROOT_BRICK), as for trees a pager loads."""
        self._L = _lib.load()
        self.pool = brickPool
        self._res = voxelizer.raw_result()
        flags = self.CONST_MODES[const_mode]
        if not root_brick:
            flags |= _lib.VOX_TREE_NO_ROOT_BRICK
        self._h = self._L.vox_octree_create(brickPool._h, C.byref(self._res), flags)
        if not self._h:
            raise MemoryError("vox_octree_create")
        err = self.getErrorMessage()
        if err:
            self.close()
            raise RuntimeError(err)

    def getErrorMessage(self) -> str:
        return self._L.vox_octree_get_error(self._h).decode()

    def view(self) -> _lib.VoxOctreeView:
        v = _lib.VoxOctreeView()
        if not self._L.vox_octree_get_view(self._h, C.byref(v)):
            raise RuntimeError(self.getErrorMessage())
        return v

    def get3DTextureDimensions(self):
        return tuple(self.view().node_dim)

    def getNodeCount(self) -> int:
        """OctTreeNodePool::getNodeCount: 8 nodes per block."""
        return 8 * self.view().node_blocks

    def nodes(self) -> np.ndarray:
        """Node texture as (z, y, x, 2) uint32."""
        v = self.view()
        n = v.node_dim[0] * v.node_dim[1] * v.node_dim[2]
        a = np.empty(2 * n, np.uint32)
        if not self._L.vox_pool_copy_to_host(a.ctypes.data, v.nodes, a.nbytes):
            raise RuntimeError("copy failed")
        return a.reshape(v.node_dim[2], v.node_dim[1], v.node_dim[0], 2)

    def node_texel(self, depth: int) -> np.ndarray:
        v = self.view()
        a = np.empty(8 ** depth, np.uint32)
        if not self._L.vox_pool_copy_to_host(a.ctypes.data, v.node_texel[depth], a.nbytes):
            raise RuntimeError("copy failed")
        return a

    def close(self):
        if getattr(self, "_h", None):
            self._L.vox_octree_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
