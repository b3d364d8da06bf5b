"""Python mirror of the reference interface cuda::Voxelizer
(VoxViz/Voxelizer/Voxelizer/Voxelizer.h 22-244) over the C ABI of
libvoxmi.so.  Method names and return conventions follow the reference:
boolean members return True/False, generateVoxelsAndOctTree returns
1 / 0 / -1, errors accumulate and are read with getErrorMessage().
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

ADDR_WRAP, ADDR_CLAMP, ADDR_MIRROR, ADDR_BORDER = 0, 1, 2, 3


class Voxelizer:
    """voxChunkDim: the chunk edge per axis (Voxelizer.h 124).  None here
    means one chunk that holds the whole voxel space; the reference's own
    default is (256, 256, 256), which its caller keeps (main.cu 4169-4170):
    pass it to get the reference's chunked results for N above 256 (the C++
    mirror, include/voxmi_voxelizer.hpp, defaults to it)."""

    def __init__(self, voxDim, brickDim, voxChunkDim=None, options: int = 0):
        self._L = _lib.load()
        n = (C.c_uint32 * 3)(*[int(x) for x in voxDim])
        b = (C.c_uint32 * 3)(*[int(x) for x in brickDim])
        c = None if voxChunkDim is None else (C.c_uint32 * 3)(*[int(x) for x in voxChunkDim])
        self._h = self._L.vox_create(n, b, c)
        if not self._h:
            raise MemoryError("vox_create failed")
        self.voxDim = tuple(int(x) for x in voxDim)
        self.brickDim = tuple(int(x) for x in brickDim)
        self._keep = []
        self._L.vox_set_options(self._h, options)

    # --- reference members -------------------------------------------------
    def setVoxelizationParams(self, p, deltaP) -> bool:
        return bool(self._L.vox_set_params(self._h, (C.c_float * 3)(*map(float, p)),
                                           (C.c_float * 3)(*map(float, deltaP))))

    def addTriangleGroup(self, verts, vertNormals=None, uvs=None, minXYZ=None, maxXYZ=None,
                         isTerrain=False, image=None, texAddressMode0=ADDR_CLAMP,
                         texAddressMode1=ADDR_CLAMP) -> bool:
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3, 3)
        vn = None if vertNormals is None else np.ascontiguousarray(vertNormals, np.float32)
        uv = None if uvs is None else np.ascontiguousarray(uvs, np.float32)
        im = None if image is None else np.ascontiguousarray(image, np.uint8)
        self._keep += [v, vn, uv, im]
        flat = v.reshape(-1, 3)
        if minXYZ is None:
            minXYZ = flat.min(axis=0) if flat.shape[0] else np.zeros(3, np.float32)
        if maxXYZ is None:
            maxXYZ = flat.max(axis=0) if flat.shape[0] else np.zeros(3, np.float32)
        g = _lib.VoxGroup()
        g.verts = v.ctypes.data
        g.vnormals = None if vn is None else vn.ctypes.data
        g.uvs = None if uv is None else uv.ctypes.data
        g.ntri = v.shape[0]
        g.bbox_min[:] = [float(x) for x in minXYZ]
        g.bbox_max[:] = [float(x) for x in maxXYZ]
        g.is_terrain = 1 if isTerrain else 0
        g.rgba8 = None if im is None else im.ctypes.data
        g.width = 0 if im is None else int(im.shape[1])
        g.height = 0 if im is None else int(im.shape[0])
        g.addr0, g.addr1 = int(texAddressMode0), int(texAddressMode1)
        return bool(self._L.vox_add_group(self._h, C.byref(g)))

    def addMesh(self, m) -> bool:
        lo, hi = m.bbox()
        return self.addTriangleGroup(m.verts, m.vnormals, m.uvs, lo, hi, m.is_terrain, m.image,
                                     m.addr[0], m.addr[1])

    def initCuda(self) -> bool:
        return bool(self._L.vox_init(self._h))

    def resetCuda(self) -> bool:
        return bool(self._L.vox_reset(self._h))

    def allocateVoxelMemory(self) -> bool:
        return bool(self._L.vox_allocate_voxel_memory(self._h))

    def deallocateVoxelMemory(self):
        self._L.vox_deallocate_voxel_memory(self._h)

    def allocateTriangleMemory(self) -> bool:
        return bool(self._L.vox_allocate_triangle_memory(self._h))

    def deallocateTriangleMemory(self):
        self._L.vox_deallocate_triangle_memory(self._h)
        self._keep = []

    def computeEdgesFaceNormalsAndBounds(self) -> bool:
        return bool(self._L.vox_compute_edges_face_normals_and_bounds(self._h))

    def generateVoxelsAndOctTree(self, outputDir=None, outputBinary=True,
                                 outputCompressed=False) -> int:
        od = None if outputDir is None else str(outputDir).encode()
        return int(self._L.vox_generate(self._h, od, 1 if outputBinary else 0,
                                        1 if outputCompressed else 0))

    def getErrorMessage(self) -> str:
        return self._L.vox_get_error(self._h).decode()

    # --- asynchronous generate (not in the reference) -------------------------
    def generateAsync(self, outputCompressed=False) -> bool:
        """Start generateVoxelsAndOctTree (no files) on a worker thread."""
        return bool(self._L.vox_generate_async(self._h, 1 if outputCompressed else 0))

    def wait(self, block: bool = True) -> int:
        """1 / 0 / -1 like generateVoxelsAndOctTree, 2 while still running
        (block=False)."""
        return int(self._L.vox_wait(self._h, 1 if block else 0))

    def releaseWorkMemory(self) -> bool:
        """Free triangles, bins and level tiles, keeping the result."""
        ok = bool(self._L.vox_release_work_memory(self._h))
        if ok:
            self._keep = []
        return ok

    # --- results -------------------------------------------------------------
    def raw_result(self) -> "_lib.VoxResult":
        r = _lib.VoxResult()
        if not self._L.vox_get_result(self._h, C.byref(r)):
            raise RuntimeError(self.getErrorMessage())
        return r

    def stats(self) -> "_lib.VoxStats":
        s = _lib.VoxStats()
        self._L.vox_get_stats(self._h, C.byref(s))
        return s

    def _d2h(self, ptr, n, dtype):
        out = np.empty(n, dtype=dtype)
        if n:
            if not self._L.vox_copy_to_host(self._h, out.ctypes.data, ptr, out.nbytes):
                raise RuntimeError(self.getErrorMessage())
        return out

    def result(self, with_bricks: bool = True) -> dict:
        """Copy the device-resident result to numpy (tests / writers)."""
        r = self.raw_result()
        B3 = [b + 2 for b in r.brick_dim]
        bvox = B3[0] * B3[1] * B3[2]
        out = dict(levels=r.levels, S=tuple(r.S), tree=[])
        out["occupancy"] = (self._d2h(r.occupancy, r.occupancy_words, np.uint32)
                            if r.occupancy else None)
        for d in range(r.levels):
            n = r.nodes[d]
            nb = r.bricks[d]
            t = dict(types=self._d2h(r.types[d], n, np.uint8),
                     const=self._d2h(r.const_colors[d], n, np.uint32),
                     bidx=self._d2h(r.brick_index[d], n, np.uint32), nb=int(nb))
            if with_bricks:
                t["bcol"] = self._d2h(r.brick_colors[d], nb * bvox, np.uint32).reshape(nb, bvox)
                t["bnrm"] = self._d2h(r.brick_normals[d], nb * bvox, np.uint32).reshape(nb, bvox)
            out["tree"].append(t)
        return out

    def level(self, k: int):
        """Dense copy of mip level k: (colors (Z,Y,X) uint32, normals (Z,Y,X,3))."""
        d = (C.c_uint32 * 3)()
        if not self._L.vox_copy_level(self._h, k, None, None, d):   # extent only
            raise RuntimeError(self.getErrorMessage())
        dims = [int(x) for x in d]
        n = dims[0] * dims[1] * dims[2]
        col = np.empty(n, np.uint32)
        nrm = np.empty(3 * n, np.float32)
        if not self._L.vox_copy_level(self._h, k, col.ctypes.data, nrm.ctypes.data, d):
            raise RuntimeError(self.getErrorMessage())
        return (col.reshape(dims[2], dims[1], dims[0]),
                nrm.reshape(dims[2], dims[1], dims[0], 3))

    def compressed(self):
        """Per depth (BC3 colour bytes, BC1 normal bytes) of the last
        generateVoxelsAndOctTree(..., outputCompressed=True), or None."""
        z = _lib.VoxCompressed()
        if not self._L.vox_get_compressed(self._h, C.byref(z)):
            return None
        out = []
        for d in range(z.levels):
            nb = z.bricks[d] * z.blocks_per_brick
            out.append((self._d2h(z.colors[d], nb * 16, np.uint8),
                        self._d2h(z.normals[d], nb * 8, np.uint8)))
        return out

    def writeFiles(self, outputDir) -> bool:
        return bool(self._L.vox_write_files(self._h, str(outputDir).encode()))

    def close(self):
        if getattr(self, "_h", None):
            self._L.vox_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def voxelize_scene(scene, options: int = 0, chunk=None, out_dir=None,
                   compressed: bool = False) -> Voxelizer:
    """Run the reference call sequence (main.cu Voxelize 3225-3341) on a
    synthetic Scene and return the voxelizer holding the device result."""
    vz = Voxelizer(scene.N, scene.B, chunk, options=options)
    ok = (vz.initCuda() and vz.allocateVoxelMemory()
          and vz.setVoxelizationParams(scene.p, scene.delta))
    for m in scene.meshes:
        ok = ok and vz.addMesh(m)
    ok = ok and vz.allocateTriangleMemory() and vz.computeEdgesFaceNormalsAndBounds()
    if not ok:
        raise RuntimeError(vz.getErrorMessage())
    st = vz.generateVoxelsAndOctTree(out_dir, True, compressed)
    if st == 0:
        raise RuntimeError(vz.getErrorMessage())
    vz.status = st
    return vz
