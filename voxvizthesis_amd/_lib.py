"""ctypes binding of libvoxmi.so (include/voxmi.h, voxmi_pool.h, voxmi_serve.h,
voxmi_batch.h, voxmi_tiler.h).

The HIP library is the product path: importing this module on a machine
without the built library raises immediately, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvoxmi.so")

VOX_MAX_LEVELS = 16
VOX_OPT_MIP_CHAIN = 1 << 0
VOX_OPT_KEEP_LEVELS = 1 << 1
VOX_OPT_NO_OCCUPANCY = 1 << 2
VOX_OPT_TIMING = 1 << 3

# Every symbol include/voxmi.h declares.
EXPORTS = [
    "vox_create", "vox_destroy", "vox_set_params", "vox_set_options", "vox_add_group",
    "vox_add_group_device", "vox_init", "vox_reset", "vox_allocate_voxel_memory",
    "vox_deallocate_voxel_memory", "vox_allocate_triangle_memory",
    "vox_deallocate_triangle_memory", "vox_compute_edges_face_normals_and_bounds",
    "vox_generate", "vox_get_error", "vox_get_result", "vox_get_stats", "vox_copy_to_host",
    "vox_copy_level", "vox_write_files", "vox_get_compressed", "vox_set_launch",
    "vox_generate_async", "vox_wait", "vox_release_work_memory", "vox_add_group_gather",
    "vox_device_alloc", "vox_device_upload", "vox_device_free", "vox_device_copy",
    "vox_device_copy_scattered",
    "vox_device_download", "vox_scan_u32", "vox_scan_u32_noalloc", "vox_scan_get_error",
]
# Every symbol include/voxmi_batch.h declares.
BATCH_EXPORTS = [
    "vox_batch_create", "vox_batch_destroy", "vox_batch_get_error", "vox_batch_capture",
    "vox_batch_launch", "vox_batch_get_stats", "vox_batch_set_voxelize",
]
# Every symbol include/voxmi_pool.h declares.
POOL_EXPORTS = [
    "vox_brick_pool_create", "vox_brick_pool_create_wide", "vox_brick_pool_create_colour",
    "vox_brick_pool_destroy", "vox_brick_pool_get_error",
    "vox_brick_pool_get_view", "vox_octree_create", "vox_octree_destroy", "vox_octree_get_error",
    "vox_octree_get_view", "vox_pool_copy_to_host",
]
# Every symbol include/voxmi_serve.h declares.
SERVE_EXPORTS = [
    "vox_raycaster_create", "vox_raycaster_destroy", "vox_raycaster_get_error",
    "vox_raycast_frame", "vox_raycaster_get_frame",
    "vox_server_create", "vox_server_destroy", "vox_server_get_error", "vox_server_frame",
    "vox_server_get_stats", "vox_server_latencies", "vox_server_state",
    "vox_server_create_shared", "vox_server_attach", "vox_server_detach",
    "vox_server_frame_multi", "vox_composite", "vox_raycast_frame_multi",
]
# Every symbol include/voxmi_terrain.h declares.
TERRAIN_EXPORTS = [
    "vox_terrain_heights", "vox_terrain_range", "vox_terrain_tile", "vox_terrain_sample",
    "vox_terrain_get_error",
]
# Every symbol include/voxmi_tiler.h declares.
TILER_EXPORTS = [
    "vox_tiler_create", "vox_tiler_destroy", "vox_tiler_get_error", "vox_tiler_assign",
    "vox_tiler_get", "vox_tiler_group_bounds",
]
VOX_TREE_REFERENCE_CONST = 1
VOX_TREE_EXACT_CONST = 2
VOX_TREE_NO_ROOT_BRICK = 4


class VoxGroup(C.Structure):
    _fields_ = [
        ("verts", C.c_void_p), ("vnormals", C.c_void_p), ("uvs", C.c_void_p),
        ("ntri", C.c_uint64), ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
        ("is_terrain", C.c_int32), ("rgba8", C.c_void_p), ("width", C.c_int32),
        ("height", C.c_int32), ("addr0", C.c_int32), ("addr1", C.c_int32),
    ]


class VoxResult(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("S", C.c_uint32 * 3), ("brick_dim", C.c_uint32 * 3),
        ("occupancy", C.c_void_p), ("occupancy_words", C.c_uint64),
        ("nodes", C.c_uint64 * VOX_MAX_LEVELS),
        ("types", C.c_void_p * VOX_MAX_LEVELS),
        ("const_colors", C.c_void_p * VOX_MAX_LEVELS),
        ("brick_index", C.c_void_p * VOX_MAX_LEVELS),
        ("bricks", C.c_uint64 * VOX_MAX_LEVELS),
        ("brick_colors", C.c_void_p * VOX_MAX_LEVELS),
        ("brick_normals", C.c_void_p * VOX_MAX_LEVELS),
        ("brick_node", C.c_void_p * VOX_MAX_LEVELS),
    ]


class VoxCompressed(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("blocks_per_brick", C.c_uint64),
        ("bricks", C.c_uint64 * VOX_MAX_LEVELS),
        ("colors", C.c_void_p * VOX_MAX_LEVELS),
        ("normals", C.c_void_p * VOX_MAX_LEVELS),
    ]


class VoxTileGrid(C.Structure):
    _fields_ = [("origin", C.c_float * 2), ("size", C.c_float * 2), ("border", C.c_float * 2),
                ("dim", C.c_uint32 * 2)]


class VoxTileLists(C.Structure):
    _fields_ = [("ntiles", C.c_uint64), ("pairs", C.c_uint64), ("tile_start", C.c_void_p),
                ("tile_tris", C.c_void_p), ("tile_zmin", C.c_void_p), ("tile_zmax", C.c_void_p)]


class VoxBatchStats(C.Structure):
    _fields_ = [("jobs", C.c_int32), ("launches", C.c_uint64), ("ms_graph_last", C.c_double),
                ("ms_check_last", C.c_double), ("ms_capture", C.c_double),
                ("graph_nodes", C.c_uint64), ("voxelize_launches", C.c_uint32),
                ("voxelize_blocks", C.c_uint32)]


class VoxStats(C.Structure):
    _fields_ = [
        ("triangles", C.c_uint64), ("pairs", C.c_uint64), ("sat_tests", C.c_uint64),
        ("active_tiles", C.c_uint64 * VOX_MAX_LEVELS), ("occupied_voxels", C.c_uint64),
        ("bricks_total", C.c_uint64), ("ms_total", C.c_double), ("nkernels", C.c_int32),
        ("kernel_name", (C.c_char * 32) * 64), ("kernel_ms", C.c_double * 64),
        ("kernel_launches", C.c_uint32 * 64),
        ("mip_source_voxels", C.c_uint64 * VOX_MAX_LEVELS), ("bins", C.c_uint64),
        ("scan_values", C.c_uint64),
    ]


class VoxBrickPoolView(C.Structure):
    _fields_ = [
        ("brick_dim", C.c_uint32 * 3), ("slots", C.c_uint32 * 3), ("dims", C.c_uint32 * 3),
        ("colors", C.c_void_p), ("normals", C.c_void_p), ("resident", C.c_uint32),
        ("wide", C.c_uint32),
    ]


class VoxOctreeView(C.Structure):
    _fields_ = [
        ("levels", C.c_int32), ("node_dim", C.c_uint32 * 3), ("node_blocks", C.c_uint64),
        ("nodes", C.c_void_p), ("node_texel", C.c_void_p * VOX_MAX_LEVELS),
    ]


class VoxCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("look", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_y_deg", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("frame", C.c_uint32)]


class VoxRaycastParams(C.Structure):
    _fields_ = [("vol_min", C.c_float * 3), ("vol_max", C.c_float * 3), ("ray_step", C.c_float),
                ("lod_scalar", C.c_float), ("jitter_seed", C.c_uint32),
                ("max_traversals", C.c_uint32), ("rect", C.c_int32 * 4)]


class VoxFrameView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgba", C.c_void_p),
                ("usage", C.c_void_p), ("requests", C.c_void_p), ("nrequests", C.c_uint32),
                ("nrequests_total", C.c_uint32), ("ms_raycast", C.c_double),
                ("ms_requests", C.c_double), ("rgba_f32", C.c_void_p)]


class VoxServerStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("requests", C.c_uint64), ("uploads", C.c_uint64),
                ("evictions", C.c_uint64), ("pending", C.c_uint32), ("resident", C.c_uint32),
                ("last_uploads", C.c_uint32), ("ms_update_last", C.c_double),
                ("ms_update_total", C.c_double), ("clock_khz", C.c_double)]


_lib = None


class DeviceArray:
    """A host array copied into device memory of the library's own runtime
    (vox_device_alloc); .ptr is the device address.  Freed by close()."""

    def __init__(self, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        self.L = load()
        self.nbytes = int(a.nbytes)
        self.ptr = self.L.vox_device_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"vox_device_alloc({self.nbytes})")
        step = 1 << 30                      # upload in 1 GiB pieces
        flat = a.reshape(-1).view(np.uint8)
        for off in range(0, self.nbytes, step):
            piece = flat[off:off + step]
            if not self.L.vox_device_upload(self.ptr, off, piece.ctypes.data, piece.nbytes):
                self.close()
                raise RuntimeError("vox_device_upload failed")

    @classmethod
    def alloc(cls, nbytes: int) -> "DeviceArray":
        """Uninitialised device memory of nbytes (for a kernel to fill)."""
        self = cls.__new__(cls)
        self.L = load()
        self.nbytes = int(nbytes)
        self.ptr = self.L.vox_device_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"vox_device_alloc({self.nbytes})")
        return self

    def close(self):
        if getattr(self, "ptr", None):
            self.L.vox_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load():
    """Load libvoxmi.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            f"g.build()'` (hipcc, gfx950). There is no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    u3 = C.POINTER(C.c_uint32)
    f3 = C.POINTER(C.c_float)
    vp = C.c_void_p
    L.vox_create.restype = vp
    L.vox_create.argtypes = [u3, u3, u3]
    L.vox_destroy.argtypes = [vp]
    L.vox_set_params.argtypes = [vp, f3, f3]
    L.vox_set_options.argtypes = [vp, C.c_uint32]
    L.vox_add_group.argtypes = [vp, C.POINTER(VoxGroup)]
    L.vox_add_group_device.argtypes = [vp, C.POINTER(VoxGroup)]
    for n in ("vox_init", "vox_reset", "vox_allocate_voxel_memory",
              "vox_allocate_triangle_memory", "vox_compute_edges_face_normals_and_bounds"):
        getattr(L, n).argtypes = [vp]
    for n in ("vox_deallocate_voxel_memory", "vox_deallocate_triangle_memory"):
        getattr(L, n).argtypes = [vp]
        getattr(L, n).restype = None
    L.vox_generate.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
    L.vox_get_error.restype = C.c_char_p
    L.vox_get_error.argtypes = [vp]
    L.vox_get_result.argtypes = [vp, C.POINTER(VoxResult)]
    L.vox_get_stats.argtypes = [vp, C.POINTER(VoxStats)]
    L.vox_copy_to_host.argtypes = [vp, vp, vp, C.c_uint64]
    L.vox_copy_level.argtypes = [vp, C.c_int32, vp, vp, u3]
    L.vox_write_files.argtypes = [vp, C.c_char_p]
    L.vox_get_compressed.argtypes = [vp, C.POINTER(VoxCompressed)]
    L.vox_set_launch.argtypes = [vp, C.c_uint32]
    L.vox_generate_async.argtypes = [vp, C.c_int32]
    L.vox_wait.argtypes = [vp, C.c_int32]
    L.vox_release_work_memory.argtypes = [vp]
    L.vox_add_group_gather.argtypes = [vp, C.POINTER(VoxGroup), vp, C.c_uint64, C.c_uint32]
    L.vox_device_alloc.restype = vp
    L.vox_device_alloc.argtypes = [C.c_uint64]
    L.vox_device_upload.argtypes = [vp, C.c_uint64, vp, C.c_uint64]
    L.vox_device_free.argtypes = [vp]
    L.vox_device_copy.argtypes = [vp, vp, C.c_uint64]
    L.vox_device_copy_scattered.argtypes = [vp, vp, C.c_uint64]
    L.vox_device_download.argtypes = [vp, vp, C.c_uint64, C.c_uint64]
    L.vox_scan_u32.argtypes = [vp, vp, C.c_uint64, C.c_int32]
    L.vox_scan_u32_noalloc.argtypes = [vp, vp, C.c_uint64, C.c_int32]
    L.vox_scan_get_error.restype = C.c_char_p
    L.vox_scan_get_error.argtypes = []
    # batch
    L.vox_batch_create.restype = vp
    L.vox_batch_create.argtypes = [C.POINTER(C.c_void_p), C.c_int32]
    L.vox_batch_destroy.argtypes = [vp]
    L.vox_batch_destroy.restype = None
    L.vox_batch_get_error.restype = C.c_char_p
    L.vox_batch_get_error.argtypes = [vp]
    L.vox_batch_capture.argtypes = [vp]
    L.vox_batch_launch.argtypes = [vp]
    L.vox_batch_get_stats.argtypes = [vp, C.POINTER(VoxBatchStats)]
    L.vox_batch_set_voxelize.argtypes = [vp, C.c_int32, C.c_uint32]
    # terrain
    L.vox_terrain_heights.argtypes = [vp, C.c_uint32, C.c_double, C.c_uint64, C.c_uint64,
                                      C.c_double]
    L.vox_terrain_range.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, vp]
    L.vox_terrain_tile.argtypes = [vp, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint32,
                                   vp, vp, vp]
    L.vox_terrain_sample.argtypes = [vp, C.c_uint32, C.c_double, vp, C.c_uint32, vp]
    L.vox_terrain_get_error.restype = C.c_char_p
    L.vox_terrain_get_error.argtypes = []
    # tiler
    L.vox_tiler_create.restype = vp
    L.vox_tiler_create.argtypes = []
    L.vox_tiler_destroy.argtypes = [vp]
    L.vox_tiler_destroy.restype = None
    L.vox_tiler_get_error.restype = C.c_char_p
    L.vox_tiler_get_error.argtypes = [vp]
    L.vox_tiler_assign.argtypes = [vp, vp, C.c_uint64, C.c_int32, C.POINTER(VoxTileGrid)]
    L.vox_tiler_get.argtypes = [vp, C.POINTER(VoxTileLists)]
    L.vox_tiler_group_bounds.argtypes = [vp, vp, C.c_uint32, vp]
    # node pool / brick pool
    L.vox_brick_pool_create.restype = vp
    L.vox_brick_pool_create.argtypes = [u3, u3]
    L.vox_brick_pool_create_wide.restype = vp
    L.vox_brick_pool_create_wide.argtypes = [u3, C.c_uint64]
    L.vox_brick_pool_create_colour.restype = vp
    L.vox_brick_pool_create_colour.argtypes = [u3, C.c_uint64]
    L.vox_brick_pool_destroy.argtypes = [vp]
    L.vox_brick_pool_destroy.restype = None
    L.vox_brick_pool_get_error.restype = C.c_char_p
    L.vox_brick_pool_get_error.argtypes = [vp]
    L.vox_brick_pool_get_view.argtypes = [vp, C.POINTER(VoxBrickPoolView)]
    L.vox_octree_create.restype = vp
    L.vox_octree_create.argtypes = [vp, C.POINTER(VoxResult), C.c_uint32]
    L.vox_octree_destroy.argtypes = [vp]
    L.vox_octree_destroy.restype = None
    L.vox_octree_get_error.restype = C.c_char_p
    L.vox_octree_get_error.argtypes = [vp]
    L.vox_octree_get_view.argtypes = [vp, C.POINTER(VoxOctreeView)]
    L.vox_pool_copy_to_host.argtypes = [vp, vp, C.c_uint64]
    # ray cast
    L.vox_raycaster_create.restype = vp
    L.vox_raycaster_create.argtypes = [vp, C.c_int32, C.c_int32, C.c_uint32]
    L.vox_raycaster_destroy.argtypes = [vp]
    L.vox_raycaster_destroy.restype = None
    L.vox_raycaster_get_error.restype = C.c_char_p
    L.vox_raycaster_get_error.argtypes = [vp]
    L.vox_raycast_frame.argtypes = [vp, C.POINTER(VoxCamera), C.POINTER(VoxRaycastParams)]
    L.vox_raycast_frame_multi.argtypes = [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(VoxCamera),
                                          C.POINTER(VoxRaycastParams), vp, vp, C.c_float,
                                          C.c_int32]
    L.vox_raycaster_get_frame.argtypes = [vp, C.POINTER(VoxFrameView)]
    # brick server
    L.vox_server_create.restype = vp
    L.vox_server_create.argtypes = [vp, C.c_uint32]
    L.vox_server_destroy.argtypes = [vp]
    L.vox_server_destroy.restype = None
    L.vox_server_get_error.restype = C.c_char_p
    L.vox_server_get_error.argtypes = [vp]
    L.vox_server_frame.argtypes = [vp, vp]
    L.vox_server_create_shared.restype = vp
    L.vox_server_create_shared.argtypes = [vp, C.c_uint32, C.c_uint64, C.c_uint32]
    L.vox_server_attach.argtypes = [vp, vp, C.POINTER(C.c_uint32)]
    L.vox_server_detach.argtypes = [vp, vp]
    L.vox_server_frame_multi.argtypes = [vp, C.POINTER(C.c_void_p), C.c_uint32]
    L.vox_composite.argtypes = [C.POINTER(C.c_void_p), C.c_uint32, vp, C.c_int32, C.c_int32,
                                C.c_int32]
    L.vox_server_get_stats.argtypes = [vp, C.POINTER(VoxServerStats)]
    L.vox_server_latencies.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
    L.vox_server_state.argtypes = [vp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
    _lib = L
    return L
