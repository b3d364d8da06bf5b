"""ctypes wrapper around the CPU oracle (oracle/ref_cpu.cpp, oracle/ref_pool.cpp).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libref_cpu.so")

REF_MIP_CHAIN = 1
REF_FP64_TWIN = 2


class RefGroup(C.Structure):
    _fields_ = [
        ("verts", C.c_void_p), ("vnormals", C.c_void_p), ("uvs", C.c_void_p),
        ("ntri", C.c_uint64), ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
        ("is_terrain", C.c_int32), ("rgba8", C.c_void_p), ("width", C.c_int32),
        ("height", C.c_int32), ("addr0", C.c_int32), ("addr1", C.c_int32),
    ]


class RefRayCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("look", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_y_deg", C.c_float), ("width", C.c_int32), ("height", C.c_int32),
                ("frame", C.c_uint32)]


class RefRayParams(C.Structure):
    _fields_ = [("vol_min", C.c_float * 3), ("vol_max", C.c_float * 3), ("ray_step", C.c_float),
                ("lod_scalar", C.c_float), ("jitter_seed", C.c_uint32),
                ("max_traversals", C.c_uint32)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.ref_run.restype = C.c_void_p
        L.ref_run.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                              C.POINTER(C.c_float), C.POINTER(C.c_float),
                              C.POINTER(RefGroup), C.c_int32, C.c_int32,
                              C.POINTER(C.c_int32)]
        L.ref_run_chunked.restype = C.c_void_p
        L.ref_run_chunked.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.POINTER(RefGroup), C.c_int32, C.c_int32,
                                      C.POINTER(C.c_int32)]
        L.ref_error.restype = C.c_char_p
        L.ref_levels.restype = C.c_int32
        L.ref_levels.argtypes = [C.c_void_p]
        L.ref_dims.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
        for name, rt in [("ref_occupancy", C.POINTER(C.c_uint32))]:
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        for name, rt in [("ref_types", C.POINTER(C.c_uint8)),
                         ("ref_const_colors", C.POINTER(C.c_uint32)),
                         ("ref_brick_index", C.POINTER(C.c_uint32))]:
            f = getattr(L, name)
            f.restype = rt
            f.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]
        L.ref_num_bricks.restype = C.c_uint64
        L.ref_num_bricks.argtypes = [C.c_void_p, C.c_int32]
        for name in ("ref_brick_colors", "ref_brick_normals"):
            f = getattr(L, name)
            f.restype = C.POINTER(C.c_uint32)
            f.argtypes = [C.c_void_p, C.c_int32]
        L.ref_level_colors.restype = C.POINTER(C.c_uint32)
        L.ref_level_colors.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32)]
        L.ref_level_normals.restype = C.POINTER(C.c_float)
        L.ref_level_normals.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint32)]
        L.ref_free.argtypes = [C.c_void_p]
        L.ref_normal_error_fp64.restype = C.c_double
        L.ref_normal_error_fp64.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
        # node pool / brick pool (ref_pool.h)
        L.ref_pool_load.restype = C.c_void_p
        L.ref_pool_load.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.c_int32]
        L.ref_pool_error.restype = C.c_char_p
        L.ref_pool_node_dims.argtypes = [C.c_void_p, C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint64)]
        L.ref_pool_brick_dims.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
        for name in ("ref_pool_nodes", "ref_pool_brick_colors", "ref_pool_brick_normals"):
            f = getattr(L, name)
            f.restype = C.POINTER(C.c_uint32)
            f.argtypes = [C.c_void_p]
        L.ref_pool_depths.restype = C.c_int32
        L.ref_pool_depths.argtypes = [C.c_void_p]
        L.ref_pool_file_bricks.restype = C.c_uint64
        L.ref_pool_file_bricks.argtypes = [C.c_void_p, C.c_int32]
        for name in ("ref_pool_file_brick_colors", "ref_pool_file_brick_normals"):
            f = getattr(L, name)
            f.restype = C.POINTER(C.c_uint32)
            f.argtypes = [C.c_void_p, C.c_int32]
        L.ref_pool_free.argtypes = [C.c_void_p]
        L.ref_pool_compressed.restype = C.c_int32
        L.ref_pool_compressed.argtypes = [C.c_void_p]
        L.ref_pool_file_compressed.restype = C.POINTER(C.c_uint8)
        L.ref_pool_file_compressed.argtypes = [C.c_void_p, C.c_int32, C.c_int32,
                                               C.POINTER(C.c_uint64)]
        # block compression (ref_bc.h)
        L.ref_bc_blocks_per_brick.restype = C.c_uint64
        L.ref_bc_blocks_per_brick.argtypes = [C.POINTER(C.c_uint32)]
        L.ref_bc_encode_bricks.restype = None
        L.ref_bc_encode_bricks.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64,
                                           C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p]
        L.ref_bc1_decode.restype = None
        L.ref_bc1_decode.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.ref_bc3_decode.restype = None
        L.ref_bc3_decode.argtypes = [C.c_void_p, C.c_void_p]
        # ray cast (ref_raycast.h)
        u3p = C.POINTER(C.c_uint32)
        L.ref_raycast.restype = C.c_int32
        L.ref_raycast.argtypes = [C.POINTER(RefRayCamera), C.POINTER(RefRayParams), C.c_int32,
                                  u3p, C.c_void_p, u3p, C.c_void_p, u3p, C.c_void_p, C.c_void_p]
        L.ref_raycast_pool.restype = C.c_int32
        L.ref_raycast_pool.argtypes = [C.POINTER(RefRayCamera), C.POINTER(RefRayParams),
                                       C.c_int32, u3p, C.c_void_p, u3p, C.c_void_p, u3p,
                                       C.c_int32, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class RefResult:
    """Materialised oracle output (numpy copies)."""

    def __init__(self, status, levels, S, occ, tree, mips, normal_err):
        self.status = status
        self.levels = levels
        self.S = S
        self.occupancy = occ
        self.tree = tree            # list over depth d: dict(types, const, bidx, bcol, bnrm, nb)
        self.mips = mips            # list over level k: (colors (S_k^3), normals (S_k^3, 3))
        self.normal_err = normal_err


def run(scene, flags: int = 0, keep_mips: bool = True, chunk=None) -> RefResult:
    """Run the oracle on a scene.  chunk (3 ints or None): the reference's
    voxChunkDim; None or a chunk covering the voxel space is the single-chunk
    run.  With smaller chunks the levels come back chunk by chunk."""
    L = lib()
    keep = []
    groups = (RefGroup * max(1, len(scene.meshes)))()
    for i, m in enumerate(scene.meshes):
        v = np.ascontiguousarray(m.verts, dtype=np.float32)
        vn = None if m.vnormals is None else np.ascontiguousarray(m.vnormals, dtype=np.float32)
        uv = None if m.uvs is None else np.ascontiguousarray(m.uvs, dtype=np.float32)
        im = None if m.image is None else np.ascontiguousarray(m.image, dtype=np.uint8)
        keep += [v, vn, uv, im]
        lo, hi = m.bbox()
        g = groups[i]
        g.verts, g.vnormals, g.uvs = _ptr(v), _ptr(vn), _ptr(uv)
        g.ntri = m.ntri
        g.bbox_min[:] = [float(x) for x in lo]
        g.bbox_max[:] = [float(x) for x in hi]
        g.is_terrain = 1 if m.is_terrain else 0
        g.rgba8 = _ptr(im)
        g.width = 0 if im is None else im.shape[1]
        g.height = 0 if im is None else im.shape[0]
        g.addr0, g.addr1 = int(m.addr[0]), int(m.addr[1])
    N = (C.c_uint32 * 3)(*scene.N)
    B = (C.c_uint32 * 3)(*scene.B)
    p = (C.c_float * 3)(*[float(np.float32(x)) for x in scene.p])
    d = (C.c_float * 3)(*[float(np.float32(x)) for x in scene.delta])
    st = C.c_int32(0)
    if chunk is None:
        h = L.ref_run(N, B, p, d, groups, len(scene.meshes), flags, C.byref(st))
    else:
        ch = (C.c_uint32 * 3)(*[int(c) for c in chunk])
        h = L.ref_run_chunked(N, B, ch, p, d, groups, len(scene.meshes), flags, C.byref(st))
    if not h:
        raise RuntimeError("oracle failed: " + L.ref_error().decode())
    try:
        levels = L.ref_levels(h)
        S3 = (C.c_uint32 * 3)()
        L.ref_dims(h, S3)
        n = C.c_uint64()
        occ = _copy(L.ref_occupancy(h, C.byref(n)), n.value, np.uint32)
        tree = []
        B3 = [b + 2 for b in scene.B]
        bvox = B3[0] * B3[1] * B3[2]
        for dd in range(levels):
            t = _copy(L.ref_types(h, dd, C.byref(n)), n.value, np.uint8)
            cc = _copy(L.ref_const_colors(h, dd, C.byref(n)), n.value, np.uint32)
            bi = _copy(L.ref_brick_index(h, dd, C.byref(n)), n.value, np.uint32)
            nb = L.ref_num_bricks(h, dd)
            bc = _copy(L.ref_brick_colors(h, dd), nb * bvox, np.uint32).reshape(nb, bvox)
            bn = _copy(L.ref_brick_normals(h, dd), nb * bvox, np.uint32).reshape(nb, bvox)
            tree.append(dict(types=t, const=cc, bidx=bi, bcol=bc, bnrm=bn, nb=int(nb)))
        mips = []
        if keep_mips:
            for k in range(levels):
                dims = (C.c_uint32 * 3)()
                cp = L.ref_level_colors(h, k, dims)
                nv = dims[0] * dims[1] * dims[2]
                col = _copy(cp, nv, np.uint32).reshape(dims[2], dims[1], dims[0])
                npn = L.ref_level_normals(h, k, dims)
                nrm = _copy(npn, 3 * nv, np.float32).reshape(dims[2], dims[1], dims[0], 3)
                mips.append((col, nrm))
        nerr = None
        if flags & REF_FP64_TWIN:
            cnt = C.c_uint64()
            nerr = (L.ref_normal_error_fp64(h, C.byref(cnt)), int(cnt.value))
        return RefResult(int(st.value), levels, tuple(S3), occ, tree, mips, nerr)
    finally:
        L.ref_free(h)


REF_POOL_STORE_CONST = 1


class RefPool:
    """Node texture and brick pool as the runtime builds them from files."""

    def __init__(self, node_dims, blocks, nodes, brick_dims, bcol, bnrm, file_bricks):
        self.node_dims = node_dims        # (x, y, z)
        self.blocks = blocks
        self.nodes = nodes                # (z, y, x, 2) uint32
        self.brick_dims = brick_dims
        self.brick_colors = bcol          # (z, y, x) uint32
        self.brick_normals = bnrm
        self.file_bricks = file_bricks    # per depth: (colors (nb, bvox), normals (nb, bvox))


def pool_load(directory: str, slots=(8, 8, 8), store_const: bool = True) -> RefPool:
    """Load dir/tree.gvx + voxels_<d>.gvb the way GigaVoxelsReader::Load,
    BrickPool::initBricks(root, 1) and OctTreeNodePool::create3DTexture do."""
    L = lib()
    s3 = (C.c_uint32 * 3)(*slots)
    h = L.ref_pool_load(directory.encode(), s3, REF_POOL_STORE_CONST if store_const else 0)
    if not h:
        raise RuntimeError("pool oracle failed: " + L.ref_pool_error().decode())
    try:
        nd = (C.c_uint32 * 3)()
        blocks = C.c_uint64()
        L.ref_pool_node_dims(h, nd, C.byref(blocks))
        ntex = nd[0] * nd[1] * nd[2]
        nodes = _copy(L.ref_pool_nodes(h), 2 * ntex, np.uint32).reshape(nd[2], nd[1], nd[0], 2)
        bd = (C.c_uint32 * 3)()
        L.ref_pool_brick_dims(h, bd)
        nb = bd[0] * bd[1] * bd[2]
        bcol = _copy(L.ref_pool_brick_colors(h), nb, np.uint32).reshape(bd[2], bd[1], bd[0])
        bnrm = _copy(L.ref_pool_brick_normals(h), nb, np.uint32).reshape(bd[2], bd[1], bd[0])
        e = [bd[a] // slots[a] for a in range(3)]     # B + 2 per axis
        bvox = e[0] * e[1] * e[2]
        fb = []
        compressed = bool(L.ref_pool_compressed(h))
        for d in range(L.ref_pool_depths(h)):
            n = int(L.ref_pool_file_bricks(h, d))
            if n == 0 or compressed:  # compressed files keep only the block bytes
                fb.append((np.zeros((0, bvox), np.uint32), np.zeros((0, bvox), np.uint32)))
                continue
            fc = _copy(L.ref_pool_file_brick_colors(h, d), n * bvox, np.uint32).reshape(n, bvox)
            fn = _copy(L.ref_pool_file_brick_normals(h, d), n * bvox, np.uint32).reshape(n, bvox)
            fb.append((fc, fn))
        out = RefPool(tuple(nd), int(blocks.value), nodes, tuple(bd), bcol, bnrm, fb)
        out.compressed = compressed
        out.file_compressed = []
        for d in range(L.ref_pool_depths(h)):
            pair = []
            for which in (0, 1):
                nbytes = C.c_uint64()
                ptr = L.ref_pool_file_compressed(h, d, which, C.byref(nbytes))
                pair.append(_copy(ptr, nbytes.value, np.uint8))
            out.file_compressed.append(tuple(pair))
        return out
    finally:
        L.ref_pool_free(h)


def raycast(camera: dict, params: dict, levels: int, B, nodes: np.ndarray, pool: np.ndarray,
            wide: bool = False):
    """Per-pixel node usage (H, W, 12) and RGBA8 (H, W) of one frame.
    nodes: (z, y, x, 2) node texture; pool: (z, y, x) brick pool colours
    (wide: a wide pool, slots stacked along z, word 1 the slot)."""
    L = lib()
    cam = RefRayCamera()
    cam.position[:] = [float(v) for v in camera["position"]]
    cam.look[:] = [float(v) for v in camera["look"]]
    cam.up[:] = [float(v) for v in camera["up"]]
    cam.fov_y_deg = float(camera["fov_y_deg"])
    cam.width, cam.height = int(camera["width"]), int(camera["height"])
    cam.frame = int(camera.get("frame", 0))
    prm = RefRayParams()
    prm.vol_min[:] = [float(v) for v in params["vol_min"]]
    prm.vol_max[:] = [float(v) for v in params["vol_max"]]
    prm.ray_step = float(params.get("ray_step", 0.0))
    prm.lod_scalar = float(params.get("lod_scalar", 0.0))
    prm.jitter_seed = int(params.get("jitter_seed", 0))
    prm.max_traversals = int(params.get("max_traversals", 0))
    nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
    pool = np.ascontiguousarray(pool, dtype=np.uint32)
    nd = (C.c_uint32 * 3)(nodes.shape[2], nodes.shape[1], nodes.shape[0])
    pd = (C.c_uint32 * 3)(pool.shape[2], pool.shape[1], pool.shape[0])
    b3 = (C.c_uint32 * 3)(*B)
    W, H = cam.width, cam.height
    usage = np.zeros((H, W, 12), np.uint32)
    rgba = np.zeros((H, W), np.uint32)
    ok = L.ref_raycast_pool(C.byref(cam), C.byref(prm), levels, b3, nodes.ctypes.data, nd,
                            pool.ctypes.data, pd, 1 if wide else 0, usage.ctypes.data,
                            rgba.ctypes.data)
    if not ok:
        raise RuntimeError("ray cast oracle: empty volume box")
    return usage, rgba


def bc_encode_bricks(col: np.ndarray, nrm: np.ndarray, e):
    """Block-compress bricks (nbricks, e^3 texels): (BC3 colour bytes, BC1
    normal bytes), blocks_per_brick * 16 / * 8 bytes per brick."""
    L = lib()
    e3 = (C.c_uint32 * 3)(*e)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    nrm = np.ascontiguousarray(nrm, dtype=np.uint32)
    nb = col.shape[0]
    blocks = int(L.ref_bc_blocks_per_brick(e3))
    oc = np.zeros(nb * blocks * 16, np.uint8)
    on = np.zeros(nb * blocks * 8, np.uint8)
    if nb:
        L.ref_bc_encode_bricks(col.ctypes.data, nrm.ctypes.data, nb, e3, oc.ctypes.data,
                               on.ctypes.data)
    return oc, on


def bc1_decode(block: bytes, four_always: bool = False) -> np.ndarray:
    out = np.zeros(16, np.uint32)
    b = np.frombuffer(bytes(block), np.uint8).copy()
    lib().ref_bc1_decode(b.ctypes.data, 1 if four_always else 0, out.ctypes.data)
    return out


def bc3_decode(block: bytes) -> np.ndarray:
    out = np.zeros(16, np.uint32)
    b = np.frombuffer(bytes(block), np.uint8).copy()
    lib().ref_bc3_decode(b.ctypes.data, out.ctypes.data)
    return out
