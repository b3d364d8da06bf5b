"""ctypes wrapper around the CPU oracle (oracle/ref_cpu.cpp).

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


def run(scene, flags: int = 0, keep_mips: bool = True) -> RefResult:
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
    h = L.ref_run(N, B, p, d, groups, len(scene.meshes), flags, C.byref(st))
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
