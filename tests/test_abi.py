"""The C-ABI library loads and exports every symbol include/voxmi.h declares
(CPU only: no compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

from voxvizthesis_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header=None):
    names = [header] if header else ["voxmi.h", "voxmi_pool.h", "voxmi_serve.h", "voxmi_batch.h",
                                        "voxmi_tiler.h", "voxmi_terrain.h"]
    out = set()
    for h in names:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        out |= set(re.findall(r"\b(vox_[a-z0-9_]+)\s*\(", txt))
    return sorted(out)


def test_header_matches_binding_list():
    assert declared_symbols("voxmi.h") == sorted(_lib.EXPORTS)
    assert declared_symbols("voxmi_pool.h") == sorted(_lib.POOL_EXPORTS)
    assert declared_symbols("voxmi_serve.h") == sorted(_lib.SERVE_EXPORTS)
    assert declared_symbols("voxmi_batch.h") == sorted(_lib.BATCH_EXPORTS)
    assert declared_symbols("voxmi_tiler.h") == sorted(_lib.TILER_EXPORTS)
    assert declared_symbols("voxmi_terrain.h") == sorted(_lib.TERRAIN_EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (vox_[a-z0-9_]+)", out))
    assert set(declared_symbols()) <= exported


def test_library_carries_gfx950_code_object():
    # the offload bundle id names the target of the embedded code object
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_create_validates_dimensions_without_gpu():
    lib = _lib.load()
    u3 = ctypes.c_uint32 * 3
    h = lib.vox_create(u3(48, 48, 48), u3(8, 8, 8), None)   # 48/8 = 6, not a power of 2
    try:
        assert b"power of two" in lib.vox_get_error(h)
        assert lib.vox_set_params(h, (ctypes.c_float * 3)(0, 0, 0),
                                  (ctypes.c_float * 3)(1, 1, 1)) == 0
    finally:
        lib.vox_destroy(h)
    # the reference's default chunk with its 1024^3 / 16^3 grid: several chunks
    h = lib.vox_create(u3(1024, 1024, 1024), u3(16, 16, 16), u3(256, 256, 256))
    try:
        assert lib.vox_get_error(h) == b""
    finally:
        lib.vox_destroy(h)
    # a chunk below S must be a multiple of 2^(levels-1) (64 here)
    h = lib.vox_create(u3(1024, 1024, 1024), u3(16, 16, 16), u3(96, 96, 96))
    try:
        assert b"multiple of 2^(levels-1)" in lib.vox_get_error(h)
    finally:
        lib.vox_destroy(h)
    h = lib.vox_create(u3(128, 128, 128), u3(8, 8, 8), u3(256, 256, 256))
    try:
        assert lib.vox_get_error(h) == b""
    finally:
        lib.vox_destroy(h)
    # bricks wider than the gather's LDS tile table (7 or more tiles per
    # axis) are accepted; (B+2)^3 at or above 2^22 is refused
    h = lib.vox_create(u3(128, 128, 128), u3(64, 64, 64), None)
    try:
        assert lib.vox_get_error(h) == b""
    finally:
        lib.vox_destroy(h)
    h = lib.vox_create(u3(400, 400, 400), u3(200, 200, 200), None)
    try:
        assert b"2^22" in lib.vox_get_error(h)
    finally:
        lib.vox_destroy(h)


def test_cpp_mirror_header_compiles_and_reports(tmp_path):
    """include/voxmi_voxelizer.hpp builds with main.cu-style argument types
    and reports construction errors like cuda::Voxelizer::getErrorMessage."""
    exe = tmp_path / "hpp_mirror_check"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "hpp_mirror_check.cpp"), "-L", libdir,
                    "-lvoxmi", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "extra 8 8 8" in r.stdout


def test_noalloc_scan_rejects_sizes_beyond_its_scratch_without_gpu():
    # the size check comes before any device work, so it runs on the CPU;
    # the pointers are aligned placeholders that are never dereferenced
    lib = _lib.load()
    assert lib.vox_scan_u32_noalloc(256, 512, (1 << 32) + 1, 0) == 0
    assert b"2^32" in lib.vox_scan_get_error()
