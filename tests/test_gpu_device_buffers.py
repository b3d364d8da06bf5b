"""Device buffers of the C ABI (GPU): upload at an offset, the plain copy
kernel and the read back give the bytes back exactly; misaligned or
unsized copies are refused."""
import ctypes as C

import numpy as np
import pytest

from voxvizthesis_amd import _lib

pytestmark = pytest.mark.gpu


def test_upload_copy_readback():
    L = _lib.load()
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2**32, size=(1 << 20) + 4, dtype=np.uint32)
    nb = a.nbytes
    src = L.vox_device_alloc(nb)
    dst = L.vox_device_alloc(nb)
    try:
        assert src and dst
        # two pieces, the second at an offset
        half = (a.size // 2) * 4
        assert L.vox_device_upload(src, 0, a.ctypes.data, half)
        assert L.vox_device_upload(src, half, a.ctypes.data + half, nb - half)
        assert L.vox_device_copy(dst, src, nb)
        back = np.empty_like(a)
        assert L.vox_pool_copy_to_host(back.ctypes.data, dst, nb)
        assert (back == a).all()
        # the copy kernel takes 16-byte multiples of 16-byte aligned pointers only
        assert not L.vox_device_copy(dst, src, nb - 4)
        assert not L.vox_device_copy(C.c_void_p(dst + 4), src, 16)
        assert L.vox_device_upload(dst, 0, None, 0)
    finally:
        assert L.vox_device_free(src) and L.vox_device_free(dst)
    assert L.vox_device_free(None)


def test_device_array_roundtrip():
    a = np.arange(3 * 1000, dtype=np.float32).reshape(1000, 3)
    d = _lib.DeviceArray(a)
    try:
        back = np.empty_like(a)
        L = _lib.load()
        assert L.vox_pool_copy_to_host(back.ctypes.data, d.ptr, a.nbytes)
        assert (back == a).all()
    finally:
        d.close()


def test_failed_allocation_leaves_later_calls_working():
    # a device allocation that cannot succeed fails cleanly, and the runtime
    # error it raised is not reported again by the next calls
    L = _lib.load()
    assert not L.vox_device_alloc(1 << 60)
    a = np.arange(64, dtype=np.uint32)
    d = _lib.DeviceArray(a)
    try:
        back = np.empty_like(a)
        assert L.vox_pool_copy_to_host(back.ctypes.data, d.ptr, a.nbytes)
        assert (back == a).all()
    finally:
        d.close()
    from voxvizthesis_amd import scenes
    from voxvizthesis_amd.voxelizer import voxelize_scene
    vz = voxelize_scene(scenes.scene_two_triangles())
    try:
        assert vz.status == 1, vz.getErrorMessage()
    finally:
        vz.close()
