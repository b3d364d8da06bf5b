"""The device-wide exclusive prefix sum (vox_scan_u32) against numpy (GPU).

The single-pass decoupled look-back scan sizes the voxelizer's bins, tile
lists and brick addresses.  Here it runs on its own: random uint32 arrays of
sizes around its 8192-value tile (one tile, a tile and a ragged tail, over
256 predecessor tiles for the look-back window), values that wrap modulo
2^32, n not a multiple of 4, and the variant that zeroes its input."""
import ctypes as C

import numpy as np
import pytest

from voxvizthesis_amd import _lib

pytestmark = pytest.mark.gpu


def _dev(a):
    return _lib.DeviceArray(np.ascontiguousarray(a))


def _down(L, d, n, dtype=np.uint32):
    out = np.empty(n, dtype)
    assert L.vox_device_download(out.ctypes.data, d.ptr, 0, out.nbytes)
    return out


def _expect(a):
    e = np.zeros(a.size + 1, np.uint64)
    np.cumsum(a.astype(np.uint64), out=e[1:])
    return (e & 0xFFFFFFFF).astype(np.uint32)


@pytest.mark.parametrize("n", [0, 1, 3, 8191, 8192, 8193, 8192 * 300 + 3])
@pytest.mark.parametrize("clear", [0, 1])
def test_scan_matches_cumsum(n, clear):
    print(f"[scan] n={n} clear={clear}", flush=True)
    L = _lib.load()
    rng = np.random.Generator(np.random.PCG64(n + 7 * clear))
    # large values so the sums wrap modulo 2^32
    a = rng.integers(0, 1 << 32, size=max(n, 1), dtype=np.uint64).astype(np.uint32)[:n]
    din = _dev(a if n else np.zeros(4, np.uint32))
    dout = _dev(np.full(n + 1, 0xDEADBEEF, np.uint32))
    try:
        assert L.vox_scan_u32(din.ptr, dout.ptr, n, clear) == 1, L.vox_scan_get_error()
        got = _down(L, dout, n + 1)
        assert np.array_equal(got, _expect(a))
        back = _down(L, din, n)
        assert np.array_equal(back, np.zeros(n, np.uint32) if clear else a)
    finally:
        din.close()
        dout.close()


def test_scan_rejects_misaligned_and_null():
    L = _lib.load()
    d = _dev(np.zeros(64, np.uint32))
    try:
        assert L.vox_scan_u32(d.ptr + 4, d.ptr + 128, 8, 0) == 0
        assert b"aligned" in L.vox_scan_get_error()
        assert L.vox_scan_u32(d.ptr, None, 8, 0) == 0
    finally:
        d.close()
