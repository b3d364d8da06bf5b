"""vox_scan_u32_noalloc (GPU): the scan with its scratch in a fixed device
array inside the library gives the same output as vox_scan_u32, whose
scratch is allocated per call, through the C ABI.  Two calls in a row, the
second smaller, so the second reads the array the first left behind; then
the input-clearing variant the same way."""
import numpy as np
import pytest

from voxvizthesis_amd import _lib

pytestmark = pytest.mark.gpu


def _dev(a):
    return _lib.DeviceArray(np.ascontiguousarray(a))


def _down(L, d, n):
    out = np.empty(n, np.uint32)
    assert L.vox_device_download(out.ctypes.data, d.ptr, 0, out.nbytes)
    return out


def _run(L, fn, a, clear):
    n = a.size
    din = _dev(a if n else np.zeros(4, np.uint32))
    dout = _dev(np.full(n + 1, 0xDEADBEEF, np.uint32))
    try:
        assert fn(din.ptr, dout.ptr, n, clear) == 1, L.vox_scan_get_error()
        return _down(L, dout, n + 1), _down(L, din, n)
    finally:
        din.close()
        dout.close()


@pytest.mark.parametrize("clear", [0, 1])
def test_noalloc_scan_equals_allocating_scan(clear):
    L = _lib.load()
    rng = np.random.Generator(np.random.PCG64(11 + clear))
    for n in (8192 * 300 + 3, 8193, 5):   # each call smaller than the one before
        print(f"[noalloc] clear={clear} n={n}", flush=True)
        a = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
        want, want_in = _run(L, L.vox_scan_u32, a, clear)
        got, got_in = _run(L, L.vox_scan_u32_noalloc, a, clear)
        assert np.array_equal(got, want)
        assert np.array_equal(got_in, want_in)
        e = np.zeros(n + 1, np.uint64)
        np.cumsum(a.astype(np.uint64), out=e[1:])
        assert np.array_equal(got, (e & 0xFFFFFFFF).astype(np.uint32))
