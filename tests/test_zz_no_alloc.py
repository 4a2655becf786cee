"""The allocation-free dcov component entry point (kpc_dcov_components_noalloc)
against the allocating one (kpc_dcov_components), through the C ABI on
caller-owned device buffers: bit-identical values, vectors, distance sums and
moments on two calls in a row, the second with a smaller input (the static
scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def entry():
    from fastkpc_amd import _lib
    fn = _lib.lib().kpc_dcov_components_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


class Hip:
    """Device buffers from the HIP runtime the library itself is linked with
    (the test allocates; the entry point under test does not)."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def test_noalloc_components_equal_the_allocating_entry():
    import fastkpc_amd as fk
    from fastkpc_amd import _lib
    assert _lib.lib().kpc_device_count() >= 1, "no GPU visible"
    fn = entry()
    hip = Hip()
    rng = np.random.default_rng(5)
    try:
        for n, count, k in [(351, 40, 35), (200, 7, 20)]:
            print("noalloc case n=%d count=%d k=%d" % (n, count, k), flush=True)
            x = np.asfortranarray(np.tanh(rng.standard_normal((n, count))) +
                                  0.3 * rng.standard_normal((n, count)))
            vals, vecs, dsum, mom = fk.dcov_components(x, k)
            dx = hip.upload(np.ascontiguousarray(x.T))
            dvals = hip.alloc(8 * count * k)
            dvecs = hip.alloc(8 * count * k * n)
            dsm = hip.alloc(8 * count)
            dmo = hip.alloc(8 * count)
            rc = fn(dx, n, count, k, dvals, dvecs, dsm, dmo, None)
            assert rc == 0, _lib.lib().kpc_last_error()
            hip.sync()
            assert np.array_equal(hip.download(dvals, (count, k)), vals)
            assert np.array_equal(hip.download(dvecs, (count, k, n)).transpose(0, 2, 1), vecs)
            assert np.array_equal(hip.download(dsm, (count,)), dsum)
            assert np.array_equal(hip.download(dmo, (count,)), mom)
    finally:
        hip.free()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from fastkpc_amd import _lib
    fn = entry()
    one = ctypes.c_void_p(16)  # never dereferenced: the size checks come first
    rc = fn(one, 351, 65537, 35, one, one, one, one, None)
    assert rc == -3
    assert b"65536" in _lib.lib().kpc_last_error()
    assert fn(None, 351, 4, 35, one, one, one, one, None) == -1
