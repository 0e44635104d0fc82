"""qd_fft_pencils_noalloc: the allocation-free variant of qd_probe_fft (its
twiddle table lives in a static device array) on caller device memory.

Bar: its output equals qd_probe_fft's on the same input within 1e-14 of the
largest output (the two compute their twiddles differently: the probe on the
host in long double, the allocation-free call with sincospi on the device),
on two calls in a row with the second smaller than the first, in both
directions; sizes past the table's 512 entries are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from qd_openmp_cuda_amd import abi, api

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _DeviceBuffer:
    """Device memory for the test, from the HIP runtime that libqdamd itself
    is bound to (looked up through the library's handle), so the allocation
    and the call share one runtime."""

    def __init__(self, nbytes):
        L = api.lib()
        self._malloc, self._free, self._memcpy, self._sync = L.hipMalloc, L.hipFree, L.hipMemcpy, L.hipDeviceSynchronize
        self._malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self._free.argtypes = [C.c_void_p]
        self._memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.ptr = C.c_void_p()
        assert self._malloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        assert a.nbytes <= self.nbytes
        assert self._memcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0      # hipMemcpyHostToDevice

    def download(self, a):
        assert self._sync() == 0
        assert self._memcpy(a.ctypes.data, self.ptr, a.nbytes, 2) == 0      # hipMemcpyDeviceToHost

    def free(self):
        self._free(self.ptr)


def test_fft_noalloc_matches_probe():
    rng = np.random.default_rng(11)
    buf = _DeviceBuffer(16 * 512 * 6)
    try:
        for n, count in ((512, 6), (96, 3)):          # the second call is smaller than the first
            x = np.ascontiguousarray(rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n)))
            for inverse in (False, True):
                _say(f"case fft noalloc n={n} count={count} inverse={inverse}")
                buf.upload(x)
                api.fft_pencils_noalloc(buf.ptr.value, n, count, inverse)
                got = np.empty_like(x)
                buf.download(got)
                ref = api.probe_fft(x, inverse=inverse)
                err = float(np.abs(got - ref).max() / np.abs(ref).max())
                _say(f"  vs qd_probe_fft {err:.2e}")
                assert err < 1e-14
    finally:
        buf.free()


def test_fft_noalloc_refuses_oversize():
    buf = _DeviceBuffer(16 * 1024)
    try:
        with pytest.raises(api.QDError) as ei:
            api.fft_pencils_noalloc(buf.ptr.value, 1024, 1, False)
        assert ei.value.code == abi.QD_ERR_CAPACITY
        assert "512" in str(ei.value)
    finally:
        buf.free()
