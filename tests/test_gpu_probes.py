"""The bandwidth probes behind the benchmark's copy and read-stream figures
(pisces_copy_probe, pisces_stream_probe) on the GPU: each launch moves the
requested bytes in a positive time, below what HBM's peak would allow."""
import ctypes

import pytest

pytestmark = pytest.mark.gpu

PEAK_BYTES_PER_S = 8e12


def test_copy_and_stream_probes(pisces):
    L = pisces.lib()
    L.pisces_copy_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]
    L.pisces_stream_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_double)]
    nbytes = 1 << 30  # well above the 256 MB memory-side cache, which can beat HBM's peak
    ms = ctypes.c_double(0)
    assert L.pisces_copy_probe(0, nbytes, 3, ctypes.byref(ms)) == 0
    rate = 2 * nbytes / (ms.value * 1e-3)
    print("copy %.0f GB/s" % (rate / 1e9))
    assert 0 < rate < PEAK_BYTES_PER_S
    for nt in (0, 1):
        assert L.pisces_stream_probe(0, nbytes, 3, nt, ctypes.byref(ms)) == 0
        rate = nbytes / (ms.value * 1e-3)
        print("read (nontemporal %d) %.0f GB/s" % (nt, rate / 1e9))
        assert 0 < rate < PEAK_BYTES_PER_S
