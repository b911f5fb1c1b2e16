"""Plain-copy probe: copies BYTES (default 1 GiB) device-to-device with a
streaming kernel and prints the achieved bandwidth.  Run under
`rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) to calibrate the byte counters:
one launch moves exactly BYTES in and BYTES out."""
import ctypes
import json
import sys

import pisces_amd as P


def main():
    nbytes = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    L = P.lib()
    L.pisces_copy_probe.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    P.check(L.pisces_copy_probe(0, nbytes, reps, ctypes.byref(ms)), "pisces_copy_probe")
    print(json.dumps({"bytes": nbytes, "launches": reps + 1, "ms_per_launch": ms.value,
                      "GBs_read_plus_write": 2 * nbytes / (ms.value * 1e-3) / 1e9}))


if __name__ == "__main__":
    main()
