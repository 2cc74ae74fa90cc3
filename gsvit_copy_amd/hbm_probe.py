"""Streaming-copy probe: attainable HBM bandwidth and counter calibration.

`python -m gsvit_copy_amd.hbm_probe [MiB] [reps]` copies a buffer of the
given size (read + write) with a plain HIP kernel and prints GB/s; run under
`rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` it gives the counter reading of a
known traffic (2 x size per copy) to scale the engine's kernel counters.
"""
import ctypes
import json
import sys

from .engine import library


def probe(mib=1024, reps=20):
    L = library()
    L.gsv_copy_probe.restype = ctypes.c_int
    L.gsv_copy_probe.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    nbytes = int(mib) << 20
    if L.gsv_copy_probe(nbytes, int(reps), ctypes.byref(ms)):
        raise RuntimeError("copy probe failed")
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms.value, "GB_s": 2 * nbytes / (ms.value * 1e-3) / 1e9}


if __name__ == "__main__":
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 20
    print(json.dumps(probe(mib, reps)))
