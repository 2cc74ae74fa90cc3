"""gsv_farfield_dft_noalloc: the far-field transform of the running DFT with
its partial sums in a fixed per-GPU scratch array, directions and result in
the caller's device memory.  Checked through the C-ABI against
gsv_farfield_dft (which allocates its buffers) on two calls in a row, the
second with fewer directions, and for the size limit (a call that needs more
than the scratch array is refused before any launch).  The caller's device
buffers come from hipMalloc looked up through the engine library's own
handle, i.e. the HIP runtime the engine itself is bound to (loading the
runtime again by another name could bring a second runtime into the
process)."""
import ctypes

import numpy as np
import pytest

from test_farfield_dft import dipole_case, directions

pytestmark = pytest.mark.gpu


def test_noalloc_equals_allocating_transform(tmp_path):
    from gsvit_copy_amd import Session, library
    L = library()
    L.gsv_farfield_dft_noalloc.restype = ctypes.c_int
    L.gsv_farfield_dft_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    H2D, D2H = 1, 2
    hip = {}
    for name, args in (("hipMalloc", [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]),
                       ("hipFree", [ctypes.c_void_p]),
                       ("hipMemcpy", [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int])):
        f = getattr(L, name)                 # dlsym on the engine's handle: its own runtime
        f.restype = ctypes.c_int
        f.argtypes = args
        hip[name] = f
    par = dipole_case(str(tmp_path / "d"), 12, nsteps=600)
    print("no-alloc far field: dipole 56^3, 600 steps", flush=True)
    s = Session(par)
    s.run()
    nw = s.dft_dims()[1]
    th, ph = directions()
    bufs = []
    try:
        for nd in (len(th), 20):                       # two calls, the second smaller
            ref = s.farfield_dft(th[:nd], ph[:nd])
            dirs = np.ascontiguousarray(np.stack([th[:nd], ph[:nd]], 1))
            d_dirs, d_out = ctypes.c_void_p(), ctypes.c_void_p()
            assert hip["hipMalloc"](ctypes.byref(d_dirs), dirs.nbytes) == 0
            bufs.append(d_dirs)
            assert hip["hipMalloc"](ctypes.byref(d_out), nd * nw * 4 * 8) == 0
            bufs.append(d_out)
            assert hip["hipMemcpy"](d_dirs, dirs.ctypes.data, dirs.nbytes, H2D) == 0
            assert L.gsv_farfield_dft_noalloc(s.h, nd, d_dirs, d_out) == 0
            out = np.zeros((nd, nw, 4))
            assert hip["hipMemcpy"](out.ctypes.data, d_out, out.nbytes, D2H) == 0
            got = out[..., 0::2] + 1j * out[..., 1::2]
            e = np.abs(got - ref).max() / np.abs(ref).max()
            print("%d directions: no-alloc vs allocating transform %.3e" % (nd, e))
            assert e <= 1e-12
        # 174763 directions x 2 wavelengths need more than the scratch array: refused
        assert L.gsv_farfield_dft_noalloc(s.h, 174763, bufs[0], bufs[1]) == 1
    finally:
        for b in bufs:
            hip["hipFree"](b)
        s.close()
