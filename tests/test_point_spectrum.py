"""OUT_POINT AllFFT (src3d/settings.c:4803-4823, src3d/output.c:517-553 and
607-640): the power spectra of Ex, Ey, Ez at a point over a step window,
written at the last step.  Parity unpinned: the reference takes the
transform from Gwyddion's gwy_data_line_fft (linear leveling, no window,
possibly resampling the trace to a size its FFT prefers), which is not
available here; this build transforms the n samples exactly.  Pinned instead:
the spectrum routine against numpy's FFT of the detrended trace, a pure
sinusoid on a bin landing in that bin with |X|^2 / n = n A^2 / 4 (to the
leveling's leakage), and on the
GPU the written log (wavelength c n dt / k, summed power, amplitudes)
against the engine's own recorded trace.
"""
import ctypes
import os

import numpy as np
import pytest

import cases
from oracle_lib import lib

C0 = 299792458.0


def spectrum(x):
    L = lib()
    L.oracle_point_spectrum.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x) // 2)
    L.oracle_point_spectrum(x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(x),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def numpy_spectrum(x):
    n = len(x)
    t = np.arange(n)
    b, a = np.polyfit(t, x, 1)
    X = np.fft.fft(x - a - b * t)
    return (np.abs(X) ** 2 / n)[: n // 2]


@pytest.mark.parametrize("n", [64, 301, 1000])
def test_spectrum_routine_against_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 0.01 * np.arange(n) + 3.0
    got, want = spectrum(x), numpy_spectrum(x)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * want.max())


def test_sinusoid_on_a_bin():
    """the detrending line fitted to a cosine is not quite zero, so the other
    bins hold a little leakage"""
    n, k, A = 512, 37, 2.5
    x = A * np.cos(2 * np.pi * k * np.arange(n) / n + 0.3)
    s = spectrum(x)
    assert abs(s[k] - n * A * A / 4) <= 1e-3 * s[k]
    s[k] = 0
    assert s.max() <= 1e-4 * n * A * A


@pytest.mark.gpu
def test_gpu_allfft_log(tmp_path):
    from gsvit_copy_amd import Session
    folder = str(tmp_path / "f")
    par = cases.small_case(folder, "spheres_cpml", nsteps=300)
    with open(par, "a") as f:
        f.write("OUT_POINT\nAllFFT 1 30 20 26 20 -1 spec.txt\n\n")
    print("AllFFT log, spheres_cpml 300 steps", flush=True)
    s = Session(par, write_outputs=True)
    s.run()
    tr = s.points()[-1]
    s.close()
    rows = np.loadtxt(os.path.join(folder, "spec.txt"))
    n = 300 - 20
    sp = [numpy_spectrum(tr[20:300, c]) for c in range(3)]
    assert rows.shape == (n // 2 - 1, 5)
    k = np.arange(1, n // 2)
    lam = rows[:, 0]
    dtv = lam * k / (C0 * n)                       # wavelength = c n dt / k
    assert np.allclose(dtv, dtv[0], rtol=1e-5)
    tot = sp[0][1:] + sp[1][1:] + sp[2][1:]
    assert np.allclose(rows[:, 1], tot, rtol=1e-5, atol=1e-6 * tot.max())
    for c in range(3):
        assert np.allclose(rows[:, 2 + c], np.sqrt(sp[c][1:]), rtol=1e-5, atol=1e-6 * np.sqrt(tot.max()))
