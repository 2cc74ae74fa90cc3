"""ctypes binding of the CPU oracle (oracle/_build/liboracle_{f64,f32}.so).

Test infrastructure only: the oracle is the checker for the MI355X engine and
is never used by the product path.
"""
import ctypes
import os
import contextlib

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(REPO, "oracle", "_build")

_libs = {}


def lib(precision="f64"):
    if precision not in _libs:
        path = os.path.join(BUILD, "liboracle_%s.so" % precision)
        if not os.path.exists(path):
            raise FileNotFoundError("oracle library %s is not built (run __graft_entry__.build())" % path)
        L = ctypes.CDLL(path)
        L.oracle_open.restype = ctypes.c_void_p
        L.oracle_open.argtypes = [ctypes.c_char_p]
        L.oracle_close.argtypes = [ctypes.c_void_p]
        L.oracle_step.restype = ctypes.c_int
        L.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_get_field.restype = ctypes.c_int
        L.oracle_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_points.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_farfield.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_pfarfield_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_get_forces.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_ltsf_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_get_ltsf.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_pfarfield.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_maps.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_get_dft.restype = ctypes.c_long
        L.oracle_get_dft.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_sums.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_flux.restype = ctypes.c_long
        L.oracle_get_flux.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_objective.restype = ctypes.c_double
        L.oracle_objective.argtypes = [ctypes.c_void_p]
        L.oracle_farfield_dft.restype = ctypes.c_int
        L.oracle_farfield_dft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double)]
        L.oracle_dft_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_cpml.restype = ctypes.c_int
        L.oracle_cpml.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_get_material.restype = ctypes.c_int
        L.oracle_get_material.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_dt.restype = ctypes.c_double
        L.oracle_dt.argtypes = [ctypes.c_void_p]
        L.oracle_real_size.restype = ctypes.c_int
        L.oracle_rc_scalar.restype = ctypes.c_int
        L.oracle_rc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.oracle_set_option.restype = ctypes.c_int
        L.oracle_tsff_waves.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oracle_tsff_waves.restype = ctypes.c_int
        L.oracle_farfield_points.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.oracle_farfield_points.restype = ctypes.c_int
        L.oracle_set_option.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        _libs[precision] = L
    return _libs[precision]


@contextlib.contextmanager
def chdir(path):
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One CPU run of a parameter file; relative paths resolve against the file's folder."""

    def __init__(self, parfile, precision="f64"):
        self.L = lib(precision)
        self.dir = os.path.dirname(os.path.abspath(parfile))
        with chdir(self.dir):
            self.h = self.L.oracle_open(os.path.basename(parfile).encode())
        if not self.h:
            raise RuntimeError("oracle could not open %s" % parfile)

    def set_option(self, name, value):
        if self.L.oracle_set_option(self.h, name.encode(), int(value)):
            raise ValueError("unknown oracle option %s" % name)

    def info(self):
        v = (ctypes.c_int * 9)()
        self.L.oracle_info(self.h, v)
        keys = ["nx", "ny", "nz", "nsteps", "npts", "nrs", "ndata", "nmat", "step"]
        return dict(zip(keys, list(v)))

    def step(self, n=None):
        inf = self.info()
        if n is None:
            n = inf["nsteps"] - inf["step"]
        with chdir(self.dir):
            return self.L.oracle_step(self.h, n)

    def field(self, comp):
        inf = self.info()
        out = np.empty((inf["nx"], inf["ny"], inf["nz"]), dtype=np.float64)
        self.L.oracle_get_field(self.h, comp, _dptr(out))
        return out

    def fields(self):
        return np.stack([self.field(c) for c in range(6)])

    def points(self):
        inf = self.info()
        out = np.zeros((max(inf["npts"], 1), inf["nsteps"] + 1, 6), dtype=np.float64)
        if inf["npts"]:
            self.L.oracle_get_points(self.h, _dptr(out))
        return out[: inf["npts"]]

    def farfield(self):
        inf = self.info()
        out = np.zeros((max(inf["nrs"], 1), 3, inf["ndata"]), dtype=np.float64)
        if inf["nrs"]:
            self.L.oracle_get_farfield(self.h, _dptr(out))
        return out[: inf["nrs"]]

    def tsff_waves(self):
        """focused-TSF plane waves [nw][5] = theta, phi, psi, corr, acorr"""
        n = self.L.oracle_tsff_waves(self.h, None)
        out = np.zeros((max(n, 1), 5), dtype=np.float64)
        self.L.oracle_tsff_waves(self.h, _dptr(out))
        return out[:n]

    def farfield_points(self, periodic=False):
        """planned far-field points [nrs][4] = ri, rj, rk, istart"""
        n = self.L.oracle_farfield_points(self.h, int(periodic), None)
        out = np.zeros((max(n, 1), 4), dtype=np.int32)
        self.L.oracle_farfield_points(self.h, int(periodic), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out[:n]

    def forces(self, nforces):
        """OUT_FORCE values [nforces][nsteps + 1][3] (without the log's dx dy)"""
        inf = self.info()
        out = np.zeros((max(nforces, 1), inf["nsteps"] + 1, 3), dtype=np.float64)
        if nforces:
            self.L.oracle_get_forces(self.h, _dptr(out))
        return out[:nforces]

    def ltsf_history(self):
        """layered-TSF 1-D model history [6: ee eh ez he hh hz][substep][z] (after the first step)"""
        d = (ctypes.c_int * 2)()
        self.L.oracle_ltsf_dims(self.h, d)
        out = np.zeros((6, max(d[0], 1), d[1]), dtype=np.float64)
        if d[0]:
            self.L.oracle_get_ltsf(self.h, _dptr(out))
        return out[:, : d[0]]

    def pfarfield(self):
        """periodic far-field bins [nrs][3][ndata] (Ex, Ey, Ez, not negated)"""
        d = (ctypes.c_int * 2)()
        self.L.oracle_pfarfield_dims(self.h, d)
        out = np.zeros((max(d[0], 1), 3, d[1]), dtype=np.float64)
        if d[0]:
            self.L.oracle_get_pfarfield(self.h, _dptr(out))
        return out[: d[0]]

    def sums(self, nsums):
        inf = self.info()
        out = np.zeros((max(nsums, 1), inf["nsteps"]), dtype=np.float64)
        if nsums:
            self.L.oracle_get_sums(self.h, _dptr(out))
        return out[:nsums]

    def objective(self):
        return self.L.oracle_objective(self.h)

    def dft(self):
        n = self.L.oracle_get_dft(self.h, None)
        out = np.zeros(max(n, 1), dtype=np.float64)
        if n:
            self.L.oracle_get_dft(self.h, _dptr(out))
        return out[:n]

    def cpml(self, face):
        """the planner's CPML vectors of face 0..5 (x0, xn, y0, yn, z0, zn):
        dict of be, ce, bh, ch, kappae, kappah (None without CPML)"""
        d = self.L.oracle_cpml(self.h, face, None)
        if d == 0:
            return None
        out = np.zeros((6, d))
        self.L.oracle_cpml(self.h, face, _dptr(out))
        return dict(zip(("be", "ce", "bh", "ch", "kappae", "kappah"), out))

    def dt(self):
        return self.L.oracle_dt(self.h)

    def dft_dims(self):
        """(face cells, wavelengths) of the running DFT"""
        d = (ctypes.c_int * 2)()
        self.L.oracle_dft_dims(self.h, d)
        return d[0], d[1]

    def farfield_dft(self, theta, phi):
        """far-field pattern of the running DFT at directions (theta, phi):
        complex [ndirs, nwavelengths, 2] (F_theta, F_phi)"""
        th = np.asarray(theta, dtype=np.float64).ravel()
        dirs = np.ascontiguousarray(np.stack([th, np.broadcast_to(np.asarray(phi, dtype=np.float64).ravel(), th.shape)], 1))
        out = np.zeros((th.size, max(self.dft_dims()[1], 1), 4), dtype=np.float64)
        if self.L.oracle_farfield_dft(self.h, th.size, _dptr(dirs), _dptr(out)):
            raise RuntimeError("no NFFF_DFT accumulators")
        return out[..., 0::2] + 1j * out[..., 1::2]

    def flux_acc(self):
        """FLUX_DFT accumulators, flat [cell][wavelength][8]."""
        n = self.L.oracle_get_flux(self.h, None)
        out = np.zeros(max(n, 1), dtype=np.float64)
        if n:
            self.L.oracle_get_flux(self.h, _dptr(out))
        return out[:n]

    def materials(self):
        """the planner's material table: dicts of type, epsilon, mu, sigma,
        sigast, omega_p, nu, poles, cp_a0, cp_omega0"""
        keys = ("type", "epsilon", "mu", "sigma", "sigast", "omega_p", "nu", "poles", "cp_a0", "cp_omega0")
        n = self.L.oracle_get_material(self.h, -1, None)
        out = []
        for m in range(n):
            v = np.zeros(10)
            self.L.oracle_get_material(self.h, m, _dptr(v))
            d = dict(zip(keys, v.tolist()))
            d["type"] = int(d["type"])
            d["poles"] = int(d["poles"])
            out.append(d)
        return out

    def rc_scalar(self, matindex, curl, d):
        """one component of recursive-convolution material `matindex` stepped
        from zero state by the oracle's own update, driven by the curl
        bracket curl[t] over spacing d; returns E^1 .. E^n"""
        c = np.ascontiguousarray(curl, dtype=np.float64)
        out = np.zeros(c.size)
        if self.L.oracle_rc_scalar(self.h, matindex, c.size, _dptr(c), d, _dptr(out)):
            raise ValueError("material %d is not a recursive-convolution material" % matindex)
        return out

    def maps(self):
        inf = self.info()
        mat = np.empty((inf["nx"], inf["ny"], inf["nz"]), dtype=np.int32)
        bnd = np.empty_like(mat)
        self.L.oracle_get_maps(self.h, mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               bnd.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return mat, bnd

    def close(self):
        if self.h:
            self.L.oracle_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
