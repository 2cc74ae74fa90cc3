"""Python binding of the MI355X engine's C-ABI session (include/gsvit_session.h).

The compute path is the HIP library gsvit_copy_amd/lib/libgsvit_amd.so; there
is no CPU fallback: if the library is missing or no GPU is present, opening a
session raises.
"""
import contextlib
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libgsvit_amd.so")
_lib = None


def library():
    """Loads libgsvit_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgsvit_amd.so is not built; run gsvit_copy_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp, fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
        L.gsv_open.restype = vp
        L.gsv_open.argtypes = [ctypes.c_char_p]
        L.gsv_run.restype = ctypes.c_int
        L.gsv_run.argtypes = [vp, ctypes.c_int]
        L.gsv_set_write_outputs.argtypes = [vp, ctypes.c_int]
        L.gsv_sync.restype = ctypes.c_int
        L.gsv_sync.argtypes = [vp]
        L.gsv_info.argtypes = [vp, ip]
        L.gsv_get_field.restype = ctypes.c_int
        L.gsv_get_field.argtypes = [vp, ctypes.c_int, fp]
        L.gsv_get_points.restype = ctypes.c_int
        L.gsv_get_points.argtypes = [vp, dp]
        L.gsv_get_farfield.restype = ctypes.c_int
        L.gsv_get_farfield.argtypes = [vp, dp]
        L.gsv_get_forces.restype = ctypes.c_long
        L.gsv_get_forces.argtypes = [vp, dp]
        L.gsv_pfarfield_dims.restype = None
        L.gsv_pfarfield_dims.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.gsv_get_pfarfield.restype = ctypes.c_int
        L.gsv_get_pfarfield.argtypes = [vp, dp]
        L.gsv_get_dft.restype = ctypes.c_long
        L.gsv_get_dft.argtypes = [vp, dp]
        L.gsv_get_maps.argtypes = [vp, ip, ip]
        L.gsv_dft_dims.restype = None
        L.gsv_dft_dims.argtypes = [vp, ip]
        L.gsv_farfield_dft.restype = ctypes.c_int
        L.gsv_farfield_dft.argtypes = [vp, ctypes.c_int, dp, dp]
        L.gsv_get_flux.restype = ctypes.c_long
        L.gsv_get_flux.argtypes = [vp, dp, dp]
        L.gsv_flux_dims.argtypes = [vp, ip]
        L.gsv_get_sums.restype = ctypes.c_long
        L.gsv_get_sums.argtypes = [vp, dp]
        L.gsv_set_timing.argtypes = [vp, ctypes.c_int]
        L.gsv_kernel_times.argtypes = [vp, dp]
        L.gsv_close.argtypes = [vp]
        L.gsv_set_launch.restype = ctypes.c_int
        L.gsv_set_launch.argtypes = [vp, ip]
        L.gsv_get_launch.argtypes = [vp, ip]
        L.gsv_autotune.restype = ctypes.c_int
        L.gsv_autotune.argtypes = [vp, ctypes.c_int, ip]
        L.gsv_design_dims.argtypes = [vp, ip]
        L.gsv_adjoint_setup.restype = ctypes.c_int
        L.gsv_adjoint_setup.argtypes = [vp, ctypes.c_longlong, ctypes.c_int]
        L.gsv_set_design_eps.restype = ctypes.c_int
        L.gsv_set_design_eps.argtypes = [vp, fp]
        L.gsv_get_design_eps.restype = ctypes.c_int
        L.gsv_get_design_eps.argtypes = [vp, fp]
        L.gsv_objective.restype = ctypes.c_int
        L.gsv_objective.argtypes = [vp, dp]
        L.gsv_objective_gradient.restype = ctypes.c_int
        L.gsv_objective_gradient.argtypes = [vp, fp]
        L.gsv_adjoint_info.argtypes = [vp, dp]
        L.gsv_batch_open.restype = vp
        L.gsv_batch_open.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.gsv_batch_run.restype = ctypes.c_int
        L.gsv_batch_run.argtypes = [vp, ctypes.c_int]
        L.gsv_batch_set_launch.restype = ctypes.c_int
        L.gsv_batch_set_launch.argtypes = [vp, ip]
        L.gsv_batch_get_launch.argtypes = [vp, ip]
        L.gsv_batch_set_graph.restype = ctypes.c_int
        L.gsv_batch_set_graph.argtypes = [vp, ctypes.c_int]
        L.gsv_batch_autotune.restype = ctypes.c_int
        L.gsv_batch_autotune.argtypes = [vp, ctypes.c_int, ip]
        L.gsv_batch_set_timing.argtypes = [vp, ctypes.c_int]
        L.gsv_batch_info.argtypes = [vp, dp]
        L.gsv_batch_close.argtypes = [vp]
        _lib = L
    return _lib


@contextlib.contextmanager
def _chdir(path):
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


INFO_KEYS = ["nx", "ny", "nz", "nsteps", "npts", "nrs", "ndata", "nmat", "step",
             "e_classes", "e_id_bytes", "h_classes", "h_id_bytes", "ade", "ade_poles",
             "threads_per_block", "z_chunk"]


class Session:
    """One GPU run of a GSvit parameter file (relative paths resolve against its folder)."""

    def __init__(self, parfile, write_outputs=False):
        self.L = library()
        self.dir = os.path.dirname(os.path.abspath(parfile))
        with _chdir(self.dir):
            self.h = self.L.gsv_open(os.path.basename(parfile).encode())
        if not self.h:
            raise RuntimeError("the MI355X engine could not open %s (see stderr)" % parfile)
        self.L.gsv_set_write_outputs(self.h, 1 if write_outputs else 0)

    def info(self):
        v = (ctypes.c_int * 17)()
        self.L.gsv_info(self.h, v)
        return dict(zip(INFO_KEYS, list(v)))

    def run(self, nsteps=None):
        inf = self.info()
        if nsteps is None:
            nsteps = inf["nsteps"] - inf["step"]
        with _chdir(self.dir):
            r = self.L.gsv_run(self.h, int(nsteps))
        if r < 0:
            raise RuntimeError("GPU step failed (see stderr)")
        return r

    def sync(self):
        if self.L.gsv_sync(self.h):
            raise RuntimeError("GPU synchronisation failed")

    def field(self, comp):
        inf = self.info()
        out = np.empty((inf["nx"], inf["ny"], inf["nz"]), dtype=np.float32)
        if self.L.gsv_get_field(self.h, int(comp), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
            raise RuntimeError("field download failed")
        return out

    def fields(self):
        return np.stack([self.field(c) for c in range(6)])

    def points(self):
        inf = self.info()
        out = np.zeros((max(inf["npts"], 1), inf["nsteps"] + 1, 6), dtype=np.float64)
        if inf["npts"] and self.L.gsv_get_points(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise RuntimeError("trace download failed")
        return out[: inf["npts"]]

    def farfield(self):
        inf = self.info()
        out = np.zeros((max(inf["nrs"], 1), 3, inf["ndata"]), dtype=np.float64)
        if inf["nrs"] and self.L.gsv_get_farfield(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise RuntimeError("far-field download failed")
        return out[: inf["nrs"]]

    def forces(self):
        """OUT_FORCE values [nforces][nsteps + 1][3] (without the log's dx dy),
        one row per force box of the parameter file"""
        inf = self.info()
        n = self.L.gsv_get_forces(self.h, None)
        out = np.zeros(max(n, 1), dtype=np.float64)
        if n > 0 and self.L.gsv_get_forces(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) < 0:
            raise RuntimeError("force download failed")
        return out[:n].reshape(-1, inf["nsteps"] + 1, 3)

    def pfarfield(self):
        """periodic far-field bins [nrs][3][ndata] (Ex, Ey, Ez, not negated)"""
        d = (ctypes.c_int * 2)()
        self.L.gsv_pfarfield_dims(self.h, d)
        out = np.zeros((max(d[0], 1), 3, d[1]), dtype=np.float64)
        if d[0] and self.L.gsv_get_pfarfield(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise RuntimeError("periodic far-field download failed")
        return out[: d[0]]

    def dft(self):
        n = self.L.gsv_get_dft(self.h, None)
        out = np.zeros(max(n, 1), dtype=np.float64)
        if n > 0 and self.L.gsv_get_dft(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) < 0:
            raise RuntimeError("DFT download failed")
        return out[:n]

    def dft_dims(self):
        """(face cells, wavelengths) of the running DFT"""
        d = (ctypes.c_int * 2)()
        self.L.gsv_dft_dims(self.h, d)
        return d[0], d[1]

    def farfield_dft(self, theta, phi):
        """far-field pattern of the running DFT at directions (theta, phi) in
        radians: complex [ndirs, wavelengths, 2] = F_theta, F_phi with
        E = F exp(-i k r) / r"""
        th = np.asarray(theta, dtype=np.float64).ravel()
        dirs = np.ascontiguousarray(np.stack([th, np.broadcast_to(np.asarray(phi, dtype=np.float64).ravel(), th.shape)], 1))
        nw = self.dft_dims()[1]
        out = np.zeros((th.size, max(nw, 1), 4), dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        if self.L.gsv_farfield_dft(self.h, th.size, dirs.ctypes.data_as(dp), out.ctypes.data_as(dp)):
            raise RuntimeError("far-field transform failed (see stderr)")
        return out[..., 0::2] + 1j * out[..., 1::2]

    def flux(self, with_acc=False):
        """FLUX_DFT spectra [nplanes, nwavelengths] (with with_acc also the raw
        accumulators [nplanes, cells per plane, nwavelengths, 8])."""
        d = (ctypes.c_int * 3)()
        self.L.gsv_flux_dims(self.h, d)
        npl, nw, per = d[0], d[1], d[2]
        if npl == 0:
            raise RuntimeError("the session has no FLUX_DFT planes")
        dp = ctypes.POINTER(ctypes.c_double)
        fl = np.zeros((npl, nw), dtype=np.float64)
        acc = np.zeros((npl, per, nw, 8), dtype=np.float64) if with_acc else None
        if self.L.gsv_get_flux(self.h, fl.ctypes.data_as(dp), acc.ctypes.data_as(dp) if with_acc else None) < 0:
            raise RuntimeError("flux download failed")
        return (fl, acc) if with_acc else fl

    def sums(self):
        """OUT_SUM values per box and step, [nsums, nsteps]."""
        inf = self.info()
        n = self.L.gsv_get_sums(self.h, None)
        out = np.zeros(max(n, 1), dtype=np.float64)
        if n > 0 and self.L.gsv_get_sums(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) < 0:
            raise RuntimeError("sum download failed")
        return out[:n].reshape(-1, inf["nsteps"]) if n > 0 else np.zeros((0, inf["nsteps"]))

    def maps(self):
        inf = self.info()
        mat = np.empty((inf["nx"], inf["ny"], inf["nz"]), dtype=np.int32)
        bnd = np.empty_like(mat)
        ptr = ctypes.POINTER(ctypes.c_int)
        self.L.gsv_get_maps(self.h, mat.ctypes.data_as(ptr), bnd.ctypes.data_as(ptr))
        return mat, bnd

    def set_timing(self, on):
        self.L.gsv_set_timing(self.h, 1 if on else 0)

    def kernel_times(self):
        v = (ctypes.c_double * 4)()
        self.L.gsv_kernel_times(self.h, v)
        return {"h_ms": v[0], "h_n": int(v[1]), "e_ms": v[2], "e_n": int(v[3])}

    def set_launch(self, shape):
        v = (ctypes.c_int * 6)(*[int(x) for x in shape])
        if self.L.gsv_set_launch(self.h, v):
            raise ValueError("invalid launch shape %r" % (shape,))

    def get_launch(self):
        v = (ctypes.c_int * 6)()
        self.L.gsv_get_launch(self.h, v)
        return list(v)

    def autotune(self, steps_per_candidate=2):
        """Times candidate launch shapes on real steps; returns (shape, steps consumed)."""
        v = (ctypes.c_int * 6)()
        with _chdir(self.dir):
            used = self.L.gsv_autotune(self.h, int(steps_per_candidate), v)
        if used < 0:
            raise RuntimeError("autotune failed")
        return list(v), used

    # ---- gradient extension (DESIGN_BOX / OBJECTIVE_DFT parameter files) ----

    def design_dims(self):
        v = (ctypes.c_int * 3)()
        self.L.gsv_design_dims(self.h, v)
        return tuple(v)

    def adjoint_setup(self, history_budget_bytes=16 << 30, ckpt_interval=100):
        if self.L.gsv_adjoint_setup(self.h, int(history_budget_bytes), int(ckpt_interval)):
            raise RuntimeError("gradient setup failed (see stderr)")

    def set_design_eps(self, eps_r):
        a = np.ascontiguousarray(eps_r, dtype=np.float32)
        if a.shape != self.design_dims():
            raise ValueError("design array shape %r, expected %r" % (a.shape, self.design_dims()))
        if self.L.gsv_set_design_eps(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
            raise RuntimeError("setting the design permittivity failed")

    def design_eps(self):
        out = np.empty(self.design_dims(), dtype=np.float32)
        if self.L.gsv_get_design_eps(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))):
            raise RuntimeError("no design region")
        return out

    def objective(self):
        """Fresh forward run over all steps; returns J."""
        J = ctypes.c_double()
        with _chdir(self.dir):
            r = self.L.gsv_objective(self.h, ctypes.byref(J))
        if r:
            raise RuntimeError("objective run failed (see stderr)")
        return J.value

    def objective_gradient(self):
        """dJ/d eps_r per design voxel, after objective()."""
        out = np.empty(self.design_dims(), dtype=np.float32)
        with _chdir(self.dir):
            r = self.L.gsv_objective_gradient(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        if r:
            raise RuntimeError("adjoint run failed (see stderr)")
        return out

    def adjoint_info(self):
        v = (ctypes.c_double * 5)()
        self.L.gsv_adjoint_info(self.h, v)
        return {"forward_ms": v[0], "backward_ms": v[1], "history_all": bool(v[2]), "ckpt_interval": int(v[3]),
                "checkpoints": int(v[4])}

    def close(self):
        if getattr(self, "h", None):
            self.L.gsv_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_INFO_KEYS = ["h_ms", "h_launches", "e_ms", "e_launches", "h_tiles", "e_tiles", "face_tiles", "tsf_runs"]


class Batch:
    """Independent sessions stepped together, one launch per kernel type for
    all of them (gsv_batch_* in include/gsvit_session.h).  The sessions stay
    open while the batch lives; read results with Session.field()."""

    def __init__(self, sessions):
        self.L = library()
        self.sessions = list(sessions)
        arr = (ctypes.c_void_p * len(self.sessions))(*[s.h for s in self.sessions])
        self.h = self.L.gsv_batch_open(arr, len(self.sessions))
        if not self.h:
            raise RuntimeError("the sessions could not be batched (see stderr)")

    def run(self, nsteps):
        r = self.L.gsv_batch_run(self.h, int(nsteps))
        if r < 0:
            raise RuntimeError("batched step failed (see stderr)")
        return r

    def set_launch(self, shape):
        v = (ctypes.c_int * 8)(*[int(x) for x in shape])
        if self.L.gsv_batch_set_launch(self.h, v):
            raise ValueError("invalid batch launch shape %r" % (list(shape),))

    def get_launch(self):
        v = (ctypes.c_int * 8)()
        self.L.gsv_batch_get_launch(self.h, v)
        return list(v)

    def set_graph(self, steps_per_graph):
        if self.L.gsv_batch_set_graph(self.h, int(steps_per_graph)):
            raise ValueError("invalid steps per graph")

    def autotune(self, steps_per_candidate=2):
        v = (ctypes.c_int * 8)()
        r = self.L.gsv_batch_autotune(self.h, int(steps_per_candidate), v)
        if r < 0:
            raise RuntimeError("batch autotune failed (see stderr)")
        return r, list(v)

    def set_timing(self, on):
        self.L.gsv_batch_set_timing(self.h, 1 if on else 0)

    def info(self):
        v = (ctypes.c_double * 8)()
        self.L.gsv_batch_info(self.h, v)
        return dict(zip(BATCH_INFO_KEYS, list(v)))

    def close(self):
        if getattr(self, "h", None):
            self.L.gsv_batch_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
