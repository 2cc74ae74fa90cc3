"""TSF beam profiles (TSF_GAUSSIAN_Z, TSF_FIBER_Z; src3d/settings.c:4308-4339,
src3d/source.c:231-251 and the face loops :558-742, 1482-1649): every face
correction of the plane wave is multiplied by a lateral profile, the Gaussian
exp(-(x-xc)^2/2rx^2) exp(-(y-yc)^2/2ry^2) or the fiber mode exp(-r^2 dx^2/w0^2)
inside the cutoff, evaluated at the cell whose line value the face uses (E
faces: i0 - 1 / j0 - 1 on the near x / y faces, the face plane elsewhere; H
faces: the face plane).

Pins (oracle): for a wave entering through one face (the other five skipped),
the first step that injects anything is clean, E only for near faces and H
first for far faces; the injection of the other family follows one step later
and is recovered by subtracting the vacuum Yee update of the fields already
there.  Profiled / plain injection at every injected cell must equal the profile
closed form (numpy) at the reference's position; a one-cell shift of that
position or the other family's position fails.  A flat Gaussian (radii 1e6)
equals the plain run.  GPU vs oracle at 1e-5 with both profiles.
"""
import math
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

EPS0 = 8.854187817620389850e-12
MU0 = 1.256637061435917295e-6
N, DX, NSTEPS = 40, 10e-9, 140
B0, B1 = 8, 31
FACES = ["i0", "in", "j0", "jn", "k0", "kn"]
ANG = {"i0": (math.pi / 2, 0.0), "in": (math.pi / 2, math.pi), "j0": (math.pi / 2, math.pi / 2),
       "jn": (math.pi / 2, 3 * math.pi / 2), "k0": (0.0, 0.0), "kn": (math.pi - 0.2, 0.3)}
GAUSS = (18.3, 21.7, 6.5, 9.1)
FIBER = (17.6, 22.4, 9.5, 20.0, 3.5, 2.0)
LAM = 500e-9


def write_par(folder, face, profile, extra=""):
    os.makedirs(folder, exist_ok=True)
    dt = cases._dt(DX, DX, DX, 0.5)
    cases.write_tsf_source(os.path.join(folder, "tsf.txt"), cases.gaussian_sine(NSTEPS, dt, 150e-9, 1.5, delay_widths=1.2))
    th, ph = ANG[face] if face else (0.7, 0.4)
    skips = "".join("TSF_SKIP\n%s\n\n" % f for f in FACES if face and f != face)
    prof = ""
    if profile == "gauss":
        prof = "TSF_GAUSSIAN_Z\n%r %r %r %r\n\n" % GAUSS
    elif profile == "flat":
        prof = "TSF_GAUSSIAN_Z\n18 21 1e6 1e6\n\n"
    elif profile == "fiber":
        prof = "TSF_FIBER_Z\n%r %r %r %r %r %r\n\n" % FIBER
    p = os.path.join(folder, "tsf_profile.par")
    with open(p, "w") as f:
        f.write("POOL\n%d %d %d %g %g %g\n\nCOMP\n%d\n\nDT_MULT\n0.5\n\nSOURCE_WAVELENGTH\n%g %g %g\n\n"
                "SOURCE_TSF\n%d %d %d %d %d %d %r %r 0.3 0 tsf.txt\n\n%s%s%s"
                % (N, N, N, DX, DX, DX, NSTEPS, 0.8 * LAM, LAM, 1.2 * LAM, B0, B0, B0, B1, B1, B1, th, ph, skips, prof, extra))
    return p


def gauss(x, y):
    xc, yc, rx, ry = (float(np.float64(v)) for v in GAUSS)
    return np.exp(-(x - xc) ** 2 / rx / rx / 2.0) * np.exp(-(y - yc) ** 2 / ry / ry / 2.0)


def fiber(x, y):
    xc, yc, radius, cutoff, ec, ecl = FIBER
    v = 2 * math.pi / LAM * radius * DX * (ec - ecl)
    w0 = (0.65 + 1.619 / v / math.sqrt(v) + 2.879 / v ** 6) * radius * DX
    r2 = (x - xc) ** 2 + (y - yc) ** 2
    return np.where(r2 < cutoff * cutoff, np.exp(-r2 * DX * DX / w0 / w0), 0.0)


def curl_h(h):
    """cb-free vacuum Yee E increment from H (the oracle's layout [i][j][k])"""
    hx, hy, hz = h
    cx = np.zeros_like(hx); cy = np.zeros_like(hx); cz = np.zeros_like(hx)
    cx[:, 1:, 1:] = (hz[:, 1:, 1:] - hz[:, :-1, 1:]) - (hy[:, 1:, 1:] - hy[:, 1:, :-1])
    cy[1:, :, 1:] = (hx[1:, :, 1:] - hx[1:, :, :-1]) - (hz[1:, :, 1:] - hz[:-1, :, 1:])
    cz[1:, 1:, :] = (hy[1:, 1:, :] - hy[:-1, 1:, :]) - (hx[1:, 1:, :] - hx[1:, :-1, :])
    return cx, cy, cz


def curl_e(e):
    """db-free vacuum Yee H increment from E"""
    ex, ey, ez = e
    cx = np.zeros_like(ex); cy = np.zeros_like(ex); cz = np.zeros_like(ex)
    cx[:, :-1, :-1] = (ey[:, :-1, 1:] - ey[:, :-1, :-1]) - (ez[:, 1:, :-1] - ez[:, :-1, :-1])
    cy[:-1, :, :-1] = (ez[1:, :, :-1] - ez[:-1, :, :-1]) - (ex[:-1, :, 1:] - ex[:-1, :, :-1])
    cz[:-1, :-1, :] = (ex[:-1, 1:, :] - ex[:-1, :-1, :]) - (ey[1:, :-1, :] - ey[:-1, :-1, :])
    return cx, cy, cz


def injections(par):
    """(E injection, H injection, which family is exact) of the first two steps
    that inject: the family injected first is the field itself, the other one
    is what remains after the vacuum update of the fields of the step before
    (to the float rounding of the oracle's update coefficients)"""
    o = Oracle(par)
    dt = o.dt()
    cb, db = dt / EPS0 / DX, dt / MU0 / DX
    for s in range(NSTEPS - 1):
        o.step(1)
        f = o.fields()
        if any(np.abs(c).max() > 0 for c in f):
            break
    e1, h1 = f[:3], f[3:]
    o.step(1)
    f2 = o.fields()
    o.close()
    if all(np.abs(c).max() == 0 for c in h1):
        # E first: E(s1) is the E injection; H(s1+1) = H update of E(s1) + H injection
        ce = curl_e(e1)
        return list(e1), [f2[3 + c] - db * ce[c] for c in range(3)], "E", max(np.abs(f2[3 + c]).max() for c in range(3))
    # H first: H(s1) is the H injection; E(s1) = E update of H(s1) + E injection
    ch = curl_h(h1)
    return [e1[c] - cb * ch[c] for c in range(3)], list(h1), "H", max(np.abs(e1[c]).max() for c in range(3))


def positions(face, fam):
    """lateral (x, y) index grids of the reference's profile position for each cell"""
    i, j = np.meshgrid(np.arange(N, dtype=np.float64), np.arange(N, dtype=np.float64), indexing="ij")
    x, y = i[:, :, None] + 0 * np.zeros(N), j[:, :, None] + 0 * np.zeros(N)
    if fam == "H":
        # H faces: the face plane; the H cells of a near face lie one cell outside it
        if face == "i0": x = x + 1
        if face == "j0": y = y + 1
    else:
        if face == "i0": x = x - 1
        if face == "j0": y = y - 1
    return x, y


@pytest.mark.parametrize("face", FACES)
@pytest.mark.parametrize("profile", ["gauss", "fiber"])
def test_face_profiles_against_closed_forms(tmp_path, face, profile):
    ea, ha, first, derived_scale = injections(write_par(str(tmp_path / "a"), face, None))
    eb, hb, _, _ = injections(write_par(str(tmp_path / "b"), face, profile))
    fn = gauss if profile == "gauss" else fiber
    checked = 0
    for fam, pa, pb in (("E", ea, eb), ("H", ha, hb)):
        x, y = positions(face, fam)
        want = fn(x, y)
        exact = fam == first
        scale = max(np.abs(c).max() for c in pa) if exact else derived_scale
        for c in range(3):
            m = np.abs(pa[c]) > (1e-9 if exact else 1e-2) * scale
            if not m.any():
                continue
            got = pb[c][m] / pa[c][m]
            err = np.abs(got - want[m]).max()
            # a one-cell shift of the profile position in x or y is far off
            shifted = min(np.abs(got - fn(x + 1, y)[m]).max(), np.abs(got - fn(x, y + 1)[m]).max(),
                          np.abs(got - fn(x - 1, y)[m]).max(), np.abs(got - fn(x, y - 1)[m]).max())
            print(face, profile, fam, c, "cells", int(m.sum()), "max |ratio - profile|", err, "shifted", shifted)
            assert err < (1e-12 if exact else 1e-5)
            assert shifted > 1e-3
            checked += 1
    assert checked >= 3


def test_flat_gaussian_equals_plain(tmp_path):
    a = Oracle(write_par(str(tmp_path / "a"), None, None))
    b = Oracle(write_par(str(tmp_path / "b"), None, "flat"))
    a.step(90); b.step(90)
    fa, fb = a.fields(), b.fields()
    for x, y in zip(fa, fb):
        assert np.linalg.norm(x - y) <= 1e-9 * np.linalg.norm(x)
    assert max(float(np.abs(x).max()) for x in fa) > 0


def test_radial_profile_is_refused(tmp_path):
    with pytest.raises(RuntimeError):
        Oracle(write_par(str(tmp_path / "r"), None, None, "TSF_RADIAL_Z\n18 21 5 5\n\n"))


@pytest.mark.gpu
@pytest.mark.parametrize("profile", ["gauss", "fiber"])
def test_gpu_tsf_profiles_against_oracle(tmp_path, profile):
    from gsvit_copy_amd import Session
    par = write_par(str(tmp_path / profile), None, profile, "BOUNDARY_ALL\ncpml 6 3 -1 0.03 4\n\n")
    print("case tsf profile", profile, flush=True)
    o = Oracle(par)
    o.step()
    ref = o.fields()
    o.close()
    s = Session(par)
    s.run()
    got = s.fields()
    s.close()
    for fam in (slice(0, 3), slice(3, 6)):
        e = float(np.sqrt(sum(np.sum((got[c] - ref[c]) ** 2) for c in range(fam.start, fam.stop))
                          / sum(np.sum(ref[c] ** 2) for c in range(fam.start, fam.stop))))
        print(profile, "GPU vs oracle rel L2", e)
        assert e <= 1e-5
