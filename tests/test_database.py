"""Optical-database materials (material type 99 in MEDIUM_VECTOR files,
src3d/plan.c:719-1312, resolved at the source's centre wavelength).

Pinned by what the formulas fix, on entries written here (the reference ships
no spectra database):
  * Sellmeier fused silica (Malitson's coefficients) gives the textbook
    n_d = 1.4585 at 587.6 nm, Cauchy and Kasarova their closed forms;
  * an n-k table (micrometre rows ascending, nanometre rows descending) gives
    the linear interpolation of n and k at the centre wavelength, eps = n^2 - k^2
    and sigma = 4 pi n k c eps0 / lambda; a lossless row pair gives sigma 0;
  * every number passes through "%g" as in the reference's rewritten
    material file, so the planner sees the rounded values: a file with a
    database entry rasterises like the same file pre-rounded by hand with the
    resolved material written directly;
  * an "ade" entry equals the same ADE material written directly (one table
    row after de-duplication), and a PLRC line in a file with database
    entries is rewritten as ADE (the reference's adjust_material does so);
  * entries resolve relative to the parameter file's folder or under
    $GSVIT_SPECTRA_DIR; a missing entry, a missing SOURCE_WAVELENGTH and a
    centre wavelength above an n-k table are errors.
On the GPU a sphere of a Sellmeier glass is checked against the oracle like
any other linear material.
"""
import math
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

C0 = 299792458.0
EPS0 = 8.854187817620389850e-12
MU0 = 1.256637061435917295e-6
F32 = np.float32

SILICA = "sellmeier 0.21 6.7\nfused silica, Malitson 1965\n1 0.6961663 0.004679148 0.4079426 0.01351206 0.8974794 97.934\n"


def g6(v):
    """a number as the reference's "%g" writes it and reads it back"""
    return float("%g" % v)


def write(folder, name, text):
    os.makedirs(folder, exist_ok=True)
    path = os.path.join(folder, name)
    with open(path, "w") as f:
        f.write(text)
    return path


def par_file(folder, vector_text, lam=(400e-9, 587.6e-9, 800e-9), n=24, extra=""):
    write(folder, "vec.txt", vector_text)
    wl = "SOURCE_WAVELENGTH\n%g %g %g\n\n" % lam if lam else ""
    return write(folder, "db.par", "POOL\n%d %d %d 1e-8 1e-8 1e-8\n\nCOMP\n2\n\n%sMEDIUM_VECTOR\nvec.txt\n\n%s"
                 % (n, n, n, wl, extra))


def sellmeier_n(p, lam):
    w = lam * 1e6
    p = [float(F32(v)) for v in p]
    return math.sqrt(p[0] + p[1] * w * w / (w * w - p[2]) + p[3] * w * w / (w * w - p[4])
                     + p[5] * w * w / (w * w - p[6]))


def test_sellmeier_fused_silica_textbook_index(tmp_path):
    d = str(tmp_path)
    write(d, "silica.txt", SILICA)
    o = Oracle(par_file(d, "4 12 12 12 6 99 silica.txt\n"))
    mats = o.materials()
    assert len(mats) == 2 and mats[1]["type"] == 1
    n = math.sqrt(mats[1]["epsilon"] / EPS0)        # tabulated materials hold absolute constants
    print("fused silica at 587.6 nm: n = %.6f" % n)
    assert abs(n - 1.4585) < 2e-4                      # n_d of fused silica
    p = [1, 0.6961663, 0.004679148, 0.4079426, 0.01351206, 0.8974794, 97.934]
    assert mats[1]["epsilon"] == float(F32(g6(sellmeier_n(p, 587.6e-9) ** 2))) * EPS0
    assert mats[1]["mu"] == MU0 and mats[1]["sigma"] == 0


def test_cauchy_and_kasarova_closed_forms(tmp_path):
    d = str(tmp_path)
    write(d, "cauchy.txt", "cauchy 0.3 2.5\nsome glass\n1.5046 0.0042 0.00011\n")
    write(d, "kas.txt", "kasarova 0.4 1.6\nsome polymer\n2.4 -0.0002 0.012 0.0003 -1.2e-5 4e-7\n")
    lam = 633e-9
    o = Oracle(par_file(d, "8 2 2 2 10 10 10 99 cauchy.txt\n8 12 12 12 20 20 20 99 kas.txt\n",
                        lam=(500e-9, lam, 700e-9)))
    mats = o.materials()
    w = lam * 1e6
    a = [float(F32(v)) for v in (1.5046, 0.0042, 0.00011)]
    n_c = a[0] + a[1] / w ** 2 + a[2] / w ** 4
    k = [float(F32(v)) for v in (2.4, -0.0002, 0.012, 0.0003, -1.2e-5, 4e-7)]
    n_k = math.sqrt(k[0] + k[1] * w ** 2 + k[2] / w ** 2 + k[3] / w ** 4 + k[4] / w ** 6 + k[5] / w ** 8)
    assert [m["type"] for m in mats] == [1, 1, 1]
    assert mats[1]["epsilon"] == float(F32(g6(n_c ** 2))) * EPS0
    assert mats[2]["epsilon"] == float(F32(g6(n_k ** 2))) * EPS0


def nk_rows(unit_nm, descending):
    lam_um = np.array([0.4, 0.5, 0.6, 0.7, 0.8])
    n = np.array([1.6, 1.55, 1.52, 1.5, 1.49])
    k = np.array([0.30, 0.20, 0.12, 0.0, 0.0])
    order = slice(None, None, -1) if descending else slice(None)
    w = lam_um * (1000 if unit_nm else 1)
    rows = "".join("%g %g %g\n" % r for r in zip(w[order], n[order], k[order]))
    return "nk 0.4 0.8\ntabulated film\n%d\n%s" % (len(w), rows), lam_um, n, k


@pytest.mark.parametrize("unit_nm,descending", [(False, False), (True, True)])
def test_nk_table_interpolation_and_conductivity(tmp_path, unit_nm, descending):
    d = str(tmp_path)
    text, lam_um, n, k = nk_rows(unit_nm, descending)
    write(d, "film.txt", text)
    lam = 555e-9
    o = Oracle(par_file(d, "8 4 4 4 20 20 20 99 film.txt\n", lam=(450e-9, lam, 650e-9)))
    m = o.materials()[1]
    ni = np.interp(lam * 1e6, lam_um, n)
    ki = np.interp(lam * 1e6, lam_um, k)
    assert m["type"] == 1
    assert m["epsilon"] / EPS0 == pytest.approx(g6(ni * ni - ki * ki), rel=1e-6)
    assert m["sigma"] == pytest.approx(g6(4 * math.pi * ni * ki * C0 / lam * EPS0), rel=1e-6)
    # a lossless stretch of the table (k = 0 at both ends of the bracket) gives no conductivity
    o2 = Oracle(par_file(str(tmp_path / "b"), "8 4 4 4 20 20 20 99 ../film.txt\n", lam=(700e-9, 750e-9, 800e-9)))
    m2 = o2.materials()[1]
    assert m2["sigma"] == 0 and m2["epsilon"] / EPS0 == pytest.approx(g6(1.495 ** 2), rel=1e-6)


def test_numbers_pass_through_g_format(tmp_path):
    """the whole vector file is rewritten with "%g" once it holds a database
    entry: a sphere of radius 4.9999996 (float 4.99999952) becomes radius 5,
    whose surface passes through 30 voxel centres (integer offsets with
    a^2 + b^2 + c^2 = 25); the raster equals that of
    the file pre-rounded by hand, and differs from that of the same numbers
    read without a database entry"""
    d = str(tmp_path)
    write(d, "silica.txt", SILICA)
    o0 = Oracle(par_file(str(tmp_path / "probe"), "4 12 12 12 6 99 ../silica.txt\n"))
    eps = o0.materials()[1]["epsilon"] / EPS0
    db = "4 12.5 12.5 12.5 4.9999996 99 silica.txt\n8 20 20 20 22 22 22 1 2.123456789 1 0 0\n"
    direct = "4 12.5 12.5 12.5 5 1 %r 1 0 0\n8 20 20 20 22 22 22 1 %g 1 0 0\n" % (eps, F32(2.123456789))
    plain = "4 12.5 12.5 12.5 4.9999996 1 %r 1 0 0\n8 20 20 20 22 22 22 1 2.123456789 1 0 0\n" % eps
    a = Oracle(par_file(d, db))
    b = Oracle(par_file(str(tmp_path / "direct"), direct))
    c = Oracle(par_file(str(tmp_path / "plain"), plain))
    ma, mb, mc = a.maps()[0], b.maps()[0], c.maps()[0]
    assert np.array_equal(ma, mb) and (ma == 1).sum() > 400
    assert (ma == 1).sum() - (mc == 1).sum() == 30
    assert [m["epsilon"] for m in a.materials()] == [m["epsilon"] for m in b.materials()]
    assert a.materials()[2]["epsilon"] == float(F32(g6(float(F32(2.123456789))))) * EPS0
    assert c.materials()[2]["epsilon"] == float(F32(2.123456789)) * EPS0


def test_ade_entry_equals_direct_and_plrc_rewritten(tmp_path):
    d = str(tmp_path)
    ade = "1 0 1.37e16 2.7e13 0 0 0 0 0 0 0 0"
    write(d, "silver.txt", "ade 0.2 2.0\nDrude silver\n%s\n" % ade)
    o = Oracle(par_file(d, "8 2 2 2 10 10 10 99 silver.txt\n8 12 12 12 20 20 20 5 %s\n"
                        "8 3 14 3 9 20 9 6 %s\n" % (ade, ade)))
    mats = o.materials()
    assert len(mats) == 2 and mats[1]["type"] == 5 and mats[1]["poles"] == 2
    assert mats[1]["omega_p"] == float(F32(g6(float(F32(1.37e16)))))
    assert mats[1]["nu"] == float(F32(g6(float(F32(2.7e13)))))
    mat, _ = o.maps()
    assert (mat == 1).sum() == 8 ** 3 + 8 ** 3 + 6 * 6 * 6
    # without a database entry the file is not rewritten and the line stays a
    # piecewise-linear recursive-convolution material (type 6)
    p = Oracle(par_file(str(tmp_path / "p"), "8 3 14 3 9 20 9 6 %s\n" % ade))
    assert [m["type"] for m in p.materials()[1:]] == [6]


def test_lookup_and_errors(tmp_path, monkeypatch):
    spectra = tmp_path / "spectra"
    write(str(spectra), "silica_env.txt", SILICA)
    monkeypatch.setenv("GSVIT_SPECTRA_DIR", str(spectra))
    o = Oracle(par_file(str(tmp_path / "a"), "4 12 12 12 6 99 silica_env.txt\n"))
    assert o.materials()[1]["type"] == 1
    monkeypatch.delenv("GSVIT_SPECTRA_DIR")
    with pytest.raises(RuntimeError):                      # no such entry
        Oracle(par_file(str(tmp_path / "b"), "4 12 12 12 6 99 silica_env.txt\n"))
    write(str(tmp_path / "c"), "silica.txt", SILICA)
    with pytest.raises(RuntimeError):                      # no SOURCE_WAVELENGTH
        Oracle(par_file(str(tmp_path / "c"), "4 12 12 12 6 99 silica.txt\n", lam=None))
    text, _, _, _ = nk_rows(False, False)
    write(str(tmp_path / "d"), "film.txt", text)
    with pytest.raises(RuntimeError):                      # above the table
        Oracle(par_file(str(tmp_path / "d"), "4 12 12 12 6 99 film.txt\n", lam=(800e-9, 900e-9, 1000e-9)))


@pytest.mark.gpu
def test_gpu_database_glass_sphere_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    d = str(tmp_path)
    write(d, "silica.txt", SILICA)
    n, steps = 40, 200
    dt = cases._dt(1e-8, 1e-8, 1e-8, 0.5)
    cases.write_point_source(os.path.join(d, "src.txt"), cases.gaussian_sine(steps, dt, 500e-9, 1.0), comp=2)
    par = par_file(d, "4 20 20 24 8 99 silica.txt\n", n=n,
                   extra="DT_MULT\n0.5\n\nBOUNDARY_ALL\ncpml 8 3 -1 0 1\n\nSOURCE_POINT\n20 20 12 0 src.txt\n\n")
    text = open(par).read().replace("COMP\n2\n", "COMP\n%d\n" % steps)
    with open(par, "w") as f:
        f.write(text)
    print("database glass sphere: %d^3, %d steps" % (n, steps), flush=True)
    with Session(par) as s:
        s.run()
        got = s.fields()
    o = Oracle(par)
    o.step()
    ref = o.fields()
    assert o.materials()[1]["type"] == 1
    e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("GPU vs oracle, database glass: rel L2 %.3e" % e)
    assert e <= 1e-5
