"""Far-field point areas and database-material sums.

NFFF_SPHERICAL_AREA, NFFF_PLANAR_AREA and PERIODIC_NFFF_SPHERICAL_AREA
(src3d/settings.c:5532-5816) add a grid of far-field points, one set each:
positions from the pool centre with C truncation toward zero, per-point logs
xnff_set<s>_<i>_<j> when savefile != 0, and one istart for all points of a set,
the first point's (src3d/farfield.c:81-91).  Pins: the planned points equal
the geometry computed here from the keyword's definition; a point of a set
holds the bins of the same cell declared alone (NFFF_RAMAHI_POINT) shifted by
the difference of the two istarts, exactly; the GPU matches the oracle and
writes the logs of the sets that ask for them.

OUT_SUMTAB (src3d/settings.c:4982-5017, src3d/plan.c:5475-5486) is OUT_SUM
whose voxel filter is a database material resolved at the centre wavelength:
its sums equal the numpy sum over the voxels of that material and the OUT_SUM
given the same permittivity in absolute units.
"""
import math
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

C = 299792458.0
EPS0 = 8.854187817e-12
N, NSTEPS = 32, 100
SPH = (3, 4, 400, 0.3, 0.2, 1.2, 2.9)          # thetares phires radius thetafrom phifrom thetato phito
PLA = (3, 2, -1e-6, -0.5e-6, 1e-6, 0.5e-6, 2, 3e-6)  # ijres jkres ijfrom jkfrom ijto jkto orientation distance
PSPH = (2, 3, 250, 0.2, 0.0, 0.7, 3.0)


def area_case(folder):
    par = cases.small_case(folder, "farfield_cpml", n=N, nsteps=NSTEPS)
    with open(par, "a") as f:
        f.write("NFFF_SPHERICAL_AREA\n%d %d %d %r %r %r %r 1\n\n" % SPH)
        f.write("NFFF_PLANAR_AREA\n%d %d %r %r %r %r %d %r 0\n\n" % PLA)
        i, j = 2, 3
        f.write("NFFF_RAMAHI_POINT\n%d %d %d alone\n\n" % sph_point(i, j))
    return par


def sph_point(i, j, spec=SPH):
    tres, pres, radius, tf, pf, tt, pt = spec
    theta = tf + float(i) * (tt - tf) / (tres - 1) if tres > 1 else tf
    phi = pf + float(j) * (pt - pf) / (pres - 1) if pres > 1 else pf
    return (int(radius * math.sin(theta) * math.cos(phi)) + N // 2,
            int(radius * math.sin(theta) * math.sin(phi)) + N // 2,
            int(radius * math.cos(theta)) + N // 2)


def pla_point(i, j):
    ijres, jkres, ijf, jkf, ijt, jkt, orient, dist = PLA
    u = ijf + (ijt - ijf) * float(i) / (float(ijres) - 1)
    v = jkf + (jkt - jkf) * float(j) / (float(jkres) - 1)
    dx = 10e-9
    a, b, d = int(u / dx) + N // 2, int(v / dx) + N // 2, int(dist / dx) + N // 2
    return {0: (d, a, b), 1: (a, d, b), 2: (a, b, d)}[orient]


def istart(p, dt, dx=10e-9):
    div = C * dt / dx
    v = math.sqrt(sum((q / div) ** 2 for q in p)) - 500
    return int(v if v > 0 else 0)


def test_area_points_follow_the_keyword_geometry(tmp_path):
    o = Oracle(area_case(str(tmp_path / "a")))
    pts = o.farfield_points()
    dt = o.dt()
    o.close()
    assert len(pts) == 2 + 12 + 6 + 1
    want = [sph_point(i, j) for i in range(3) for j in range(4)] + [pla_point(i, j) for i in range(3) for j in range(2)]
    got = [tuple(int(v) for v in p[:3]) for p in pts[2:20]]
    assert got == want
    # one istart per set (the first point's); single points keep their own
    assert all(p[3] == istart(want[0], dt) for p in pts[2:14])
    assert all(p[3] == istart(want[12], dt) for p in pts[14:20])
    assert len({istart(w, dt) for w in want[:12]}) > 1          # the sharing matters here
    assert pts[20][3] == istart(sph_point(2, 3), dt)
    assert pts[20][3] != pts[13][3]


def test_set_point_bins_are_the_single_point_bins_shifted(tmp_path):
    o = Oracle(area_case(str(tmp_path / "a")))
    o.step()
    ff = o.farfield()
    pts = o.farfield_points()
    o.close()
    a, b = ff[13], ff[20]                      # set 1 point (2, 3) and the same cell alone
    d = int(pts[13][3] - pts[20][3])           # bins[m] of the set point = bins[m + d] of the single one
    nd = a.shape[1]
    m = np.arange(max(0, -d), min(nd, nd - d))
    assert np.abs(b).max() > 0
    assert np.array_equal(a[:, m], b[:, m + d])
    assert not np.any(a[:, : max(0, -d)]) and not np.any(a[:, min(nd, nd - d):])


def test_periodic_spherical_area_points(tmp_path):
    folder = str(tmp_path / "p")
    par = cases.small_case(folder, "farfield_cpml", n=N, nsteps=20)
    with open(par, "a") as f:
        f.write("PERIODIC_NFFF\n8 8 8 24 24 24 -1 -1 2 2\n\n")
        f.write("PERIODIC_NFFF_SPHERICAL_AREA\n%d %d %d %r %r %r %r 1\n\n" % PSPH)
    o = Oracle(par)
    pts = o.farfield_points(periodic=True)
    dt = o.dt()
    o.close()
    want = [sph_point(i, j, PSPH) for i in range(2) for j in range(3)]
    assert [tuple(int(v) for v in p[:3]) for p in pts] == want
    # the periodic far field keeps one istart per point (src3d/pfarfield.c:50-56)
    assert [int(p[3]) for p in pts] == [istart(w, dt) for w in want]


def test_planar_area_with_one_point_rows_is_refused(tmp_path):
    folder = str(tmp_path / "r")
    par = cases.small_case(folder, "farfield_cpml", n=N, nsteps=20)
    with open(par, "a") as f:
        f.write("NFFF_PLANAR_AREA\n1 3 0 0 1e-6 1e-6 2 3e-6 0\n\n")
    with pytest.raises(RuntimeError):
        Oracle(par)


GLASS = "cauchy 0.2 2.0\nconstant index sqrt(2.5)\n1.58113883 0 0\n"
BLOCK = (10, 9, 8, 22, 24, 20)          # LINTAB eps 2.5 box, material index 1


def sumtab_case(folder):
    os.makedirs(folder, exist_ok=True)
    par = cases.small_case(folder, "vacuum_none", n=N, nsteps=60)
    with open(os.path.join(folder, "glass25.txt"), "w") as f:
        f.write(GLASS)
    with open(os.path.join(folder, "mats.txt"), "w") as f:
        f.write("8 %d %d %d %d %d %d 1 2.5 1 0 0\n" % BLOCK)
    with open(par, "a") as f:
        f.write("SOURCE_WAVELENGTH\n400e-9 600e-9 800e-9\n\nMEDIUM_VECTOR\nmats.txt\n\n")
        f.write("OUT_SUMTAB\nAll 10 4 4 4 28 28 28 glass25.txt sumtab\n\n")
        f.write("OUT_SUM\nAll 10 4 4 4 28 28 28 %.12g 1 0 0 sumabs\n\n" % (2.5 * EPS0))
        f.write("OUT_SUM\nAll 10 4 4 4 28 28 28 -1 1 0 0 sumall\n\n")
    return par


def test_out_sumtab_sums_the_database_material(tmp_path):
    o = Oracle(sumtab_case(str(tmp_path / "s")))
    o.step()
    sums = o.sums(3)
    e = o.fields()[:3]
    mat, _ = o.maps()
    o.close()
    e2 = e[0] ** 2 + e[1] ** 2 + e[2] ** 2
    sl = (slice(4, 28),) * 3
    assert (mat == 1).sum() == (mat[sl] == 1).sum() > 0      # the block lies inside the sum box
    want = np.sum(e2[sl][mat[sl] == 1])
    print("OUT_SUMTAB", sums[0][-1], "numpy", want, "OUT_SUM absolute eps", sums[1][-1], "all", sums[2][-1])
    assert want > 0 and sums[2][-1] > want * (1 + 1e-6)
    assert sums[0][-1] == pytest.approx(want, rel=1e-12)
    assert np.array_equal(sums[0], sums[1])


def test_out_sumtab_without_the_database_entry_is_refused(tmp_path):
    folder = str(tmp_path / "m")
    par = sumtab_case(folder)
    os.remove(os.path.join(folder, "glass25.txt"))
    with pytest.raises(RuntimeError):
        Oracle(par)


@pytest.mark.gpu
def test_gpu_areas_and_sumtab_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    folder = str(tmp_path / "a")
    par = area_case(folder)
    print("case nfff areas", flush=True)
    o = Oracle(par)
    o.step()
    ref = o.farfield()
    o.close()
    s = Session(par, write_outputs=True)
    s.run()
    got = s.farfield()
    s.close()
    e = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print("area far field GPU vs oracle rel L2", e)
    assert e <= 1e-5
    names = sorted(n for n in os.listdir(folder) if n.startswith("xnff_set"))
    assert names == ["xnff_set001_%03d_%03d" % (i, j) for i in range(3) for j in range(4)]
    rows = np.loadtxt(os.path.join(folder, "xnff_set001_002_003"))
    assert np.array_equal(rows[:, 0], np.arange(got.shape[2]))
    np.testing.assert_allclose(rows[:, 1:], -got[13].T, rtol=1e-5, atol=1e-6 * np.abs(got[13]).max())
    print("case sumtab", flush=True)
    par = sumtab_case(str(tmp_path / "s"))
    o = Oracle(par)
    o.step()
    rs = o.sums(3)
    o.close()
    s = Session(par)
    s.run()
    gs = s.sums()
    s.close()
    for q in range(3):
        eq = float(np.abs(gs[q][: rs.shape[1]] - rs[q]).max() / np.abs(rs[q]).max())
        print("sum", q, "GPU vs oracle max rel", eq)
        assert eq <= 1e-5
