"""Planner checks against independent restatements (brute force and closed forms).

The planner (gsvit_copy_amd/csrc/plan.c) is product code shared with the
oracle; these tests pin it without using it: sphere rasterisation against a
float32 brute force of the voxel-centre test (src3d/plan.c:2136-2143,
4378-4391), the internal-boundary map against a numpy 26-neighbour
comparison (src3d/plan.c:5381-5417), material de-duplication rules
(src3d/plan.c:2318-2367), CPML grading against the closed form
(src3d/boundary.c:33-99) through its effect on a 1-D trace, and the time step
(src3d/plan.c:5452).
"""
import math
import os

import numpy as np
import pytest

from oracle_lib import Oracle


def write(folder, name, text):
    os.makedirs(folder, exist_ok=True)
    p = os.path.join(folder, name)
    with open(p, "w") as f:
        f.write(text)
    return p


def sphere_raster(n, spheres):
    """float32 brute force of the voxel-centre test; later entities overwrite earlier ones"""
    mat = np.zeros((n, n, n), dtype=np.int32)
    idx = np.arange(n, dtype=np.float32) + np.float32(0.5)
    X, Y, Z = np.meshgrid(idx, idx, idx, indexing="ij")
    for m, (cx, cy, cz, r) in spheres:
        cx, cy, cz, r = (np.float32(v) for v in (cx, cy, cz, r))
        d = (X - cx) * (X - cx) + (Y - cy) * (Y - cy) + (Z - cz) * (Z - cz)
        mat[d <= r * r] = m
    return mat


def boundary_map(mat):
    n = mat.shape
    b = np.zeros(n, dtype=np.int32)
    b[0, :, :] = b[-1, :, :] = 1
    b[:, 0, :] = b[:, -1, :] = 1
    b[:, :, 0] = b[:, :, -1] = 1
    core = mat[1:-1, 1:-1, 1:-1]
    diff = np.zeros(core.shape, dtype=bool)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for dk in (-1, 0, 1):
                nb = mat[1 + di:n[0] - 1 + di, 1 + dj:n[1] - 1 + dj, 1 + dk:n[2] - 1 + dk]
                diff |= nb != core
    b[1:-1, 1:-1, 1:-1] |= diff.astype(np.int32)
    return b


def test_sphere_raster_dedup_and_boundary_map(tmp_path):
    rng = np.random.default_rng(11)
    n = 40
    lines = []
    spheres = []
    eps_list = [2.0, 3.0, 2.0005, 1.0004]   # 2.0005 dedups to 2.0 (rel 1e-3), 1.0004 to vacuum (index 0)
    expect_index = [1, 2, 1, 0]
    for e, m in zip(eps_list, expect_index):
        c = rng.uniform(8, 32, size=3)
        r = rng.uniform(3, 9)
        c = np.round(c, 3)
        r = round(r, 3)
        lines.append("4 %.3f %.3f %.3f %.3f 1 %g 1 0 0" % (c[0], c[1], c[2], r, e))
        spheres.append((m, (c[0], c[1], c[2], r)))
    write(str(tmp_path), "s.txt", "\n".join(lines) + "\n")
    par = write(str(tmp_path), "s.par", "POOL\n%d %d %d 1e-8 1e-8 1e-8\n\nCOMP\n2\n\nMEDIUM_VECTOR\ns.txt\n\n" % (n, n, n))
    o = Oracle(par)
    mat, bnds = o.maps()
    assert o.info()["nmat"] == 3
    assert np.array_equal(mat, sphere_raster(n, spheres))
    assert np.array_equal(bnds, boundary_map(mat))


def test_voxel_entity_and_linear_material_map(tmp_path):
    n = 24
    write(str(tmp_path), "v.txt", "8 3 4 5 20 18 30 1 4 1 0 0\n8 10 10 10 12 13 14 0 2 1 0.5 0\n")
    par = write(str(tmp_path), "v.par", "POOL\n%d %d %d 1e-8 1e-8 1e-8\n\nCOMP\n2\n\nMEDIUM_VECTOR\nv.txt\n\n" % (n, n, n))
    o = Oracle(par)
    mat, bnds = o.maps()
    ref = np.zeros((n, n, n), dtype=np.int32)
    ref[3:20, 4:18, 5:24] = 1
    ref[10:12, 10:13, 10:14] = 0   # linear entity: voxel data, tabulated index back to 0
    assert np.array_equal(mat, ref)


def test_unsupported_keyword_is_an_error(tmp_path):
    par = write(str(tmp_path), "bad.par", "POOL\n10 10 10 1e-8 1e-8 1e-8\n\nCOMP\n2\n\nSOURCE_LTSF\n1 2 3\n")
    with pytest.raises(RuntimeError):
        Oracle(par)


CPML_PAR = """POOL
30 32 34 1e-8 2e-8 3e-8

COMP
1

DT_MULT
0.7

BOUNDARY_ALL
cpml 10 3 -1 0.03 4

BOUNDARY_YN
cpml 7 2 -1 0.01 2.5

BOUNDARY_Z0
cpml 6 4 2.5e5 0 1

BOUNDARY_ZN
cpml 8 3 -1 0.02 1.5
"""


def grading(near, depth, m, sigma, alpha, kappa, dt):
    """The CPML grading of src3d/boundary.c:33-99 written out independently,
    as fractions of the layer: sigma and kappa grow as f^m and alpha as g,
    with, at vector index i,
      near faces (becebhch0(i)):      E  f = (d - i) / d,      g = i / d
                                      H  f = (d - i - 1/2) / d, g = (i - 1/2) / d
      far faces (becebhch(d - i)):    E  f = i / d,            g = (d - i) / d
                                      H  f = (i + 1/2) / d,     g = (d - i + 1/2) / d
    (the H samples' alpha takes the opposite half-cell offset from their
    sigma and kappa, as the reference has it); b = exp(-(sigma / kappa + alpha)
    dt / eps0), c = sigma / (sigma + kappa alpha) / kappa (b - 1), 0 where
    sigma = 0."""
    eps0 = 8.854187817620389850e-12
    i = np.arange(depth, dtype=float)
    d = float(depth)
    if near:
        frac = {"e": ((d - i) / d, i / d), "h": ((d - i - 0.5) / d, (i - 0.5) / d)}
    else:
        frac = {"e": (i / d, (d - i) / d), "h": ((i + 0.5) / d, (d - i + 0.5) / d)}
    out = {}
    for side, (f, g) in frac.items():
        p = f ** m
        sg, kp, al = sigma * p, 1 + (kappa - 1) * p, alpha * g
        b = np.exp(-(sg / kp + al) * dt / eps0)
        with np.errstate(invalid="ignore", divide="ignore"):
            c = np.where(sg == 0, 0.0, sg / (sg + kp * al) / kp * (b - 1))
        out["b" + side], out["c" + side], out["kappa" + side] = b, c, kp
    return out


def test_cpml_grading_closed_form(tmp_path):
    """The planner's CPML vectors on all six faces against the closed form,
    for a non-cubic cell where BOUNDARY_ALL and the per-face keywords take
    their default sigma = 0.8 m / (377 d) with d = dx for BOUNDARY_ALL and the
    face's own spacing for BOUNDARY_YN / ZN (src3d/settings.c:5151, 5277,
    5337); dt from the same closed form as test_time_step_formula."""
    par = write(str(tmp_path), "cpml.par", CPML_PAR)
    o = Oracle(par)
    dx, dy, dz = 1e-8, 2e-8, 3e-8
    dt = 0.7 / (299792458.0 * math.sqrt(1 / dx ** 2 + 1 / dy ** 2 + 1 / dz ** 2))
    faces = {0: (True, 10, 3, 0.8 * 3 / (377 * dx), 0.03, 4.0), 1: (False, 10, 3, 0.8 * 3 / (377 * dx), 0.03, 4.0),
             2: (True, 10, 3, 0.8 * 3 / (377 * dx), 0.03, 4.0), 3: (False, 7, 2, 0.8 * 2 / (377 * dy), 0.01, 2.5),
             4: (True, 6, 4, 2.5e5, 0.0, 1.0), 5: (False, 8, 3, 0.8 * 3 / (377 * dz), 0.02, 1.5)}
    for f, (near, depth, m, sigma, alpha, kappa) in faces.items():
        got = o.cpml(f)
        want = grading(near, depth, m, sigma, alpha, kappa, dt)
        for key in ("be", "ce", "bh", "ch", "kappae", "kappah"):
            w = want[key]
            assert got[key].shape == w.shape, (f, key)
            assert np.allclose(got[key], w, rtol=1e-12, atol=1e-300), (f, key, got[key], w)
        # the grading reaches sigma_max at the outer edge and 0 at the inner one
        assert (got["ce"][0] < 0) if near else (got["ce"][0] == 0)
    o.close()


def test_time_step_formula(tmp_path):
    """dt = DT_MULT / (c sqrt(1/dx^2 + 1/dy^2 + 1/dz^2)) (src3d/plan.c:5452)"""
    par = write(str(tmp_path), "dt.par", "POOL\n12 14 16 1e-8 2e-8 3e-8\n\nCOMP\n1\n\nDT_MULT\n0.7\n\n")
    o = Oracle(par)
    want = 0.7 / (299792458.0 * math.sqrt(1e16 + 1 / 4e-16 + 1 / 9e-16))
    assert abs(o.dt() - want) <= 1e-15 * want
    o.close()
