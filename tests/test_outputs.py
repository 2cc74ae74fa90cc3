"""Plane and volume outputs (OUT_PLANE, OUT_VOLUME), src3d/output.c:712-905 and
:1100-1263: written by the engine's host code after copyfrom on output steps,
in the reference's formats (VTK STRUCTURED_POINTS "%g", ASCII "%g" / "%10.10g",
raw doubles, file names base_%.4d[.vtk|.raw]), checked on the GPU against the
FP64 oracle's fields at the same steps; the accumulated volumes (Sumall,
Maxall) against the same sums and maxima over the oracle's fields.
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

NSTEPS = 160
OUTPUTS = """
OUT_VOLUME
Ex 50 -1 -1 2 vol_ex

OUT_VOLUME
Hy 40 60 120 0 vol_hy

OUT_VOLUME
Epsilon 1000 -1 -1 1 vol_eps

OUT_VOLUME
Sumall 20 0 -1 1 sumall

OUT_VOLUME
Maxall 20 0 -1 0 maxall

OUT_PLANE
Ez 30 0 200 1 -1 -1 24 pl_ez

OUT_PLANE
Hx 50 0 200 0 10 -1 -1 pl_hx
"""


def output_case(folder):
    par = cases.small_case(folder, "spheres_cpml", nsteps=NSTEPS)
    with open(par, "a") as f:
        f.write(OUTPUTS)
    return par


def test_planner_rejects_a_plane_with_two_fixed_indices(tmp_path):
    par = cases.small_case(str(tmp_path / "c"), "spheres_cpml", nsteps=10)
    with open(par, "a") as f:
        f.write("OUT_PLANE\nEz 1 0 10 1 5 6 -1 bad\n\n")
    with pytest.raises(RuntimeError):
        Oracle(par)


def read_vtk(path, n):
    with open(path) as f:
        head = [next(f) for _ in range(10)]
        vals = np.array(f.read().split(), dtype=np.float64)
    assert head[0].startswith("# vtk DataFile Version 2.0") and head[3].startswith("DATASET STRUCTURED_POINTS")
    assert head[4].split() == ["DIMENSIONS", str(n), str(n), str(n)]
    return vals.reshape(n, n, n).transpose(2, 1, 0)     # file is x fastest


def close(got, ref, rel=1e-5):
    """max |got - ref| relative to max |ref| (an all-zero reference needs zeros)"""
    scale = np.abs(ref).max()
    if scale == 0:
        return bool(np.abs(got).max() == 0), 0.0
    err = float(np.abs(got - ref).max() / scale)
    return err <= rel, err


@pytest.mark.gpu
def test_plane_and_volume_outputs_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    folder = str(tmp_path / "out")
    par = output_case(folder)
    s = Session(par, write_outputs=True)
    s.run()
    s.close()
    n = 48
    o = Oracle(par)
    fields = {}
    acc_sum = np.zeros((n, n, n))
    acc_max = np.zeros((n, n, n))
    want = {0, 50, 100, 150, 159, 80, 120, 30, 60, 90} | set(range(0, NSTEPS, 20)) | {NSTEPS - 1}
    for st in range(NSTEPS):
        o.step(1)
        if st in want:
            f = o.fields()
            fields[st] = f
            if st % 20 == 0 or st == NSTEPS - 1:
                e2 = f[0] ** 2 + f[1] ** 2 + f[2] ** 2
                acc_sum += e2
                acc_max = np.maximum(acc_max, e2)
    mats, _ = o.maps()
    o.close()
    errs = {}
    for st in (0, 50, 100, 150, 159):                           # VTK, every 50 steps and the last
        ok, errs["vtk Ex %d" % st] = close(read_vtk(os.path.join(folder, "vol_ex_%.4d.vtk" % st), n), fields[st][0])
        assert ok
    for st in (80, 120):                                        # raw doubles inside the window 60..120
        raw = np.fromfile(os.path.join(folder, "vol_hy_%.4d.raw" % st), dtype=np.float64).reshape(n, n, n)
        ok, errs["raw Hy %d" % st] = close(raw.transpose(2, 1, 0), fields[st][4])
        assert ok
    assert not os.path.exists(os.path.join(folder, "vol_hy_0040.raw"))
    assert not os.path.exists(os.path.join(folder, "vol_hy_0159.raw"))
    # ASCII epsilon: header, then rows of x for each y, blank line after each z plane
    with open(os.path.join(folder, "vol_eps_0000")) as f:
        assert f.readline().split() == ["48", "48", "48", "7"]
        eps = np.array(f.read().split(), dtype=np.float64).reshape(n, n, n).transpose(2, 1, 0)
    # spheres of tabulated dielectrics (material index per voxel), then a linear
    # box of eps 2.5 over (4..11, 4..11, 4..9) that sets its voxels directly
    lines = [ln.split() for ln in open(os.path.join(folder, "mats.txt"))]
    table = [1.0] + [float(v[6]) for v in lines if v[0] == "4"]
    want_eps = np.array(table)[mats]
    want_eps[4:12, 4:12, 4:10] = 2.5
    assert np.allclose(eps, want_eps, rtol=1e-5)
    # accumulated |E|^2 over the output steps (every 20 and the last)
    with open(os.path.join(folder, "sumall")) as f:
        f.readline()
        sumall = np.array(f.read().split(), dtype=np.float64).reshape(n, n, n).transpose(2, 1, 0)
    ok, errs["Sumall"] = close(sumall, acc_sum)
    assert ok
    maxall = np.fromfile(os.path.join(folder, "maxall.raw"), dtype=np.float64).reshape(n, n, n).transpose(2, 1, 0)
    ok, errs["Maxall"] = close(maxall, acc_max)
    assert ok
    # planes: Ez at k = 24 (rows j, columns i, "%10.10g"); Hx at i = 10 (rows k, columns j, raw)
    for st in (30, 60, 90, 120, 150, 159):
        pl = np.loadtxt(os.path.join(folder, "pl_ez_%.4d" % st))
        ok, errs["plane Ez %d" % st] = close(pl.T, fields[st][2][:, :, 24])
        assert ok
    assert not os.path.exists(os.path.join(folder, "pl_ez_0000"))   # the window's start is exclusive
    for st in (50, 100, 150, 159):
        raw = np.fromfile(os.path.join(folder, "pl_hx_%.4d.raw" % st), dtype=np.float64).reshape(n, n)
        ok, errs["plane Hx %d" % st] = close(raw.T, fields[st][3][10, :, :])
        assert ok
    print("max relative errors", errs)
