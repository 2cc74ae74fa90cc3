"""OUT_SUM box sums (src3d/output.c:259-306).

The oracle's per-step sums are pinned at the last step against the same
quantity computed in numpy from the oracle's final E fields and its material
map (different code path: field download + masking here, accumulation inside
the step loop there).  Modes: Ex only, |E|^2, sigma |E|^2, max |E|^2, and a
permittivity-restricted |E|^2 (vacuum voxels only).
"""
import os

import numpy as np

import cases
from oracle_lib import Oracle

EPS0 = 8.854187817e-12
N = 40
SIGMA = 1000.0
BOX = (14, 12, 10, 26, 30, 22)          # dielectric block (LINTAB eps 2.5, sigma SIGMA)
SUMS = [  # (mode, box, layered_epsilon)
    ("Ex", (5, 6, 7, 33, 31, 29), -1),
    ("All", (5, 6, 7, 33, 31, 29), -1),
    ("Abs", (10, 10, 8, 30, 34, 26), -1),
    ("Max", (3, 3, 3, 37, 37, 37), -1),
    ("All", (5, 6, 7, 33, 31, 29), EPS0),
]


def sums_case(folder, nsteps=120):
    os.makedirs(folder, exist_ok=True)
    par = cases.small_case(folder, "spheres_cpml", n=N, nsteps=nsteps)
    # replace the sphere geometry by one block with known sigma
    with open(os.path.join(folder, "mats.txt"), "w") as f:
        f.write("8 %d %d %d %d %d %d 1 2.5 1 %g 0\n" % (BOX + (SIGMA,)))
    with open(par, "a") as f:
        for q, (mode, b, le) in enumerate(SUMS):
            f.write("OUT_SUM\n%s 10 %d %d %d %d %d %d %.10g 1 0 0 sum%d\n\n" % ((mode,) + b + (le, q)))
    return par


def expected_last(fields, mat):
    ex2, ey2, ez2 = fields[0] ** 2, fields[1] ** 2, fields[2] ** 2
    e2 = ex2 + ey2 + ez2
    sigma = np.where(mat == 1, SIGMA, 0.0)
    out = []
    for mode, (i0, j0, k0, i1, j1, k1), le in SUMS:
        sl = (slice(i0, i1), slice(j0, j1), slice(k0, k1))
        keep = np.ones_like(e2, dtype=bool) if le == -1 else (mat != 1)
        if mode == "Ex":
            out.append(np.sum(ex2[sl][keep[sl]]))
        elif mode == "All":
            out.append(np.sum(e2[sl][keep[sl]]))
        elif mode == "Abs":
            out.append(np.sum((sigma * e2)[sl][keep[sl]]))
        else:
            out.append(np.max(e2[sl][keep[sl]]))
    return np.array(out)


def test_oracle_sums_match_numpy_at_last_step(tmp_path):
    par = sums_case(str(tmp_path / "sums"))
    o = Oracle(par)
    o.step()
    sums = o.sums(len(SUMS))
    fields = o.fields()
    mat, _ = o.maps()
    o.close()
    assert (mat == 1).sum() == np.prod(np.array(BOX[3:]) - np.array(BOX[:3]))
    want = expected_last(fields, mat)
    got = sums[:, -1]
    assert np.all(want > 0)
    rel = np.abs(got - want) / want
    assert rel.max() < 1e-12, (got, want)
    # the restricted sum excludes the block, so it is smaller than the full one
    assert sums[4, -1] < sums[1, -1]

