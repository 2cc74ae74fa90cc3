"""Running-DFT near-to-far-field accumulators (new on this path, no reference).

The oracle's accumulators are pinned against an independent computation: a
numpy DFT of the point traces that the oracle records at the same face cells
and at their neighbours below along the face normal (E sampled at (s+1) dt on
the face cell, H at (s+1/2) dt as the mean of the two cells, which puts it on
the face plane).  A wrong time offset, exponent sign, component choice, cell
ordering or neighbour fails this check.  The GPU test then compares the
engine's accumulators with the oracle's.
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

N = 48
# (i, j, k, face) on the farfield_cpml box 12..36
FACE_POINTS = [(12, 20, 30, 0), (36, 25, 17, 1), (20, 12, 30, 2), (30, 36, 22, 3), (17, 23, 12, 4), (25, 30, 36, 5),
               (12, 12, 12, 0), (35, 35, 36, 5)]
TANGENTIAL = {0: (1, 2, 4, 5), 1: (1, 2, 4, 5), 2: (0, 2, 3, 5), 3: (0, 2, 3, 5), 4: (0, 1, 3, 4), 5: (0, 1, 3, 4)}


def below(i, j, k, face):
    """the neighbour along the face normal whose H is averaged with the cell's"""
    return (i - 1, j, k) if face < 2 else (i, j - 1, k) if face < 4 else (i, j, k - 1)


def dft_case(folder):
    par = cases.small_case(folder, "farfield_cpml", n=N)
    with open(par, "a") as f:
        for q, (i, j, k, _) in enumerate(FACE_POINTS):
            f.write("OUT_POINT\nAll 1 %d %d %d face%d\n\n" % (i, j, k, q))
        for q, (i, j, k, face) in enumerate(FACE_POINTS):
            f.write("OUT_POINT\nAll 1 %d %d %d below%d\n\n" % (below(i, j, k, face) + (q,)))
    return par


def dft_index(i, j, k, face, box=(12, 12, 12, 36, 36, 36)):
    i0, j0, k0, i1, j1, k1 = box
    ix, iy, iz = i1 - i0, j1 - j0, k1 - k0
    far = face % 2
    if face < 2:
        return 2 * ((j - j0) * iz + (k - k0)) + far
    if face < 4:
        return 2 * iy * iz + 2 * ((i - i0) * iz + (k - k0)) + far
    return 2 * iy * iz + 2 * ix * iz + 2 * ((i - i0) * iy + (j - j0)) + far


def dt_of(par):
    # dt = dtmult / (c sqrt(3) / dx), src3d/plan.c:5452 (cubic cells, DT_MULT 0.5)
    return 0.5 / (299792458.0 * np.sqrt(3.0 / (10e-9) ** 2))


def test_oracle_dft_equals_numpy_dft_of_traces(tmp_path):
    par = dft_case(str(tmp_path / "dft"))
    o = Oracle(par)
    o.step()
    inf = o.info()
    nsteps = inf["nsteps"]
    npt = len(FACE_POINTS)
    allp = o.points()
    tr, trb = allp[-2 * npt:-npt], allp[-npt:]
    acc = o.dft().reshape(-1, 2, 8)
    o.close()
    dt = dt_of(par)
    omegas = 2 * np.pi * 299792458.0 / np.array([300e-9, 450e-9])
    s = np.arange(nsteps)
    errs = []
    for q, (i, j, k, face) in enumerate(FACE_POINTS):
        got = acc[dft_index(i, j, k, face)]
        for w, om in enumerate(omegas):
            want = []
            for slot, comp in enumerate(TANGENTIAL[face]):
                t = (s + 1.0) * dt if comp < 3 else (s + 0.5) * dt
                v = tr[q, :nsteps, comp] if comp < 3 else 0.5 * (tr[q, :nsteps, comp] + trb[q, :nsteps, comp])
                z = np.sum(v * np.exp(-1j * om * t))
                want += [z.real, z.imag]
            want = np.array(want)
            assert np.linalg.norm(want) > 0
            errs.append(np.linalg.norm(got[w] - want) / np.linalg.norm(want))
    assert max(errs) < 1e-10, errs
