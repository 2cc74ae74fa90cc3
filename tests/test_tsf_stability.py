"""Stability of the TSF auxiliary line.

The reference's 1-D line (src3d/1dpool.c:83-162, stepped in the order of
src3d/source.c:1469-1473: E update, Mur conditions on E and H at both ends,
hard source on E at index 0, H update) applies the Mur condition to the H
sample at the source end and then the Yee update to the same sample.  That
scheme grows without bound once the line's Courant number c dt / (dx / corr)
exceeds 0.75: at the default time step an oblique plane wave at 45 degrees
runs the line at corr / sqrt 3 = 0.816.  The oracle restates the reference
and therefore reproduces the growth (so does the engine, which is held to the
oracle); the planner warns, and the oblique workloads of the tests and the
bench run at DT_MULT 0.9 (line Courant 0.735).  Pinned here on the oracle:
the vacuum plane wave diverges at DT_MULT 1 and stays bounded at 0.9, and the
planner's warning appears only for the former.
"""
import re

import numpy as np

import cases
from oracle_lib import Oracle


def oblique_vacuum(folder, dtmult, nsteps=900):
    par = cases.rough_surface(folder, n=32, nsteps=nsteps, depth=6, nfff=False)
    text = open(par).read()
    text = re.sub(r"MEDIUM_VECTOR\nmetal.txt\n\nMEDIUM_VTK_DATA\nsurface.vtk\n\n", "", text)
    text = text.replace("DT_MULT\n%g\n" % cases.OBLIQUE_DT_MULT, "DT_MULT\n%g\n" % dtmult)
    assert "DT_MULT\n%g\n" % dtmult in text and "MEDIUM" not in text
    with open(par, "w") as f:
        f.write(text)
    return par


def test_oblique_tsf_line_diverges_above_courant_limit(tmp_path, capfd):
    peaks = {}
    for dtmult in (1.0, cases.OBLIQUE_DT_MULT):
        o = Oracle(oblique_vacuum(str(tmp_path / ("dt%g" % dtmult)), dtmult))
        err = capfd.readouterr().err
        warned = "TSF auxiliary line runs at Courant number" in err
        o.step()
        peaks[dtmult] = float(np.abs(o.fields()[:3]).max())
        o.close()
        assert warned == (dtmult == 1.0)
    print("max |E| after 900 steps per DT_MULT", peaks)
    assert peaks[1.0] > 1e3                       # the reference's line has diverged
    assert peaks[cases.OBLIQUE_DT_MULT] < 10.0    # a unit-amplitude pulse stays bounded
