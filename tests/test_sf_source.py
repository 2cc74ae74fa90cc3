"""Scattered-field source (SOURCE_SF; src3d/source.c:1319-1338,
src3d/yee.c:1139-1237, src3d/plan.c:1985-2040): the whole pool holds the
scattered field, and the components the PEC rule would zero are set to minus
the incident field, read from the source's own 1-D line.

Pin (the oracle's restatement): at normal incidence in a cell periodic in x
and y the grid carries exactly the line's plane wave, so an SF run's field
equals the scattered field of a TSF run with the same PEC plate (below the
TSF box, which holds the scattered field only) -- two steps later.  The
reference advances the SF line at the end of a step (twice at step 0, its
"synchronize with TSF" step) and reads it during the next E update, which
sets E^(n+1) on the plate from the incident field of E^(n-1).  The equality
holds to round-off until reflections from the far end of the SF line (which
is only sqrt(nx^2 + ny^2 + nz^2 + 10) cells long) reach the plate.  With the
sign of the incident field flipped, as a planted mistake, the fields differ
by their own size.  On the GPU the engine's SF run is held to the oracle.
"""
import numpy as np
import pytest

import cases
from oracle_lib import Oracle

PTS = [(6, 6, 15), (6, 6, 60)]


def run(par):
    o = Oracle(par)
    o.step()
    f, p = o.fields(), o.points()
    o.close()
    return f, p


@pytest.fixture(scope="module")
def plate_runs(tmp_path_factory):
    out = {}
    for src in ("tsf", "sf"):
        out[src] = run(cases.pec_plate_wave(str(tmp_path_factory.mktemp(src)), src, points=PTS))
    return out


def test_sf_equals_tsf_scattered_field_two_steps_later(plate_runs):
    tsf = plate_runs["tsf"][1][0][:-1]            # trace below the TSF box (the last row is never recorded)
    sf = plate_runs["sf"][1][0][:-1]
    scale = np.abs(tsf).max()
    early = np.abs(tsf[:398] - sf[2:400]).max() / scale
    whole = np.abs(tsf[:-2] - sf[2:]).max() / scale
    same = np.abs(tsf - sf).max() / scale
    print("SF vs TSF scattered field two steps earlier: first 400 steps %.2e, whole run %.2e; unshifted %.2e"
          % (early, whole, same))
    assert scale > 0.05
    assert early <= 1e-10
    assert whole <= 1e-4
    assert same > 1e-2


def test_sf_fills_the_pool_with_the_scattered_field(plate_runs):
    """in front of the plate the SF run holds the reflected wave only, the TSF
    run the incident wave as well"""
    tsf = plate_runs["tsf"][1][1][:-1]
    sf = plate_runs["sf"][1][1][:-1]
    assert np.abs(tsf).max() > 2 * np.abs(sf).max() > 0


@pytest.mark.gpu
def test_gpu_sf_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors, trace_errors
    par = cases.pec_plate_wave(str(tmp_path / "sf"), "sf", points=PTS)
    print("SF source: PEC plate, periodic 12 x 12 x 140", flush=True)
    ref, rp = run(par)
    s = Session(par)
    s.run()
    got, gp = s.fields(), s.points()
    s.close()
    errs = field_errors(got, ref)
    terr = trace_errors(gp, rp)
    print("GPU vs oracle", errs, "traces", terr)
    assert max(errs.values()) <= 1e-5
    assert max(terr) <= 1e-5
