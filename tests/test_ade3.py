"""Three-pole ADE materials (type 7: Drude + three critical points, e.g. two
critical points and a Lorentz term) -- an extension with no reference
counterpart, needed for the thin-film workload's silicon.

The third pole uses the reference's pole formulas (src3d/plan.c:2736-2762) and
ADE branch (src3d/yee.c:914-1108).  What pins it: the ADE update is linear in a
pole's amplitude and the poles enter as a sum, so a pole of amplitude A equals
two identical poles of amplitude A/2 -- the oracle with pole 1 split over
poles 1 and 3 must reproduce the two-pole material; and a third pole of zero
amplitude must change nothing.  The engine is then held to the oracle.
"""
import numpy as np
import pytest

import cases
from oracle_lib import Oracle


def _fields(folder, kind, precision="f64"):
    par = cases.small_case(str(folder / kind), kind)
    o = Oracle(par, precision)
    o.step()
    return par, o.fields(), o.points()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_split_pole_equals_two_pole_material(tmp_path):
    _, two, tr2 = _fields(tmp_path, "ade3_two")
    _, split, trs = _fields(tmp_path, "ade3_split")
    e = rel(split, two)
    print("split vs two-pole rel L2", e)
    assert e < 1e-12
    assert rel(trs, tr2) < 1e-12
    # the split material really runs three poles: a third pole of the same
    # amplitude as the halved first one but different dynamics differs
    _, lor, _ = _fields(tmp_path, "ade3_lorentz")
    assert rel(lor, two) > 1e-4


def test_idle_third_pole_changes_nothing(tmp_path):
    _, two, _ = _fields(tmp_path, "ade3_two")
    _, idle, _ = _fields(tmp_path, "ade3_idle")
    assert rel(idle, two) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["ade3_lorentz", "ade3_split"])
def test_three_pole_engine_against_oracle(tmp_path, kind):
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors, trace_errors
    par, ref, rtr = _fields(tmp_path, kind)
    with Session(par) as s:
        assert s.info()["ade_poles"] == 3
        s.run()
        got, gtr = s.fields(), s.points()
    errs = field_errors(got, ref)
    print(kind, "GPU vs FP64 oracle", errs)
    assert max(errs.values()) <= 1e-5
    assert max(trace_errors(gtr, rtr)) <= 1e-5
