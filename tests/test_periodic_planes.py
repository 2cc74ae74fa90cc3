"""Interior periodic planes (MBOUNDARY_*, src3d/boundary.c:304-346 after the
H step, :889-932 after the E step): tangential H of plane b0pos-1 is copied
from plane bnpos-1 and tangential E of plane bnpos from plane b0pos, which
makes the window b0pos <= i, j < bnpos an exactly periodic cell, coupled to
the rest of the pool through nothing else.  The oracle's restatement is
pinned by what that fixes, independently of its own copy code:

  * under a normally incident plane wave the window's fields are uniform in
    x and y, bit for bit, whatever scatters outside it;
  * moving a scatterer inside the window by (s, t) cells moves the window's
    fields by (s, t) cells, cyclically, bit for bit.

Both fail with a planted mistake in the oracle's copies (H taken from plane
bnpos instead of bnpos-1; E copied from bnpos to b0pos instead of the other
way), which the tests check as well.  On the GPU, the engine's copies are
held to the oracle in the same two cases.
"""
import numpy as np
import pytest

import cases
from oracle_lib import Oracle

W0, W1 = 8, 32
OUTSIDE = (1, 10, 12, 6, 30, 26, 4.0)          # between the Liao face and the window, in x
INSIDE = (12, 14, 14, 17, 19, 22, 3.0)
SHIFT = (5, 3)


def run_oracle(par, fault=0):
    o = Oracle(par)
    if fault:
        o.set_option("mb_fault", fault)
    o.step()
    f = o.fields()
    o.close()
    return f


def window(f):
    return f[:, W0:W1, W0:W1, :]


def is_uniform(f):
    w = window(f)
    return bool(np.all(w == w[:, :1, :1, :]))


def uniform_case(folder):
    return cases.periodic_window(folder, window=(W0, W1), blocks=(OUTSIDE,))


def shifted(block, s, t):
    i0, j0, k0, i1, j1, k1, e = block
    return (i0 + s, j0 + t, k0, i1 + s, j1 + t, k1, e)


def test_window_is_uniform_under_normal_incidence(tmp_path):
    f = run_oracle(uniform_case(str(tmp_path / "u")))
    assert np.abs(window(f)[:3]).max() > 0                    # the pulse is there
    outside = f[:, 2:W0, W0:W1, :]
    assert not np.all(outside == outside[:, :1, :1, :])      # the block outside makes a difference there
    assert is_uniform(f)


def test_scatterer_translation_moves_fields_cyclically(tmp_path):
    a = run_oracle(cases.periodic_window(str(tmp_path / "a"), blocks=(INSIDE,)))
    b = run_oracle(cases.periodic_window(str(tmp_path / "b"), blocks=(shifted(INSIDE, *SHIFT),)))
    wa, wb = window(a), window(b)
    assert not np.array_equal(wa, np.roll(wa, SHIFT, axis=(1, 2)))
    assert np.array_equal(np.roll(wa, SHIFT, axis=(1, 2)), wb)


@pytest.mark.parametrize("fault", [1, 2])
def test_planted_copy_mistakes_break_both_invariants(tmp_path, fault):
    f = run_oracle(uniform_case(str(tmp_path / "u")), fault)
    assert not is_uniform(f)
    a = run_oracle(cases.periodic_window(str(tmp_path / "a"), blocks=(INSIDE,)), fault)
    b = run_oracle(cases.periodic_window(str(tmp_path / "b"), blocks=(shifted(INSIDE, *SHIFT),)), fault)
    assert not np.array_equal(np.roll(window(a), SHIFT, axis=(1, 2)), window(b))


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["uniform", "inside"])
def test_gpu_periodic_planes_against_oracle(tmp_path, kind):
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors
    par = (uniform_case(str(tmp_path / kind)) if kind == "uniform"
           else cases.periodic_window(str(tmp_path / kind), blocks=(INSIDE,)))
    print("periodic window", kind, flush=True)
    ref = run_oracle(par)
    s = Session(par)
    s.run()
    got = s.fields()
    s.close()
    errs = field_errors(got, ref)
    print(kind, errs)
    assert max(errs.values()) <= 1e-5
    if kind == "uniform":
        w = window(got)
        spread = np.abs(w - w[:, :1, :1, :]).max() / np.abs(w).max()
        print("GPU window spread", spread)
        assert spread <= 1e-6


def test_mboundary_all_equals_the_six_keywords(tmp_path):
    """MBOUNDARY_ALL periodic x0 xn y0 yn z0 zn (src3d/settings.c:5109-5128)
    plans the same six interior periodic planes as the six MBOUNDARY_* keys"""
    import os
    par = cases.periodic_window(str(tmp_path / "a"), n=40, nsteps=60, blocks=(INSIDE,))
    text = open(par).read()
    six = "".join("MBOUNDARY_%s\nperiodic %d\n\n" % (f, p) for f, p in (("Z0", 6), ("ZN", 34)))
    one = text
    for f in ("X0", "XN", "Y0", "YN"):
        one = one.replace("MBOUNDARY_%s\nperiodic %d\n\n" % (f, W0 if f.endswith("0") else W1), "")
    one += "MBOUNDARY_ALL\nperiodic %d %d %d %d 6 34\n\n" % (W0, W1, W0, W1)
    pa = os.path.join(os.path.dirname(par), "six.par")
    pb = os.path.join(os.path.dirname(par), "all.par")
    open(pa, "w").write(text + six)
    open(pb, "w").write(one)
    assert "MBOUNDARY_X0" not in one
    fa, fb = run_oracle(pa), run_oracle(pb)
    assert all(np.array_equal(x, y) for x, y in zip(fa, fb))
    assert max(float(np.abs(x).max()) for x in fa) > 0
