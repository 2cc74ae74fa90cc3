"""The fused step (gsv_set_step_mode 1: H, TSF H faces and E of a time step in
one kernel, fields double-buffered) against the FP64 oracle, on the small
cases it supports and the reduced rough-surface and textured-film workloads,
to the parity bar of the two-pass step (relative L2 <= 1e-5, traces too).
"""
import numpy as np
import pytest

import cases
from oracle_lib import Oracle
from test_gpu_parity import field_errors, trace_errors

pytestmark = pytest.mark.gpu

FUSED_KINDS = ["vacuum_none", "spheres_cpml", "ade_cpml", "tsf_ade_cpml", "farfield_cpml", "solids_cpml"]


def run_fused(par, rows=4, kchunk=16):
    from gsvit_copy_amd import Session
    s = Session(par)
    assert s.step_mode()["fused_available"]
    s.set_step_mode(True, rows, kchunk)
    s.run()
    return s


@pytest.mark.parametrize("kind", FUSED_KINDS)
def test_fused_small_cases(tmp_path, kind):
    par = cases.small_case(str(tmp_path / kind), kind)
    print("fused", kind, flush=True)
    o = Oracle(par)
    o.step()
    s = run_fused(par, rows=3, kchunk=7)          # uneven rows and chunks exercise the block edges
    errs = field_errors(s.fields(), o.fields())
    terr = trace_errors(s.points(), o.points())
    print(kind, errs, max(terr))
    assert max(errs.values()) <= 1e-5 and max(terr) <= 1e-5
    s.close()
    o.close()


def test_fused_rough_surface_and_film(tmp_path):
    for name, par in (("rough", cases.rough_surface(str(tmp_path / "r"), n=64, nsteps=400, depth=8, nfff=False)),
                      ("film", cases.textured_film(str(tmp_path / "f"), nx=40, ny=40, nz=150, nsteps=600, nw=4,
                                                   au=10, si=30, sio2=8))):
        o = Oracle(par)
        o.step()
        s = run_fused(par)
        errs = field_errors(s.fields(), o.fields())
        print(name, errs)
        assert max(errs.values()) <= 1e-5
        s.close()
        o.close()


def test_fused_mode_is_refused_where_unsupported(tmp_path):
    from gsvit_copy_amd import Session
    par = cases.small_case(str(tmp_path / "liao"), "magnetic_liao")
    with Session(par) as s:
        assert not s.step_mode()["fused"] and not s.step_mode()["fused_available"]
        with pytest.raises(ValueError):
            s.set_step_mode(True)
