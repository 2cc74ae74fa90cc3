"""Dispersive (ADE) cells inside CPML slabs.

The reference updates the CPML psi of an ADE component twice per step --
inline in its ADE branch (src3d/yee.c:950-1057) and again in boundary_estep
(src3d/boundary.c:691-887, whose E correction the component's flag skips) --
and applies the inline correction with the dt/eps of the material's eps/sigma
instead of the ADE update's own curl coefficient, with no correction at all on
the x0 face (its cb is zero there).  With a Drude metal reaching into the
slabs (the rough-surface workload: the metal half space crosses the side and
bottom slabs) this grows without bound once the pulse has reached the metal:
the oracle, which restates the reference, shows it at 48^3 (CPML 8) from step ~2500.

ADE_CPML_CONSISTENT 1 (no reference counterpart) updates each such psi once,
applies it with c0, and corrects x0 too; elsewhere it changes nothing.  Pins:
  * the reference scheme grows by orders of magnitude after the pulse, the
    consistent one decays, on a rough and on a flat interface;
  * pools without dispersive cells in the slabs are bit-identical in both
    modes (the keyword only touches dispersive components);
  * on the GPU the engine's consistent mode equals the oracle's.
"""
import numpy as np
import pytest

import cases
from oracle_lib import Oracle


def peak_then_late(par, t_peak, t_late, depth):
    """max |E| over the grid at t_peak and at t_late"""
    o = Oracle(par)
    o.step(t_peak)
    peak = max(np.abs(o.field(c)).max() for c in range(3))
    o.step(t_late - t_peak)
    late = max(np.abs(o.field(c)).max() for c in range(3))
    o.close()
    return peak, late


@pytest.mark.parametrize("rms_nm", [20.0, 0.001])
def test_reference_scheme_grows_and_consistent_scheme_decays(tmp_path, rms_nm):
    n, depth = 48, 8
    res = {}
    for mode in (False, True):
        par = cases.rough_surface(str(tmp_path / ("m%d" % mode)), n=n, depth=depth, nsteps=3250, nfff=False,
                                  rms_nm=rms_nm, ade_cpml_consistent=mode)
        res[mode] = peak_then_late(par, 2000, 3250, depth)
    print("rms %g nm: reference peak %.3e -> %.3e, consistent peak %.3e -> %.3e"
          % (rms_nm, res[False][0], res[False][1], res[True][0], res[True][1]))
    assert res[False][1] > 100 * res[False][0]
    assert res[True][1] < 0.05 * res[True][0]


@pytest.mark.parametrize("kind", ["spheres_cpml", "solids_cpml"])
def test_consistent_mode_leaves_pools_without_slab_dispersive_cells_alone(tmp_path, kind):
    """spheres_cpml has no dispersive material; solids_cpml keeps its Drude
    cone out of the slabs"""
    out = []
    for mode in (0, 1):
        par = cases.small_case(str(tmp_path / ("%s%d" % (kind, mode))), kind, nsteps=120)
        if mode:
            with open(par, "a") as f:
                f.write("ADE_CPML_CONSISTENT\n1\n\n")
        o = Oracle(par)
        o.step()
        out.append(o.fields())
        o.close()
    assert np.array_equal(out[0], out[1])


def test_consistent_mode_changes_dispersive_slab_cells(tmp_path):
    out = []
    for mode in (0, 1):
        par = cases.small_case(str(tmp_path / ("a%d" % mode)), "ade_cpml", nsteps=160)
        if mode:
            with open(par, "a") as f:
                f.write("ADE_CPML_CONSISTENT\n1\n\n")
        o = Oracle(par)
        o.step()
        out.append(o.fields())
        o.close()
    d = np.abs(out[0] - out[1]).max() / np.abs(out[0]).max()
    print("ade_cpml: consistent vs reference scheme, max difference / field scale", d)
    assert d > 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["ade_cpml", "tsf_ade_cpml", "rough"])
def test_gpu_consistent_mode_against_oracle(tmp_path, kind):
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors
    if kind == "rough":
        par = cases.rough_surface(str(tmp_path / kind), n=48, depth=8, nsteps=2400, nfff=False,
                                  ade_cpml_consistent=True)
    else:
        par = cases.small_case(str(tmp_path / kind), kind, nsteps=300)
        with open(par, "a") as f:
            f.write("ADE_CPML_CONSISTENT\n1\n\n")
    print("ADE_CPML_CONSISTENT", kind, flush=True)
    o = Oracle(par)
    o.step()
    ref = o.fields()
    o.close()
    s = Session(par)
    s.run()
    got = s.fields()
    s.close()
    errs = field_errors(got, ref)
    print(kind, errs)
    assert max(errs.values()) <= 1e-5
