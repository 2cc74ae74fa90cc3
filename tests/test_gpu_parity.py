"""GPU parity: the HIP engine (through the C-ABI session) against the CPU oracle.

Bar (floating point): every field component and every point trace agrees with
the FP64 oracle to relative L2 <= 1e-5 after the full step count; the error of
the FP32-storage oracle against the FP64 oracle is printed beside it.
"""
import numpy as np
import pytest

import cases
from oracle_lib import Oracle
from gsvit_copy_amd import Session
from selftest_recipe import write_selftest, read_golden
from test_running_dft import dft_case
import test_box_sums

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def field_errors(got, ref):
    """Relative L2 of the E and H vector fields, and per component relative to
    max(|component|, |family| / sqrt(3)) so that a component that is zero by
    symmetry is judged against the field it belongs to."""
    out = {}
    for name, sl in (("E", slice(0, 3)), ("H", slice(3, 6))):
        fam = np.linalg.norm(ref[sl])
        out[name] = float(np.linalg.norm(got[sl] - ref[sl]) / fam)
        for c in range(sl.start, sl.stop):
            den = max(np.linalg.norm(ref[c]), fam / np.sqrt(3))
            out["c%d" % c] = float(np.linalg.norm(got[c] - ref[c]) / den)
    return out


def trace_errors(got, ref):
    """per point, E and H families separately"""
    errs = []
    for n in range(ref.shape[0]):
        for sl in (slice(0, 3), slice(3, 6)):
            den = np.linalg.norm(ref[n][:, sl])
            if den > 0:
                errs.append(float(np.linalg.norm(got[n][:, sl] - ref[n][:, sl]) / den))
    return errs


def run_both(par, oracle_options=None):
    o = Oracle(par)
    for k, v in (oracle_options or {}).items():
        o.set_option(k, v)
    o.step()
    s = Session(par)
    s.run()
    return o, s


@pytest.mark.parametrize("kind", cases.SMALL_KINDS)
def test_small_case_fields_and_traces(kind, tmp_path):
    par = cases.small_case(str(tmp_path / kind), kind)
    print("case", kind, flush=True)
    o, s = run_both(par)
    ref, got = o.fields(), s.fields()
    errs = field_errors(got, ref)
    o32 = Oracle(par, "f32")
    o32.step()
    f32 = field_errors(o32.fields(), ref)
    print(kind, "GPU vs FP64", errs, "FP32 oracle vs FP64", f32)
    assert max(errs.values()) <= TOL, errs
    terr = trace_errors(s.points(), o.points())
    print(kind, "trace errors", terr)
    assert max(terr) <= TOL, terr
    if kind == "farfield_cpml":
        rf, gf = o.farfield(), s.farfield()
        e = rel_l2(gf, rf)
        print("time-domain far field rel L2", e)
        assert e <= TOL
    o.close()
    s.close()


@pytest.mark.parametrize("test_id", ["10000000", "10021100", "10022100", "10024100", "10026000", "10027000",
                                     "10028000", "10029000", "10035011"])
def test_gpu_reproduces_reference_selftest(test_id, tmp_path):
    """The reference's own golden traces (printed with %g), recipe src3d/tests.c:92-257."""
    par, nsteps = write_selftest(str(tmp_path / test_id), test_id)
    ff = int(test_id[6])
    s = Session(par)
    s.run()
    gold = np.array(read_golden(test_id, nsteps, ff))
    if ff:
        b = s.farfield()[0]
        got = np.sqrt(b[0] ** 2 + b[1] ** 2 + b[2] ** 2)[: len(gold)]
    else:
        got = s.points()[0][: len(gold), 0]
    e = rel_l2(got, gold)
    print(test_id, "GPU vs golden rel L2", e)
    assert e <= 2e-5
    s.close()


def test_sphere_scatter_reduced_against_oracle(tmp_path):
    """The 40-sphere dipole workload's geometry at 64^3, 12 spheres, 300 steps; maps bit-exact."""
    par = cases.sphere_scatter(str(tmp_path / "c1"), n=64, nsteps=300, nspheres=12)
    o, s = run_both(par)
    ref, got = o.fields(), s.fields()
    errs = field_errors(got, ref)
    print("sphere scatter 64^3", errs, s.info())
    assert max(errs.values()) <= TOL
    assert max(trace_errors(s.points(), o.points())) <= TOL
    mo, bo = o.maps()
    ms, bs = s.maps()
    assert np.array_equal(mo, ms) and np.array_equal(bo, bs)


def test_sphere_scatter_full_size_against_oracle(tmp_path):
    """The plain-FDTD correctness workload at its full size: 128^3, 10 nm cells,
    Courant 0.5, CPML 10, 40 dielectric spheres, Ez dipole, 1000 steps, all six
    field components and 16 point traces against the FP64 oracle; maps bit-exact."""
    par = cases.sphere_scatter(str(tmp_path / "c1full"))
    print("sphere scatter 128^3 x 1000 steps", flush=True)
    o, s = run_both(par)
    ref, got = o.fields(), s.fields()
    errs = field_errors(got, ref)
    terr = trace_errors(s.points(), o.points())
    print("sphere scatter 128^3: fields", errs, "traces max", max(terr), "points", len(terr) // 2)
    assert max(errs.values()) <= TOL
    assert len(terr) >= 16 and max(terr) <= TOL
    mo, bo = o.maps()
    ms, bs = s.maps()
    assert np.array_equal(mo, ms) and np.array_equal(bo, bs)
    o.close()
    s.close()


def test_rough_surface_reduced_against_oracle(tmp_path):
    """The rough-surface workload (Drude metal below a rough interface, oblique
    p-polarised TSF plane wave, CPML) at 64^3 with CPML 8, 400 steps; maps
    bit-exact.  The FP32-storage oracle (the reference's own update order with
    float state) is far less accurate here than the engine, whose ADE state is
    kept in increment form; both errors against FP64 are printed."""
    par = cases.rough_surface(str(tmp_path / "c2"), n=64, nsteps=400, depth=8, nfff=False)
    o, s = run_both(par)
    ref, got = o.fields(), s.fields()
    errs = field_errors(got, ref)
    o32 = Oracle(par, "f32")
    o32.step()
    f32 = field_errors(o32.fields(), ref)
    print("rough surface 64^3: GPU vs FP64", errs, "FP32-storage oracle vs FP64", f32)
    assert max(errs.values()) <= TOL
    mo, bo = o.maps()
    ms, bs = s.maps()
    assert np.array_equal(mo, ms) and np.array_equal(bo, bs)


def test_running_dft_against_oracle(tmp_path):
    """Running-DFT accumulators on all six NFFF box faces at two wavelengths."""
    par = dft_case(str(tmp_path / "dft"))
    o, s = run_both(par)
    ref, got = o.dft(), s.dft()
    assert ref.shape == got.shape and ref.size > 0
    e = rel_l2(got, ref)
    print("running DFT rel L2", e)
    assert e <= TOL
    o.close()
    s.close()


def test_box_sums_against_oracle(tmp_path):
    """OUT_SUM in every mode, every step; the log files in the reference format."""
    folder = tmp_path / "sums"
    par = test_box_sums.sums_case(str(folder))
    o = Oracle(par)
    o.step()
    s = Session(par, write_outputs=True)
    s.run()
    ref, got = o.sums(len(test_box_sums.SUMS)), s.sums()
    assert got.shape == ref.shape
    errs = [rel_l2(got[q], ref[q]) for q in range(ref.shape[0])]
    print("box sums rel L2", errs)
    assert max(errs) <= TOL
    vol = (10e-9) ** 3
    for q, (mode, _, _) in enumerate(test_box_sums.SUMS):
        rows = np.loadtxt(str(folder / ("sum%d" % q)))
        assert rows.shape == (ref.shape[1], 2)
        assert np.array_equal(rows[:, 0], np.arange(ref.shape[1]))
        scale = 1.0 if mode == "Max" else vol
        assert np.allclose(rows[:, 1], got[q] * scale, rtol=1e-5, atol=0)
    o.close()
    s.close()
