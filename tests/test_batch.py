"""Batched runs: several independent sessions of uneven grid sizes stepped
together (gsv_batch_*), one launch per kernel type for all of them.

Bar: a batched run's six field components equal the same run stepped alone
bit for bit (the batch runs the same per-cell arithmetic), for every launch
shape, workgroup count and graph capture length; one batched run is also
checked against the FP64 oracle to relative L2 <= 1e-5.  The sweep generator
itself is checked on the CPU.
"""
import os

import numpy as np
import pytest

import cases

SHAPES = [(48, 40, 56), (80, 48, 40), (40, 72, 48), (64, 56, 64)]


def small_sweep(folder, nsteps=120):
    """Four small rough-surface TSF runs of uneven sizes (CPML 6)."""
    paths = []
    rng = np.random.default_rng(3)
    for r, shape in enumerate(SHAPES):
        paths.append(cases.sweep_run(os.path.join(folder, "run%d" % r), shape, seed=100 + r,
                                     rms_nm=float(rng.uniform(5, 30)), theta_deg=float(rng.uniform(0, 70)),
                                     lam_center=float(rng.uniform(450e-9, 750e-9)), nsteps=nsteps, depth=6,
                                     tsf_margin=3))
    return paths


def test_sweep_params_are_seeded_and_in_range():
    a = cases.sweep_params(64, seed=11)
    b = cases.sweep_params(64, seed=11)
    assert a == b
    assert cases.sweep_params(64, seed=12) != a
    for p in a:
        assert all(e in cases.SWEEP_EDGES for e in p["shape"])
        assert 5 <= p["rms_nm"] <= 30 and 0 <= p["theta_deg"] <= 70 and 450e-9 <= p["lam_center"] <= 750e-9
    # every edge value occurs on every axis over 64 runs
    for ax in range(3):
        assert {p["shape"][ax] for p in a} == set(cases.SWEEP_EDGES)


def test_sweep_run_writes_uneven_grid(tmp_path):
    par = cases.sweep_run(str(tmp_path / "r"), (64, 96, 128), seed=5, rms_nm=10.0, theta_deg=30.0,
                          lam_center=600e-9, nsteps=50)
    text = open(par).read()
    assert "POOL\n64 96 128 " in text
    assert "cpml 10 3" in text and "SOURCE_TSF" in text and "NFFF" not in text
    hdr = open(tmp_path / "r" / "surface.vtk", "rb").read(400).decode(errors="replace")
    assert "DIMENSIONS 64 96 128" in hdr


def _single(paths):
    """each run alone, through the same per-cell column functions as the
    batched kernels"""
    from gsvit_copy_amd import Session
    out = []
    for p in paths:
        with Session(p) as s:
            s.run()
            out.append(s.fields())
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("launch,graph", [
    (None, 0),
    ([128, 2, 8, 64, 8, 32, 0, 0], 8),
    ([64, 1, 128, 256, 2, 16, 7, 5], 5),
])
def test_batch_equals_single_runs(tmp_path, launch, graph):
    from gsvit_copy_amd import Batch, Session
    paths = small_sweep(str(tmp_path))
    print("batch launch", launch, "graph", graph, flush=True)
    ref = _single(paths)
    sessions = [Session(p) for p in paths]
    with Batch(sessions) as b:
        if launch:
            b.set_launch(launch)
            assert b.get_launch() == launch
        b.set_graph(graph)
        assert b.run(50) == 50
        assert b.run(1000) == 70          # stops at the runs' step count
        assert b.run(10) == 0
        inf = b.info()
        assert inf["tsf_runs"] == len(paths) and inf["face_tiles"] > 0
    for r, s in enumerate(sessions):
        got = s.fields()
        assert s.info()["step"] == 120
        for c in range(6):
            assert np.array_equal(got[c], ref[r][c]), (r, c, float(np.abs(got[c] - ref[r][c]).max()))
        assert np.abs(got).max() > 0
        s.close()


@pytest.mark.gpu
def test_batch_continues_single_steps_and_autotunes(tmp_path):
    from gsvit_copy_amd import Batch, Session
    paths = small_sweep(str(tmp_path), nsteps=160)
    ref = _single(paths)
    sessions = [Session(p) for p in paths]
    start = [0, 17, 3, 0]                     # runs may start from different steps
    for s, n in zip(sessions, start):
        s.run(n)
    with Batch(sessions) as b:
        used, best = b.autotune(1)
        print("autotune used", used, "best", best, flush=True)
        assert used > 0 and best == b.get_launch()
        b.set_graph(4)
        ran = b.run(1000)                     # until the furthest run reaches its step count
        assert used + ran == 160 - 17
    for r, s in enumerate(sessions):
        assert s.info()["step"] == start[r] + used + ran
        s.run()                               # the rest alone
        got = s.fields()
        for c in range(6):
            assert np.array_equal(got[c], ref[r][c]), (r, c)
        s.close()


@pytest.mark.gpu
def test_batched_run_against_oracle(tmp_path):
    from gsvit_copy_amd import Batch, Session
    from oracle_lib import Oracle
    from test_gpu_parity import field_errors
    paths = small_sweep(str(tmp_path))
    sessions = [Session(p) for p in paths]
    with Batch(sessions) as b:
        b.set_graph(6)
        b.run(1000)
    o = Oracle(paths[1])
    o.step()
    errs = field_errors(sessions[1].fields(), o.fields())
    print("batched run 1 vs FP64 oracle", errs)
    assert max(errs.values()) <= 1e-5, errs
    for s in sessions:
        s.close()


@pytest.mark.gpu
def test_batch_rejects_unsupported_runs(tmp_path):
    from gsvit_copy_amd import Batch, Session
    paths = small_sweep(str(tmp_path))
    point = cases.small_case(str(tmp_path / "point"), "spheres_cpml")
    sessions = [Session(paths[0]), Session(point)]
    with pytest.raises(RuntimeError):
        Batch(sessions)
    with pytest.raises(RuntimeError):
        Batch([sessions[0], sessions[0]])
    with Batch([sessions[0]]) as b:
        with pytest.raises(ValueError):
            b.set_launch([96, 4, 16, 64, 4, 16, 0, 0])   # x width not a multiple of 64
    for s in sessions:
        s.close()


@pytest.mark.gpu
def test_batch_of_periodic_runs_equals_single_runs(tmp_path):
    """Domain-periodic runs (normal-incidence plane wave through a three-pole
    slab in periodic cells of two sizes) batched with a non-periodic Drude-only
    run.  A batch runs every grid through the widest build (here three poles,
    periodic); runs of that build equal their single runs bit for bit, the
    Drude-only run, whose single run uses the Drude-only non-periodic build,
    to the parity bar (the builds may contract the E-update sums into fused
    multiply-adds differently, and the dispersive recursion carries such
    last-bit differences: ~1e-6 after 300 steps); that run is also held to
    the FP64 oracle."""
    from oracle_lib import Oracle
    from test_gpu_parity import field_errors
    from gsvit_copy_amd import Batch, Session
    slab = "7 2 0 1.2e16 1e14 0.8 -1.2 4e15 4e14 1.5 -1.1 4.2e15 2.4e15 1.1 0 3.3e15 5e14"
    paths = [cases.plane_wave_flux(str(tmp_path / "p16"), n=16, nsteps=300, planes=(), slab=(50, 62, slab)),
             cases.plane_wave_flux(str(tmp_path / "p24"), n=24, nz=100, nsteps=300, planes=(), slab=(40, 50, slab)),
             small_sweep(str(tmp_path / "sw"), nsteps=300)[0]]
    ref = _single(paths)
    sessions = [Session(p) for p in paths]
    with Batch(sessions) as b:
        b.set_graph(4)
        assert b.run(1000) == 300
    for r, s in enumerate(sessions):
        got = s.fields()
        for c in range(6):
            if r < 2:
                assert np.array_equal(got[c], ref[r][c]), (r, c, float(np.abs(got[c] - ref[r][c]).max()))
            else:
                e = float(np.linalg.norm(got[c] - ref[r][c]) / np.linalg.norm(ref[r][c]))
                assert e <= 1e-5, (r, c, e)
        assert np.abs(got).max() > 0
        if r == 2:
            o = Oracle(paths[2])
            o.step()
            errs = field_errors(got, o.fields())
            print("mixed-build batched run vs FP64 oracle", errs)
            assert max(errs.values()) <= 1e-5
            o.close()
        s.close()
