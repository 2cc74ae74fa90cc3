"""Flux spectra (FLUX_DFT) and domain-periodic x / y boundaries -- extensions
with no reference counterpart, used by the textured thin-film workload (per-
layer absorbed power under a normally incident plane wave in a periodic cell).

Pins (CPU, on the oracle): the flux accumulators equal a numpy DFT of point
traces recorded at the same cells; in vacuum the flux through two planes is
the same (energy conservation) and equals |E|^2 / (2 eta0) times the area for
the plane wave (so the sign, the factor 1/2 and the conjugation are right); a
periodic pool keeps a normally incident plane wave uniform across x and y, and
moving a scatterer by s cells moves the fields by s cells exactly (the wrap
joins the far plane to the near plane).  The engine is then held to the
oracle on the GPU.
"""
import math
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

ETA0 = 376.730313668
DX = 5e-9
DT = 1.0 / 299792458.0 / math.sqrt(3.0 / DX ** 2)


def flux_from_acc(acc, nplanes, nw):
    a = acc.reshape(nplanes, -1, nw, 8)
    return 0.5 * DX * DX * DT * DT * (a[..., 0] * a[..., 6] + a[..., 1] * a[..., 7]
                                        - a[..., 2] * a[..., 4] - a[..., 3] * a[..., 5]).sum(axis=1)


def test_flux_accumulators_equal_dft_of_traces(tmp_path):
    pts = [(5, 7, 40), (5, 7, 39), (9, 3, 70), (9, 3, 69)]
    par = cases.plane_wave_flux(str(tmp_path / "c"), nsteps=900, points=pts)
    o = Oracle(par)
    o.step()
    acc = o.flux_acc().reshape(2, 16, 16, 3, 8)       # plane, j, i, wavelength, 8
    tr = o.points()                                   # [npts][nsteps+1][6]
    lam = np.array([450e-9, 500e-9, 550e-9])
    w = 2 * np.pi * 299792458.0 / lam
    n = np.arange(900)
    pe = np.exp(-1j * np.outer(w, (n + 1) * DT))
    ph = np.exp(-1j * np.outer(w, (n + 0.5) * DT))
    for p, (top, below) in enumerate(((0, 1), (2, 3))):
        i, j = pts[top][0], pts[top][1]
        ex, ey = tr[top][:900, 0], tr[top][:900, 1]
        hx = 0.5 * (tr[top][:900, 3] + tr[below][:900, 3])
        hy = 0.5 * (tr[top][:900, 4] + tr[below][:900, 4])
        ref = np.stack([pe @ ex, pe @ ey, ph @ hx, ph @ hy], axis=1)      # [w][4]
        got = acc[p, j, i].reshape(3, 4, 2)
        got = got[..., 0] + 1j * got[..., 1]
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print("plane", p, "accumulator vs numpy DFT of traces", err)
        assert err < 1e-10


def test_plane_wave_flux_conserved_and_closed_form(tmp_path):
    o = Oracle(cases.plane_wave_flux(str(tmp_path / "c")))
    o.step()
    acc = o.flux_acc()
    F = flux_from_acc(acc, 2, 3)
    a = acc.reshape(2, -1, 3, 8)
    e2 = 0.5 * DX * DX * DT * DT * (a[..., 0] ** 2 + a[..., 1] ** 2 + a[..., 2] ** 2 + a[..., 3] ** 2).sum(axis=1) / ETA0
    print("flux", F, "ratio to |E|^2/(2 eta0)", F / e2)
    assert np.all(F > 0)                                  # the wave travels towards +z
    assert np.abs(F[0] - F[1]).max() / np.abs(F).max() < 1e-4
    assert np.all(np.abs(F / e2 - 1) < 2e-3)              # Yee impedance error at 40+ cells per wavelength


def test_periodic_pool_keeps_plane_wave_uniform_and_translates(tmp_path):
    base = cases.plane_wave_flux(str(tmp_path / "u"), nsteps=700)
    o = Oracle(base)
    o.step()
    f = o.fields()
    assert np.abs(f).max() > 1e-3
    assert np.array_equal(f, np.broadcast_to(f[:, :1, :1, :], f.shape))
    fields = []
    for s, x in ((0, 5), (1, 10)):
        par = cases.plane_wave_flux(str(tmp_path / ("t%d" % s)), nsteps=700)
        with open(os.path.join(os.path.dirname(par), "sph.txt"), "w") as fh:
            fh.write("4 %d 6 60 3.5 1 4 1 0 0\n" % x)
        with open(par, "a") as fh:
            fh.write("MEDIUM_VECTOR\nsph.txt\n\n")
        o = Oracle(par)
        o.step()
        fields.append(o.fields())
    moved = np.roll(fields[0], 5, axis=1)
    print("scattered field max", np.abs(fields[0] - f).max())
    assert np.abs(fields[0] - f).max() > 1e-3
    assert np.array_equal(moved, fields[1])


def test_film_media_are_passive_and_absorb(tmp_path):
    """The textured film's dispersive media: Im chi >= 0 at every frequency up
    to 1e18 rad/s (chi as the ADE line defines it; a critical point with phi = 0
    reduces to the textbook Lorentz form), and a 12-cell slab of each in the
    periodic plane-wave cell absorbs (0 < flux behind < flux before) with the
    fields decayed once the pulse has passed.  Media with a gain band fail the
    last two: earlier synthetic lines with large critical-point phases left
    |E| ~ 0.75 after 4000 steps and a negative flux."""
    w = cases.PASSIVITY_OMEGA
    a, om, g = 1.3, 4e15, 3e14
    lor = cases.ade_susceptibility("7 1 0 0 0 %g 0 %g %g 0 0 0 0 0 0 0 0" % (a, om, g), w)
    assert np.allclose(lor, 2 * a * om ** 2 / (om ** 2 + g ** 2 - w ** 2 - 2j * g * w), rtol=1e-10)
    cases.textured_film(str(tmp_path / "tf"), nx=8, ny=8, nz=150, nsteps=10, nw=4, au=10, si=30, sio2=8)
    lines = [ln.split(None, 7)[7] for ln in open(str(tmp_path / "tf" / "mats.txt"))]
    ade = [ln for ln in lines if ln.split()[0] in ("5", "7")]
    assert len(ade) == 2
    for q, line in enumerate(ade):
        assert cases.ade_susceptibility(line, w).imag.min() >= 0, line
        par = cases.plane_wave_flux(str(tmp_path / ("s%d" % q)), n=4, nsteps=4000, slab=(50, 62, line),
                                    planes=(40, 70))
        o = Oracle(par)
        o.step()
        F = flux_from_acc(o.flux_acc(), 2, 3)
        tail = float(np.abs(o.fields()[:3]).max())
        print(line.split()[0], "flux before / behind the slab", F, "max |E| at the end", tail)
        assert np.all(F[1] > 0) and np.all(F[1] < F[0]) and tail < 1e-2
        o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("slab", [None, (50, 62, "7 2 0 1.2e16 1e14 0.8 -1.2 4e15 4e14 1.5 -1.1 4.2e15 2.4e15 "
                                              "1.1 0 3.3e15 5e14")])
def test_flux_and_periodic_engine_against_oracle(tmp_path, slab):
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors
    par = cases.plane_wave_flux(str(tmp_path / "c"), nsteps=1500, slab=slab, planes=(40, 70, 90))
    print("slab", slab, flush=True)
    o = Oracle(par)
    o.step()
    with Session(par) as s:
        s.run()
        got = s.fields()
        fl, acc = s.flux(with_acc=True)
    ref = o.fields()
    errs = field_errors(got, ref)
    racc = o.flux_acc()
    e_acc = float(np.linalg.norm(acc.reshape(-1) - racc) / np.linalg.norm(racc))
    rfl = flux_from_acc(racc, 3, 3)
    e_fl = float(np.abs(fl - rfl).max() / np.abs(rfl).max())
    print("fields", errs, "flux accumulators", e_acc, "flux spectra", e_fl, "\nflux", fl)
    assert max(errs.values()) <= 1e-5 and e_acc <= 1e-5 and e_fl <= 1e-5
    if slab:
        assert np.all(fl[1] < fl[0])          # the dispersive slab between the planes absorbs


@pytest.mark.gpu
def test_textured_film_reduced_against_oracle(tmp_path):
    """The thin-film workload (periodic x / y, CPML z, SiO2 on pyramid-textured
    three-pole silicon on a Drude + 2 CP back contact, flux planes) at 40 x 40 x
    150 for 600 steps; maps bit-exact."""
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors
    par = cases.textured_film(str(tmp_path / "tf"), nx=40, ny=40, nz=150, nsteps=600, nw=4, au=10, si=30, sio2=8)
    o = Oracle(par)
    o.step()
    with Session(par) as s:
        assert s.info()["ade_poles"] == 3
        s.run()
        got = s.fields()
        fl, acc = s.flux(with_acc=True)
        ms, bs = s.maps()
    errs = field_errors(got, o.fields())
    racc = o.flux_acc()
    e_acc = float(np.linalg.norm(acc.reshape(-1) - racc) / np.linalg.norm(racc))
    print("textured film: fields", errs, "flux accumulators", e_acc)
    assert max(errs.values()) <= 1e-5 and e_acc <= 1e-5
    mo, bo = o.maps()
    assert np.array_equal(mo, ms) and np.array_equal(bo, bs)
