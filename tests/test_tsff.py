"""Focused TSF (SOURCE_TSFF, TSFF_SHIFT, TSFF_SKIP; src3d/settings.c:4605-4697,
src3d/plan.c:1855-1979, src3d/source.c:167-194,773-930,1892-2060).

The focused source is nip x mip plane waves: Gauss-Legendre nodes in theta
over [0, thetamax], midpoints in phi, psi = phi + pol, each with its own
vacuum 1-D line, amplitude fs_amplitude and its phase origin at the pool
centre plus the shift.  Pins of the oracle:
* the planned waves against numpy's Gauss-Legendre rule and the amplitude
  formula (1e-13);
* a single wave (nip = mip = 1) with the shift chosen along the wave so that
  its line is read at the same distances as a plain TSF's equals acorr times
  that plain TSF run (whose faces are pinned by the reference's goldens and
  the leakage tests) to rounding, in E and H, everywhere;
* a focused beam's time-integrated |E|^2 peaks on the axis through the pool
  centre (its axial position, 6 cells before the centre in this geometry, is
  the reference's) and moves with TSFF_SHIFT, to a cell.
GPU vs oracle at 1e-5 with several waves, skipped faces and a shift.
"""
import math
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

C = 299792458.0
N, DX = 40, 10e-9
B0, B1 = 8, 31


def write_par(folder, body, nsteps=120, lam=150e-9, mode_line=None, extra=""):
    os.makedirs(folder, exist_ok=True)
    dt = cases._dt(DX, DX, DX, 0.5)
    cases.write_tsf_source(os.path.join(folder, "tsf.txt"), cases.gaussian_sine(nsteps, dt, lam, 1.5, delay_widths=1.2))
    p = os.path.join(folder, "tsff.par")
    with open(p, "w") as f:
        f.write("POOL\n%d %d %d %g %g %g\n\nCOMP\n%d\n\nDT_MULT\n0.5\n\n%s%s" % (N, N, N, DX, DX, DX, nsteps, body, extra))
    return p


def tsff_block(thetamax, fdist, pol, nip, mip, shift=None, skips=()):
    s = "SOURCE_TSFF\n%d %d %d %d %d %d %r %r %r %d %d 0 tsf.txt\n\n" % (B0, B0, B0, B1, B1, B1, thetamax, fdist, pol, nip, mip)
    if shift is not None:
        s += "TSFF_SHIFT\n%r %r %r\n\n" % tuple(float(v) for v in shift)
    for f in skips:
        s += "TSFF_SKIP\n%s\n\n" % f
    return s


def test_planned_waves_follow_gauss_legendre_and_the_amplitude_formula(tmp_path):
    tm, fd, pol, nip, mip = 1.1, 800.0, 0.25, 5, 7
    o = Oracle(write_par(str(tmp_path / "w"), tsff_block(tm, fd, pol, nip, mip), nsteps=4))
    w = o.tsff_waves()
    o.close()
    x, wt = np.polynomial.legendre.leggauss(nip)
    th = np.repeat((x + 1) * tm / 2, mip)
    ph = np.tile((np.arange(mip) + 0.5) * 2 * np.pi / mip, nip)
    corr = 1 / np.sqrt(np.sin(th) ** 4 * (np.cos(ph) ** 4 + np.sin(ph) ** 4) + np.cos(th) ** 4)
    acorr = 1e15 * np.repeat(wt * tm / 2, mip) * (2 * np.pi / mip) * (fd * DX / (2 * np.pi * C)) * np.sqrt(np.cos(th)) * np.sin(th)
    np.testing.assert_allclose(w[:, 0], th, rtol=1e-13)
    np.testing.assert_allclose(w[:, 1], ph, rtol=1e-13)
    np.testing.assert_allclose(w[:, 2], ph + pol, rtol=1e-13)
    np.testing.assert_allclose(w[:, 3], corr, rtol=1e-13)
    np.testing.assert_allclose(w[:, 4], acorr, rtol=1e-13)
    # the rule integrates polynomials of degree 2 nip - 1 exactly
    assert abs(np.sum(wt * x ** 8) - 2 / 9) < 1e-14


def test_single_wave_equals_the_plain_tsf_times_its_amplitude(tmp_path):
    tm, fd, pol = 0.8, 600.0, 0.3
    th, ph, ps = tm / 2, math.pi, math.pi + pol
    a = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    c = np.array([N // 2, N // 2, N // 2], dtype=float)
    # TSFF reads its line at corr (250 - a.(c + s - r)), the plain TSF at corr (10 + a.(r - (N, 0, 0)))
    t = 240 - a @ c + a[0] * N
    pf = write_par(str(tmp_path / "f"), tsff_block(tm, fd, pol, 1, 1, shift=t * a))
    pp = write_par(str(tmp_path / "p"), "SOURCE_TSF\n%d %d %d %d %d %d %r %r %r 0 tsf.txt\n\n" % (B0, B0, B0, B1, B1, B1, th, ph, ps))
    of, op = Oracle(pf), Oracle(pp)
    w = of.tsff_waves()
    assert np.allclose(w[0, :3], [th, ph, ps], rtol=1e-15)
    acorr = w[0, 4]
    of.step(); op.step()
    ff, fp = of.fields(), op.fields()
    of.close(); op.close()
    for q in range(6):
        scale = np.abs(fp[q]).max()
        err = np.abs(ff[q] - acorr * fp[q]).max() / (acorr * scale)
        print("component", q, "max |TSFF - acorr TSF| / max", err)
        assert scale > 0 and err < 1e-9


def focus_peak(folder, shift):
    body = tsff_block(0.9, 500.0, 0.0, 4, 8, shift=shift) + "BOUNDARY_ALL\ncpml 6 3 -1 0.03 4\n\n"
    # the phase origin sits 250 line cells from the line's source: the pulse
    # reaches the focus after about 250 / (c dt corr / dx) = 870 steps
    o = Oracle(write_par(folder, body, nsteps=1100, lam=100e-9))
    acc = np.zeros((N, N, N))
    for s in range(1100):
        o.step(1)
        if s >= 700:
            e = o.fields()[:3]
            acc += e[0] ** 2 + e[1] ** 2 + e[2] ** 2
    o.close()
    inner = np.zeros_like(acc)
    inner[B0 + 2:B1 - 1, B0 + 2:B1 - 1, B0 + 2:B1 - 1] = acc[B0 + 2:B1 - 1, B0 + 2:B1 - 1, B0 + 2:B1 - 1]
    return np.array(np.unravel_index(np.argmax(inner), inner.shape))


def test_focused_beam_peaks_on_axis_and_follows_its_shift(tmp_path):
    """the time-integrated |E|^2 of a 4 x 8-wave beam (NA 0.78, 10 cells per
    wavelength) peaks on the axis through the pool centre, and moves with
    TSFF_SHIFT; a sign error in the phase origin would move it the other way"""
    p0 = focus_peak(str(tmp_path / "f0"), (0, 0, 0))
    print("no shift: energy peak", p0)
    assert abs(p0[0] - N // 2) <= 1 and abs(p0[1] - N // 2) <= 1
    for shift in ((3, -2, 2), (0, 0, 5)):
        p = focus_peak(str(tmp_path / ("f%d%d%d" % shift)), shift)
        print("shift", shift, "energy peak", p, "moved by", p - p0)
        assert np.abs(p - p0 - np.array(shift)).max() <= 1


@pytest.mark.gpu
def test_gpu_tsff_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    body = tsff_block(0.7, 700.0, 0.4, 3, 5, shift=(1.5, -2.0, 0.75), skips=("jn",)) + "BOUNDARY_ALL\ncpml 6 3 -1 0.03 4\n\n"
    body += "MEDIUM_VECTOR\nmats.txt\n\n"
    folder = str(tmp_path / "g")
    par = write_par(folder, body, nsteps=1000)
    with open(os.path.join(folder, "mats.txt"), "w") as f:
        f.write("4 20 21 19 5 1 2.2 1 0 0\n")
    print("case tsff", flush=True)
    o = Oracle(par)
    o.step()
    ref = o.fields()
    o.close()
    s = Session(par)
    s.run()
    got = s.fields()
    s.close()
    for fam in (slice(0, 3), slice(3, 6)):
        e = float(np.sqrt(sum(np.sum((got[c] - ref[c]) ** 2) for c in range(fam.start, fam.stop))
                          / sum(np.sum(ref[c] ** 2) for c in range(fam.start, fam.stop))))
        print("TSFF GPU vs oracle rel L2", e)
        assert e <= 1e-5
