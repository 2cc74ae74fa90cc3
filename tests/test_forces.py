"""Force boxes (OUT_FORCE, src3d/output.c:355-512 and the log at :908-938).

The oracle's restatement is pinned against an independent numpy statement of
the textbook Maxwell stress tensor, T = eps E E + mu H H - (eps E^2 + mu H^2)/2,
with the local absolute eps and mu (here vacuum and the case's dielectric
spheres), applied to the oracle's own fields at several steps and summed over
the six faces with the reference's normals (+x at i0, -x at in, ... : they
point into the box).  A dropped off-diagonal term, a wrong half or a swapped
normal fails.  The GPU kernel and the log writer (force x dx dy and its mean
over the preceding 5 lambda_center / (c dt) steps, "%g") are checked against
the oracle on the GPU.
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

BOXES = [(10, 14, 14, 14, 34, 34, 34, "force0"), (0, 20, 18, 22, 30, 33, 31, "force1")]


def force_case(folder, nsteps=160):
    par = cases.small_case(folder, "spheres_cpml", nsteps=nsteps)
    with open(par, "a") as f:
        f.write("SOURCE_WAVELENGTH\n100e-9 150e-9 200e-9\n\n")
        for b in BOXES:
            f.write("OUT_FORCE\n%d %d %d %d %d %d %d %s\n\n" % b)
    return par


def numpy_force(F, eps, mu, box):
    """textbook stress tensor summed over the box faces, inward normals"""
    _, i0, j0, k0, i1, j1, k1, _ = box
    E, H = F[:3], F[3:]
    tot = np.zeros(3)
    for axis, (a0, a1) in enumerate(((i0, i1), (j0, j1), (k0, k1))):
        for pos, sign in ((a0, 1.0), (a1, -1.0)):
            sl = [slice(i0, i1), slice(j0, j1), slice(k0, k1)]
            sl[axis] = pos
            e = [c[tuple(sl)] for c in E]
            h = [c[tuple(sl)] for c in H]
            ep, m = eps[tuple(sl)], mu[tuple(sl)]
            w = (ep * sum(c * c for c in e) + m * sum(c * c for c in h)) / 2.0
            for r in range(3):
                t = ep * e[r] * e[axis] + m * h[r] * h[axis] - (w if r == axis else 0.0)
                tot[r] += sign * t.sum()
    return tot


def test_oracle_forces_match_textbook_stress_tensor(tmp_path):
    par = force_case(str(tmp_path / "f"))
    o = Oracle(par)
    mats, _ = o.maps()
    lines = [ln.split() for ln in open(os.path.join(str(tmp_path / "f"), "mats.txt"))]
    eps0, mu0 = 8.854187817e-12, 4e-7 * np.pi
    table = [1.0] + [float(v[6]) for v in lines if v[0] == "4"]
    eps = np.array(table)[np.maximum(mats, 0)] * eps0
    eps[4:12, 4:12, 4:10] = 2.5 * eps0                 # the case's linear box sets its voxels directly
    mu = np.full(eps.shape, mu0)
    done = 0
    for stop in (40, 100, 159):
        o.step(stop + 1 - done)
        done = stop + 1
        F = o.fields()
        got = o.forces(2)
        for b in range(2):
            want = numpy_force(F, eps, mu, BOXES[b])
            e = float(np.abs(got[b, stop] - want).max() / np.abs(want).max())
            print("step", stop, "box", b, got[b, stop], want, e)
            # the planner keeps material permittivities in float as the reference
            # does (per-voxel cubes and material rows), numpy uses exact doubles
            assert np.abs(want).max() > 0 and e < 1e-6
    o.close()


@pytest.mark.gpu
def test_gpu_forces_and_log_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    folder = str(tmp_path / "g")
    par = force_case(folder)
    print("force boxes", flush=True)
    o = Oracle(par)
    o.step()
    ref = o.forces(2)
    o.close()
    s = Session(par, write_outputs=True)
    s.run()
    got = s.forces()
    assert got.shape[0] == 2
    s.close()
    for b in range(2):
        for c in range(3):
            e = float(np.linalg.norm(got[b, :, c] - ref[b, :, c]) / np.linalg.norm(ref[b, :, c]))
            print("box", b, "component", c, "GPU vs FP64 oracle", e)
            assert e <= 1e-5
    # the log of box 0 (written every 10 steps and at the end); box 1 has step 0: no log
    log = np.loadtxt(os.path.join(folder, "force0"))
    assert not os.path.exists(os.path.join(folder, "force1"))
    n = ref.shape[1] - 1
    area = 10e-9 * 10e-9
    assert log.shape == (n - 1, 7) and np.array_equal(log[:, 0], np.arange(1, n))
    want = ref[0, 1:n] * area
    assert np.abs(log[:, 1:4] - want).max() <= 1e-5 * np.abs(want).max()
    dtv = 10e-9 / 299792458.0 / np.sqrt(3.0) * 0.5
    wsteps = 5.0 * 150e-9 / dtv / 299792458.0
    means = np.array([ref[0, int(max(0.0, i - wsteps)):i].mean(axis=0) for i in range(1, n)]) * area
    assert np.abs(log[:, 4:7] - means).max() <= 1e-5 * np.abs(means).max()


def radiation_case(folder, slab, nsteps=2800):
    """A normally incident pulse (+z, injected at k = 20) in a 12 x 12 periodic
    cell onto a dielectric slab (eps_r 12, k = 70..76, about a quarter
    wavelength thick at 500 nm, so that it reflects strongly) filling the cell,
    or no slab; a force box around the slab; traces in the scattered-field
    region (k = 14, reflected field only), in front of the slab (k = 64) and
    behind it (k = 90)"""
    n, k0, k1 = 12, 70, 77
    par = cases.plane_wave_flux(folder, n=n, nz=140, nsteps=nsteps, planes=(),
                                slab=(k0, k1, "0 12 1 0 0") if slab else None,
                                points=[(n // 2, n // 2, 14), (n // 2, n // 2, k0 - 6), (n // 2, n // 2, 90)])
    with open(par, "a") as f:
        f.write("OUT_FORCE\n1 2 2 %d %d %d %d force\n\n" % (k0 - 6, n - 2, n - 2, k1 + 4))
    return par, n


def test_radiation_pressure_momentum_balance(tmp_path):
    """Momentum balance, independent of the stress-tensor code: the impulse a
    pulse gives a slab is (incident + reflected - transmitted fluence) / c per
    area, the fluences being integrals of E^2 / eta0 of the plane waves
    (incident: in front of the slab's position without the slab; reflected:
    in the scattered-field region; transmitted: behind the slab), and the
    same box with no slab gets no net impulse.  The reference sums T . n over
    the faces with inward normals, so the impulse along +z comes out negative;
    the transverse components vanish.  The box's front face sees the standing
    wave of incident and reflected pulse, whose interference term the
    staggered E and H sample half a cell and half a step apart; the impulse,
    a difference of the two faces' momentum fluxes, comes out 2.4% above the
    fluence balance here (2.0% with an eps_r 4 slab, whose impulse is a
    smaller part of the face fluxes), hence the 4% bound.  A dropped 1/2 in the
    energy term, a dropped H term or a flipped normal moves it by 50-200%.
    The empty box stays below 1e-5 of the impulse: momentum conservation."""
    c0, eta0 = 299792458.0, 376.730313668
    out = {}
    for slab in (True, False):
        par, n = radiation_case(str(tmp_path / ("s%d" % slab)), slab)
        o = Oracle(par)
        o.step()
        dx, dt = 5e-9, o.dt()
        F = o.forces(1)[0] * dx * dx                 # the log's dx dy factor
        flu = [float((p[:, :3] ** 2).sum()) / eta0 * dt for p in o.points()[-3:]]
        out[slab] = (F.sum(axis=0) * dt, flu)
        o.close()
    (imp, (f_ref, _, f_tr)), (imp0, (_, f_inc, _)) = out[True], out[False]
    area = (n - 4) ** 2 * dx * dx
    want = -(f_inc + f_ref - f_tr) / c0 * area
    print("fluences: incident %.4g reflected %.4g transmitted %.4g" % (f_inc, f_ref, f_tr))
    print("impulse", imp, "expected z", want, "empty box", imp0)
    assert f_ref > 0.1 * f_inc and f_tr > 0.1 * f_inc          # both waves matter here
    assert abs(imp[2] - want) <= 4e-2 * abs(want)
    assert abs(imp[0]) <= 1e-6 * abs(want) and abs(imp[1]) <= 1e-6 * abs(want)
    assert np.abs(imp0).max() <= 1e-4 * abs(want)
