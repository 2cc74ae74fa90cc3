"""External plane source (SOURCE_EXT; src3d/settings.c:4147-4186,
src3d/source.c:1340-1399, 2206-2266).

Every step the planes <base>_%.4d of step - shift are read (extyres rows of
extxres values each); the window [iextfrom, iextto) x [jextfrom, jextto) drives
the z plane k from (ijstart, jkstart) on: E on plane k from the stored H
values (ex += dt/dx/eps hy, ey -= dt/dx/eps hx), H on plane k - 1 from the
stored E values (hx -= dt/dx/mu0 ey, hy += dt/dx/mu0 ex); a step whose files
are missing is skipped.  Pins (oracle): at the first step with files the H
field equals its injection exactly and the E injection (E minus the vacuum
update of that H) equals the closed form over a dielectric-spanning window,
the missing steps inject nothing; the window, file layout, shift, planes and
signs all enter.  GPU vs oracle at 1e-5.
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

EPS0 = 8.854187817620389850e-12
MU0 = 1.256637061435917295e-6
N, DX = 40, 10e-9
XR, YR = 14, 12                     # stored plane size
WIN = (2, 1, 12, 11)                # iextfrom jextfrom iextto jextto
START = (10, 12)                    # ijstart jkstart
K, SHIFT = 20, 3


def write_case(folder, nsteps, extra=""):
    os.makedirs(folder, exist_ok=True)
    rng = np.random.default_rng(5)
    pat = rng.standard_normal((6, YR, XR))
    for s in range(nsteps - SHIFT):
        amp = np.sin(0.3 * (s + 1)) * np.exp(-((s - 25) / 15.0) ** 2)
        for c, base in enumerate(["ex", "ey", "ez", "hx", "hy", "hz"]):
            scale = 1.0 if c < 3 else 1.0 / 376.73
            np.savetxt(os.path.join(folder, "%s_%.4d" % (base, s)), (amp * scale * pat[c]).astype(np.float32), fmt="%.7g")
    with open(os.path.join(folder, "mats.txt"), "w") as f:
        f.write("8 14 5 16 30 30 24 1 3.0 1 0 0\n")
    p = os.path.join(folder, "ext.par")
    with open(p, "w") as f:
        f.write("POOL\n%d %d %d %g %g %g\n\nCOMP\n%d\n\nDT_MULT\n0.5\n\nMEDIUM_VECTOR\nmats.txt\n\n" % (N, N, N, DX, DX, DX, nsteps))
        f.write("SOURCE_EXT\n-1 -1 %d %d %d %d %d %d %d %d %d %d ex ey ez hx hy hz\n\n"
                % ((K,) + START + (XR, YR) + WIN + (SHIFT,)))
        f.write(extra)
    return p


def planes(folder, base, s):
    return np.loadtxt(os.path.join(folder, "%s_%.4d" % (base, s)), dtype=np.float32).astype(np.float64)


def test_first_injection_matches_the_closed_form(tmp_path):
    folder = str(tmp_path / "x")
    o = Oracle(write_case(folder, 10))
    dt = o.dt()
    for s in range(SHIFT):
        o.step(1)
        assert all(np.abs(f).max() == 0 for f in o.fields())   # files -3 .. -1 are missing: skipped
    o.step(1)
    f = o.fields()
    mat, _ = o.maps()
    o.close()
    i0, j0, i1, j1 = WIN
    ni, nj = i1 - i0, j1 - j0
    si, sj = START
    ex, ey = planes(folder, "ex", 0), planes(folder, "ey", 0)
    hx, hy = planes(folder, "hx", 0), planes(folder, "hy", 0)
    hcor, ecor0 = dt / DX / MU0, dt / DX / EPS0
    want_hx = np.zeros((N, N, N)); want_hy = np.zeros((N, N, N))
    want_hx[si:si + ni, sj:sj + nj, K - 1] = -hcor * ey[j0:j1, i0:i1].T
    want_hy[si:si + ni, sj:sj + nj, K - 1] = hcor * ex[j0:j1, i0:i1].T
    assert np.array_equal(f[3], want_hx) and np.array_equal(f[4], want_hy) and not np.any(f[5])
    # E = vacuum update of that H (+ the stored H values on plane k over eps)
    cb = dt / EPS0 / DX
    ex_upd = np.zeros((N, N, N)); ey_upd = np.zeros((N, N, N))
    ex_upd[:, 1:, 1:] = -(f[4][:, 1:, 1:] - f[4][:, 1:, :-1])
    ey_upd[1:, :, 1:] = f[3][1:, :, 1:] - f[3][1:, :, :-1]
    eps_r = np.where(mat[si:si + ni, sj:sj + nj, K] == 1, 3.0, 1.0)
    assert (eps_r == 3.0).any() and (eps_r == 1.0).any()
    inj_x = f[0][si:si + ni, sj:sj + nj, K] - (cb / np.where(mat[si:si + ni, sj:sj + nj, K] == 1, 3.0, 1.0)) * ex_upd[si:si + ni, sj:sj + nj, K]
    inj_y = f[1][si:si + ni, sj:sj + nj, K] - (cb / eps_r) * ey_upd[si:si + ni, sj:sj + nj, K]
    np.testing.assert_allclose(inj_x, ecor0 / eps_r * hy[j0:j1, i0:i1].T, rtol=1e-5, atol=1e-9 * np.abs(inj_x).max())
    # Ey's update coefficient is averaged over the x interface of the block (i = 14): leave that column out
    keep = mat[si - 1:si + ni - 1, sj:sj + nj, K] == mat[si:si + ni, sj:sj + nj, K]
    assert keep.sum() == ni * nj - nj
    np.testing.assert_allclose(inj_y[keep], (-ecor0 / eps_r * hx[j0:j1, i0:i1].T)[keep], rtol=1e-5,
                               atol=1e-9 * np.abs(inj_y).max())


def test_region_outside_the_pool_is_refused(tmp_path):
    folder = str(tmp_path / "r")
    par = write_case(folder, 5)
    s = open(par).read().replace("-1 -1 %d %d %d" % (K, START[0], START[1]), "-1 -1 %d %d %d" % (K, 35, START[1]))
    open(par, "w").write(s)
    with pytest.raises(RuntimeError):
        Oracle(par)


@pytest.mark.gpu
def test_gpu_ext_source_against_oracle(tmp_path):
    from gsvit_copy_amd import Session
    par = write_case(str(tmp_path / "g"), 90, "BOUNDARY_ALL\ncpml 6 3 -1 0.03 4\n\n")
    print("case external plane source", flush=True)
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
        print("external source GPU vs oracle rel L2", e)
        assert e <= 1e-5
