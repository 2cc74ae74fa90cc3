"""MEDIUM_CROP (src3d/settings.c:1898-1921, src3d/plan.c:5244-5285): voxels
on or outside the faces of the crop box take the crop medium.

Pins: with per-voxel data (MEDIUM_LINEAR) a crop equals writing the crop
medium into the voxel file on and outside the box faces, field for field
after a run to 1e-8 of the field scale (the two round eps0 x eps_r to float
differently; skipping the crop changes the fields by 1e-3); with tabulated materials the material map is unchanged
strictly inside the box and one new tabulated material everywhere else, and
that material's sigma is zero whatever the keyword gives (the reference adds
it with sigma = sigma* = 0): runs with crop sigma 0 and 1e4 are identical.
The tabulated crop material is stored in absolute units like every other
table row (the reference's multme, src3d/plan.c:2439-2442): a vacuum crop
(eps_r 1, mu_r 1) around a case is transparent, its fields equal the
uncropped run's to float round-off at the crop faces, shell included.  With MATMODE_CHECK and a
voxel file holding electric data only, the reference returns from its
loader before the crop (src3d/plan.c:4162-4165), so the keyword changes
nothing there.
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

CROP = (9, 8, 10, 38, 36, 39)


def outside(shape, box=CROP):
    i0, j0, k0, i1, j1, k1 = box
    i, j, k = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    return (i <= i0) | (i >= i1) | (j <= j0) | (j >= j1) | (k <= k0) | (k >= k1)


def run(par, steps=None):
    o = Oracle(par)
    o.step(steps or 10 ** 9)
    f, m = o.fields(), o.maps()[0]
    o.close()
    return f, m


def voxel_case(folder, crop, n=48, nsteps=120):
    par = cases.small_case(folder, "vacuum_none", n=n, nsteps=nsteps)
    rng = np.random.default_rng(5)
    eps = rng.uniform(1.0, 3.0, (n, n, n)).astype(np.float32)
    sig = rng.uniform(0.0, 1e3, (n, n, n)).astype(np.float32)
    if crop == "explicit":
        out = outside(eps.shape)
        eps[out] = np.float32(2.25)
        sig[out] = np.float32(300.0)
    cases.write_linear_medium(os.path.join(folder, "lin.bin"), eps, sigma=sig)
    with open(par, "a") as f:
        f.write("MEDIUM_LINEAR\nlin.bin\n\n")
        if crop == "keyword":
            f.write("MEDIUM_CROP\n%d %d %d %d %d %d 2.25 1 300 0\n\n" % CROP)
    return par


def test_crop_equals_explicit_voxel_data(tmp_path):
    a, _ = run(voxel_case(str(tmp_path / "kw"), "keyword"))
    b, _ = run(voxel_case(str(tmp_path / "ex"), "explicit"))
    c, _ = run(voxel_case(str(tmp_path / "no"), None))
    # the voxel file's eps is scaled by eps0 in float, the crop's in double and
    # then rounded (as the reference's (gfloat)(crop_epsilon * EPSILON_0)):
    # one float ulp of material apart
    scale = np.abs(b).max()
    print("crop vs explicit", np.abs(a - b).max() / scale, "crop vs none", np.abs(a - c).max() / scale)
    assert np.abs(a - b).max() <= 1e-8 * scale
    assert np.abs(a - c).max() > 1e-4 * scale


def tab_case(folder, sigma, crop=True, eps_r=2.25):
    par = cases.small_case(folder, "spheres_cpml", nsteps=120)
    if crop:
        with open(par, "a") as f:
            f.write("MEDIUM_CROP\n%d %d %d %d %d %d %g 1 %g 0\n\n" % (CROP + (eps_r, sigma)))
    return par


def shell_vacuum_case(folder, crop):
    """a dielectric sphere strictly inside the crop box, vacuum in the shell, CPML"""
    par = cases.small_case(folder, "vacuum_none", nsteps=120)
    with open(os.path.join(folder, "sphere.txt"), "w") as f:
        f.write("4 30 26 22 5 1 3 1 0 0\n")
    with open(par, "a") as f:
        f.write("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\nMEDIUM_VECTOR\nsphere.txt\n\n")
        if crop:
            f.write("MEDIUM_CROP\n%d %d %d %d %d %d 1 1 0 0\n\n" % CROP)
    return par


def test_vacuum_crop_of_tabulated_materials_is_transparent(tmp_path):
    fv, mv = run(shell_vacuum_case(str(tmp_path / "v"), True))
    fn, mn = run(shell_vacuum_case(str(tmp_path / "n"), False))
    out = outside(mv.shape)
    assert len(np.unique(mv[out])) == 1 and mv[out][0] > mn.max()   # a new material, not vacuum's index
    scale = np.abs(fn).max()
    shell = np.abs(fv[:, out]).max() / np.abs(fn[:, out]).max()
    print("vacuum crop vs none", np.abs(fv - fn).max() / scale, "shell amplitude ratio", shell)
    # the crop faces become internal boundaries (a new material index), where
    # the coefficients come from permittivities averaged in float: 1e-10 of the
    # scale.  With relative constants in the table the shell froze (ratio ~0).
    assert np.abs(fv - fn).max() <= 1e-8 * scale
    assert 0.99 < shell < 1.01


def test_crop_skipped_on_electric_voxel_file(tmp_path):
    def case(folder, crop):
        n = 40
        par = cases.small_case(folder, "vacuum_none", n=n, nsteps=80)
        eps = np.random.default_rng(3).uniform(1.0, 3.0, (n, n, n)).astype(np.float32)
        cases.write_linear_medium(os.path.join(folder, "lin.bin"), eps)
        with open(par, "a") as f:
            f.write("MATMODE_CHECK\n1\n\nMEDIUM_LINEAR\nlin.bin\n\n")
            if crop:
                f.write("MEDIUM_CROP\n9 8 10 30 28 31 2.25 1 300 0\n\n")
        return par
    a, _ = run(case(str(tmp_path / "kw"), True))
    b, _ = run(case(str(tmp_path / "no"), False))
    assert np.array_equal(a, b)


def test_crop_of_tabulated_materials(tmp_path):
    f0, m0 = run(tab_case(str(tmp_path / "a"), 0.0))
    f1, m1 = run(tab_case(str(tmp_path / "b"), 1e4))
    fn, mn = run(tab_case(str(tmp_path / "c"), 0.0, crop=False))
    out = outside(m0.shape)
    assert np.array_equal(m0[~out], mn[~out])
    new = np.unique(m0[out])
    assert len(new) == 1 and new[0] > mn.max()
    assert np.array_equal(f0, f1)                    # the crop material's sigma is not used
    assert not np.array_equal(f0, fn)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["voxel", "tabulated"])
def test_gpu_crop_against_oracle(tmp_path, kind):
    from gsvit_copy_amd import Session
    from test_gpu_parity import field_errors
    par = voxel_case(str(tmp_path / "v"), "keyword") if kind == "voxel" else tab_case(str(tmp_path / "t"), 50.0)
    print("crop", kind, flush=True)
    ref, _ = run(par)
    s = Session(par)
    s.run()
    errs = field_errors(s.fields(), ref)
    s.close()
    print(kind, errs)
    assert max(errs.values()) <= 1e-5
