"""Periodic near-to-far-field transform (PERIODIC_NFFF*, src3d/pfarfield.c,
src3d/settings.c:3424-3499, log format src3d/output.c:1335-1354).

No golden of the reference covers it, so the oracle's restatement
(src3d/pfarfield.c:120-401) is pinned by what the mathematics fixes:
* one image (pi, pj in [0, 1)) is the plain time-domain far field restricted to
  the z faces of the same box -- the plain NFFF with its four side faces
  skipped, including a partial skip rectangle on each z face -- bin for bin
  (the plain NFFF is pinned by the reference's selftest 10035011 far-field
  golden); a flipped derivative sign, a wrong obliquity or a wrong area fails;
* the image sum over pi, pj in [-1, 2) equals the sum of one-image transforms
  for the observation point moved by -(pi (in - i0), pj (jn - j0), 0): the
  period is the box width and the images are shifted the right way;
* the postprocess mode (fields stored from ppstart, integrated at the last
  step) equals the direct mode when the z planes are still dark before
  ppstart, and differs once ppstart is after the pulse's arrival;
* skipping all of z0 plus skipping all of zn adds up to skipping nothing.
The GPU kernel is checked against the oracle in both modes (GPU test below).
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

BOX = "12 12 12 28 28 28"


def run_oracle(par):
    o = Oracle(par)
    o.step()
    pf, ff = o.pfarfield(), o.farfield()
    o.close()
    return pf, ff


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_one_image_equals_plain_nfff_z_faces(tmp_path):
    pts = [(20, 20, 70), (31, 9, 55), (14, 26, -30)]
    blocks = ["PERIODIC_NFFF\n%s 0 0 1 1" % BOX,
              "PERIODIC_NFFF_SKIP\nk0 14 15 19 22", "PERIODIC_NFFF_SKIP\nkn 20 13 26 17",
              "NFFF\n%s 0 0" % BOX,
              "NFFF_SKIP\ni0 0 0 1000 1000", "NFFF_SKIP\nin 0 0 1000 1000",
              "NFFF_SKIP\nj0 0 0 1000 1000", "NFFF_SKIP\njn 0 0 1000 1000",
              "NFFF_SKIP\nk0 14 15 19 22", "NFFF_SKIP\nkn 20 13 26 17"]
    for q, p in enumerate(pts):
        blocks.append("PERIODIC_NFFF_RAMAHI_POINT\n%d %d %d pff%d" % (p + (q,)))
        blocks.append("NFFF_RAMAHI_POINT\n%d %d %d ff%d" % (p + (q,)))
    pf, ff = run_oracle(cases.pfarfield_case(str(tmp_path / "c"), blocks))
    nd = pf.shape[2]
    assert ff.shape[2] >= nd and np.all(ff[:, :, nd:] == 0)
    for q in range(len(pts)):
        e = rel(pf[q], ff[q, :, :nd])
        print("point", pts[q], "periodic vs plain z faces", e, "peak", np.abs(pf[q]).max())
        assert np.abs(pf[q]).max() > 0 and e < 1e-12


def test_image_sum_equals_shifted_points(tmp_path):
    p = (23, 18, 75)
    w = 16                                   # box width along x and y
    imgs = [(a, b) for a in (-1, 0, 1) for b in (-1, 0, 1)]
    many = cases.pfarfield_case(str(tmp_path / "imgs"), [
        "PERIODIC_NFFF\n%s -1 -1 2 2" % BOX, "PERIODIC_NFFF_RAMAHI_POINT\n%d %d %d pf" % p])
    blocks = ["PERIODIC_NFFF\n%s 0 0 1 1" % BOX]
    for q, (a, b) in enumerate(imgs):
        blocks.append("PERIODIC_NFFF_RAMAHI_POINT\n%d %d %d s%d" % (p[0] - a * w, p[1] - b * w, p[2], q))
    single = cases.pfarfield_case(str(tmp_path / "single"), blocks)
    pf_many, _ = run_oracle(many)
    pf_single, _ = run_oracle(single)
    o = Oracle(single)                       # every istart is 0 at these distances: bins align
    o.close()
    want = pf_single.sum(axis=0)
    e = rel(pf_many[0], want)
    print("image sum vs shifted points", e)
    assert e < 1e-12
    assert rel(pf_many[0], pf_single[4]) > 0.1   # the images matter


@pytest.mark.parametrize("ppstart", [4, 70])
def test_postprocess_mode_against_direct_mode(tmp_path, ppstart):
    pts = [(20, 20, 70), (31, 9, -40)]
    base = ["PERIODIC_NFFF\n%s -1 0 2 1" % BOX, "PERIODIC_NFFF_SKIP\nk0 13 13 16 18"]
    base += ["PERIODIC_NFFF_RAMAHI_POINT\n%d %d %d p%d" % (p + (q,)) for q, p in enumerate(pts)]
    direct, _ = run_oracle(cases.pfarfield_case(str(tmp_path / "d"), base))
    post, _ = run_oracle(cases.pfarfield_case(str(tmp_path / "p"), base + ["PERIODIC_NFFF_POSTPROCESS\n1 %d" % ppstart]))
    e = rel(post, direct)
    print("ppstart", ppstart, "postprocess vs direct", e)
    if ppstart == 4:        # the planes 8 cells from the source are still exactly zero before step 4
        assert e < 1e-12
    else:
        assert e > 1e-2


def test_skipped_planes_add_up(tmp_path):
    pt = "PERIODIC_NFFF_RAMAHI_POINT\n25 17 66 p"
    head = "PERIODIC_NFFF\n%s 0 -1 1 2" % BOX
    full, _ = run_oracle(cases.pfarfield_case(str(tmp_path / "a"), [head, pt]))
    no0, _ = run_oracle(cases.pfarfield_case(str(tmp_path / "b"), [head, pt, "PERIODIC_NFFF_SKIP\nk0 0 0 100 100"]))
    non, _ = run_oracle(cases.pfarfield_case(str(tmp_path / "c"), [head, pt, "PERIODIC_NFFF_SKIP\nkn 0 0 100 100"]))
    e = rel(no0 + non, full)
    print("z0 + zn vs both", e)
    assert e < 1e-12 and rel(no0, full) > 1e-2 and rel(non, full) > 1e-2


def test_planner_checks_and_ignored_areas(tmp_path, capfd):
    bad = cases.pfarfield_case(str(tmp_path / "bad"), ["PERIODIC_NFFF\n12 12 0 28 28 28 0 0 1 1",
                                                        "PERIODIC_NFFF_RAMAHI_POINT\n20 20 60 p"])
    with pytest.raises(RuntimeError):
        Oracle(bad)
    late = cases.pfarfield_case(str(tmp_path / "late"), ["PERIODIC_NFFF\n%s 0 0 1 1" % BOX,
                                                          "PERIODIC_NFFF_POSTPROCESS\n1 200",
                                                          "PERIODIC_NFFF_RAMAHI_POINT\n20 20 60 p"])
    with pytest.raises(RuntimeError):
        Oracle(late)
    ok = cases.pfarfield_case(str(tmp_path / "ok"), ["PERIODIC_NFFF\n%s 0 0 1 1" % BOX,
                                                      "PERIODIC_NFFF_SPHERICAL_AREA\n10 10 1000 0 0 10 10 0",
                                                      "PERIODIC_NFFF_RAMAHI_POINT\n20 20 60 p"], nsteps=5)
    o = Oracle(ok)
    o.close()
    assert "PERIODIC_NFFF_SPHERICAL_AREA" in capfd.readouterr().err


@pytest.mark.gpu
@pytest.mark.parametrize("post", [0, 1])
def test_gpu_periodic_farfield_against_oracle(tmp_path, post):
    from gsvit_copy_amd import Session
    pts = [(20, 20, 70), (31, 9, 55), (14, 26, -30)]
    blocks = ["PERIODIC_NFFF\n%s -2 -1 2 2" % BOX, "PERIODIC_NFFF_SKIP\nk0 14 15 19 22",
              "PERIODIC_NFFF_SKIP\nkn 20 13 26 17"]
    blocks += ["PERIODIC_NFFF_RAMAHI_POINT\n%d %d %d pff%d" % (p + (q,)) for q, p in enumerate(pts)]
    if post:
        blocks.append("PERIODIC_NFFF_POSTPROCESS\n1 30")
    folder = str(tmp_path / "g")
    par = cases.pfarfield_case(folder, blocks)
    print("periodic far field, postprocess", post, flush=True)
    ref, _ = run_oracle(par)
    s = Session(par, write_outputs=True)
    s.run()
    got = s.pfarfield()
    s.close()
    for q in range(len(pts)):
        e = rel(got[q], ref[q])
        print("point", pts[q], "GPU vs FP64 oracle", e)
        assert e <= 1e-5
        # the log: "%d %g %g %g" of the negated bins, src3d/output.c:1347-1352
        log = np.loadtxt(os.path.join(folder, "pff%d" % q))
        assert log.shape == (ref.shape[2], 4) and np.array_equal(log[:, 0], np.arange(ref.shape[2]))
        assert rel(-log[:, 1:].T, ref[q]) <= 1e-5
