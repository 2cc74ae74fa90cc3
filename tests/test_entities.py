"""Geometric entities of MEDIUM_VECTOR files: cylinder (7), tetrahedron (12),
cone (107), rounded cone (108) and TetGen meshes (21), rasterised at voxel
centres in float arithmetic as src3d/plan.c:2062-2240 and :4393-4587 do.

Pins (CPU, through the planner that both the engine and the oracle use):
each inside test against an independent double-precision statement of the
solid (axis projection and radial distance, barycentric coordinates) within
the reference's voxel loop (the end points' box widened by radius + 1, which
clips the cone: its radius grows with the distance from the apex), with
disagreement allowed only where a voxel centre lies within 1e-4 cell of the
surface (float rounding); the six tetrahedra of a box subdivision tile the
box exactly; a TetGen mesh equals its tetrahedra written out (node numbering
from 0 or 1, comments, scaling and shifting in double rounded to float);
entities are drawn in file order, the later one winning.
"""
import os

import numpy as np

from oracle_lib import Oracle

N = 40


def entity_maps(folder, lines, n=N):
    """material-index map of a pool holding the given MEDIUM_VECTOR lines"""
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "ents.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    par = os.path.join(folder, "ents.par")
    with open(par, "w") as f:
        f.write("POOL\n%d %d %d 1e-8 1e-8 1e-8\n\nCOMP\n1\n\nMEDIUM_VECTOR\nents.txt\n\n" % (n, n, n))
    o = Oracle(par)
    mat, _ = o.maps()
    o.close()
    return mat


def lintab(q):
    """a distinct tabulated material per entity (index q + 1 after vacuum)"""
    return "1 %.3f 1 0 0" % (1.5 + 0.25 * q)


def centres(n=N):
    g = np.arange(n) + 0.5
    return np.meshgrid(g, g, g, indexing="ij")


def axis_frame(p1, p2):
    x, y, z = centres()
    a = np.asarray(p1, dtype=float)
    u = np.asarray(p2, dtype=float) - a
    length = np.linalg.norm(u)
    u /= length
    rel = np.stack([x - a[0], y - a[1], z - a[2]], axis=-1)
    t = rel @ u
    radial = np.linalg.norm(rel - t[..., None] * u, axis=-1)
    return t, radial, np.linalg.norm(rel, axis=-1), length


def check(mat, inside, margin, q):
    got = mat == q + 1
    bad = got != inside
    print("entity %d: %d voxels, %d disagree, smallest margin there %s"
          % (q, int(inside.sum()), int(bad.sum()), float(margin[bad].min()) if bad.any() else None))
    assert inside.sum() > 50
    assert np.all(margin[bad] < 1e-4)


def test_cylinder_cone_rcone_against_closed_forms(tmp_path):
    rng = np.random.default_rng(4)
    solids = []
    for kind in ("cyl", "cone", "rcone", "cyl", "cone", "rcone"):
        p1 = rng.uniform(8, 32, 3)
        p2 = p1 + rng.uniform(-1, 1, 3) * 14
        r = float(rng.uniform(3, 7))
        solids.append((kind, p1, p2, r, float(rng.uniform(1, 3))))
    for q, (kind, p1, p2, r, r1) in enumerate(solids):
        pts = " ".join("%.6f" % v for v in list(p1) + list(p2))
        if kind == "cyl":
            line = "7 %s %.6f %s" % (pts, r, lintab(q))
        elif kind == "cone":
            line = "107 %s %.6f %s" % (pts, r, lintab(q))
        else:
            line = "108 %s %.6f %.6f %s" % (pts, r1, r, lintab(q))
        mat = entity_maps(str(tmp_path / ("s%d" % q)), [line])
        # the solids in double precision, from the values the file holds
        p1f = np.array([float("%.6f" % v) for v in p1])
        p2f = np.array([float("%.6f" % v) for v in p2])
        rf, r1f = float("%.6f" % r), float("%.6f" % r1)
        t, radial, pos, length = axis_frame(p1f, p2f)
        slab = np.minimum(t, length - t)                 # within the end planes when >= 0
        if kind == "cyl":
            lateral = rf - radial
        elif kind == "cone":                              # radius grows with the distance from p1
            lateral = pos * rf / length - radial
        else:
            lateral = r1f + t * (rf - r1f) / length - radial
        # the voxel loop spans the end points' box widened by the (largest) radius + 1; the
        # cone's criterion (radius grows with the distance from p1, not along the axis)
        # reaches beyond it near the base, where the box clips it
        big = max(rf, r1f) if kind == "rcone" else rf
        x, y, z = centres()
        inbox = np.ones(x.shape, dtype=bool)
        for c, g in ((0, x), (1, y), (2, z)):
            lo = int(max(0.0, min(p1f[c], p2f[c]) - big - 1))
            hi = min(N, max(p1f[c], p2f[c]) + big + 1)
            inbox &= (g - 0.5 >= lo) & (g - 0.5 < hi)
        inside = (slab >= 0) & (lateral > 0) & inbox
        margin = np.minimum(np.abs(slab), np.abs(lateral))
        check(mat, inside, margin, 0)


def tet_inside(p):
    """barycentric inside test in double and the smallest |coordinate|"""
    x, y, z = centres()
    pts = np.stack([x, y, z], axis=-1)
    v0 = np.asarray(p[0], float)
    m = np.stack([np.asarray(p[k], float) - v0 for k in (1, 2, 3)], axis=1)
    lam = np.linalg.solve(m, (pts - v0).reshape(-1, 3).T).T.reshape(pts.shape)
    bary = np.concatenate([1 - lam.sum(axis=-1, keepdims=True), lam], axis=-1)
    return bary.min(axis=-1) >= 0, np.abs(bary).min(axis=-1)


def test_tetrahedra_against_barycentric_and_tiling_a_box(tmp_path):
    lo, hi = np.array([5.3, 7.6, 4.2]), np.array([31.7, 28.4, 33.1])
    corner = lambda b: lo + (hi - lo) * np.array(b, float)
    # Kuhn subdivision: one tetrahedron per ordering of the axes along the main diagonal
    tets = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        b = [0, 0, 0]
        path = [corner(b)]
        for a in perm:
            b[a] = 1
            path.append(corner(b))
        tets.append(path)
    lines = ["12 %s %s" % (" ".join("%.6f" % v for pt in tet for v in pt), lintab(q)) for q, tet in enumerate(tets)]
    union = np.zeros((N, N, N), dtype=bool)
    for q, line in enumerate(lines):
        mat = entity_maps(str(tmp_path / ("t%d" % q)), [line])
        inside, margin = tet_inside([[float("%.6f" % v) for v in pt] for pt in tets[q]])
        check(mat, inside, margin, 0)
        union |= mat == 1
    x, y, z = centres()
    box = (x >= lo[0]) & (x <= hi[0]) & (y >= lo[1]) & (y <= hi[1]) & (z >= lo[2]) & (z <= hi[2])
    assert np.array_equal(union, box)
    # one file, later entities overwrite earlier ones where they overlap
    allm = entity_maps(str(tmp_path / "all"), lines)
    assert np.array_equal(allm > 0, box)


def test_file_order_decides_overlaps(tmp_path):
    """a box and a sphere that overlap in part; material indices follow first
    appearance, the later entity owns the overlap"""
    a = "8 4 4 4 17 30 30 %s" % lintab(0)
    b = "4 17.0 17.0 17.0 9.0 %s" % lintab(1)
    sphere = entity_maps(str(tmp_path / "b"), [b]) > 0
    box = entity_maps(str(tmp_path / "a"), [a]) > 0
    both, only = sphere & box, sphere & ~box
    assert both.sum() > 500 and only.sum() > 500
    m1 = entity_maps(str(tmp_path / "ab"), [a, b])      # a -> 1, b -> 2; b drawn last
    m2 = entity_maps(str(tmp_path / "ba"), [b, a])      # b -> 1, a -> 2; a drawn last
    assert np.all(m1[sphere] == 2) and np.all(m1[box & ~sphere] == 1)
    assert np.all(m2[both] == 2) and np.all(m2[only] == 1)


def write_tetgen(folder, base, nodes, tets, first=1, natt_node=1, natt_ele=0, comments=True):
    with open(os.path.join(folder, base + ".node"), "w") as f:
        if comments:
            f.write("# nodes\n\n")
        f.write("%d 3 %d 0\n" % (len(nodes), natt_node))
        if comments:
            f.write("# index x y z attributes\n")
        for q, v in enumerate(nodes):
            f.write("%d %.7g %.7g %.7g%s\n" % (q + first, v[0], v[1], v[2], " 7" * natt_node))
    with open(os.path.join(folder, base + ".ele"), "w") as f:
        if comments:
            f.write("# elements\n")
        f.write("%d 4 %d\n" % (len(tets), natt_ele))
        for q, t in enumerate(tets):
            f.write("%d %d %d %d %d%s\n" % (q + first, t[0] + first, t[1] + first, t[2] + first, t[3] + first,
                                            " 1" * natt_ele))


def test_tetgen_mesh_equals_its_tetrahedra(tmp_path):
    rng = np.random.default_rng(9)
    nodes = rng.uniform(0, 3, (8, 3)).astype(np.float32)
    tets = [(0, 1, 2, 3), (4, 5, 6, 7), (1, 3, 5, 7)]
    shift, mult = (6.25, 4.5, 8.0), (9.5, 10.25, 8.75)
    mat_line = lintab(0)
    explicit = []
    for t in tets:
        pts = []
        for n in t:
            pts += [float(np.float32(float(nodes[n][a]) * mult[a] + shift[a])) for a in range(3)]
        explicit.append("12 %s %s" % (" ".join(repr(v) for v in pts), mat_line))
    ref = entity_maps(str(tmp_path / "explicit"), explicit)
    assert (ref > 0).sum() > 100
    for first in (0, 1):
        folder = str(tmp_path / ("mesh%d" % first))
        os.makedirs(folder, exist_ok=True)
        write_tetgen(folder, "m", nodes, tets, first=first)
        got = entity_maps(folder, ["21 m -1 -1 %g %g %g %g %g %g %s" % (shift + mult + (mat_line,))])
        assert np.array_equal(got, ref)
    # with element attributes and a filter on, nothing passes (reference behaviour); off, all do
    folder = str(tmp_path / "att")
    os.makedirs(folder, exist_ok=True)
    write_tetgen(folder, "m", nodes, tets, natt_ele=1)
    off = entity_maps(folder, ["21 m -1 -1 %g %g %g %g %g %g %s" % (shift + mult + (mat_line,))])
    on = entity_maps(folder, ["21 m 0 1 %g %g %g %g %g %g %s" % (shift + mult + (mat_line,))])
    assert np.array_equal(off, ref) and not on.any()
