"""Parameter files of the reference's self-tests.

Restates the recipe of make_test (src3d/tests.c:92-257): for a self-test id
"RRRBMCFS" (resolution, boundary, material, matcheck, farfield, source) it
writes tmp_selftest.par, tmp_selftest.material and tmp_selftest.source exactly
as the reference writes them ("%g" samples), so that the run can be compared
with the golden trace tests/golden/ref/selftests/<id>, which holds the first
rows of the reference's own golden output (the reference compares the first
nsteps rows, src3d/tests.c:294-317).
"""
import math
import os

MATERIAL_LINES = {
    # src3d/tests.c:165-185 (entity 8 = box)
    1: "8 %d %d %d %d %d %d 0 2 1 0 0\n",
    2: "8 %d %d %d %d %d %d 0 1 2 0 0\n",
    3: "8 %d %d %d %d %d %d 1 2 1 0 0\n",
    4: "8 %d %d %d %d %d %d 1 1 2 0 0\n",
    5: "8 %d %d %d %d %d %d 10\n",
    6: "8 %d %d %d %d %d %d 2 5.0 9.5 0.0987\n",
    7: "8 %d %d %d %d %d %d 4  1.1431 1.3202e16 1.0805e14  0.26698 -1.2371 3.8711e15 4.4642e14  3.0834 -1.0968 4.1684e15 2.3555e15\n",
    8: "8 %d %d %d %d %d %d 5  1.1431 0 1.3202e16 1.0805e14  0.26698 -1.2371 3.8711e15 4.4642e14  3.0834 -1.0968 4.1684e15 2.3555e15\n",
    9: "8 %d %d %d %d %d %d 6  1.1431 0 1.3202e16 1.0805e14  0.26698 -1.2371 3.8711e15 4.4642e14  3.0834 -1.0968 4.1684e15 2.3555e15\n",
}


def g(v):
    return "%g" % v


def write_selftest(folder, test_id):
    res = int(test_id[:3])
    bnd, mat, matcheck, farfield, source = (int(c) for c in test_id[3:])
    nsteps = 301 if res > 101 else 201
    os.makedirs(folder, exist_ok=True)
    par = []
    if mat in (6, 8, 9, 7):
        par.append("POOL\n%d %d %d 1e-9 1e-9 1e-9\n\n" % (res, res, res))
    else:
        par.append("POOL\n%d %d %d 1e-6 1e-6 1e-6\n\n" % (res, res, res))
    par.append("COMP\n%d\n\n" % nsteps)
    if mat in (8, 9, 6):
        par.append("DT_MULT\n0.5\n\n")
    par.append("MATMODE_CHECK\n%d\n\n" % (1 if matcheck == 1 else 0))
    par.append("THREADS\n1\n\n")
    if mat > 0:
        par.append("MEDIUM_VECTOR\ntmp_selftest.material\n\n")
        with open(os.path.join(folder, "tmp_selftest.material"), "w") as f:
            f.write(MATERIAL_LINES[mat] % (res // 2 + 10, 30, 30, res - 30, res - 30, res - 30))
    skip = 100
    if farfield == 0:
        if mat in (7, 9, 8, 6):
            par.append("OUT_POINT\nEx %d %d %d %d %s\n\n" % (skip, res // 2, res // 2 + 30, res // 2 + 30, test_id))
        else:
            par.append("OUT_POINT\nEx %d %d %d %d %s\n\n" % (skip, res // 2 + 30, res // 2 + 30, res // 2 + 30, test_id))
    elif farfield == 1:
        par.append("NFFF\n%d %d %d %d %d %d 1 10000000\n\nNFFF_RAMAHI_POINT\n%d %d %d %s\n\n"
                   % (16, 16, 16, res - 16, res - 16, res - 16, res - 12, res - 12, res - 12, test_id))
    else:
        raise ValueError("periodic far field self-tests are not supported")
    if source == 0:
        size = 10
        lines = ["%d\n" % nsteps]
        for i in range(nsteps):
            val = 5 * math.sin(0.25 * i)
            damp = (i * 1e-6 - 4 * size) / math.sqrt(2.0) / size
            val *= math.exp(-damp * damp)
            lines.append("%d %s %s %s %s %s %s\n" % (i, g(val), g(0.0), g(0.0), g(0.0), g(0.0), g(0.0)))
        with open(os.path.join(folder, "tmp_selftest.source"), "w") as f:
            f.writelines(lines)
        par.append("SOURCE_POINT\n%d %d %d 0 tmp_selftest.source\n\n" % (res // 2, res // 2, res // 2))
    elif source == 1:
        par.append("SOURCE_TSF\n20 20 20 %d %d %d 1.57079632679 1.57079632679 0 2 20e-6 10 5\n\n"
                   % (res - 20, res - 20, res - 20))
    else:
        raise ValueError("source type %d is not supported" % source)
    if bnd == 2:
        par.append("BOUNDARY_ALL\nliao\n\n")
    elif bnd == 3:
        par.append("BOUNDARY_ALL\ncpml 10 3 -1 0.03 4\n\n")
    path = os.path.join(folder, "tmp_selftest.par")
    with open(path, "w") as f:
        f.write("".join(par))
    return path, nsteps


def read_golden(test_id, nrows, farfield):
    """Returns the golden trace: magnitude for far-field tests, the value otherwise."""
    here = os.path.dirname(os.path.abspath(__file__))
    rows = []
    with open(os.path.join(here, "ref", "selftests", test_id)) as f:
        for line in f:
            p = line.split()
            if not p:
                continue
            if farfield:
                a, b, c = (float(x) for x in p[1:4])
                rows.append(math.sqrt(a * a + b * b + c * c))
            else:
                rows.append(float(p[1]))
            if len(rows) == nrows:
                break
    return rows
