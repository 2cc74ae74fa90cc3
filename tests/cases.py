"""Seeded parameter-file workloads for parity tests and the bench.

Every function writes a GSvit parameter file plus its input files into a
folder and returns the parameter-file path.  The workloads:
  sphere_scatter  plain 3-D FDTD, 40 dielectric spheres, CPML, Ez dipole (128^3)
  rough_surface   Drude rough-surface scattering, TSF plane wave at 45 degrees,
                  running-DFT near-to-far-field box (512^3)
  inverse_design  continuous-permittivity design region, CW Ez dipole below
                  it, DFT objective on a plane above it (gradient extension)
  sweep           64 seeded rough-surface runs of uneven grid sizes (batching)
  textured_film   dispersive textured thin-film stack, periodic in x and y,
                  three-pole silicon, flux spectra per layer (768 x 768 x 384)
and small feature cases (CPML + ADE, TSF oblique, Liao + PEC, periodic, far field).
"""
import math
import os

import numpy as np

C0 = 299792458.0
ETA0 = 376.73


def _dt(dx, dy, dz, dtmult):
    return 1.0 / C0 / math.sqrt(1.0 / dx ** 2 + 1.0 / dy ** 2 + 1.0 / dz ** 2) * dtmult


def gaussian_sine(nsteps, dt, lam, width_periods, delay_widths=3.0, amplitude=1.0):
    """Gaussian-modulated sine: FWHM of the envelope = width_periods periods."""
    period = lam / C0
    sigma = width_periods * period / 2.3548
    t0 = delay_widths * width_periods * period
    t = np.arange(nsteps) * dt
    return amplitude * np.sin(2 * math.pi * t / period) * np.exp(-0.5 * ((t - t0) / sigma) ** 2)


def write_point_source(path, vals, comp=2):
    with open(path, "w") as f:
        f.write("%d\n" % len(vals))
        for n, v in enumerate(vals):
            row = [0.0] * 6
            row[comp] = v
            f.write("%d %s\n" % (n, " ".join("%.17g" % x for x in row)))


def write_tsf_source(path, vals):
    with open(path, "w") as f:
        f.write("%d\n" % len(vals))
        for n, v in enumerate(vals):
            f.write("%d %.17g\n" % (n, v))


def place_spheres(rng, n, lo, hi, rmin, rmax):
    """Non-overlapping spheres with centres uniform in [lo+r, hi-r]."""
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place %d spheres" % n)
        r = rng.uniform(rmin, rmax)
        c = rng.uniform(lo + r, hi - r, size=3)
        if all(np.linalg.norm(c - np.array(o[:3])) > r + o[3] + 1 for o in out):
            out.append((c[0], c[1], c[2], r))
    return out


def sphere_scatter(folder, n=128, nsteps=1000, nspheres=40, seed=1, npoints=16):
    """Plain 3-D FDTD: 10 nm cells, Courant 0.5, CPML 10 (m=3,
    sigma_max = 0.8 (m+1)/(eta0 dx), kappa 1, alpha 0), 40 LINTAB spheres of
    radius U[5,15] cells and eps U[1,4], Ez dipole at the centre driven by a
    Gaussian-modulated sine (lambda0 = 500 nm, width 8 periods)."""
    os.makedirs(folder, exist_ok=True)
    rng = np.random.default_rng(seed)
    dx = 10e-9
    depth = 10
    sigma = 0.8 * 4 / (ETA0 * dx)
    spheres = place_spheres(rng, nspheres, depth, n - depth, 5, 15) if nspheres else []
    eps = rng.uniform(1, 4, size=len(spheres))
    with open(os.path.join(folder, "spheres.txt"), "w") as f:
        for (x, y, z, r), e in zip(spheres, eps):
            f.write("4 %.6f %.6f %.6f %.6f 1 %.6f 1 0 0\n" % (x, y, z, r, e))
    dt = _dt(dx, dx, dx, 0.5)
    write_point_source(os.path.join(folder, "dipole.txt"), gaussian_sine(nsteps, dt, 500e-9, 8), comp=2)
    pts = rng.integers(depth + 2, n - depth - 2, size=(npoints, 3))
    c = n // 2
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, dx, dx, dx), "COMP\n%d\n\n" % nsteps, "DT_MULT\n0.5\n\n",
             "BOUNDARY_ALL\ncpml %d 3 %.10g 0 1\n\n" % (depth, sigma),
             "SOURCE_POINT\n%d %d %d 0 dipole.txt\n\n" % (c, c, c)]
    if spheres:
        lines.append("MEDIUM_VECTOR\nspheres.txt\n\n")
    for q, p in enumerate(pts):
        lines.append("OUT_POINT\nAll 1 %d %d %d point%02d\n\n" % (p[0], p[1], p[2], q))
    path = os.path.join(folder, "sphere_scatter.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def rough_height(n, rms, corr, seed, ny=None):
    """Gaussian-correlated random height field (cells) of n x ny (default n x n):
    white noise filtered with a Gaussian of correlation length `corr` cells,
    scaled to the requested rms."""
    ny = n if ny is None else ny
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, ny))
    kx, ky = np.meshgrid(np.fft.fftfreq(n), np.fft.fftfreq(ny), indexing="ij")
    filt = np.exp(-(math.pi * corr) ** 2 * (kx ** 2 + ky ** 2) / 2)
    h = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    h -= h.mean()
    return h * (rms / h.std())


def write_vtk_ids(path, ids):
    """ASCII STRUCTURED_POINTS integer map, x fastest (src3d/plan.c:3877-3960)."""
    nx, ny, nz = ids.shape
    flat = np.ascontiguousarray(ids.transpose(2, 1, 0)).reshape(-1)
    if flat.max() > 9 or flat.min() < 0:
        raise ValueError("write_vtk_ids handles single-digit ids")
    buf = np.empty(2 * flat.size, dtype=np.uint8)
    buf[0::2] = flat.astype(np.uint8) + ord("0")
    buf[1::2] = ord("\n")
    with open(path, "wb") as f:
        f.write(b"# vtk DataFile Version 2.0\nmaterial ids\nASCII\nDATASET STRUCTURED_POINTS\n")
        f.write(("DIMENSIONS %d %d %d\nORIGIN 0 0 0\nSPACING 1 1 1\nPOINT_DATA %d\n" % (nx, ny, nz, flat.size)).encode())
        f.write(b"SCALARS mat int 1\nLOOKUP_TABLE default\n")
        f.write(buf.tobytes())


DRUDE_SILVER = "5 1 0 1.37e16 2.7e13 0 0 0 0 0 0 0 0"

# Time-step multiplier of the oblique-incidence workloads.  The TSF auxiliary
# line (src3d/1dpool.c: Mur condition on H at the source end, then the Yee
# update of the same H) diverges when its Courant number c dt / (dx / corr)
# exceeds 0.75; at the default dt that number is corr / sqrt(3), 0.816 for
# theta = 45 deg (corr = sqrt 2), so the line grows by ~1e10 per 250 steps.
# 0.9 keeps every incidence angle with phi = 0 below 0.75 (worst: 0.735).
OBLIQUE_DT_MULT = 0.9


def rough_surface(folder, n=512, nsteps=4000, dx=5e-9, depth=16, seed=2, rms_nm=20.0, corr_nm=100.0,
                  theta_deg=45.0, lam_center=600e-9, dft_lambdas=(500e-9, 600e-9, 700e-9), tsf_margin=4,
                  nfff=True, ade_cpml_consistent=False):
    """Rough-surface scattering: lower half a Drude metal
    (eps_inf 1, omega_p 1.37e16 rad/s, gamma 2.7e13 1/s, ADE) below a
    Gaussian-correlated interface (rms 20 nm, correlation 100 nm); CPML 16;
    a p-polarised TSF plane wave at 45 degrees travelling down onto the metal
    with a Gaussian pulse centred at 600 nm (bottom TSF face skipped); running
    DFT of the tangential fields on a closed box at 500/600/700 nm."""
    os.makedirs(folder, exist_ok=True)
    h = rough_height(n, rms_nm * 1e-9 / dx, corr_nm * 1e-9 / dx, seed)
    zc = n / 2.0 + h
    k = np.arange(n) + 0.5
    ids = (k[None, None, :] < zc[:, :, None]).astype(np.int32)
    write_vtk_ids(os.path.join(folder, "surface.vtk"), ids)
    with open(os.path.join(folder, "metal.txt"), "w") as f:
        f.write("8 0 0 0 1 1 1 %s\n" % DRUDE_SILVER)
    dt = _dt(dx, dx, dx, OBLIQUE_DT_MULT)
    write_tsf_source(os.path.join(folder, "pulse.txt"), gaussian_sine(nsteps, dt, lam_center, 4, delay_widths=2.0))
    b0 = depth + tsf_margin
    b1 = n - depth - tsf_margin - 1
    theta = math.pi - math.radians(theta_deg)   # travelling towards -z
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "DT_MULT\n%g\n\n" % OBLIQUE_DT_MULT, "MATMODE_CHECK\n1\n\n",
             "BOUNDARY_ALL\ncpml %d 3 -1 0 1\n\n" % depth,
             "MEDIUM_VECTOR\nmetal.txt\n\n", "MEDIUM_VTK_DATA\nsurface.vtk\n\n",
             "SOURCE_TSF\n%d %d %d %d %d %d %.17g 0 %.17g 0 pulse.txt\n\n" % (b0, b0, b0, b1, b1, b1, theta, math.pi / 2),
             "TSF_SKIP\nk0\n\n"]
    if ade_cpml_consistent:
        lines.append("ADE_CPML_CONSISTENT\n1\n\n")
    if nfff:
        f0 = b0 + 6
        f1 = b1 - 6
        zlow = int(n / 2 + 4 * rms_nm * 1e-9 / dx) + 8
        lines.append("NFFF\n%d %d %d %d %d %d 0 0\n\n" % (f0, f0, zlow, f1, f1, f1))
        lines.append("NFFF_DFT\n%d %s dft.txt\n\n" % (len(dft_lambdas), " ".join("%.6g" % l for l in dft_lambdas)))
    path = os.path.join(folder, "rough_surface.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


SWEEP_EDGES = (64, 96, 128, 192, 256)


def sweep_params(nruns=64, seed=11, edges=SWEEP_EDGES):
    """Seeded parameters of a batched sweep: per run and per axis a grid edge
    drawn from `edges`, a roughness seed, rms U[5, 30] nm, incidence angle
    U[0, 70] degrees and centre wavelength U[450, 750] nm."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nruns):
        shape = tuple(int(e) for e in rng.choice(edges, size=3))
        out.append(dict(shape=shape, seed=int(rng.integers(1, 1 << 30)), rms_nm=float(rng.uniform(5, 30)),
                        theta_deg=float(rng.uniform(0, 70)), lam_center=float(rng.uniform(450e-9, 750e-9))))
    return out


def sweep_run(folder, shape, seed, rms_nm, theta_deg, lam_center, nsteps=2000, dx=5e-9, depth=10,
              corr_nm=100.0, tsf_margin=4):
    """One run of a batched sweep: the rough-surface geometry (Drude metal below
    a Gaussian-correlated interface at mid height) on an nx x ny x nz grid,
    CPML `depth` on all faces, a p-polarised TSF plane wave at theta_deg from
    the normal travelling down onto the metal (bottom face skipped) with a
    Gaussian pulse centred at lam_center.  No far-field box."""
    os.makedirs(folder, exist_ok=True)
    nx, ny, nz = shape
    h = rough_height(nx, rms_nm * 1e-9 / dx, corr_nm * 1e-9 / dx, seed, ny=ny)
    zc = nz / 2.0 + h
    k = np.arange(nz) + 0.5
    ids = (k[None, None, :] < zc[:, :, None]).astype(np.int32)
    write_vtk_ids(os.path.join(folder, "surface.vtk"), ids)
    with open(os.path.join(folder, "metal.txt"), "w") as f:
        f.write("8 0 0 0 1 1 1 %s\n" % DRUDE_SILVER)
    dt = _dt(dx, dx, dx, OBLIQUE_DT_MULT)
    write_tsf_source(os.path.join(folder, "pulse.txt"), gaussian_sine(nsteps, dt, lam_center, 4, delay_widths=2.0))
    b0 = [depth + tsf_margin] * 3
    b1 = [n - depth - tsf_margin - 1 for n in shape]
    theta = math.pi - math.radians(theta_deg)
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (nx, ny, nz, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "DT_MULT\n%g\n\n" % OBLIQUE_DT_MULT, "MATMODE_CHECK\n1\n\n",
             "BOUNDARY_ALL\ncpml %d 3 -1 0 1\n\n" % depth,
             "MEDIUM_VECTOR\nmetal.txt\n\n", "MEDIUM_VTK_DATA\nsurface.vtk\n\n",
             "SOURCE_TSF\n%d %d %d %d %d %d %.17g 0 %.17g 0 pulse.txt\n\n" % (*b0, *b1, theta, math.pi / 2),
             "TSF_SKIP\nk0\n\n"]
    path = os.path.join(folder, "sweep_run.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def sweep(folder, nruns=64, seed=11, nsteps=2000, edges=SWEEP_EDGES):
    """Write the runs of a seeded sweep into folder/runNN; returns the
    parameter-file paths and the per-run parameters."""
    params = sweep_params(nruns, seed, edges)
    paths = [sweep_run(os.path.join(folder, "run%02d" % r), nsteps=nsteps, **p) for r, p in enumerate(params)]
    return paths, params


def plane_wave_flux(folder, n=16, nz=120, nsteps=2200, dx=5e-9, depth=10, lam=500e-9,
                    planes=(40, 70), lambdas=(450e-9, 500e-9, 550e-9), slab=None, points=()):
    """Normal-incidence plane wave (theta = 0, travelling +z) in an n x n cell
    periodic in x and y (domain-periodic boundaries), CPML in z; the TSF box
    spans the cell from k = 20 up to nz - 20 and injects through its bottom
    face only; FLUX_DFT on `planes` (none if empty).  `slab` = (k0, k1, material line) fills
    k0 <= k < k1 with a MEDIUM_VECTOR material; `points` adds OUT_POINT traces
    at (i, j, k)."""
    os.makedirs(folder, exist_ok=True)
    dt = _dt(dx, dx, dx, 1.0)
    write_tsf_source(os.path.join(folder, "pulse.txt"), gaussian_sine(nsteps, dt, lam, 2, delay_widths=3.0))
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, nz, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "BOUNDARY_X0\nperiodic\n\nBOUNDARY_XN\nperiodic\n\nBOUNDARY_Y0\nperiodic\n\nBOUNDARY_YN\nperiodic\n\n",
             "BOUNDARY_Z0\ncpml %d 3 -1 0 1\n\nBOUNDARY_ZN\ncpml %d 3 -1 0 1\n\n" % (depth, depth),
             "SOURCE_TSF\n0 0 %d %d %d %d 0 0 0 0 pulse.txt\n\n" % (20, n, n, nz - 20),
             "TSF_SKIP\nkn\n\n"]
    if planes:
        lines.append("FLUX_DFT\n0 0 %d %d %d %s %d %s flux.txt\n\n"
                     % (n, n, len(planes), " ".join(str(k) for k in planes), len(lambdas),
                        " ".join("%.9g" % l for l in lambdas)))
    for q, (i, j, k) in enumerate(points):
        lines.append("OUT_POINT\nAll 1 %d %d %d pt%d\n\n" % (i, j, k, q))
    if slab:
        with open(os.path.join(folder, "slab.txt"), "w") as f:
            f.write("8 0 0 %d %d %d %d %s\n" % (slab[0], n, n, slab[1], slab[2]))
        lines.append("MEDIUM_VECTOR\nslab.txt\n\n")
    path = os.path.join(folder, "plane_wave_flux.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def pec_plate_wave(folder, source, n=12, nz=140, nsteps=900, dx=5e-9, depth=10, lam=500e-9, plate=(80, 84),
                   theta=0.0, points=()):
    """A plane-wave pulse (theta from +z, phi = psi = 0) in an n x n cell
    periodic in x and y, CPML in z, onto a PEC plate filling the cell at
    plate[0] <= k < plate[1].  source "tsf": a TSF box from k = 20 to nz - 20
    injecting through its bottom face (below it the scattered field only);
    source "sf": a scattered-field source (SOURCE_SF), the whole pool holding
    the scattered field."""
    os.makedirs(folder, exist_ok=True)
    dt = _dt(dx, dx, dx, 1.0)
    write_tsf_source(os.path.join(folder, "pulse.txt"), gaussian_sine(nsteps, dt, lam, 2, delay_widths=2.5))
    with open(os.path.join(folder, "plate.txt"), "w") as f:
        f.write("8 0 0 %d %d %d %d 10\n" % (plate[0], n, n, plate[1]))
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, nz, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "BOUNDARY_X0\nperiodic\n\nBOUNDARY_XN\nperiodic\n\nBOUNDARY_Y0\nperiodic\n\nBOUNDARY_YN\nperiodic\n\n",
             "BOUNDARY_Z0\ncpml %d 3 -1 0 1\n\nBOUNDARY_ZN\ncpml %d 3 -1 0 1\n\n" % (depth, depth),
             "MEDIUM_VECTOR\nplate.txt\n\n"]
    if source == "tsf":
        lines += ["SOURCE_TSF\n0 0 %d %d %d %d %.17g 0 0 0 pulse.txt\n\n" % (20, n, n, nz - 20, theta), "TSF_SKIP\nkn\n\n"]
    else:
        lines.append("SOURCE_SF\n%.17g 0 0 0 pulse.txt\n\n" % theta)
    for q, (i, j, k) in enumerate(points):
        lines.append("OUT_POINT\nAll 1 %d %d %d pt%d\n\n" % (i, j, k, q))
    path = os.path.join(folder, "plate_%s.par" % source)
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


# gold-like Drude + 2 critical points (eps_inf, sigma, omega_D, gamma_D, then
# A, phi, Omega, Gamma per pole; rad/s), in the range of published Drude-CP
# fits of gold; passive (Im eps > 0 at every frequency)
AU_ADE = "5 1.1431 0 1.3202e16 1.0805e14 0.26698 -1.2371 3.8711e15 4.4642e14 3.0834 -1.0968 4.1684e15 2.3555e15"

PASSIVITY_OMEGA = np.logspace(11, 18, 20000)


def ade_susceptibility(line, omega):
    """chi(omega) of an ADE material line (type 5 or 7) for exp(-i omega t):
    -omega_D^2 / (omega^2 + i gamma_D omega) plus, per pole,
    A Omega (e^{i phi} / (Omega - omega - i Gamma) + e^{-i phi} / (Omega + omega + i Gamma))."""
    v = [float(x) for x in line.split()]
    wd, gd = v[3], v[4]
    w = np.asarray(omega, dtype=float)
    chi = np.zeros(w.shape, dtype=complex)
    if wd:
        chi -= wd ** 2 / (w ** 2 + 1j * gd * w)
    for q in range((len(v) - 5) // 4):
        a, phi, om, g = v[5 + 4 * q: 9 + 4 * q]
        chi += a * om * (np.exp(1j * phi) / (om - w - 1j * g) + np.exp(-1j * phi) / (om + w + 1j * g))
    return chi


def silicon_3pole(rng):
    """Seeded three-pole silicon-like ADE line (type 7): two critical points in
    the near UV and a Lorentz term (a critical point with phi = 0) at the blue
    edge, no Drude term; eps ~ 12 in the near infrared.  A critical point with
    phase phi has a 1/detuning tail of sign -phi above and +phi below its
    resonance, which only the Lorentzian part (Gamma / detuning^2) offsets, so
    the phases are drawn small and negative (|phi| <= Gamma / Omega) and the
    draw is repeated until Im chi >= 0 at every frequency up to 1e18 rad/s: an
    active medium would grow without bound over a long run."""
    for _ in range(100):
        p = []
        for a, om, g in ((rng.uniform(2.5, 3.5), rng.uniform(5.0e15, 5.4e15), rng.uniform(1.5e14, 3e14)),
                         (rng.uniform(1.5, 2.5), rng.uniform(6.2e15, 6.6e15), rng.uniform(5e14, 8e14))):
            p.append((a, -rng.uniform(0.0, 1.0) * g / om, om, g))
        p.append((rng.uniform(0.5, 1.0), 0.0, rng.uniform(3.8e15, 4.2e15), rng.uniform(3e14, 6e14)))
        line = "7 1 0 0 0 " + " ".join("%.9g" % v for q in p for v in q)
        if ade_susceptibility(line, PASSIVITY_OMEGA).imag.min() >= 0:
            return line
    raise RuntimeError("no passive draw")


def pyramid_height(nx, ny, npyr, base_lo, base_hi, rng):
    """Height field (cells) of random upright square pyramids with 54.7 degree
    facets (height = base / 2 * tan 54.7), positions uniform, wrapped
    periodically in x and y; the texture is the maximum over the pyramids."""
    t = math.tan(math.radians(54.7))
    x = np.arange(nx)[:, None] + 0.5
    y = np.arange(ny)[None, :] + 0.5
    h = np.zeros((nx, ny))
    for _ in range(npyr):
        b = rng.uniform(base_lo, base_hi)
        cx, cy = rng.uniform(0, nx), rng.uniform(0, ny)
        dx = np.abs((x - cx + nx / 2) % nx - nx / 2)
        dy = np.abs((y - cy + ny / 2) % ny - ny / 2)
        h = np.maximum(h, b / 2 * t - t * np.maximum(dx, dy))
    return h


def textured_film(folder, nx=768, ny=768, nz=384, dx=4e-9, nsteps=6000, seed=3, nw=32, depth=16,
                  au=25, si=75, sio2=20, npyr=None, media=True, gap=8):
    """Dispersive textured thin-film stack, periodic in x and y, CPML in z.
    From the illuminated side: SiO2 (eps 2.13, `sio2` cells thick, conformal
    on the texture), silicon (three-pole ADE, `si` cells plus random
    pyramids of base U[100, 300] nm with 54.7 degree facets) and a gold-like
    Drude + 2 critical-point back contact (`au` cells), followed by `gap` air
    cells before the CPML (a dispersive medium continued into the CPML takes
    the reference's ADE-in-CPML update, which lets this metal's fields grow
    over thousands of steps).  A normally incident broadband pulse (400-1100 nm) enters through
    the TSF box's near-z face (the stack is stored upside down: the wave
    travels +z); FLUX_DFT planes in air before the stack, in the silicon just
    before the back contact, inside the contact and in the air behind it give
    the reflected, silicon-absorbed, gold-absorbed and transmitted power at
    nw wavelengths.  media=False
    writes the same cell, source and planes in vacuum (the incident power
    that normalises the spectra)."""
    os.makedirs(folder, exist_ok=True)
    rng = np.random.default_rng(seed)
    if npyr is None:
        npyr = int(1.2 * nx * ny * dx * dx / (200e-9) ** 2)
    h = pyramid_height(nx, ny, npyr, 100e-9 / dx, 300e-9 / dx, rng)
    # natural orientation: z up, light from the top; k_nat = nz - 1 - k
    kn = np.arange(nz)[None, None, :] + 0.5
    z_au = depth + gap + au
    z_si = z_au + si
    surf = z_si + h[:, :, None]
    ids = np.zeros((nx, ny, nz), dtype=np.int32)
    ids[(kn >= surf) & (kn < surf + sio2)] = 1
    ids[(kn >= z_au) & (kn < surf)] = 2
    ids[:, :, depth + gap:z_au] = 3
    ids = ids[:, :, ::-1]
    si_line = silicon_3pole(rng)
    if media:
        write_vtk_ids(os.path.join(folder, "film.vtk"), ids)
        with open(os.path.join(folder, "mats.txt"), "w") as f:
            f.write("8 0 0 0 1 1 1 1 2.13 1 0 0\n")
            f.write("8 1 0 0 2 1 1 %s\n" % si_line)
            f.write("8 2 0 0 3 1 1 %s\n" % AU_ADE)
    del ids
    dt = _dt(dx, dx, dx, 1.0)
    write_tsf_source(os.path.join(folder, "pulse.txt"), gaussian_sine(nsteps, dt, 600e-9, 1.0, delay_widths=3.0))
    top = nz - 1 - int(math.ceil(z_si + h.max() + sio2))     # lowest stack cell after the flip
    k_tsf = depth + 8
    planes = (max(k_tsf + 4, (k_tsf + top) // 2), nz - 1 - z_au - 2, nz - 1 - z_au + au // 2,
              nz - 1 - depth - gap // 2)
    lambdas = np.linspace(400e-9, 1100e-9, nw)
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (nx, ny, nz, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "MATMODE_CHECK\n1\n\n",
             "BOUNDARY_X0\nperiodic\n\nBOUNDARY_XN\nperiodic\n\nBOUNDARY_Y0\nperiodic\n\nBOUNDARY_YN\nperiodic\n\n",
             "BOUNDARY_Z0\ncpml %d 3 -1 0 1\n\nBOUNDARY_ZN\ncpml %d 3 -1 0 1\n\n" % (depth, depth),
             "MEDIUM_VECTOR\nmats.txt\n\nMEDIUM_VTK_DATA\nfilm.vtk\n\n" if media else "",
             "SOURCE_TSF\n0 0 %d %d %d %d 0 0 0 0 pulse.txt\n\n" % (k_tsf, nx, ny, nz - depth - 4),
             "TSF_SKIP\nkn\n\n",
             "FLUX_DFT\n0 0 %d %d %d %s %d %s flux.txt\n\n"
             % (nx, ny, len(planes), " ".join(str(k) for k in planes), nw, " ".join("%.6g" % l for l in lambdas))]
    path = os.path.join(folder, "textured_film.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def small_case(folder, kind, n=48, nsteps=160, seed=7):
    """Small feature cases for GPU/oracle parity."""
    os.makedirs(folder, exist_ok=True)
    rng = np.random.default_rng(seed)
    dx = 10e-9
    dt = _dt(dx, dx, dx, 0.5)
    c = n // 2
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, dx, dx, dx), "COMP\n%d\n\n" % nsteps, "DT_MULT\n0.5\n\n"]
    pts = [(c + 7, c - 5, c + 3), (6, c, c), (c, n - 7, 9), (c + 2, c + 2, n - 8)]
    for q, p in enumerate(pts):
        lines.append("OUT_POINT\nAll 1 %d %d %d pt%d\n\n" % (p[0], p[1], p[2], q))
    pulse = gaussian_sine(nsteps, dt, 150e-9, 1.5, delay_widths=1.2)
    write_point_source(os.path.join(folder, "src.txt"), pulse, comp=2)
    tsf_pulse = pulse
    write_tsf_source(os.path.join(folder, "tsf.txt"), tsf_pulse)
    mats = []
    if kind == "vacuum_none":
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c))
    elif kind == "spheres_cpml":
        sph = place_spheres(rng, 4, 8, n - 8, 3, 6)
        for (x, y, z, r) in sph:
            mats.append("4 %.5f %.5f %.5f %.5f 1 %.4f 1 %.4g 0" % (x, y, z, r, rng.uniform(1.5, 4), rng.uniform(0, 2e3)))
        mats.append("8 %d %d %d %d %d %d 0 2.5 1 0 0" % (4, 4, 4, 12, 12, 10))
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c))
    elif kind == "solids_cpml":
        # every supported solid: cylinder, cone, rounded cone and tetrahedron of
        # tabulated dielectrics, a Drude + 2 CP rounded cone, a PEC cylinder
        mats.append("7 10.3 12.1 9.2 18.6 15.4 11.7 3.6 1 3.1 1 40 0")
        mats.append("107 34.2 11.5 30.3 27.7 20.4 36.1 5.2 1 2.2 1 0 0")
        mats.append("108 12.4 33.6 12.2 20.8 36.3 21.5 1.5 4.1 %s" % AU_ADE)
        mats.append("12 28.3 30.1 9.4 38.6 34.2 13.3 31.7 39.1 18.8 35.2 31.4 20.6 1 4.5 1 0 0")
        mats.append("7 30.5 8.5 36.5 36.5 8.5 36.5 2.4 10")
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c))
    elif kind == "magnetic_liao":
        mats.append("8 %d %d %d %d %d %d 0 1 2 0 300" % (c + 4, 10, 10, n - 10, n - 10, n - 10))
        mats.append("4 %d %d %d 5 1 1 3 0 0" % (c - 8, c - 8, c - 8))
        lines.append("BOUNDARY_ALL\nliao\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c))
    elif kind == "ade_cpml":
        mats.append("8 %d %d %d %d %d %d %s" % (0, 0, 0, n, n, c - 4, DRUDE_SILVER))
        mats.append("4 %d %d %d 4 5 2 0 1.2e16 1e14 0.8 -1.2 4e15 4e14 1.5 -1.1 4.2e15 2.4e15" % (c + 8, c + 6, c + 2))
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c + 6))
    elif kind.startswith("ade3_"):
        # a sphere of a three-pole ADE material (type 7) over a Drude slab; the
        # variants pin the third pole: "ade3_split" splits pole 1 into two equal
        # poles of half amplitude (poles 1 and 3), "ade3_two" is the same medium as
        # a two-pole type-5 material, "ade3_idle" has a third pole of zero
        # amplitude, "ade3_lorentz" a genuine third pole (a Lorentz term: phi = 0)
        p1, p2 = (0.8, -1.2, 4e15, 4e14), (1.5, -1.1, 4.2e15, 2.4e15)
        half = (p1[0] / 2,) + p1[1:]
        fmt = lambda poles: " ".join("%.9g" % v for q in poles for v in q)
        med = {"ade3_split": "7 2 0 1.2e16 1e14 " + fmt((half, p2, half)),
               "ade3_two": "5 2 0 1.2e16 1e14 " + fmt((p1, p2)),
               "ade3_idle": "7 2 0 1.2e16 1e14 " + fmt((p1, p2, (0.0, 0.3, 3e15, 5e14))),
               "ade3_lorentz": "7 2 0 1.2e16 1e14 " + fmt((p1, p2, (1.1, 0.0, 3.3e15, 5e14)))}[kind]
        mats.append("8 %d %d %d %d %d %d %s" % (0, 0, 0, n, n, c - 4, DRUDE_SILVER))
        mats.append("4 %d %d %d 5 %s" % (c + 8, c + 6, c + 2, med))
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c + 6))
    elif kind in ("rc_cpml", "rc_cpml_aniso"):
        # the three recursive-convolution types together: a PLRC (type 6, with
        # conductivity) half space through every side slab, a Drude + 2 CP (type
        # 4) sphere and an eV-scaled Drude (type 2) box through the x0 slab;
        # "rc_cpml_aniso" has 10 x 8 x 6 nm cells (the types take their
        # spacings differently)
        poles = "0.26698 -1.2371 3.8711e15 4.4642e14 3.0834 -1.0968 4.1684e15 2.3555e15"
        mats.append("8 %d %d %d %d %d %d 6 1.1431 3e5 1.3202e16 1.0805e14 %s" % (0, 0, 0, n, n, c - 4, poles))
        mats.append("4 %d %d %d 5 4 1.1431 1.3202e16 1.0805e14 %s" % (c + 8, c + 6, c + 2, poles))
        mats.append("8 %d %d %d %d %d %d 2 3.0 9.5 0.0987" % (2, c + 4, c + 6, 14, n - 4, c + 12))
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c + 6))
        if kind == "rc_cpml_aniso":
            lines[0] = "POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, 10e-9, 8e-9, 6e-9)
    elif kind == "tsf_pec_liao":
        mats.append("4 %d %d %d 7 10" % (c, c, c))
        lines.append("BOUNDARY_ALL\nliao\n\n")
        lines.append("SOURCE_TSF\n8 8 8 %d %d %d 2.356194490192345 0.6 0.4 0 tsf.txt\n\n" % (n - 9, n - 9, n - 9))
    elif kind == "tsf_ade_cpml":
        mats.append("8 %d %d %d %d %d %d %s" % (0, 0, 0, n, n, c - 6, DRUDE_SILVER))
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0 1\n\n")
        lines.append("SOURCE_TSF\n10 10 10 %d %d %d 2.356194490192345 0.3 1.5707963267948966 0 tsf.txt\n\n" % (n - 11, n - 11, n - 11))
        lines.append("TSF_SKIP\nk0\n\n")
    elif kind == "periodic_liao":
        mats.append("4 %d %d %d 5 1 2.2 1 0 0" % (c, c + 3, c))
        lines.append("BOUNDARY_ALL\nliao\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c))
        lines.append("MBOUNDARY_X0\nperiodic 8\n\nMBOUNDARY_XN\nperiodic %d\n\n" % (n - 8))
        lines.append("MBOUNDARY_Y0\nperiodic 8\n\nMBOUNDARY_YN\nperiodic %d\n\n" % (n - 8))
    elif kind == "farfield_cpml":
        lines.append("BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n")
        lines.append("SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c))
        lines.append("NFFF\n12 12 12 %d %d %d 0 0\n\n" % (n - 12, n - 12, n - 12))
        lines.append("NFFF_RAMAHI_POINT\n%d %d %d ff0\n\n" % (n - 10, n - 10, n - 10))
        lines.append("NFFF_RAMAHI_POINT\n%d %d %d ff1\n\n" % (2 * n, c, c))
        lines.append("NFFF_DFT\n2 300e-9 450e-9 dft.txt\n\n")
    else:
        raise ValueError(kind)
    if mats:
        with open(os.path.join(folder, "mats.txt"), "w") as f:
            f.write("\n".join(mats) + "\n")
        lines.append("MEDIUM_VECTOR\nmats.txt\n\n")
    path = os.path.join(folder, kind + ".par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def periodic_window(folder, n=40, nsteps=200, window=(8, 32), blocks=(), tsf=True):
    """Interior periodic planes (MBOUNDARY_X0/XN/Y0/YN periodic at window[0],
    window[1]) in a pool of n^3 10 nm cells with Liao faces; a normally
    incident (+z) plane-wave pulse from a TSF box that spans the pool in x and y
    and cells 8..n-9 in z.  `blocks`: (i0, j0, k0, i1, j1, k1, eps_r) boxes of
    linear dielectric."""
    os.makedirs(folder, exist_ok=True)
    dx = 10e-9
    dt = _dt(dx, dx, dx, 0.5)
    write_tsf_source(os.path.join(folder, "tsf.txt"), gaussian_sine(nsteps, dt, 150e-9, 1.5, delay_widths=1.2))
    w0, w1 = window
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, dx, dx, dx), "COMP\n%d\n\n" % nsteps, "DT_MULT\n0.5\n\n",
             "BOUNDARY_ALL\nliao\n\n",
             "MBOUNDARY_X0\nperiodic %d\n\nMBOUNDARY_XN\nperiodic %d\n\n" % (w0, w1),
             "MBOUNDARY_Y0\nperiodic %d\n\nMBOUNDARY_YN\nperiodic %d\n\n" % (w0, w1)]
    if tsf:
        lines.append("SOURCE_TSF\n2 2 8 %d %d %d 0 0 0 0 tsf.txt\n\n" % (n - 3, n - 3, n - 9))
    if blocks:
        with open(os.path.join(folder, "blocks.txt"), "w") as f:
            for b in blocks:
                f.write("8 %d %d %d %d %d %d 0 %g 1 0 0\n" % b)
        lines.append("MEDIUM_VECTOR\nblocks.txt\n\n")
    path = os.path.join(folder, "window.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


SMALL_KINDS = ["vacuum_none", "spheres_cpml", "magnetic_liao", "ade_cpml", "tsf_pec_liao", "tsf_ade_cpml",
               "periodic_liao", "farfield_cpml", "solids_cpml", "rc_cpml", "rc_cpml_aniso"]


def pfarfield_case(folder, blocks, n=40, nsteps=200, seed=3):
    """A point source (Ez, centre) among two dielectric spheres, CPML 8,
    with the given parameter-file blocks (PERIODIC_NFFF*, NFFF*) appended."""
    os.makedirs(folder, exist_ok=True)
    rng = np.random.default_rng(seed)
    dx = 10e-9
    dt = _dt(dx, dx, dx, 0.5)
    c = n // 2
    pulse = gaussian_sine(nsteps, dt, 150e-9, 1.5, delay_widths=1.2)
    write_point_source(os.path.join(folder, "src.txt"), pulse, comp=2)
    mats = ["4 %.3f %.3f %.3f 3.5 1 %.3f 1 0 0" % (c + 5, c - 4, c + 2, rng.uniform(2, 4)),
            "4 %.3f %.3f %.3f 2.5 1 %.3f 1 0 0" % (c - 5, c + 3, c - 3, rng.uniform(2, 4))]
    with open(os.path.join(folder, "mats.txt"), "w") as f:
        f.write("\n".join(mats) + "\n")
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, dx, dx, dx), "COMP\n%d\n\n" % nsteps, "DT_MULT\n0.5\n\n",
             "BOUNDARY_ALL\ncpml 8 3 -1 0.03 4\n\n", "SOURCE_POINT\n%d %d %d 0 src.txt\n\n" % (c, c, c),
             "MEDIUM_VECTOR\nmats.txt\n\n"]
    lines += [b.rstrip("\n") + "\n\n" for b in blocks]
    path = os.path.join(folder, "pff.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def ltsf_case(folder, shape=(24, 24, 80), nsteps=400, theta=0.0, phi=0.0, psi=0.0, layers=(), box=None, skips=(),
              halfspace=None, lam=400e-9, dx=10e-9, dtmult=0.5, points=(), extra=(), boundary="cpml 8 3 -1 0 1"):
    """Layered TSF (SOURCE_LTSF) case: a Gaussian-modulated sine sampled at dt/3
    (3 nsteps samples), layers given as (z position, eps_r), optionally a
    dielectric half space (z from halfspace[0] to the end, eps_r halfspace[1])
    in the 3-D pool itself; angles in degrees."""
    os.makedirs(folder, exist_ok=True)
    nx, ny, nz = shape
    dt = _dt(dx, dx, dx, dtmult)
    write_tsf_source(os.path.join(folder, "ltsf.txt"), gaussian_sine(3 * nsteps, dt / 3, lam, 1.2, delay_widths=1.6))
    if box is None:
        box = (8 + 3, 8 + 3, 8 + 3, nx - 12, ny - 12, nz - 12)
    lay = " ".join("%d %.6g 1 0 0" % (z, e) for z, e in layers)
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (nx, ny, nz, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "DT_MULT\n%g\n\n" % dtmult, "BOUNDARY_ALL\n%s\n\n" % boundary,
             "SOURCE_LTSF\n%d %d %d %d %d %d %.17g %.17g %.17g %d %s 0 ltsf.txt\n\n"
             % (tuple(box) + (math.radians(theta), math.radians(phi), math.radians(psi), len(layers), lay))]
    lines += ["LTSF_SKIP\n%s\n\n" % s for s in skips]
    if halfspace is not None:
        with open(os.path.join(folder, "mats.txt"), "w") as f:
            f.write("8 -5 -5 %d %d %d %d 1 %.6g 1 0 0\n" % (halfspace[0], nx + 5, ny + 5, nz + 5, halfspace[1]))
        lines.append("MEDIUM_VECTOR\nmats.txt\n\n")
    for q, p in enumerate(points):
        lines.append("OUT_POINT\nAll 1 %d %d %d pt%d\n\n" % (tuple(p) + (q,)))
    lines += [b.rstrip("\n") + "\n\n" for b in extra]
    path = os.path.join(folder, "ltsf.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


def write_linear_medium(path, eps_r, sigma=None, mu_r=None, sigast=None):
    """MEDIUM_LINEAR file: 3 int32 dims, then float32 eps_r, sigma, mu_r, sigma* ([i][j][k], k fastest)."""
    eps_r = np.asarray(eps_r, dtype=np.float32)
    shape = eps_r.shape
    parts = [eps_r,
             np.zeros(shape, np.float32) if sigma is None else np.asarray(sigma, np.float32),
             np.ones(shape, np.float32) if mu_r is None else np.asarray(mu_r, np.float32),
             np.zeros(shape, np.float32) if sigast is None else np.asarray(sigast, np.float32)]
    with open(path, "wb") as f:
        f.write(np.array(shape, dtype=np.int32).tobytes())
        for a in parts:
            f.write(np.ascontiguousarray(a).tobytes())


def cw_ramp(nsteps, dt, lam, ramp_periods=3.0, amplitude=1.0):
    """CW sine with a smooth (raised-cosine) turn-on over ramp_periods periods."""
    period = lam / C0
    t = np.arange(nsteps) * dt
    ramp = np.clip(t / (ramp_periods * period), 0.0, 1.0)
    env = 0.5 - 0.5 * np.cos(math.pi * ramp)
    return amplitude * env * np.sin(2 * math.pi * t / period)


def inverse_design(folder, n=256, nsteps=3000, design=(64, 64, 32), dx=10e-9, cpml=12, seed=5, lam=1.0e-6,
                   eps_range=(1.0, 12.0), gap=8, magnetic_voxel=False, dtmult=1.0, dispersive=None):
    """Differentiable design problem: a design box of per-voxel eps_r ~ U[eps_range]
    (MEDIUM_LINEAR) centred in x, y; an Ez dipole `gap` cells below it driven by
    a CW ramp at lam; the objective is the DFT |E|^2 at lam summed over a plane
    of the design's x-y extent `gap` cells above it.  CPML on all faces.
    magnetic_voxel puts mu_r = 2 in one voxel far from the design, which makes
    the planner average the E coefficients at material interfaces.
    dispersive = (voxel box i0 j0 k0 i1 j1 k1, material text) adds a MEDIUM_VECTOR
    voxel entity of that material (e.g. an ADE metal block beside the design)."""
    os.makedirs(folder, exist_ok=True)
    rng = np.random.default_rng(seed)
    dnx, dny, dnz = design
    c = n // 2
    i0, j0 = c - dnx // 2, c - dny // 2
    k0 = c - dnz // 2
    box = (i0, j0, k0, i0 + dnx, j0 + dny, k0 + dnz)
    eps = np.ones((n, n, n), dtype=np.float32)
    eps[box[0]:box[3], box[1]:box[4], box[2]:box[5]] = rng.uniform(eps_range[0], eps_range[1], size=design)
    mu = None
    if magnetic_voxel:
        mu = np.ones((n, n, n), dtype=np.float32)
        mu[cpml + 2, cpml + 2, cpml + 2] = 2.0
    write_linear_medium(os.path.join(folder, "design.lin"), eps, mu_r=mu)
    dt = _dt(dx, dx, dx, dtmult)
    write_point_source(os.path.join(folder, "dipole.txt"), cw_ramp(nsteps, dt, lam), comp=2)
    sk = box[2] - gap
    mk = box[5] + gap
    if sk < cpml + 1 or mk > n - cpml - 2:
        raise ValueError("design, source and monitor do not fit between the CPML slabs")
    lines = ["POOL\n%d %d %d %g %g %g\n\n" % (n, n, n, dx, dx, dx), "COMP\n%d\n\n" % nsteps,
             "DT_MULT\n%g\n\n" % dtmult,
             "MEDIUM_LINEAR\ndesign.lin\n\n",
             "BOUNDARY_ALL\ncpml %d 3 %.9g 0 1\n\n" % (cpml, 0.8 * 4 / (ETA0 * dx)),
             "SOURCE_POINT\n%d %d %d 0 dipole.txt\n\n" % (c, c, sk),
             "DESIGN_BOX\n%d %d %d %d %d %d\n\n" % box,
             "OBJECTIVE_DFT\n%d %d %d %d %d %d %.9g\n\n" % (box[0], box[1], mk, box[3], box[4], mk + 1, lam)]
    if dispersive:
        vb, mtext = dispersive
        with open(os.path.join(folder, "dispersive.txt"), "w") as f:
            f.write("8 %d %d %d %d %d %d %s\n" % (tuple(vb) + (mtext,)))
        lines.insert(4, "MEDIUM_VECTOR\ndispersive.txt\n\n")
    path = os.path.join(folder, "inverse_design.par")
    with open(path, "w") as f:
        f.write("".join(lines))
    return path, box, eps[box[0]:box[3], box[1]:box[4], box[2]:box[5]].copy()
