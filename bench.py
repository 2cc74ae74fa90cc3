"""Benchmark of the MI355X FDTD engine on the headline workload.

Workload (one GPU): rough-surface scattering at 512^3, 5 nm cells, lower half
a Drude metal (ADE), CPML 16, TSF p-polarised plane wave at 45 degrees,
running-DFT near-to-far-field box at 3 wavelengths.  A "step" is one full
time step of the pool (H and E updates with CPML, dispersive currents, TSF
faces, DFT accumulation).

Output: one JSON line with Gcell-updates/s (whole job), the roofline of the
dominant kernel (algorithmic HBM bytes per launch / launch time, both
measured live with HIP events on the engine's stream) and the CPU baseline
(the oracle's FP64 restatement of the reference CPU path on a bounded sample).
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s + % HBM peak at 512\u00b3; adjoint-grad time \u00f7 plain forward"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(info, shape, depth, n_ade_cells, engine=False):
    """Algorithmic HBM bytes of one step, per kernel, by the reference
    formulation's per-cell figures.

    Yee core per half-step: read E(3) + H(3), write the updated field (3) in
    FP32 = 36 B/cell, plus a 1-byte class id = 37 B/cell (74 B per step).
    CPML: per stretched direction inside a slab, 2 psi read + 2 written per
    half-step = 16 B/cell/half-step.  Drude ADE: the reference's state E(n-1),
    dP, dP(n-1) per component read and written in the E half-step = 72 B per
    dispersive cell.  With engine=True the E count is the engine's own: a
    Drude-only pool keeps one state per component (Q = cj J + c2 E(n-1), 24 B
    per dispersive cell read and written), so its kernels move fewer bytes
    than the reference-formulation count; the measured traffic is reported
    beside both.
    """
    nx, ny, nz = (shape, shape, shape) if np.isscalar(shape) else shape
    cells = float(nx) * ny * nz
    slab_dir_cells = 2.0 * depth * (float(ny) * nz + float(nx) * nz + float(nx) * ny)   # cell-directions in the six slabs
    h_bytes = 37.0 * cells + 16.0 * slab_dir_cells
    e_bytes = 37.0 * cells + 16.0 * slab_dir_cells + (24.0 if engine else 72.0) * n_ade_cells
    return h_bytes, e_bytes


def copy_probe(nbytes, reps, variant):
    """ms per copy of the streaming-copy probe (bench_probe.hip, built as
    gsvit_copy_amd/lib/libgsv_probe.so): reads nbytes and writes nbytes"""
    import ctypes
    L = ctypes.CDLL(os.path.join(HERE, "gsvit_copy_amd", "lib", "libgsv_probe.so"))
    L.gsv_probe_copy.restype = ctypes.c_int
    L.gsv_probe_copy.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    ms = ctypes.c_double()
    if L.gsv_probe_copy(int(nbytes), int(reps), int(variant), ctypes.byref(ms)):
        raise RuntimeError("copy probe failed")
    return ms.value


def attainable_bandwidth():
    """Attainable HBM bandwidth: the best of the probe's copy variants over
    2 GiB (read 2 GiB + write 2 GiB per copy, 20 timed copies each)"""
    nbytes = 2 << 30
    best = None
    for v in range(3):
        ms = copy_probe(nbytes, 20, v)
        gbs = 2 * nbytes / (ms * 1e-3) / 1e9
        log("bench: copy probe variant %d: %.3f ms per 2 GiB copy, %.1f GB/s" % (v, ms, gbs))
        if best is None or gbs > best["GB_s"]:
            best = {"GB_s": round(gbs, 1), "variant": v, "bytes_per_copy": 2 * nbytes}
    return best


def adjoint_bench(args, work, shape):
    """Adjoint-gradient cost on the inverse-design workload: 256^3 cells of
    10 nm, CPML 12, a 64x64x32 design region of eps_r ~ U[1,12], CW Ez dipole
    at 1 um, 3000 steps, DFT |E|^2 objective on a 64x64 plane.  Ratio = (forward
    run + backward run) / one plain forward run of the same problem, both
    wall-clock with all data resident on the GPU.  Measured as the workload
    specifies it -- full field checkpoints every 100 steps, each segment
    recomputed before its backward pass -- and, beside it, with the whole
    run's design-region history kept in HBM (no recomputation)."""
    import cases
    from gsvit_copy_amd import Session
    n, nsteps = args.adjoint_n, args.adjoint_steps
    design = (n // 4, n // 4, n // 8)
    t0 = time.time()
    par, box, eps = cases.inverse_design(os.path.join(work, "inverse"), n=n, nsteps=nsteps, design=design)
    log("bench: inverse-design workload %d^3 written in %.1f s" % (n, time.time() - t0))
    s = Session(par)
    s.set_launch(shape)
    s.run(5)
    s.sync()
    s.close()
    s = Session(par)
    s.set_launch(shape)
    t0 = time.perf_counter()
    s.run(nsteps)
    s.sync()
    plain = time.perf_counter() - t0
    s.close()
    modes = {}
    for name, budget, interval in (("checkpointed", 0, 100), ("history", 16 << 30, 100)):
        s = Session(par)
        s.set_launch(shape)
        s.adjoint_setup(history_budget_bytes=budget, ckpt_interval=interval)
        t0 = time.perf_counter()
        J = s.objective()
        grad = s.objective_gradient()
        total = time.perf_counter() - t0
        info = s.adjoint_info()
        s.close()
        log("bench: adjoint (%s): plain forward %.3f s, forward+gradient %.3f s (ratio %.2f), J %.6g, |grad| %.4g"
            % (name, plain, total, total / plain, J, float(np.linalg.norm(grad))))
        modes[name] = {"gradient_s": round(total, 4), "ratio": round(total / plain, 3),
                       "forward_ms": round(info["forward_ms"], 2), "backward_ms": round(info["backward_ms"], 2),
                       "history_all": info["history_all"], "ckpt_interval": info["ckpt_interval"],
                       "checkpoints": info["checkpoints"]}
    ck = modes["checkpointed"]
    return {"workload": "inverse design: %d^3, 10 nm cells, CPML 12, %dx%dx%d eps_r design region, CW Ez dipole "
                        "1 um, DFT |E|^2 objective on a %dx%d plane" % (n, design[0], design[1], design[2],
                                                                        design[0], design[1]),
            "steps": nsteps, "plain_forward_s": round(plain, 4), "gradient_s": ck["gradient_s"],
            "ratio": ck["ratio"], "mode": "checkpoints every %d steps, recomputation" % ck["ckpt_interval"],
            "modes": modes, "metric": "adjoint-grad time / plain forward", "higher_is_better": False}


def batch_bench(args, work):
    """Batched sweep: 64 seeded rough-surface runs (Drude half space, TSF
    p-polarised at U[0,70] deg, centre wavelength U[450,750] nm, rms U[5,30] nm,
    CPML 10, 5 nm cells) with grid edges drawn from {64,96,128,192,256} per axis.
    Sequential: each run alone through the plain session path (default launch
    shape), `batch_steps` timed steps per run after 5 untimed ones, times
    summed.  Batched: all 64 in one batch (autotuned tile shapes, steps captured
    as HIP graphs of 16), the same number of timed steps.  Throughput is total
    cell updates / time; the ratio batched / sequential."""
    import cases
    from gsvit_copy_amd import Batch, Session
    S = args.batch_steps
    t0 = time.time()
    paths, params = cases.sweep(os.path.join(work, "sweep"), nruns=64, seed=11, nsteps=2000)
    cells = sum(p["shape"][0] * p["shape"][1] * p["shape"][2] for p in params)
    log("bench: 64-run sweep (%.1f Mcells) written in %.1f s" % (cells / 1e6, time.time() - t0))
    t_seq, seq = None, None
    single_h = single_e = 0.0
    if not args.batch_only:
        t_seq = 0.0
        for p in paths:
            with Session(p) as s:
                s.run(5)
                s.sync()
                t0 = time.perf_counter()
                s.run(S)
                s.sync()
                t_seq += time.perf_counter() - t0
                # the same volume kernels per run, timed with events (after the timed steps)
                s.set_timing(True)
                s.run(5)
                kt = s.kernel_times()
                single_h += kt["h_ms"] / max(kt["h_n"], 1)
                single_e += kt["e_ms"] / max(kt["e_n"], 1)
        seq = cells * S / t_seq / 1e9
        log("bench: sequential plain runs %.2f Gcell-updates/s; their volume kernels H %.3f ms E %.3f ms per sweep step"
            % (seq, single_h, single_e))
    t0 = time.time()
    sessions = [Session(p) for p in paths]
    t_open = time.time() - t0
    # algorithmic bytes of one batched step (the single-run per-cell figures, summed over the runs)
    h_bytes = e_bytes = 0.0
    for s, p in zip(sessions, params):
        mat, _ = s.maps()
        hb, eb = algorithmic_bytes(None, p["shape"], 10, int(dilate(mat == 1).sum()))
        h_bytes += hb
        e_bytes += eb
        del mat
    b = Batch(sessions)
    if args.batch_launch:
        best, used = [int(v) for v in args.batch_launch.split(",")], 0
        b.set_launch(best)
    else:
        used, best = b.autotune(2)
    b.set_graph(16)
    b.run(16)
    t0 = time.perf_counter()
    b.run(S)
    t_b = time.perf_counter() - t0
    b.set_graph(0)
    b.set_timing(True)
    b.run(5)
    inf = b.info()
    b.close()
    for s in sessions:
        s.close()
    bat = cells * S / t_b / 1e9
    h_gbs = h_bytes / (inf["h_ms"] * 1e-3) / 1e9
    e_gbs = e_bytes / (inf["e_ms"] * 1e-3) / 1e9
    log("bench: batched %.2f Gcell-updates/s (x%s), launch %s, H %.3f ms (%.0f GB/s) E %.3f ms (%.0f GB/s) per step"
        % (bat, "%.2f" % (bat / seq) if seq else "-", best, inf["h_ms"], h_gbs, inf["e_ms"], e_gbs))
    return {"workload": "64-run rough-surface sweep, grid edges from {64,96,128,192,256} per axis, Drude ADE, "
                        "CPML 10, TSF p-pol U[0,70] deg, lambda U[450,750] nm",
            "runs": 64, "cells": cells, "timed_steps": S,
            "sequential_gcells": round(seq, 3) if seq else None, "batched_gcells": round(bat, 3),
            "ratio": round(bat / seq, 3) if seq else None,
            "sequential_ms_per_step": round(1e3 * t_seq / S, 3) if seq else None,
            "batched_ms_per_step": round(1e3 * t_b / S, 3),
            "open_64_sessions_s": round(t_open, 2), "autotune_steps": used, "launch": best, "graph_steps": 16,
            "h_ms": round(inf["h_ms"], 4), "e_ms": round(inf["e_ms"], 4),
            "h_algorithmic_bytes": h_bytes, "e_algorithmic_bytes": e_bytes,
            "single_run_kernels_h_ms": round(single_h, 4) if seq else None,
            "single_run_kernels_e_ms": round(single_e, 4) if seq else None,
            "h_frac": round(h_gbs / HBM_PEAK_GBS, 4), "e_frac": round(e_gbs / HBM_PEAK_GBS, 4),
            "metric": "batched / sequential throughput", "higher_is_better": True}


def kernel_table(kt, h_bytes, e_bytes, traffic=None):
    """Per-kernel times (HIP events) against the H and E passes' algorithmic bytes."""
    traffic = traffic or {}
    frac = lambda b, ms: round(b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None
    h_ms = kt["h_ms"] / max(kt["h_n"], 1)
    e_ms = kt["e_ms"] / max(kt["e_n"], 1)
    tab = {"k_hstep": {"ms": round(h_ms, 4), "algorithmic_bytes": h_bytes, "frac": frac(h_bytes, h_ms),
                       "traffic": traffic.get("k_hstep")},
           "k_estep": {"ms": round(e_ms, 4), "algorithmic_bytes": e_bytes, "frac": frac(e_bytes, e_ms),
                       "traffic": traffic.get("k_estep")}}
    if e_ms >= h_ms:
        return tab, "k_estep", e_ms, e_bytes
    return tab, "k_hstep", h_ms, h_bytes


def dilate(mask):
    """cells that take a material's dispersive branch: the material's own
    cells and the +1 neighbours whose averaged coefficients it feeds"""
    out = mask.copy()
    out[1:, :, :] |= mask[:-1, :, :]
    out[:, 1:, :] |= mask[:, :-1, :]
    out[:, :, 1:] |= mask[:, :, :-1]
    out[1:, 1:, :] |= mask[:-1, :-1, :]
    out[:, 1:, 1:] |= mask[:, :-1, :-1]
    out[1:, :, 1:] |= mask[:-1, :, :-1]
    return out


def film_bench(args, work):
    """Dispersive textured thin film: 768 x 768 x 384 cells of 4 nm, periodic
    in x and y, CPML 16 in z; SiO2 / three-pole silicon with seeded pyramids /
    Drude + 2 CP gold; normally incident 400-1100 nm pulse; flux DFT on three
    planes at 32 wavelengths.  Timed: `film_steps` steps after 20 untimed ones,
    at the fixed launch shape `film_launch` (or autotuned with `film_tune`).  Algorithmic bytes by the reference formulation: 74 B
    per cell and step, 32 B per CPML cell, and per dispersive cell the
    reference's state arrays (Drude + 2 CP: 21 arrays = +168 B; three poles:
    27 arrays = +216 B).  With `film_full` the whole 6000-step run is made and
    the flux spectra are turned into reflectance, the fraction absorbed in
    the silicon (and the lossless SiO2) between the first two planes, the
    fraction absorbed in the gold contact and the fraction transmitted behind
    it, normalised by the incident power of the same source through the same
    planes in vacuum."""
    import cases
    from gsvit_copy_amd import Session
    nx, ny, nz, depth, dx = 768, 768, 384, 16, 4e-9
    t0 = time.time()
    par = cases.textured_film(os.path.join(work, "film"), nx=nx, ny=ny, nz=nz, dx=dx, nsteps=6000)
    log("bench: textured film written in %.1f s" % (time.time() - t0))
    t0 = time.time()
    s = Session(par)
    info = s.info()
    t_open = time.time() - t0
    mat, _ = s.maps()
    si = dilate(mat == 2)
    au = dilate(mat == 3) & ~si
    n_si, n_au = int(si.sum()), int(au.sum())
    del mat, si, au
    cells = float(nx * ny * nz)
    cpml_cells = 2.0 * depth * nx * ny
    h_bytes = 37.0 * cells + 16.0 * cpml_cells
    e_bytes = 37.0 * cells + 16.0 * cpml_cells + 216.0 * n_si + 168.0 * n_au
    if args.film_tune:
        shape, used = s.autotune(2)
    else:
        shape, used = [int(v) for v in args.film_launch.split(",")], 0
        s.set_launch(shape)
    s.run(20)
    s.sync()
    S = args.film_steps
    t0 = time.perf_counter()
    s.run(S)
    s.sync()
    el = time.perf_counter() - t0
    s.set_timing(True)
    s.run(10)
    kt = s.kernel_times()
    s.set_timing(False)
    gc = cells * S / el / 1e9
    traffic = {}
    try:
        with open(os.path.join(HERE, "profiles", "hbm_traffic_film768.json")) as f:
            tr = json.load(f)
        if tr["launch_shape"] == list(shape):
            traffic = {k: round(v["traffic_bytes"]) for k, v in tr["kernels"].items()}
    except (OSError, ValueError, KeyError):
        pass
    log("bench: textured film %.3f ms/step, %.2f Gcell-updates/s; kernels %s"
        % (1e3 * el / S, gc, {k: v["ms"] for k, v in kernel_table(kt, h_bytes, e_bytes)[0].items()}))
    out = {"workload": "dispersive textured thin film 768x768x384, 4 nm, periodic x/y, CPML 16 z, SiO2 / 3-pole Si "
                       "with pyramids / Drude+2CP Au, TSF normal 400-1100 nm, flux DFT 3 planes x 32 wavelengths",
           "cells": int(cells), "si_ade_cells": n_si, "au_ade_cells": n_au, "ade_poles": info["ade_poles"],
           "timed_steps": S, "ms_per_step": round(1e3 * el / S, 4), "gcells": round(gc, 3),
           "open_s": round(t_open, 2), "autotune_steps": used, "launch": shape,
           "kernels": kernel_table(kt, h_bytes, e_bytes, traffic)[0],
           "traffic_source": "profiles/hbm_traffic_film768.md" if traffic else None,
           "metric": "Gcell-updates/s", "higher_is_better": True}
    if args.film_full:
        t0 = time.perf_counter()
        s.run()
        s.sync()
        t_rest = time.perf_counter() - t0
        steps = s.info()["step"]
        fl = s.flux()
        f = s.fields()
        finite = bool(np.isfinite(f).all())
        del f
        # incident power per wavelength: the same source and planes in vacuum;
        # the normally incident field is uniform in x and y, so a 16 x 16 column
        # of the periodic cell scaled by area gives it
        vpar = cases.textured_film(os.path.join(work, "film_vacuum"), nx=16, ny=16, nz=nz, dx=dx, nsteps=6000,
                                   media=False)
        with Session(vpar) as v:
            v.run()
            pin = v.flux()[0] * (nx * ny) / (16.0 * 16.0)
        lam = np.linspace(400e-9, 1100e-9, fl.shape[1])
        out["full_run"] = {"steps": steps, "rest_s": round(t_rest, 2), "fields_finite": finite,
                           "wavelength_nm": [round(float(v) * 1e9, 1) for v in lam],
                           "reflectance": [round(float(v), 4) for v in 1 - fl[0] / pin],
                           "absorbed_si": [round(float(v), 4) for v in (fl[0] - fl[1]) / pin],
                           "absorbed_au": [round(float(v), 4) for v in (fl[1] - fl[3]) / pin],
                           "transmitted": [round(float(v), 6) for v in fl[3] / pin]}
        log("bench: textured film full run: %d steps, %.1f s after the timed part, finite %s" % (steps, t_rest, finite))
    s.close()
    return out


_BATCH_E2E = """
import sys
sys.path.insert(0, sys.argv[1])
from gsvit_copy_amd import Batch, Session
sessions = [Session(p) for p in sys.argv[2:]]
b = Batch(sessions)
b.autotune(2)
b.set_graph(16)
b.run(1 << 30)
b.close()
for s in sessions:
    s.close()
"""


def batch_e2e(work):
    """End-to-end sweep time (--batch-e2e): the 64 runs of the batched sweep,
    all 2000 steps, as one process per run through the command-line driver
    (the reference's usage: parse, plan, upload, step), against one process
    that opens all 64 and steps them as a batch (its autotuning steps count
    towards the 2000).  Wall-clock of the child processes, GPU initialisation
    included in both."""
    import subprocess
    paths = [os.path.join(work, "sweep", "run%02d" % r, "sweep_run.par") for r in range(64)]
    exe = os.path.join(HERE, "gsvit_copy_amd", "bin", "gsvit_amd")
    t0 = time.perf_counter()
    for r, p in enumerate(paths):
        subprocess.run([exe, os.path.basename(p)], cwd=os.path.dirname(p), stdout=subprocess.DEVNULL, check=True)
        if r % 16 == 15:
            log("bench: %d runs, one process each, %.1f s" % (r + 1, time.perf_counter() - t0))
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    subprocess.run([sys.executable, "-c", _BATCH_E2E, HERE] + paths, check=True)
    t_bat = time.perf_counter() - t0
    log("bench: end to end, 64 x 2000 steps: one process per run %.1f s, one batch %.1f s (x%.2f)"
        % (t_seq, t_bat, t_seq / t_bat))
    return {"steps": 2000, "one_process_per_run_s": round(t_seq, 2), "batched_process_s": round(t_bat, 2),
            "ratio": round(t_seq / t_bat, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--probe-only", type=int, default=None, metavar="VARIANT",
                    help="only run the streaming-copy probe (1 GiB, 3 copies) of the given variant and exit "
                         "(counter calibration under rocprofv3)")
    ap.add_argument("--kernel-steps", type=int, default=50)
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--tune-steps", type=int, default=2, help="steps per launch-shape candidate (0 = no tuning)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient measurement")
    ap.add_argument("--adjoint-n", type=int, default=256)
    ap.add_argument("--adjoint-steps", type=int, default=3000)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-sweep measurement")
    ap.add_argument("--batch-steps", type=int, default=400, help="timed steps of the batched-sweep comparison")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched leg, without the sequential runs (profiling); prints its JSON object")
    ap.add_argument("--batch-launch", default=None,
                    help="fixed batch launch shape (8 numbers as gsv_batch_set_launch; implies no batch tuning)")
    ap.add_argument("--batch-e2e", action="store_true",
                    help="also time the whole 64-run sweep end to end (one process per run vs one batch)")
    ap.add_argument("--no-film", action="store_true", help="skip the textured thin-film measurement")
    ap.add_argument("--film-steps", type=int, default=300, help="timed steps of the textured thin film")
    ap.add_argument("--film-launch", default="256,2,8,256,1,8",
                    help="launch shape of the textured film (as --launch; the autotuner's usual choice, and the "
                         "shape of the committed counter summary)")
    ap.add_argument("--film-tune", action="store_true", help="autotune the textured film's launch shape instead")
    ap.add_argument("--film-full", action="store_true",
                    help="also run the textured thin film to its 6000 steps and report its spectra")
    ap.add_argument("--launch", default=None,
                    help="fixed launch shape hx,hy,hz,ex,ey,ez (tile x width, rows, z-march depth of the H and E "
                         "kernels); implies no tuning")
    args = ap.parse_args()

    if args.probe_only is not None:
        ms = copy_probe(1 << 30, 3, args.probe_only)
        print(json.dumps({"probe_variant": args.probe_only, "bytes_read": 1 << 30, "bytes_written": 1 << 30,
                          "ms": ms, "GB_s": 2 * (1 << 30) / (ms * 1e-3) / 1e9}), flush=True)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path (a z-slab decomposition "
              "with halo exchange is the planned sharding); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.batch_only:
        work = args.workdir or tempfile.mkdtemp(prefix="gsv_bench_")
        batch = batch_bench(args, work)
        if args.batch_e2e:
            batch["end_to_end"] = batch_e2e(work)
        print(json.dumps(batch), flush=True)
        return

    import cases
    from gsvit_copy_amd import Session

    n = args.n
    depth = 16
    ncand = 45
    nsteps = ncand * args.tune_steps + args.warmup + args.steps + args.kernel_steps
    work = args.workdir or tempfile.mkdtemp(prefix="gsv_bench_")
    t0 = time.time()
    log("bench: writing %d^3 rough-surface workload to %s" % (n, work))
    par = cases.rough_surface(work, n=n, nsteps=nsteps)
    log("bench: workload written in %.1f s; planning" % (time.time() - t0))
    t0 = time.time()
    s = Session(par)
    info = s.info()
    log("bench: plan + upload %.1f s; info %s" % (time.time() - t0, info))

    # ADE cells (dispersive cells and their +1 neighbours take the ADE branch)
    mat, bnds = s.maps()
    n_ade = int(dilate(mat == 1).sum())
    del mat, bnds

    shape = s.get_launch()
    if args.launch:
        shape = [int(v) for v in args.launch.split(",")]
        s.set_launch(shape)
    elif args.tune_steps > 0:
        shape, used = s.autotune(args.tune_steps)
        log("bench: autotuned in %d steps: launch shape %s" % (used, shape))
    log("bench: warmup %d steps" % args.warmup)
    s.run(args.warmup)
    s.sync()
    log("bench: timing %d steps" % args.steps)
    t0 = time.perf_counter()
    s.run(args.steps)
    s.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    gcells = float(n) ** 3 * args.steps / elapsed / 1e9
    log("bench: %.3f ms/step, %.2f Gcell-updates/s" % (ms_per_step, gcells))

    # per-kernel times with HIP events on the engine stream, in steps after the
    # timed region (the per-launch event synchronisation would perturb it)
    s.set_timing(True)
    s.run(args.kernel_steps)
    kt = s.kernel_times()
    s.close()
    h_bytes, e_bytes = algorithmic_bytes(info, n, depth, n_ade)
    ktab, dom, dom_ms, dom_bytes = kernel_table(kt, h_bytes, e_bytes)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_bytes = h_bytes + e_bytes
    log("bench: kernels %s" % ({k: (v["ms"], v["frac"]) for k, v in ktab.items()},))
    kshape = list(shape)

    # HBM traffic of the dominant kernel from the committed counter summary
    # (rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by the copy-kernel
    # calibration), used only when it describes this workload and launch shape
    traffic, traffic_src, traffic_shape = None, None, None
    try:
        with open(os.path.join(HERE, "profiles", "hbm_traffic_rough512.json")) as f:
            tr = json.load(f)
        if n == 512 and dom in tr["kernels"]:
            traffic = round(tr["kernels"][dom]["traffic_bytes"])
            traffic_src = "profiles/hbm_traffic_rough512.md"
            traffic_shape = tr["launch_shape"]   # the counters' launch shape (the halo re-reads depend on it)
    except (OSError, ValueError, KeyError):
        pass
    engine_bytes = algorithmic_bytes(info, n, depth, n_ade, engine=True)[1 if dom == "k_estep" else 0]

    attainable = None
    try:
        attainable = attainable_bandwidth()
    except Exception as exc:  # reported beside the roofline, never required
        log("bench: copy probe skipped: %r" % (exc,))

    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_bench(args, work, shape)

    batch = None
    if not args.no_batch:
        batch = batch_bench(args, work)
        if args.batch_e2e:
            batch["end_to_end"] = batch_e2e(work)

    film = None
    if not args.no_film:
        film = film_bench(args, work)

    cpu = None
    if not args.no_cpu_baseline:
        try:
            from oracle_lib import Oracle
            nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            cn = args.cpu_n
            log("bench: CPU baseline (FP64 oracle, %d threads), %d steps at %d^3" % (nthreads, args.cpu_steps, cn))
            cpar = cases.rough_surface(os.path.join(work, "cpu%d" % cn), n=cn, nsteps=args.cpu_steps + 1)
            t0 = time.time()
            o = Oracle(cpar)
            log("bench: oracle planned in %.1f s" % (time.time() - t0))
            o.step(1)
            t0 = time.perf_counter()
            o.step(args.cpu_steps)
            ce = time.perf_counter() - t0
            o.close()
            cpu = {"value": float(cn) ** 3 * args.cpu_steps / ce / 1e9, "unit": "Gcell-updates/s", "cores": nthreads,
                   "kind": "port",
                   "sample": "%d steps (after 1 untimed) of the same workload at %d^3, FP64 fields, float materials, "
                             "OpenMP over x (the reference CPU path's layout and loop order)" % (args.cpu_steps, cn)}
            log("bench: CPU %.4f Gcell-updates/s" % cpu["value"])
        except Exception as exc:  # the baseline is reported, never required
            log("bench: CPU baseline skipped: %r" % (exc,))

    out = {
        "metric": METRIC,
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "rough-surface scattering, Drude ADE half space, CPML 16, TSF 45 deg p-pol, "
                               "running-DFT NFFF at 3 wavelengths",
                   "grid": [n, n, n], "cell_nm": 5, "ade_cells": n_ade,
                   "e_classes": info["e_classes"], "h_classes": info["h_classes"],
                   "launch_shape": kshape},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel_ms": round(dom_ms, 4), "algorithmic_bytes": dom_bytes,
                     "step_algorithmic_bytes": step_bytes,
                     "kernels": ktab,
                     "step_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "attainable": attainable,
                     "traffic_launch_shape": traffic_shape,
                     "traffic_GB_s": round(traffic / (dom_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_frac_of_attainable": (round(traffic / (dom_ms * 1e-3) / 1e9 / attainable["GB_s"], 4)
                                                    if traffic and attainable else None),
                     "engine_bytes": engine_bytes,
                     "engine_frac": round(engine_bytes / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "batch": batch,
        "film": film,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
