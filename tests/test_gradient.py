"""Gradient extension: objective J = sum |DFT_lambda(E)|^2 over a monitor box
and its adjoint gradient with respect to the design voxels' eps_r.

The oracle's objective is pinned against a numpy DFT of point traces that the
oracle records on every monitor cell.  The GPU objective is compared with the
oracle's; the GPU adjoint gradient with central finite differences of the
FP64 oracle on 16 seeded voxels (h = 1e-3 on eps_r), with and without the
interface averaging of the E coefficients, with a Drude or Drude + 2 CP metal
block in the pool (the transposed ADE step), and with torch autograd through a
pure-torch FDTD of the same scheme at 32^3; the checkpointed backward pass
with the one that keeps the whole history.
"""
import os

import numpy as np
import pytest

import cases
from oracle_lib import Oracle

N, NSTEPS, DESIGN, CPML, GAP = 28, 300, (8, 8, 4), 6, 4
C0 = 299792458.0


def small_design(folder, magnetic=False, nsteps=NSTEPS):
    return cases.inverse_design(folder, n=N, nsteps=nsteps, design=DESIGN, cpml=CPML, gap=GAP,
                                magnetic_voxel=magnetic)


def oracle_J(par):
    o = Oracle(par)
    o.step()
    J = o.objective()
    o.close()
    return J


def rewrite_design(folder, box, eps_design, magnetic):
    eps = np.ones((N, N, N), dtype=np.float32)
    eps[box[0]:box[3], box[1]:box[4], box[2]:box[5]] = eps_design
    mu = None
    if magnetic:
        mu = np.ones((N, N, N), dtype=np.float32)
        mu[CPML + 2, CPML + 2, CPML + 2] = 2.0
    cases.write_linear_medium(os.path.join(folder, "design.lin"), eps, mu_r=mu)


def fd_gradient(folder, par, box, eps, voxels, magnetic, h=1e-3):
    out = []
    for (a, b, c) in voxels:
        vals = []
        for sgn in (1, -1):
            e = eps.copy()
            e[a, b, c] += sgn * h
            rewrite_design(folder, box, e, magnetic)
            vals.append(oracle_J(par))
        out.append((vals[0] - vals[1]) / (2 * h))
    rewrite_design(folder, box, eps, magnetic)
    return np.array(out)


def test_oracle_objective_equals_numpy_dft_of_traces(tmp_path):
    folder = str(tmp_path / "obj")
    par, box, eps = small_design(folder)
    mk = box[5] + GAP
    with open(par, "a") as f:
        for i in range(box[0], box[3]):
            for j in range(box[1], box[4]):
                f.write("OUT_POINT\nAll 1 %d %d %d m_%d_%d\n\n" % (i, j, mk, i, j))
    o = Oracle(par)
    o.step()
    J = o.objective()
    tr = o.points()
    o.close()
    dt = 1.0 / C0 / np.sqrt(3.0 / (10e-9) ** 2)
    w = 2 * np.pi * C0 / 1.0e-6
    t = (np.arange(NSTEPS) + 1.0) * dt
    F = np.sum(tr[:, :NSTEPS, :3] * np.exp(-1j * w * t)[None, :, None], axis=1)
    Jnp = float(np.sum(np.abs(F) ** 2))
    assert J > 0
    assert abs(J - Jnp) / Jnp < 1e-10, (J, Jnp)


def seeded_voxels(n, seed=11):
    rng = np.random.default_rng(seed)
    return [tuple(int(v) for v in (rng.integers(0, DESIGN[0]), rng.integers(0, DESIGN[1]), rng.integers(0, DESIGN[2])))
            for _ in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("magnetic", [False, True], ids=["electric", "averaged"])
def test_adjoint_gradient_against_central_differences(tmp_path, magnetic):
    from gsvit_copy_amd import Session
    folder = str(tmp_path / ("fd_%d" % magnetic))
    par, box, eps = small_design(folder, magnetic)
    J_oracle = oracle_J(par)
    s = Session(par)
    s.adjoint_setup()
    J = s.objective()
    grad = s.objective_gradient()
    s.close()
    print("J gpu", J, "oracle", J_oracle)
    assert abs(J - J_oracle) / J_oracle < 1e-5
    vox = seeded_voxels(16)
    fd = fd_gradient(folder, par, box, eps, vox, magnetic)
    ad = np.array([grad[v] for v in vox])
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print("adjoint", ad, "\nFD", fd, "\nrel err", err)
    assert err < 1e-3, err


# a Drude metal (silver-like) and a Drude + 2 critical-point metal (gold-like), rad/s
DISPERSIVE = {"drude": "5 1 0 1.37e16 2.7e13 0 0 0 0 0 0 0 0",
              "drude_2cp": "5 1.1431 0 1.3202e16 1.0805e14 0.26698 -1.2371 3.8711e15 4.4642e14 "
                           "3.0834 -1.0968 4.1684e15 2.3555e15"}


def dispersive_design(folder, material):
    """the design problem with a metal block beside the design box (x just past
    the design's coefficient cells, outside the CPML slabs, a cell taller than
    the design in z), whose dispersive cells the adjoint steps through"""
    c = N // 2
    i0, j0, k0 = c - DESIGN[0] // 2, c - DESIGN[1] // 2, c - DESIGN[2] // 2
    box = (i0, j0, k0, i0 + DESIGN[0], j0 + DESIGN[1], k0 + DESIGN[2])
    block = (box[3] + 1, box[1], box[2] - 1, box[3] + 3, box[4], box[5] + 1)
    return cases.inverse_design(folder, n=N, nsteps=NSTEPS, design=DESIGN, cpml=CPML, gap=GAP,
                                dispersive=(block, DISPERSIVE[material]))


@pytest.mark.gpu
@pytest.mark.parametrize("material", ["drude", "drude_2cp"])
def test_adjoint_through_dispersive_cells_against_central_differences(tmp_path, material):
    """the transposed ADE step (Drude current, critical-point poles): GPU
    adjoint gradient against central differences of the FP64 oracle with a
    metal block in the pool; the checkpointed backward pass (which restores
    the dispersive state with the fields) equals the full-history one"""
    from gsvit_copy_amd import Session
    folder = str(tmp_path / material)
    par, box, eps = dispersive_design(folder, material)
    print("adjoint through %s cells: %d^3, %d steps" % (material, N, NSTEPS), flush=True)
    J_oracle = oracle_J(par)
    s = Session(par)
    s.adjoint_setup()
    J = s.objective()
    grad = s.objective_gradient()
    s.close()
    s = Session(par)
    s.adjoint_setup(history_budget_bytes=0, ckpt_interval=37)
    J2 = s.objective()
    g2 = s.objective_gradient()
    s.close()
    print("J gpu", J, "oracle", J_oracle)
    assert abs(J - J_oracle) / J_oracle < 1e-5
    assert J2 == J and np.linalg.norm(g2 - grad) / np.linalg.norm(grad) < 1e-6
    vox = seeded_voxels(16)
    fd = fd_gradient(folder, par, box, eps, vox, False)
    ad = np.array([grad[v] for v in vox])
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    # the same voxels without the metal: the block changes the gradient, so the check sees the transpose
    plain = small_design(str(tmp_path / "plain"))[0]
    s = Session(plain)
    s.adjoint_setup()
    s.objective()
    g0 = s.objective_gradient()
    s.close()
    ad0 = np.array([g0[v] for v in vox])
    print("%s: adjoint vs FD rel err %.3e; gradient without the block differs by %.3e"
          % (material, err, np.linalg.norm(ad0 - fd) / np.linalg.norm(fd)))
    assert err < 1e-3, err
    assert np.linalg.norm(ad0 - fd) / np.linalg.norm(fd) > 1e-2


@pytest.mark.gpu
def test_checkpointed_backward_matches_full_history(tmp_path):
    from gsvit_copy_amd import Session
    par, box, eps = small_design(str(tmp_path / "ck"))
    s1 = Session(par)
    s1.adjoint_setup()
    J1 = s1.objective()
    g1 = s1.objective_gradient()
    info1 = s1.adjoint_info()
    s1.close()
    s2 = Session(par)
    s2.adjoint_setup(history_budget_bytes=0, ckpt_interval=37)
    J2 = s2.objective()
    g2 = s2.objective_gradient()
    info2 = s2.adjoint_info()
    s2.close()
    print(info1, info2)
    assert info1["history_all"] and not info2["history_all"] and info2["checkpoints"] == (NSTEPS + 36) // 37
    assert J1 == J2
    assert np.linalg.norm(g1 - g2) / np.linalg.norm(g1) < 1e-6


@pytest.mark.gpu
def test_set_design_eps_equals_file(tmp_path):
    """Setting eps_r through the API gives the same J as writing it in the file."""
    from gsvit_copy_amd import Session
    folder = str(tmp_path / "set")
    par, box, eps = small_design(folder)
    s = Session(par)
    s.adjoint_setup()
    # eps is stored as float(eps_r) * float(eps0), as the MEDIUM_LINEAR reader does
    assert np.allclose(s.design_eps(), eps, rtol=1e-6, atol=0)
    e2 = eps.copy()
    e2[2:5, 3:6, 1:3] += 0.75
    s.set_design_eps(e2)
    J_api = s.objective()
    s.close()
    rewrite_design(folder, box, e2, False)
    s = Session(par)
    s.adjoint_setup()
    J_file = s.objective()
    s.close()
    assert J_api == J_file


def torch_problem(folder, n=32, nsteps=240, design=(8, 8, 4), cpml=6, gap=3):
    """The inverse-design problem at 32^3, a size the pure-torch FDTD differentiates quickly."""
    import torch_fdtd
    par, box, eps = cases.inverse_design(folder, n=n, nsteps=nsteps, design=design, cpml=cpml, gap=gap)
    c = n // 2
    dx = 10e-9
    dt = 1.0 / C0 / np.sqrt(3.0 / dx ** 2)
    src = cases.cw_ramp(nsteps, dt, 1.0e-6)
    # the source file is written with %.17g and read back as double
    src = [float("%.17g" % v) for v in src]
    mk = box[5] + gap
    sigma = float("%.9g" % (0.8 * 4 / (376.73 * dx)))   # as written in the parameter file
    prob = torch_fdtd.Problem(n, dx, cpml, 3, sigma, 1.0, box, (c, c, box[2] - gap), src,
                              (box[0], box[1], mk, box[3], box[4], mk + 1), 1.0e-6)
    return par, box, eps, prob


def test_torch_fdtd_reproduces_oracle_objective(tmp_path):
    """The pure-torch FDTD is the same discrete scheme: J agrees with the FP64 oracle."""
    import torch
    par, box, eps, prob = torch_problem(str(tmp_path / "t"))
    J_oracle = oracle_J(par)
    # the CPU path stores float32(eps_r) * float32(eps0) per voxel
    e32 = (eps.astype(np.float32) * np.float32(8.854187817620389850e-12)).astype(np.float64) / 8.854187817620389850e-12
    bg = float(np.float32(1.0) * np.float32(8.854187817620389850e-12))
    with torch.no_grad():
        J_torch = float(prob.objective(torch.tensor(e32, dtype=torch.float64), background_eps=bg))
    assert J_oracle > 0
    # fields agree to ~1e-10; the DFT over ~1.2 periods cancels, so J is held to 1e-7
    # (leaving out the interface averaging, for instance, moves J by 5e-4)
    assert abs(J_torch - J_oracle) / J_oracle < 1e-7, (J_torch, J_oracle)


@pytest.mark.gpu
def test_adjoint_gradient_against_torch_autograd(tmp_path):
    """Engine adjoint (FP32) vs torch autograd through the pure-torch FDTD (FP64), whole design box."""
    import torch
    from gsvit_copy_amd import Session
    from gsvit_copy_amd.autograd import DesignObjective
    par, box, eps, prob = torch_problem(str(tmp_path / "ta"))
    e = torch.tensor(eps.astype(np.float64), requires_grad=True)
    J_t = prob.objective(e)
    J_t.backward()
    g_torch = e.grad.numpy()
    s = Session(par)
    s.adjoint_setup()
    e_gpu = torch.tensor(eps, dtype=torch.float32, requires_grad=True)
    J_g = DesignObjective.apply(e_gpu, s)
    J_g.backward()
    g_gpu = e_gpu.grad.numpy().astype(np.float64)
    s.close()
    print("J torch", float(J_t), "gpu", float(J_g))
    err = np.linalg.norm(g_gpu - g_torch) / np.linalg.norm(g_torch)
    print("gradient rel L2 vs torch autograd", err)
    assert abs(float(J_g) - float(J_t)) / float(J_t) < 1e-5
    assert err < 1e-3, err
