"""Gradient extension: objective J = sum |DFT_lambda(E)|^2 over a monitor box
and its adjoint gradient with respect to the design voxels' eps_r.

The oracle's objective is pinned against a numpy DFT of point traces that the
oracle records on every monitor cell.  The GPU objective is compared with the
oracle's; the GPU adjoint gradient with central finite differences of the
FP64 oracle (h = 1e-3 on eps_r), with and without the interface averaging of
the E coefficients; the checkpointed backward pass with the one that keeps the
whole history.
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
    vox = seeded_voxels(12)
    fd = fd_gradient(folder, par, box, eps, vox, magnetic)
    ad = np.array([grad[v] for v in vox])
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print("adjoint", ad, "\nFD", fd, "\nrel err", err)
    assert err < 1e-3, err


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
