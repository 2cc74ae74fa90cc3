"""A small pure-torch FDTD of the gradient extension's problem class, used by
the tests as an independent gradient: torch autograd through the whole time
loop (FP64, CPU) instead of the engine's hand-written adjoint.

TEST INFRASTRUCTURE ONLY.  It restates, for cubic cells, non-dispersive
per-voxel permittivity with sigma = 0 (a MEDIUM_LINEAR file), CPML on all
faces with kappa = 1 and alpha = 0, and one hard Ez point source, the same
discrete scheme as the CPU path:
  H update ............. src3d/yee.c:145-265
  CPML H psi ........... src3d/boundary.c:347-518 (far-plane zeroing included)
  E update ............. src3d/yee.c:528-1247 (cb = dt / (eps d), eps averaged
                         per component over the voxels src3d/yee.c:1396-1491
                         names where the boundary map src3d/plan.c:5381-5417
                         marks the cell)
  CPML E psi ........... src3d/boundary.c:691-887 (near-plane zeroing included)
  CPML grading ......... src3d/boundary.c:33-99 (becebhch0 / becebhch)
  hard point source .... src3d/source.c:405-437
and the objective J = sum over the monitor box of |DFT_lambda(E)|^2 with E
after step n sampled at (n+1) dt.  The tests pin it against the oracle's J.
"""
import math

import numpy as np
import torch

C0 = 299792458.0
MU0 = 1.256637061435917295e-6
EPS0 = 8.854187817620389850e-12


def grading(depth, m, sigma, a, kappa, dt, near):
    """E-side (b, c) and H-side (b, c) per slab index (src3d/boundary.c:33-99)."""
    be, ce, bh, ch = (np.zeros(depth) for _ in range(4))
    for i in range(depth):
        arg = i if near else depth - i
        off = -0.5 if near else 0.5
        pe = ((depth - arg) / depth) ** m
        ph = ((depth - arg + off) / depth) ** m
        sve, kve, ave = sigma * pe, 1 + (kappa - 1) * pe, a * (arg / depth)
        svh, kvh, avh = sigma * ph, 1 + (kappa - 1) * ph, a * ((arg + off) / depth)
        be[i] = math.exp(-(sve / kve + ave) * dt / EPS0)
        ce[i] = 0.0 if sve == 0 else sve / (sve + kve * ave) / kve * (be[i] - 1.0)
        bh[i] = math.exp(-(svh / kvh + avh) * dt / EPS0)
        ch[i] = 0.0 if svh == 0 else svh / (svh + kvh * avh) / kvh * (bh[i] - 1.0)
    return be, ce, bh, ch


class Problem:
    """n^3 cubic cells of size d, CPML depth `cpml` (m, sigma, a = 0, kappa = 1)
    on all faces, background eps_r = 1, design box of eps_r, hard Ez source
    values `src` at cell `src_pos`, monitor box and wavelength."""

    def __init__(self, n, d, cpml, m, sigma, dtmult, box, src_pos, src, mon_box, lam):
        self.n, self.d, self.D = n, d, cpml
        self.dt = 1.0 / C0 / math.sqrt(3.0 / d ** 2) * dtmult
        self.box, self.src_pos, self.src, self.mon, self.lam = box, src_pos, src, mon_box, lam
        nb, nc, nbh, nch = grading(cpml, m, sigma, 0.0, 1.0, self.dt, True)
        fb, fc, fbh, fch = grading(cpml, m, sigma, 0.0, 1.0, self.dt, False)
        # coefficient per index along an axis: near slab [0, D), far slab [n - D, n)
        self.be, self.ced, self.bh, self.chd = (np.zeros(n) for _ in range(4))
        self.in_e = np.zeros(n, bool)   # E-side psi cells: [1, D) and [n - D + 1, n)
        self.in_h = np.zeros(n, bool)   # H-side psi cells: [0, D - 1) and [n - D + 1, n - 1)
        for r in range(cpml):
            self.be[r], self.ced[r], self.bh[r], self.chd[r] = nb[r], nc[r] / d, nbh[r], nch[r] / d
            q = n - cpml + r
            self.be[q], self.ced[q], self.bh[q], self.chd[q] = fb[r], fc[r] / d, fbh[r], fch[r] / d
        self.in_e[1:cpml] = True
        self.in_e[n - cpml + 1:] = True
        self.in_h[:cpml - 1] = True
        self.in_h[n - cpml + 1:n - 1] = True

    def objective(self, eps_design, background_eps=EPS0):
        """J for the design box's eps_r (a torch tensor, differentiable).
        background_eps: absolute permittivity outside the design box (the CPU
        path stores float32(1) * float32(eps0) there)."""
        n, d, dt = self.n, self.d, self.dt
        t64 = dict(dtype=torch.float64)
        eps = torch.full((n, n, n), background_eps, **t64)
        b = self.box
        eps = eps.clone()
        eps[b[0]:b[3], b[1]:b[4], b[2]:b[5]] = eps_design * EPS0
        # boundary map: domain faces, or any of the 26 neighbours with another permittivity
        ev = eps.detach()
        pad = torch.nn.functional.pad(ev[None, None], (1, 1, 1, 1, 1, 1), mode="replicate")[0, 0]
        differs = torch.zeros((n, n, n), dtype=torch.bool)
        for a in range(3):
            for bb in range(3):
                for c in range(3):
                    differs |= pad[a:a + n, bb:bb + n, c:c + n] != ev
        idx0 = torch.arange(n)
        face = (idx0 == 0) | (idx0 == n - 1)
        bnds = differs | face.view(-1, 1, 1) | face.view(1, -1, 1) | face.view(1, 1, -1)
        # per-component averages over self and the -1 neighbours that exist
        def sh(f, di, dj, dk):
            out = torch.zeros_like(f)
            out[di:, dj:, dk:] = f[:n - di, :n - dj, :n - dk]
            return out
        def has(di, dj, dk):
            m = torch.zeros((n, n, n), dtype=torch.float64)
            m[di:, dj:, dk:] = 1.0
            return m
        ex_s = eps + sh(eps, 0, 1, 0) + sh(eps, 0, 0, 1) + sh(eps, 0, 1, 1)
        ex_n = 1 + has(0, 1, 0) + has(0, 0, 1) + has(0, 1, 1)
        ey_s = eps + sh(eps, 1, 0, 0) + sh(eps, 0, 0, 1) + sh(eps, 1, 0, 1)
        ey_n = 1 + has(1, 0, 0) + has(0, 0, 1) + has(1, 0, 1)
        ez_s = eps + sh(eps, 1, 0, 0) + sh(eps, 0, 1, 0) + sh(eps, 1, 1, 0)
        ez_n = 1 + has(1, 0, 0) + has(0, 1, 0) + has(1, 1, 0)
        cbx = dt / (torch.where(bnds, ex_s / ex_n, eps) * d)
        cby = dt / (torch.where(bnds, ey_s / ey_n, eps) * d)
        cbz = dt / (torch.where(bnds, ez_s / ez_n, eps) * d)
        db = dt / MU0 / d
        ax = lambda v: torch.tensor(v, **t64)
        be = [ax(self.be).view(-1, 1, 1), ax(self.be).view(1, -1, 1), ax(self.be).view(1, 1, -1)]
        ced = [ax(self.ced).view(-1, 1, 1), ax(self.ced).view(1, -1, 1), ax(self.ced).view(1, 1, -1)]
        bh = [ax(self.bh).view(-1, 1, 1), ax(self.bh).view(1, -1, 1), ax(self.bh).view(1, 1, -1)]
        chd = [ax(self.chd).view(-1, 1, 1), ax(self.chd).view(1, -1, 1), ax(self.chd).view(1, 1, -1)]
        me = [torch.tensor(self.in_e).view(-1, 1, 1), torch.tensor(self.in_e).view(1, -1, 1),
              torch.tensor(self.in_e).view(1, 1, -1)]
        mh = [torch.tensor(self.in_h).view(-1, 1, 1), torch.tensor(self.in_h).view(1, -1, 1),
              torch.tensor(self.in_h).view(1, 1, -1)]
        idx = torch.arange(n)
        inner = (idx >= 1) & (idx < n - 1)
        I, J, K = inner.view(-1, 1, 1), inner.view(1, -1, 1), inner.view(1, 1, -1)
        nz0 = idx != 0
        NI, NJ, NK = nz0.view(-1, 1, 1), nz0.view(1, -1, 1), nz0.view(1, 1, -1)
        # H-side psi exist where the other two coordinates are inner; E-side everywhere in the slab
        mh_x = mh[0] & J & K
        mh_y = mh[1] & I & K
        mh_z = mh[2] & I & J
        z = lambda: torch.zeros((n, n, n), **t64)
        ex, ey, ez, hx, hy, hz = z(), z(), z(), z(), z(), z()
        pex1, pex2, pey1, pey2, pez1, pez2 = z(), z(), z(), z(), z(), z()
        phx1, phx2, phy1, phy2, phz1, phz2 = z(), z(), z(), z(), z(), z()
        w = 2 * math.pi * C0 / self.lam
        m0 = self.mon
        Fr = torch.zeros((3,) + tuple(m0[3 + q] - m0[q] for q in range(3)), **t64)
        Fi = torch.zeros_like(Fr)
        # forward differences with a zero beyond the last cell, backward with a zero before the first
        fwd = lambda f, a: torch.cat([f.narrow(a, 1, n - 1) - f.narrow(a, 0, n - 1), -f.narrow(a, n - 1, 1)], a)
        bwd = lambda f, a: torch.cat([f.narrow(a, 0, 1), f.narrow(a, 1, n - 1) - f.narrow(a, 0, n - 1)], a)
        last = idx == n - 1
        LI, LJ, LK = last.view(-1, 1, 1), last.view(1, -1, 1), last.view(1, 1, -1)
        ux_h, uy_h, uz_h = ~(LJ | LK), ~(LI | LK), ~(LI | LJ)
        first = idx == 0
        FI, FJ, FK = first.view(-1, 1, 1), first.view(1, -1, 1), first.view(1, 1, -1)
        ux_e, uy_e, uz_e = ~(FJ | FK), ~(FI | FK), ~(FI | FJ)
        zero_h = ((K & ((LI & J) | (LJ & I))) | (LK & I & J))
        zero_e = FI | FJ | FK
        for step in range(len(self.src)):
            # H update (da = 1 in vacuum), src3d/yee.c:145-265
            dey_z, dez_y = fwd(ey, 2), fwd(ez, 1)
            dez_x, dex_z = fwd(ez, 0), fwd(ex, 2)
            dex_y, dey_x = fwd(ex, 1), fwd(ey, 0)
            hx = hx + torch.where(ux_h, db * (dey_z - dez_y), 0.0)
            hy = hy + torch.where(uy_h, db * (dez_x - dex_z), 0.0)
            hz = hz + torch.where(uz_h, db * (dex_y - dey_x), 0.0)
            # CPML H, src3d/boundary.c:347-518 (x and y faces inside 1 <= k < n-1, z faces inside i, j)
            phx1 = torch.where(mh_x, bh[0] * phx1 + chd[0] * dey_x, phx1)
            phx2 = torch.where(mh_x, bh[0] * phx2 + chd[0] * dez_x, phx2)
            hy = hy + torch.where(mh_x, d * db * phx2, 0.0)
            hz = hz - torch.where(mh_x, d * db * phx1, 0.0)
            phy1 = torch.where(mh_y, bh[1] * phy1 + chd[1] * dez_y, phy1)
            phy2 = torch.where(mh_y, bh[1] * phy2 + chd[1] * dex_y, phy2)
            hx = hx - torch.where(mh_y, d * db * phy1, 0.0)
            hz = hz + torch.where(mh_y, d * db * phy2, 0.0)
            phz1 = torch.where(mh_z, bh[2] * phz1 + chd[2] * dex_z, phz1)
            phz2 = torch.where(mh_z, bh[2] * phz2 + chd[2] * dey_z, phz2)
            hy = hy - torch.where(mh_z, d * db * phz1, 0.0)
            hx = hx + torch.where(mh_z, d * db * phz2, 0.0)
            hx, hy, hz = (torch.where(zero_h, 0.0, f) for f in (hx, hy, hz))
            # E update (ca = 1, sigma = 0), src3d/yee.c:528-1247
            dhz_y, dhy_z = bwd(hz, 1), bwd(hy, 2)
            dhx_z, dhz_x = bwd(hx, 2), bwd(hz, 0)
            dhy_x, dhx_y = bwd(hy, 0), bwd(hx, 1)
            ex = ex + torch.where(ux_e, cbx * (dhz_y - dhy_z), 0.0)
            ey = ey + torch.where(uy_e, cby * (dhx_z - dhz_x), 0.0)
            ez = ez + torch.where(uz_e, cbz * (dhy_x - dhx_y), 0.0)
            # CPML E, src3d/boundary.c:691-887
            pex1 = torch.where(me[0], be[0] * pex1 + ced[0] * dhz_x, pex1)
            pex2 = torch.where(me[0], be[0] * pex2 + ced[0] * dhy_x, pex2)
            ey = ey - torch.where(me[0] & NJ & NK, d * cby * pex1, 0.0)
            ez = ez + torch.where(me[0] & NJ & NK, d * cbz * pex2, 0.0)
            pey1 = torch.where(me[1], be[1] * pey1 + ced[1] * dhz_y, pey1)
            pey2 = torch.where(me[1], be[1] * pey2 + ced[1] * dhx_y, pey2)
            ex = ex + torch.where(me[1] & NI & NK, d * cbx * pey1, 0.0)
            ez = ez - torch.where(me[1] & NI & NK, d * cbz * pey2, 0.0)
            pez1 = torch.where(me[2], be[2] * pez1 + ced[2] * dhx_z, pez1)
            pez2 = torch.where(me[2], be[2] * pez2 + ced[2] * dhy_z, pez2)
            ey = ey + torch.where(me[2] & NI & NJ, d * cby * pez1, 0.0)
            ex = ex - torch.where(me[2] & NI & NJ, d * cbx * pez2, 0.0)
            ex, ey, ez = (torch.where(zero_e, 0.0, f) for f in (ex, ey, ez))
            # hard source, src3d/source.c:405-437
            v = self.src[step]
            if v != 0:
                s = self.src_pos
                ez = ez.clone()
                ez[s[0], s[1], s[2]] = v
            # objective DFT
            t = (step + 1) * dt
            co, si = math.cos(w * t), math.sin(w * t)
            sl = (slice(m0[0], m0[3]), slice(m0[1], m0[4]), slice(m0[2], m0[5]))
            Em = torch.stack([ex[sl], ey[sl], ez[sl]])
            Fr = Fr + Em * co
            Fi = Fi - Em * si
        return torch.sum(Fr ** 2 + Fi ** 2)
