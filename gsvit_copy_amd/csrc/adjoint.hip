/*
 * adjoint.hip -- gradient of a running-DFT objective with respect to the
 * relative permittivity of the design voxels, by the discrete adjoint of the
 * engine's own FP32 time step (no reference counterpart: GSvit has no
 * gradient mode).
 *
 * Objective (OBJECTIVE_DFT box, wavelength lambda, omega = 2 pi c / lambda):
 *   J = sum_{cells in box} sum_{c = x,y,z} |F_c|^2,
 *   F_c = sum_{n=0}^{N-1} E_c^{n+1} exp(-i omega t_n),  t_n = (n+1) dt,
 * with E^{n+1} the field after step n (after its point sources).
 *
 * Scope: CPML or plain boundaries, hard E point sources, and dispersive (ADE)
 * media outside the CPML slabs and away from the design cells.  One forward step is x^{n+1} = S_n M_E M_H x^n + s_n for the state
 * x = (E, H, psi_E, psi_H); the backward pass applies the transposed steps in
 * reverse order (M_H^T M_E^T S_n^T), in two bulk kernels per step (with the
 * CPML psi terms in four slab kernels around them, see k_slab_*):
 *   k_adj_e  (E cells)  adds the H-step transpose of step n+1 (curl^T of r),
 *            the objective's source, the point-source
 *            projection, accumulates the coefficient gradient of the design
 *            cells and applies the local part of the E-step transpose
 *            (q = cb m, new adjoint E = ca m);
 *   k_adj_h  (H cells)  adds the E-step transpose (curl^T of q) and applies
 *            the local part of the H-step transpose (r = db h, new adjoint
 *            H = da h).
 * The adjoint of psi is stored without its decay factor b, which the next
 * (earlier) step applies when it reads it.
 *
 * Dispersive cells (kernels.hip, ade_cell, increment form): per component c
 * the step computes A = the ADE update (c0 curl / d + cj J + sum_q (pcj Jq +
 * pcp Pq) + c2 Ep + c3 E) for the components it owns, A = E for the others,
 * advances J' = dj J + da (Ep + 2 E + A), Jq' = pj Jq + ps Pq + b2 Ep + b3 E +
 * b4 A, Pq' = Pq + Jq', Ep' = E, and then gives the other components the
 * ordinary update E' = ca A + cb curl.  The transpose keeps an adjoint of
 * each state (Ep, J, Jq, Pq) in the forward state's layout: with u_q the
 * adjoint of Jq' (its own plus that of Pq') and a_A = (ca or 1) E'bar + da Jbar'
 * + sum b4 u_q, it feeds c0 / d a_A (or cb E'bar) to the curl transpose and
 * returns Ebar = (c3 a_A or a_A) + 2 da Jbar' + sum b3 u_q + Epbar',
 * Epbar = c2 a_A + da Jbar' + sum b2 u_q, Jbar = cj a_A + dj Jbar',
 * Jqbar = pcj a_A + pj u_q, Pqbar = pcp a_A + ps u_q + Pqbar' (the a_A terms
 * only for components the ADE update owns).
 *
 * The gradient needs, per design cell and step, E^n and the coefficient-free
 * increment D^n = (E^{n+1} - ca E^n) / cb of the E update.  They are recorded
 * for the whole run when they fit the history budget; otherwise the forward
 * pass keeps full checkpoints every K steps and each K-step segment is run
 * forward again (recording) just before its backward pass.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <omp.h>
#include "engine_host.h"
#include "device_access.h"

/* ------------------------------------------------------------------ */
/* device parameter block                                               */
/* ------------------------------------------------------------------ */

typedef struct {
    /* objective monitor box */
    int mi0, mj0, mk0, mnx, mny, mnz;
    double omega, dt;
    double *macc;             /* [cell][3][2] running DFT (re, im) of E */
    /* adjoint state, same layout as the fields; allocation starts (guard plane k = -1) */
    float *ebx, *eby, *ebz, *hbx, *hby, *hbz;
    float *qx, *qy, *qz, *rx, *ry, *rz;
    float2 *te[6], *th[6];    /* psi adjoints per face (E side, H side) */
    float *xad, *xap;         /* adjoints of the ADE state, the layout of GsvGrid::ade_d / ade_p */
    /* design coefficient box: history and gradient accumulators */
    long long ncoef;
    float *hist;              /* [slot][6][ncoef]: E^n x y z, D^n x y z */
    double *gca, *gcb;        /* [3][ncoef] */
    /* hard E point sources */
    int nsrc;
    const long long *spos;
    const float *svals;       /* [nsteps][nsrc][6] */
} GsvAdj;

/* ------------------------------------------------------------------ */
/* forward-side kernels                                                 */
/* ------------------------------------------------------------------ */

/* objective DFT: acc += E(cell) exp(-i omega (step+1) dt) over the monitor box */
__global__ __launch_bounds__(256) void
k_obj_dft(GsvGrid g, GsvAdj a, int step)
{
    const long long n = (long long)a.mnx * a.mny * a.mnz;
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    const int i = a.mi0 + (int)(q % a.mnx);
    const int j = a.mj0 + (int)((q / a.mnx) % a.mny);
    const int k = a.mk0 + (int)(q / ((long long)a.mnx * a.mny));
    const long long c = i + (long long)g.px * j + g.plane * k;
    double s, co;
    sincos(a.omega * (step + 1) * a.dt, &s, &co);
    double *acc = a.macc + q * 6;
    const float e[3] = { g.ex[c], g.ey[c], g.ez[c] };
    for (int comp = 0; comp < 3; comp++) {
        acc[2 * comp] += e[comp] * co;
        acc[2 * comp + 1] -= e[comp] * s;
    }
}

/* history, before the E update: E^n of the design coefficient cells */
__global__ __launch_bounds__(256) void
k_rec_pre(GsvGrid g, GsvAdj a, int slot)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.ncoef)
        return;
    const int i = g.pc_i0 + (int)(q % g.pc_nx);
    const int j = g.pc_j0 + (int)((q / g.pc_nx) % g.pc_ny);
    const int k = g.pc_k0 + (int)(q / ((long long)g.pc_nx * g.pc_ny));
    const long long c = i + (long long)g.px * j + g.plane * k;
    float *h = a.hist + (long long)slot * 6 * a.ncoef + q;
    h[0] = g.ex[c];
    h[a.ncoef] = g.ey[c];
    h[2 * a.ncoef] = g.ez[c];
}

/* history, after the E update (before the point sources): D^n = (E^{n+1} - ca E^n) / cb */
__global__ __launch_bounds__(256) void
k_rec_post(GsvGrid g, GsvAdj a, int slot)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.ncoef)
        return;
    const int i = g.pc_i0 + (int)(q % g.pc_nx);
    const int j = g.pc_j0 + (int)((q / g.pc_nx) % g.pc_ny);
    const int k = g.pc_k0 + (int)(q / ((long long)g.pc_nx * g.pc_ny));
    const long long c = i + (long long)g.px * j + g.plane * k;
    const GsvECoef ec = g.pctab[q];
    float *h = a.hist + (long long)slot * 6 * a.ncoef + q;
    const float e1[3] = { g.ex[c], g.ey[c], g.ez[c] };
    for (int comp = 0; comp < 3; comp++) {
        const float e0 = h[comp * a.ncoef];
        h[(3 + comp) * a.ncoef] = ec.cb[comp] != 0.f ? (e1[comp] - ec.ca[comp] * e0) / ec.cb[comp] : 0.f;
    }
}

/* ------------------------------------------------------------------ */
/* backward kernels                                                     */
/* ------------------------------------------------------------------ */

/* Both kernels march a chunk of z planes per (i, j) column like the forward
 * volume kernels, address every per-cell array by one 32-bit byte offset from
 * its guard plane and carry the neighbour plane they share in registers. */

/* transpose of ade_cell's step for one dispersive cell (see the file head):
 * e[c] holds the adjoint of E' on entry and the adjoint of E on return, q[c]
 * the curl coefficient's adjoint; xd / xp the cell's adjoint state */
template <int NP, typename OFF>
__device__ __forceinline__ void
adj_ade_cell(const GsvGrid &g, const GsvECoef &ec, float *xd, OFF la, float *xp, OFF lp, int i, int j, int k,
             unsigned fl, float e[3], float q[3])
{
    const GsvAde &m = g.adem[(fl >> GSV_E_SLOT_SHIFT) & 0xF];
    const bool t[3] = { (fl & GSV_E_TBCX) != 0, (fl & GSV_E_TBCY) != 0, (fl & GSV_E_TBCZ) != 0 };
    /* the update applies (either kind) where the forward step's bounds allow it */
    const bool upd[3] = { j > 0 && k > 0, i > 0 && k > 0, i > 0 && j > 0 };
    const float dinv[3] = { 1.f / g.dx, 1.f / g.dy, 1.f / g.dz };
    const bool dr = NP == 0 || m.drude;
    const int npl = NP ? m.poles : 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float aep = ldsn(xd, la, 256 * c);
        const float aj = dr ? ldsn(xd, la, 768 + 256 * c) : 0.f;
        float u[3] = {}, apq[3] = {};
#pragma unroll
        for (int p = 0; p < NP; p++)
            if (p < npl) {
                apq[p] = ldsn(xp, lp, (3 * p + c) * 256);
                u[p] = ldsn(xp, lp, (3 * NP + 3 * p + c) * 256) + apq[p];
            }
        /* the ordinary update after the ADE branch */
        const bool ord = t[c] && upd[c], own = !t[c] && upd[c];
        float abar = ord ? ec.ca[c] * e[c] : e[c];
        q[c] = ord ? ec.cb[c] * e[c] : 0.f;
        /* A feeds the state updates of this step */
        abar += m.da * aj;
#pragma unroll
        for (int p = 0; p < NP; p++)
            if (p < npl)
                abar += m.b4[p] * u[p];
        float ebar = own ? m.c3 * abar : abar;
        float epbar = own ? m.c2 * abar : 0.f;
        float jbar = own ? m.cj * abar : 0.f;
        if (own)
            q[c] = m.c0 * dinv[c] * abar;
        ebar += 2.f * m.da * aj + aep;
        epbar += m.da * aj;
        jbar += m.dj * aj;
#pragma unroll
        for (int p = 0; p < NP; p++)
            if (p < npl) {
                ebar += m.b3[p] * u[p];
                epbar += m.b2[p] * u[p];
                stsn(xp, lp, (3 * NP + 3 * p + c) * 256, (own ? m.pcj[p] * abar : 0.f) + m.pj[p] * u[p]);
                stsn(xp, lp, (3 * p + c) * 256, (own ? m.pcp[p] * abar : 0.f) + m.ps[p] * u[p] + apq[p]);
            }
        stsn(xd, la, 256 * c, epbar);
        if (dr)
            stsn(xd, la, 768 + 256 * c, jbar);
        e[c] = ebar;
    }
}

/* E cells: H-step transpose of step n+1 (gathers of r and the psi_H
 * adjoints), objective source and point-source projection of step n,
 * coefficient gradient, local part of the E-step transpose of step n */
template <typename ID, typename OFF, int ADE>
__global__ __launch_bounds__(512) void
k_adj_e(GsvGrid g, GsvAdj a, int step, int slot, int kchunk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= nx || j >= ny)
        return;
    const int k0 = blockIdx.z * kchunk;
    const int k1 = min(k0 + kchunk, nz);
    const int px = g.px;
    const OFF dP = (OFF)g.plane * 4, dJ = (OFF)px * 4, dPc = (OFF)g.plane * sizeof(ID);
    OFF o = ((OFF)i + (OFF)px * j + (OFF)g.plane * (k0 + 1)) * 4;
    OFF oc = ((OFF)i + (OFF)px * j + (OFF)g.plane * k0) * sizeof(ID);
    /* lane offsets in the blocked ADE state planes (as estep_col) */
    constexpr int NP = ADE >= 2 ? ADE : 0;
    const OFF la = ADE ? ((OFF)j * g.pxa * GSV_ADE_D + (OFF)(i >> 6) * 64 * GSV_ADE_D + (i & 63)) * 4 : 0;
    const OFF lp = NP ? ((OFF)j * g.pxa * GSV_ADE_P(NP) + (OFF)(i >> 6) * 64 * GSV_ADE_P(NP) + (i & 63)) * 4 : 0;

    const float rhx = g.rkhx[i], rhy = g.rkhy[j];
    const float rhx1 = i > 0 ? g.rkhx[i - 1] : 0.f, rhy1 = j > 0 ? g.rkhy[j - 1] : 0.f;
    const bool zplane_xy = (g.f[0].on && i == 0) || (g.f[2].on && j == 0), zplane_z = g.f[4].on;
    const bool in_mon_ij = i >= a.mi0 && i < a.mi0 + a.mnx && j >= a.mj0 && j < a.mj0 + a.mny;
    const bool in_pc_ij = a.ncoef && i >= g.pc_i0 && i < g.pc_i0 + g.pc_nx && j >= g.pc_j0 && j < g.pc_j0 + g.pc_ny;

    /* r of plane k-1 (the guard plane below k = 0 is zero) */
    float rxm = ldf(a.rx, o - dP), rym = ldf(a.ry, o - dP);
    /* class id one plane ahead: the coefficient gather starts with the plane's loads */
    unsigned idn = ldid<ID>(g.ecls, oc);
    for (int k = k0; k < k1; ++k, o += dP, oc += dPc) {
        float ex = ldn(a.ebx, o), ey = ldn(a.eby, o), ez = ldn(a.ebz, o);
        const float rxc = ldf(a.rx, o), ryc = ldf(a.ry, o), rzc = ldf(a.rz, o);
        GsvECoef ec = ldec(g.etab + idn);
        if (k + 1 < k1)
            idn = ldid<ID>(g.ecls, oc + dPc);
        const float rhz = g.rkhz[k], rhz1 = k > 0 ? g.rkhz[k - 1] : 0.f;
        /* transpose of the H curl of step n+1 (H-grid 1/kappa) */
        ex += rhz * ryc - rhy * rzc - rhz1 * rym;
        ey += rhx * rzc - rhz * rxc + rhz1 * rxm;
        ez += rhy * rxc - rhx * ryc;
        if (j > 0) {
            ex += rhy1 * ldf(a.rz, o - dJ);
            ez -= rhy1 * ldf(a.rx, o - dJ);
        }
        if (i > 0) {
            ey -= rhx1 * ldf(a.rz, o - 4);
            ez += rhx1 * ldf(a.ry, o - 4);
        }
        rxm = rxc;
        rym = ryc;
        /* objective source: dJ/dE^{n+1} = 2 (Re F cos - Im F sin) */
        if (in_mon_ij && k >= a.mk0 && k < a.mk0 + a.mnz) {
            const long long q = (i - a.mi0) + (long long)a.mnx * ((j - a.mj0) + (long long)a.mny * (k - a.mk0));
            const double *F = a.macc + q * 6;
            double s, co;
            sincos(a.omega * (step + 1) * a.dt, &s, &co);
            ex += (float)(2.0 * (F[0] * co - F[1] * s));
            ey += (float)(2.0 * (F[2] * co - F[3] * s));
            ez += (float)(2.0 * (F[4] * co - F[5] * s));
        }
        /* hard point sources overwrite E: their cells carry no adjoint back */
        if (a.nsrc) {
            const long long c = i + (long long)px * j + g.plane * k;
            for (int s = 0; s < a.nsrc; s++)
                if (a.spos[s] == c) {
                    const float *v = a.svals + ((long long)step * a.nsrc + s) * 6;
                    if (v[0] != 0.f) ex = 0.f;
                    if (v[1] != 0.f) ey = 0.f;
                    if (v[2] != 0.f) ez = 0.f;
                }
        }

        /* E-step transpose: plane zeroing, then the update and the psi_E corrections */
        if (zplane_xy || (zplane_z && k == 0)) {
            ex = 0.f;
            ey = 0.f;
            ez = 0.f;
        }
        if (ec.flags & GSV_E_PERCELL)
            ec = g.pctab[(i - g.pc_i0) + g.pc_nx * ((j - g.pc_j0) + g.pc_ny * (k - g.pc_k0))];
        const unsigned fl = ec.flags;
        if (in_pc_ij && k >= g.pc_k0 && k < g.pc_k0 + g.pc_nz) {
            /* design coefficient cell: dJ/dca += adj(E^{n+1}) E^n, dJ/dcb += adj(E^{n+1}) D^n */
            const long long q = (i - g.pc_i0) + (long long)g.pc_nx * ((j - g.pc_j0) + (long long)g.pc_ny * (k - g.pc_k0));
            const float *h = a.hist + (long long)slot * 6 * a.ncoef + q;
            a.gca[q] += (double)ex * h[0];
            a.gca[a.ncoef + q] += (double)ey * h[a.ncoef];
            a.gca[2 * a.ncoef + q] += (double)ez * h[2 * a.ncoef];
            a.gcb[q] += (double)ex * h[3 * a.ncoef];
            a.gcb[a.ncoef + q] += (double)ey * h[4 * a.ncoef];
            a.gcb[2 * a.ncoef + q] += (double)ez * h[5 * a.ncoef];
        }
        if (ADE && (fl & GSV_E_ADE)) {
            /* dispersive cell (no PEC, no CPML slab, no design cell here) */
            float e[3] = { ex, ey, ez }, q[3];
            adj_ade_cell<NP>(g, ec, a.xad + (size_t)k * g.ade_plane_d, la, NP ? a.xap + (size_t)k * g.ade_plane_p : nullptr,
                             lp, i, j, k, fl, e, q);
            stf(a.qx, o, q[0]);
            stf(a.qy, o, q[1]);
            stf(a.qz, o, q[2]);
            stn(a.ebx, o, e[0]);
            stn(a.eby, o, e[1]);
            stn(a.ebz, o, e[2]);
            continue;
        }
        const bool tx = fl & GSV_E_TBCX, ty = fl & GSV_E_TBCY, tz = fl & GSV_E_TBCZ;
        /* PEC zeroing acts on the curl update only (the corrections come after it) */
        const float mx = (fl & GSV_E_PECX) ? 0.f : ex, my = (fl & GSV_E_PECY) ? 0.f : ey;
        const float mz = (fl & GSV_E_PECZ) ? 0.f : ez;
        const bool ux = tx && j > 0 && k > 0, uy = ty && i > 0 && k > 0, uz = tz && i > 0 && j > 0;
        stf(a.qx, o, ux ? ec.cb[0] * mx : 0.f);
        stf(a.qy, o, uy ? ec.cb[1] * my : 0.f);
        stf(a.qz, o, uz ? ec.cb[2] * mz : 0.f);
        stn(a.ebx, o, ux ? ec.ca[0] * mx : mx);
        stn(a.eby, o, uy ? ec.ca[1] * my : my);
        stn(a.ebz, o, uz ? ec.ca[2] * mz : mz);
    }
}

/* H cells: E-step transpose (gathers of q and the psi_E adjoints), then the
 * local part of the H-step transpose */
template <typename ID, typename OFF>
__global__ __launch_bounds__(512) void
k_adj_h(GsvGrid g, GsvAdj a, int kchunk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    if (i >= nx || j >= ny)
        return;
    const int k0 = blockIdx.z * kchunk;
    const int k1 = min(k0 + kchunk, nz);
    const int px = g.px;
    const OFF dP = (OFF)g.plane * 4, dJ = (OFF)px * 4, dPc = (OFF)g.plane * sizeof(ID);
    OFF o = ((OFF)i + (OFF)px * j + (OFF)g.plane * (k0 + 1)) * 4;
    OFF oc = ((OFF)i + (OFF)px * j + (OFF)g.plane * k0) * sizeof(ID);

    const float rex = g.rkex[i], rey = g.rkey[j];
    const float rex1 = i + 1 < nx ? g.rkex[i + 1] : 0.f, rey1 = j + 1 < ny ? g.rkey[j + 1] : 0.f;
    const bool ip = i + 1 < nx, jp = j + 1 < ny;
    const bool iin = (i >= 1 && i < nx - 1), jin = (j >= 1 && j < ny - 1), ijin = iin && jin;
    const bool zx = g.f[1].on && i == nx - 1 && jin, zy = g.f[3].on && j == ny - 1 && iin, zzn = g.f[5].on && ijin;

    /* q of plane k (carried), the plane above is loaded each step (guard plane above nz-1 is zero) */
    float qxk = ldf(a.qx, o), qyk = ldf(a.qy, o);
    /* class id one plane ahead: the coefficient gather starts with the plane's loads */
    unsigned idn = ldid<ID>(g.hcls, oc);
    for (int k = k0; k < k1; ++k, o += dP, oc += dPc) {
        float hx = ldn(a.hbx, o), hy = ldn(a.hby, o), hz = ldn(a.hbz, o);
        const GsvHCoef hc = ldhc(g.htab + idn);
        if (k + 1 < k1)
            idn = ldid<ID>(g.hcls, oc + dPc);
        const float qzc = ldf(a.qz, o);
        const float qxk1 = ldf(a.qx, o + dP), qyk1 = ldf(a.qy, o + dP);
        const float rez = g.rkez[k], rez1 = k + 1 < nz ? g.rkez[k + 1] : 0.f;
        /* transpose of the E curl (E-grid 1/kappa) */
        hx += rez * qyk - rey * qzc - rez1 * qyk1;
        hy += rex * qzc - rez * qxk + rez1 * qxk1;
        hz += rey * qxk - rex * qyk;
        if (jp) {
            hx += rey1 * ldf(a.qz, o + dJ);
            hz -= rey1 * ldf(a.qx, o + dJ);
        }
        if (ip) {
            hy -= rex1 * ldf(a.qz, o + 4);
            hz += rex1 * ldf(a.qy, o + 4);
        }
        qxk = qxk1;
        qyk = qyk1;
        /* H-step transpose: zeroing, then the update and the psi_H corrections */
        const bool kin = (k >= 1 && k < nz - 1);
        if ((kin && (zx || zy)) || (zzn && k == nz - 1)) {
            hx = 0.f;
            hy = 0.f;
            hz = 0.f;
        }
        const bool ux = !(j == ny - 1 || k == nz - 1), uy = !(i == nx - 1 || k == nz - 1);
        const bool uz = !(i == nx - 1 || j == ny - 1);
        stf(a.rx, o, ux ? hc.db[0] * hx : 0.f);
        stf(a.ry, o, uy ? hc.db[1] * hy : 0.f);
        stf(a.rz, o, uz ? hc.db[2] * hz : 0.f);
        stn(a.hbx, o, ux ? hc.da * hx : hx);
        stn(a.hby, o, uy ? hc.da * hy : hy);
        stn(a.hbz, o, uz ? hc.da * hz : hz);
    }
}

/* ------------------------------------------------------------------ */
/* CPML slabs of the backward step                                      */
/* ------------------------------------------------------------------ */

/* The psi terms of the transposed step run in four slab kernels around the
 * two bulk kernels above, so that the bulk kernels carry no CPML state:
 *   k_slab_e_pre   adds the psi_H transpose of step n+1 to the adjoint E;
 *   k_slab_e_post  psi_E adjoint of step n from the new adjoint E (which holds
 *                  ca m, or m where the update does not apply: m = E'' without PEC);
 *   k_slab_h_pre   adds the psi_E transpose to the adjoint H;
 *   k_slab_h_post  psi_H adjoint from the new adjoint H (da h'' or h'').
 * Per face axis: psi pair 1 / 2 correct E components ec and are fed by H
 * components es (E side); they correct H components hc and are fed by E
 * components hs (H side); 0, 1, 2 = x, y, z.  The pre kernels run once per
 * axis (near and far faces of one axis never share a cell). */

__device__ __forceinline__ float *
fcomp(float *x, float *y, float *z, int c)
{
    return c == 0 ? x : (c == 1 ? y : z);
}

/* cell of thread q in the windows [0, dn) and [n - df, n) along AXIS */
template <int AXIS>
__device__ __forceinline__ bool
window_cell(const GsvGrid &g, long long q, int dn, int df, int &i, int &j, int &k, int &pos)
{
    const int w = dn + df;
    int t;
    if (AXIS == 0) {
        t = (int)(q % w);
        const long long rest = q / w;
        j = (int)(rest % g.ny);
        k = (int)(rest / g.ny);
        if (k >= g.nz) return false;
    } else if (AXIS == 1) {
        i = (int)(q % g.nx);
        const long long rest = q / g.nx;
        t = (int)(rest % w);
        k = (int)(rest / w);
        if (k >= g.nz) return false;
    } else {
        i = (int)(q % g.nx);
        const long long rest = q / g.nx;
        j = (int)(rest % g.ny);
        t = (int)(rest / g.ny);
        if (t >= w) return false;
    }
    const int n = AXIS == 0 ? g.nx : (AXIS == 1 ? g.ny : g.nz);
    pos = t < dn ? t : n - df + (t - dn);
    if (AXIS == 0) i = pos; else if (AXIS == 1) j = pos; else k = pos;
    return true;
}

/* psi pair index of slab cell r of a face along AXIS: x [k][j][r], y [k][r][i], z [r][j][i] */
template <int AXIS>
__device__ __forceinline__ long long
psi_idx(const GsvGrid &g, int depth, int r, int i, int j, int k)
{
    if (AXIS == 0) return ((long long)k * g.ny + j) * depth + r;
    if (AXIS == 1) return ((long long)k * depth + r) * g.px + i;
    return ((long long)r * g.ny + j) * g.px + i;
}

template <int AXIS>
__global__ __launch_bounds__(256) void
k_slab_e_pre(GsvGrid g, GsvAdj a)
{
    const GsvCpmlFace &Fn = g.f[2 * AXIS], &Ff = g.f[2 * AXIS + 1];
    const int dn = Fn.on ? Fn.depth : 0, df = Ff.on ? Ff.depth : 0;
    int i = 0, j = 0, k = 0, pos;
    if (!window_cell<AXIS>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, dn, df, i, j, k, pos))
        return;
    const bool iin = i >= 1 && i < g.nx - 1, jin = j >= 1 && j < g.ny - 1, kin = k >= 1 && k < g.nz - 1;
    if (!(AXIS == 0 ? (jin && kin) : (AXIS == 1 ? (iin && kin) : (iin && jin))))
        return;
    const int n = AXIS == 0 ? g.nx : (AXIS == 1 ? g.ny : g.nz);
    /* E components feeding psi_H 1 / 2 */
    const int c1 = AXIS == 0 ? 1 : (AXIS == 1 ? 2 : 0), c2 = AXIS == 0 ? 2 : (AXIS == 1 ? 0 : 1);
    float v1 = 0.f, v2 = 0.f;
    for (int d = 0; d < 2; d++) {
        const int pp = pos - d;   /* psi_H cell at pos (sign -) and pos - 1 (sign +) */
        const float sg = d ? 1.f : -1.f;
        if (Fn.on && pp >= 0 && pp < Fn.depth - 1) {
            const float cc = sg * Fn.ch[pp].y;
            const float2 t = a.th[2 * AXIS][psi_idx<AXIS>(g, Fn.depth, pp, i, j, k)];
            v1 += cc * t.x;
            v2 += cc * t.y;
        }
        if (Ff.on && pp >= n - Ff.depth + 1 && pp < n - 1) {
            const int r = pp - (n - Ff.depth);
            const float cc = sg * Ff.ch[r].y;
            const float2 t = a.th[2 * AXIS + 1][psi_idx<AXIS>(g, Ff.depth, r, i, j, k)];
            v1 += cc * t.x;
            v2 += cc * t.y;
        }
    }
    const long long c = i + (long long)g.px * j + g.plane * k;
    float *e1 = fcomp(a.ebx, a.eby, a.ebz, c1) + g.plane, *e2 = fcomp(a.ebx, a.eby, a.ebz, c2) + g.plane;
    e1[c] += v1;
    e2[c] += v2;
}

template <int AXIS>
__global__ __launch_bounds__(256) void
k_slab_h_pre(GsvGrid g, GsvAdj a)
{
    const GsvCpmlFace &Fn = g.f[2 * AXIS], &Ff = g.f[2 * AXIS + 1];
    const int dn = Fn.on ? Fn.depth : 0, df = Ff.on ? Ff.depth : 0;
    int i = 0, j = 0, k = 0, pos;
    if (!window_cell<AXIS>(g, (long long)blockIdx.x * blockDim.x + threadIdx.x, dn, df, i, j, k, pos))
        return;
    const int n = AXIS == 0 ? g.nx : (AXIS == 1 ? g.ny : g.nz);
    /* H components feeding psi_E 1 / 2 */
    const int c1 = AXIS == 0 ? 2 : (AXIS == 1 ? 2 : 0), c2 = AXIS == 0 ? 1 : (AXIS == 1 ? 0 : 1);
    float v1 = 0.f, v2 = 0.f;
    for (int d = 0; d < 2; d++) {
        const int pp = pos + d;   /* psi_E cell at pos (sign +) and pos + 1 (sign -) */
        const float sg = d ? -1.f : 1.f;
        if (Fn.on && pp >= 1 && pp < Fn.depth) {
            const float cc = sg * Fn.ce[pp].y;
            const float2 t = a.te[2 * AXIS][psi_idx<AXIS>(g, Fn.depth, pp, i, j, k)];
            v1 += cc * t.x;
            v2 += cc * t.y;
        }
        if (Ff.on && pp >= n - Ff.depth + 1 && pp < n) {
            const int r = pp - (n - Ff.depth);
            const float cc = sg * Ff.ce[r].y;
            const float2 t = a.te[2 * AXIS + 1][psi_idx<AXIS>(g, Ff.depth, r, i, j, k)];
            v1 += cc * t.x;
            v2 += cc * t.y;
        }
    }
    const long long c = i + (long long)g.px * j + g.plane * k;
    float *h1 = fcomp(a.hbx, a.hby, a.hbz, c1) + g.plane, *h2 = fcomp(a.hbx, a.hby, a.hbz, c2) + g.plane;
    h1[c] += v1;
    h2[c] += v2;
}

/* slab cell (r, i, j, k) of face f from a flat psi index q; false for padding */
__device__ __forceinline__ bool
face_cell(const GsvGrid &g, int f, long long q, int &r, int &i, int &j, int &k)
{
    const int axis = f / 2, depth = g.f[f].depth;
    if (axis == 0) {
        r = (int)(q % depth);
        j = (int)((q / depth) % g.ny);
        k = (int)(q / ((long long)depth * g.ny));
        if (k >= g.nz) return false;
        i = (f & 1) ? g.nx - depth + r : r;
    } else if (axis == 1) {
        i = (int)(q % g.px);
        r = (int)((q / g.px) % depth);
        k = (int)(q / ((long long)g.px * depth));
        if (i >= g.nx || k >= g.nz) return false;
        j = (f & 1) ? g.ny - depth + r : r;
    } else {
        i = (int)(q % g.px);
        j = (int)((q / g.px) % g.ny);
        r = (int)(q / ((long long)g.px * g.ny));
        if (i >= g.nx || r >= depth) return false;
        k = (f & 1) ? g.nz - depth + r : r;
    }
    return true;
}

/* psi_E adjoint of the step: T = b T(later) + P^T E'' at every E-side slab cell */
template <typename ID>
__global__ __launch_bounds__(256) void
k_slab_e_post(GsvGrid g, GsvAdj a)
{
    const int f = blockIdx.y;
    const GsvCpmlFace &F = g.f[f];
    int r, i, j, k;
    if (!F.on || !face_cell(g, f, (long long)blockIdx.x * blockDim.x + threadIdx.x, r, i, j, k) || r < 1)
        return;
    const int axis = f / 2;
    const long long c = i + (long long)g.px * j + g.plane * k;
    GsvECoef ec = g.etab[((const ID *)g.ecls)[c]];
    if (ec.flags & GSV_E_PERCELL)
        ec = g.pctab[(i - g.pc_i0) + g.pc_nx * ((j - g.pc_j0) + g.pc_ny * (k - g.pc_k0))];
    const unsigned fl = ec.flags;
    /* E'' of the components this face corrects (m = E'' here: no PEC in gradient mode) */
    const int c1 = axis == 1 ? 0 : 1, c2 = axis == 0 ? 2 : (axis == 1 ? 2 : 0);
    const bool u[3] = { (fl & GSV_E_TBCX) && j > 0 && k > 0, (fl & GSV_E_TBCY) && i > 0 && k > 0,
                        (fl & GSV_E_TBCZ) && i > 0 && j > 0 };
    const float e1 = fcomp(a.ebx, a.eby, a.ebz, c1)[g.plane + c], e2 = fcomp(a.ebx, a.eby, a.ebz, c2)[g.plane + c];
    const float m1 = u[c1] ? e1 / ec.ca[c1] : e1, m2 = u[c2] ? e2 / ec.ca[c2] : e2;
    /* the forward corrections: x faces -dx cb_y, +dx cb_z; y faces +dy cb_x, -dy cb_z;
     * z faces +dx cb_y, -dx cb_x, gated on the flags of src3d/boundary.c:691-887 */
    const bool nz2 = axis == 0 ? (j != 0 && k != 0) : (axis == 1 ? (i != 0 && k != 0) : (i != 0 && j != 0));
    const unsigned g1 = axis == 1 ? GSV_E_TBCX : GSV_E_TBCY;
    const unsigned g2 = axis == 2 ? (f == 4 ? GSV_E_TBCZ : GSV_E_TBCX) : GSV_E_TBCZ;
    const float d = axis == 1 ? g.dy : g.dx;
    const float s1 = axis == 0 ? -d : d, s2 = axis == 0 ? d : -d;
    const long long p = axis == 0 ? psi_idx<0>(g, F.depth, r, i, j, k)
                                  : (axis == 1 ? psi_idx<1>(g, F.depth, r, i, j, k) : psi_idx<2>(g, F.depth, r, i, j, k));
    const float b = F.ce[r].x;
    float2 t = a.te[f][p];
    t.x *= b;
    t.y *= b;
    if (nz2) {
        if (fl & g1) t.x += s1 * ec.cb[c1] * m1;
        if (fl & g2) t.y += s2 * ec.cb[c2] * m2;
    }
    a.te[f][p] = t;
}

/* psi_H adjoint of the step: T = b T(later) + P_h^T H'' at every H-side slab cell */
template <typename ID>
__global__ __launch_bounds__(256) void
k_slab_h_post(GsvGrid g, GsvAdj a)
{
    const int f = blockIdx.y;
    const GsvCpmlFace &F = g.f[f];
    int r, i, j, k;
    if (!F.on || !face_cell(g, f, (long long)blockIdx.x * blockDim.x + threadIdx.x, r, i, j, k))
        return;
    const int axis = f / 2;
    const int n = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.nz);
    const int pos = axis == 0 ? i : (axis == 1 ? j : k);
    const bool inside = (f & 1) ? (pos >= n - F.depth + 1 && pos < n - 1) : (pos < F.depth - 1);
    const bool iin = i >= 1 && i < g.nx - 1, jin = j >= 1 && j < g.ny - 1, kin = k >= 1 && k < g.nz - 1;
    if (!inside || !(axis == 0 ? (jin && kin) : (axis == 1 ? (iin && kin) : (iin && jin))))
        return;
    const long long c = i + (long long)g.px * j + g.plane * k;
    const GsvHCoef hc = g.htab[((const ID *)g.hcls)[c]];
    /* H'' of the components this face corrects: x faces z, y; y faces x, z; z faces y, x */
    const int c1 = axis == 0 ? 2 : (axis == 1 ? 0 : 1), c2 = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
    const bool u[3] = { !(j == g.ny - 1 || k == g.nz - 1), !(i == g.nx - 1 || k == g.nz - 1),
                        !(i == g.nx - 1 || j == g.ny - 1) };
    const float h1 = fcomp(a.hbx, a.hby, a.hbz, c1)[g.plane + c], h2 = fcomp(a.hbx, a.hby, a.hbz, c2)[g.plane + c];
    const float m1 = u[c1] ? h1 / hc.da : h1, m2 = u[c2] ? h2 / hc.da : h2;
    const float d = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
    const long long p = axis == 0 ? psi_idx<0>(g, F.depth, r, i, j, k)
                                  : (axis == 1 ? psi_idx<1>(g, F.depth, r, i, j, k) : psi_idx<2>(g, F.depth, r, i, j, k));
    const float b = F.ch[r].x;
    float2 t = a.th[f][p];
    a.th[f][p] = make_float2(b * t.x - d * hc.dbc * m1, b * t.y + d * hc.dbc * m2);
}

/* ------------------------------------------------------------------ */
/* host driver                                                          */
/* ------------------------------------------------------------------ */

struct GsvAdjoint {
    GsvAdj a;
    int nsteps;
    int K;                      /* checkpoint interval (steps per segment) */
    int history_all;            /* the whole run's history fits: no recomputation */
    int hist_slots;
    int recording, record_base; /* recording on, step of history slot 0 */
    int accumulate;             /* objective DFT accumulation on */
    std::vector<float *> ckpt;  /* per segment: 6 fields then the psi pairs */
    size_t ckpt_floats;
    std::vector<float> src_vals_host;
    double ms_forward, ms_backward;
};

static size_t
psi_pairs(const GsvCpmlFace &F, int axis, const GsvGrid &g)
{
    if (!F.on)
        return 0;
    if (axis == 0) return (size_t)F.depth * g.ny * g.nz;
    if (axis == 1) return (size_t)g.px * F.depth * g.nz;
    return (size_t)g.px * g.ny * F.depth;
}

/* floats of the forward ADE state (Drude blocks, then the pole blocks) */
static size_t
ade_floats(const GsvGrid &g)
{
    if (!g.ade)
        return 0;
    return (size_t)g.ade_plane_d * g.nz + (g.ade_p ? (size_t)g.ade_plane_p * g.nz : 0);
}

void
gsv_adjoint_free(Engine *E)
{
    if (E->adj) {
        delete E->adj;
        E->adj = NULL;
    }
}

/* device buffers of the gradient state (allocated once per engine) */
extern "C" int
gsv_engine_adjoint_setup(void *eng, long long history_budget_bytes, int ckpt_interval)
{
    Engine *E = (Engine *)eng;
    SvSet *set = E->set;
    GsvGrid &g = E->g;
    if (!E->ready)
        return 1;
    if (E->adj)
        return 0;
    if (!set->sa.design || !set->sa.objective) {
        fprintf(stderr, "Error: the gradient needs DESIGN_BOX and OBJECTIVE_DFT in the parameter file\n");
        return 1;
    }
    if (!E->forces.empty() || E->has_pff) {
        fprintf(stderr, "Error: the gradient does not support force boxes or the periodic far field\n");
        return 1;
    }
    if (E->has_tsf || E->has_ltsf || E->has_tsff || E->has_ext || E->has_sf || E->has_liao || E->has_per || g.perx || g.pery) {
        fprintf(stderr, "Error: the gradient supports CPML and point sources only\n");
        return 1;
    }
    if (g.ade && g.ade_cpml) {
        fprintf(stderr, "Error: the gradient does not support ADE_CPML_CONSISTENT\n");
        return 1;
    }
    if (g.rc) {
        fprintf(stderr, "Error: the gradient supports ADE dispersive materials (types 5, 7), not the recursive-convolution "
                        "types 2, 4, 6\n");
        return 1;
    }
    if (g.ade) {
        /* dispersive cells stay out of the CPML slabs (E side) and off the design cells;
         * the class ids come back from the device (the host copy is freed after set-up) */
        const size_t ib = E->idbytes_e;
        std::vector<uint8_t> ids((size_t)g.plane * g.nz * ib);
        CK(hipSetDevice(E->dev));
        CK(hipMemcpy(ids.data(), g.ecls, ids.size(), hipMemcpyDeviceToHost));
        for (int k = 0; k < g.nz; k++)
            for (int j = 0; j < g.ny; j++)
                for (int i = 0; i < g.nx; i++) {
                    const size_t c = (size_t)i + (size_t)g.px * j + (size_t)g.plane * k;
                    const unsigned cls = ib == 1 ? ids[c] : ((const uint16_t *)ids.data())[c];
                    if (!(E->etab[cls].flags & GSV_E_ADE))
                        continue;
                    const bool slab = (g.f[0].on && i < g.f[0].depth) || (g.f[1].on && i >= g.nx - g.f[1].depth)
                                      || (g.f[2].on && j < g.f[2].depth) || (g.f[3].on && j >= g.ny - g.f[3].depth)
                                      || (g.f[4].on && k < g.f[4].depth) || (g.f[5].on && k >= g.nz - g.f[5].depth);
                    const bool pc = i >= g.pc_i0 && i < g.pc_i0 + g.pc_nx && j >= g.pc_j0 && j < g.pc_j0 + g.pc_ny
                                    && k >= g.pc_k0 && k < g.pc_k0 + g.pc_nz;
                    if (slab || pc) {
                        fprintf(stderr, "Error: the gradient needs dispersive cells outside the CPML slabs and the design "
                                        "region (cell %d %d %d)\n", i, j, k);
                        return 1;
                    }
                }
    }
    if (E->has_src)
        for (size_t q = 0; q < (size_t)E->src.nsteps * E->src.n; q++)
            for (int c = 3; c < 6; c++)
                if (E->src_vals_host[q * 6 + c] != 0.f) {
                    fprintf(stderr, "Error: the gradient supports E point sources only\n");
                    return 1;
                }
    /* the slab kernels recover E'' = (adjoint E) / ca and H'' = (adjoint H) / da */
    for (const GsvECoef &c : E->etab)
        if (!(c.flags & GSV_E_PERCELL)
            && ((c.flags & (GSV_E_PECX | GSV_E_PECY | GSV_E_PECZ))
                || (!(c.flags & GSV_E_ADE) && (c.ca[0] == 0.f || c.ca[1] == 0.f || c.ca[2] == 0.f)))) {
            fprintf(stderr, "Error: the gradient supports no PEC cells and needs nonzero E update factors\n");
            return 1;
        }
    for (const GsvHCoef &c : E->htab)
        if (c.da == 0.f) {
            fprintf(stderr, "Error: the gradient needs nonzero H update factors\n");
            return 1;
        }
    for (int ax = 0; ax < 3; ax++) {
        const int n = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
        const int w = (g.f[2 * ax].on ? g.f[2 * ax].depth : 0) + (g.f[2 * ax + 1].on ? g.f[2 * ax + 1].depth : 0);
        if (w > n) {
            fprintf(stderr, "Error: the CPML slabs of one axis overlap\n");
            return 1;
        }
    }
    CK(hipSetDevice(E->dev));
    GsvAdjoint *A = new GsvAdjoint();
    GsvAdj &a = A->a;
    memset(&a, 0, sizeof(a));
    const SvSetAdjoint &sa = set->sa;
    a.mi0 = sa.obj_i0; a.mj0 = sa.obj_j0; a.mk0 = sa.obj_k0;
    a.mnx = sa.obj_in - sa.obj_i0; a.mny = sa.obj_jn - sa.obj_j0; a.mnz = sa.obj_kn - sa.obj_k0;
    a.omega = 2 * M_PI * SV_LIGHT_SPEED / sa.obj_lambda;
    a.dt = set->plan.dt;
    a.ncoef = (long long)g.pc_nx * g.pc_ny * g.pc_nz;
    A->nsteps = set->sc.nsteps;
    if (dmalloc(E, &a.macc, (size_t)a.mnx * a.mny * a.mnz * 6)) { delete A; return 1; }
    float **fa[12] = { &a.ebx, &a.eby, &a.ebz, &a.hbx, &a.hby, &a.hbz, &a.qx, &a.qy, &a.qz, &a.rx, &a.ry, &a.rz };
    for (int q = 0; q < 12; q++) {
        float *p;
        if (dmalloc(E, &p, E->field_floats)) { delete A; return 1; }
        CK(hipMemsetAsync(p, 0, E->field_floats * 4, E->st));
        *fa[q] = p;
    }
    for (int f = 0; f < 6; f++) {
        size_t n = psi_pairs(g.f[f], f / 2, g);
        if (n) {
            if (dmalloc(E, &a.te[f], n) || dmalloc(E, &a.th[f], n)) { delete A; return 1; }
        }
    }
    if (dmalloc(E, &a.gca, 3 * (size_t)a.ncoef) || dmalloc(E, &a.gcb, 3 * (size_t)a.ncoef)) { delete A; return 1; }
    if (g.ade) {
        if (dmalloc(E, &a.xad, (size_t)g.ade_plane_d * g.nz)) { delete A; return 1; }
        if (g.ade_p && dmalloc(E, &a.xap, (size_t)g.ade_plane_p * g.nz)) { delete A; return 1; }
    }
    /* history: all steps if they fit the budget, else one segment of K steps */
    size_t per_step = 6 * (size_t)a.ncoef * 4;
    A->K = ckpt_interval > 0 ? ckpt_interval : 100;
    A->history_all = (long long)per_step * A->nsteps <= history_budget_bytes;
    A->hist_slots = A->history_all ? A->nsteps : (A->K < A->nsteps ? A->K : A->nsteps);
    if (dmalloc(E, &a.hist, per_step / 4 * A->hist_slots)) { delete A; return 1; }
    if (!A->history_all) {
        size_t nf = 6 * E->field_floats + ade_floats(g);
        for (int f = 0; f < 6; f++)
            nf += 2 * 2 * psi_pairs(g.f[f], f / 2, g);
        A->ckpt_floats = nf;
        int nseg = (A->nsteps + A->K - 1) / A->K;
        for (int s = 0; s < nseg; s++) {
            float *p;
            if (dmalloc(E, &p, nf)) { delete A; return 1; }
            A->ckpt.push_back(p);
        }
    }
    if (E->has_src) {
        a.nsrc = E->src.n;
        a.spos = E->src.pos;
        a.svals = E->src.vals;
    }
    CK(hipStreamSynchronize(E->st));
    E->adj = A;
    return 0;
}

/* state copy between the live fields / psi and a checkpoint buffer */
static int
ckpt_copy(Engine *E, float *buf, int save)
{
    GsvGrid &g = E->g;
    size_t off = 0;
    for (int q = 0; q < 6; q++) {
        if (save)
            CK(hipMemcpyAsync(buf + off, E->fbase[q], E->field_floats * 4, hipMemcpyDeviceToDevice, E->st));
        else
            CK(hipMemcpyAsync(E->fbase[q], buf + off, E->field_floats * 4, hipMemcpyDeviceToDevice, E->st));
        off += E->field_floats;
    }
    for (int f = 0; f < 6; f++) {
        size_t n = psi_pairs(g.f[f], f / 2, g);
        if (!n)
            continue;
        float2 *ps[2] = { g.f[f].pe, g.f[f].ph };
        for (int s = 0; s < 2; s++) {
            if (save)
                CK(hipMemcpyAsync(buf + off, ps[s], n * 8, hipMemcpyDeviceToDevice, E->st));
            else
                CK(hipMemcpyAsync(ps[s], buf + off, n * 8, hipMemcpyDeviceToDevice, E->st));
            off += 2 * n;
        }
    }
    if (g.ade) {
        /* the dispersive state travels with the fields */
        const size_t nd = (size_t)g.ade_plane_d * g.nz, np = g.ade_p ? (size_t)g.ade_plane_p * g.nz : 0;
        if (save)
            CK(hipMemcpyAsync(buf + off, g.ade_d, nd * 4, hipMemcpyDeviceToDevice, E->st));
        else
            CK(hipMemcpyAsync(g.ade_d, buf + off, nd * 4, hipMemcpyDeviceToDevice, E->st));
        off += nd;
        if (np) {
            if (save)
                CK(hipMemcpyAsync(buf + off, g.ade_p, np * 4, hipMemcpyDeviceToDevice, E->st));
            else
                CK(hipMemcpyAsync(g.ade_p, buf + off, np * 4, hipMemcpyDeviceToDevice, E->st));
            off += np;
        }
    }
    return 0;
}

/* hooks called by the step entry points (gpu_engine.hip) */
int
gsv_adjoint_pre_e(Engine *E, int step)
{
    GsvAdjoint *A = E->adj;
    if (!A || !A->recording)
        return 0;
    const long long n = A->a.ncoef;
    hipLaunchKernelGGL(k_rec_pre, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, E->st, E->g, A->a,
                       step - A->record_base);
    CK(hipGetLastError());
    return 0;
}

int
gsv_adjoint_post_e(Engine *E, int step)
{
    GsvAdjoint *A = E->adj;
    if (!A || !A->recording)
        return 0;
    const long long n = A->a.ncoef;
    hipLaunchKernelGGL(k_rec_post, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, E->st, E->g, A->a,
                       step - A->record_base);
    CK(hipGetLastError());
    return 0;
}

int
gsv_adjoint_after_step(Engine *E, int step)
{
    GsvAdjoint *A = E->adj;
    if (!A || !A->accumulate)
        return 0;
    const long long n = (long long)A->a.mnx * A->a.mny * A->a.mnz;
    hipLaunchKernelGGL(k_obj_dft, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, E->st, E->g, A->a, step);
    CK(hipGetLastError());
    return 0;
}

/* zero fields, psi and the step counter (a fresh run of the same plan) */
int
gsv_engine_reset_state(Engine *E)
{
    GsvGrid &g = E->g;
    for (int q = 0; q < 6; q++)
        CK(hipMemsetAsync(E->fbase[q], 0, E->field_floats * 4, E->st));
    for (int f = 0; f < 6; f++) {
        size_t n = psi_pairs(g.f[f], f / 2, g);
        if (n) {
            CK(hipMemsetAsync(g.f[f].pe, 0, n * 8, E->st));
            CK(hipMemsetAsync(g.f[f].ph, 0, n * 8, E->st));
        }
    }
    if (g.ade) {
        CK(hipMemsetAsync(g.ade_d, 0, (size_t)g.ade_plane_d * g.nz * 4, E->st));
        if (g.ade_p)
            CK(hipMemsetAsync(g.ade_p, 0, (size_t)g.ade_plane_p * g.nz * 4, E->st));
    }
    CK(hipStreamSynchronize(E->st));
    return 0;
}

/* the step sequence of one forward step (src3d/svit.c:239-257) */
typedef int (*GsvStepFn)(void *ctx);

/* Forward pass for the gradient: fresh state, every step, objective DFT,
 * checkpoints (or the full history); returns J. */
extern "C" int
gsv_engine_adjoint_forward(void *eng, int (*step_fn)(void *), void *ctx, double *J)
{
    Engine *E = (Engine *)eng;
    GsvAdjoint *A = E->adj;
    SvSet *set = E->set;
    if (!A)
        return 1;
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    if (gsv_engine_reset_state(E))
        return 1;
    const size_t nm = (size_t)A->a.mnx * A->a.mny * A->a.mnz * 6;
    CK(hipMemsetAsync(A->a.macc, 0, nm * 8, E->st));
    set->sc.step_act = 0;
    A->accumulate = 1;
    A->recording = A->history_all;
    A->record_base = 0;
    CK(hipEventRecord(e0, E->st));
    for (int n = 0; n < A->nsteps; n++) {
        if (!A->history_all && n % A->K == 0 && ckpt_copy(E, A->ckpt[n / A->K], 1))
            return 1;
        if (step_fn(ctx))
            return 1;
    }
    CK(hipEventRecord(e1, E->st));
    A->accumulate = 0;
    A->recording = 0;
    std::vector<double> F(nm);
    CK(hipMemcpyAsync(F.data(), A->a.macc, nm * 8, hipMemcpyDeviceToHost, E->st));
    CK(hipStreamSynchronize(E->st));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    A->ms_forward = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    double j = 0;
    for (size_t q = 0; q < nm; q++)
        j += F[q] * F[q];
    *J = j;
    return 0;
}

/* one backward step: slab terms around the two bulk kernels (the bulk
 * kernels use the E kernel's launch shape and 32-bit offsets, which the design
 * region already requires) */
template <typename ID>
static int
launch_adj(Engine *E, int step, int slot)
{
    GsvGrid &g = E->g;
    const GsvAdj &a = E->adj->a;
    const int bx = E->bx_e, by = E->by_e, kc = E->kchunk_e;
    dim3 blk(bx, by, 1), grd((g.nx + bx - 1) / bx, (g.ny + by - 1) / by, (g.nz + kc - 1) / kc);
    long long win[3], fmax = 0;
    for (int ax = 0; ax < 3; ax++) {
        const int w = (g.f[2 * ax].on ? g.f[2 * ax].depth : 0) + (g.f[2 * ax + 1].on ? g.f[2 * ax + 1].depth : 0);
        win[ax] = ax == 0 ? (long long)w * g.ny * g.nz : (ax == 1 ? (long long)g.nx * w * g.nz : (long long)g.nx * g.ny * w);
    }
    for (int f = 0; f < 6; f++)
        if (g.f[f].on) {
            const long long d = g.f[f].depth;
            const long long n = f < 2 ? d * g.ny * g.nz : (f < 4 ? (long long)g.px * d * g.nz : (long long)g.px * g.ny * d);
            fmax = n > fmax ? n : fmax;
        }
#define GSV_WIN(n) dim3((unsigned)(((n) + 255) / 256)), dim3(256), 0, E->st
    if (win[0]) hipLaunchKernelGGL(k_slab_e_pre<0>, GSV_WIN(win[0]), g, a);
    if (win[1]) hipLaunchKernelGGL(k_slab_e_pre<1>, GSV_WIN(win[1]), g, a);
    if (win[2]) hipLaunchKernelGGL(k_slab_e_pre<2>, GSV_WIN(win[2]), g, a);
    /* the dispersive branch as in the forward E kernel: none, Drude only, 2 or 3 poles */
    const int ade = g.ade ? (g.ade_poles ? g.ade_poles : 1) : 0;
    if (ade == 0)
        hipLaunchKernelGGL((k_adj_e<ID, uint32_t, 0>), grd, blk, 0, E->st, g, a, step, slot, kc);
    else if (ade == 1)
        hipLaunchKernelGGL((k_adj_e<ID, uint32_t, 1>), grd, blk, 0, E->st, g, a, step, slot, kc);
    else if (ade == 2)
        hipLaunchKernelGGL((k_adj_e<ID, uint32_t, 2>), grd, blk, 0, E->st, g, a, step, slot, kc);
    else
        hipLaunchKernelGGL((k_adj_e<ID, uint32_t, 3>), grd, blk, 0, E->st, g, a, step, slot, kc);
    if (fmax) hipLaunchKernelGGL(k_slab_e_post<ID>, dim3((unsigned)((fmax + 255) / 256), 6), dim3(256), 0, E->st, g, a);
    if (win[0]) hipLaunchKernelGGL(k_slab_h_pre<0>, GSV_WIN(win[0]), g, a);
    if (win[1]) hipLaunchKernelGGL(k_slab_h_pre<1>, GSV_WIN(win[1]), g, a);
    if (win[2]) hipLaunchKernelGGL(k_slab_h_pre<2>, GSV_WIN(win[2]), g, a);
    hipLaunchKernelGGL((k_adj_h<ID, uint32_t>), grd, blk, 0, E->st, g, a, kc);
    if (fmax) hipLaunchKernelGGL(k_slab_h_post<ID>, dim3((unsigned)((fmax + 255) / 256), 6), dim3(256), 0, E->st, g, a);
#undef GSV_WIN
    CK(hipGetLastError());
    return 0;
}

/* Backward pass: coefficient gradients accumulated on the device; the host
 * chains them through the permittivity averaging to the design voxels.
 * grad: dJ / d eps_r per design voxel, [i][j][k] order over the design box. */
extern "C" int
gsv_engine_adjoint_backward(void *eng, int (*step_fn)(void *), void *ctx, float *grad)
{
    Engine *E = (Engine *)eng;
    GsvAdjoint *A = E->adj;
    SvSet *set = E->set;
    GsvGrid &g = E->g;
    if (!A)
        return 1;
    if (E->idbytes_e != E->idbytes_h) {
        fprintf(stderr, "Error: the gradient needs E and H class ids of the same width\n");
        return 1;
    }
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    GsvAdj &a = A->a;
    for (float *p : { a.ebx, a.eby, a.ebz, a.hbx, a.hby, a.hbz, a.qx, a.qy, a.qz, a.rx, a.ry, a.rz })
        CK(hipMemsetAsync(p, 0, E->field_floats * 4, E->st));
    for (int f = 0; f < 6; f++) {
        size_t n = psi_pairs(g.f[f], f / 2, g);
        if (n) {
            CK(hipMemsetAsync(a.te[f], 0, n * 8, E->st));
            CK(hipMemsetAsync(a.th[f], 0, n * 8, E->st));
        }
    }
    CK(hipMemsetAsync(a.gca, 0, 3 * a.ncoef * 8, E->st));
    CK(hipMemsetAsync(a.gcb, 0, 3 * a.ncoef * 8, E->st));
    if (a.xad)
        CK(hipMemsetAsync(a.xad, 0, (size_t)g.ade_plane_d * g.nz * 4, E->st));
    if (a.xap)
        CK(hipMemsetAsync(a.xap, 0, (size_t)g.ade_plane_p * g.nz * 4, E->st));
    CK(hipEventRecord(e0, E->st));
    const int nseg = A->history_all ? 1 : (A->nsteps + A->K - 1) / A->K;
    for (int s = nseg - 1; s >= 0; s--) {
        int n0 = 0, n1 = A->nsteps;
        if (!A->history_all) {
            /* rebuild the segment's history from its checkpoint */
            n0 = s * A->K;
            n1 = n0 + A->K < A->nsteps ? n0 + A->K : A->nsteps;
            if (ckpt_copy(E, A->ckpt[s], 0))
                return 1;
            A->recording = 1;
            A->record_base = n0;
            set->sc.step_act = n0;
            for (int n = n0; n < n1; n++)
                if (step_fn(ctx))
                    return 1;
            A->recording = 0;
        }
        for (int n = n1 - 1; n >= n0; n--) {
            const int slot = A->history_all ? n : n - n0;
            if (E->idbytes_e == 1 ? launch_adj<uint8_t>(E, n, slot) : launch_adj<uint16_t>(E, n, slot))
                return 1;
        }
    }
    CK(hipEventRecord(e1, E->st));
    std::vector<double> gca(3 * a.ncoef), gcb(3 * a.ncoef);
    CK(hipMemcpyAsync(gca.data(), a.gca, gca.size() * 8, hipMemcpyDeviceToHost, E->st));
    CK(hipMemcpyAsync(gcb.data(), a.gcb, gcb.size() * 8, hipMemcpyDeviceToHost, E->st));
    CK(hipStreamSynchronize(E->st));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    A->ms_backward = ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    set->sc.step_act = A->nsteps;

    /* chain rule through ca, cb of each coefficient cell to the averaged
     * permittivity and on to the voxels (sv_pool_cabs_vjp, plan.c) */
    const SvSetAdjoint &sa = set->sa;
    const int dnx = sa.design_in - sa.design_i0, dny = sa.design_jn - sa.design_j0, dnz = sa.design_kn - sa.design_k0;
    std::vector<double> gv((size_t)dnx * dny * dnz, 0.0);
    for (int k = 0; k < g.pc_nz; k++)
        for (int j = 0; j < g.pc_ny; j++)
            for (int i = 0; i < g.pc_nx; i++) {
                const size_t q = (size_t)i + (size_t)g.pc_nx * (j + (size_t)g.pc_ny * k);
                double ga[3] = { gca[q], gca[a.ncoef + q], gca[2 * a.ncoef + q] };
                double gb[3] = { gcb[q], gcb[a.ncoef + q], gcb[2 * a.ncoef + q] };
                sv_pool_cabs_vjp(E->mp, i + g.pc_i0, j + g.pc_j0, k + g.pc_k0, ga, gb, sa.design_i0, sa.design_j0,
                                 sa.design_k0, dnx, dny, dnz, gv.data());
            }
    for (size_t q = 0; q < gv.size(); q++)
        grad[q] = (float)(gv[q] * SV_EPSILON_0);
    return 0;
}

/* relative permittivity of the design voxels, [i][j][k] over the design box */
extern "C" int
gsv_engine_set_design_eps(void *eng, const float *eps_r)
{
    Engine *E = (Engine *)eng;
    SvSet *set = E->set;
    const SvSetAdjoint &sa = set->sa;
    if (!sa.design || !E->ready)
        return 1;
    const int dnx = sa.design_in - sa.design_i0, dny = sa.design_jn - sa.design_j0, dnz = sa.design_kn - sa.design_k0;
    SvFCube *ep = E->mp->d->epsilon;
    for (int i = 0; i < dnx; i++)
        for (int j = 0; j < dny; j++)
            for (int k = 0; k < dnz; k++) {
                const float v = eps_r[((size_t)i * dny + j) * dnz + k];
                if (!(v > 0)) {
                    fprintf(stderr, "Error: design permittivity must be positive\n");
                    return 1;
                }
                /* the same float product as the MEDIUM_LINEAR reader */
                ep->data[i + sa.design_i0][j + sa.design_j0][k + sa.design_k0] = v * (float)SV_EPSILON_0;
            }
    if (gsv_build_pctab(E))
        return 1;
    CK(hipMemcpyAsync((void *)E->g.pctab, E->pctab_host.data(), E->pctab_host.size() * sizeof(GsvECoef),
                      hipMemcpyHostToDevice, E->st));
    CK(hipStreamSynchronize(E->st));
    return 0;
}

extern "C" int
gsv_engine_get_design_eps(void *eng, float *eps_r)
{
    Engine *E = (Engine *)eng;
    const SvSetAdjoint &sa = E->set->sa;
    if (!sa.design)
        return 1;
    const int dnx = sa.design_in - sa.design_i0, dny = sa.design_jn - sa.design_j0, dnz = sa.design_kn - sa.design_k0;
    SvFCube *ep = E->mp->d->epsilon;
    for (int i = 0; i < dnx; i++)
        for (int j = 0; j < dny; j++)
            for (int k = 0; k < dnz; k++)
                eps_r[((size_t)i * dny + j) * dnz + k] =
                    (float)(ep->data[i + sa.design_i0][j + sa.design_j0][k + sa.design_k0] / SV_EPSILON_0);
    return 0;
}

extern "C" void
gsv_engine_adjoint_info(void *eng, double *out)
{
    Engine *E = (Engine *)eng;
    GsvAdjoint *A = E->adj;
    out[0] = A ? A->ms_forward : 0;
    out[1] = A ? A->ms_backward : 0;
    out[2] = A ? A->history_all : 0;
    out[3] = A ? A->K : 0;
    out[4] = A ? (double)A->ckpt.size() : 0;
}
