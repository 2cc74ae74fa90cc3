/*
 * kernels.hip -- CDNA4 (gfx950) kernels of the FDTD time step.
 *
 * Volume kernels (one wave64 covers 64 consecutive x cells of one row, a block
 * covers BX x BY columns and marches up a chunk of z planes, keeping the
 * previous plane's values in registers):
 *   k_hstep   Yee H update (src3d/yee.c:145-265) fused with the CPML H psi
 *             update, H correction and far-plane zeroing (src3d/boundary.c:347-518)
 *   k_estep   Yee E update (src3d/yee.c:528-1247: per-component averaged
 *             coefficients via classes, ADE branch with its inline CPML, PEC)
 *             fused with the CPML E psi update, E correction and near-plane
 *             zeroing (src3d/boundary.c:691-887)
 * Surface kernels:
 *   k_tsf_line            1-D auxiliary line step, src3d/1dpool.c:83-162
 *   k_tsf_h / k_tsf_e     TSF face corrections, src3d/source.c:547-783, 1464-1657
 *   k_tsff_line / k_tsff_face   focused TSF: one line per plane wave, face
 *                         corrections summed over the waves, src3d/source.c:773-930, 1892-2060
 *   k_point_src           hard point sources, src3d/source.c:405-437, 1441-1462
 *   k_liao_store / k_liao_face   Liao ABC, src3d/boundary.c:549-689, 938-1066
 *   k_zero_planes         CPML near-plane zeroing after Liao on mixed boundaries
 *   k_periodic_h / _e     interior periodic plane copies, src3d/boundary.c:304-346, 889-932
 *   k_record_points       point traces, src3d/output.c:340-352
 *   k_farfield_td         time-domain Kirchhoff bins, src3d/farfield.c:174-407
 *   k_ltsf_line           layered-TSF split-field 1-D model, src3d/zpool.c:172-338
 *   k_ltsf_face           layered-TSF face corrections, src3d/source.c:921-1150, 1700-1900
 *   k_force               stress tensor on force-box faces, src3d/output.c:355-512
 *   k_pfarfield           periodic Kirchhoff bins (z planes, images), src3d/pfarfield.c:120-401
 *   k_dft_sample / k_dft_flush   running DFT of tangential fields on the NFFF box
 *   k_box_sum             OUT_SUM box sums of E^2, src3d/output.c:259-306
 * Batched variants (k_*_batch) run the same column and face-cell functions
 * over a tile list that covers every grid of a batch of independent runs.
 */
#include <hip/hip_runtime.h>
#include "gsvit_settings.h"
#include "engine.h"
#include "kernels.h"
#include "device_access.h"



/* ------------------------------------------------------------------ */
/* H update                                                             */
/* ------------------------------------------------------------------ */

/* column (i, j) of the H update over the planes of z chunk kz (kchunk planes each) */
template <typename ID, typename OFF, bool PER>
__device__ __forceinline__ void
hstep_col(const GsvGrid &g, int i, int j, int kz, int kchunk)
{
    if (i >= g.nx || j >= g.ny)
        return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int k0 = kz * kchunk;
    const int k1 = min(k0 + kchunk, nz);
    const int px = g.px;
    const OFF dP = (OFF)g.plane * 4, dJ = (OFF)px * 4;
    const OFF dPc = (OFF)g.plane * sizeof(ID);
    /* o: field byte offset of (i, j, k) from the guard plane; oc: class id offset */
    OFF o = ((OFF)i + (OFF)px * j + (OFF)g.plane * (k0 + 1)) * 4;
    OFF oc = ((OFF)i + (OFF)px * j + (OFF)g.plane * k0) * sizeof(ID);

    const float rkx = gp(g.rkhx)[i], rky = gp(g.rkhy)[j];
    /* a periodic axis has no boundary planes: every column is interior, and the
     * +x / +y neighbour of the far plane is the near plane.  Pools without
     * periodic axes use the PER = false build, whose neighbour offsets are the
     * constants 4 and dJ (immediate and scalar operands of the loads). */
    const bool perx = PER && g.perx, pery = PER && g.pery;
    const bool iin = (i >= 1 && i < nx - 1) || perx, jin = (j >= 1 && j < ny - 1) || pery;
    const bool xe = i == nx - 1 && !perx, ye = j == ny - 1 && !pery;
    const OFF dIp = (perx && i == nx - 1) ? (OFF)0 - (OFF)(nx - 1) * 4 : (OFF)4;
    const OFF dJp = (pery && j == ny - 1) ? (OFF)0 - (OFF)(ny - 1) * dJ : dJ;
    const GsvCpmlFace &F0 = g.f[0], &F1 = g.f[1], &F2 = g.f[2], &F3 = g.f[3], &F4 = g.f[4], &F5 = g.f[5];
    const bool sx0 = F0.on && i < F0.depth - 1 && jin;
    const int irn = i - (nx - F1.depth);
    const bool sxn = F1.on && i >= nx - F1.depth + 1 && i < nx - 1 && jin;
    const bool sy0 = F2.on && j < F2.depth - 1 && iin;
    const int jrn = j - (ny - F3.depth);
    const bool syn = F3.on && j >= ny - F3.depth + 1 && j < ny - 1 && iin;
    const bool zx = F1.on && i == nx - 1 && jin;
    const bool zy = F3.on && j == ny - 1 && iin;
    const bool ijin = iin && jin;

    float exk = ldf(g.exg, o), eyk = ldf(g.eyg, o);
    /* class id one plane ahead: the coefficient gather starts with the plane's loads */
    unsigned idn = ldid<ID>(g.hcls, oc);
    for (int k = k0; k < k1; ++k, o += dP, oc += dPc) {
        const float ezk = ldf(g.ezg, o);
        const float exk1 = ldf(g.exg, o + dP), eyk1 = ldf(g.eyg, o + dP);
        const float ezj1 = ldf(g.ezg, o + dJp), exj1 = ldf(g.exg, o + dJp);
        const float ezi1 = ldf(g.ezg, o + dIp), eyi1 = ldf(g.eyg, o + dIp);
        float hx = ldn(g.hxg, o), hy = ldn(g.hyg, o), hz = ldn(g.hzg, o);
        const GsvHCoef hc = ldhc(g.htab + idn);
        if (k + 1 < k1)
            idn = ldid<ID>(g.hcls, oc + dPc);
        const float rkz = gp(g.rkhz)[k];
        const bool kin = (k >= 1 && k < nz - 1);

        if (!(ye || k == nz - 1))
            hx = hc.da * hx + hc.db[0] * ((eyk1 - eyk) * rkz - (ezj1 - ezk) * rky);
        if (!(xe || k == nz - 1))
            hy = hc.da * hy + hc.db[1] * ((ezi1 - ezk) * rkx - (exk1 - exk) * rkz);
        if (!(xe || ye))
            hz = hc.da * hz + hc.db[2] * ((exj1 - exk) * rky - (eyi1 - eyk) * rkx);

        /* CPML H psi and correction, src3d/boundary.c:347-518 */
        if (kin) {
            if (sx0) {
                const OFF p = (OFF)(((unsigned)k * ny + j) * F0.depth + i) * 8;
                const float2 bc = ldg2(F0.ch + i);
                const float2 ps = ld2(F0.ph, p);
                const float a1 = bc.x * ps.x + bc.y * (eyi1 - eyk);
                const float a2 = bc.x * ps.y + bc.y * (ezi1 - ezk);
                st2(F0.ph, p, a1, a2);
                hy += g.dx * hc.dbc * a2;
                hz -= g.dx * hc.dbc * a1;
            }
            if (sxn) {
                const OFF p = (OFF)(((unsigned)k * ny + j) * F1.depth + irn) * 8;
                const float2 bc = ldg2(F1.ch + irn);
                const float2 ps = ld2(F1.ph, p);
                const float a1 = bc.x * ps.x + bc.y * (eyi1 - eyk);
                const float a2 = bc.x * ps.y + bc.y * (ezi1 - ezk);
                st2(F1.ph, p, a1, a2);
                hy += g.dx * hc.dbc * a2;
                hz -= g.dx * hc.dbc * a1;
            }
            if (sy0) {
                const OFF p = (OFF)(((unsigned)k * F2.depth + j) * px + i) * 8;
                const float2 bc = ldg2(F2.ch + j);
                const float2 ps = ld2(F2.ph, p);
                const float a1 = bc.x * ps.x + bc.y * (ezj1 - ezk);
                const float a2 = bc.x * ps.y + bc.y * (exj1 - exk);
                st2(F2.ph, p, a1, a2);
                hx -= g.dy * hc.dbc * a1;
                hz += g.dy * hc.dbc * a2;
            }
            if (syn) {
                const OFF p = (OFF)(((unsigned)k * F3.depth + jrn) * px + i) * 8;
                const float2 bc = ldg2(F3.ch + jrn);
                const float2 ps = ld2(F3.ph, p);
                const float a1 = bc.x * ps.x + bc.y * (ezj1 - ezk);
                const float a2 = bc.x * ps.y + bc.y * (exj1 - exk);
                st2(F3.ph, p, a1, a2);
                hx -= g.dy * hc.dbc * a1;
                hz += g.dy * hc.dbc * a2;
            }
        }
        if (ijin) {
            if (F4.on && k < F4.depth - 1) {
                const OFF p = (OFF)(((unsigned)k * ny + j) * px + i) * 8;
                const float2 bc = ldg2(F4.ch + k);
                const float2 ps = ld2(F4.ph, p);
                const float a1 = bc.x * ps.x + bc.y * (exk1 - exk);
                const float a2 = bc.x * ps.y + bc.y * (eyk1 - eyk);
                st2(F4.ph, p, a1, a2);
                hy -= g.dz * hc.dbc * a1;
                hx += g.dz * hc.dbc * a2;
            }
            if (F5.on && k >= nz - F5.depth + 1 && k < nz - 1) {
                const int kr = k - (nz - F5.depth);
                const OFF p = (OFF)(((unsigned)kr * ny + j) * px + i) * 8;
                const float2 bc = ldg2(F5.ch + kr);
                const float2 ps = ld2(F5.ph, p);
                const float a1 = bc.x * ps.x + bc.y * (exk1 - exk);
                const float a2 = bc.x * ps.y + bc.y * (eyk1 - eyk);
                st2(F5.ph, p, a1, a2);
                hy -= g.dz * hc.dbc * a1;
                hx += g.dz * hc.dbc * a2;
            }
        }
        if ((kin && (zx || zy)) || (F5.on && k == nz - 1 && ijin)) {
            hx = 0.f;
            hy = 0.f;
            hz = 0.f;
        }
        stn(g.hxg, o, hx);
        stn(g.hyg, o, hy);
        stn(g.hzg, o, hz);
        exk = exk1;
        eyk = eyk1;
    }
}

template <typename ID, typename OFF, bool PER>
__global__ __launch_bounds__(512) void
k_hstep(GsvGrid g, int kchunk)
{
    hstep_col<ID, OFF, PER>(g, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y,
                       blockIdx.z, kchunk);
}

/* ------------------------------------------------------------------ */
/* E update                                                             */
/* ------------------------------------------------------------------ */

/* ADE branch of one cell (src3d/yee.c:914-1108); updates the components whose
 * flag is clear, runs the branch's own CPML psi update and advances the
 * polarisation state (increment form, see GsvAde).  Taken in dispersive cells
 * and their +1 neighbours.  ad/ap: this plane's ADE state blocks, la/lp: the
 * lane's offset in them; state q of a cell sits q * 256 bytes after its first
 * state. */
/* the Drude states a dispersive cell reads whatever its material: states 0-2
 * (E^(n-1), or Q in the Drude-only form) and, in Drude-only builds that keep
 * J, states 3-5; loaded by the caller with the plane's field loads */
struct AdePre {
    float s[6];
};

template <int NP, bool QF, typename OFF>
__device__ __forceinline__ AdePre
ade_preload(const float *ad, OFF la)
{
    AdePre p;
#pragma unroll
    for (int q = 0; q < 3; q++)
        p.s[q] = ldsn(ad, la, q * 256);
#pragma unroll
    for (int q = 3; q < 6; q++)
        p.s[q] = (NP == 0 && !QF) ? ldsn(ad, la, q * 256) : 0.f;
    return p;
}

template <int NP, bool PER, bool QF, typename OFF>
__device__ __forceinline__ void
ade_cell(const GsvGrid &g, float *ad, OFF la, float *ap, OFF lp, const AdePre &pre, int i, int j, int k, unsigned fl,
         float &ex, float &ey, float &ez, float hx, float hy, float hz, float hxkm1, float hykm1, float hzi, float hyi,
         float hzj, float hxj, float rkx, float rky, float rkz)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, px = g.px;
    const GSV_GLOBAL GsvAde &m = gp(g.adem)[(fl >> GSV_E_SLOT_SHIFT) & 0xF];
    /* the Drude-only form's coefficients in one 16-byte gather and one load,
     * instead of one dependent load per coefficient and component (the inline
     * CPML's cb, read in slab cells only, stays where it is used) */
    struct {
        float c0, c3, dj, qa, qb;
    } mq = {};
    if (QF) {
        const gsv_v4f a = *(const GSV_GLOBAL gsv_v4f *)(const void *)&m;
        mq.c0 = a.x;
        mq.c3 = a.y;
        mq.dj = a.z;
        mq.qa = a.w;
        mq.qb = m.qb;
    }
    const bool tx = fl & GSV_E_TBCX, ty = fl & GSV_E_TBCY, tz = fl & GSV_E_TBCZ;
    const bool xi = i > 0 || (PER && g.perx), yi = j > 0 || (PER && g.pery);   /* not on a (non-periodic) near plane */
    const float vexp = ex, veyp = ey, vezp = ez;
    /* states 0-2: E^(n-1), or Q^n in the Drude-only form (QF, engine.h) */
    const float expv = pre.s[0], eypv = pre.s[1], ezpv = pre.s[2];
    /* state that stays zero (an undriven Drude current, the poles beyond the
     * material's own in a pool of wider materials) is neither read nor written */
    const bool dr = !QF && (NP == 0 || m.drude);   /* Drude-only builds always carry J (or Q) */
    const int npl = NP ? m.poles : 0;
    float jx = 0.f, jy = 0.f, jz = 0.f;
    if (NP == 0) {
        jx = pre.s[3];
        jy = pre.s[4];
        jz = pre.s[5];
    } else if (dr) {
        jx = ldsn(ad, la, 768);
        jy = ldsn(ad, la, 1024);
        jz = ldsn(ad, la, 1280);
    }
    /* pole state (NP poles): P of pole q, component c at slot 3q + c, J at slot 3 NP + 3q + c */
    float pp[3][3] = {}, jj[3][3] = {};
#pragma unroll
    for (int q = 0; q < NP; q++)
        if (q < npl) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                pp[q][c] = ldsn(ap, lp, (3 * q + c) * 256);
                jj[q][c] = ldsn(ap, lp, (3 * NP + 3 * q + c) * 256);
            }
        }
    float cx = m.cj * jx, cy = m.cj * jy, cz = m.cj * jz;   /* dispersive current terms of the E update */
#pragma unroll
    for (int q = 0; q < NP; q++) {
        cx += m.pcj[q] * jj[q][0] + m.pcp[q] * pp[q][0];
        cy += m.pcj[q] * jj[q][1] + m.pcp[q] * pp[q][1];
        cz += m.pcj[q] * jj[q][2] + m.pcp[q] * pp[q][2];
    }

    if (QF) {
        if (yi && k > 0 && !tx)
            ex = mq.c0 * ((hz - hzj) * rky - (hy - hykm1) * rkz) / g.dx + expv + mq.c3 * ex;
        if (xi && k > 0 && !ty)
            ey = mq.c0 * ((hx - hxkm1) * rkz - (hz - hzi) * rkx) / g.dy + eypv + mq.c3 * ey;
        if (xi && yi && !tz)
            ez = mq.c0 * ((hy - hyi) * rkx - (hx - hxj) * rky) / g.dz + ezpv + mq.c3 * ez;
    } else {
        if (yi && k > 0 && !tx)
            ex = m.c0 * ((hz - hzj) * rky - (hy - hykm1) * rkz) / g.dx + cx + m.c2 * expv + m.c3 * ex;
        if (xi && k > 0 && !ty)
            ey = m.c0 * ((hx - hxkm1) * rkz - (hz - hzi) * rkx) / g.dy + cy + m.c2 * eypv + m.c3 * ey;
        if (xi && yi && !tz)
            ez = m.c0 * ((hy - hyi) * rkx - (hx - hxj) * rky) / g.dz + cz + m.c2 * ezpv + m.c3 * ez;
    }

    const GsvCpmlFace &F0 = g.f[0], &F1 = g.f[1], &F2 = g.f[2], &F3 = g.f[3], &F4 = g.f[4], &F5 = g.f[5];
    if (g.ade_cpml) {
        /* ADE_CPML_CONSISTENT: the psi of each dispersive component, once per
         * step over the slab ranges of the ordinary update, applied with c0 */
        const float c0 = QF ? mq.c0 : m.c0;
        if (F0.on && i >= 1 && i < F0.depth) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * F0.depth + i) * 8;
            const float4 bc = ldg4(F0.ce + i);
            const float2 ps = ld2(F0.pe, p);
            const float a1 = ty ? ps.x : bc.x * ps.x + bc.y * (hz - hzi);
            const float a2 = tz ? ps.y : bc.x * ps.y + bc.y * (hy - hyi);
            st2(F0.pe, p, a1, a2);
            if (yi && k != 0) {
                if (!ty) ey -= c0 * a1;
                if (!tz) ez += c0 * a2;
            }
        }
        if (F1.on && i >= nx - F1.depth + 1) {
            const int ir = i - (nx - F1.depth);
            const OFF p = (OFF)(((unsigned)k * ny + j) * F1.depth + ir) * 8;
            const float4 bc = ldg4(F1.ce + ir);
            const float2 ps = ld2(F1.pe, p);
            const float a1 = ty ? ps.x : bc.x * ps.x + bc.y * (hz - hzi);
            const float a2 = tz ? ps.y : bc.x * ps.y + bc.y * (hy - hyi);
            st2(F1.pe, p, a1, a2);
            if (yi && k != 0) {
                if (!ty) ey -= c0 * a1;
                if (!tz) ez += c0 * a2;
            }
        }
        if (F2.on && j >= 1 && j < F2.depth) {
            const OFF p = (OFF)(((unsigned)k * F2.depth + j) * px + i) * 8;
            const float4 bc = ldg4(F2.ce + j);
            const float2 ps = ld2(F2.pe, p);
            const float a1 = tx ? ps.x : bc.x * ps.x + bc.y * (hz - hzj);
            const float a2 = tz ? ps.y : bc.x * ps.y + bc.y * (hx - hxj);
            st2(F2.pe, p, a1, a2);
            if (xi && k != 0) {
                if (!tx) ex += c0 * a1;
                if (!tz) ez -= c0 * a2;
            }
        }
        if (F3.on && j >= ny - F3.depth + 1) {
            const int jr = j - (ny - F3.depth);
            const OFF p = (OFF)(((unsigned)k * F3.depth + jr) * px + i) * 8;
            const float4 bc = ldg4(F3.ce + jr);
            const float2 ps = ld2(F3.pe, p);
            const float a1 = tx ? ps.x : bc.x * ps.x + bc.y * (hz - hzj);
            const float a2 = tz ? ps.y : bc.x * ps.y + bc.y * (hx - hxj);
            st2(F3.pe, p, a1, a2);
            if (xi && k != 0) {
                if (!tx) ex += c0 * a1;
                if (!tz) ez -= c0 * a2;
            }
        }
        if (F4.on && k >= 1 && k < F4.depth) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * px + i) * 8;
            const float4 bc = ldg4(F4.ce + k);
            const float2 ps = ld2(F4.pe, p);
            const float a1 = ty ? ps.x : bc.x * ps.x + bc.y * (hx - hxkm1);
            const float a2 = tx ? ps.y : bc.x * ps.y + bc.y * (hy - hykm1);
            st2(F4.pe, p, a1, a2);
            if (xi && yi) {
                if (!ty) ey += c0 * a1;
                if (!tx) ex -= c0 * a2;
            }
        }
        if (F5.on && k >= nz - F5.depth + 1) {
            const int kr = k - (nz - F5.depth);
            const OFF p = (OFF)(((unsigned)kr * ny + j) * px + i) * 8;
            const float4 bc = ldg4(F5.ce + kr);
            const float2 ps = ld2(F5.pe, p);
            const float a1 = ty ? ps.x : bc.x * ps.x + bc.y * (hx - hxkm1);
            const float a2 = tx ? ps.y : bc.x * ps.y + bc.y * (hy - hykm1);
            st2(F5.pe, p, a1, a2);
            if (xi && yi) {
                if (!ty) ey += c0 * a1;
                if (!tx) ex -= c0 * a2;
            }
        }
    } else {
    /* the branch's own CPML update (src3d/yee.c:950-1057); its x0 correction
     * multiplies by a zero cb */
    if (F0.on && i > 1 && i < F0.depth) {
        const OFF p = (OFF)(((unsigned)k * ny + j) * F0.depth + i) * 8;
        const float4 bc = ldg4(F0.ce + i);
        const float2 ps = ld2(F0.pe, p);
        st2(F0.pe, p, bc.x * ps.x + bc.z * (hz - hzi) / g.dy, bc.x * ps.y + bc.z * (hy - hyi) / g.dz);
    }
    if (F2.on && j > 1 && j < F2.depth) {
        const OFF p = (OFF)(((unsigned)k * F2.depth + j) * px + i) * 8;
        const float4 bc = ldg4(F2.ce + j);
        const float2 ps = ld2(F2.pe, p);
        const float a1 = bc.x * ps.x + bc.z * (hz - hzj) / g.dz;
        const float a2 = bc.x * ps.y + bc.z * (hx - hxj) / g.dx;
        st2(F2.pe, p, a1, a2);
        if ((xi && k != 0) && !tx) ex += g.dx * m.cbm[0] * a1;
        if ((xi && k != 0) && !tz) ez -= g.dz * m.cbm[2] * a2;
    }
    if (F4.on && k > 1 && k < F4.depth) {
        const OFF p = (OFF)(((unsigned)k * ny + j) * px + i) * 8;
        const float4 bc = ldg4(F4.ce + k);
        const float2 ps = ld2(F4.pe, p);
        const float a1 = bc.x * ps.x + bc.z * (hx - hxkm1) / g.dy;
        const float a2 = bc.x * ps.y + bc.z * (hy - hykm1) / g.dx;
        st2(F4.pe, p, a1, a2);
        if ((xi && yi) && !ty) ey += g.dy * m.cbm[1] * a1;
        if ((xi && yi) && !tx) ex -= g.dx * m.cbm[0] * a2;
    }
    if (F1.on && i >= nx - F1.depth + 1) {
        const int ir = i - (nx - F1.depth);
        const OFF p = (OFF)(((unsigned)k * ny + j) * F1.depth + ir) * 8;
        const float4 bc = ldg4(F1.ce + ir);
        const float2 ps = ld2(F1.pe, p);
        const float a1 = bc.x * ps.x + bc.z * (hz - hzi) / g.dy;
        const float a2 = bc.x * ps.y + bc.z * (hy - hyi) / g.dz;
        st2(F1.pe, p, a1, a2);
        if ((yi && k != 0) && !ty) ey -= g.dy * m.cbm[1] * a1;
        if ((yi && k != 0) && !tz) ez += g.dz * m.cbm[2] * a2;
    }
    if (F3.on && j >= ny - F3.depth + 1) {
        const int jr = j - (ny - F3.depth);
        const OFF p = (OFF)(((unsigned)k * F3.depth + jr) * px + i) * 8;
        const float4 bc = ldg4(F3.ce + jr);
        const float2 ps = ld2(F3.pe, p);
        const float a1 = bc.x * ps.x + bc.z * (hz - hzj) / g.dx;
        const float a2 = bc.x * ps.y + bc.z * (hx - hxj) / g.dz;
        st2(F3.pe, p, a1, a2);
        if ((xi && k != 0) && !tx) ex += g.dx * m.cbm[0] * a1;
        if ((xi && k != 0) && !tz) ez -= g.dz * m.cbm[2] * a2;
    }
    if (F5.on && k >= nz - F5.depth + 1) {
        const int kr = k - (nz - F5.depth);
        const OFF p = (OFF)(((unsigned)kr * ny + j) * px + i) * 8;
        const float4 bc = ldg4(F5.ce + kr);
        const float2 ps = ld2(F5.pe, p);
        const float a1 = bc.x * ps.x + bc.z * (hx - hxkm1) / g.dy;
        const float a2 = bc.x * ps.y + bc.z * (hy - hykm1) / g.dx;
        st2(F5.pe, p, a1, a2);
        if ((xi && yi) && !ty) ey += g.dy * m.cbm[1] * a1;
        if ((xi && yi) && !tx) ex -= g.dx * m.cbm[0] * a2;
    }
    }

    /* polarisation state (all three components, with the values at this point) */
    if (QF) {
        stsn(ad, la, 0, mq.dj * expv + mq.qa * vexp + mq.qb * ex);
        stsn(ad, la, 256, mq.dj * eypv + mq.qa * veyp + mq.qb * ey);
        stsn(ad, la, 512, mq.dj * ezpv + mq.qa * vezp + mq.qb * ez);
        return;
    }
    if (dr) {
        stsn(ad, la, 768, m.dj * jx + m.da * (expv + 2.f * vexp + ex));
        stsn(ad, la, 1024, m.dj * jy + m.da * (eypv + 2.f * veyp + ey));
        stsn(ad, la, 1280, m.dj * jz + m.da * (ezpv + 2.f * vezp + ez));
    }
    {
        const float ev[3] = { ex, ey, ez }, epv[3] = { expv, eypv, ezpv }, evp[3] = { vexp, veyp, vezp };
#pragma unroll
        for (int q = 0; q < NP; q++)
            if (q < npl) {
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float nj = m.pj[q] * jj[q][c] + m.ps[q] * pp[q][c] + m.b2[q] * epv[c]
                                     + m.b3[q] * evp[c] + m.b4[q] * ev[c];
                    stsn(ap, lp, (3 * q + c) * 256, pp[q][c] + nj);
                    stsn(ap, lp, (3 * NP + 3 * q + c) * 256, nj);
                }
            }
    }
    stsn(ad, la, 0, vexp);
    stsn(ad, la, 256, veyp);
    stsn(ad, la, 512, vezp);
}

/* Recursive-convolution branch of one cell (src3d/yee.c:620-905): types 2
 * (Drude), 4 (Drude + 2 critical points) and 6 (piecewise-linear, with its
 * own inline CPML); coefficient forms in GsvRc.  Like the ADE branch it is
 * taken in dispersive cells and their +1 neighbours, updates the components
 * whose flag is clear and advances the accumulators of all three components.
 * State: the Drude accumulator of component c at slot c of the Drude block,
 * pole q's complex accumulator at slots 3q + c (real) and 6 + 3q + c
 * (imaginary) of the pole block. */
template <bool PER, typename OFF>
__device__ __forceinline__ void
rc_cell(const GsvGrid &g, float *ad, OFF la, float *ap, OFF lp, int i, int j, int k, unsigned fl, float &ex,
        float &ey, float &ez, float hx, float hy, float hz, float hxkm1, float hykm1, float hzi, float hyi, float hzj,
        float hxj, float rkx, float rky, float rkz)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz, px = g.px;
    const GSV_GLOBAL GsvRc &m = gp(g.rcm)[(fl >> GSV_E_SLOT_SHIFT) & 0xF];
    const bool tx = fl & GSV_E_TBCX, ty = fl & GSV_E_TBCY, tz = fl & GSV_E_TBCZ;
    const bool xi = i > 0 || (PER && g.perx), yi = j > 0 || (PER && g.pery);
    const bool plrc = m.kind == SV_MAT_PLRC, pre = m.kind == SV_MAT_CP, pl = m.poles != 0;
    const float e0[3] = { ex, ey, ez };
    float d[3], pr[2][3] = {}, pi[2][3] = {};
#pragma unroll
    for (int c = 0; c < 3; c++)
        d[c] = ldsn(ad, la, c * 256);
    if (pl) {
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                pr[q][c] = ldsn(ap, lp, (3 * q + c) * 256);
                pi[q][c] = ldsn(ap, lp, (6 + 3 * q + c) * 256);
            }
    }
    /* d' = dd d + d0 a + d1 b, p' = rot p + q0 a + q1 b */
    auto advance = [&](const float *a, const float *b) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            d[c] = m.dd * d[c] + m.d0 * a[c] + m.d1 * b[c];
            if (pl) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    const float r = pr[q][c], im = pi[q][c];
                    pr[q][c] = m.rr[q] * r - m.ri[q] * im + m.q0r[q] * a[c] + m.q1r[q] * b[c];
                    pi[q][c] = m.rr[q] * im + m.ri[q] * r + m.q0i[q] * a[c] + m.q1i[q] * b[c];
                }
            }
        }
    };
    if (pre)
        advance(e0, e0);   /* q1 = d1 = 0 for this type */
    /* only type 6 divides the H differences by kappa */
    const float kx = plrc ? rkx : 1.f, ky = plrc ? rky : 1.f, kz = plrc ? rkz : 1.f;
    if (yi && k > 0 && !tx)
        ex = m.ca * ex + m.cs * (d[0] + pr[0][0] + pr[1][0]) + m.cb[0] * ((hz - hzj) * ky - (hy - hykm1) * kz);
    if (xi && k > 0 && !ty)
        ey = m.ca * ey + m.cs * (d[1] + pr[0][1] + pr[1][1]) + m.cb[1] * ((hx - hxkm1) * kz - (hz - hzi) * kx);
    if (xi && yi && !tz)
        ez = m.ca * ez + m.cs * (d[2] + pr[0][2] + pr[1][2]) + m.cb[2] * ((hy - hyi) * kx - (hx - hxj) * ky);

    if (plrc) {
        /* the type-6 branch's own CPML update (src3d/yee.c:752-856): psi of the
         * slab cell advanced here and again by the slab code, the material's cb
         * on all six faces (x0 included) */
        const GsvCpmlFace &F0 = g.f[0], &F1 = g.f[1], &F2 = g.f[2], &F3 = g.f[3], &F4 = g.f[4], &F5 = g.f[5];
        if (F0.on && i > 1 && i < F0.depth) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * F0.depth + i) * 8;
            const float4 bc = ldg4(F0.ce + i);
            const float2 ps = ld2(F0.pe, p);
            const float a1 = bc.x * ps.x + bc.z * (hz - hzi) / g.dy;
            const float a2 = bc.x * ps.y + bc.z * (hy - hyi) / g.dz;
            st2(F0.pe, p, a1, a2);
            if ((yi && k != 0) && !ty) ey -= g.dy * m.cbm[1] * a1;
            if ((yi && k != 0) && !tz) ez += g.dz * m.cbm[2] * a2;
        }
        if (F2.on && j > 1 && j < F2.depth) {
            const OFF p = (OFF)(((unsigned)k * F2.depth + j) * px + i) * 8;
            const float4 bc = ldg4(F2.ce + j);
            const float2 ps = ld2(F2.pe, p);
            const float a1 = bc.x * ps.x + bc.z * (hz - hzj) / g.dx;
            const float a2 = bc.x * ps.y + bc.z * (hx - hxj) / g.dz;
            st2(F2.pe, p, a1, a2);
            if ((xi && k != 0) && !tx) ex += g.dx * m.cbm[0] * a1;
            if ((xi && k != 0) && !tz) ez -= g.dz * m.cbm[2] * a2;
        }
        if (F4.on && k > 1 && k < F4.depth) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * px + i) * 8;
            const float4 bc = ldg4(F4.ce + k);
            const float2 ps = ld2(F4.pe, p);
            const float a1 = bc.x * ps.x + bc.z * (hx - hxkm1) / g.dy;
            const float a2 = bc.x * ps.y + bc.z * (hy - hykm1) / g.dx;
            st2(F4.pe, p, a1, a2);
            if ((xi && yi) && !ty) ey += g.dy * m.cbm[1] * a1;
            if ((xi && yi) && !tx) ex -= g.dx * m.cbm[0] * a2;
        }
        if (F1.on && i >= nx - F1.depth + 1) {
            const int ir = i - (nx - F1.depth);
            const OFF p = (OFF)(((unsigned)k * ny + j) * F1.depth + ir) * 8;
            const float4 bc = ldg4(F1.ce + ir);
            const float2 ps = ld2(F1.pe, p);
            const float a1 = bc.x * ps.x + bc.z * (hz - hzi) / g.dy;
            const float a2 = bc.x * ps.y + bc.z * (hy - hyi) / g.dz;
            st2(F1.pe, p, a1, a2);
            if ((yi && k != 0) && !ty) ey -= g.dy * m.cbm[1] * a1;
            if ((yi && k != 0) && !tz) ez += g.dz * m.cbm[2] * a2;
        }
        if (F3.on && j >= ny - F3.depth + 1) {
            const int jr = j - (ny - F3.depth);
            const OFF p = (OFF)(((unsigned)k * F3.depth + jr) * px + i) * 8;
            const float4 bc = ldg4(F3.ce + jr);
            const float2 ps = ld2(F3.pe, p);
            const float a1 = bc.x * ps.x + bc.z * (hz - hzj) / g.dx;
            const float a2 = bc.x * ps.y + bc.z * (hx - hxj) / g.dz;
            st2(F3.pe, p, a1, a2);
            if ((xi && k != 0) && !tx) ex += g.dx * m.cbm[0] * a1;
            if ((xi && k != 0) && !tz) ez -= g.dz * m.cbm[2] * a2;
        }
        if (F5.on && k >= nz - F5.depth + 1) {
            const int kr = k - (nz - F5.depth);
            const OFF p = (OFF)(((unsigned)kr * ny + j) * px + i) * 8;
            const float4 bc = ldg4(F5.ce + kr);
            const float2 ps = ld2(F5.pe, p);
            const float a1 = bc.x * ps.x + bc.z * (hx - hxkm1) / g.dy;
            const float a2 = bc.x * ps.y + bc.z * (hy - hykm1) / g.dx;
            st2(F5.pe, p, a1, a2);
            if ((xi && yi) && !ty) ey += g.dy * m.cbm[1] * a1;
            if ((xi && yi) && !tx) ex -= g.dx * m.cbm[0] * a2;
        }
    }
    if (!pre) {
        const float e1[3] = { ex, ey, ez };
        advance(e0, e1);
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
        stsn(ad, la, c * 256, d[c]);
    if (pl) {
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                stsn(ap, lp, (3 * q + c) * 256, pr[q][c]);
                stsn(ap, lp, (6 + 3 * q + c) * 256, pi[q][c]);
            }
    }
}

template <typename ID, int ADE, typename OFF, bool PC, bool PER>
__device__ __forceinline__ void
estep_col(const GsvGrid &g, int i, int j, int kz, int kchunk)
{
    if (i >= g.nx || j >= g.ny)
        return;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int k0 = kz * kchunk;
    const int k1 = min(k0 + kchunk, nz);
    const int px = g.px;
    const OFF dP = (OFF)g.plane * 4, dJ = (OFF)px * 4;
    const OFF dPc = (OFF)g.plane * sizeof(ID);
    OFF o = ((OFF)i + (OFF)px * j + (OFF)g.plane * (k0 + 1)) * 4;
    OFF oc = ((OFF)i + (OFF)px * j + (OFF)g.plane * k0) * sizeof(ID);

    /* lane offsets in the blocked ADE state planes, [j][i/64][state][64] */
    const OFF la = ADE ? ((OFF)j * g.pxa * GSV_ADE_D + (OFF)(i >> 6) * 64 * GSV_ADE_D + (i & 63)) * 4 : 0;
    /* critical-point poles: ADE 2 -> 2, ADE 3 -> 3; ADE 4 (recursive convolution) -> 2 */
    constexpr int NP = ADE == 4 ? 2 : ADE >= 2 ? ADE : 0;
    const OFF lp = NP ? ((OFF)j * g.pxa * GSV_ADE_P(NP) + (OFF)(i >> 6) * 64 * GSV_ADE_P(NP) + (i & 63)) * 4 : 0;

    const float rkx = gp(g.rkex)[i], rky = gp(g.rkey)[j];
    const GsvCpmlFace &F0 = g.f[0], &F1 = g.f[1], &F2 = g.f[2], &F3 = g.f[3], &F4 = g.f[4], &F5 = g.f[5];
    const bool sx0 = F0.on && i >= 1 && i < F0.depth;
    const int irn = i - (nx - F1.depth);
    const bool sxn = F1.on && i >= nx - F1.depth + 1;
    const bool sy0 = F2.on && j >= 1 && j < F2.depth;
    const int jrn = j - (ny - F3.depth);
    const bool syn = F3.on && j >= ny - F3.depth + 1;
    const bool zplane_xy = (F0.on && i == 0) || (F2.on && j == 0);
    /* a periodic axis has no near plane: the -x / -y neighbour of the near plane is the far plane */
    const bool perx = PER && g.perx, pery = PER && g.pery;
    const bool xi = i > 0 || perx, yi = j > 0 || pery;
    const OFF dIm = (perx && i == 0) ? (OFF)(nx - 1) * 4 : (OFF)0 - (OFF)4;
    const OFF dJm = (pery && j == 0) ? (OFF)(ny - 1) * dJ : (OFF)0 - dJ;
    const bool cons = ADE && g.ade_cpml;

    constexpr bool QF = ADE == 1 && !PC;   /* Drude-only state form (engine.h) */
    float hxkm1 = ldf(g.hxg, o - dP), hykm1 = ldf(g.hyg, o - dP);
    /* the class id is loaded one plane ahead, so that the coefficient gather
     * and (in a dispersive class: ids from ade_id0 up) the state loads start
     * with the plane's field loads instead of after them */
    unsigned idn = ldid<ID>(g.ecls, oc);
    for (int k = k0; k < k1; ++k, o += dP, oc += dPc) {
        const unsigned id = idn;
        const float hx = ldn(g.hxg, o), hy = ldn(g.hyg, o), hz = ldn(g.hzg, o);
        const float hzj = ldf(g.hzg, o + dJm), hxj = ldf(g.hxg, o + dJm);
        const float hzi = ldf(g.hzg, o + dIm), hyi = ldf(g.hyg, o + dIm);
        float ex = ldn(g.exg, o), ey = ldn(g.eyg, o), ez = ldn(g.ezg, o);
        GsvECoef ec = ldec(g.etab + id);
        const bool dis = ADE != 0 && ADE != 4 && id >= (unsigned)g.ade_id0;
        float *adk = ADE ? g.ade_d + (size_t)k * g.ade_plane_d : nullptr;
        AdePre pre = {};
        if (dis)
            pre = ade_preload<NP, QF>(adk, la);
        if (k + 1 < k1)
            idn = ldid<ID>(g.ecls, oc + dPc);
        const float rkz = gp(g.rkez)[k];
        if (PC && (ec.flags & GSV_E_PERCELL))
            ec = ldec(g.pctab + ((i - g.pc_i0) + g.pc_nx * ((j - g.pc_j0) + g.pc_ny * (k - g.pc_k0))));
        const unsigned fl = ec.flags;
        const bool tx = fl & GSV_E_TBCX, ty = fl & GSV_E_TBCY, tz = fl & GSV_E_TBCZ;

        if (ADE == 4 && (fl & GSV_E_ADE))
            rc_cell<PER>(g, adk, la, g.ade_p ? g.ade_p + (size_t)k * g.ade_plane_p : nullptr, lp, i, j, k, fl, ex, ey,
                         ez, hx, hy, hz, hxkm1, hykm1, hzi, hyi, hzj, hxj, rkx, rky, rkz);
        else if (dis)
            ade_cell<NP, PER, QF>(g, adk, la, NP ? g.ade_p + (size_t)k * g.ade_plane_p : nullptr, lp, pre, i, j, k, fl,
                                  ex, ey, ez, hx, hy, hz, hxkm1, hykm1, hzi, hyi, hzj, hxj, rkx, rky, rkz);

        if (tx && yi && k > 0)
            ex = ec.ca[0] * ex + ec.cb[0] * ((hz - hzj) * rky - (hy - hykm1) * rkz);
        if (ty && xi && k > 0)
            ey = ec.ca[1] * ey + ec.cb[1] * ((hx - hxkm1) * rkz - (hz - hzi) * rkx);
        if (tz && xi && yi)
            ez = ec.ca[2] * ez + ec.cb[2] * ((hy - hyi) * rkx - (hx - hxj) * rky);

        if (fl & (GSV_E_PECX | GSV_E_PECY | GSV_E_PECZ)) {
            if (fl & GSV_E_PECX) ex = 0.f;
            if (fl & GSV_E_PECY) ey = 0.f;
            if (fl & GSV_E_PECZ) ez = 0.f;
        }

        /* CPML E psi and correction, src3d/boundary.c:691-887; under
         * ADE_CPML_CONSISTENT the psi of dispersive components (flag clear)
         * were advanced by the ADE branch */
        if (sx0) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * F0.depth + i) * 8;
            const float4 bc = ldg4(F0.ce + i);
            const float2 ps = ld2(F0.pe, p);
            const float a1 = (cons && !ty) ? ps.x : bc.x * ps.x + bc.y * (hz - hzi);
            const float a2 = (cons && !tz) ? ps.y : bc.x * ps.y + bc.y * (hy - hyi);
            st2(F0.pe, p, a1, a2);
            if (yi && k != 0) {
                if (ty) ey -= g.dx * ec.cb[1] * a1;
                if (tz) ez += g.dx * ec.cb[2] * a2;
            }
        }
        if (sy0) {
            const OFF p = (OFF)(((unsigned)k * F2.depth + j) * px + i) * 8;
            const float4 bc = ldg4(F2.ce + j);
            const float2 ps = ld2(F2.pe, p);
            const float a1 = (cons && !tx) ? ps.x : bc.x * ps.x + bc.y * (hz - hzj);
            const float a2 = (cons && !tz) ? ps.y : bc.x * ps.y + bc.y * (hx - hxj);
            st2(F2.pe, p, a1, a2);
            if (xi && k != 0) {
                if (tx) ex += g.dy * ec.cb[0] * a1;
                if (tz) ez -= g.dy * ec.cb[2] * a2;
            }
        }
        if (F4.on && k >= 1 && k < F4.depth) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * px + i) * 8;
            const float4 bc = ldg4(F4.ce + k);
            const float2 ps = ld2(F4.pe, p);
            const float a1 = (cons && !ty) ? ps.x : bc.x * ps.x + bc.y * (hx - hxkm1);
            const float a2 = (cons && !tx) ? ps.y : bc.x * ps.y + bc.y * (hy - hykm1);
            st2(F4.pe, p, a1, a2);
            if (xi && yi) {
                if (ty) ey += g.dx * ec.cb[1] * a1;
                /* the reference gates this Ex correction on the z flag
                 * (src3d/boundary.c:800-801); the x flag under ADE_CPML_CONSISTENT */
                if (cons ? tx : tz) ex -= g.dx * ec.cb[0] * a2;
            }
        }
        if (sxn) {
            const OFF p = (OFF)(((unsigned)k * ny + j) * F1.depth + irn) * 8;
            const float4 bc = ldg4(F1.ce + irn);
            const float2 ps = ld2(F1.pe, p);
            const float a1 = (cons && !ty) ? ps.x : bc.x * ps.x + bc.y * (hz - hzi);
            const float a2 = (cons && !tz) ? ps.y : bc.x * ps.y + bc.y * (hy - hyi);
            st2(F1.pe, p, a1, a2);
            if (yi && k != 0) {
                if (ty) ey -= g.dx * ec.cb[1] * a1;
                if (tz) ez += g.dx * ec.cb[2] * a2;
            }
        }
        if (syn) {
            const OFF p = (OFF)(((unsigned)k * F3.depth + jrn) * px + i) * 8;
            const float4 bc = ldg4(F3.ce + jrn);
            const float2 ps = ld2(F3.pe, p);
            const float a1 = (cons && !tx) ? ps.x : bc.x * ps.x + bc.y * (hz - hzj);
            const float a2 = (cons && !tz) ? ps.y : bc.x * ps.y + bc.y * (hx - hxj);
            st2(F3.pe, p, a1, a2);
            if (xi && k != 0) {
                if (tx) ex += g.dy * ec.cb[0] * a1;
                if (tz) ez -= g.dy * ec.cb[2] * a2;
            }
        }
        if (F5.on && k >= nz - F5.depth + 1) {
            const int kr = k - (nz - F5.depth);
            const OFF p = (OFF)(((unsigned)kr * ny + j) * px + i) * 8;
            const float4 bc = ldg4(F5.ce + kr);
            const float2 ps = ld2(F5.pe, p);
            const float a1 = (cons && !ty) ? ps.x : bc.x * ps.x + bc.y * (hx - hxkm1);
            const float a2 = (cons && !tx) ? ps.y : bc.x * ps.y + bc.y * (hy - hykm1);
            st2(F5.pe, p, a1, a2);
            if (xi && yi) {
                if (ty) ey += g.dx * ec.cb[1] * a1;
                if (tx) ex -= g.dx * ec.cb[0] * a2;
            }
        }
        if (zplane_xy || (F4.on && k == 0)) {
            ex = 0.f;
            ey = 0.f;
            ez = 0.f;
        }
        stn(g.exg, o, ex);
        stn(g.eyg, o, ey);
        stn(g.ezg, o, ez);
        hxkm1 = hx;
        hykm1 = hy;
    }
}

template <typename ID, int ADE, typename OFF, bool PC, bool PER>
__global__ __launch_bounds__(512) void
k_estep(GsvGrid g, int kchunk)
{
    estep_col<ID, ADE, OFF, PC, PER>(g, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y,
                                blockIdx.z, kchunk);
}

/* ------------------------------------------------------------------ */
/* TSF plane wave                                                       */
/* ------------------------------------------------------------------ */

__device__ __forceinline__ double
line_val(const double *data, int res, double x)
{
    /* cubic interpolation of the 1-D line, src3d/dline.c:89-112; samples outside the line are 0 */
    const int l = (int)floor(x);
    const double a = x - (double)l;
#define DV(q) (((q) >= 0 && (q) < res) ? gp(data)[q] : 0.0)
    if (x >= 1 && x < (res - 1)) {
        double w1 = a + 1, w2 = a, w3 = 1 - a, w4 = 2 - a;
        w1 = 4 - 8 * w1 + 5 * w1 * w1 - w1 * w1 * w1;
        w2 = 1 - 2 * w2 * w2 + w2 * w2 * w2;
        w3 = 1 - 2 * w3 * w3 + w3 * w3 * w3;
        w4 = 4 - 8 * w4 + 5 * w4 * w4 - w4 * w4 * w4;
        return w1 * DV(l - 1) + w2 * DV(l) + w3 * DV(l + 1) + w4 * DV(l + 2);
    } else if ((x < 1 && x >= 0) || (x >= (res - 1) && x < res))
        return (1 - a) * DV(l) + a * DV(l + 1);
#undef DV
    return 0;
}

__device__ __forceinline__ double
tsf_d(const GsvTsf &t, int i, int j, int k)
{
    return t.corr * (10.0 + t.ax * (double)(i - t.oi) + t.ay * (double)(j - t.oj) + t.az * (double)(k - t.ok));
}

/* one step of the 1-D line (double precision), src3d/1dpool.c:83-162; one
 * block, `step` is read by thread 0 only */
__device__ __forceinline__ void
tsf_line_step(const GsvTsf &t, int step)
{
    const int res = t.res;
    GSV_GLOBAL double *e = gp(t.je), *h = gp(t.jh), *b = gp(t.jb);
    for (int q = threadIdx.x + 1; q < res; q += blockDim.x)
        e[q] += t.multe * (h[q - 1] - h[q]);
    __syncthreads();
    if (threadIdx.x == 0) {
        e[0] = b[1] + t.coef * (e[1] - b[0]);
        h[0] = b[3] + t.coef * (h[1] - b[2]);
        e[res - 1] = b[5] + t.coef * (e[res - 2] - b[4]);
        h[res - 1] = b[7] + t.coef * (h[res - 2] - b[6]);
        b[0] = e[0]; b[1] = e[1]; b[2] = h[0]; b[3] = h[1];
        b[4] = e[res - 1]; b[5] = e[res - 2]; b[6] = h[res - 1]; b[7] = h[res - 2];
        e[0] = (step < t.ndata) ? gp(t.wave)[step] : 0.0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < res - 1; q += blockDim.x)
        h[q] += t.multh * (e[q] - e[q + 1]);
}

__global__ __launch_bounds__(1024) void
k_tsf_line(GsvTsf t, int step)
{
    tsf_line_step(t, step);
}

/* focused TSF: every plane wave's line in one launch, one workgroup each */
__global__ __launch_bounds__(256) void
k_tsff_line(const GsvTsf *__restrict__ lines, int step)
{
    const GsvTsf t = lines[blockIdx.x];
    tsf_line_step(t, step);
}

/* focused TSF faces normal to AXIS (both sides): per face cell the plain TSF
 * correction of every wave, summed in double and added once.  With A, B the
 * in-plane axes in cyclic order after AXIS and p the wave's polarisation
 * factors, F_A += s p_B v and F_B -= s p_A v, s = +1 on the near E face and
 * the far H face, -1 on the others (src3d/source.c:800-915, 1921-2037); E
 * faces read the H line one cell outside the near face, H faces the E line
 * on the face plane, and write the H cells one outside the near face. */
template <int AXIS>
__global__ __launch_bounds__(256) void
k_tsff_face(GsvGrid g, GsvTsff f, int hside)
{
    constexpr int A = (AXIS + 1) % 3, B = (AXIS + 2) % 3;
    const int lo[3] = { f.i0, f.j0, f.k0 }, hi[3] = { f.in, f.jn, f.kn };
    const int na = hi[A] - lo[A] + 1, nb = hi[B] - lo[B] + 1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 2 * na * nb)
        return;
    const int side = q / (na * nb), r = q - side * na * nb, a = r % na, b = r / na;
    const int face = 2 * AXIS + side;
    if (f.skip[face])
        return;
    int c[3], e[3];
    c[A] = e[A] = lo[A] + a;
    c[B] = e[B] = lo[B] + b;
    const int plane = side ? hi[AXIS] : lo[AXIS];
    e[AXIS] = (!side && !hside) ? plane - 1 : plane;     /* where the line is read */
    c[AXIS] = (!side && hside) ? plane - 1 : plane;      /* the cell corrected */
    const double *line = hside ? f.je : f.jh;
    double sa = 0.0, sb = 0.0;
    for (int w = 0; w < f.nw; w++) {
        const GsvTsffWave &W = f.waves[w];
        const double d = W.corr * (250.0 - (W.ax * (W.cx - e[0]) + W.ay * (W.cy - e[1]) + W.az * (W.cz - e[2])));
        const double v = line_val(line + (size_t)w * f.res, f.res, d);
        const double *p = hside ? W.ph : W.pe;
        sa += p[B] * v;
        sb += p[A] * v;
    }
    const double s = ((side != 0) != (hside != 0)) ? -1.0 : 1.0;
    const double coef = hside ? (double)f.hcor : (double)gp(f.coefE[face])[a + na * b];
    const long long o = c[0] + (long long)g.px * c[1] + g.plane * c[2];
    float *fa, *fb;
    if (hside) {
        fa = A == 0 ? g.hx : A == 1 ? g.hy : g.hz;
        fb = B == 0 ? g.hx : B == 1 ? g.hy : g.hz;
    } else {
        fa = A == 0 ? g.ex : A == 1 ? g.ey : g.ez;
        fb = B == 0 ? g.ex : B == 1 ? g.ey : g.ez;
    }
    /* an E component exists below the box end along its own axis, an H
     * component below the box end along the other in-plane axis */
    const bool wa = hside ? c[B] < hi[B] : c[A] < hi[A];
    const bool wb = hside ? c[A] < hi[A] : c[B] < hi[B];
    if (wa)
        gp(fa)[o] += (float)(s * coef * sa);
    if (wb)
        gp(fb)[o] -= (float)(s * coef * sb);
}

/* external plane source (src3d/source.c:1377-1394, 2247-2259): E on plane k
 * from the stored H plane values, H on plane k - 1 from the stored E values */
__global__ __launch_bounds__(256) void
k_ext_plane(GsvGrid g, GsvExt x, int hside)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= x.ni * x.nj)
        return;
    const int i = x.ijstart + q % x.ni, j = x.jkstart + q / x.ni;
    const float vx = gp(x.px)[q], vy = gp(x.py)[q];
    if (hside) {
        const long long o = i + (long long)g.px * j + g.plane * (x.k - 1);
        gp(g.hx)[o] -= x.hcor * vy;
        gp(g.hy)[o] += x.hcor * vx;
    } else {
        const long long o = i + (long long)g.px * j + g.plane * x.k;
        const float c = gp(x.ce)[q];
        gp(g.ex)[o] += c * vy;
        gp(g.ey)[o] -= c * vx;
    }
}

/* scattered-field source on the PEC rule's components: minus the incident
 * field read from the SF line at the cell's distance along the wave
 * (src3d/yee.c:1146-1230; every component of a cell reads the same sample) */
__global__ __launch_bounds__(256) void
k_sf_pec(GsvGrid g, GsvTsf t, const int4 *cells, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    const int4 c = cells[q];
    const double ve = line_val(t.je, t.res, tsf_d(t, c.x, c.y, c.z));
    const long long o = c.x + (long long)g.px * c.y + g.plane * c.z;
    if (c.w & 1) gp(g.ex)[o] = (float)(-t.gex * ve);
    if (c.w & 2) gp(g.ey)[o] = (float)(-t.gey * ve);
    if (c.w & 4) gp(g.ez)[o] = (float)(-t.gez * ve);
}

/* H correction of face cell (a, b) of the face normal to `axis` (far: 0 near, 1 far) */
__device__ __forceinline__ void
tsf_h_cell(const GsvGrid &g, const GsvTsf &t, int axis, int far, int a, int b)
{
    int na, nb;
    tsf_face_extent(t, g.nx, g.ny, axis, &na, &nb);
    if (a >= na || b >= nb)
        return;
    const float cf = gp(t.coefH[axis * 2 + far])[a + (long long)na * b];
    if (cf == 0.f)
        return;
    const long long P = g.plane;
    const int px = g.px;
    if (axis == 0) {
        const int j = t.j0 + a, k = t.k0 + b;
        const int ip = far ? t.in : t.i0;
        const double ve = line_val(t.je, t.res, tsf_d(t, ip, j, k));
        const int iw = far ? t.in : t.i0 - 1;
        const long long c = iw + (long long)px * j + P * k;
        const float s = far ? 1.f : -1.f;
        if (k < t.kn) gp(g.hy)[c] += s * (float)(cf * t.gez * ve);
        if (j < t.jn) gp(g.hz)[c] -= s * (float)(cf * t.gey * ve);
    } else if (axis == 1) {
        const int i = t.i0 + a, k = t.k0 + b;
        const int jp = far ? t.jn : t.j0;
        const double ve = line_val(t.je, t.res, tsf_d(t, i, jp, k));
        const int jw = far ? t.jn : t.j0 - 1;
        const long long c = i + (long long)px * jw + P * k;
        const float s = far ? 1.f : -1.f;
        if (k < t.kn) gp(g.hx)[c] -= s * (float)(cf * t.gez * ve);
        if (i < t.in) gp(g.hz)[c] += s * (float)(cf * t.gex * ve);
    } else {
        const int i = t.i0 + a, j = t.j0 + b;
        const int kp = far ? t.kn : t.k0;
        const double ve = line_val(t.je, t.res, tsf_d(t, i, j, kp));
        const int kw = far ? t.kn : t.k0 - 1;
        const long long c = i + (long long)px * j + P * kw;
        const float s = far ? 1.f : -1.f;
        if (j < t.jn) gp(g.hx)[c] += s * (float)(cf * t.gey * ve);
        if (i < t.in) gp(g.hy)[c] -= s * (float)(cf * t.gex * ve);
    }
}

/* E correction of face cell (a, b) of the face normal to `axis` */
__device__ __forceinline__ void
tsf_e_cell(const GsvGrid &g, const GsvTsf &t, int axis, int far, int a, int b)
{
    int na, nb;
    tsf_face_extent(t, g.nx, g.ny, axis, &na, &nb);
    if (a >= na || b >= nb)
        return;
    const float cf = gp(t.coefE[axis * 2 + far])[a + (long long)na * b];
    if (cf == 0.f)
        return;
    const long long P = g.plane;
    const int px = g.px;
    const float s = far ? -1.f : 1.f;
    if (axis == 0) {
        const int j = t.j0 + a, k = t.k0 + b;
        const int ip = far ? t.in : t.i0 - 1;
        const double vh = line_val(t.jh, t.res, tsf_d(t, ip, j, k));
        const long long c = (far ? t.in : t.i0) + (long long)px * j + P * k;
        if (j < t.jn) gp(g.ey)[c] += s * (float)(cf * t.ghz * vh);
        if (k < t.kn) gp(g.ez)[c] -= s * (float)(cf * t.ghy * vh);
    } else if (axis == 1) {
        const int i = t.i0 + a, k = t.k0 + b;
        const int jp = far ? t.jn : t.j0 - 1;
        const double vh = line_val(t.jh, t.res, tsf_d(t, i, jp, k));
        const long long c = i + (long long)px * (far ? t.jn : t.j0) + P * k;
        if (i < t.in) gp(g.ex)[c] -= s * (float)(cf * t.ghz * vh);
        if (k < t.kn) gp(g.ez)[c] += s * (float)(cf * t.ghx * vh);
    } else {
        const int i = t.i0 + a, j = t.j0 + b;
        const int kp = far ? t.kn : t.k0 - 1;
        const double vh = line_val(t.jh, t.res, tsf_d(t, i, j, kp));
        const long long c = i + (long long)px * j + P * (far ? t.kn : t.k0);
        if (i < t.in) gp(g.ex)[c] += s * (float)(cf * t.ghy * vh);
        if (j < t.jn) gp(g.ey)[c] -= s * (float)(cf * t.ghx * vh);
    }
}

/* H corrections on the two faces normal to `axis` (blockIdx.z: 0 near, 1 far) */
__global__ __launch_bounds__(256) void
k_tsf_h(GsvGrid g, GsvTsf t, int axis)
{
    tsf_h_cell(g, t, axis, blockIdx.z, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y);
}

/* E corrections on the two faces normal to `axis` */
__global__ __launch_bounds__(256) void
k_tsf_e(GsvGrid g, GsvTsf t, int axis)
{
    tsf_e_cell(g, t, axis, blockIdx.z, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y);
}

/* ------------------------------------------------------------------ */
/* point sources, traces, boundaries                                    */
/* ------------------------------------------------------------------ */

__global__ void
k_point_src(GsvGrid g, GsvPointSrc s, int step, int hside)
{
    /* sources in order, one thread: later sources win on a shared cell */
    if (threadIdx.x != 0 || blockIdx.x != 0 || step >= s.nsteps)
        return;
    for (int n = 0; n < s.n; n++) {
        const long long c = s.pos[n];
        const float *v = s.vals + ((long long)step * s.n + n) * 6 + (hside ? 3 : 0);
        float *f0 = hside ? g.hx : g.ex, *f1 = hside ? g.hy : g.ey, *f2 = hside ? g.hz : g.ez;
        if (v[0] != 0.f) f0[c] = v[0];
        if (v[1] != 0.f) f1[c] = v[1];
        if (v[2] != 0.f) f2[c] = v[2];
    }
}

__global__ void
k_record_points(GsvGrid g, const long long *pos, int n, float *rec)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n)
        return;
    const long long c = pos[q];
    float *r = rec + (long long)q * 6;
    r[0] = g.ex[c];
    r[1] = g.ey[c];
    r[2] = g.ez[c];
    r[3] = g.hx[c];
    r[4] = g.hy[c];
    r[5] = g.hz[c];
}

__device__ __forceinline__ long long
face_cell(const GsvGrid &g, int axis, int plane_index, int a, int b)
{
    if (axis == 0) return plane_index + (long long)g.px * a + g.plane * b;
    if (axis == 1) return a + (long long)g.px * plane_index + g.plane * b;
    return a + (long long)g.px * b + g.plane * plane_index;
}

/* store the two outer planes of each Liao face, src3d/boundary.c:938-1066 */
__global__ void
k_liao_store(GsvGrid g, GsvLiao L, int axis)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    int na, nb, n;
    if (axis == 0) { na = g.ny; nb = g.nz; n = g.nx; }
    else if (axis == 1) { na = g.nx; nb = g.nz; n = g.ny; }
    else { na = g.nx; nb = g.ny; n = g.nz; }
    if (a >= na || b >= nb)
        return;
    float *F[6] = { g.ex, g.ey, g.ez, g.hx, g.hy, g.hz };
    float *st = L.store[axis];
    const long long pl = (long long)na * nb, s = (long long)a * nb + b;
    const long long c0 = face_cell(g, axis, 0, a, b), c1 = face_cell(g, axis, 1, a, b);
    const long long cn = face_cell(g, axis, n - 1, a, b), cn1 = face_cell(g, axis, n - 2, a, b);
    for (int q = 0; q < 6; q++) {
        if (L.on[axis * 2]) {
            st[(0 + q) * pl + s] = F[q][c0];
            st[(12 + q) * pl + s] = F[q][c1];
        }
        if (L.on[axis * 2 + 1]) {
            st[(6 + q) * pl + s] = F[q][cn];
            st[(18 + q) * pl + s] = F[q][cn1];
        }
    }
}

__global__ void
k_liao_face(GsvGrid g, GsvLiao L, int axis, int far)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    int na, nb, n;
    if (axis == 0) { na = g.ny; nb = g.nz; n = g.nx; }
    else if (axis == 1) { na = g.nx; nb = g.nz; n = g.ny; }
    else { na = g.nx; nb = g.ny; n = g.nz; }
    if (a >= na || b >= nb)
        return;
    float *F[6] = { g.ex, g.ey, g.ez, g.hx, g.hy, g.hz };
    const float *st = L.store[axis];
    const long long pl = (long long)na * nb, s = (long long)a * nb + b;
    const long long c = face_cell(g, axis, far ? n - 1 : 0, a, b);
    const long long ci = face_cell(g, axis, far ? n - 2 : 1, a, b);
    const float *cf = L.coef[axis * 2 + far] + s * 3;
    for (int q = 0; q < 6; q++) {
        const int p0 = far ? 18 + q : 12 + q, p00 = far ? 6 + q : q;
        F[q][c] = st[p0 * pl + s] + cf[q % 3] * (F[q][ci] - st[p00 * pl + s]);
    }
}

/* near-plane zeroing of CPML faces, re-applied after Liao faces on mixed boundaries */
__global__ void
k_zero_planes(GsvGrid g, int axis)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    int na, nb;
    if (axis == 0) { na = g.ny; nb = g.nz; }
    else if (axis == 1) { na = g.nx; nb = g.nz; }
    else { na = g.nx; nb = g.ny; }
    if (a >= na || b >= nb)
        return;
    const long long c = face_cell(g, axis, 0, a, b);
    g.ex[c] = 0.f;
    g.ey[c] = 0.f;
    g.ez[c] = 0.f;
}

__global__ void
k_periodic_h(GsvGrid g, GsvPeriodic p)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    const long long P = g.plane;
    const int px = g.px;
    if (p.x) {
        const int j = p.y0 + a, k = p.z0 + b;
        if (j < p.yn && k < p.zn) {
            const long long d = (p.x0 - 1) + (long long)px * j + P * k, s = (p.xn - 1) + (long long)px * j + P * k;
            g.hy[d] = g.hy[s];
            g.hz[d] = g.hz[s];
        }
    }
    if (p.y) {
        const int i = p.x0 + a, k = p.z0 + b;
        if (i < p.xn && k < p.zn) {
            const long long d = i + (long long)px * (p.y0 - 1) + P * k, s = i + (long long)px * (p.yn - 1) + P * k;
            g.hx[d] = g.hx[s];
            g.hz[d] = g.hz[s];
        }
    }
    if (p.z) {
        const int i = p.x0 + a, j = p.y0 + b;
        if (i < p.xn && j < p.yn) {
            const long long d = i + (long long)px * j + P * (p.z0 - 1), s = i + (long long)px * j + P * (p.zn - 1);
            g.hx[d] = g.hx[s];
            g.hy[d] = g.hy[s];
        }
    }
}

__global__ void
k_periodic_e(GsvGrid g, GsvPeriodic p)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y * blockDim.y + threadIdx.y;
    const long long P = g.plane;
    const int px = g.px;
    if (p.x) {
        const int j = p.y0 + a, k = p.z0 + b;
        if (j < p.yn && k < p.zn) {
            const long long d = p.xn + (long long)px * j + P * k, s = p.x0 + (long long)px * j + P * k;
            g.ey[d] = g.ey[s];
            g.ez[d] = g.ez[s];
        }
    }
    if (p.y) {
        const int i = p.x0 + a, k = p.z0 + b;
        if (i < p.xn && k < p.zn) {
            const long long d = i + (long long)px * p.yn + P * k, s = i + (long long)px * p.y0 + P * k;
            g.ex[d] = g.ex[s];
            g.ez[d] = g.ez[s];
        }
    }
    if (p.z) {
        const int i = p.x0 + a, j = p.y0 + b;
        if (i < p.xn && j < p.yn) {
            const long long d = i + (long long)px * j + P * p.zn, s = i + (long long)px * j + P * p.z0;
            g.ex[d] = g.ex[s];
            g.ey[d] = g.ey[s];
        }
    }
}

/* ------------------------------------------------------------------ */
/* far field                                                            */
/* ------------------------------------------------------------------ */

/* time-domain Kirchhoff accumulation: one thread per (point, surface cell) */
__global__ __launch_bounds__(256) void
k_farfield_td(GsvGrid g, GsvFarfieldTD f, int step)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (q >= f.ncells)
        return;
    int face, a, b;
    long long r = q;
    const long long sxy = (long long)f.ix * f.iy, sxz = (long long)f.ix * f.iz, syz = (long long)f.iy * f.iz;
    if (r < 2 * syz) { face = (r < syz) ? 2 : 3; r %= syz; a = (int)(r / f.iz); b = (int)(r % f.iz); }
    else if ((r -= 2 * syz) < 2 * sxz) { face = (r < sxz) ? 0 : 1; r %= sxz; a = (int)(r / f.iz); b = (int)(r % f.iz); }
    else { r -= 2 * sxz; face = (r < sxy) ? 4 : 5; r %= sxy; a = (int)(r / f.iy); b = (int)(r % f.iy); }
    int i, j, k, ni = 0, nj = 0, nk = 0;
    double size;
    /* faces: 2/3 x0/xn (a=j, b=k), 0/1 y0/yn (a=i, b=k), 4/5 z0/zn (a=i, b=j) */
    if (face == 2 || face == 3) {
        j = f.j0 + a; k = f.k0 + b; i = (face == 2) ? f.i0 : f.i1; ni = 1; size = f.dy * f.dz;
        const int *sk = f.skip[face == 2 ? 0 : 1];
        if (sk[0] && j >= sk[1] && k >= sk[2] && j <= sk[3] && k <= sk[4]) return;
    } else if (face == 0 || face == 1) {
        i = f.i0 + a; k = f.k0 + b; j = (face == 0) ? f.j0 : f.j1; nj = 1; size = f.dx * f.dz;
        const int *sk = f.skip[face == 0 ? 2 : 3];
        if (sk[0] && i >= sk[1] && k >= sk[2] && i <= sk[3] && k <= sk[4]) return;
    } else {
        i = f.i0 + a; j = f.j0 + b; k = (face == 4) ? f.k0 : f.k1; nk = 1; size = f.dx * f.dy;
        const int *sk = f.skip[face == 4 ? 4 : 5];
        if (sk[0] && i >= sk[1] && j >= sk[2] && i <= sk[3] && j <= sk[4]) return;
    }
    const int x = f.ri[n], y = f.rj[n], z = f.rk[n];
    const double ddx = x - i, ddy = y - j, ddz = z - k;
    const double rc = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    double nvx = 0, nvy = 0, nvz = 0;
    if (face == 0) nvy = -1; else if (face == 1) nvy = 1; else if (face == 2) nvx = -1;
    else if (face == 3) nvx = 1; else if (face == 4) nvz = -1; else nvz = 1;
    const double ct = (ddx * nvx + ddy * nvy + ddz * nvz) / rc;
    const double rr = rc * f.dx;
    const int nn = (int)(step + 1 + rr / (f.c * f.dt)) - f.istart[n];
    if (nn < 3 || nn >= f.ndata)
        return;
    const double A = 1.0 / (4.0 * M_PI * rr);
    const double B = -1.0 / (4.0 * M_PI) * ct / (rr * rr);
    const double C = -1.0 / (4.0 * M_PI) * ct / (f.c * rr);
    const double sgn = (face == 0 || face == 2 || face == 4) ? -1.0 : 1.0;
    const long long c = i + (long long)g.px * j + g.plane * k;
    const long long off = ni + (long long)g.px * nj + g.plane * nk;
    const float *FF[3] = { g.ex, g.ey, g.ez };
    double *bins = f.bins + (long long)n * 3 * f.ndata;
    for (int qq = 0; qq < 3; qq++) {
        const double v = FF[qq][c], vp = FF[qq][c + off], vm = FF[qq][c - off];
        double *bq = bins + (long long)qq * f.ndata;
        atomicAdd(bq + nn, -C / 2.0 / f.dt * v * size);
        atomicAdd(bq + nn - 1, (sgn * A * (vp - vm) / 2.0 / f.dx + B * v) * size);
        atomicAdd(bq + nn - 2, C / 2.0 / f.dt * v * size);
    }
}

/* layered TSF 1-D model, src3d/zpool.c:172-338 in the order of
 * src3d/source.c:1680-1689, one workgroup: every substep saves the end
 * values, updates E (vacuum line e and the split fields ee, eh, ez), injects
 * the line's h[9] into ee / eh at cell 10, applies the Mur ends to e, ee, eh
 * and the hard source at e[0], updates H, injects e[10] into he / hh at cell
 * 9 and stores the six split fields.  The line's reads beyond its ends (h[-1],
 * e[res]) are zero guard cells.  Serial end work runs on thread 0 between
 * barriers. */
__global__ __launch_bounds__(1024) void
k_ltsf_line(GsvLtsfLine L)
{
    extern __shared__ double lm[];
    const int res = L.res, n2 = res + 2, tid = threadIdx.x, nt = blockDim.x;
    double *e = lm + 1, *h = lm + n2 + 1;
    double *ee = lm + 2 * n2, *eh = ee + n2, *ez = eh + n2, *he = ez + n2, *hh = he + n2, *hz = hh + n2;
    for (int i = tid; i < 8 * n2; i += nt)
        lm[i] = 0.0;
    __syncthreads();
    for (int s = 0; s < L.nrows; s++) {
        double fe0 = 0, fe1 = 0, fen0 = 0, fen1 = 0, fee0 = 0, fee1 = 0, feen0 = 0, feen1 = 0;
        double feh0 = 0, feh1 = 0, fehn0 = 0, fehn1 = 0;
        if (tid == 0) {
            fe0 = e[0]; fe1 = e[1]; fen0 = e[res - 1]; fen1 = e[res - 2];
            fee0 = ee[0]; fee1 = ee[1]; feen0 = ee[res - 1]; feen1 = ee[res - 2];
            feh0 = eh[0]; feh1 = eh[1]; fehn0 = eh[res - 1]; fehn1 = eh[res - 2];
        }
        __syncthreads();
        for (int i = tid; i < res; i += nt) {
            e[i] += 1.0 * L.multe * (h[i - 1] - h[i]);
            if (i > 0) {
                const double ep = L.eps[i];
                ee[i] = ee[i] + 1.0 / ep * L.multe * (he[i - 1] - he[i]);
                eh[i] = eh[i] + 1.0 / (ep - L.st * L.st) * L.multe * (hh[i - 1] - hh[i]);
                ez[i] = -ez[i] - L.cez / ep * he[i];
            }
        }
        __syncthreads();
        if (tid == 0) {
            ee[10] -= L.inj_e_s * h[9];
            eh[10] -= L.inj_e_c * h[9];
            e[0] = fe1 + L.m0 * (e[1] - fe0);
            e[res - 1] = fen1 + L.m1 * (e[res - 2] - fen0);
            ee[0] = fee1 + L.m0 * (ee[1] - fee0);
            ee[res - 1] = feen1 + L.m1 * (ee[res - 2] - feen0);
            eh[0] = feh1 + L.m0 * (eh[1] - feh0);
            eh[res - 1] = fehn1 + L.m1 * (eh[res - 2] - fehn0);
            const int ip = s - 12;
            e[0] = (ip > 0 && ip < L.ndata - 1) ? L.wave[ip] : 0.0;
        }
        __syncthreads();
        for (int i = tid; i < res; i += nt) {
            h[i] += 1.0 * L.multh * (e[i] - e[i + 1]);
            if (i < res - 1) {
                const double ep = L.eps[i], m = L.mu[i];
                he[i] = he[i] + 1.0 / m * ep / (ep - L.st * L.st) * L.multh * (ee[i] - ee[i + 1]);
                hh[i] = hh[i] + 1.0 / m * L.multh * (eh[i] - eh[i + 1]);
                hz[i] = -hz[i] + L.chz * eh[i];
            }
        }
        __syncthreads();
        if (tid == 0) {
            he[9] -= L.inj_h_s * e[10];
            hh[9] -= L.inj_h_c * e[10];
        }
        __syncthreads();
        const double *F[6] = { ee, eh, ez, he, hh, hz };
        for (int q = 0; q < 6; q++) {
            double *dst = L.hist + ((long long)q * L.nrows + s) * res;
            for (int i = tid; i < res; i += nt)
                dst[i] = F[q][i];
        }
    }
}

/* the stored component q at z cell k for a face cell of delay qd (steps),
 * src3d/source.c:286-322: timepos = 3 (step - qd) substeps, tp truncated
 * toward zero then clamped at 0, linear between tp and tp + 1 */
__device__ __forceinline__ double
ltsf_at(const GsvLtsf &L, int q, int k, double tpos)
{
    int tp = (int)tpos;
    const double atp = tpos - (double)tp;
    if (tp < 0)
        tp = 0;
    const double *d = L.hist + (long long)q * L.nrows * L.res;
    return (1.0 - atp) * d[(long long)tp * L.res + k] + atp * d[(long long)(tp + 1) * L.res + k];
}

/* layered TSF faces of one axis, src3d/source.c:921-1150 (E, hside 0) and
 * 1700-1900 (H, hside 1): one thread per (face, a, b); E faces sample the
 * model's H outside the near faces (i0 - 1, j0 - 1, k0 - 1), H faces sample
 * its E at the face cells */
__global__ __launch_bounds__(256) void
k_ltsf_face(GsvGrid g, GsvLtsf L, int axis, int hside, int step)
{
    const int na = axis == 0 ? L.j1 - L.j0 + 1 : L.i1 - L.i0 + 1;
    const int nb = axis == 2 ? L.j1 - L.j0 + 1 : L.k1 - L.k0 + 1;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 2 * na * nb)
        return;
    const int f = q / (na * nb), a = (q / nb) % na, b = q % nb;
    if (L.skip[2 * axis + f])
        return;
    const double tpos = 3.0 * ((double)step - L.q[hside][axis][q]);
    const double cf = L.coef[hside][axis][q];
    const double sph = L.sph, cph = L.cph;
    const long long px = g.px, pl = g.plane;
    int i, j, k, ks;
    if (axis == 0) { i = f ? L.i1 : L.i0; j = L.j0 + a; k = L.k0 + b; ks = k; }
    else if (axis == 1) { i = L.i0 + a; j = f ? L.j1 : L.j0; k = L.k0 + b; ks = k; }
    else { i = L.i0 + a; j = L.j0 + b; k = f ? L.k1 : L.k0; ks = (f || hside) ? k : k - 1; }
    if (!hside) {
        const double vhh = ltsf_at(L, 4, ks, tpos), vhe = ltsf_at(L, 3, ks, tpos);
        const long long c = i + px * j + pl * k;
        const double sg = f ? -1.0 : 1.0;
        if (axis == 0) {
            const double hz = ltsf_at(L, 5, ks, tpos), hy = -(vhh * sph - vhe * cph);
            if (j < L.j1) g.ey[c] = (float)((double)g.ey[c] + sg * (cf * hz));
            if (k < L.k1) g.ez[c] = (float)((double)g.ez[c] - sg * (cf * hy));
        } else if (axis == 1) {
            const double hz = ltsf_at(L, 5, ks, tpos), hx = -(vhh * cph + vhe * sph);
            if (i < L.i1) g.ex[c] = (float)((double)g.ex[c] - sg * (cf * hz));
            if (k < L.k1) g.ez[c] = (float)((double)g.ez[c] + sg * (cf * hx));
        } else {
            const double hx = -(vhh * cph + vhe * sph), hy = -(vhh * sph - vhe * cph);
            if (i < L.i1) g.ex[c] = (float)((double)g.ex[c] + sg * (cf * hy));
            if (j < L.j1) g.ey[c] = (float)((double)g.ey[c] - sg * (cf * hx));
        }
    } else {
        const double veh = ltsf_at(L, 1, ks, tpos), vee = ltsf_at(L, 0, ks, tpos);
        const double sg = f ? -1.0 : 1.0;
        if (axis == 0) {
            const long long c = (f ? i : i - 1) + px * j + pl * k;
            const double ez = ltsf_at(L, 2, ks, tpos), ey = vee * sph + veh * cph;
            if (k < L.k1) g.hy[c] = (float)((double)g.hy[c] - sg * (cf * ez));
            if (j < L.j1) g.hz[c] = (float)((double)g.hz[c] + sg * (cf * ey));
        } else if (axis == 1) {
            const long long c = i + px * (f ? j : j - 1) + pl * k;
            const double ez = ltsf_at(L, 2, ks, tpos), ex = vee * cph - veh * sph;
            if (k < L.k1) g.hx[c] = (float)((double)g.hx[c] + sg * (cf * ez));
            if (i < L.i1) g.hz[c] = (float)((double)g.hz[c] - sg * (cf * ex));
        } else {
            const long long c = i + px * j + pl * (f ? k : k - 1);
            const double ex = vee * cph - veh * sph, ey = vee * sph + veh * cph;
            if (j < L.j1) g.hx[c] = (float)((double)g.hx[c] - sg * (cf * ey));
            if (i < L.i1) g.hy[c] = (float)((double)g.hy[c] + sg * (cf * ex));
        }
    }
}

/* stress tensor times the inward normal, summed per block (wave64
 * reductions, then LDS) and added to the step's three values */
__global__ __launch_bounds__(256) void
k_force(GsvGrid g, GsvForce f, int step)
{
    __shared__ double part[4][3];
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double a0 = 0, a1 = 0, a2 = 0;
    if (q < f.ncells) {
        int i, j, k, nx, ny, nz;
        force_cell_of(f, q, i, j, k, nx, ny, nz);
        const long long c = i + (long long)g.px * j + g.plane * k;
        const double ex = g.ex[c], ey = g.ey[c], ez = g.ez[c], hx = g.hx[c], hy = g.hy[c], hz = g.hz[c];
        const double eps = f.em[q].x, mu = f.em[q].y;
        const double val = (eps * (ex * ex + ey * ey + ez * ez) + mu * (hx * hx + hy * hy + hz * hz)) / 2.0;
        const double t00 = eps * ex * ex + mu * hx * hx - val, t11 = eps * ey * ey + mu * hy * hy - val;
        const double t22 = eps * ez * ez + mu * hz * hz - val;
        const double t01 = eps * ex * ey + mu * hx * hy, t02 = eps * ex * ez + mu * hx * hz, t12 = eps * ey * ez + mu * hy * hz;
        a0 = t00 * nx + t01 * ny + t02 * nz;
        a1 = t01 * nx + t11 * ny + t12 * nz;
        a2 = t02 * nx + t12 * ny + t22 * nz;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_down(a0, o, 64);
        a1 += __shfl_down(a1, o, 64);
        a2 += __shfl_down(a2, o, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        part[w][0] = a0; part[w][1] = a1; part[w][2] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0;
        for (int v = 0; v < (int)(blockDim.x >> 6); v++)
            s += part[v][threadIdx.x];
        atomicAdd(f.out + (long long)step * 3 + threadIdx.x, s);
    }
}

/* periodic far field, src3d/pfarfield.c:120-401: one thread per (image,
 * plane, i, j) and point (blockIdx.y), i fastest so that a wave reads 64
 * consecutive cells.  The distance is formed in float with explicit
 * round-to-nearest operations (no contraction), as the reference's
 * ((float)ri - (i + pri)) squares are; everything after it in double. */
__global__ __launch_bounds__(256) void
k_pfarfield(GsvGrid g, GsvPFarfield f, int step)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (q >= f.nwork || (f.post && step < f.ppstart))
        return;
    const long long pl = (long long)f.ix * f.iy;
    const int i = f.i0 + (int)(q % f.ix);
    const int j = f.j0 + (int)((q / f.ix) % f.iy);
    const int s = (int)((q / pl) & 1);
    const long long img = q / (2 * pl);
    const int npj = f.pjmax - f.pjmin;
    const int pi = f.pimin + (int)(img / npj), pj = f.pjmin + (int)(img % npj);
    const int *sk = f.skip[s];
    if (sk[0] && i >= sk[1] && j >= sk[2] && i <= sk[3] && j <= sk[4])
        return;
    const int k = s ? f.k1 : f.k0;
    const int rk = f.rk[n];
    const float ddx = __fsub_rn((float)f.ri[n], (float)(i + f.ix * pi));
    const float ddy = __fsub_rn((float)f.rj[n], (float)(j + f.iy * pj));
    const float ddz = __fsub_rn((float)rk, (float)k);
    const float rs = __fadd_rn(__fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy)), __fmul_rn(ddz, ddz));
    double r = sqrt((double)rs);
    const double ct = (s ? (double)(-(f.k1 - rk)) : (double)(f.k0 - rk)) / r;
    r *= f.dx;
    const double mult = 1.0 / (4.0 * M_PI), size = f.dx * f.dy;
    int nn;
    double b, cvala, aval;
    if (f.post) {
        const double nred = f.ppstart + 1 + r / (f.c * f.dt);
        nn = (int)((step - f.ppstart) + nred) - f.istart[n];
        b = -mult * ct / (r * r) * size;
        aval = mult / r / 2.0 / f.dx * size;
    } else {
        nn = (int)(step + 1 + r / (f.c * f.dt)) - f.istart[n];
        b = -mult * ct / (r * r);
        aval = mult / r / 2.0 / f.dx;
    }
    if (nn < 3 || nn >= f.ndata)
        return;
    cvala = -mult * ct / (f.c * r) / 2.0 / f.dt * size;
    if (!s)
        aval = -aval;
    const long long c = i + (long long)g.px * j + g.plane * k;
    const float *FF[3] = { g.ex, g.ey, g.ez };
    double *bins = f.bins + (long long)n * 3 * f.ndata;
    for (int qq = 0; qq < 3; qq++) {
        const double e = FF[qq][c];
        const double de = (double)FF[qq][c + g.plane] - (double)FF[qq][c - g.plane];
        const double cvalb = cvala * e;
        double *bq = bins + (long long)qq * f.ndata;
        atomicAdd(bq + nn, -cvalb);
        atomicAdd(bq + nn - 1, f.post ? aval * de + b * e : (aval * de + b * e) * size);
        atomicAdd(bq + nn - 2, cvalb);
    }
}

/* running DFT, sampling half: the four tangential components of every NFFF
 * box face cell into ring slot `slot`.  E is taken on the face plane; H is the
 * mean of the two planes either side of it (the cell and its neighbour along
 * the face normal), which puts H on the face plane as well (face codes 0-5:
 * x, y, z faces; 6: FLUX_DFT z planes) */
__global__ __launch_bounds__(256) void
k_dft_sample(GsvGrid g, GsvDft d, int slot)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.ncells)
        return;
    const long long c = d.cell[q];
    const int face = d.face[q];
    float v0, v1, v2, v3;
    if (face < 2) {
        v0 = g.ey[c]; v1 = g.ez[c];
        v2 = 0.5f * (g.hy[c] + g.hy[c - 1]); v3 = 0.5f * (g.hz[c] + g.hz[c - 1]);
    } else if (face < 4) {
        v0 = g.ex[c]; v1 = g.ez[c];
        v2 = 0.5f * (g.hx[c] + g.hx[c - g.px]); v3 = 0.5f * (g.hz[c] + g.hz[c - g.px]);
    } else if (face < 6) {
        v0 = g.ex[c]; v1 = g.ey[c];
        v2 = 0.5f * (g.hx[c] + g.hx[c - g.plane]); v3 = 0.5f * (g.hy[c] + g.hy[c - g.plane]);
    } else {
        v0 = g.ex[c];
        v1 = g.ey[c];
        v2 = 0.5f * (g.hx[c] + g.hx[c - g.plane]);
        v3 = 0.5f * (g.hy[c] + g.hy[c - g.plane]);
    }
    float *r = d.ring + (long long)slot * 4 * d.ncells + q;
    r[0] = v0;
    r[d.ncells] = v1;
    r[2 * d.ncells] = v2;
    r[3 * d.ncells] = v3;
}

/* running DFT, folding half: acc += F(t) exp(-i w t) for the pending samples
 * of steps first_step .. first_step + n - 1 (slot = step % T), E sampled at
 * (step+1) dt and H at (step+1/2) dt; the phases of the block's steps are
 * tabulated in LDS (T * nw * 4 doubles, at most 48 KiB) */
__global__ __launch_bounds__(256) void
k_dft_flush(GsvDft d, int first_step, int n)
{
    extern __shared__ double ph[];   /* [n][nw][4]: cos E, sin E, cos H, sin H */
    for (int t = threadIdx.x; t < n * d.nw; t += blockDim.x) {
        const int s = t / d.nw, w = t % d.nw;
        const double te = (first_step + s + 1) * d.dt, th = (first_step + s + 0.5) * d.dt;
        double se, ce, sh, ch;
        sincos(-d.omega[w] * te, &se, &ce);
        sincos(-d.omega[w] * th, &sh, &ch);
        ph[4 * t] = ce; ph[4 * t + 1] = se; ph[4 * t + 2] = ch; ph[4 * t + 3] = sh;
    }
    __syncthreads();
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= d.ncells)
        return;
    for (int w = 0; w < d.nw; w++) {
        double *a = d.acc + (q * d.nw + w) * 8;
        double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6], a7 = a[7];
        for (int s = 0; s < n; s++) {
            const int slot = (first_step + s) % d.T;
            const float *r = d.ring + (long long)slot * 4 * d.ncells + q;
            const double v0 = r[0], v1 = r[d.ncells], v2 = r[2 * d.ncells], v3 = r[3 * d.ncells];
            const double *p = ph + 4 * (s * d.nw + w);
            a0 += v0 * p[0]; a1 += v0 * p[1];
            a2 += v1 * p[0]; a3 += v1 * p[1];
            a4 += v2 * p[2]; a5 += v2 * p[3];
            a6 += v3 * p[2]; a7 += v3 * p[3];
        }
        a[0] = a0; a[1] = a1; a[2] = a2; a[3] = a3; a[4] = a4; a[5] = a5; a[6] = a6; a[7] = a7;
    }
}

/* box sum of E^2 (modes of SvSumType); double accumulation, one atomic per block */
__global__ __launch_bounds__(256) void
k_box_sum(GsvGrid g, GsvBoxSum b, int step)
{
    __shared__ double red[256];
    const long long n = (long long)b.ni * b.nj * b.nk;
    const int t = threadIdx.x;
    double acc = 0.0;
    for (long long q = (long long)blockIdx.x * blockDim.x + t; q < n; q += (long long)gridDim.x * blockDim.x) {
        const float w = b.w ? b.w[q] : 1.f;
        if (w == 0.f)
            continue;
        const int i = (int)(q % b.ni);
        const long long r = q / b.ni;
        const int j = (int)(r % b.nj), k = (int)(r / b.nj);
        const long long c = (long long)(i + b.i0) + (long long)g.px * (j + b.j0) + g.plane * (k + b.k0);
        double v = 0.0;
        if (b.mode != SV_SUM_EY && b.mode != SV_SUM_EZ) { const double e = g.ex[c]; v += e * e; }
        if (b.mode != SV_SUM_EX && b.mode != SV_SUM_EZ) { const double e = g.ey[c]; v += e * e; }
        if (b.mode != SV_SUM_EX && b.mode != SV_SUM_EY) { const double e = g.ez[c]; v += e * e; }
        if (b.mode == SV_SUM_MAX)
            acc = v > acc ? v : acc;
        else
            acc += (double)w * v;
    }
    red[t] = acc;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (t < h)
            red[t] = (b.mode == SV_SUM_MAX) ? fmax(red[t], red[t + h]) : red[t] + red[t + h];
        __syncthreads();
    }
    if (t == 0) {
        if (b.mode == SV_SUM_MAX)   /* non-negative doubles order like their bit patterns */
            atomicMax((unsigned long long *)(b.out + step), (unsigned long long)__double_as_longlong(red[0]));
        else
            atomicAdd(b.out + step, red[0]);
    }
}

/* ------------------------------------------------------------------ */
/* launchers                                                            */
/* ------------------------------------------------------------------ */

/* 32-bit byte offsets cover every per-cell array when one field array
 * (guard planes included) is below 4 GiB */
static bool
offsets32(const GsvGrid *g)
{
    return (unsigned long long)g->plane * (g->nz + 2) * 4ull < (1ull << 32);
}

extern "C" hipError_t
gsv_launch_hstep(const GsvGrid *g, int idbytes, int bx, int by, int kchunk, hipStream_t st)
{
    dim3 blk(bx, by, 1), grd((g->nx + bx - 1) / bx, (g->ny + by - 1) / by, (g->nz + kchunk - 1) / kchunk);
    const bool o32 = offsets32(g), per = g->perx || g->pery;
#define GSV_HSTEP(ID, OFF)                                                                      \
    do {                                                                                         \
        if (per) hipLaunchKernelGGL((k_hstep<ID, OFF, true>), grd, blk, 0, st, *g, kchunk);    \
        else hipLaunchKernelGGL((k_hstep<ID, OFF, false>), grd, blk, 0, st, *g, kchunk);       \
    } while (0)
    if (idbytes == 1) {
        if (o32) GSV_HSTEP(uint8_t, uint32_t);
        else GSV_HSTEP(uint8_t, uint64_t);
    } else {
        if (o32) GSV_HSTEP(uint16_t, uint32_t);
        else GSV_HSTEP(uint16_t, uint64_t);
    }
#undef GSV_HSTEP
    return hipGetLastError();
}

/* the per-cell coefficient branch (design region) and the periodic
 * neighbour offsets are compiled in only for pools that need them, so the
 * production kernels carry none of their registers */
template <typename ID, typename OFF, bool PC, bool PER>
static void
launch_estep(const GsvGrid *g, int ade, dim3 grd, dim3 blk, int kchunk, hipStream_t st)
{
    if (ade == 0)
        hipLaunchKernelGGL((k_estep<ID, 0, OFF, PC, PER>), grd, blk, 0, st, *g, kchunk);
    else if (ade == 1)
        hipLaunchKernelGGL((k_estep<ID, 1, OFF, PC, PER>), grd, blk, 0, st, *g, kchunk);
    else if (ade == 2)
        hipLaunchKernelGGL((k_estep<ID, 2, OFF, PC, PER>), grd, blk, 0, st, *g, kchunk);
    else if (ade == 3)
        hipLaunchKernelGGL((k_estep<ID, 3, OFF, PC, PER>), grd, blk, 0, st, *g, kchunk);
    else
        hipLaunchKernelGGL((k_estep<ID, 4, OFF, PC, PER>), grd, blk, 0, st, *g, kchunk);
}

/* dispersive build of a grid: 0 none, 1 Drude only, 2 / 3 Drude + 2 / 3
 * critical-point poles, 4 recursive-convolution materials */
static int
ade_mode(const GsvGrid *g)
{
    if (g->rc)
        return 4;
    return g->ade ? (g->ade_poles ? g->ade_poles : 1) : 0;
}

extern "C" hipError_t
gsv_launch_estep(const GsvGrid *g, int idbytes, int bx, int by, int kchunk, hipStream_t st)
{
    dim3 blk(bx, by, 1), grd((g->nx + bx - 1) / bx, (g->ny + by - 1) / by, (g->nz + kchunk - 1) / kchunk);
    /* the dispersive branch is compiled in only when the pool has ADE cells,
     * and the critical-point terms only when some material has them */
    const int ade = ade_mode(g);
    const bool o32 = offsets32(g), pc = g->pc_nx > 0, per = g->perx || g->pery;
    if (pc && (!o32 || per))
        return hipErrorInvalidValue;   /* design regions need 32-bit offsets (field arrays below 4 GiB), no periodic axes */
#define GSV_ESTEP(ID)                                                                               \
    do {                                                                                            \
        if (pc) launch_estep<ID, uint32_t, true, false>(g, ade, grd, blk, kchunk, st);             \
        else if (o32 && per) launch_estep<ID, uint32_t, false, true>(g, ade, grd, blk, kchunk, st); \
        else if (o32) launch_estep<ID, uint32_t, false, false>(g, ade, grd, blk, kchunk, st);      \
        else if (per) launch_estep<ID, uint64_t, false, true>(g, ade, grd, blk, kchunk, st);       \
        else launch_estep<ID, uint64_t, false, false>(g, ade, grd, blk, kchunk, st);               \
    } while (0)
    if (idbytes == 1)
        GSV_ESTEP(uint8_t);
    else
        GSV_ESTEP(uint16_t);
#undef GSV_ESTEP
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_tsf_line(const GsvTsf *t, int step, hipStream_t st)
{
    hipLaunchKernelGGL(k_tsf_line, dim3(1), dim3(1024), 0, st, *t, step);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_ext(const GsvGrid *g, const GsvExt *x, int hside, hipStream_t st)
{
    hipLaunchKernelGGL(k_ext_plane, dim3((x->ni * x->nj + 255) / 256), dim3(256), 0, st, *g, *x, hside);
    return hipGetLastError();
}

hipError_t
gsv_launch_tsff_line(const GsvTsff *t, int step, hipStream_t st)
{
    hipLaunchKernelGGL(k_tsff_line, dim3(t->nw), dim3(256), 0, st, t->lines, step);
    return hipGetLastError();
}

hipError_t
gsv_launch_tsff_faces(const GsvGrid *g, const GsvTsff *t, int hside, hipStream_t st)
{
    const int nx = t->in - t->i0 + 1, ny = t->jn - t->j0 + 1, nz = t->kn - t->k0 + 1;
    hipLaunchKernelGGL(k_tsff_face<0>, dim3((2 * ny * nz + 255) / 256), dim3(256), 0, st, *g, *t, hside);
    hipLaunchKernelGGL(k_tsff_face<1>, dim3((2 * nx * nz + 255) / 256), dim3(256), 0, st, *g, *t, hside);
    hipLaunchKernelGGL(k_tsff_face<2>, dim3((2 * nx * ny + 255) / 256), dim3(256), 0, st, *g, *t, hside);
    return hipGetLastError();
}

hipError_t
gsv_launch_tsf_faces(const GsvGrid *g, const GsvTsf *t, int hside, hipStream_t st)
{
    for (int axis = 0; axis < 3; axis++) {
        int na, nb;
        tsf_face_extent(*t, g->nx, g->ny, axis, &na, &nb);
        dim3 blk(64, 4, 1), grd((na + 63) / 64, (nb + 3) / 4, 2);
        if (hside)
            hipLaunchKernelGGL(k_tsf_h, grd, blk, 0, st, *g, *t, axis);
        else
            hipLaunchKernelGGL(k_tsf_e, grd, blk, 0, st, *g, *t, axis);
    }
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_point_src(const GsvGrid *g, const GsvPointSrc *s, int step, int hside, hipStream_t st)
{
    hipLaunchKernelGGL(k_point_src, dim3(1), dim3(64), 0, st, *g, *s, step, hside);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_record(const GsvGrid *g, const long long *pos, int n, float *rec, hipStream_t st)
{
    hipLaunchKernelGGL(k_record_points, dim3((n + 63) / 64), dim3(64), 0, st, *g, pos, n, rec);
    return hipGetLastError();
}

static void
axis_dims(const GsvGrid *g, int axis, int *na, int *nb)
{
    if (axis == 0) { *na = g->ny; *nb = g->nz; }
    else if (axis == 1) { *na = g->nx; *nb = g->nz; }
    else { *na = g->nx; *nb = g->ny; }
}

extern "C" hipError_t
gsv_launch_liao_store(const GsvGrid *g, const GsvLiao *L, hipStream_t st)
{
    for (int axis = 0; axis < 3; axis++) {
        if (!L->on[2 * axis] && !L->on[2 * axis + 1])
            continue;
        int na, nb;
        axis_dims(g, axis, &na, &nb);
        hipLaunchKernelGGL(k_liao_store, dim3((na + 63) / 64, (nb + 3) / 4), dim3(64, 4), 0, st, *g, *L, axis);
    }
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_liao_faces(const GsvGrid *g, const GsvLiao *L, hipStream_t st)
{
    for (int axis = 0; axis < 3; axis++)
        for (int far = 0; far < 2; far++) {
            if (!L->on[2 * axis + far])
                continue;
            int na, nb;
            axis_dims(g, axis, &na, &nb);
            hipLaunchKernelGGL(k_liao_face, dim3((na + 63) / 64, (nb + 3) / 4), dim3(64, 4), 0, st, *g, *L, axis, far);
        }
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_zero_planes(const GsvGrid *g, int axis, hipStream_t st)
{
    int na, nb;
    axis_dims(g, axis, &na, &nb);
    hipLaunchKernelGGL(k_zero_planes, dim3((na + 63) / 64, (nb + 3) / 4), dim3(64, 4), 0, st, *g, axis);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_periodic(const GsvGrid *g, const GsvPeriodic *p, int hside, hipStream_t st)
{
    int m = g->nx > g->ny ? g->nx : g->ny;
    if (g->nz > m) m = g->nz;
    dim3 blk(64, 4), grd((m + 63) / 64, (m + 3) / 4);
    if (hside)
        hipLaunchKernelGGL(k_periodic_h, grd, blk, 0, st, *g, *p);
    else
        hipLaunchKernelGGL(k_periodic_e, grd, blk, 0, st, *g, *p);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_ltsf_line(const GsvLtsfLine *l, hipStream_t st)
{
    const int nt = l->res >= 1024 ? 1024 : ((l->res + 63) / 64) * 64;
    const size_t lds = 8 * (size_t)(l->res + 2) * sizeof(double);
    if (lds > 160 * 1024)
        return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        hipError_t err = hipFuncSetAttribute((const void *)k_ltsf_line, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (err != hipSuccess)
            return err;
    }
    hipLaunchKernelGGL(k_ltsf_line, dim3(1), dim3(nt), lds, st, *l);
    return hipGetLastError();
}

/* one launch per axis, x then y then z, as the reference's loops: faces of
 * different axes correct the same components on the box edges */
extern "C" hipError_t
gsv_launch_ltsf_faces(const GsvGrid *g, const GsvLtsf *l, int hside, int step, hipStream_t st)
{
    for (int axis = 0; axis < 3; axis++) {
        const int na = axis == 0 ? l->j1 - l->j0 + 1 : l->i1 - l->i0 + 1;
        const int nb = axis == 2 ? l->j1 - l->j0 + 1 : l->k1 - l->k0 + 1;
        const int n = 2 * na * nb;
        hipLaunchKernelGGL(k_ltsf_face, dim3((n + 255) / 256), dim3(256), 0, st, *g, *l, axis, hside, step);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess)
            return err;
    }
    return hipSuccess;
}

extern "C" hipError_t
gsv_launch_force(const GsvGrid *g, const GsvForce *f, int step, hipStream_t st)
{
    hipLaunchKernelGGL(k_force, dim3((unsigned)((f->ncells + 255) / 256)), dim3(256), 0, st, *g, *f, step);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_pfarfield(const GsvGrid *g, const GsvPFarfield *f, int step, hipStream_t st)
{
    if (f->nwork <= 0 || f->nrs <= 0)
        return hipSuccess;
    dim3 grd((unsigned)((f->nwork + 255) / 256), (unsigned)f->nrs);
    hipLaunchKernelGGL(k_pfarfield, grd, dim3(256), 0, st, *g, *f, step);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_farfield_td(const GsvGrid *g, const GsvFarfieldTD *f, int step, hipStream_t st)
{
    dim3 grd((unsigned)((f->ncells + 255) / 256), f->nrs);
    hipLaunchKernelGGL(k_farfield_td, grd, dim3(256), 0, st, *g, *f, step);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_dft_sample(const GsvGrid *g, const GsvDft *d, int slot, hipStream_t st)
{
    hipLaunchKernelGGL(k_dft_sample, dim3((unsigned)((d->ncells + 255) / 256)), dim3(256), 0, st, *g, *d, slot);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_dft_flush(const GsvDft *d, int first_step, int nsamples, hipStream_t st)
{
    if (nsamples <= 0)
        return hipSuccess;
    const size_t lds = (size_t)nsamples * d->nw * 4 * sizeof(double);
    hipLaunchKernelGGL(k_dft_flush, dim3((unsigned)((d->ncells + 255) / 256)), dim3(256), lds, st, *d, first_step,
                       nsamples);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* batched stepping: the same column and face-cell updates, one launch per */
/* kernel type for every grid of a batch; a workgroup reads its tile, then */
/* the tile's grid parameters (uniform, scalar loads)                       */
/* ------------------------------------------------------------------ */

template <typename ID, typename OFF, bool PER>
__global__ __launch_bounds__(512) void
k_hstep_batch(const GsvGrid *__restrict__ grids, const GsvTile *__restrict__ tiles, int ntiles, int kchunk)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const GsvTile e = tiles[t];
        hstep_col<ID, OFF, PER>(grids[e.run], e.tx * blockDim.x + threadIdx.x, e.ty * blockDim.y + threadIdx.y, e.kz, kchunk);
    }
}

template <typename ID, int ADE, typename OFF, bool PER>
__global__ __launch_bounds__(512) void
k_estep_batch(const GsvGrid *__restrict__ grids, const GsvTile *__restrict__ tiles, int ntiles, int kchunk)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const GsvTile e = tiles[t];
        estep_col<ID, ADE, OFF, false, PER>(grids[e.run], e.tx * blockDim.x + threadIdx.x,
                                            e.ty * blockDim.y + threadIdx.y, e.kz, kchunk);
    }
}

/* one workgroup per tile: without the resident loop the E column needs fewer
 * registers (80 against 83 VGPRs for the Drude build, six waves per SIMD
 * instead of five; a seven-wave bound would spill) */
template <typename ID, int ADE, typename OFF, bool PER>
__global__ __launch_bounds__(512) void
k_estep_tile(const GsvGrid *__restrict__ grids, const GsvTile *__restrict__ tiles, int kchunk)
{
    const GsvTile e = tiles[blockIdx.x];
    estep_col<ID, ADE, OFF, false, PER>(grids[e.run], e.tx * blockDim.x + threadIdx.x, e.ty * blockDim.y + threadIdx.y,
                                        e.kz, kchunk);
}

/* one block per grid with a TSF line; the grid's step counter lives in device
 * memory so that a captured step sequence replays with advancing steps */
__global__ __launch_bounds__(1024) void
k_tsf_line_batch(const GsvTsf *__restrict__ tsfs, const int *__restrict__ runs, int *steps)
{
    const int r = runs[blockIdx.x];
    const int step = threadIdx.x == 0 ? steps[r] : 0;
    tsf_line_step(tsfs[r], step);
    if (threadIdx.x == 0)
        steps[r] = step + 1;
}

__global__ __launch_bounds__(256) void
k_tsf_h_batch(const GsvGrid *__restrict__ grids, const GsvTsf *__restrict__ tsfs, const GsvFaceTile *__restrict__ tiles)
{
    const GsvFaceTile e = tiles[blockIdx.x];
    tsf_h_cell(grids[e.run], tsfs[e.run], e.face >> 1, e.face & 1, e.ta * blockDim.x + threadIdx.x,
               e.tb * blockDim.y + threadIdx.y);
}

__global__ __launch_bounds__(256) void
k_tsf_e_batch(const GsvGrid *__restrict__ grids, const GsvTsf *__restrict__ tsfs, const GsvFaceTile *__restrict__ tiles)
{
    const GsvFaceTile e = tiles[blockIdx.x];
    tsf_e_cell(grids[e.run], tsfs[e.run], e.face >> 1, e.face & 1, e.ta * blockDim.x + threadIdx.x,
               e.tb * blockDim.y + threadIdx.y);
}

static dim3
batch_grid(int ntiles, int nblocks)
{
    return dim3((unsigned)(nblocks > 0 && nblocks < ntiles ? nblocks : ntiles));
}

extern "C" hipError_t
gsv_launch_hstep_batch(const GsvGrid *grids, const GsvTile *tiles, int ntiles, int idbytes, int per, int bx, int by,
                       int kchunk, int nblocks, hipStream_t st)
{
    if (ntiles <= 0)
        return hipSuccess;
    const dim3 blk(bx, by, 1), grd = batch_grid(ntiles, nblocks);
#define GSV_HB(ID, PER) hipLaunchKernelGGL((k_hstep_batch<ID, uint32_t, PER>), grd, blk, 0, st, grids, tiles, ntiles, kchunk)
    if (idbytes == 1) {
        if (per) GSV_HB(uint8_t, true);
        else GSV_HB(uint8_t, false);
    } else {
        if (per) GSV_HB(uint16_t, true);
        else GSV_HB(uint16_t, false);
    }
#undef GSV_HB
    return hipGetLastError();
}

template <typename ID, bool PER>
static void
launch_estep_batch(const GsvGrid *grids, const GsvTile *tiles, int ntiles, int ade, dim3 grd, dim3 blk, int kchunk,
                   hipStream_t st)
{
    if ((int)grd.x == ntiles) {
        if (ade == 0)
            hipLaunchKernelGGL((k_estep_tile<ID, 0, uint32_t, PER>), grd, blk, 0, st, grids, tiles, kchunk);
        else if (ade == 1)
            hipLaunchKernelGGL((k_estep_tile<ID, 1, uint32_t, PER>), grd, blk, 0, st, grids, tiles, kchunk);
        else if (ade == 2)
            hipLaunchKernelGGL((k_estep_tile<ID, 2, uint32_t, PER>), grd, blk, 0, st, grids, tiles, kchunk);
        else
            hipLaunchKernelGGL((k_estep_tile<ID, 3, uint32_t, PER>), grd, blk, 0, st, grids, tiles, kchunk);
        return;
    }
    if (ade == 0)
        hipLaunchKernelGGL((k_estep_batch<ID, 0, uint32_t, PER>), grd, blk, 0, st, grids, tiles, ntiles, kchunk);
    else if (ade == 1)
        hipLaunchKernelGGL((k_estep_batch<ID, 1, uint32_t, PER>), grd, blk, 0, st, grids, tiles, ntiles, kchunk);
    else if (ade == 2)
        hipLaunchKernelGGL((k_estep_batch<ID, 2, uint32_t, PER>), grd, blk, 0, st, grids, tiles, ntiles, kchunk);
    else
        hipLaunchKernelGGL((k_estep_batch<ID, 3, uint32_t, PER>), grd, blk, 0, st, grids, tiles, ntiles, kchunk);
}

extern "C" hipError_t
gsv_launch_estep_batch(const GsvGrid *grids, const GsvTile *tiles, int ntiles, int idbytes, int ade, int per, int bx,
                       int by, int kchunk, int nblocks, hipStream_t st)
{
    if (ntiles <= 0)
        return hipSuccess;
    const dim3 blk(bx, by, 1), grd = batch_grid(ntiles, nblocks);
    if (idbytes == 1) {
        if (per) launch_estep_batch<uint8_t, true>(grids, tiles, ntiles, ade, grd, blk, kchunk, st);
        else launch_estep_batch<uint8_t, false>(grids, tiles, ntiles, ade, grd, blk, kchunk, st);
    } else {
        if (per) launch_estep_batch<uint16_t, true>(grids, tiles, ntiles, ade, grd, blk, kchunk, st);
        else launch_estep_batch<uint16_t, false>(grids, tiles, ntiles, ade, grd, blk, kchunk, st);
    }
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_tsf_line_batch(const GsvTsf *tsfs, const int *runs, int nruns, int *steps, hipStream_t st)
{
    if (nruns <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_tsf_line_batch, dim3(nruns), dim3(1024), 0, st, tsfs, runs, steps);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_tsf_faces_batch(const GsvGrid *grids, const GsvTsf *tsfs, const GsvFaceTile *tiles, int ntiles, int hside,
                           hipStream_t st)
{
    if (ntiles <= 0)
        return hipSuccess;
    if (hside)
        hipLaunchKernelGGL(k_tsf_h_batch, dim3(ntiles), dim3(64, 4), 0, st, grids, tsfs, tiles);
    else
        hipLaunchKernelGGL(k_tsf_e_batch, dim3(ntiles), dim3(64, 4), 0, st, grids, tsfs, tiles);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* HBM probe: a plain streaming copy of known size, used to calibrate the
 * FETCH_SIZE / WRITE_SIZE counters and to measure attainable bandwidth    */
/* ------------------------------------------------------------------ */

extern "C" hipError_t
gsv_launch_sf_pec(const GsvGrid *g, const GsvTsf *t, const int4 *cells, int n, hipStream_t st)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_sf_pec, dim3((n + 255) / 256), dim3(256), 0, st, *g, *t, cells, n);
    return hipGetLastError();
}

extern "C" hipError_t
gsv_launch_box_sum(const GsvGrid *g, const GsvBoxSum *b, int step, hipStream_t st)
{
    const long long n = (long long)b->ni * b->nj * b->nk;
    if (n <= 0)
        return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 2048)
        blocks = 2048;
    hipLaunchKernelGGL(k_box_sum, dim3((unsigned)blocks), dim3(256), 0, st, *g, *b, step);
    return hipGetLastError();
}
