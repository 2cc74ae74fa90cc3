/*
 * engine.h -- device-side data layout of the MI355X FDTD engine, shared by the
 * HIP kernels (kernels.hip) and the host driver (gpu_engine.hip).
 *
 * Layout in HBM
 *   fields      six float arrays Ex..Hz, x fastest: idx = i + px*(j + ny*k),
 *               px = nx rounded up to 16 floats (64 B rows); each array has one
 *               guard plane below k = 0 and above k = nz-1 so that every
 *               neighbour load of the stencils stays inside the allocation.
 *   classes     one uint8 (or uint16) coefficient-class id per cell for the E
 *               update and one for the H update; the class tables hold the
 *               per-component update coefficients, the dispersive flags and
 *               the PEC zero mask, computed on the host in double and rounded.
 *   CPML psi    stored only on the boundary slabs (x slabs [k][j][depth],
 *               y slabs [k][depth][px], z slabs [depth][ny][px]).
 *   ADE state   full-grid float arrays with the same guard planes, touched
 *               only in dispersive cells.
 * The volume kernels address every per-cell array by one 32-bit byte offset
 * from the guard-plane start (one offset register shared by all arrays);
 * grids whose field arrays exceed 4 GiB use 64-bit offsets.
 */
#ifndef GSV_ENGINE_H
#define GSV_ENGINE_H

#include <stdint.h>

/* flags of an E class */
#define GSV_E_TBCX 0x1u      /* Ex takes the ordinary (non-dispersive) update */
#define GSV_E_TBCY 0x2u
#define GSV_E_TBCZ 0x4u
#define GSV_E_ADE 0x8u       /* at least one component takes the ADE update */
#define GSV_E_SLOT_SHIFT 4   /* bits 4..7: ADE material slot */
#define GSV_E_PECX 0x100u    /* PEC zeroing of Ex after the update */
#define GSV_E_PECY 0x200u
#define GSV_E_PECZ 0x400u

#define GSV_MAX_ADE 16

typedef struct {
    float ca[3];
    float cb[3];
    uint32_t flags;
    float pad;
} GsvECoef;

typedef struct {
    float da;
    float db[3];     /* H update, per axis */
    float dbc;       /* CPML H correction, src3d/boundary.c uses the dx-based db */
    float pad[3];
} GsvHCoef;

/* ADE material (Drude + 2 critical points), src3d/plan.c:2736-2762 */
typedef struct {
    float c0, c1, c2, c3;
    float a0, a1, a2;
    float bp0[2], bp1[2], bp2[2], bp3[2], bp4[2];
    float cbm[3];    /* material cb used by the ADE branch's inline CPML, src3d/yee.c:952-1057 */
    int poles;       /* nonzero if the critical-point terms are present */
} GsvAde;

/* one CPML face on the device */
typedef struct {
    int on;
    int depth;
    float *be, *ce_d;       /* E side: b, c / spacing of the stretched axis (boundary update) */
    float *bh, *ch_d;       /* H side: b, c / spacing */
    float *ce_raw;          /* E side c (the ADE inline update divides by another spacing) */
    float *pe1, *pe2;       /* E psi */
    float *ph1, *ph2;       /* H psi */
} GsvCpmlFace;

typedef struct {
    int nx, ny, nz;
    int px;                 /* row pitch (floats) */
    long long plane;        /* px * ny */
    float *ex, *ey, *ez, *hx, *hy, *hz;   /* base (k = 0) pointers */
    float *exg, *eyg, *ezg, *hxg, *hyg, *hzg;   /* allocation starts (guard plane k = -1) */
    const void *ecls;       /* uint8_t or uint16_t per cell */
    const void *hcls;
    const GsvECoef *etab;
    const GsvHCoef *htab;
    int netab, nhtab;
    const float *rkex, *rkey, *rkez;      /* 1/kappa on the E grid, per axis */
    const float *rkhx, *rkhy, *rkhz;      /* 1/kappa on the H grid, per axis */
    GsvCpmlFace f[6];                     /* x0, xn, y0, yn, z0, zn */
    float dx, dy, dz;
    /* ADE */
    int ade, ade_poles;
    const GsvAde *adem;     /* ADE material table in device memory */
    /* ADE state: same layout as the fields including the guard planes; these
     * are allocation starts, so cell (i,j,k) is at the fields' guard offset */
    float *exp_, *eyp, *ezp, *dpx, *dpy, *dpz, *dpxp, *dpyp, *dpzp;
    float *px0, *px1, *py0, *py1, *pz0, *pz1, *pxp0, *pxp1, *pyp0, *pyp1, *pzp0, *pzp1;
} GsvGrid;

#endif
