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
 *   ADE state   blocked per z plane (see GSV_ADE_D), touched only in
 *               dispersive cells.
 * The volume kernels address every per-cell array by one 32-bit byte offset
 * from the guard-plane start (one offset register shared by all arrays);
 * grids whose field arrays exceed 4 GiB use 64-bit offsets.
 */
#ifndef GSV_ENGINE_H
#define GSV_ENGINE_H

#include <stdint.h>
#include <hip/hip_runtime.h>

/* flags of an E class */
#define GSV_E_TBCX 0x1u      /* Ex takes the ordinary (non-dispersive) update */
#define GSV_E_TBCY 0x2u
#define GSV_E_TBCZ 0x4u
#define GSV_E_ADE 0x8u       /* at least one component takes the ADE update */
#define GSV_E_SLOT_SHIFT 4   /* bits 4..7: ADE material slot */
#define GSV_E_PECX 0x100u    /* PEC zeroing of Ex after the update */
#define GSV_E_PECY 0x200u
#define GSV_E_PECZ 0x400u
#define GSV_E_PERCELL 0x800u /* coefficients come from the per-cell table of the design region */

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

/* ADE material (Drude + 2 critical points), src3d/plan.c:2736-2762 and the
 * ADE branch of src3d/yee.c:914-1108, carried in increment form.  With
 * J^n = P^n - P^(n-1), the reference's polarisation recursion
 *     P^(n+1) = a0 P^(n-1) + a1 P^n + a2 E^(n-1) + a3 E^n + a4 E^(n+1)
 * is  J^(n+1) = -a0 J^n + s P^n + a2 E^(n-1) + a3 E^n + a4 E^(n+1),  s = a0 + a1 - 1,
 * and its E-update term -a0 P^(n-1) + (1 - a1) P^n is a0 J^n - s P^n: the same
 * update in exact arithmetic, but in FP32 the small increment is stored and
 * multiplied directly instead of arising from the cancellation of two nearly
 * equal P values with separately rounded coefficients (whose rounding, with
 * a0 + a1 - 1 no longer zero, also acts as a spurious restoring term).  The
 * Drude term has s = 0 identically (a0 = (nu dt - 2)/(nu dt + 2), a1 = 4/(nu
 * dt + 2)), so its state is J alone.
 *
 * Drude-only form.  Without critical points c2 = -a2 / den and cj = a0 / den,
 * so cj da = dj c2 exactly, and Q^n = cj J^n + c2 E^(n-1) obeys
 *     E^(n+1) = c0 curl / d + Q^n + c3 E^n,
 *     Q^(n+1) = dj Q^n + qa E^n + qb E^(n+1),   qa = 2 cj da + c2, qb = cj da:
 * one state per component instead of two (E^(n-1) and J).  The E kernels of
 * Drude-only pools without a design region carry Q in states 0-2; the
 * design-region builds (whose state the adjoint transposes and checkpoints)
 * keep E^(n-1) and J. */
typedef struct {
    float c0, c3, dj;  /* E update: c0 curl / d + ... + c2 E^(n-1) + c3 E^n; Drude: J' = dj J + ..., dj = -a0 */
    float qa, qb;      /* Drude-only form: Q' = dj Q + qa E^n + qb E^(n+1) */
    float cbm[3];    /* material cb used by the ADE branch's inline CPML, src3d/yee.c:952-1057 */
    /* (the 32 bytes above are all the Drude-only form reads: GsvAdeQ) */
    float c2;
    float cj;          /* E-update coefficient of the Drude J: c1 a0 */
    float da;          /* Drude: J' = dj J + da (E^(n-1) + 2 E^n + E^(n+1)); da = a2 */
    float pcj[3];      /* E-update coefficients of a pole's J and P: c1 bp0, -c1 s */
    float pcp[3];
    float pj[3], ps[3];           /* pole J recursion coefficients: -bp0, s */
    float b2[3], b3[3], b4[3];    /* pole E coefficients: bp2, bp3, bp4 */
    int poles;       /* critical-point poles with nonzero terms: 0, 2 or 3 (the third has no reference counterpart) */
    int drude;       /* the Drude current is driven (da != 0); without it J stays 0 and is neither read nor written */
} GsvAde;

/* the leading 32 bytes of a GsvAde: the coefficients of the Drude-only form */
typedef struct {
    float c0, c3, dj, qa, qb;
    float cbm[3];
} GsvAdeQ;

/* recursive-convolution material (types 2 Drude, 4 Drude + 2 critical points,
 * 6 piecewise-linear), src3d/yee.c:266-455, 620-905, with the coefficients the
 * reference forms per cell and step combined once on the host in double.
 * Per component the state is a Drude accumulator d and, per critical point,
 * a complex accumulator p; the E update is
 *     E' = ca E + cs (d + Re p_0 + Re p_1) + cb[c] curl
 * with curl the bracket of H differences (over kappa for type 6 only).  Type 4
 * advances d and p with E^n before the E update, types 2 and 6 after it:
 *     d' = dd d + d0 E^n + d1 E',   p' = rot p + q0 E^n + q1 E'  (complex). */
typedef struct {
    int kind;            /* SV_MAT_DRUDE, SV_MAT_CP or SV_MAT_PLRC */
    int poles;           /* 2 for types 4 and 6 (critical-point accumulators carried), 0 for type 2 */
    float ca, cs;
    float cb[3];         /* per component, the spacing's division included */
    float dd, d0, d1;
    float rr[2], ri[2];  /* per-step rotation of each pole's accumulator */
    float q0r[2], q0i[2], q1r[2], q1i[2];
    float cbm[3];        /* type 6: the material cb of its inline CPML correction, src3d/yee.c:752-856 */
} GsvRc;

/* one CPML face on the device: grading coefficients per index along the
 * stretched axis, packed so that one load brings all of them, and the two
 * psi values of a slab cell stored as a pair */
typedef struct {
    int on;
    int depth;
    const float4 *ce;       /* E side: {b, c / spacing of the stretched axis, c, 0} */
    const float2 *ch;       /* H side: {b, c / spacing} */
    float2 *pe;             /* E psi pairs per slab cell */
    float2 *ph;             /* H psi pairs per slab cell */
} GsvCpmlFace;

/* ADE state blocks: per z plane [j][i/64][state][64], x pitch pxa (a
 * multiple of 64), so one base per plane plus the lane's offset reaches
 * every state of a cell through immediate offsets (state * 256 bytes) */
#define GSV_ADE_D 6      /* E(n-1) x y z, Drude J x y z */
#define GSV_ADE_P(np) (6 * (np))   /* critical points: P of each pole x y z, then J of each pole x y z */

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
    int ade_id0;            /* E classes with the dispersive flag are numbered from ade_id0 up */
    const float *rkex, *rkey, *rkez;      /* 1/kappa on the E grid, per axis */
    const float *rkhx, *rkhy, *rkhz;      /* 1/kappa on the H grid, per axis */
    GsvCpmlFace f[6];                     /* x0, xn, y0, yn, z0, zn */
    float dx, dy, dz;
    int perx, pery;         /* domain-periodic x / y: the far plane's neighbour is the near plane */
    /* per-cell E coefficients of the design region: cells whose class carries
     * GSV_E_PERCELL read pctab[(i-pc_i0) + pc_nx*((j-pc_j0) + pc_ny*(k-pc_k0))] */
    int pc_i0, pc_j0, pc_k0, pc_nx, pc_ny, pc_nz;
    const GsvECoef *pctab;
    /* ADE */
    int ade, ade_poles;     /* ADE cells present; poles of the widest ADE material (0, 2 or 3) */
    const GsvAde *adem;     /* ADE material table in device memory */
    int pxa;                /* x pitch of the ADE state blocks */
    float *ade_d;           /* Drude state blocks, GSV_ADE_D states */
    float *ade_p;           /* critical-point state blocks, GSV_ADE_P(ade_poles) states (NULL without poles) */
    long long ade_plane_d;  /* floats per z plane of ade_d = ny * pxa * GSV_ADE_D */
    long long ade_plane_p;  /* floats per z plane of ade_p */
    int rc;                 /* the dispersive cells are recursive-convolution materials (table rcm, state in
                               the ADE blocks: d in Drude states 0-2, pole q's p as P / J slots 3q + c) */
    const GsvRc *rcm;
    int ade_cpml;           /* ADE_CPML_CONSISTENT: the ADE branch updates and applies the psi of its own
                               components once per step (curl coefficient c0, all six faces) and the
                               slab code leaves them alone; 0: the reference's double update */
} GsvGrid;

#endif
