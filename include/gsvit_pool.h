/*
 * gsvit_pool.h -- computational pool (SvPool) and the drop-in GPU entry points.
 *
 * The entry points keep the names, argument lists and call order of GSvit's
 * GPU driver, declared in src3d/pool.h:98-124,167 and called by the step loop
 * of src3d/svit.c:203-257.  Each returns 0 on success; unlike the reference,
 * which prints CUDA errors and returns 0 (src3d/gpu.c:811-817), a HIP error
 * makes the call return a nonzero code and print the HIP error string.
 *
 * Host containers keep the reference's [i][j][k] accessor (src3d/dcube.h:35-43,
 * src3d/icube.h:32-40) but are backed by one contiguous block each.
 */
#ifndef GSVIT_POOL_H
#define GSVIT_POOL_H

#include "gsvit_settings.h"

#ifdef __cplusplus
extern "C" {
#endif

/* synchronisation selector, src3d/pool.h:49-54 */
typedef enum {
    SV_SYNC_ALL = 0,
    SV_SYNC_EH = 1,
    SV_SYNC_E = 2,
    SV_SYNC_H = 3
} SvGpuSync;

typedef struct { int xres, yres, zres; double ***data; double *block; } SvDCube;
typedef struct { int xres, yres, zres; float ***data; float *block; } SvFCube;
typedef struct { int xres, yres, zres; int ***data; int *block; } SvICube;

SvDCube *sv_dcube_new(int xres, int yres, int zres);
SvFCube *sv_fcube_new(int xres, int yres, int zres);
SvICube *sv_icube_new(int xres, int yres, int zres);
void sv_dcube_free(SvDCube *c);
void sv_fcube_free(SvFCube *c);
void sv_icube_free(SvICube *c);
void sv_fcube_fill(SvFCube *c, float v);
void sv_icube_fill(SvICube *c, int v);

/* field and material storage, src3d/yee.h:45-101 (dispersive state excluded:
 * the device holds it) */
typedef struct {
    SvDCube *ex, *ey, *ez, *hx, *hy, *hz;
    SvFCube *epsilon, *mu, *sigma, *sigast;   /* absolute values; NULL unless matmode needs them */
    SvICube *mat;                             /* tabulated material index, NULL if no table */
    SvICube *bnds;                            /* 1 where E coefficients are averaged */
    SvICube *sourceskip;                      /* TSF faces skipped near materials, or NULL */
    double dx, dy, dz, dt;
    int xres, yres, zres;
} SvYeeData;

/* point source samples, src3d/source.h (SvSourcePoint / sdata) */
typedef struct {
    int i, j, k;
    int ndata;
    int *layered_zpos;
    double *ex, *ey, *ez, *hx, *hy, *hz;
} SvSourcePoint;

/* TSF plane-wave source, src3d/source.h (SvSourceTSF), 1-D line per src3d/plan.c:1584-1595 */
typedef struct {
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    int skip_i0, skip_j0, skip_k0, skip_in, skip_jn, skip_kn;
    double ia_theta, ia_phi, ia_psi;
    double corr;
    int ndata;
    double *e;                  /* incident waveform, one value per step */
    int jp_res;                 /* 1-D auxiliary line length */
    double jp_dx;               /* 1-D line spacing dx/corr */
    double jp_epsilon, jp_mu;   /* relative material of the 1-D line */
} SvSourceTSF;

/* layered TSF source, src3d/source.h (SvSourceLTSF) and src3d/plan.c:1600-1685:
 * the waveform (one sample per dt/3), the angular correction and the 1-D
 * split-field model along z (res = zres, spacing dx/corr, step dt/3) whose
 * fields are kept for nrows substeps: 3 nsteps plus a margin for the face
 * cells whose lateral delay is negative (they read up to a few substeps past
 * 3 (nsteps - 1); the reference reads past its stored history there) */
typedef struct {
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    double ia_theta, ia_phi, ia_psi;
    double corr;
    int ndata;
    double *e;
    int nrows;
    double zdx, zdt;
    double *zeps, *zmu;         /* relative material of the 1-D model cells */
} SvSourceLTSF;

/* focused TSF (src3d/source.h SvSourceTSFF, src3d/plan.c:1855-1979): nw plane
 * waves, each an SvSourceTSF with its own angles, corr and vacuum 1-D line of
 * (int) sqrt(xres^2 + yres^2 + zres^2) + 250 cells (sharing one waveform),
 * amplitude acorr (fs_amplitude, src3d/source.c:167-171) and distance
 * corr (250 - a . (centre + shift - r)) (srdcomp, src3d/source.c:173-194) */
typedef struct {
    int nw;
    SvSourceTSF *w;
    double *acorr;
    double xshift, yshift, zshift;
} SvSourceTSFF;

typedef struct {
    int npnts;
    SvSourcePoint *sp;
    SvSourceTSF *tsf;
    SvSourceTSFF *tsff;
    SvSourceLTSF *ltsf;
    /* scattered-field source (src3d/plan.c:1985-2040): waveform and 1-D line
     * of (int) sqrt(xres^2 + yres^2 + zres^2 + 10) cells of dx; corr = 1 and
     * the box fields are unused */
    SvSourceTSF *sf;
} SvSource;

/* recorded traces, src3d/output.h (outpointdata[i][step*6 + comp]) */
typedef struct {
    double **outpointdata;
    double **outsumdata;
    double **outforcedata;  /* OUT_FORCE: [force][step * 3 + component], src3d/output.c:121-126 */
    SvDCube **outabsdata;   /* OUT_VOLUME ABS / SUMALL / MAXALL accumulators, one per such output */
    int nabs;
} SvOutput;

/* CPML grading per face, src3d/boundary.h:70-152 (be, ce, bh, ch, kappae, kappah) */
typedef struct {
    double *be, *ce, *bh, *ch, *kappae, *kappah;
    int depth;
} SvCpmlFace;

typedef struct {
    SvCpmlFace x0, xn, y0, yn, z0, zn;
} SvBoundary;

/* time-domain Kirchhoff far field, src3d/farfield.h; bins as src3d/farfield.c:56-102 */
typedef struct {
    int istart;
    double *ex, *ey, *ez;
} SvRPoint;

typedef struct {
    int ndata;
    int nrs;
    SvRPoint *rpoints;
    /* running DFT of tangential fields on the box faces (no reference counterpart):
     * dft[f] holds, per face cell, 4 tangential components x ndft x (re, im) */
    int ndft;
    double *dft_omega;
    double *dft_data;
    long dft_ncells;
    /* far-field transform of the running DFT (NFFF_DFT_PATTERN), attached by
     * whatever holds the accumulators: dirs [ndirs][2] = theta, phi; out
     * [ndirs][ndft][4] = re, im of F_theta and F_phi; returns 0 on success */
    int (*dft_transform)(void *ctx, int ndirs, const double *dirs, double *out);
    void *dft_ctx;
    /* FLUX_DFT planes (no reference counterpart): per plane cell (plane, j, i
     * order), Ex, Ey, mean Hx, mean Hy x nflux wavelengths x (re, im) */
    int nflux;
    double *flux_omega;
    double *flux_data;
    long flux_ncells;
} SvFarfield;

/* periodic far field, src3d/pfarfield.h and src3d/pfarfield.c:36-62: per
 * point E bins, ndata = 1000 + 2 nsteps + diagonal / (c dt / dx), istart as
 * the plain far field */
typedef struct {
    int ndata;
    int nrs;
    SvRPoint *prpoints;
} SvPFarfield;

typedef struct _SvPool {
    SvYeeData *d;
    SvMatProp *mats;
    int nmat;
    SvSource *src;
    SvOutput *out;
    SvBoundary *bnd;
    SvFarfield *farfield;
    SvPFarfield *pfarfield;
    void *gpu;                  /* device engine, created by sv_gpu_new */
    SvSet *set;
} SvPool;

/* Planner: src3d/plan.c:5442-5604 (init_and_make_plan).  Returns NULL on error. */
SvPool *init_and_make_plan(SvSet *set);
void sv_pool_free(SvPool *mp, SvSet *set);

/* Point/far-field/sum log writers, src3d/output.c:51-99 (is_output), :556-710, :1266-1271 */
int is_output(SvPool *mp, SvSet *set);
void sv_pool_output(SvPool *mp, SvSet *set);

/* ---------------- drop-in GPU entry points ---------------- */

/* src3d/pool.h:167, impl src3d/gpu.c:1959.  Returns NULL if usegpu == 0 or no
 * device; otherwise sets set->plan.gpumode = SV_GPUMODE_FULL. */
void *sv_gpu_new(SvPool *mp, SvSet *set);
/* src3d/pool.h:98 -- allocate device state */
int sv_pool_gpu_init(SvPool *mp, SvSet *set);
/* src3d/pool.h:104 -- host -> device (fields, materials, coefficients) */
int sv_pool_gpu_copyto(SvPool *mp, SvSet *set, SvGpuSync what);
/* src3d/pool.h:105 -- device -> host (fields, traces, far-field bins) */
int sv_pool_gpu_copyfrom(SvPool *mp, SvSet *set, SvGpuSync what);
/* src3d/pool.h:102 -- Yee H update (fused with the CPML H psi update) */
int sv_pool_gpu_ystep_h(SvPool *mp, SvSet *set);
/* src3d/pool.h:103 -- Yee E update (fused with dispersive, PEC, CPML E and Liao) */
int sv_pool_gpu_ystep_e(SvPool *mp, SvSet *set);
/* src3d/pool.h:106 */
int sv_pool_gpu_getsum(SvPool *mp, SvSet *set);
/* src3d/pool.h:107 -- Liao plane store */
int sv_pool_gpu_boundary_copy(SvPool *mp, SvSet *set);
/* src3d/pool.h:108 -- interior periodic H plane copies */
int sv_pool_gpu_boundary_hstep(SvPool *mp, SvSet *set);
/* src3d/pool.h:109 -- no-op: the E boundary work runs inside ystep_e */
int sv_pool_gpu_boundary_estep(SvPool *mp, SvSet *set);
/* src3d/pool.h:99 -- point H sources, 1-D line steps, TSF, layered and focused TSF H faces */
int sv_pool_gpu_apply_source_hstep(SvPool *mp, SvSet *set);
/* src3d/pool.h:100 -- point E sources, TSF E faces */
int sv_pool_gpu_apply_source_estep(SvPool *mp, SvSet *set);
/* src3d/pool.h:113 -- far-field accumulation (time-domain bins and running DFT) */
int sv_pool_gpu_apply_farfield(SvPool *mp, SvSet *set);
/* src3d/pool.h:119 -- periodic far field (PERIODIC_NFFF, src3d/pfarfield.c): Kirchhoff bins over the
 * z planes and their periodic images, both modes; returns 0 when unused */
int sv_pool_gpu_apply_pfarfield(SvPool *mp, SvSet *set);
/* src3d/pool.h:114 -- record point traces */
int sv_pool_gpu_getpoints(SvPool *mp, SvSet *set);
/* src3d/pool.h:115 -- Maxwell-stress force boxes (OUT_FORCE, src3d/output.c:355-512): face sums of the
 * stress tensor per step; returns 0 when unused */
int sv_pool_gpu_getforce(SvPool *mp, SvSet *set);

#ifdef __cplusplus
}
#endif

#endif /* GSVIT_POOL_H */
