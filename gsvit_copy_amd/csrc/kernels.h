/*
 * kernels.h -- parameter blocks of the surface kernels and the launcher
 * entry points of kernels.hip (all launches go to the engine's stream).
 */
#ifndef GSV_KERNELS_H
#define GSV_KERNELS_H

#include <hip/hip_runtime.h>
#include "engine.h"

/* TSF plane wave: the 1-D line in double, per-face coefficient arrays */
typedef struct {
    int i0, j0, k0, in, jn, kn;
    /* d(i,j,k) = corr * (10 + ax (i - oi) + ay (j - oj) + az (k - ok)), src3d/source.c:53-107 */
    double corr, ax, ay, az;
    int oi, oj, ok;
    /* polarisation factors of gex..ghz, src3d/source.c:130-165 */
    double gex, gey, gez, ghx, ghy, ghz;
    int res;                  /* 1-D line length */
    double *je, *jh, *jb;     /* line fields and the 8 boundary scalars */
    const double *wave;       /* incident waveform */
    int ndata;
    double multe, multh, coef;
    /* per face (x0, xn, y0, yn, z0, zn): dt/dx/eps (E) and dt/dx/mu0 (H) per face
     * cell, 0 where the face is skipped or (z0) the cell is PEC */
    const float *coefE[6];
    const float *coefH[6];
} GsvTsf;

/* focused TSF plane wave: line distance d(i,j,k) = corr (250 - (ax (cx - i) +
 * ay (cy - j) + az (cz - k))) with c = pool centre + shift (srdcomp,
 * src3d/source.c:173-194); polarisation factors times the wave's amplitude */
typedef struct {
    double corr, ax, ay, az, cx, cy, cz;
    double pe[3];             /* acorr (ghx, ghy, ghz): E faces, from the H line */
    double ph[3];             /* acorr (gex, gey, gez): H faces, from the E line */
} GsvTsffWave;

/* focused TSF: nw plane waves sharing one box; lines as [nw][res] */
typedef struct {
    int i0, j0, k0, in, jn, kn;
    int skip[6];              /* x0, xn, y0, yn, z0, zn */
    int nw, res;
    const GsvTsffWave *waves;
    const GsvTsf *lines;      /* per-wave line descriptors for the line step */
    const double *je, *jh;
    float hcor;               /* dt / dx / mu0 */
    const float *coefE[6];    /* dt / dx / eps of the face cell */
} GsvTsff;

/* external plane source window: ni x nj cells from (ijstart, jkstart) on plane k */
typedef struct {
    int k, ni, nj, ijstart, jkstart;
    float hcor;               /* dt / dx / mu0 */
    const float *ce;          /* dt / dx / eps per window cell on plane k */
    const float *px, *py;     /* this step's x and y plane values over the window */
} GsvExt;

/* cells of a TSF face along its two in-plane axes; a box that spans a
 * periodic axis ends at the pool's last plane */
static __host__ __device__ __forceinline__ void
tsf_face_extent(const GsvTsf &t, int nx, int ny, int axis, int *na, int *nb)
{
    /* written as selects: the if/else-if form of this function was compiled
     * for gfx950 (ROCm 7.2 hipcc, kernel-argument GsvTsf) into code that read
     * k0 in place of j0 for axis 2 */
    const int in = t.in < nx - 1 ? t.in : nx - 1, jn = t.jn < ny - 1 ? t.jn : ny - 1;
    const int i0 = t.i0, j0 = t.j0, k0 = t.k0, kn = t.kn;
    *na = (axis == 0) ? jn - j0 + 1 : in - i0 + 1;
    *nb = (axis == 2) ? jn - j0 + 1 : kn - k0 + 1;
}

typedef struct {
    int n;                    /* number of point sources */
    int nsteps;
    const long long *pos;     /* cell index per source */
    const float *vals;        /* [nsteps][n][6]: Ex Ey Ez Hx Hy Hz, 0 = no write */
} GsvPointSrc;

typedef struct {
    int on[6];                /* x0, xn, y0, yn, z0, zn use Liao */
    float *store[3];          /* [24][na][nb] per axis */
    const float *coef[6];     /* per face cell, 3 coefficients (dx, dy, dz based) */
} GsvLiao;

typedef struct {
    int x, y, z;              /* periodic along this axis */
    int x0, xn, y0, yn, z0, zn;
} GsvPeriodic;

typedef struct {
    int i0, j0, k0, i1, j1, k1;
    int ix, iy, iz;
    long long ncells;
    int skip[6][5];           /* per face (x0, xn, y0, yn, z0, zn): on, amin, bmin, amax, bmax */
    int nrs;
    const int *ri, *rj, *rk, *istart;
    int ndata;
    double *bins;             /* [nrs][3][ndata] */
    double dx, dy, dz, dt, c;
} GsvFarfieldTD;

/* layered TSF, the split-field 1-D model along z (one workgroup, fields in
 * LDS): res cells, nrows substeps of dt/3, every substep's ee eh ez he hh hz
 * stored into hist [6][nrows][res] */
typedef struct {
    int res, nrows, ndata;
    double multe, multh;      /* zdt / (eps0 zdx) / corr, zdt / (mu0 zdx) / corr */
    double m0, m1;            /* Mur factors at the two ends */
    double st;                /* sin(theta) */
    double inj_e_s, inj_e_c;  /* zdt / zdx / (eps0 eps[10]) sin(psi), cos(psi) */
    double inj_h_s, inj_h_c;  /* zdt / zdx / mu0 sin(psi), cos(psi) */
    double cez, chz;          /* 2 sqrt(mu0 / eps0) sin(theta), 2 sqrt(eps0 / mu0) sin(theta) */
    const double *eps, *mu;   /* relative material per cell */
    const double *wave;       /* waveform, one sample per substep (shifted by 12) */
    double *hist;
} GsvLtsfLine;

/* layered TSF face corrections of one axis (faces 0 / 1 = near / far) for E
 * (hside 0) or H (hside 1): per face cell the lateral delay q (steps) and the
 * coefficient (E: dt / dx x profile / eps of the cell, H: dt / dx / mu0 x profile) */
typedef struct {
    int i0, j0, k0, i1, j1, k1;
    int skip[6];              /* i0 in j0 jn k0 kn */
    int res, nrows;
    double sph, cph;
    const double *hist;
    const double *q[2][3];    /* [hside][axis]: [face][a][b] */
    const double *coef[2][3];
} GsvLtsf;

/* periodic far field: the z0 / zn planes of the box, (pimax - pimin) x
 * (pjmax - pjmin) images shifted by the box width; post != 0 bins only steps
 * >= ppstart with the postprocess mode's operation order */
typedef struct {
    int i0, j0, k0, i1, j1, k1;
    int ix, iy;
    int pimin, pimax, pjmin, pjmax;
    int skip[2][5];           /* z0, zn: on, imin, jmin, imax, jmax */
    long long nwork;          /* images x 2 planes x ix x iy */
    int nrs;
    const int *ri, *rj, *rk, *istart;
    int ndata;
    double *bins;             /* [nrs][3][ndata] */
    double dx, dy, dt, c;
    int post, ppstart;
} GsvPFarfield;

/* running DFT: each step the four tangential components of every face cell
 * (face 0..5 = x0, xn, y0, yn, z0, zn of the NFFF box; face 6 = a FLUX_DFT
 * plane cell: Ex, Ey and the means of Hx, Hy over the cell and the cell
 * below)
 * are stored in FP32 into slot `step % T` of a sample ring; every T steps (and
 * before results are read) the pending samples are folded into the double
 * accumulators, so the accumulators move once per T steps instead of every step */
typedef struct {
    long long ncells;
    const long long *cell;    /* cell index of each face cell */
    const int *face;          /* 0/1 x faces, 2/3 y faces, 4/5 z faces, 6 flux plane */
    int nw;
    const double *omega;
    double *acc;              /* [ncells][nw][8] */
    double dt;
    int T;                    /* ring slots */
    float *ring;              /* [T][4][ncells] */
} GsvDft;

/* one OUT_FORCE box: per step the stress tensor summed over the six faces,
 * src3d/output.c:355-512; per face cell its absolute eps and mu */
typedef struct {
    int i0, j0, k0, i1, j1, k1;
    long long ncells;             /* 2 (ny nz + nx nz + nx ny) face cells */
    const double2 *em;            /* eps, mu per face cell in the kernel's cell order */
    double *out;                  /* [step][3] */
} GsvForce;

/* one OUT_SUM box: per step the sum (or maximum) of selected E components
 * squared over the box, src3d/output.c:259-306 */
typedef struct {
    int i0, j0, k0, ni, nj, nk;   /* box origin and extents */
    int mode;                     /* SvSumType */
    const float *w;               /* weight per box cell (i fastest): 0 = excluded voxel,
                                   * sigma for SV_SUM_ABS, else 1; NULL = 1 everywhere */
    double *out;                  /* this box's value per step */
} GsvBoxSum;

/* batched stepping of independent grids, one launch per kernel type for all
 * of them: a volume tile is one block of a grid's columns (tx, ty) over z
 * chunk kz; a face tile is one 64 x 4 block (ta, tb) of a TSF face
 * (face = 2 * axis + far) */
typedef struct {
    uint16_t run, tx, ty, kz;
} GsvTile;

typedef struct {
    uint16_t run, face, ta, tb;
} GsvFaceTile;

#ifdef __cplusplus
extern "C" {
#endif
hipError_t gsv_launch_hstep(const GsvGrid *g, int idbytes, int bx, int by, int kchunk, hipStream_t st);
hipError_t gsv_launch_estep(const GsvGrid *g, int idbytes, int bx, int by, int kchunk, hipStream_t st);
hipError_t gsv_launch_tsf_line(const GsvTsf *t, int step, hipStream_t st);
hipError_t gsv_launch_tsff_line(const GsvTsff *t, int step, hipStream_t st);
hipError_t gsv_launch_ext(const GsvGrid *g, const GsvExt *x, int hside, hipStream_t st);
hipError_t gsv_launch_tsff_faces(const GsvGrid *g, const GsvTsff *t, int hside, hipStream_t st);
/* scattered-field source, src3d/yee.c:1139-1237: the listed PEC-rule cells
 * (i, j, k, mask 1 Ex / 2 Ey / 4 Ez) take minus the incident field of line t */
hipError_t gsv_launch_sf_pec(const GsvGrid *g, const GsvTsf *t, const int4 *cells, int n, hipStream_t st);
hipError_t gsv_launch_tsf_faces(const GsvGrid *g, const GsvTsf *t, int hside, hipStream_t st);
hipError_t gsv_launch_point_src(const GsvGrid *g, const GsvPointSrc *s, int step, int hside, hipStream_t st);
hipError_t gsv_launch_record(const GsvGrid *g, const long long *pos, int n, float *rec, hipStream_t st);
hipError_t gsv_launch_liao_store(const GsvGrid *g, const GsvLiao *L, hipStream_t st);
hipError_t gsv_launch_liao_faces(const GsvGrid *g, const GsvLiao *L, hipStream_t st);
hipError_t gsv_launch_zero_planes(const GsvGrid *g, int axis, hipStream_t st);
hipError_t gsv_launch_periodic(const GsvGrid *g, const GsvPeriodic *p, int hside, hipStream_t st);
hipError_t gsv_launch_farfield_td(const GsvGrid *g, const GsvFarfieldTD *f, int step, hipStream_t st);
hipError_t gsv_launch_ltsf_line(const GsvLtsfLine *l, hipStream_t st);
hipError_t gsv_launch_ltsf_faces(const GsvGrid *g, const GsvLtsf *l, int hside, int step, hipStream_t st);
hipError_t gsv_launch_force(const GsvGrid *g, const GsvForce *f, int step, hipStream_t st);
hipError_t gsv_launch_pfarfield(const GsvGrid *g, const GsvPFarfield *f, int step, hipStream_t st);
hipError_t gsv_launch_dft_sample(const GsvGrid *g, const GsvDft *d, int slot, hipStream_t st);
hipError_t gsv_launch_dft_flush(const GsvDft *d, int first_step, int nsamples, hipStream_t st);
hipError_t gsv_launch_box_sum(const GsvGrid *g, const GsvBoxSum *b, int step, hipStream_t st);
/* batched launches: grids / tsfs / tiles / runs / steps are device arrays;
 * nblocks 0 launches one workgroup per tile, otherwise nblocks workgroups
 * stride over the tile list */
hipError_t gsv_launch_hstep_batch(const GsvGrid *grids, const GsvTile *tiles, int ntiles, int idbytes, int per, int bx,
                                  int by, int kchunk, int nblocks, hipStream_t st);
hipError_t gsv_launch_estep_batch(const GsvGrid *grids, const GsvTile *tiles, int ntiles, int idbytes, int ade, int per,
                                  int bx, int by, int kchunk, int nblocks, hipStream_t st);
hipError_t gsv_launch_tsf_line_batch(const GsvTsf *tsfs, const int *runs, int nruns, int *steps, hipStream_t st);
hipError_t gsv_launch_tsf_faces_batch(const GsvGrid *grids, const GsvTsf *tsfs, const GsvFaceTile *tiles, int ntiles,
                                      int hside, hipStream_t st);
#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
/* face cell q of a force box: faces x0, xn (a = j, b = k), y0, yn (a = i,
 * b = k), z0, zn (a = i, b = j); the normal points into the box */
static inline __host__ __device__ void
force_cell_of(const GsvForce &f, long long q, int &i, int &j, int &k, int &nx, int &ny, int &nz)
{
    const long long sx = (long long)(f.j1 - f.j0) * (f.k1 - f.k0), sy = (long long)(f.i1 - f.i0) * (f.k1 - f.k0);
    const long long sz = (long long)(f.i1 - f.i0) * (f.j1 - f.j0);
    nx = ny = nz = 0;
    if (q < 2 * sx) {
        const int far = q >= sx;
        const long long r = q - far * sx;
        i = far ? f.i1 : f.i0; j = f.j0 + (int)(r / (f.k1 - f.k0)); k = f.k0 + (int)(r % (f.k1 - f.k0));
        nx = far ? -1 : 1;
    } else if (q < 2 * sx + 2 * sy) {
        const long long r0 = q - 2 * sx;
        const int far = r0 >= sy;
        const long long r = r0 - far * sy;
        i = f.i0 + (int)(r / (f.k1 - f.k0)); j = far ? f.j1 : f.j0; k = f.k0 + (int)(r % (f.k1 - f.k0));
        ny = far ? -1 : 1;
    } else {
        const long long r0 = q - 2 * sx - 2 * sy;
        const int far = r0 >= sz;
        const long long r = r0 - far * sz;
        i = f.i0 + (int)(r / (f.j1 - f.j0)); j = f.j0 + (int)(r % (f.j1 - f.j0)); k = far ? f.k1 : f.k0;
        nz = far ? -1 : 1;
    }
}
#endif

#endif
