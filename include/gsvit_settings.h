/*
 * gsvit_settings.h -- parameter-file settings (SvSet) for the MI355X FDTD engine.
 *
 * Field names and enum values follow GSvit's own settings structure so that
 * code written against the reference reads the same here:
 *   enums        src3d/settings.h:38-176  (SV_BOUNDARY_*, SV_COMP_*, SV_MAT_*, SV_MATMODE_*)
 *   SvSet        src3d/settings.h:866-878 (sp, sc, sb, smb, ss, so, sf, sm, plan)
 *   SvMatProp    src3d/settings.h:921-964
 * Only the members that the 3-D time-stepping path reads are present.
 * The parser that fills this structure lives in gsvit_copy_amd/csrc/settings.c.
 */
#ifndef GSVIT_SETTINGS_H
#define GSVIT_SETTINGS_H

#ifdef __cplusplus
extern "C" {
#endif

#define SV_MAX_GPUS 16
#define SV_NAME_LEN 512

/* physical constants, src3d/constants.h:28-35 */
#define SV_LIGHT_SPEED 299792458.0
#define SV_MU_0 1.256637061435917295e-6
#define SV_EPSILON_0 8.854187817620389850e-12
#define SV_EV_TO_J 1.60217646e-19

typedef enum {
    SV_BOUNDARY_NONE = 0,
    SV_BOUNDARY_PEC = 1,
    SV_BOUNDARY_LIAO = 2,
    SV_BOUNDARY_CPML = 3,
    SV_BOUNDARY_PERIODIC = 4
} SvBoundaryType;

typedef enum {
    SV_COMP_EX = 0,
    SV_COMP_EY = 1,
    SV_COMP_EZ = 2,
    SV_COMP_HX = 3,
    SV_COMP_HY = 4,
    SV_COMP_HZ = 5,
    SV_COMP_ALL = 6,
    SV_COMP_ALLFFT = 21   /* point spectra of Ex, Ey, Ez at the last step (src3d/settings.h:69) */
} SvComponent;

typedef enum {
    SV_MAT_LINEAR = 0,
    SV_MAT_LINTAB = 1,
    SV_MAT_DRUDE = 2,
    SV_MAT_CP3 = 3,
    SV_MAT_CP = 4,
    SV_MAT_ADE = 5,
    SV_MAT_PLRC = 6,
    /* no reference counterpart: ADE with a third critical-point pole (e.g. a
     * Lorentz term, which is a critical point with phi = 0); a type-7 line reads
     * as type 5 plus four more numbers and is stored as an SV_MAT_ADE material
     * with ade_npoles = 3 */
    SV_MAT_ADE3 = 7,
    SV_MAT_PEC = 10
} SvMatType;

typedef enum {
    SV_MATMODE_NONE = 0,
    SV_MATMODE_ELECTRIC = 1,
    SV_MATMODE_MAGNETIC = 2,
    SV_MATMODE_FULL = 3,
    SV_MATMODE_ERROR = 999
} SvMatMode;

typedef enum {
    SV_GPUMODE_NONE = 0,
    SV_GPUMODE_FULL = 1
} SvGpuMode;

/* geometry entities of a MEDIUM_VECTOR file, src3d/plan.h:41-51 (SV_ENTITY_*) */
typedef enum {
    SV_ENTITY_SPHERE = 4,
    SV_ENTITY_CYLINDER = 7,
    SV_ENTITY_VOXEL = 8,
    SV_ENTITY_TETRAHEDRON = 12,
    SV_ENTITY_TETGEN = 21,
    SV_ENTITY_CONE = 107,
    SV_ENTITY_RCONE = 108
} SvEntityType;

/* material description, src3d/settings.h:921-964 */
typedef struct {
    int type;
    double epsilon;
    double mu;
    double sigma;
    double sigast;
    double drude_omega_p;
    double drude_nu;
    double cp3_a[3];
    double cp3_phi[3];
    double cp3_omega[3];
    double cp3_gamma[3];
    double ade_a0;
    double ade_a1;
    double ade_a2;
    double ade_bp0[3];
    double ade_bp1[3];
    double ade_bp2[3];
    double ade_bp3[3];
    double ade_bp4[3];
    int ade_npoles;       /* critical-point poles of an ADE material: 2 (type 5) or 3 (type 7) */
    double ade_c0;
    double ade_c1;
    double ade_c2;
    double ade_c3;
    /* piecewise-linear recursive convolution (type 6) susceptibility
     * integrals: Drude term and two critical-point poles (real and imaginary
     * parts of the increments), src3d/plan.c:2765-2824 */
    double plrc_d_chi, plrc_d_xi, plrc_d_dchi, plrc_d_dxi;
    double plrc_p_chi[2], plrc_p_xi[2];
    double plrc_p_dchir[2], plrc_p_dxir[2], plrc_p_dchii[2], plrc_p_dxii[2];
    int pos;
} SvMatProp;

typedef struct {
    int xres, yres, zres;
    double dx, dy, dz;
} SvSetPool;

typedef struct {
    int nsteps;
    int step_act;
    int verbose;
    double dtmult;
    int usegpu;
    int ugpu[SV_MAX_GPUS];
    int devicequery;
    int nthreads;
    char killer[SV_NAME_LEN];
} SvSetComp;

typedef struct {
    int bx0, bxn, by0, byn, bz0, bzn;
    int depth_bx0, depth_bxn, depth_by0, depth_byn, depth_bz0, depth_bzn;
    int m_bx0, m_bxn, m_by0, m_byn, m_bz0, m_bzn;
    double sigma_bx0, sigma_bxn, sigma_by0, sigma_byn, sigma_bz0, sigma_bzn;
    double a_bx0, a_bxn, a_by0, a_byn, a_bz0, a_bzn;
    double kappa_bx0, kappa_bxn, kappa_by0, kappa_byn, kappa_bz0, kappa_bzn;
    /* ADE_CPML_CONSISTENT 1 (no reference counterpart): inside CPML slabs the
     * psi of a component that takes the dispersive (ADE) update is updated
     * once per step and applied with that update's own curl coefficient, on
     * all six faces.  0 (default): the reference's scheme, whose psi of such
     * components is updated twice per step and applied with the coefficient
     * of the material's eps/sigma, skipping x0 (src3d/yee.c:950-1057,
     * src3d/boundary.c:691-887) -- a dispersive metal in the slabs then grows
     * without bound after some thousand steps. */
    int ade_cpml_consistent;
} SvSetBoundary;

typedef struct {
    int bx0, bxn, by0, byn, bz0, bzn;
    int bx0pos, bxnpos, by0pos, bynpos, bz0pos, bznpos;
} SvSetMediumBoundary;

typedef struct {
    int point_origin_position_i;
    int point_origin_position_j;
    int point_origin_position_k;
    int source_mode;              /* 0 = file, 1 = sine, 2 = pulse */
    double source_wl;
    double source_amplitude;
    int source_pulsewidth;
    double point_origin_theta;
    double point_origin_phi;
    char source_filename[SV_NAME_LEN];
} SvSetSourcePoint;

typedef struct {
    int valid;
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    double ia_theta, ia_phi, ia_psi;
    int source_mode;
    double source_wl;
    double source_amplitude;
    int source_pulsewidth;
    char source_filename[SV_NAME_LEN];
    double layered_epsilon, layered_mu, layered_sigma, layered_sigast;
    int box_boundary_skipi0, box_boundary_skipj0, box_boundary_skipk0;
    int box_boundary_skipin, box_boundary_skipjn, box_boundary_skipkn;
    int box_boundary_skipdepth;
    /* beam profiles of the face corrections, TSF_GAUSSIAN_Z and TSF_FIBER_Z
     * (src3d/settings.c:4308-4339, src3d/source.c:231-251) */
    int gaussian;
    double gaussian_fxpos, gaussian_fypos, gaussian_rx, gaussian_ry;
    int fiber;
    double fiber_fxpos, fiber_fypos, fiber_radius, fiber_cutoff, fiber_epsilon_core, fiber_epsilon_cladding;
} SvSetSourceTSF;

/* layered total/scattered-field source, src3d/settings.h (SvSetSourceLTSF)
 * and src3d/settings.c:4342-4480: a plane wave precomputed by a split-field
 * 1-D model along z with layers starting at layered_zpos (cells) */
#define SV_LTSF_MAX_LAYERS 100
typedef struct {
    int valid;
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    double ia_theta, ia_phi, ia_psi;
    int layered_count;
    int layered_zpos[SV_LTSF_MAX_LAYERS];
    double layered_epsilon[SV_LTSF_MAX_LAYERS], layered_mu[SV_LTSF_MAX_LAYERS];
    double layered_sigma[SV_LTSF_MAX_LAYERS], layered_sigast[SV_LTSF_MAX_LAYERS];
    /* a layer given as 99 and an optical-database entry: its name (empty otherwise) */
    char layered_material[SV_LTSF_MAX_LAYERS][SV_NAME_LEN];
    int source_mode;
    char source_filename[SV_NAME_LEN];
    /* generated waveform (source_mode 1 sine, 2 pulse), sampled at dt/3 */
    double source_wl;
    double source_amplitude;
    int source_pulsewidth;
    int box_boundary_skipi0, box_boundary_skipj0, box_boundary_skipk0;
    int box_boundary_skipin, box_boundary_skipjn, box_boundary_skipkn;
    int box_boundary_skipdepth;
    int gaussian;
    double gaussian_fxpos, gaussian_fypos, gaussian_rx, gaussian_ry;
    int fiber;
    double fiber_fxpos, fiber_fypos, fiber_radius, fiber_cutoff, fiber_epsilon_core, fiber_epsilon_cladding;
} SvSetSourceLTSF;

/* focused TSF, src3d/settings.c:4605-4697 (SOURCE_TSFF i0 j0 k0 in jn kn
 * thetamax fdist pol nip mip mode file-or-parameters, TSFF_SHIFT, TSFF_SKIP):
 * nip x mip plane waves over the aperture cone theta <= thetamax, their phase
 * origin at the pool centre plus the shift (cells) */
typedef struct {
    SvSetSourceTSF w;         /* box, skipped faces and waveform; the angles are unused */
    double focused_thetamax, focused_fdist, focused_pol;
    int focused_nip, focused_mip;
    double xshift, yshift, zshift;
} SvSetSourceTSFF;

typedef struct {
    int npnts;
    SvSetSourcePoint *pnts;
    SvSetSourceTSF tsf;
    SvSetSourceTSFF tsff;
    /* external plane source, src3d/settings.c:4147-4186 (SOURCE_EXT i j k
     * ijstart jkstart extxres extyres iextfrom jextfrom iextto jextto shift,
     * then the six file bases): per step the text planes <base>_%.4d of step
     * - shift (extyres rows of extxres values) drive the z plane k; only
     * i = j = -1 is a mode the reference applies (src3d/source.c:1377, 2251) */
    struct {
        int valid;
        int i, j, k, ijstart, jkstart, extxres, extyres, iextfrom, jextfrom, iextto, jextto, shift;
        char filebase[6][SV_NAME_LEN];   /* ex ey ez hx hy hz */
    } ext;
    /* scattered-field source, src3d/settings.c:4698-4760 (SOURCE_SF theta phi
     * psi mode file-or-parameters); the box fields are unused */
    SvSetSourceTSF sf;
    SvSetSourceLTSF ltsf;
    double lambda_min, lambda_center, lambda_max;
} SvSetSource;

typedef struct {
    int component;
    int step;
    int i, j, k;
    int start, stop;        /* AllFFT: the steps the spectrum covers */
    char filebase[SV_NAME_LEN];
} SvOutputPar;

/* sum modes, src3d/settings.h:94-99 */
typedef enum {
    SV_SUM_EX = 0,          /* sum of Ex^2 over the box */
    SV_SUM_EY = 1,
    SV_SUM_EZ = 2,
    SV_SUM_ALL = 3,         /* sum of |E|^2 */
    SV_SUM_ABS = 4,         /* sum of sigma |E|^2 */
    SV_SUM_MAX = 5          /* max of |E|^2 */
} SvSumType;

/* OUT_SUM, src3d/settings.h:596-610 */
typedef struct {
    int component;          /* SvSumType */
    int step;               /* log written every `step` steps (0 = never) */
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    double layered_epsilon; /* absolute permittivity a voxel must match within EPSILON_0/100; -1 = every voxel */
    double layered_mu, layered_sigma, layered_sigast;
    char filename[SV_NAME_LEN];
    /* OUT_SUMTAB (src3d/settings.c:4982-5017): the voxel filter is the
     * optical-database material `string`, resolved at the centre wavelength
     * when the pool is planned (src3d/plan.c:5475-5486) */
    int stringbased;
    char string[SV_NAME_LEN];
} SvOutputSum;

/* OUT_VOLUME components, src3d/settings.h:73-90 */
typedef enum {
    SV_OVOLUME_EX = 0,
    SV_OVOLUME_EY = 1,
    SV_OVOLUME_EZ = 2,
    SV_OVOLUME_HX = 3,
    SV_OVOLUME_HY = 4,
    SV_OVOLUME_HZ = 5,
    SV_OVOLUME_ALL = 6,        /* |E| */
    SV_OVOLUME_EPSILON = 7,    /* relative */
    SV_OVOLUME_SIGMA = 8,
    SV_OVOLUME_MU = 9,         /* relative */
    SV_OVOLUME_SIGAST = 10,
    SV_OVOLUME_MAT = 11,       /* material index */
    SV_OVOLUME_MATTYPE = 12,   /* material type */
    SV_OVOLUME_ABS = 13,       /* sum over output steps of sigma |E|^2 */
    SV_OVOLUME_SUMALL = 14,    /* sum over output steps of |E|^2 */
    SV_OVOLUME_MAXALL = 15     /* maximum over output steps of |E|^2 */
} SvOutputVolumeType;

/* OUT_PLANE and OUT_VOLUME, src3d/settings.c:4869-4936 */
typedef struct {
    int component;          /* plane: SV_COMP_EX..HZ; volume: SvOutputVolumeType */
    int step;               /* every `step` steps (and the last step) */
    int start, stop;        /* step window */
    int format;             /* plane: 0 raw doubles, else ASCII; volume: 0 raw doubles, 1 ASCII, 2 VTK */
    int i, j, k;            /* plane: the fixed index, -1 on the two in-plane axes;
                             * volume: k = accumulator slot of ABS / SUMALL / MAXALL */
    char filebase[SV_NAME_LEN];
} SvOutputRegion;

/* OUT_FORCE, src3d/settings.c:5018-5040: Maxwell stress tensor integrated
 * over the six faces of a box every step, log written every `step` steps */
typedef struct {
    int step;
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    char filename[SV_NAME_LEN];
} SvOutputForce;

typedef struct {
    int npnts;
    SvOutputPar *pnts;
    int nsums;
    SvOutputSum *sums;
    int nforces;
    SvOutputForce *forces;
    int nplns;
    SvOutputRegion *plns;
    int ncubs;
    SvOutputRegion *cubs;
} SvSetOutput;

typedef struct {
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    int sumfrom, sumto;
    int box_boundary_skipi0, box_boundary_skipin, box_boundary_skipj0;
    int box_boundary_skipjn, box_boundary_skipk0, box_boundary_skipkn;
    int skipi0_jmin, skipi0_kmin, skipi0_jmax, skipi0_kmax;
    int skipin_jmin, skipin_kmin, skipin_jmax, skipin_kmax;
    int skipj0_imin, skipj0_kmin, skipj0_imax, skipj0_kmax;
    int skipjn_imin, skipjn_kmin, skipjn_imax, skipjn_kmax;
    int skipk0_imin, skipk0_jmin, skipk0_imax, skipk0_jmax;
    int skipkn_imin, skipkn_jmin, skipkn_imax, skipkn_jmax;
    int nrs;
    int *ri, *rj, *rk;
    int *individual;        /* 0: a single point; s > 0: point of area set s (its bins share one istart) */
    char (*source_filename)[SV_NAME_LEN];  /* "" = no log file */
    int nsets;              /* NFFF_SPHERICAL_AREA / NFFF_PLANAR_AREA sets, counted from 1 */
    /* running-DFT accumulation on the same box (no reference counterpart) */
    int ndft;
    double *dft_lambda;
    char dft_filename[SV_NAME_LEN];
    /* NFFF_DFT_PATTERN ntheta nphi file (no reference counterpart): far-field
     * pattern of the running DFT on ntheta polar angles 0..pi (inclusive) x
     * nphi azimuths 0..2 pi (exclusive), written at the last step */
    int dft_pattern_ntheta, dft_pattern_nphi;
    char dft_pattern_filename[SV_NAME_LEN];
} SvSetFarfield;

typedef struct {
    char in_voxel_filename[SV_NAME_LEN];
    char in_vector_filename[SV_NAME_LEN];
    char in_vtk_data_filename[SV_NAME_LEN];
    int matmode_check;
    /* MEDIUM_CROP, src3d/settings.c:1898-1921: everything on or outside the
     * box faces becomes the crop medium (relative eps, mu; sigma, sigma*) */
    int crop;
    int crop_i0, crop_j0, crop_k0, crop_in, crop_jn, crop_kn;
    double crop_epsilon, crop_mu, crop_sigma, crop_sigast;
} SvSetMat;

typedef struct {
    double dt;
    int matmode;
    int gpumode;
} SvSetPlan;

/* Gradient extension (no reference counterpart): DESIGN_BOX i0 j0 k0 in jn kn
 * marks voxels whose relative permittivity is a design parameter (given per
 * voxel by MEDIUM_LINEAR); OBJECTIVE_DFT i0 j0 k0 in jn kn lambda defines
 * J = sum over the box cells and the three E components of |DFT_lambda(E)|^2,
 * the DFT running over every step with E sampled at (step+1) dt. */
typedef struct {
    int design;
    int design_i0, design_j0, design_k0, design_in, design_jn, design_kn;
    int objective;
    int obj_i0, obj_j0, obj_k0, obj_in, obj_jn, obj_kn;
    double obj_lambda;
} SvSetAdjoint;

/* Flux spectra (no reference counterpart): FLUX_DFT i0 j0 i1 j1 n k_1 .. k_n
 * m lambda_1 .. lambda_m file -- the net z-directed power flux spectrum
 * through the rectangle [i0, i1) x [j0, j1) of each plane k (E grid; H taken
 * as the mean of planes k-1 and k), F = 1/2 Re sum (Ex Hy* - Ey Hx*) dx dy with
 * the DFTs of E at (step+1) dt and of H at (step+1/2) dt times dt.  The
 * difference of two planes is the power absorbed between them, e.g. per
 * layer of a stack. */
typedef struct {
    int i0, j0, i1, j1;
    int nplanes;
    int *k;
    int nw;
    double *lambda;
    char filename[SV_NAME_LEN];
} SvSetFlux;

/* periodic near-to-far-field transform, src3d/settings.h (SvSetPFarfield) and
 * src3d/settings.c:3424-3499: the Kirchhoff integral of E over the z0 and zn
 * planes of the box, repeated over the periodic images pi in [pimin, pimax),
 * pj in [pjmin, pjmax) shifted by the box width; postprocess != 0 stores the
 * plane fields from step ppstart on and integrates them at the last step */
typedef struct {
    int box_i0, box_j0, box_k0, box_in, box_jn, box_kn;
    int pimin, pjmin, pimax, pjmax;
    int postprocess, ppstart;
    int box_boundary_skipk0, box_boundary_skipkn;
    int skipk0_imin, skipk0_jmin, skipk0_imax, skipk0_jmax;
    int skipkn_imin, skipkn_jmin, skipkn_imax, skipkn_jmax;
    int nrs;
    int *ri, *rj, *rk;
    char (*source_filename)[SV_NAME_LEN];  /* "" = no log file */
    int nsets;              /* PERIODIC_NFFF_SPHERICAL_AREA sets, counted from 1 */
} SvSetPFarfield;

typedef struct {
    SvSetPool sp;
    SvSetComp sc;
    SvSetBoundary sb;
    SvSetMediumBoundary smb;
    SvSetSource ss;
    SvSetOutput so;
    SvSetFarfield sf;
    SvSetPFarfield spf;
    SvSetMat sm;
    SvSetPlan plan;
    SvSetAdjoint sa;
    SvSetFlux sx;
} SvSet;

/* src3d/settings.h:967,974 (clear_settings src3d/settings.c:570, parse_settings :3822).
 * clear_settings(set, 1) frees what a parse allocated, then resets the defaults.
 * parse_settings returns 1 on success and 0 on error, like the reference
 * (src3d/settings.c:5831-5836, checked by src3d/svit.c:186); the third
 * argument is accepted for signature compatibility. */
void clear_settings(SvSet *set, int remove);
int parse_settings(const char *filename, SvSet *set, int called_from_gsvit);

#ifdef __cplusplus
}
#endif

#endif /* GSVIT_SETTINGS_H */
