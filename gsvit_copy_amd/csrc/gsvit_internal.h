/*
 * gsvit_internal.h -- declarations shared by the host-side modules of the
 * MI355X engine (parameter reader, planner, log writers, device driver).
 * Not part of the public C-ABI (include/ holds that).
 */
#ifndef GSVIT_INTERNAL_H
#define GSVIT_INTERNAL_H

#include <stdio.h>
#include "gsvit_settings.h"
#include "gsvit_pool.h"

#ifdef __cplusplus
extern "C" {
#endif

/* settings.c */
int sv_generate_waveform(int type, int nsteps, double wavelength, int pulse_width, double dt,
                         double amplitude, double *vals);
double sv_round_g6(double v);
/* external plane source: the three planes of one family (0: ex ey ez, 1: hx hy
 * hz) for `step`, read from <base>_%.4d of step - shift into out[3][extyres][extxres];
 * 1 (with a message) when a file is missing or short, as the reference skips that step */
int sv_ext_read_planes(const SvSet *set, int family, int step, float *out);
int sv_generate_waveform_ltsf(int type, int nsteps, double wavelength, int pulse_width, double dt,
                              double amplitude, double *vals);

/* plan.c */
int sv_material_is_tabulated(int type);   /* LINTAB, DRUDE, PLRC, ADE, CP, CP3 */
int sv_material_is_dispersive(int type);  /* DRUDE, CP3, CP, ADE, PLRC */
/* per-voxel epsilon/sigma as seen by the E update, src3d/yee.c:1317-1350 */
int sv_pool_epsilon_sigma(const SvPool *mp, int i, int j, int k, double *eps, double *sig);
void sv_pool_cabs_vjp(const SvPool *mp, int i, int j, int k, const double gca[3], const double gcb[3], int b_i0, int b_j0,
                      int b_k0, int bnx, int bny, int bnz, double *grad);
/* per-voxel mu/sigast as seen by the H update, src3d/yee.c:1615-1645 */
int sv_pool_mu_sigast(const SvPool *mp, int i, int j, int k, double *mu, double *sigast);
/* component-wise averaged E coefficients, src3d/yee.c:1396-1492 */
void sv_pool_get_cabs(const SvPool *mp, int i, int j, int k,
                      double *cax, double *cay, double *caz, double *cbx, double *cby, double *cbz);
/* dispersive-neighbour flags, src3d/yee.c:453-526; returns the material index used */
int sv_pool_get_tbc(const SvPool *mp, int i, int j, int k, int *xtbc, int *ytbc, int *ztbc);
/* PEC test of a tabulated voxel, src3d/yee.c:1250-1263 */
int sv_pool_is_pec(const SvPool *mp, int i, int j, int k);
/* TSF plane-wave geometry, src3d/source.c:53-165 */
double sv_dcomp(int i, int j, int k, double theta, double phi, int i0, int i1, int j0, int j1, int k0, int k1);
double sv_gex(double f, double theta, double phi, double psi);
double sv_gey(double f, double theta, double phi, double psi);
double sv_gez(double f, double theta, double phi, double psi);
double sv_ghx(double f, double theta, double phi, double psi);
double sv_ghy(double f, double theta, double phi, double psi);
double sv_ghz(double f, double theta, double phi, double psi);
/* TSF face test with the skip-near-material map, src3d/source.c:253-264 */
int sv_tsf_face_active(const SvPool *mp, const SvSet *set, int i, int j, int k);
/* far-field geometry helpers, src3d/farfield.c (dist, angle) */
double sv_ff_dist(int x, int y, int z, int i, int j, int k);
double sv_ff_angle(int i, int j, int k, int x, int y, int z, int face);
/* FLUX_DFT spectra [plane][wavelength] from the accumulators (output.c) */
void sv_flux_spectra(const SvPool *mp, const SvSet *set, double *out);
/* AllFFT point spectrum: out[k] = |DFT of the detrended trace|^2 / n, k < n/2 */
void sv_point_spectrum(const double *x, int n, double *out);
/* the MEDIUM_VECTOR file to read: the file itself, or an anonymous temporary
 * file holding its rewrite when it names optical-database materials (type
 * 99, database.c); NULL if it cannot be opened or resolved */
FILE *sv_open_vector_file(const SvSet *set, const char *fname);
/* a non-dispersive optical-database entry as relative eps, mu and sigma,
 * sigma* (as a vector-file entry resolves it); 1 on error */
int sv_database_linear(const SvSet *set, const char *name, double *eps, double *mu, double *sigma, double *sigast);

#ifdef __cplusplus
}
#endif

#endif
