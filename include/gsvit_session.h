/*
 * gsvit_session.h -- plain C-ABI session over the drop-in GPU entry points.
 *
 * No counterpart in the reference: GSvit drives the entry points from its own
 * main loop (src3d/svit.c:203-257).  This header is for callers that only
 * have a parameter file: the Python bindings, the bench and the tests.  All
 * arguments are plain pointers and sizes; arrays are caller-allocated.
 */
#ifndef GSVIT_SESSION_H
#define GSVIT_SESSION_H

#ifdef __cplusplus
extern "C" {
#endif

/* parse + plan + sv_gpu_new + sv_pool_gpu_init + sv_pool_gpu_copyto(SV_SYNC_ALL);
 * relative file names inside the parameter file resolve against the current
 * directory, as in the reference.  NULL on error (message on stderr). */
void *gsv_open(const char *parfile);
/* run up to nsteps steps of the reference GPU step order; returns the number
 * of steps done so far, or -1 on a HIP error */
int gsv_run(void *session, int nsteps);
/* write the point / far-field logs on output steps, as src3d/svit.c:253-255 */
void gsv_set_write_outputs(void *session, int on);
int gsv_sync(void *session);
/* info[0..16]: nx ny nz nsteps npts nrs ndata nmat step_act,
 * E classes, E id bytes, H classes, H id bytes, ADE, ADE poles, threads per block, z chunk */
void gsv_info(void *session, int *info);
/* one field component (0..5 = Ex..Hz) as float, [i][j][k] order (k fastest) */
int gsv_get_field(void *session, int comp, float *out);
/* point traces as [npts][nsteps+1][6] */
int gsv_get_points(void *session, double *out);
/* time-domain far-field bins as [nrs][3][ndata] */
int gsv_get_farfield(void *session, double *out);
/* periodic far field (PERIODIC_NFFF, src3d/pfarfield.c): dims[0..1] = number of
 * points, bins per component; bins as [nrs][3][ndata] (Ex, Ey, Ez; the log file
 * holds them negated, src3d/output.c:1347-1352) */
/* OUT_FORCE values as [nforces][nsteps+1][3] (before the log's dx dy factor);
 * returns the element count (out may be NULL, to size the buffer), -1 on error */
long gsv_get_forces(void *session, double *out);
void gsv_pfarfield_dims(void *session, int *dims);
int gsv_get_pfarfield(void *session, double *out);
/* running-DFT accumulators; returns the element count (out may be NULL) */
long gsv_get_dft(void *session, double *out);
/* dims[0..1]: running-DFT face cells, wavelengths (NFFF_DFT) */
void gsv_dft_dims(void *session, int *dims);
/* far-field pattern of the running DFT (no reference counterpart; the
 * frequency-domain counterpart of the time-domain far field of
 * src3d/farfield.c:174-407): dirs [ndirs][2] = theta, phi (radians, host
 * memory), out [ndirs][wavelengths][4] = re, im of F_theta and F_phi (host
 * memory), E = F exp(-i k r) / r with r from the grid origin.  Surface
 * equivalence on the NFFF box faces (J = n x H, M = E x n), NFFF_SKIP
 * rectangles left out.  Device buffers are allocated for the call.  Returns 0
 * on success. */
int gsv_farfield_dft(void *session, int ndirs, const double *dirs, double *out);
/* gsv_farfield_dft without any allocation -- no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  dirs
 * and out are DEVICE memory of the session's GPU, same layouts.  The partial
 * sums use a fixed per-GPU scratch array of 16 MiB, so ndirs x wavelengths
 * must not exceed 174762 (larger calls return 1 with a message).  Calls on
 * one GPU share that array: calls that overlap in time on one GPU race.
 * Synchronises the session's stream before returning; 0 on success. */
int gsv_farfield_dft_noalloc(void *session, int ndirs, const double *dirs, double *out);
/* OUT_SUM values as [nsums][nsteps] (before the cell-volume scaling of the
 * log file); returns the element count (out may be NULL), -1 on error */
long gsv_get_sums(void *session, double *out);
/* FLUX_DFT spectra (no reference counterpart) as [nplanes][nwavelengths] into
 * flux and, when acc is not NULL, the raw accumulators [cell][wavelength][8]
 * (cells plane, j, i order; Ex, Ey, mean Hx, mean Hy as re, im); returns the
 * accumulator count, -1 on error */
long gsv_get_flux(void *session, double *flux, double *acc);
/* dims[0..2]: FLUX_DFT planes, wavelengths, cells per plane (0 without FLUX_DFT) */
void gsv_flux_dims(void *session, int *dims);
/* planner maps, [i][j][k]: tabulated material index (-1 without a table) and internal-boundary flags */
void gsv_get_maps(void *session, int *mat, int *bnds);
/* per-launch HIP-event timing of the H and E volume kernels */
void gsv_set_timing(void *session, int on);
/* out[0..3]: H total ms, H launches, E total ms, E launches */
void gsv_kernel_times(void *session, double *out);
/* launch shape of the volume kernels: shape[0..5] = H tile x width, H rows,
 * H z-march depth, E tile x width, E rows, E z-march depth (x width * rows <= 512) */
int gsv_set_launch(void *session, const int *shape);
void gsv_get_launch(void *session, int *shape);
/* pick the fastest launch shape for each volume kernel by timing candidate
 * shapes on real steps (steps_per_candidate each); writes the chosen shape to
 * best[0..5] and returns the number of steps consumed, -1 on error. */
int gsv_autotune(void *session, int steps_per_candidate, int *best);
/* Gradient extension (no reference counterpart).  The parameter file
 * declares DESIGN_BOX i0 j0 k0 in jn kn (voxels with per-voxel permittivity
 * from MEDIUM_LINEAR) and OBJECTIVE_DFT i0 j0 k0 in jn kn lambda, the
 * objective J = sum over the box of |DFT_lambda(E)|^2 over the whole run.
 * Design arrays are [i][j][k] (k fastest) over the design box. */
void gsv_design_dims(void *session, int *dims);
/* allocate the gradient state: the whole run's design-region history is kept
 * when it fits history_budget_bytes, otherwise full checkpoints every
 * ckpt_interval steps with recomputation */
int gsv_adjoint_setup(void *session, long long history_budget_bytes, int ckpt_interval);
int gsv_set_design_eps(void *session, const float *eps_r);
int gsv_get_design_eps(void *session, float *eps_r);
/* fresh forward run over all steps; *J = objective */
int gsv_objective(void *session, double *J);
/* after gsv_objective: grad = dJ / d eps_r per design voxel */
int gsv_objective_gradient(void *session, float *grad);
/* out[0..4]: forward ms, backward ms, whole history kept, checkpoint interval, checkpoints */
void gsv_adjoint_info(void *session, double *out);
void gsv_close(void *session);
/* Batched runs (no reference counterpart: GSvit runs one simulation per
 * process).  A batch steps n open sessions together, one launch per kernel
 * type for all of them (runs of uneven grid sizes, e.g. an incidence-angle,
 * wavelength or roughness sweep); each run's results equal those of stepping
 * it alone.  A batched session has CPML (or no absorbing boundary) and at
 * most a TSF source, no point sources, point or sum outputs, far field, DFT,
 * design region or adjoint, and its output writing off; the sessions stay
 * open (and are not stepped on their own) while the batch exists, and their
 * fields are read with gsv_get_field.  NULL on error (message on stderr). */
void *gsv_batch_open(void *const *sessions, int n);
/* run nsteps steps of every session (fewer if a session reaches its step
 * count); returns the number of steps run, -1 on error */
int gsv_batch_run(void *batch, int nsteps);
/* shape[0..7]: H x width, rows, z chunk, E x width, rows, z chunk, H and E
 * workgroups (0: one per tile, otherwise that many stride over the tiles);
 * x widths are multiples of 64 */
int gsv_batch_set_launch(void *batch, const int *shape);
void gsv_batch_get_launch(void *batch, int *shape);
/* capture steps_per_graph steps as a HIP graph and replay it (0: launch
 * every kernel from the host) */
int gsv_batch_set_graph(void *batch, int steps_per_graph);
/* as gsv_autotune over the batch's tile lists, then the workgroup count;
 * writes best[0..7] and returns the number of steps consumed, -1 on error */
int gsv_batch_autotune(void *batch, int steps_per_candidate, int *best);
/* out[0..7]: mean H ms, timed H launches, mean E ms, timed E launches, H tiles, E tiles, TSF face tiles, TSF runs */
void gsv_batch_set_timing(void *batch, int on);
void gsv_batch_info(void *batch, double *out);
void gsv_batch_close(void *batch);

#ifdef __cplusplus
}
#endif

#endif
