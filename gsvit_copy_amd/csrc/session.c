/*
 * session.c -- plain C-ABI session over the drop-in entry points, for callers
 * that do not keep their own SvSet/SvPool (the Python bindings, the bench and
 * the tests).  A session is parse -> plan -> sv_gpu_new -> sv_pool_gpu_init ->
 * sv_pool_gpu_copyto(SV_SYNC_ALL), then the GPU branch of the reference step
 * loop (src3d/svit.c:239-257) for as many steps as requested.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "gsvit_internal.h"
#include "gsvit_session.h"

/* engine helpers, gpu_engine.hip */
int gsv_engine_sync(void *eng);
void gsv_engine_set_timing(void *eng, int on);
void gsv_engine_kernel_times(void *eng, double *out);
void gsv_engine_info(void *eng, int *info);
int gsv_engine_get_field(void *eng, int comp, float *out);
int gsv_engine_fetch_aux(void *eng);
void gsv_engine_free(void *eng);
int gsv_engine_set_launch(void *eng, const int *shape);
void gsv_engine_get_launch(void *eng, int *shape);
int gsv_engine_adjoint_setup(void *eng, long long history_budget_bytes, int ckpt_interval);
int gsv_engine_adjoint_forward(void *eng, int (*step_fn)(void *), void *ctx, double *J);
int gsv_engine_adjoint_backward(void *eng, int (*step_fn)(void *), void *ctx, float *grad);
int gsv_engine_set_design_eps(void *eng, const float *eps_r);
int gsv_engine_get_design_eps(void *eng, float *eps_r);
void gsv_engine_adjoint_info(void *eng, double *out);
void *gsv_engine_batch_new(void *const *engines, int n);
int gsv_engine_batch_run(void *b, int nsteps, const int *steps);
int gsv_engine_batch_set_launch(void *b, const int *shape);
void gsv_engine_batch_get_launch(void *b, int *shape);
int gsv_engine_batch_set_graph(void *b, int steps_per_graph);
void gsv_engine_batch_set_timing(void *b, int on);
void gsv_engine_batch_info(void *b, double *out);
void gsv_engine_batch_free(void *b);
int gsv_engine_farfield_dft(void *eng, int ndirs, const double *dirs, double *out);

typedef struct {
    SvSet set;
    SvPool *mp;
    int write_outputs;
} GsvSession;

void *
gsv_open(const char *parfile)
{
    GsvSession *s = (GsvSession *)calloc(1, sizeof(GsvSession));
    if (!s)
        return NULL;
    clear_settings(&s->set, 0);
    if (!parse_settings(parfile, &s->set, 1)) {
        fprintf(stderr, "Settings error!\n");
        free(s);
        return NULL;
    }
    s->set.sc.usegpu = 1;
    s->mp = init_and_make_plan(&s->set);
    if (!s->mp) {
        clear_settings(&s->set, 1);
        free(s);
        return NULL;
    }
    s->mp->set = &s->set;
    if (!sv_gpu_new(s->mp, &s->set) || sv_pool_gpu_init(s->mp, &s->set)
        || sv_pool_gpu_copyto(s->mp, &s->set, SV_SYNC_ALL)) {
        gsv_close(s);
        return NULL;
    }
    return s;
}

void
gsv_set_write_outputs(void *h, int on)
{
    ((GsvSession *)h)->write_outputs = on;
}

int
gsv_run(void *h, int nsteps)
{
    GsvSession *s = (GsvSession *)h;
    SvPool *mp = s->mp;
    SvSet *set = &s->set;
    int n, rc = 0;
    for (n = 0; n < nsteps && set->sc.step_act < set->sc.nsteps; n++) {
        rc |= sv_pool_gpu_boundary_copy(mp, set);
        rc |= sv_pool_gpu_ystep_h(mp, set);
        rc |= sv_pool_gpu_apply_source_hstep(mp, set);
        rc |= sv_pool_gpu_boundary_hstep(mp, set);
        rc |= sv_pool_gpu_ystep_e(mp, set);
        rc |= sv_pool_gpu_apply_source_estep(mp, set);
        rc |= sv_pool_gpu_boundary_estep(mp, set);
        rc |= sv_pool_gpu_getsum(mp, set);
        rc |= sv_pool_gpu_apply_farfield(mp, set);
        rc |= sv_pool_gpu_apply_pfarfield(mp, set);
        rc |= sv_pool_gpu_getpoints(mp, set);
        rc |= sv_pool_gpu_getforce(mp, set);
        if (rc)
            return -1;
        if (s->write_outputs && (is_output(mp, set) || set->sc.step_act == (set->sc.nsteps - 1))) {
            if (sv_pool_gpu_copyfrom(mp, set, SV_SYNC_EH))
                return -1;
            sv_pool_output(mp, set);
        }
        set->sc.step_act++;
    }
    return set->sc.step_act;
}

int
gsv_sync(void *h)
{
    GsvSession *s = (GsvSession *)h;
    return gsv_engine_sync(s->mp->gpu);
}

void
gsv_info(void *h, int *info)
{
    GsvSession *s = (GsvSession *)h;
    info[0] = s->set.sp.xres;
    info[1] = s->set.sp.yres;
    info[2] = s->set.sp.zres;
    info[3] = s->set.sc.nsteps;
    info[4] = s->set.so.npnts;
    info[5] = s->set.sf.nrs;
    info[6] = s->mp->farfield->ndata;
    info[7] = s->mp->nmat;
    info[8] = s->set.sc.step_act;
    gsv_engine_info(s->mp->gpu, info + 9);
}

void
gsv_dft_dims(void *h, int *dims)
{
    GsvSession *s = (GsvSession *)h;
    dims[0] = (int)s->mp->farfield->dft_ncells;
    dims[1] = s->mp->farfield->ndft;
}

void *
gsv_session_engine(void *h)
{
    return h ? ((GsvSession *)h)->mp->gpu : NULL;
}

int
gsv_farfield_dft(void *h, int ndirs, const double *dirs, double *out)
{
    return gsv_engine_farfield_dft(((GsvSession *)h)->mp->gpu, ndirs, dirs, out);
}

int
gsv_get_field(void *h, int comp, float *out)
{
    GsvSession *s = (GsvSession *)h;
    return gsv_engine_get_field(s->mp->gpu, comp, out);
}

int
gsv_get_points(void *h, double *out)
{
    GsvSession *s = (GsvSession *)h;
    int n, nst = s->set.sc.nsteps + 1;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return 1;
    for (n = 0; n < s->set.so.npnts; n++)
        memcpy(out + (size_t)n * nst * 6, s->mp->out->outpointdata[n], (size_t)nst * 6 * sizeof(double));
    return 0;
}

int
gsv_get_farfield(void *h, double *out)
{
    GsvSession *s = (GsvSession *)h;
    int n, nd = s->mp->farfield->ndata;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return 1;
    for (n = 0; n < s->set.sf.nrs; n++)
        memcpy(out + (size_t)n * 3 * nd, s->mp->farfield->rpoints[n].ex, 3 * (size_t)nd * sizeof(double));
    return 0;
}

long
gsv_get_forces(void *h, double *out)
{
    GsvSession *s = (GsvSession *)h;
    const long per = 3L * (s->set.sc.nsteps + 1);
    int n;
    if (!out)
        return per * s->set.so.nforces;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return -1;
    for (n = 0; n < s->set.so.nforces; n++)
        memcpy(out + (size_t)n * per, s->mp->out->outforcedata[n], (size_t)per * sizeof(double));
    return per * s->set.so.nforces;
}

void
gsv_pfarfield_dims(void *h, int *dims)
{
    GsvSession *s = (GsvSession *)h;
    dims[0] = s->mp->pfarfield->nrs;
    dims[1] = s->mp->pfarfield->ndata;
}

int
gsv_get_pfarfield(void *h, double *out)
{
    GsvSession *s = (GsvSession *)h;
    int n, nd = s->mp->pfarfield->ndata;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return 1;
    for (n = 0; n < s->mp->pfarfield->nrs; n++)
        memcpy(out + (size_t)n * 3 * nd, s->mp->pfarfield->prpoints[n].ex, 3 * (size_t)nd * sizeof(double));
    return 0;
}

long
gsv_get_dft(void *h, double *out)
{
    GsvSession *s = (GsvSession *)h;
    SvFarfield *ff = s->mp->farfield;
    long cnt = (long)ff->dft_ncells * ff->ndft * 8;
    if (!out)
        return cnt;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return -1;
    memcpy(out, ff->dft_data, (size_t)cnt * sizeof(double));
    return cnt;
}

long
gsv_get_sums(void *h, double *out)
{
    GsvSession *s = (GsvSession *)h;
    int n, nst = s->set.sc.nsteps;
    long cnt = (long)s->set.so.nsums * nst;
    if (!out || !cnt)
        return cnt;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return -1;
    for (n = 0; n < s->set.so.nsums; n++)
        memcpy(out + (size_t)n * nst, s->mp->out->outsumdata[n], (size_t)nst * sizeof(double));
    return cnt;
}

void
gsv_flux_dims(void *h, int *dims)
{
    const SvSetFlux *x = &((GsvSession *)h)->set.sx;
    dims[0] = x->nplanes;
    dims[1] = x->nplanes ? x->nw : 0;
    dims[2] = x->nplanes ? (x->i1 - x->i0) * (x->j1 - x->j0) : 0;
}

long
gsv_get_flux(void *h, double *flux, double *acc)
{
    GsvSession *s = (GsvSession *)h;
    SvFarfield *ff = s->mp->farfield;
    long cnt = ff ? ff->flux_ncells * ff->nflux * 8 : 0;
    if (!cnt || (!flux && !acc))
        return cnt;
    if (gsv_engine_fetch_aux(s->mp->gpu))
        return -1;
    if (flux)
        sv_flux_spectra(s->mp, &s->set, flux);
    if (acc)
        memcpy(acc, ff->flux_data, (size_t)cnt * sizeof(double));
    return cnt;
}

void
gsv_get_maps(void *h, int *mat, int *bnds)
{
    GsvSession *s = (GsvSession *)h;
    SvYeeData *d = s->mp->d;
    size_t q, n = (size_t)d->xres * d->yres * d->zres;
    for (q = 0; q < n; q++) {
        mat[q] = d->mat ? d->mat->block[q] : -1;
        bnds[q] = d->bnds->block[q];
    }
}

void
gsv_set_timing(void *h, int on)
{
    gsv_engine_set_timing(((GsvSession *)h)->mp->gpu, on);
}

void
gsv_kernel_times(void *h, double *out)
{
    gsv_engine_kernel_times(((GsvSession *)h)->mp->gpu, out);
}

int
gsv_set_launch(void *h, const int *shape)
{
    return gsv_engine_set_launch(((GsvSession *)h)->mp->gpu, shape);
}

void
gsv_get_launch(void *h, int *shape)
{
    gsv_engine_get_launch(((GsvSession *)h)->mp->gpu, shape);
}

/* Launch-shape autotuner: each candidate tile (x width, rows, z-march depth)
 * runs `steps_per_candidate` real steps of the simulation with per-launch
 * event timing; the fastest shape is kept for the H and for the E kernel
 * independently.  Returns the number of steps consumed, or -1 on error. */
int
gsv_autotune(void *h, int steps_per_candidate, int *best)
{
    static const int bxs[] = { 64, 128, 256 };
    static const int bys[] = { 1, 2, 4, 8 };
    static const int kcs[] = { 8, 16, 32, 64, 128 };
    GsvSession *s = (GsvSession *)h;
    void *eng = s->mp->gpu;
    double bh = 1e30, be = 1e30, t[4];
    int a, b, c, used = 0, shape[6], start = s->set.sc.step_act;
    gsv_engine_get_launch(eng, best);
    for (a = 0; a < 3; a++)
        for (b = 0; b < 4; b++)
            for (c = 0; c < 5; c++) {
                int nthr = bxs[a] * bys[b];
                if (nthr < 64 || nthr > 512)
                    continue;
                if (s->set.sc.step_act + steps_per_candidate > s->set.sc.nsteps)
                    goto done;
                shape[0] = shape[3] = bxs[a];
                shape[1] = shape[4] = bys[b];
                shape[2] = shape[5] = kcs[c];
                if (gsv_engine_set_launch(eng, shape))
                    continue;
                gsv_engine_set_timing(eng, 1);
                if (gsv_run(h, steps_per_candidate) < 0)
                    return -1;
                gsv_engine_kernel_times(eng, t);
                gsv_engine_set_timing(eng, 0);
                if (t[1] > 0 && t[0] / t[1] < bh) {
                    bh = t[0] / t[1];
                    best[0] = shape[0]; best[1] = shape[1]; best[2] = shape[2];
                }
                if (t[3] > 0 && t[2] / t[3] < be) {
                    be = t[2] / t[3];
                    best[3] = shape[3]; best[4] = shape[4]; best[5] = shape[5];
                }
            }
done:
    gsv_engine_set_launch(eng, best);
    used = s->set.sc.step_act - start;
    if (s->set.sc.verbose > 0)
        printf("autotune: H %dx%d z%d (%.3f ms), E %dx%d z%d (%.3f ms), %d steps\n", best[0], best[1], best[2], bh,
               best[3], best[4], best[5], be, used);
    return used;
}

/* ---- gradient extension ---- */

static int
one_step(void *h)
{
    GsvSession *s = (GsvSession *)h;
    int save = s->write_outputs, r;
    s->write_outputs = 0;
    r = gsv_run(h, 1) < 0;
    s->write_outputs = save;
    return r;
}

void
gsv_design_dims(void *h, int *dims)
{
    const SvSetAdjoint *a = &((GsvSession *)h)->set.sa;
    dims[0] = a->design ? a->design_in - a->design_i0 : 0;
    dims[1] = a->design ? a->design_jn - a->design_j0 : 0;
    dims[2] = a->design ? a->design_kn - a->design_k0 : 0;
}

int
gsv_adjoint_setup(void *h, long long history_budget_bytes, int ckpt_interval)
{
    return gsv_engine_adjoint_setup(((GsvSession *)h)->mp->gpu, history_budget_bytes, ckpt_interval);
}

int
gsv_set_design_eps(void *h, const float *eps_r)
{
    return gsv_engine_set_design_eps(((GsvSession *)h)->mp->gpu, eps_r);
}

int
gsv_get_design_eps(void *h, float *eps_r)
{
    return gsv_engine_get_design_eps(((GsvSession *)h)->mp->gpu, eps_r);
}

int
gsv_objective(void *h, double *J)
{
    return gsv_engine_adjoint_forward(((GsvSession *)h)->mp->gpu, one_step, h, J);
}

int
gsv_objective_gradient(void *h, float *grad)
{
    return gsv_engine_adjoint_backward(((GsvSession *)h)->mp->gpu, one_step, h, grad);
}

void
gsv_adjoint_info(void *h, double *out)
{
    gsv_engine_adjoint_info(((GsvSession *)h)->mp->gpu, out);
}

void
gsv_close(void *h)
{
    GsvSession *s = (GsvSession *)h;
    if (!s)
        return;
    if (s->mp) {
        gsv_engine_free(s->mp->gpu);
        s->mp->gpu = NULL;
        sv_pool_free(s->mp, &s->set);
    }
    clear_settings(&s->set, 1);
    free(s);
}

/* ---- batches of independent runs ---- */

typedef struct {
    int n;
    GsvSession **s;
    int *steps;
    void *eng;
} GsvBatchSession;

void
gsv_batch_close(void *h)
{
    GsvBatchSession *b = (GsvBatchSession *)h;
    if (!b)
        return;
    gsv_engine_batch_free(b->eng);
    free(b->s);
    free(b->steps);
    free(b);
}

void *
gsv_batch_open(void *const *sessions, int n)
{
    GsvBatchSession *b;
    void **eng;
    int r, q;
    if (n <= 0) {
        fprintf(stderr, "Error: empty batch\n");
        return NULL;
    }
    for (r = 0; r < n; r++) {
        if (!sessions[r]) {
            fprintf(stderr, "Error: batch run %d is not an open session\n", r);
            return NULL;
        }
        if (((GsvSession *)sessions[r])->write_outputs) {
            fprintf(stderr, "Error: batch run %d writes outputs; batched runs are read with gsv_get_field\n", r);
            return NULL;
        }
        for (q = 0; q < r; q++)
            if (sessions[q] == sessions[r]) {
                fprintf(stderr, "Error: session of batch run %d repeats run %d\n", r, q);
                return NULL;
            }
    }
    b = (GsvBatchSession *)calloc(1, sizeof(GsvBatchSession));
    eng = (void **)calloc((size_t)n, sizeof(void *));
    if (b) {
        b->s = (GsvSession **)calloc((size_t)n, sizeof(GsvSession *));
        b->steps = (int *)calloc((size_t)n, sizeof(int));
    }
    if (!b || !eng || !b->s || !b->steps) {
        free(eng);
        gsv_batch_close(b);
        return NULL;
    }
    b->n = n;
    for (r = 0; r < n; r++) {
        b->s[r] = (GsvSession *)sessions[r];
        eng[r] = b->s[r]->mp->gpu;
    }
    b->eng = gsv_engine_batch_new(eng, n);
    free(eng);
    if (!b->eng) {
        gsv_batch_close(b);
        return NULL;
    }
    return b;
}

int
gsv_batch_run(void *h, int nsteps)
{
    GsvBatchSession *b = (GsvBatchSession *)h;
    int r;
    for (r = 0; r < b->n; r++) {
        int left = b->s[r]->set.sc.nsteps - b->s[r]->set.sc.step_act;
        if (left < nsteps)
            nsteps = left;
    }
    if (nsteps <= 0)
        return 0;
    for (r = 0; r < b->n; r++)
        b->steps[r] = b->s[r]->set.sc.step_act;
    if (gsv_engine_batch_run(b->eng, nsteps, b->steps))
        return -1;
    for (r = 0; r < b->n; r++)
        b->s[r]->set.sc.step_act += nsteps;
    return nsteps;
}

int
gsv_batch_set_launch(void *h, const int *shape)
{
    return gsv_engine_batch_set_launch(((GsvBatchSession *)h)->eng, shape);
}

void
gsv_batch_get_launch(void *h, int *shape)
{
    gsv_engine_batch_get_launch(((GsvBatchSession *)h)->eng, shape);
}

int
gsv_batch_set_graph(void *h, int steps_per_graph)
{
    return gsv_engine_batch_set_graph(((GsvBatchSession *)h)->eng, steps_per_graph);
}

void
gsv_batch_set_timing(void *h, int on)
{
    gsv_engine_batch_set_timing(((GsvBatchSession *)h)->eng, on);
}

void
gsv_batch_info(void *h, double *out)
{
    gsv_engine_batch_info(((GsvBatchSession *)h)->eng, out);
}

/* Batch autotuner: the single-run candidate tiles timed on real batched steps
 * (H and E chosen independently), then the workgroup count of each kernel
 * (one per tile, or a resident set striding over the tile list). */
int
gsv_batch_autotune(void *h, int steps_per_candidate, int *best)
{
    static const int bxs[] = { 64, 128, 256 };
    static const int bys[] = { 1, 2, 4, 8 };
    static const int kcs[] = { 8, 16, 32, 64, 128 };
    static const int blocks[] = { 1024, 2048, 4096, 8192 };
    GsvBatchSession *b = (GsvBatchSession *)h;
    double bh = 1e30, be = 1e30, t[8];
    int a, c, d, used = 0, shape[8], left = 1 << 30;
    for (a = 0; a < b->n; a++)
        if (b->s[a]->set.sc.nsteps - b->s[a]->set.sc.step_act < left)
            left = b->s[a]->set.sc.nsteps - b->s[a]->set.sc.step_act;
    gsv_engine_batch_get_launch(b->eng, best);
    for (a = 0; a < 3; a++)
        for (c = 0; c < 4; c++)
            for (d = 0; d < 5; d++) {
                int nthr = bxs[a] * bys[c];
                if (nthr > 512)
                    continue;
                if (used + steps_per_candidate > left)
                    goto shapes_done;
                shape[0] = shape[3] = bxs[a];
                shape[1] = shape[4] = bys[c];
                shape[2] = shape[5] = kcs[d];
                shape[6] = shape[7] = 0;
                if (gsv_engine_batch_set_launch(b->eng, shape))
                    continue;
                gsv_engine_batch_set_timing(b->eng, 1);
                if (gsv_batch_run(h, steps_per_candidate) < 0)
                    return -1;
                used += steps_per_candidate;
                gsv_engine_batch_info(b->eng, t);
                gsv_engine_batch_set_timing(b->eng, 0);
                if (t[1] > 0 && t[0] < bh) {
                    bh = t[0];
                    best[0] = shape[0]; best[1] = shape[1]; best[2] = shape[2]; best[6] = 0;
                }
                if (t[3] > 0 && t[2] < be) {
                    be = t[2];
                    best[3] = shape[3]; best[4] = shape[4]; best[5] = shape[5]; best[7] = 0;
                }
            }
shapes_done:
    memcpy(shape, best, sizeof(shape));
    for (a = 0; a < 4; a++) {
        if (used + steps_per_candidate > left)
            break;
        shape[6] = shape[7] = blocks[a];
        if (gsv_engine_batch_set_launch(b->eng, shape))
            continue;
        gsv_engine_batch_set_timing(b->eng, 1);
        if (gsv_batch_run(h, steps_per_candidate) < 0)
            return -1;
        used += steps_per_candidate;
        gsv_engine_batch_info(b->eng, t);
        gsv_engine_batch_set_timing(b->eng, 0);
        if (t[1] > 0 && t[0] < bh) {
            bh = t[0];
            best[6] = blocks[a];
        }
        if (t[3] > 0 && t[2] < be) {
            be = t[2];
            best[7] = blocks[a];
        }
    }
    if (gsv_engine_batch_set_launch(b->eng, best))
        return -1;
    if (b->s[0]->set.sc.verbose > 0)
        printf("batch autotune: H %dx%d z%d wg %d (%.3f ms), E %dx%d z%d wg %d (%.3f ms), %d steps\n", best[0], best[1],
               best[2], best[6], bh, best[3], best[4], best[5], best[7], be, used);
    return used;
}
