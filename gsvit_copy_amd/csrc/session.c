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
    clear_settings(&s->set);
    if (!parse_settings(parfile, &s->set)) {
        fprintf(stderr, "Settings error!\n");
        free(s);
        return NULL;
    }
    s->set.sc.usegpu = 1;
    s->mp = init_and_make_plan(&s->set);
    if (!s->mp) {
        free_settings(&s->set);
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
    free_settings(&s->set);
    free(s);
}
