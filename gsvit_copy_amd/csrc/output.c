/*
 * output.c -- text logs of the step loop: point traces and time-domain far
 * field, in the reference formats (src3d/output.c:51-99 is_output,
 * :645-681 point logs "%d %g" / "%d %g %g %g %g %g %g", :690-709 sum logs, :1266-1271 far-field
 * log "%d %g %g %g" of the negated bins).  The running-DFT face data, which
 * has no reference counterpart, is written as one line per face cell.
 */
#include <stdio.h>
#include <stdlib.h>
#include "gsvit_internal.h"

int
is_output(SvPool *mp, SvSet *set)
{
    int i;
    (void)mp;
    for (i = 0; i < set->so.npnts; i++) {
        if (set->so.pnts[i].step == 0)
            return 0;
        if (set->sc.step_act % set->so.pnts[i].step == 0)
            return 1;
    }
    for (i = 0; i < set->so.nsums; i++) {
        if (set->so.sums[i].step == 0)
            return 0;
        if (set->sc.step_act % set->so.sums[i].step == 0)
            return 1;
    }
    return 0;
}

/* FLUX_DFT spectra from the accumulators: out[p * nw + w] = 1/2 Re sum over the
 * plane's cells of (Ex Hy* - Ey Hx*) dx dy, the DFTs scaled by dt */
void
sv_flux_spectra(const SvPool *mp, const SvSet *set, double *out)
{
    const SvFarfield *ff = mp->farfield;
    const SvSetFlux *x = &set->sx;
    const long per = (long)(x->i1 - x->i0) * (x->j1 - x->j0);
    const double scale = 0.5 * set->sp.dx * set->sp.dy * set->plan.dt * set->plan.dt;
    int p, w;
    for (p = 0; p < x->nplanes; p++)
        for (w = 0; w < ff->nflux; w++) {
            double s = 0;
            long c;
            for (c = 0; c < per; c++) {
                const double *a = ff->flux_data + ((size_t)(p * per + c) * ff->nflux + w) * 8;
                s += a[0] * a[6] + a[1] * a[7] - a[2] * a[4] - a[3] * a[5];
            }
            out[p * ff->nflux + w] = scale * s;
        }
}

void
sv_pool_output(SvPool *mp, SvSet *set)
{
    int i, j, n;
    FILE *fh;

    for (i = 0; i < set->so.npnts; i++) {
        SvOutputPar *o = set->so.pnts + i;
        if (o->step == 0)
            continue;
        if (set->sc.step_act % o->step != 0 && set->sc.step_act < (set->sc.nsteps - 1))
            continue;
        fh = fopen(o->filebase, "w");
        if (!fh) {
            fprintf(stderr, "Error: point output: cannot open file %s for writing\n", o->filebase);
            continue;
        }
        for (j = 0; j < set->sc.step_act; j++) {
            const double *v = mp->out->outpointdata[i] + (size_t)j * 6;
            if (o->component == SV_COMP_ALL)
                fprintf(fh, "%d %g %g %g %g %g %g\n", j, v[0], v[1], v[2], v[3], v[4], v[5]);
            else
                fprintf(fh, "%d %g\n", j, v[o->component]);
        }
        fclose(fh);
    }

    /* sum logs "%d %g", every step of the run, scaled by the cell volume
     * except for the maximum (src3d/output.c:690-709; the reference tests
     * the `step` of a stale index there, the log's own `step` is used here) */
    for (n = 0; n < set->so.nsums; n++) {
        SvOutputSum *o = set->so.sums + n;
        double vol = set->sp.dx * set->sp.dy * set->sp.dz;
        if (o->step == 0)
            continue;
        if (set->sc.step_act % o->step != 0 && set->sc.step_act < (set->sc.nsteps - 1))
            continue;
        fh = fopen(o->filename, "w");
        if (!fh) {
            fprintf(stderr, "Error: sum: cannot open file %s for writing\n", o->filename);
            continue;
        }
        for (j = 0; j < set->sc.nsteps; j++)
            fprintf(fh, "%d %g\n", j, mp->out->outsumdata[n][j] * (o->component == SV_SUM_MAX ? 1.0 : vol));
        fclose(fh);
    }

    /* far-field logs are written at the last step only */
    if (set->sc.step_act != set->sc.nsteps - 1)
        return;
    for (n = 0; n < set->sf.nrs; n++) {
        SvRPoint *r = mp->farfield->rpoints + n;
        fh = fopen(set->sf.source_filename[n], "w");
        if (!fh) {
            fprintf(stderr, "Error: farfield: cannot open file %s for writing\n", set->sf.source_filename[n]);
            continue;
        }
        for (i = 0; i < mp->farfield->ndata; i++)
            fprintf(fh, "%d %g %g %g\n", i, -r->ex[i], -r->ey[i], -r->ez[i]);
        fclose(fh);
    }
    if (set->sf.ndft && set->sf.dft_filename[0]) {
        SvFarfield *ff = mp->farfield;
        long c;
        int q;
        fh = fopen(set->sf.dft_filename, "w");
        if (!fh) {
            fprintf(stderr, "Error: farfield DFT: cannot open file %s for writing\n", set->sf.dft_filename);
            return;
        }
        fprintf(fh, "# running DFT of tangential fields on the NFFF box: cell, then per wavelength "
                    "(re, im) of the 4 tangential components\n");
        for (c = 0; c < ff->dft_ncells; c++) {
            fprintf(fh, "%ld", c);
            for (q = 0; q < ff->ndft * 8; q++)
                fprintf(fh, " %g", ff->dft_data[(size_t)c * ff->ndft * 8 + q]);
            fprintf(fh, "\n");
        }
        fclose(fh);
    }
    if (set->sx.nplanes && set->sx.filename[0]) {
        SvFarfield *ff = mp->farfield;
        double *fl = (double *)malloc((size_t)set->sx.nplanes * ff->nflux * sizeof(double));
        int w, p;
        fh = fopen(set->sx.filename, "w");
        if (!fh || !fl) {
            fprintf(stderr, "Error: flux DFT: cannot open file %s for writing\n", set->sx.filename);
            free(fl);
            if (fh)
                fclose(fh);
            return;
        }
        sv_flux_spectra(mp, set, fl);
        fprintf(fh, "# net z flux spectrum through each plane: lambda, then one column per plane (k =");
        for (p = 0; p < set->sx.nplanes; p++)
            fprintf(fh, " %d", set->sx.k[p]);
        fprintf(fh, ")\n");
        for (w = 0; w < ff->nflux; w++) {
            fprintf(fh, "%g", set->sx.lambda[w]);
            for (p = 0; p < set->sx.nplanes; p++)
                fprintf(fh, " %g", fl[p * ff->nflux + w]);
            fprintf(fh, "\n");
        }
        fclose(fh);
        free(fl);
    }
}
