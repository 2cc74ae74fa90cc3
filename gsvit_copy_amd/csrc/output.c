/*
 * output.c -- text logs of the step loop: point traces and time-domain far
 * field, in the reference formats (src3d/output.c:51-99 is_output,
 * :645-681 point logs "%d %g" / "%d %g %g %g %g %g %g", :690-709 sum logs, :1266-1271 far-field
 * log "%d %g %g %g" of the negated bins), plane and volume outputs
 * (:712-905 volumes: ASCII "%g", VTK STRUCTURED_POINTS, raw doubles;
 * :1100-1263 planes: ASCII "%10.10g", raw doubles).  The running-DFT face data,
 * which has no reference counterpart, is written as one line per face cell,
 * and its far-field pattern as one line per wavelength and direction.
 */
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
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
    for (i = 0; i < set->so.nforces; i++) {
        if (set->so.forces[i].step == 0)
            return 0;
        if (set->sc.step_act % set->so.forces[i].step == 0)
            return 1;
    }
    for (i = 0; i < set->so.ncubs; i++) {
        const SvOutputRegion *o = set->so.cubs + i;
        if (o->step == 0)
            return 0;
        if (set->sc.step_act % o->step == 0 && !(set->sc.step_act < o->start || set->sc.step_act > o->stop))
            return 1;
    }
    for (i = 0; i < set->so.nplns; i++) {
        const SvOutputRegion *o = set->so.plns + i;
        if (o->step == 0)
            return 0;
        if (set->sc.step_act % o->step == 0 && !(set->sc.step_act < o->start || set->sc.step_act > o->stop))
            return 1;
    }
    return 0;
}

static void
vtk_header(FILE *fw, int xres, int yres, int zres, const char *description)
{
    /* src3d/output.c:139-151 */
    fprintf(fw, "# vtk DataFile Version 2.0\n%s\nASCII\nDATASET STRUCTURED_POINTS\n", description);
    fprintf(fw, "DIMENSIONS %d %d %d\nASPECT_RATIO 1 1 1\nORIGIN 0 0 0\nPOINT_DATA %d\n", xres, yres, zres,
            xres * yres * zres);
    fprintf(fw, "SCALARS volume_scalars double 1\nLOOKUP_TABLE default\n");
}

/* value of volume component c at (i, j, k); *is_int for integer outputs */
static double
volume_value(const SvPool *mp, const SvOutputRegion *o, int i, int j, int k, int *is_int)
{
    const SvYeeData *d = mp->d;
    double eps, sig, mu, sigast;
    int mi = d->mat ? d->mat->data[i][j][k] : 0;
    *is_int = 0;
    switch (o->component) {
    case SV_OVOLUME_EX: return d->ex->data[i][j][k];
    case SV_OVOLUME_EY: return d->ey->data[i][j][k];
    case SV_OVOLUME_EZ: return d->ez->data[i][j][k];
    case SV_OVOLUME_HX: return d->hx->data[i][j][k];
    case SV_OVOLUME_HY: return d->hy->data[i][j][k];
    case SV_OVOLUME_HZ: return d->hz->data[i][j][k];
    case SV_OVOLUME_ALL:
        return sqrt(d->ex->data[i][j][k] * d->ex->data[i][j][k] + d->ey->data[i][j][k] * d->ey->data[i][j][k]
                    + d->ez->data[i][j][k] * d->ez->data[i][j][k]);
    case SV_OVOLUME_EPSILON:
        sv_pool_epsilon_sigma(mp, i, j, k, &eps, &sig);
        return eps / SV_EPSILON_0;
    case SV_OVOLUME_SIGMA:
        sv_pool_epsilon_sigma(mp, i, j, k, &eps, &sig);
        return sig;
    case SV_OVOLUME_MU:
        sv_pool_mu_sigast(mp, i, j, k, &mu, &sigast);
        return mu / SV_MU_0;
    case SV_OVOLUME_SIGAST:
        sv_pool_mu_sigast(mp, i, j, k, &mu, &sigast);
        return sigast;
    case SV_OVOLUME_MAT:
        *is_int = 1;
        return mi;
    case SV_OVOLUME_MATTYPE:
        *is_int = 1;
        return mp->mats[mi].type;
    default:
        return mp->out->outabsdata[o->k]->data[i][j][k];
    }
}

/* OUT_VOLUME at this step, src3d/output.c:712-905 */
static void
write_volumes(SvPool *mp, SvSet *set)
{
    const SvYeeData *d = mp->d;
    const int X = set->sp.xres, Y = set->sp.yres, Z = set->sp.zres;
    static const char *vtk_names[] = { "Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "All", "Epsilon", "Sigma", "Mu",
                                       "Sigast", "Material", "Mattype", "Abs", "Sumall", "Maxall" };
    char filename[SV_NAME_LEN + 32];
    int n, i, j, k, is_int;
    for (n = 0; n < set->so.ncubs; n++) {
        const SvOutputRegion *o = set->so.cubs + n;
        const int acc = o->component >= SV_OVOLUME_ABS;
        const int table = o->component == SV_OVOLUME_MAT || o->component == SV_OVOLUME_MATTYPE;
        FILE *fh;
        if (o->step == 0)
            continue;
        if (set->sc.step_act % o->step != 0 && set->sc.step_act < (set->sc.nsteps - 1))
            continue;
        if (set->sc.step_act < o->start || set->sc.step_act > o->stop)
            continue;
        if (acc) {
            /* accumulated over the output steps: sigma |E|^2, |E|^2 or max |E|^2 */
            SvDCube *a = mp->out->outabsdata[o->k];
            for (i = 0; i < X; i++)
                for (j = 0; j < Y; j++)
                    for (k = 0; k < Z; k++) {
                        const double e2 = d->ex->data[i][j][k] * d->ex->data[i][j][k]
                                          + d->ey->data[i][j][k] * d->ey->data[i][j][k]
                                          + d->ez->data[i][j][k] * d->ez->data[i][j][k];
                        double w = 1, eps;
                        if (o->component == SV_OVOLUME_ABS)
                            sv_pool_epsilon_sigma(mp, i, j, k, &eps, &w);
                        if (o->component == SV_OVOLUME_MAXALL)
                            a->data[i][j][k] = a->data[i][j][k] > e2 ? a->data[i][j][k] : e2;
                        else
                            a->data[i][j][k] += w * e2;
                    }
        }
        if (o->format == 1 || o->format == 2) {
            if (acc)
                snprintf(filename, sizeof(filename), o->format == 2 ? "%s.vtk" : "%s", o->filebase);
            else
                snprintf(filename, sizeof(filename), o->format == 2 ? "%s_%.4d.vtk" : "%s_%.4d", o->filebase,
                         set->sc.step_act);
            fh = fopen(filename, "w");
            if (!fh) {
                fprintf(stderr, "Error: volume: cannot open file %s for writing\n", filename);
                continue;
            }
            if (o->format == 1) {
                /* ASCII: the reference prints the E components for Hx, Hy, Hz here */
                SvOutputRegion oe = *o;
                if (oe.component >= SV_OVOLUME_HX && oe.component <= SV_OVOLUME_HZ)
                    oe.component -= 3;
                fprintf(fh, "%d %d %d %d\n", X, Y, Z, o->component);
                if (table && mp->nmat == 0)
                    fprintf(fh, "No data available\n");
                else {
                    for (k = 0; k < Z; k++) {
                        for (j = 0; j < Y; j++) {
                            for (i = 0; i < X; i++) {
                                const double v = volume_value(mp, &oe, i, j, k, &is_int);
                                if (is_int)
                                    fprintf(fh, "%d ", (int)v);
                                else
                                    fprintf(fh, "%g ", v);
                            }
                            fprintf(fh, "\n");
                        }
                        fprintf(fh, "\n");
                    }
                }
            } else {
                vtk_header(fh, X, Y, Z, vtk_names[o->component]);
                for (k = 0; k < Z; k++)
                    for (j = 0; j < Y; j++)
                        for (i = 0; i < X; i++) {
                            double v;
                            if (table && !d->mat) {
                                fprintf(fh, "0\n");
                                continue;
                            }
                            v = volume_value(mp, o, i, j, k, &is_int);
                            if (is_int)
                                fprintf(fh, "%d\n", (int)v);
                            else
                                fprintf(fh, "%g\n", v);
                        }
            }
            fclose(fh);
        } else {
            /* raw doubles, x fastest; the accumulated outputs only once their window closes */
            if (acc && !(set->sc.step_act > (o->stop - o->step)))
                continue;
            if (acc)
                snprintf(filename, sizeof(filename), "%s.raw", o->filebase);
            else
                snprintf(filename, sizeof(filename), "%s_%.4d.raw", o->filebase, set->sc.step_act);
            fh = fopen(filename, "wb");
            if (!fh) {
                fprintf(stderr, "Error: volume: cannot open file %s for writing\n", filename);
                continue;
            }
            if (table && mp->nmat == 0)
                fprintf(stderr, "No data available for binary output.\n");
            else
                for (k = 0; k < Z; k++)
                    for (j = 0; j < Y; j++)
                        for (i = 0; i < X; i++) {
                            const double v = volume_value(mp, o, i, j, k, &is_int);
                            fwrite(&v, sizeof(double), 1, fh);
                        }
            fclose(fh);
        }
    }
}

/* OUT_PLANE at this step, src3d/output.c:1100-1263: a plane of one field
 * component, the plane's fixed index the one that is not -1 */
static void
write_planes(SvPool *mp, SvSet *set)
{
    const SvYeeData *d = mp->d;
    char filename[SV_NAME_LEN + 32];
    int n;
    for (n = 0; n < set->so.nplns; n++) {
        const SvOutputRegion *o = set->so.plns + n;
        const SvDCube *c = NULL;
        int a, b, na, nb, axis;
        FILE *fh;
        if (o->step == 0)
            continue;
        if (set->sc.step_act % o->step != 0 && set->sc.step_act < (set->sc.nsteps - 1))
            continue;
        if (set->sc.step_act <= o->start || set->sc.step_act > o->stop)
            continue;
        if (o->format)
            snprintf(filename, sizeof(filename), "%s_%.4d", o->filebase, set->sc.step_act);
        else
            snprintf(filename, sizeof(filename), "%s_%.4d.raw", o->filebase, set->sc.step_act);
        fh = fopen(filename, o->format ? "w" : "wb");
        if (!fh) {
            fprintf(stderr, "Error: plane: cannot open file %s for writing\n", filename);
            continue;
        }
        switch (o->component) {
        case SV_COMP_EX: c = d->ex; break;
        case SV_COMP_EY: c = d->ey; break;
        case SV_COMP_EZ: c = d->ez; break;
        case SV_COMP_HX: c = d->hx; break;
        case SV_COMP_HY: c = d->hy; break;
        case SV_COMP_HZ: c = d->hz; break;
        default:
            fprintf(stderr, "Error: Unsupported output type for plane output\n");
        }
        for (axis = 0; axis < 3 && c; axis++) {
            /* i fixed: rows k, columns j; j fixed: rows k, columns i; k fixed: rows j, columns i */
            if ((axis == 0 && !(o->j == -1 && o->k == -1)) || (axis == 1 && !(o->i == -1 && o->k == -1))
                || (axis == 2 && !(o->i == -1 && o->j == -1)))
                continue;
            nb = axis == 2 ? set->sp.yres : set->sp.zres;
            na = axis == 0 ? set->sp.yres : set->sp.xres;
            for (b = 0; b < nb; b++) {
                for (a = 0; a < na; a++) {
                    const double v = axis == 0 ? c->data[o->i][a][b] : axis == 1 ? c->data[a][o->j][b] : c->data[a][b][o->k];
                    if (o->format)
                        fprintf(fh, "%10.10g ", v);
                    else
                        fwrite(&v, sizeof(double), 1, fh);
                }
                if (o->format)
                    fprintf(fh, "\n");
            }
        }
        fclose(fh);
    }
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

/* Power spectrum of a point trace for the AllFFT log (src3d/output.c:517-553,
 * 607-640): the trace of n samples less its least-squares straight line,
 * its discrete Fourier transform X_k, and out[k] = |X_k|^2 / n for
 * k = 0 .. n/2 - 1.  The reference takes the transform from Gwyddion (which
 * may resample the trace to a size its FFT prefers); this is a direct
 * transform of the n samples, summed in double with the phases rotated by
 * exact angle steps. */
void
sv_point_spectrum(const double *x, int n, double *out)
{
    double sx = 0, sxx = 0, sy = 0, sxy = 0, a = 0, b = 0;
    double *y;
    int i, k;
    if (n < 2)
        return;
    y = (double *)malloc((size_t)n * sizeof(double));
    if (!y)
        return;
    for (i = 0; i < n; i++) {
        sx += i;
        sxx += (double)i * i;
        sy += x[i];
        sxy += i * x[i];
    }
    b = (n * sxy - sx * sy) / (n * sxx - sx * sx);
    a = (sy - b * sx) / n;
    for (i = 0; i < n; i++)
        y[i] = x[i] - a - b * i;
    #pragma omp parallel for private(i) schedule(static)
    for (k = 0; k < n / 2; k++) {
        double re = 0, im = 0;
        for (i = 0; i < n; i++) {
            /* the angle 2 pi k i / n reduced exactly in integers */
            const double ph = 2 * M_PI * (double)(((long long)k * i) % n) / n;
            re += y[i] * cos(ph);
            im -= y[i] * sin(ph);
        }
        out[k] = (re * re + im * im) / n;
    }
    free(y);
}

/* NFFF_DFT_PATTERN: the far-field pattern F (E = F exp(-i k r) / r) of the
 * running DFT on the requested angle grid, through the transform attached to
 * the far-field data by the engine that holds the accumulators.  One line per
 * wavelength and direction: lambda, theta, phi (degrees), re / im F_theta,
 * re / im F_phi. */
static void
write_dft_pattern(SvPool *mp, SvSet *set)
{
    SvFarfield *ff = mp->farfield;
    const int nt = set->sf.dft_pattern_ntheta, np = set->sf.dft_pattern_nphi, nd = nt * np;
    double *dirs, *out;
    int a, b, w;
    FILE *fh;
    if (!ff->dft_transform) {
        fprintf(stderr, "Error: farfield DFT pattern: no transform available\n");
        return;
    }
    dirs = (double *)malloc((size_t)nd * 2 * sizeof(double));
    out = (double *)malloc((size_t)nd * ff->ndft * 4 * sizeof(double));
    if (!dirs || !out) {
        free(dirs);
        free(out);
        fprintf(stderr, "Error: farfield DFT pattern: out of memory\n");
        return;
    }
    for (a = 0; a < nt; a++)
        for (b = 0; b < np; b++) {
            dirs[2 * (a * np + b)] = nt > 1 ? M_PI * a / (nt - 1) : 0.5 * M_PI;
            dirs[2 * (a * np + b) + 1] = 2 * M_PI * b / np;
        }
    if (ff->dft_transform(ff->dft_ctx, nd, dirs, out)) {
        fprintf(stderr, "Error: farfield DFT pattern: transform failed\n");
    } else if (!(fh = fopen(set->sf.dft_pattern_filename, "w"))) {
        fprintf(stderr, "Error: farfield DFT pattern: cannot open file %s for writing\n", set->sf.dft_pattern_filename);
    } else {
        fprintf(fh, "# far-field pattern of the running DFT, E = F exp(-i k r) / r: lambda theta phi (deg) "
                    "re(F_theta) im(F_theta) re(F_phi) im(F_phi)\n");
        for (w = 0; w < ff->ndft; w++)
            for (a = 0; a < nd; a++) {
                const double *o = out + ((size_t)a * ff->ndft + w) * 4;
                fprintf(fh, "%g %g %g %g %g %g %g\n", set->sf.dft_lambda[w], dirs[2 * a] * 180 / M_PI,
                        dirs[2 * a + 1] * 180 / M_PI, o[0], o[1], o[2], o[3]);
            }
        fclose(fh);
    }
    free(dirs);
    free(out);
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
        if (o->component == SV_COMP_ALLFFT) {
            /* src3d/output.c:607-640: written at the last step only, one line per
             * frequency bin k = 1 .. n/2 - 1: wavelength c n dt / k, the summed
             * power of Ex, Ey, Ez, then each component's amplitude */
            const int n = o->stop - o->start;
            if (set->sc.step_act == set->sc.nsteps - 1 && n >= 4) {
                double *tr = (double *)malloc((size_t)n * sizeof(double));
                double *sp = (double *)calloc((size_t)3 * (n / 2), sizeof(double));
                int c;
                if (tr && sp) {
                    for (c = 0; c < 3; c++) {
                        for (j = 0; j < n; j++)
                            tr[j] = mp->out->outpointdata[i][(size_t)(j + o->start) * 6 + c];
                        sv_point_spectrum(tr, n, sp + (size_t)c * (n / 2));
                    }
                    for (j = 1; j < n / 2; j++)
                        fprintf(fh, "%g %g %g %g %g\n", SV_LIGHT_SPEED / ((double)j / ((double)n / 2.0 * set->plan.dt * 2.0)),
                                sp[j] + sp[n / 2 + j] + sp[2 * (n / 2) + j], sqrt(sp[j]), sqrt(sp[n / 2 + j]),
                                sqrt(sp[2 * (n / 2) + j]));
                }
                free(tr);
                free(sp);
            }
            fclose(fh);
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

    /* force logs, src3d/output.c:908-938: per step the force times dx dy and
     * its mean over the preceding 5 lambda_center / (c dt) steps (the
     * reference tests forces[i].step with a stale index there; the force's
     * own step is used here) */
    for (n = 0; n < set->so.nforces; n++) {
        const SvOutputForce *o = set->so.forces + n;
        const double *fd = mp->out->outforcedata[n], area = set->sp.dx * set->sp.dy;
        if (o->step == 0)
            continue;
        if (set->sc.step_act % o->step != 0 && set->sc.step_act < set->sc.nsteps - 1)
            continue;
        fh = fopen(o->filename, "w");
        if (!fh) {
            fprintf(stderr, "Error: force: cannot open file %s for writing\n", o->filename);
            continue;
        }
        for (i = 1; i < set->sc.nsteps; i++) {
            double avx = 0, avy = 0, avz = 0, w = i - (5.0 * set->ss.lambda_center / set->plan.dt / SV_LIGHT_SPEED);
            int j, nav = 0;
            for (j = (int)(w > 0 ? w : 0); j < i; j++) {
                avx += fd[3 * j];
                avy += fd[3 * j + 1];
                avz += fd[3 * j + 2];
                nav++;
            }
            fprintf(fh, "%d %g %g %g %g %g %g\n", i, fd[3 * i] * area, fd[3 * i + 1] * area, fd[3 * i + 2] * area,
                    avx / nav * area, avy / nav * area, avz / nav * area);
        }
        fclose(fh);
    }

    write_volumes(mp, set);
    write_planes(mp, set);

    /* far-field logs are written at the last step only */
    if (set->sc.step_act != set->sc.nsteps - 1)
        return;
    for (n = 0; n < set->sf.nrs; n++) {
        SvRPoint *r = mp->farfield->rpoints + n;
        if (!set->sf.source_filename[n][0])
            continue;
        fh = fopen(set->sf.source_filename[n], "w");
        if (!fh) {
            fprintf(stderr, "Error: farfield: cannot open file %s for writing\n", set->sf.source_filename[n]);
            continue;
        }
        for (i = 0; i < mp->farfield->ndata; i++)
            fprintf(fh, "%d %g %g %g\n", i, -r->ex[i], -r->ey[i], -r->ez[i]);
        fclose(fh);
    }
    /* periodic far field, src3d/output.c:1335-1354 (the reference rewrites these
     * files on every output step; the content of the last one is the result) */
    for (n = 0; n < set->spf.nrs; n++) {
        SvRPoint *r = mp->pfarfield->prpoints + n;
        if (!set->spf.source_filename[n][0])
            continue;
        fh = fopen(set->spf.source_filename[n], "w");
        if (!fh) {
            fprintf(stderr, "Error: periodic farfield: cannot open file %s for writing\n", set->spf.source_filename[n]);
            continue;
        }
        for (i = 0; i < mp->pfarfield->ndata; i++)
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
    if (set->sf.ndft && set->sf.dft_pattern_filename[0])
        write_dft_pattern(mp, set);
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
