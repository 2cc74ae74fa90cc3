/*
 * settings.c -- GSvit parameter-file reader for the 3-D time-stepping path.
 *
 * Accepts the keyword format of src3d/settings.c:3822-5840 (whitespace separated
 * keywords, '#' comments, values read with the same integer/double rules as
 * get_int/get_double, src3d/settings.c:103-135).  Keywords outside the hot path
 * (sub-gridding, local and focused layered sources, geometry modifiers) are
 * rejected with an error instead of being silently ignored, so a parameter
 * file never runs with part of its physics dropped.  Keywords whose only
 * product is a Gwyddion container (OUT_IMAGE, OUT_FILE) are consumed and
 * ignored with a warning; the far-field areas (NFFF_SPHERICAL_AREA,
 * NFFF_PLANAR_AREA, PERIODIC_NFFF_SPHERICAL_AREA) add their points and
 * per-point logs, only the Gwyddion image of each set is not written.
 *
 * Generated sources (types 1 and 2, src3d/settings.c:444-548) are produced here
 * with the reference's waveform formulas and then rounded to 6 significant
 * digits, because the reference writes them with "%g" and reads them back
 * (src3d/settings.c:515-538, src3d/plan.c:1449-1512).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "gsvit_settings.h"
#include "gsvit_internal.h"

/* remove != 0 frees what an earlier parse allocated, as the reference's
 * clear_settings(set, TRUE) does; remove == 0 expects an uninitialised SvSet */
void
clear_settings(SvSet *set, int remove)
{
    if (remove) {
        free(set->ss.pnts);
        free(set->so.pnts);
        free(set->so.sums);
        free(set->so.forces);
        free(set->so.plns);
        free(set->so.cubs);
        free(set->sf.ri);
        free(set->sf.rj);
        free(set->sf.rk);
        free(set->sf.individual);
        free(set->sf.source_filename);
        free(set->sf.dft_lambda);
        free(set->spf.ri);
        free(set->spf.rj);
        free(set->spf.rk);
        free(set->spf.source_filename);
        free(set->sx.k);
        free(set->sx.lambda);
    }
    memset(set, 0, sizeof(SvSet));
    /* defaults from src3d/settings.c:570-990 and src3d/defaults.h */
    set->sp.xres = set->sp.yres = set->sp.zres = 100;
    set->sp.dx = set->sp.dy = set->sp.dz = 1e-8;
    set->sc.nsteps = 100;
    set->sc.dtmult = 1.0;
    set->sc.nthreads = -1;
    set->sc.verbose = 0;
    set->sm.matmode_check = 0;
    set->ss.tsf.box_i0 = set->ss.tsf.box_j0 = set->ss.tsf.box_k0 = 10;
    set->ss.tsf.box_in = set->ss.tsf.box_jn = set->ss.tsf.box_kn = 90;
    set->ss.tsf.layered_epsilon = 1;
    set->ss.tsf.layered_mu = 1;
    set->ss.tsf.box_boundary_skipdepth = -1;
    set->ss.tsf.source_mode = 1;
    set->ss.tsf.source_wl = 0.633e-6;
    set->ss.tsf.source_amplitude = 1.0;
    set->ss.tsf.source_pulsewidth = 100;
    set->ss.ltsf.box_i0 = set->ss.ltsf.box_j0 = set->ss.ltsf.box_k0 = 10;
    set->ss.ltsf.box_in = set->ss.ltsf.box_jn = set->ss.ltsf.box_kn = 90;
    set->ss.ltsf.box_boundary_skipdepth = -1;
    set->ss.sf.layered_epsilon = 1;
    set->ss.sf.layered_mu = 1;
    set->sb.depth_bx0 = set->sb.depth_bxn = set->sb.depth_by0 = 10;
    set->sb.depth_byn = set->sb.depth_bz0 = set->sb.depth_bzn = 10;
    set->smb.bxnpos = set->smb.bynpos = set->smb.bznpos = -1;
    set->sf.sumfrom = 0;
    set->sf.sumto = 0;
    /* src3d/settings.c:945-957, src3d/defaults.h:167-179 */
    set->spf.box_i0 = set->spf.box_j0 = set->spf.box_k0 = 10;
    set->spf.box_in = set->spf.box_jn = set->spf.box_kn = 90;
    set->spf.pimin = set->spf.pjmin = 0;
    set->spf.pimax = set->spf.pjmax = 1;
}

static int
get_token(FILE *fr, char *buf, size_t n)
{
    char fmt[16];
    snprintf(fmt, sizeof(fmt), "%%%ds", (int)n - 1);
    return fscanf(fr, fmt, buf) == 1 ? 0 : 1;
}

/* src3d/settings.c:103-119: integers may be written in exponential form */
static int
get_int(FILE *fr, int *val, const char *key)
{
    char value[100], *err;
    if (get_token(fr, value, sizeof(value))) {
        fprintf(stderr, "Error: missing integer value (key %s)\n", key);
        return 1;
    }
    *val = (int)strtol(value, &err, 10);
    if (*err != '\0') {
        *val = (int)strtod(value, &err);
        if (*err != '\0') {
            fprintf(stderr, "Error parsing integer value (key %s)\n", key);
            return 1;
        }
    }
    return 0;
}

/* src3d/settings.c:121-135 */
static int
get_double(FILE *fr, double *val, const char *key)
{
    char value[100], *err;
    if (get_token(fr, value, sizeof(value))) {
        fprintf(stderr, "Error: missing double value (key %s)\n", key);
        return 1;
    }
    *val = strtod(value, &err);
    if (*err != '\0') {
        fprintf(stderr, "Error parsing double value (key %s)\n", key);
        return 1;
    }
    return 0;
}

/* rest of the line, whitespace stripped (the reference's "%100[^\n\r]" + g_strstrip) */
static int
get_rest_of_line(FILE *fr, char *out, size_t n)
{
    char buf[SV_NAME_LEN];
    char *s, *e;
    int c;

    /* skip blanks on this line, then continue onto the next non-empty line,
     * as fscanf("%[^\n]") would after the previous token */
    do {
        c = fgetc(fr);
    } while (c == ' ' || c == '\t' || c == '\n' || c == '\r');
    if (c == EOF) {
        out[0] = 0;
        return 1;
    }
    ungetc(c, fr);
    if (!fgets(buf, sizeof(buf), fr)) {
        out[0] = 0;
        return 1;
    }
    s = buf;
    while (*s == ' ' || *s == '\t')
        s++;
    e = s + strlen(s);
    while (e > s && (e[-1] == '\n' || e[-1] == '\r' || e[-1] == ' ' || e[-1] == '\t'))
        *--e = 0;
    snprintf(out, n, "%s", s);
    return out[0] == 0;
}

/* src3d/settings.c:137-170 */
static int
get_component(FILE *fr, int *val, const char *key)
{
    char c[32];
    if (get_token(fr, c, sizeof(c)))
        return 1;
    if (!strcmp(c, "Ex") || !strcmp(c, "ex")) *val = SV_COMP_EX;
    else if (!strcmp(c, "Ey") || !strcmp(c, "ey")) *val = SV_COMP_EY;
    else if (!strcmp(c, "Ez") || !strcmp(c, "ez")) *val = SV_COMP_EZ;
    else if (!strcmp(c, "Hx") || !strcmp(c, "hx")) *val = SV_COMP_HX;
    else if (!strcmp(c, "Hy") || !strcmp(c, "hy")) *val = SV_COMP_HY;
    else if (!strcmp(c, "Hz") || !strcmp(c, "hz")) *val = SV_COMP_HZ;
    else if (!strcmp(c, "All") || !strcmp(c, "all")) *val = SV_COMP_ALL;
    else if (!strcmp(c, "AllFFT") || !strcmp(c, "allfft")) *val = SV_COMP_ALLFFT;
    else {
        fprintf(stderr, "Error: field component %s is not supported here (key %s)\n", c, key);
        return 1;
    }
    return 0;
}

/* OUT_VOLUME component names, src3d/settings.c:180-221 */
static int
get_volume_component(FILE *fr, int *val, const char *key)
{
    static const char *names[] = { "Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "All", "Epsilon", "Sigma", "Mu", "Sigast",
                                   "Material", "Mattype", "Abs", "Sumall", "Maxall" };
    char c[32];
    int q;
    if (get_token(fr, c, sizeof(c)))
        return 1;
    for (q = 0; q < (int)(sizeof(names) / sizeof(names[0])); q++) {
        char lower[32];
        int n;
        for (n = 0; names[q][n]; n++)
            lower[n] = (char)(names[q][n] | (n == 0 ? 0x20 : 0));
        lower[n] = 0;
        if (!strcmp(c, names[q]) || !strcmp(c, lower)) {
            *val = q;
            return 0;
        }
    }
    fprintf(stderr, "Error: volume component %s is not supported (key %s)\n", c, key);
    return 1;
}

static double
round_g6(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%g", v);
    return strtod(b, NULL);
}

/* Waveform of a generated source, src3d/settings.c:444-512 (SRC_SINE = 1,
 * SRC_PULSE = 2); values rounded as the reference's "%g" round trip. */
static int
generate_waveform(int type, int nsteps, double wavelength, int pulse_width, double dt, double amplitude, int dampval,
                  double *vals)
{
    int i;
    double wavefreq = SV_LIGHT_SPEED * 2 * M_PI / wavelength;
    double dpulse_width = pulse_width * dt;

    if (type != 1 && type != 2) {
        fprintf(stderr, "Error: generated source type %d is not supported (use 1 = sine, 2 = pulse, or a file)\n", type);
        return 1;
    }
    for (i = 0; i < nsteps; i++) {
        double v = amplitude * sin(wavefreq * dt * (double)i);
        if (i < dampval)
            v *= (1.0 / (double)dampval) * ((double)i);
        if (type == 2) {
            double damp = ((double)i * dt - 4.0 * dpulse_width) / sqrt(2.0) / dpulse_width;
            v *= exp(-damp * damp);
        }
        vals[i] = v;
    }
    return 0;
}

int
sv_generate_waveform(int type, int nsteps, double wavelength, int pulse_width, double dt,
                     double amplitude, double *vals)
{
    return generate_waveform(type, nsteps, wavelength, pulse_width, dt, amplitude, 100, vals);
}

/* the layered-TSF form (src3d/settings.c:460-465): 3 nsteps samples at dt / 3,
 * pulse width and ramp in substeps (tripled) */
int
sv_generate_waveform_ltsf(int type, int nsteps, double wavelength, int pulse_width, double dt,
                          double amplitude, double *vals)
{
    return generate_waveform(type, 3 * nsteps, wavelength, 3 * pulse_width, dt / 3, amplitude, 300, vals);
}

static double
plan_dt(const SvSet *set)
{
    /* src3d/plan.c:5452 */
    return 1.0 / SV_LIGHT_SPEED / sqrt(1.0 / set->sp.dx / set->sp.dx + 1.0 / set->sp.dy / set->sp.dy
                                       + 1.0 / set->sp.dz / set->sp.dz) * set->sc.dtmult;
}

/* one face's boundary keyword; `spacing` is the cell size the default CPML
 * sigma uses: dx for BOUNDARY_ALL and the x faces, dy for the y faces, dz for
 * the z faces (src3d/settings.c:5151, 5186, 5216, 5246, 5277, 5307, 5337) */
static int
parse_cpml_face(FILE *fr, const char *key, double spacing, int *type, int *depth, int *m,
                double *sigma, double *a, double *kappa)
{
    char value[64];
    if (get_token(fr, value, sizeof(value)))
        return 1;
    if (!strcmp(value, "none"))
        *type = SV_BOUNDARY_NONE;
    else if (!strcmp(value, "pec"))
        *type = SV_BOUNDARY_PEC;
    else if (!strcmp(value, "liao"))
        *type = SV_BOUNDARY_LIAO;
    else if (!strcmp(value, "periodic"))
        /* domain-periodic axis (no reference counterpart): the far plane's
         * neighbour is the near plane; both faces of the axis must say so */
        *type = SV_BOUNDARY_PERIODIC;
    else if (!strcmp(value, "cpml")) {
        *type = SV_BOUNDARY_CPML;
        if (get_int(fr, depth, key) || get_int(fr, m, key) || get_double(fr, sigma, key))
            return 1;
        /* sigma == -1 selects 0.8 m / (377 spacing) */
        if (*sigma == -1)
            *sigma = 0.8 * (double)(*m) / (377 * spacing);
        if (get_double(fr, a, key) || get_double(fr, kappa, key))
            return 1;
    } else {
        fprintf(stderr, "Error: unsupported boundary condition (%s)\n", value);
        return 1;
    }
    if (*type == SV_BOUNDARY_PEC) {
        fprintf(stderr, "Error: PEC outer boundary is not supported on this path\n");
        return 1;
    }
    return 0;
}

static int
skip_output_keyword(FILE *fr, const char *key)
{
    char buf[SV_NAME_LEN];
    fprintf(stderr, "Warning: %s output is not produced by this engine, ignored\n", key);
    return get_rest_of_line(fr, buf, sizeof(buf)) && 0;
}

/* room for `add` more far-field points; individual may be NULL (periodic set) */
static int
far_points_grow(int **ri, int **rj, int **rk, int **individual, char (**fn)[SV_NAME_LEN], int nrs, int add)
{
    int *a = (int *)realloc(*ri, (size_t)(nrs + add) * sizeof(int));
    if (a) *ri = a;
    int *b = (int *)realloc(*rj, (size_t)(nrs + add) * sizeof(int));
    if (b) *rj = b;
    int *c = (int *)realloc(*rk, (size_t)(nrs + add) * sizeof(int));
    if (c) *rk = c;
    char (*f)[SV_NAME_LEN] = (char (*)[SV_NAME_LEN])realloc(*fn, (size_t)(nrs + add) * SV_NAME_LEN);
    if (f) *fn = f;
    int ok = a && b && c && f;
    if (individual) {
        int *d = (int *)realloc(*individual, (size_t)(nrs + add) * sizeof(int));
        if (d) *individual = d;
        ok = ok && d;
    }
    if (!ok)
        fprintf(stderr, "Error: cannot allocate %d far-field points\n", nrs + add);
    return !ok;
}

/* NFFF_SPHERICAL_AREA / PERIODIC_NFFF_SPHERICAL_AREA thetares phires radius
 * thetafrom phifrom thetato phito savefile (src3d/settings.c:5532-5620,
 * 5621-5704): thetares x phires points at `radius` cells from the pool centre,
 * theta and phi (radians) stepped evenly from `from` to `to`, cell indices
 * truncated toward zero; logs xnff_set<set>_<i>_<j> when savefile != 0 */
static int
parse_spherical_area(FILE *fr, const char *key, const SvSet *set, int *nrs, int *nsets, int **ri, int **rj, int **rk,
                     int **individual, char (**fn)[SV_NAME_LEN])
{
    int tres, pres, radius, savefile, i, j, n;
    double tf, pf, tt, pt;
    if (get_int(fr, &tres, key) || get_int(fr, &pres, key) || get_int(fr, &radius, key) || get_double(fr, &tf, key)
        || get_double(fr, &pf, key) || get_double(fr, &tt, key) || get_double(fr, &pt, key) || get_int(fr, &savefile, key))
        return 1;
    if (tres < 1 || pres < 1) {
        fprintf(stderr, "Error: %s needs at least one point per direction\n", key);
        return 1;
    }
    if (far_points_grow(ri, rj, rk, individual, fn, *nrs, tres * pres))
        return 1;
    n = *nrs;
    (*nsets)++;
    fprintf(stderr, "Warning: %s set %d: its total-power image (a Gwyddion container) is not written, its points are\n",
            key, *nsets);
    for (i = 0; i < tres; i++)
        for (j = 0; j < pres; j++) {
            double theta = tres > 1 ? tf + (double)i * (tt - tf) / (tres - 1) : tf;
            double phi = pres > 1 ? pf + (double)j * (pt - pf) / (pres - 1) : pf;
            (*ri)[n] = (int)(radius * sin(theta) * cos(phi)) + set->sp.xres / 2;
            (*rj)[n] = (int)(radius * sin(theta) * sin(phi)) + set->sp.yres / 2;
            (*rk)[n] = (int)(radius * cos(theta)) + set->sp.zres / 2;
            if (individual)
                (*individual)[n] = *nsets;
            if (savefile)
                snprintf((*fn)[n], SV_NAME_LEN, "xnff_set%03d_%03d_%03d", *nsets, i, j);
            else
                (*fn)[n][0] = 0;
            n++;
        }
    *nrs = n;
    return 0;
}

/* NFFF_PLANAR_AREA ijres jkres ijfrom jkfrom ijto jkto orientation distance
 * savefile (src3d/settings.c:5705-5816): an ijres x jkres grid of points on
 * the plane normal to x (0), y (1) or z (2) at `distance` (m) from the pool
 * centre, in-plane coordinates (m) stepped evenly from `from` to `to`, each
 * divided by the cell size and truncated toward zero.  The reference divides
 * by ijres - 1 and jkres - 1 unguarded; one-point rows are refused here. */
static int
parse_planar_area(FILE *fr, const char *key, SvSet *set)
{
    SvSetFarfield *f = &set->sf;
    int ijres, jkres, orientation, savefile, i, j, n;
    double ijfrom, jkfrom, ijto, jkto, distance;
    if (get_int(fr, &ijres, key) || get_int(fr, &jkres, key) || get_double(fr, &ijfrom, key) || get_double(fr, &jkfrom, key)
        || get_double(fr, &ijto, key) || get_double(fr, &jkto, key) || get_int(fr, &orientation, key)
        || get_double(fr, &distance, key) || get_int(fr, &savefile, key))
        return 1;
    if (ijres < 2 || jkres < 2) {
        fprintf(stderr, "Error: %s needs at least two points per direction\n", key);
        return 1;
    }
    if (orientation < 0 || orientation > 2) {
        fprintf(stderr, "Error: unsupported NFFF area orientation (%d)\n", orientation);
        return 1;
    }
    if (far_points_grow(&f->ri, &f->rj, &f->rk, &f->individual, &f->source_filename, f->nrs, ijres * jkres))
        return 1;
    n = f->nrs;
    f->nsets++;
    fprintf(stderr, "Warning: %s set %d: its total-power image (a Gwyddion container) is not written, its points are\n",
            key, f->nsets);
    for (i = 0; i < ijres; i++)
        for (j = 0; j < jkres; j++) {
            const double u = ijfrom + (ijto - ijfrom) * ((double)i) / ((double)ijres - 1);
            const double v = jkfrom + (jkto - jkfrom) * ((double)j) / ((double)jkres - 1);
            if (orientation == 0) {
                f->ri[n] = (int)(distance / set->sp.dx) + set->sp.xres / 2;
                f->rj[n] = (int)(u / set->sp.dy) + set->sp.yres / 2;
                f->rk[n] = (int)(v / set->sp.dz) + set->sp.zres / 2;
            } else if (orientation == 1) {
                f->ri[n] = (int)(u / set->sp.dx) + set->sp.xres / 2;
                f->rj[n] = (int)(distance / set->sp.dy) + set->sp.yres / 2;
                f->rk[n] = (int)(v / set->sp.dz) + set->sp.zres / 2;
            } else {
                f->ri[n] = (int)(u / set->sp.dx) + set->sp.xres / 2;
                f->rj[n] = (int)(v / set->sp.dy) + set->sp.yres / 2;
                f->rk[n] = (int)(distance / set->sp.dz) + set->sp.zres / 2;
            }
            f->individual[n] = f->nsets;
            if (savefile)
                snprintf(f->source_filename[n], SV_NAME_LEN, "xnff_set%03d_%03d_%03d", f->nsets, i, j);
            else
                f->source_filename[n][0] = 0;
            n++;
        }
    f->nrs = n;
    return 0;
}

int
parse_settings(const char *filename, SvSet *set, int called_from_gsvit)
{
    (void)called_from_gsvit;   /* the reference uses it to pick message styles only */
    char key[128];
    char value[64];
    FILE *fr = fopen(filename, "r");

    if (!fr) {
        fprintf(stderr, "Error: cannot open parameter file %s\n", filename);
        return 0;
    }

    while (fscanf(fr, "%127s", key) == 1) {
        if (strchr(key, '#')) {
            char buf[256];
            if (!fgets(buf, sizeof(buf), fr))
                break;
            continue;
        }
        if (!strcmp(key, "DIMENSIONALITY")) {
            int dim;
            if (get_int(fr, &dim, key))
                goto fail;
            if (dim != 3) {
                fprintf(stderr, "Error: parameter file dimensionality should be 3\n");
                goto fail;
            }
        } else if (!strcmp(key, "VERBOSE")) {
            if (get_int(fr, &set->sc.verbose, key))
                goto fail;
        } else if (!strcmp(key, "POOL")) {
            if (get_int(fr, &set->sp.xres, key) || get_int(fr, &set->sp.yres, key) || get_int(fr, &set->sp.zres, key)
                || get_double(fr, &set->sp.dx, key) || get_double(fr, &set->sp.dy, key) || get_double(fr, &set->sp.dz, key))
                goto fail;
        } else if (!strcmp(key, "COMP")) {
            if (get_int(fr, &set->sc.nsteps, key))
                goto fail;
        } else if (!strcmp(key, "DT_MULT")) {
            if (get_double(fr, &set->sc.dtmult, key))
                goto fail;
        } else if (!strcmp(key, "GPU")) {
            if (get_int(fr, &set->sc.usegpu, key))
                goto fail;
        } else if (!strcmp(key, "UGPU")) {
            int v;
            if (get_int(fr, &v, key))
                goto fail;
            if (v >= 0 && v < SV_MAX_GPUS)
                set->sc.ugpu[v] = 1;
        } else if (!strcmp(key, "GPU_QUERY")) {
            if (get_int(fr, &set->sc.devicequery, key))
                goto fail;
        } else if (!strcmp(key, "THREADS")) {
            if (get_int(fr, &set->sc.nthreads, key))
                goto fail;
        } else if (!strcmp(key, "MATMODE_CHECK")) {
            if (get_int(fr, &set->sm.matmode_check, key))
                goto fail;
        } else if (!strcmp(key, "MEDIUM_CROP")) {
            SvSetMat *m = &set->sm;
            if (get_int(fr, &m->crop_i0, key) || get_int(fr, &m->crop_j0, key) || get_int(fr, &m->crop_k0, key)
                || get_int(fr, &m->crop_in, key) || get_int(fr, &m->crop_jn, key) || get_int(fr, &m->crop_kn, key)
                || get_double(fr, &m->crop_epsilon, key) || get_double(fr, &m->crop_mu, key)
                || get_double(fr, &m->crop_sigma, key) || get_double(fr, &m->crop_sigast, key))
                goto fail;
            m->crop = 1;
        } else if (!strcmp(key, "MEDIUM_LINEAR")) {
            if (get_rest_of_line(fr, set->sm.in_voxel_filename, SV_NAME_LEN))
                goto fail;
        } else if (!strcmp(key, "MEDIUM_VECTOR")) {
            if (get_rest_of_line(fr, set->sm.in_vector_filename, SV_NAME_LEN))
                goto fail;
        } else if (!strcmp(key, "MEDIUM_VTK_DATA")) {
            if (get_rest_of_line(fr, set->sm.in_vtk_data_filename, SV_NAME_LEN))
                goto fail;
        } else if (!strcmp(key, "SOURCE_POINT")) {
            SvSetSourcePoint *p;
            int srctype;
            set->ss.pnts = (SvSetSourcePoint *)realloc(set->ss.pnts, (set->ss.npnts + 1) * sizeof(SvSetSourcePoint));
            p = set->ss.pnts + set->ss.npnts;
            memset(p, 0, sizeof(*p));
            if (get_int(fr, &p->point_origin_position_i, key) || get_int(fr, &p->point_origin_position_j, key)
                || get_int(fr, &p->point_origin_position_k, key) || get_int(fr, &srctype, key))
                goto fail;
            p->source_mode = srctype;
            if (srctype == 0) {
                if (get_rest_of_line(fr, p->source_filename, SV_NAME_LEN))
                    goto fail;
            } else if (srctype == 1 || srctype == 2) {
                /* src3d/settings.c:4090-4140 */
                p->source_pulsewidth = 100;
                if (get_double(fr, &p->source_wl, key))
                    goto fail;
                if (srctype == 2 && get_int(fr, &p->source_pulsewidth, key))
                    goto fail;
                if (get_double(fr, &p->source_amplitude, key) || get_double(fr, &p->point_origin_theta, key)
                    || get_double(fr, &p->point_origin_phi, key))
                    goto fail;
            } else {
                fprintf(stderr, "Error: point source type %d is not supported\n", srctype);
                goto fail;
            }
            set->ss.npnts++;
        } else if (!strcmp(key, "SOURCE_TSF")) {
            SvSetSourceTSF *t = &set->ss.tsf;
            int srctype;
            if (get_int(fr, &t->box_i0, key) || get_int(fr, &t->box_j0, key) || get_int(fr, &t->box_k0, key)
                || get_int(fr, &t->box_in, key) || get_int(fr, &t->box_jn, key) || get_int(fr, &t->box_kn, key)
                || get_double(fr, &t->ia_theta, key) || get_double(fr, &t->ia_phi, key) || get_double(fr, &t->ia_psi, key)
                || get_int(fr, &srctype, key))
                goto fail;
            t->source_mode = srctype;
            if (srctype == 0) {
                if (get_rest_of_line(fr, t->source_filename, SV_NAME_LEN))
                    goto fail;
            } else if (srctype == 1 || srctype == 2) {
                /* src3d/settings.c:4253-4281 */
                if (get_double(fr, &t->source_wl, key))
                    goto fail;
                if (srctype == 2 && get_int(fr, &t->source_pulsewidth, key))
                    goto fail;
                if (get_double(fr, &t->source_amplitude, key))
                    goto fail;
            } else {
                fprintf(stderr, "Error: TSF source type %d is not supported\n", srctype);
                goto fail;
            }
            t->valid = 1;
        } else if (!strcmp(key, "SOURCE_TSFF")) {
            SvSetSourceTSFF *f = &set->ss.tsff;
            SvSetSourceTSF *t = &f->w;
            int srctype;
            if (get_int(fr, &t->box_i0, key) || get_int(fr, &t->box_j0, key) || get_int(fr, &t->box_k0, key)
                || get_int(fr, &t->box_in, key) || get_int(fr, &t->box_jn, key) || get_int(fr, &t->box_kn, key)
                || get_double(fr, &f->focused_thetamax, key) || get_double(fr, &f->focused_fdist, key)
                || get_double(fr, &f->focused_pol, key) || get_int(fr, &f->focused_nip, key)
                || get_int(fr, &f->focused_mip, key) || get_int(fr, &srctype, key))
                goto fail;
            t->source_mode = srctype;
            if (srctype == 0) {
                if (get_rest_of_line(fr, t->source_filename, SV_NAME_LEN))
                    goto fail;
            } else if (srctype == 1 || srctype == 2) {
                if (get_double(fr, &t->source_wl, key))
                    goto fail;
                if (srctype == 2 && get_int(fr, &t->source_pulsewidth, key))
                    goto fail;
                if (get_double(fr, &t->source_amplitude, key))
                    goto fail;
            } else {
                fprintf(stderr, "Error: TSFF source type %d is not supported\n", srctype);
                goto fail;
            }
            t->valid = 1;
        } else if (!strcmp(key, "SOURCE_EXT")) {
            int *v[12] = { &set->ss.ext.i, &set->ss.ext.j, &set->ss.ext.k, &set->ss.ext.ijstart, &set->ss.ext.jkstart,
                           &set->ss.ext.extxres, &set->ss.ext.extyres, &set->ss.ext.iextfrom, &set->ss.ext.jextfrom,
                           &set->ss.ext.iextto, &set->ss.ext.jextto, &set->ss.ext.shift };
            int q;
            for (q = 0; q < 12; q++)
                if (get_int(fr, v[q], key))
                    goto fail;
            for (q = 0; q < 6; q++)
                if (get_token(fr, set->ss.ext.filebase[q], SV_NAME_LEN))
                    goto fail;
            set->ss.ext.valid = 1;
        } else if (!strcmp(key, "TSFF_SHIFT")) {
            SvSetSourceTSFF *f = &set->ss.tsff;
            if (get_double(fr, &f->xshift, key) || get_double(fr, &f->yshift, key) || get_double(fr, &f->zshift, key))
                goto fail;
        } else if (!strcmp(key, "TSFF_SKIP")) {
            SvSetSourceTSF *t = &set->ss.tsff.w;
            if (get_token(fr, value, sizeof(value)))
                goto fail;
            if (!strcmp(value, "i0")) t->box_boundary_skipi0 = 1;
            else if (!strcmp(value, "in")) t->box_boundary_skipin = 1;
            else if (!strcmp(value, "j0")) t->box_boundary_skipj0 = 1;
            else if (!strcmp(value, "jn")) t->box_boundary_skipjn = 1;
            else if (!strcmp(value, "k0")) t->box_boundary_skipk0 = 1;
            else if (!strcmp(value, "kn")) t->box_boundary_skipkn = 1;
            else if (!strcmp(value, "depth")) {
                /* the focused faces never read the skip depth (src3d/source.c:773-930, 1892-2060) */
                int depth;
                if (get_int(fr, &depth, key))
                    goto fail;
            }
        } else if (!strcmp(key, "SOURCE_SF")) {
            /* src3d/settings.c:4698-4760: theta phi psi, then a waveform file (0)
             * or generated-waveform parameters (1: wavelength amplitude, 2:
             * wavelength pulse width amplitude) */
            SvSetSourceTSF *t = &set->ss.sf;
            int srctype;
            if (get_double(fr, &t->ia_theta, key) || get_double(fr, &t->ia_phi, key) || get_double(fr, &t->ia_psi, key)
                || get_int(fr, &srctype, key))
                goto fail;
            t->source_mode = srctype;
            if (srctype == 0) {
                if (get_rest_of_line(fr, t->source_filename, SV_NAME_LEN))
                    goto fail;
            } else if (srctype == 1 || srctype == 2) {
                if (get_double(fr, &t->source_wl, key))
                    goto fail;
                if (srctype == 2 && get_int(fr, &t->source_pulsewidth, key))
                    goto fail;
                if (get_double(fr, &t->source_amplitude, key))
                    goto fail;
            } else {
                fprintf(stderr, "Error: SF source type %d is not supported\n", srctype);
                goto fail;
            }
            t->valid = 1;
        } else if (!strcmp(key, "TSF_SKIP")) {
            SvSetSourceTSF *t = &set->ss.tsf;
            if (get_token(fr, value, sizeof(value)))
                goto fail;
            if (!strcmp(value, "i0")) t->box_boundary_skipi0 = 1;
            else if (!strcmp(value, "in")) t->box_boundary_skipin = 1;
            else if (!strcmp(value, "j0")) t->box_boundary_skipj0 = 1;
            else if (!strcmp(value, "jn")) t->box_boundary_skipjn = 1;
            else if (!strcmp(value, "k0")) t->box_boundary_skipk0 = 1;
            else if (!strcmp(value, "kn")) t->box_boundary_skipkn = 1;
            else if (!strcmp(value, "depth")) {
                if (get_int(fr, &t->box_boundary_skipdepth, key))
                    goto fail;
            }
        } else if (!strcmp(key, "SOURCE_LTSF")) {
            /* src3d/settings.c:4342-4433: box, theta phi psi, layer count, per layer
             * z position and eps mu sigma sigma* (or 99 and a database material),
             * source type and file */
            SvSetSourceLTSF *t = &set->ss.ltsf;
            int srctype, q;
            if (get_int(fr, &t->box_i0, key) || get_int(fr, &t->box_j0, key) || get_int(fr, &t->box_k0, key)
                || get_int(fr, &t->box_in, key) || get_int(fr, &t->box_jn, key) || get_int(fr, &t->box_kn, key)
                || get_double(fr, &t->ia_theta, key) || get_double(fr, &t->ia_phi, key) || get_double(fr, &t->ia_psi, key)
                || get_int(fr, &t->layered_count, key))
                goto fail;
            if (t->layered_count < 0 || t->layered_count > SV_LTSF_MAX_LAYERS) {
                fprintf(stderr, "Error: SOURCE_LTSF layer count %d outside 0..%d\n", t->layered_count, SV_LTSF_MAX_LAYERS);
                goto fail;
            }
            for (q = 0; q < t->layered_count; q++) {
                char *end;
                if (get_int(fr, &t->layered_zpos[q], key) || get_token(fr, value, sizeof(value)))
                    goto fail;
                t->layered_material[q][0] = 0;
                if (!strcmp(value, "99")) {
                    /* a database entry, resolved at plan time */
                    if (get_token(fr, t->layered_material[q], SV_NAME_LEN))
                        goto fail;
                    t->layered_epsilon[q] = t->layered_mu[q] = 1;
                    t->layered_sigma[q] = t->layered_sigast[q] = 0;
                    continue;
                }
                t->layered_epsilon[q] = strtod(value, &end);
                if (*end != '\0') {
                    fprintf(stderr, "Error parsing double value (key %s)\n", key);
                    goto fail;
                }
                if (get_double(fr, &t->layered_mu[q], key) || get_double(fr, &t->layered_sigma[q], key)
                    || get_double(fr, &t->layered_sigast[q], key))
                    goto fail;
            }
            if (get_int(fr, &srctype, key))
                goto fail;
            t->source_mode = srctype;
            if (srctype == 1 || srctype == 2) {
                /* generated at plan time, src3d/settings.c:4400-4428 */
                if (set->sc.nsteps == 0) {
                    fprintf(stderr, "Error: COMP parameter (number of steps) should precede generated source command.\n");
                    goto fail;
                }
                if (get_double(fr, &t->source_wl, key) || (srctype == 2 && get_int(fr, &t->source_pulsewidth, key))
                    || get_double(fr, &t->source_amplitude, key))
                    goto fail;
            } else if (srctype != 0) {
                fprintf(stderr, "Error: SOURCE_LTSF source type %d is not supported (0 file, 1 sine, 2 pulse; the "
                                "broadband type 3 needs Gwyddion's FFT)\n", srctype);
                goto fail;
            } else if (get_rest_of_line(fr, t->source_filename, SV_NAME_LEN))
                goto fail;
            t->valid = 1;
        } else if (!strcmp(key, "LTSF_SKIP")) {
            SvSetSourceLTSF *t = &set->ss.ltsf;
            if (get_token(fr, value, sizeof(value)))
                goto fail;
            if (!strcmp(value, "i0")) t->box_boundary_skipi0 = 1;
            else if (!strcmp(value, "in")) t->box_boundary_skipin = 1;
            else if (!strcmp(value, "j0")) t->box_boundary_skipj0 = 1;
            else if (!strcmp(value, "jn")) t->box_boundary_skipjn = 1;
            else if (!strcmp(value, "k0")) t->box_boundary_skipk0 = 1;
            else if (!strcmp(value, "kn")) t->box_boundary_skipkn = 1;
            else if (!strcmp(value, "depth")) {
                if (get_int(fr, &t->box_boundary_skipdepth, key))
                    goto fail;
            }
        } else if (!strcmp(key, "TSF_GAUSSIAN_Z")) {
            SvSetSourceTSF *t = &set->ss.tsf;
            t->gaussian = 1;
            if (get_double(fr, &t->gaussian_fxpos, key) || get_double(fr, &t->gaussian_fypos, key)
                || get_double(fr, &t->gaussian_rx, key) || get_double(fr, &t->gaussian_ry, key))
                goto fail;
        } else if (!strcmp(key, "TSF_FIBER_Z")) {
            SvSetSourceTSF *t = &set->ss.tsf;
            t->fiber = 1;
            if (get_double(fr, &t->fiber_fxpos, key) || get_double(fr, &t->fiber_fypos, key)
                || get_double(fr, &t->fiber_radius, key) || get_double(fr, &t->fiber_cutoff, key)
                || get_double(fr, &t->fiber_epsilon_core, key) || get_double(fr, &t->fiber_epsilon_cladding, key))
                goto fail;
        } else if (!strcmp(key, "TSF_RADIAL_Z")) {
            /* the reference's CPU faces evaluate this profile at a loop index
             * that is never set on the i faces (src3d/source.c:574-580) and its
             * GPU path leaves it out (src3d/gpu.c:2405) */
            fprintf(stderr, "Error: TSF_RADIAL_Z is not supported by the MI355X engine\n");
            goto fail;
        } else if (!strcmp(key, "LTSF_GAUSSIAN")) {
            SvSetSourceLTSF *t = &set->ss.ltsf;
            t->gaussian = 1;
            if (get_double(fr, &t->gaussian_fxpos, key) || get_double(fr, &t->gaussian_fypos, key)
                || get_double(fr, &t->gaussian_rx, key) || get_double(fr, &t->gaussian_ry, key))
                goto fail;
        } else if (!strcmp(key, "LTSF_FIBER")) {
            SvSetSourceLTSF *t = &set->ss.ltsf;
            t->fiber = 1;
            if (get_double(fr, &t->fiber_fxpos, key) || get_double(fr, &t->fiber_fypos, key)
                || get_double(fr, &t->fiber_radius, key) || get_double(fr, &t->fiber_cutoff, key)
                || get_double(fr, &t->fiber_epsilon_core, key) || get_double(fr, &t->fiber_epsilon_cladding, key))
                goto fail;
        } else if (!strcmp(key, "LTSF_RADIAL")) {
            /* parsed by the reference but not used by its LTSF faces */
            double v[4];
            if (get_double(fr, v, key) || get_double(fr, v + 1, key) || get_double(fr, v + 2, key)
                || get_double(fr, v + 3, key))
                goto fail;
        } else if (!strcmp(key, "TSFSOURCE_MATERIAL")) {
            SvSetSourceTSF *t = &set->ss.tsf;
            if (get_double(fr, &t->layered_epsilon, key) || get_double(fr, &t->layered_mu, key)
                || get_double(fr, &t->layered_sigma, key) || get_double(fr, &t->layered_sigast, key))
                goto fail;
        } else if (!strcmp(key, "SOURCE_WAVELENGTH")) {
            if (get_double(fr, &set->ss.lambda_min, key) || get_double(fr, &set->ss.lambda_center, key)
                || get_double(fr, &set->ss.lambda_max, key))
                goto fail;
        } else if (!strcmp(key, "OUT_POINT")) {
            SvOutputPar *o;
            set->so.pnts = (SvOutputPar *)realloc(set->so.pnts, (set->so.npnts + 1) * sizeof(SvOutputPar));
            o = set->so.pnts + set->so.npnts;
            memset(o, 0, sizeof(*o));
            if (get_component(fr, &o->component, key) || get_int(fr, &o->step, key) || get_int(fr, &o->i, key)
                || get_int(fr, &o->j, key) || get_int(fr, &o->k, key))
                goto fail;
            if (o->component == SV_COMP_ALLFFT) {
                /* src3d/settings.c:4803-4823: step window, -1 for the whole run,
                 * cropped to the step count known at this point of the file */
                if (get_int(fr, &o->start, key) || get_int(fr, &o->stop, key))
                    goto fail;
                if (o->start == -1)
                    o->start = 0;
                if (o->stop == -1)
                    o->stop = set->sc.nsteps;
                if (o->start > set->sc.nsteps - 4) {
                    o->start = set->sc.nsteps - 4;
                    fprintf(stderr, "Warning: requesting FFT output outside of the data range, cropped to %d\n", o->start);
                }
                if (o->stop > set->sc.nsteps) {
                    o->stop = set->sc.nsteps;
                    fprintf(stderr, "Warning: requesting FFT output outside of the data range, cropped to %d\n", o->stop);
                }
            }
            if (get_rest_of_line(fr, o->filebase, SV_NAME_LEN))
                goto fail;
            set->so.npnts++;
        } else if (!strcmp(key, "OUT_FORCE")) {
            /* src3d/settings.c:5018-5040: step, box, file */
            SvOutputForce *o;
            set->so.forces = (SvOutputForce *)realloc(set->so.forces, (set->so.nforces + 1) * sizeof(SvOutputForce));
            o = set->so.forces + set->so.nforces;
            memset(o, 0, sizeof(*o));
            if (get_int(fr, &o->step, key) || get_int(fr, &o->box_i0, key) || get_int(fr, &o->box_j0, key)
                || get_int(fr, &o->box_k0, key) || get_int(fr, &o->box_in, key) || get_int(fr, &o->box_jn, key)
                || get_int(fr, &o->box_kn, key) || get_rest_of_line(fr, o->filename, SV_NAME_LEN))
                goto fail;
            set->so.nforces++;
        } else if (!strcmp(key, "OUT_SUM") || !strcmp(key, "OUT_SUMTAB")) {
            /* src3d/settings.c:4937-4981: component step i0 j0 k0 in jn kn eps mu sigma sigast file;
             * OUT_SUMTAB (:4982-5017) names a database material instead of eps mu sigma sigast */
            SvOutputSum *o;
            char comp[16];
            set->so.sums = (SvOutputSum *)realloc(set->so.sums, (set->so.nsums + 1) * sizeof(SvOutputSum));
            o = set->so.sums + set->so.nsums;
            memset(o, 0, sizeof(*o));
            if (fscanf(fr, "%15s", comp) != 1)
                goto fail;
            if (!strcmp(comp, "Ex") || !strcmp(comp, "ex")) o->component = SV_SUM_EX;
            else if (!strcmp(comp, "Ey") || !strcmp(comp, "ey")) o->component = SV_SUM_EY;
            else if (!strcmp(comp, "Ez") || !strcmp(comp, "ez")) o->component = SV_SUM_EZ;
            else if (!strcmp(comp, "All") || !strcmp(comp, "all")) o->component = SV_SUM_ALL;
            else if (!strcmp(comp, "Abs") || !strcmp(comp, "abs")) o->component = SV_SUM_ABS;
            else if (!strcmp(comp, "Max") || !strcmp(comp, "max")) o->component = SV_SUM_MAX;
            else {
                fprintf(stderr, "Error parsing field component (key %s)\n", key);
                goto fail;
            }
            if (get_int(fr, &o->step, key) || get_int(fr, &o->box_i0, key) || get_int(fr, &o->box_j0, key)
                || get_int(fr, &o->box_k0, key) || get_int(fr, &o->box_in, key) || get_int(fr, &o->box_jn, key)
                || get_int(fr, &o->box_kn, key))
                goto fail;
            if (!strcmp(key, "OUT_SUMTAB")) {
                o->stringbased = 1;
                if (get_token(fr, o->string, sizeof(o->string)))
                    goto fail;
            } else if (get_double(fr, &o->layered_epsilon, key) || get_double(fr, &o->layered_mu, key)
                       || get_double(fr, &o->layered_sigma, key) || get_double(fr, &o->layered_sigast, key))
                goto fail;
            if (get_rest_of_line(fr, o->filename, SV_NAME_LEN))
                goto fail;
            set->so.nsums++;
        } else if (!strcmp(key, "OUT_PLANE")) {
            /* src3d/settings.c:4869-4901: component step start stop format i j k file */
            SvOutputRegion *o;
            set->so.plns = (SvOutputRegion *)realloc(set->so.plns, (set->so.nplns + 1) * sizeof(SvOutputRegion));
            o = set->so.plns + set->so.nplns;
            memset(o, 0, sizeof(*o));
            if (get_component(fr, &o->component, key) || get_int(fr, &o->step, key) || get_int(fr, &o->start, key)
                || get_int(fr, &o->stop, key) || get_int(fr, &o->format, key) || get_int(fr, &o->i, key)
                || get_int(fr, &o->j, key) || get_int(fr, &o->k, key))
                goto fail;
            if (get_rest_of_line(fr, o->filebase, SV_NAME_LEN))
                goto fail;
            set->so.nplns++;
        } else if (!strcmp(key, "OUT_VOLUME")) {
            /* src3d/settings.c:4902-4936: component step start stop format file; start -1 = 0,
             * stop -1 = the step count read so far */
            SvOutputRegion *o;
            int q, nabs = 0;
            set->so.cubs = (SvOutputRegion *)realloc(set->so.cubs, (set->so.ncubs + 1) * sizeof(SvOutputRegion));
            o = set->so.cubs + set->so.ncubs;
            memset(o, 0, sizeof(*o));
            if (get_volume_component(fr, &o->component, key) || get_int(fr, &o->step, key)
                || get_int(fr, &o->start, key) || get_int(fr, &o->stop, key) || get_int(fr, &o->format, key))
                goto fail;
            if (o->start == -1)
                o->start = 0;
            if (o->stop == -1)
                o->stop = set->sc.nsteps;
            if (get_rest_of_line(fr, o->filebase, SV_NAME_LEN))
                goto fail;
            for (q = 0; q < set->so.ncubs; q++)
                if (set->so.cubs[q].component >= SV_OVOLUME_ABS)
                    nabs++;
            if (o->component >= SV_OVOLUME_ABS)
                o->k = nabs;
            set->so.ncubs++;
        } else if (!strcmp(key, "OUT_IMAGE") || !strcmp(key, "OUT_FILE")) {
            skip_output_keyword(fr, key);
        } else if (!strcmp(key, "NFFF_SPHERICAL_AREA")) {
            SvSetFarfield *f = &set->sf;
            if (parse_spherical_area(fr, key, set, &f->nrs, &f->nsets, &f->ri, &f->rj, &f->rk, &f->individual,
                                     &f->source_filename))
                goto fail;
        } else if (!strcmp(key, "PERIODIC_NFFF_SPHERICAL_AREA")) {
            SvSetPFarfield *f = &set->spf;
            if (parse_spherical_area(fr, key, set, &f->nrs, &f->nsets, &f->ri, &f->rj, &f->rk, NULL, &f->source_filename))
                goto fail;
        } else if (!strcmp(key, "NFFF_PLANAR_AREA")) {
            if (parse_planar_area(fr, key, set))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_ALL")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dx, &b->bx0, &b->depth_bx0, &b->m_bx0, &b->sigma_bx0, &b->a_bx0, &b->kappa_bx0))
                goto fail;
            b->bxn = b->by0 = b->byn = b->bz0 = b->bzn = b->bx0;
            b->depth_bxn = b->depth_by0 = b->depth_byn = b->depth_bz0 = b->depth_bzn = b->depth_bx0;
            b->m_bxn = b->m_by0 = b->m_byn = b->m_bz0 = b->m_bzn = b->m_bx0;
            b->sigma_bxn = b->sigma_by0 = b->sigma_byn = b->sigma_bz0 = b->sigma_bzn = b->sigma_bx0;
            b->a_bxn = b->a_by0 = b->a_byn = b->a_bz0 = b->a_bzn = b->a_bx0;
            b->kappa_bxn = b->kappa_by0 = b->kappa_byn = b->kappa_bz0 = b->kappa_bzn = b->kappa_bx0;
        } else if (!strcmp(key, "ADE_CPML_CONSISTENT")) {
            /* no reference counterpart: consistent CPML of dispersive cells */
            if (get_int(fr, &set->sb.ade_cpml_consistent, key))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_X0")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dx, &b->bx0, &b->depth_bx0, &b->m_bx0, &b->sigma_bx0, &b->a_bx0, &b->kappa_bx0))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_XN")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dx, &b->bxn, &b->depth_bxn, &b->m_bxn, &b->sigma_bxn, &b->a_bxn, &b->kappa_bxn))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_Y0")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dy, &b->by0, &b->depth_by0, &b->m_by0, &b->sigma_by0, &b->a_by0, &b->kappa_by0))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_YN")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dy, &b->byn, &b->depth_byn, &b->m_byn, &b->sigma_byn, &b->a_byn, &b->kappa_byn))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_Z0")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dz, &b->bz0, &b->depth_bz0, &b->m_bz0, &b->sigma_bz0, &b->a_bz0, &b->kappa_bz0))
                goto fail;
        } else if (!strcmp(key, "BOUNDARY_ZN")) {
            SvSetBoundary *b = &set->sb;
            if (parse_cpml_face(fr, key, set->sp.dz, &b->bzn, &b->depth_bzn, &b->m_bzn, &b->sigma_bzn, &b->a_bzn, &b->kappa_bzn))
                goto fail;
        } else if (!strcmp(key, "MBOUNDARY_ALL")) {
            /* src3d/settings.c:5109-5128: one condition, then the x0 xn y0 yn z0 zn plane positions */
            SvSetMediumBoundary *m = &set->smb;
            if (get_token(fr, value, sizeof(value)) || get_int(fr, &m->bx0pos, key) || get_int(fr, &m->bxnpos, key)
                || get_int(fr, &m->by0pos, key) || get_int(fr, &m->bynpos, key) || get_int(fr, &m->bz0pos, key)
                || get_int(fr, &m->bznpos, key))
                goto fail;
            if (strcmp(value, "periodic")) {
                fprintf(stderr, "Error: unsupported medium boundary condition (%s)\n", value);
                goto fail;
            }
            m->bx0 = m->bxn = m->by0 = m->byn = m->bz0 = m->bzn = SV_BOUNDARY_PERIODIC;
        } else if (!strncmp(key, "MBOUNDARY_", 10)) {
            /* src3d/settings.c:5047-5128 */
            int pos, *type = NULL, *ppos = NULL;
            if (get_token(fr, value, sizeof(value)) || get_int(fr, &pos, key))
                goto fail;
            if (strcmp(value, "periodic")) {
                fprintf(stderr, "Error: unsupported medium boundary condition (%s)\n", value);
                goto fail;
            }
            if (!strcmp(key, "MBOUNDARY_X0")) { type = &set->smb.bx0; ppos = &set->smb.bx0pos; }
            else if (!strcmp(key, "MBOUNDARY_XN")) { type = &set->smb.bxn; ppos = &set->smb.bxnpos; }
            else if (!strcmp(key, "MBOUNDARY_Y0")) { type = &set->smb.by0; ppos = &set->smb.by0pos; }
            else if (!strcmp(key, "MBOUNDARY_YN")) { type = &set->smb.byn; ppos = &set->smb.bynpos; }
            else if (!strcmp(key, "MBOUNDARY_Z0")) { type = &set->smb.bz0; ppos = &set->smb.bz0pos; }
            else if (!strcmp(key, "MBOUNDARY_ZN")) { type = &set->smb.bzn; ppos = &set->smb.bznpos; }
            else {
                fprintf(stderr, "Error: unsupported keyword %s\n", key);
                goto fail;
            }
            *type = SV_BOUNDARY_PERIODIC;
            *ppos = pos;
        } else if (!strcmp(key, "NFFF")) {
            SvSetFarfield *f = &set->sf;
            if (get_int(fr, &f->box_i0, key) || get_int(fr, &f->box_j0, key) || get_int(fr, &f->box_k0, key)
                || get_int(fr, &f->box_in, key) || get_int(fr, &f->box_jn, key) || get_int(fr, &f->box_kn, key)
                || get_int(fr, &f->sumfrom, key) || get_int(fr, &f->sumto, key))
                goto fail;
        } else if (!strcmp(key, "NFFF_SKIP")) {
            SvSetFarfield *f = &set->sf;
            int *a = NULL;
            if (get_token(fr, value, sizeof(value)))
                goto fail;
            if (!strcmp(value, "i0")) { f->box_boundary_skipi0 = 1; a = &f->skipi0_jmin; }
            else if (!strcmp(value, "in")) { f->box_boundary_skipin = 1; a = &f->skipin_jmin; }
            else if (!strcmp(value, "j0")) { f->box_boundary_skipj0 = 1; a = &f->skipj0_imin; }
            else if (!strcmp(value, "jn")) { f->box_boundary_skipjn = 1; a = &f->skipjn_imin; }
            else if (!strcmp(value, "k0")) { f->box_boundary_skipk0 = 1; a = &f->skipk0_imin; }
            else if (!strcmp(value, "kn")) { f->box_boundary_skipkn = 1; a = &f->skipkn_imin; }
            if (a && (get_int(fr, a, key) || get_int(fr, a + 1, key) || get_int(fr, a + 2, key) || get_int(fr, a + 3, key)))
                goto fail;
        } else if (!strcmp(key, "NFFF_RAMAHI_POINT")) {
            SvSetFarfield *f = &set->sf;
            int n = f->nrs + 1;
            f->ri = (int *)realloc(f->ri, n * sizeof(int));
            f->rj = (int *)realloc(f->rj, n * sizeof(int));
            f->rk = (int *)realloc(f->rk, n * sizeof(int));
            f->individual = (int *)realloc(f->individual, n * sizeof(int));
            f->source_filename = (char (*)[SV_NAME_LEN])realloc(f->source_filename, n * SV_NAME_LEN);
            if (get_int(fr, &f->ri[n - 1], key) || get_int(fr, &f->rj[n - 1], key) || get_int(fr, &f->rk[n - 1], key))
                goto fail;
            if (get_rest_of_line(fr, f->source_filename[n - 1], SV_NAME_LEN))
                goto fail;
            f->individual[n - 1] = 0;
            f->nrs = n;
        } else if (!strcmp(key, "PERIODIC_NFFF")) {
            /* src3d/settings.c:3424-3446: box vertices, then pimin pjmin pimax pjmax */
            SvSetPFarfield *f = &set->spf;
            if (get_int(fr, &f->box_i0, key) || get_int(fr, &f->box_j0, key) || get_int(fr, &f->box_k0, key)
                || get_int(fr, &f->box_in, key) || get_int(fr, &f->box_jn, key) || get_int(fr, &f->box_kn, key)
                || get_int(fr, &f->pimin, key) || get_int(fr, &f->pjmin, key) || get_int(fr, &f->pimax, key)
                || get_int(fr, &f->pjmax, key))
                goto fail;
        } else if (!strcmp(key, "PERIODIC_NFFF_POSTPROCESS")) {
            if (get_int(fr, &set->spf.postprocess, key) || get_int(fr, &set->spf.ppstart, key))
                goto fail;
        } else if (!strcmp(key, "PERIODIC_NFFF_SKIP")) {
            /* src3d/settings.c:3452-3474: k0 or kn, then imin jmin imax jmax */
            SvSetPFarfield *f = &set->spf;
            int *a = NULL;
            if (get_token(fr, value, sizeof(value)))
                goto fail;
            if (!strcmp(value, "k0")) { f->box_boundary_skipk0 = 1; a = &f->skipk0_imin; }
            else if (!strcmp(value, "kn")) { f->box_boundary_skipkn = 1; a = &f->skipkn_imin; }
            if (a && (get_int(fr, a, key) || get_int(fr, a + 1, key) || get_int(fr, a + 2, key) || get_int(fr, a + 3, key)))
                goto fail;
        } else if (!strcmp(key, "PERIODIC_NFFF_RAMAHI_POINT")) {
            /* src3d/settings.c:3475-3499 */
            SvSetPFarfield *f = &set->spf;
            int n = f->nrs + 1;
            f->ri = (int *)realloc(f->ri, n * sizeof(int));
            f->rj = (int *)realloc(f->rj, n * sizeof(int));
            f->rk = (int *)realloc(f->rk, n * sizeof(int));
            f->source_filename = (char (*)[SV_NAME_LEN])realloc(f->source_filename, n * SV_NAME_LEN);
            if (get_int(fr, &f->ri[n - 1], key) || get_int(fr, &f->rj[n - 1], key) || get_int(fr, &f->rk[n - 1], key))
                goto fail;
            if (get_rest_of_line(fr, f->source_filename[n - 1], SV_NAME_LEN))
                goto fail;
            f->nrs = n;
        } else if (!strcmp(key, "NFFF_DFT")) {
            /* running-DFT accumulation on the NFFF box: NFFF_DFT n lambda_1 .. lambda_n file */
            SvSetFarfield *f = &set->sf;
            int n, q;
            if (get_int(fr, &n, key) || n <= 0 || n > 4096)
                goto fail;
            f->dft_lambda = (double *)realloc(f->dft_lambda, n * sizeof(double));
            for (q = 0; q < n; q++)
                if (get_double(fr, &f->dft_lambda[q], key))
                    goto fail;
            f->ndft = n;
            if (get_rest_of_line(fr, f->dft_filename, SV_NAME_LEN))
                goto fail;
        } else if (!strcmp(key, "NFFF_DFT_PATTERN")) {
            /* far-field pattern of the running DFT (no reference counterpart):
             * NFFF_DFT_PATTERN ntheta nphi file */
            SvSetFarfield *f = &set->sf;
            if (get_int(fr, &f->dft_pattern_ntheta, key) || get_int(fr, &f->dft_pattern_nphi, key)
                || f->dft_pattern_ntheta < 1 || f->dft_pattern_nphi < 1
                || (long)f->dft_pattern_ntheta * f->dft_pattern_nphi > (1L << 22))
                goto fail;
            if (get_rest_of_line(fr, f->dft_pattern_filename, SV_NAME_LEN))
                goto fail;
        } else if (!strcmp(key, "FLUX_DFT")) {
            /* flux spectra through planes (no reference counterpart):
             * FLUX_DFT i0 j0 i1 j1 n k_1 .. k_n m lambda_1 .. lambda_m file */
            SvSetFlux *x = &set->sx;
            int n, q;
            if (get_int(fr, &x->i0, key) || get_int(fr, &x->j0, key) || get_int(fr, &x->i1, key)
                || get_int(fr, &x->j1, key) || get_int(fr, &n, key) || n <= 0 || n > 256)
                goto fail;
            x->k = (int *)realloc(x->k, n * sizeof(int));
            for (q = 0; q < n; q++)
                if (get_int(fr, &x->k[q], key))
                    goto fail;
            x->nplanes = n;
            if (get_int(fr, &n, key) || n <= 0 || n > 1024)
                goto fail;
            x->lambda = (double *)realloc(x->lambda, n * sizeof(double));
            for (q = 0; q < n; q++)
                if (get_double(fr, &x->lambda[q], key) || x->lambda[q] <= 0)
                    goto fail;
            x->nw = n;
            if (get_rest_of_line(fr, x->filename, SV_NAME_LEN))
                goto fail;
        } else if (!strcmp(key, "DESIGN_BOX")) {
            SvSetAdjoint *a = &set->sa;
            if (get_int(fr, &a->design_i0, key) || get_int(fr, &a->design_j0, key) || get_int(fr, &a->design_k0, key)
                || get_int(fr, &a->design_in, key) || get_int(fr, &a->design_jn, key) || get_int(fr, &a->design_kn, key))
                goto fail;
            a->design = 1;
        } else if (!strcmp(key, "OBJECTIVE_DFT")) {
            SvSetAdjoint *a = &set->sa;
            if (get_int(fr, &a->obj_i0, key) || get_int(fr, &a->obj_j0, key) || get_int(fr, &a->obj_k0, key)
                || get_int(fr, &a->obj_in, key) || get_int(fr, &a->obj_jn, key) || get_int(fr, &a->obj_kn, key)
                || get_double(fr, &a->obj_lambda, key) || a->obj_lambda <= 0)
                goto fail;
            a->objective = 1;
        } else {
            fprintf(stderr, "Error: keyword %s is not supported by the MI355X engine\n", key);
            goto fail;
        }
    }

    /* src3d/settings.c:5823-5828 */
    if (set->smb.bxnpos == -1) set->smb.bxnpos = set->sp.xres;
    if (set->smb.bynpos == -1) set->smb.bynpos = set->sp.yres;
    if (set->smb.bznpos == -1) set->smb.bznpos = set->sp.zres;
    set->plan.dt = plan_dt(set);

    fclose(fr);
    return 1;

fail:
    fclose(fr);
    return 0;
}

/* exported for the planner: generated-source samples after the "%g" round trip */
int
sv_generated_source_samples(int type, int nsteps, double wavelength, int pulse_width, double dt,
                            double amplitude, double *out)
{
    int i;
    if (sv_generate_waveform(type, nsteps, wavelength, pulse_width, dt, amplitude, out))
        return 1;
    for (i = 0; i < nsteps; i++)
        out[i] = round_g6(out[i]);
    return 0;
}

double
sv_round_g6(double v)
{
    return round_g6(v);
}
