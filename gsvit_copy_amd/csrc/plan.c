/*
 * plan.c -- planner for the 3-D time-stepping path: time step, sources,
 * material rasterisation, internal-boundary map, CPML grading and far-field
 * geometry.  Restates the host-side preparation that GSvit performs before its
 * step loop (src3d/plan.c:5442-5604, init_and_make_plan) for the entities and
 * materials that the hot path consumes.
 *
 * Deliberately reproduced reference behaviour (each cited where it happens):
 *   - materials are scanned as float and widened (src3d/plan.c:3112-3300);
 *   - material de-duplication with almost_eq, vacuum at index 0, including the
 *     cp3_phi[2] / cp3_a[2] comparison (src3d/plan.c:2318-2367, 2409-2428);
 *   - sphere voxel test at cell centres in float (src3d/plan.c:2136-2143, 4378-4391);
 *   - matmode selection including the electric+magnetic -> NONE case of
 *     test_script_file (src3d/plan.c:573-711, 5489-5518);
 *   - internal-boundary map on faces and 26-neighbour differences
 *     (src3d/plan.c:5381-5417);
 *   - a vector file naming optical-database materials is read through its
 *     rewrite (every number through float and "%g", database.c).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include "gsvit_internal.h"

#define EPS0 SV_EPSILON_0
#define MU0 SV_MU_0

int
sv_material_is_tabulated(int type)
{
    return type == SV_MAT_LINTAB || type == SV_MAT_DRUDE || type == SV_MAT_PLRC || type == SV_MAT_ADE
           || type == SV_MAT_CP || type == SV_MAT_CP3;
}

int
sv_material_is_dispersive(int type)
{
    return type == SV_MAT_DRUDE || type == SV_MAT_CP3 || type == SV_MAT_CP || type == SV_MAT_ADE
           || type == SV_MAT_PLRC;
}

/* ------------------------------------------------------------------ */
/* material accessors (src3d/yee.c:1265-1682)                          */
/* ------------------------------------------------------------------ */

int
sv_pool_epsilon_sigma(const SvPool *mp, int i, int j, int k, double *eps, double *sig)
{
    const SvYeeData *d = mp->d;
    int matindex = (mp->nmat > 0) ? d->mat->data[i][j][k] : -1;
    if (matindex > 0 && sv_material_is_tabulated(mp->mats[matindex].type)) {
        *eps = mp->mats[matindex].epsilon;
        *sig = mp->mats[matindex].sigma;
    } else if (d->epsilon && d->sigma) {
        *eps = d->epsilon->data[i][j][k];
        *sig = d->sigma->data[i][j][k];
    } else {
        *eps = EPS0;
        *sig = 0;
    }
    return matindex;
}

int
sv_pool_mu_sigast(const SvPool *mp, int i, int j, int k, double *mu, double *sigast)
{
    const SvYeeData *d = mp->d;
    const SvSet *set = mp->set;
    int matindex = (mp->nmat > 0) ? d->mat->data[i][j][k] : -1;
    if (matindex > 0 && sv_material_is_tabulated(mp->mats[matindex].type)) {
        *mu = mp->mats[matindex].mu;
        *sigast = mp->mats[matindex].sigast;
    } else if (set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_MAGNETIC) {
        *mu = d->mu->data[i][j][k];
        *sigast = d->sigast->data[i][j][k];
    } else {
        *mu = MU0;
        *sigast = 0;
    }
    return matindex;
}

void
sv_pool_get_cabs(const SvPool *mp, int i, int j, int k,
                 double *cax, double *cay, double *caz, double *cbx, double *cby, double *cbz)
{
    const SvYeeData *d = mp->d;
    double dt = d->dt;
    double ex, ey, ez, e, sx, sy, sz, s;
    int nx, ny, nz;

    sv_pool_epsilon_sigma(mp, i, j, k, &e, &s);
    ex = ey = ez = e;
    sx = sy = sz = s;

    if (d->bnds->data[i][j][k] > 0) {
        nx = ny = nz = 1;
        if (i > 0) {
            sv_pool_epsilon_sigma(mp, i - 1, j, k, &e, &s);
            ey += e; sy += s; ez += e; sz += s; ny++; nz++;
        }
        if (j > 0) {
            sv_pool_epsilon_sigma(mp, i, j - 1, k, &e, &s);
            ex += e; sx += s; ez += e; sz += s; nx++; nz++;
        }
        if (k > 0) {
            sv_pool_epsilon_sigma(mp, i, j, k - 1, &e, &s);
            ex += e; sx += s; ey += e; sy += s; nx++; ny++;
        }
        if (j > 0 && k > 0) {
            sv_pool_epsilon_sigma(mp, i, j - 1, k - 1, &e, &s);
            ex += e; sx += s; nx++;
        }
        if (i > 0 && k > 0) {
            sv_pool_epsilon_sigma(mp, i - 1, j, k - 1, &e, &s);
            ey += e; sy += s; ny++;
        }
        if (i > 0 && j > 0) {
            sv_pool_epsilon_sigma(mp, i - 1, j - 1, k, &e, &s);
            ez += e; sz += s; nz++;
        }
        ex /= nx; sx /= nx;
        ey /= ny; sy /= ny;
        ez /= nz; sz /= nz;
    }

    *cax = (1 - sx * dt / 2 / ex) / (1 + sx * dt / 2 / ex);
    *cbx = (dt / ex / d->dx) / (1 + sx * dt / 2 / ex);
    if (d->bnds->data[i][j][k] > 0 || (d->dx != d->dy || d->dx != d->dz)) {
        *cay = (1 - sy * dt / 2 / ey) / (1 + sy * dt / 2 / ey);
        *cby = (dt / ey / d->dy) / (1 + sy * dt / 2 / ey);
        *caz = (1 - sz * dt / 2 / ez) / (1 + sz * dt / 2 / ez);
        *cbz = (dt / ez / d->dz) / (1 + sz * dt / 2 / ez);
    } else {
        *cay = *caz = *cax;
        *cby = *cbz = *cbx;
    }
}

/* Vector-Jacobian product of sv_pool_get_cabs with respect to the voxel
 * permittivities (sigma held fixed): given dJ/dca and dJ/dcb of cell
 * (i,j,k) per component, adds dJ/d(eps_v) to grad[v] for the voxels v of the
 * box (b_i0, b_j0, b_k0) + (bnx, bny, bnz), [i][j][k] order.  Voxels outside
 * the box are not parameters and are skipped. */
void
sv_pool_cabs_vjp(const SvPool *mp, int i, int j, int k, const double gca[3], const double gcb[3], int b_i0, int b_j0,
                 int b_k0, int bnx, int bny, int bnz, double *grad)
{
    const SvYeeData *d = mp->d;
    const double dt = d->dt;
    const double dn[3] = { d->dx, d->dy, d->dz };
    /* voxels averaged per component (Ex: self, j-1, k-1, j-1 k-1; ...), or self only */
    int vox[3][4][3], nv[3];
    double eavg[3], savg[3];
    int c, q;
    if (d->bnds->data[i][j][k] > 0) {
        static const int offs[3][4][3] = {
            { { 0, 0, 0 }, { 0, -1, 0 }, { 0, 0, -1 }, { 0, -1, -1 } },
            { { 0, 0, 0 }, { -1, 0, 0 }, { 0, 0, -1 }, { -1, 0, -1 } },
            { { 0, 0, 0 }, { -1, 0, 0 }, { 0, -1, 0 }, { -1, -1, 0 } },
        };
        for (c = 0; c < 3; c++) {
            nv[c] = 0;
            for (q = 0; q < 4; q++) {
                int vi = i + offs[c][q][0], vj = j + offs[c][q][1], vk = k + offs[c][q][2];
                if (vi < 0 || vj < 0 || vk < 0)
                    continue;
                vox[c][nv[c]][0] = vi; vox[c][nv[c]][1] = vj; vox[c][nv[c]][2] = vk;
                nv[c]++;
            }
        }
    } else
        for (c = 0; c < 3; c++) {
            nv[c] = 1;
            vox[c][0][0] = i; vox[c][0][1] = j; vox[c][0][2] = k;
        }
    for (c = 0; c < 3; c++) {
        double e, s, es = 0, ss = 0, half, dca, dcb, g;
        for (q = 0; q < nv[c]; q++) {
            sv_pool_epsilon_sigma(mp, vox[c][q][0], vox[c][q][1], vox[c][q][2], &e, &s);
            es += e;
            ss += s;
        }
        eavg[c] = es / nv[c];
        savg[c] = ss / nv[c];
        /* ca = (eps - h)/(eps + h), cb = dt / (d (eps + h)), h = sigma dt / 2 */
        half = savg[c] * dt / 2;
        dca = 2 * half / ((eavg[c] + half) * (eavg[c] + half));
        dcb = -dt / (dn[c] * (eavg[c] + half) * (eavg[c] + half));
        g = (gca[c] * dca + gcb[c] * dcb) / nv[c];
        for (q = 0; q < nv[c]; q++) {
            int vi = vox[c][q][0] - b_i0, vj = vox[c][q][1] - b_j0, vk = vox[c][q][2] - b_k0;
            if (vi < 0 || vj < 0 || vk < 0 || vi >= bnx || vj >= bny || vk >= bnz)
                continue;
            grad[((size_t)vi * bny + vj) * bnz + vk] += g;
        }
    }
}

static int
disp_at(const SvPool *mp, int i, int j, int k)
{
    int m = mp->d->mat->data[i][j][k];
    return m > 0 && sv_material_is_dispersive(mp->mats[m].type);
}

int
sv_pool_get_tbc(const SvPool *mp, int i, int j, int k, int *xtbc, int *ytbc, int *ztbc)
{
    const SvYeeData *d = mp->d;
    int matindex;

    *xtbc = *ytbc = *ztbc = 1;
    if (mp->nmat <= 0)
        return -1;
    matindex = d->mat->data[i][j][k];
    if (matindex > 0 && sv_material_is_dispersive(mp->mats[matindex].type)) {
        *xtbc = *ytbc = *ztbc = 0;
        return matindex;
    }
    /* dispersive neighbour on the minus side: the last match wins, src3d/yee.c:474-522 */
    if (i > 0 && disp_at(mp, i - 1, j, k)) { matindex = d->mat->data[i - 1][j][k]; *ytbc = *ztbc = 0; }
    if (j > 0 && disp_at(mp, i, j - 1, k)) { matindex = d->mat->data[i][j - 1][k]; *xtbc = *ztbc = 0; }
    if (k > 0 && disp_at(mp, i, j, k - 1)) { matindex = d->mat->data[i][j][k - 1]; *xtbc = *ytbc = 0; }
    if (i > 0 && j > 0 && disp_at(mp, i - 1, j - 1, k)) { matindex = d->mat->data[i - 1][j - 1][k]; *ztbc = 0; }
    if (j > 0 && k > 0 && disp_at(mp, i, j - 1, k - 1)) { matindex = d->mat->data[i][j - 1][k - 1]; *xtbc = 0; }
    if (i > 0 && k > 0 && disp_at(mp, i - 1, j, k - 1)) { matindex = d->mat->data[i - 1][j][k - 1]; *ytbc = 0; }
    return matindex;
}

int
sv_pool_is_pec(const SvPool *mp, int i, int j, int k)
{
    int m;
    if (mp->nmat <= 0)
        return 0;
    m = mp->d->mat->data[i][j][k];
    return m > 0 && mp->mats[m].type == SV_MAT_PEC;
}

/* ------------------------------------------------------------------ */
/* TSF geometry (src3d/source.c:53-165)                                */
/* ------------------------------------------------------------------ */

double
sv_dcomp(int i, int j, int k, double theta, double phi, int i0, int i1, int j0, int j1, int k0, int k1)
{
    double rx = 0, ry = 0, rz = 0;
    double ax = sin(theta) * cos(phi);
    double ay = sin(theta) * sin(phi);
    double az = cos(theta);

    if (theta >= 0 && theta <= (M_PI / 2.0)) {
        if (phi >= 0 && phi <= (M_PI / 2.0)) { rx = i - i0; ry = j - j0; rz = k - k0; }
        else if (phi > (M_PI / 2.0) && phi <= M_PI) { rx = i - i1; ry = j - j0; rz = k - k0; }
        else if (phi > M_PI && phi <= (3.0 * M_PI / 2.0)) { rx = i - i1; ry = j - j1; rz = k - k0; }
        else { rx = i - i0; ry = j - j1; rz = k - k0; }
    } else if (theta < M_PI && theta > (M_PI / 2.0)) {
        /* the reference leaves rx = 0 in the second quadrant (src3d/source.c:89-91) */
        if (phi >= 0 && phi <= (M_PI / 2.0)) { rx = i - i0; ry = j - j0; rz = k - k1; }
        else if (phi > (M_PI / 2.0) && phi <= M_PI) { ry = j - j0; rz = k - k1; }
        else if (phi > M_PI && phi <= (3.0 * M_PI / 2.0)) { rx = i - i1; ry = j - j1; rz = k - k1; }
        else { rx = i - i0; ry = j - j1; rz = k - k1; }
    } else {
        return -1;
    }
    return 10 + (ax * rx + ay * ry + az * rz);
}

double sv_gex(double f, double t, double p, double s) { return f * (cos(s) * sin(p) - sin(s) * cos(t) * cos(p)); }
double sv_gey(double f, double t, double p, double s) { return f * (-cos(s) * cos(p) - sin(s) * cos(t) * sin(p)); }
double sv_gez(double f, double t, double p, double s) { return f * (sin(s) * sin(t)); }
double sv_ghx(double f, double t, double p, double s) { return f * (sin(s) * sin(p) + cos(s) * cos(t) * cos(p)); }
double sv_ghy(double f, double t, double p, double s) { return f * (-sin(s) * cos(p) + cos(s) * cos(t) * sin(p)); }
double sv_ghz(double f, double t, double p, double s) { return f * (-cos(s) * sin(t)); }

int
sv_tsf_face_active(const SvPool *mp, const SvSet *set, int i, int j, int k)
{
    /* src3d/source.c:253-264: one map for the TSF and the layered TSF */
    if (set->ss.tsf.box_boundary_skipdepth == -1 && set->ss.ltsf.box_boundary_skipdepth == -1)
        return 1;
    if (mp->d->sourceskip && mp->d->sourceskip->data[i][j][k])
        return 0;
    return 1;
}

/* ------------------------------------------------------------------ */
/* far-field geometry (src3d/farfield.c:32-54)                         */
/* ------------------------------------------------------------------ */

double
sv_ff_dist(int x, int y, int z, int i, int j, int k)
{
    return sqrt((double)(x - i) * (x - i) + (double)(y - j) * (y - j) + (double)(z - k) * (z - k));
}

static double
angvec(double x1, double y1, double z1, double x2, double y2, double z2)
{
    return (x1 * x2 + y1 * y2 + z1 * z2) / (sqrt((x1 * x1 + y1 * y1 + z1 * z1) * (x2 * x2 + y2 * y2 + z2 * z2)));
}

double
sv_ff_angle(int x1, int y1, int z1, int x2, int y2, int z2, int side)
{
    double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    switch (side) {
        case 0: return angvec(dx, dy, dz, 0, -1, 0);
        case 1: return angvec(dx, dy, dz, 0, 1, 0);
        case 2: return angvec(dx, dy, dz, -1, 0, 0);
        case 3: return angvec(dx, dy, dz, 1, 0, 0);
        case 4: return angvec(dx, dy, dz, 0, 0, -1);
        default: return angvec(dx, dy, dz, 0, 0, 1);
    }
}

/* ------------------------------------------------------------------ */
/* sources (src3d/plan.c:1430-1597)                                    */
/* ------------------------------------------------------------------ */

static int
read_point_source_file(const char *fname, SvSourcePoint *sp, int nsteps)
{
    FILE *fr = fopen(fname, "r");
    int n, q;
    if (!fr) {
        fprintf(stderr, "Error: cannot open file %s for reading\n", fname);
        return 1;
    }
    if (fscanf(fr, "%d", &n) != 1 || n <= 0) {
        fprintf(stderr, "Error: cannot read number of values from file %s\n", fname);
        fclose(fr);
        return 1;
    }
    if (n < nsteps)
        fprintf(stderr, "Warning: less data (%d) in \"%s\" than scheduled computation steps (%d)\n", n, fname, nsteps);
    sp->ndata = n;
    sp->layered_zpos = (int *)malloc(n * sizeof(int));
    sp->ex = (double *)malloc(6 * (size_t)n * sizeof(double));
    sp->ey = sp->ex + n;
    sp->ez = sp->ey + n;
    sp->hx = sp->ez + n;
    sp->hy = sp->hx + n;
    sp->hz = sp->hy + n;
    for (q = 0; q < n; q++) {
        if (fscanf(fr, "%d %lf %lf %lf %lf %lf %lf", &sp->layered_zpos[q], &sp->ex[q], &sp->ey[q], &sp->ez[q],
                   &sp->hx[q], &sp->hy[q], &sp->hz[q]) != 7) {
            fprintf(stderr, "Error: cannot read point source sample %d from %s\n", q, fname);
            fclose(fr);
            return 1;
        }
    }
    fclose(fr);
    return 0;
}

/* the waveform of a plane-wave source (TSF, SF): a file of "count" then
 * "index value" lines, or a generated waveform rounded through "%g" as the
 * reference writes and re-reads it (src3d/settings.c:444-548, src3d/plan.c:1430-1520) */
static int
load_plane_waveform(const SvSet *set, const SvSetSourceTSF *t, const char *what, SvSourceTSF *ts)
{
    int q;
    if (t->source_mode == 0) {
        FILE *fr = fopen(t->source_filename, "r");
        int n;
        if (!fr) {
            fprintf(stderr, "Cannot load source file %s\n", t->source_filename);
            return 1;
        }
        if (fscanf(fr, "%d", &n) != 1 || n <= 0) {
            fprintf(stderr, "Error: cannot read number of values from file %s\n", t->source_filename);
            fclose(fr);
            return 1;
        }
        ts->ndata = n;
        ts->e = (double *)calloc(n > set->sc.nsteps ? n : set->sc.nsteps, sizeof(double));
        for (q = 0; q < n; q++) {
            int pos;
            if (fscanf(fr, "%d %lf", &pos, &ts->e[q]) != 2) {
                fprintf(stderr, "Error: cannot read %s sample %d from %s\n", what, q, t->source_filename);
                fclose(fr);
                return 1;
            }
        }
        fclose(fr);
        if (n < set->sc.nsteps)
            fprintf(stderr, "Warning: less data (%d) in \"%s\" than scheduled computation steps (%d) in %s source; "
                    "missing samples are zero\n", n, t->source_filename, set->sc.nsteps, what);
    } else {
        int n = set->sc.nsteps;
        ts->ndata = n;
        ts->e = (double *)calloc(n, sizeof(double));
        if (sv_generate_waveform(t->source_mode, n, t->source_wl, t->source_pulsewidth, set->plan.dt,
                                 t->source_amplitude, ts->e))
            return 1;
        for (q = 0; q < n; q++)
            ts->e[q] = sv_round_g6(ts->e[q]);
    }
    return 0;
}

/* Gauss-Legendre nodes (ascending) and weights on [-1, 1]: Newton's method on
 * P_n from the three-term recurrence, started at cos(pi (i + 3/4) / (n + 1/2))
 * (Abramowitz & Stegun 25.4.29; the reference's get_gl_coeffs,
 * src3d/plan.c:1402-1427, is the same textbook method) */
static void
gauss_legendre(int n, double *x, double *w)
{
    int i, j, it;
    for (i = 0; i < (n + 1) / 2; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = 0.0, zn;
            for (j = 1; j <= n; j++) {          /* p0 = P_j(z), p1 = P_{j-1}(z) */
                const double pm = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            zn = z - p0 / dp;
            if (fabs(zn - z) <= 1e-15) {
                z = zn;
                break;
            }
            z = zn;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* focused TSF, src3d/plan.c:1855-1979: Gauss-Legendre nodes in theta over
 * [0, thetamax], midpoints in phi; psi = phi + pol */
static int
load_tsff(SvPool *mp, SvSet *set)
{
    const SvSetSourceTSFF *f = &set->ss.tsff;
    const SvSetSourceTSF *t = &f->w;
    const int nip = f->focused_nip, mip = f->focused_mip;
    SvSourceTSFF *ff;
    SvSourceTSF base;
    double *xs, *ws;
    int a, b;
    if (nip < 1 || mip < 1 || nip * mip > 4096) {
        fprintf(stderr, "Error: TSFF needs 1 .. 4096 plane waves (nip x mip = %d x %d)\n", nip, mip);
        return 1;
    }
    if (t->box_i0 < 1 || t->box_j0 < 1 || t->box_k0 < 1 || t->box_in >= set->sp.xres || t->box_jn >= set->sp.yres
        || t->box_kn >= set->sp.zres || t->box_in <= t->box_i0 || t->box_jn <= t->box_j0 || t->box_kn <= t->box_k0) {
        fprintf(stderr, "Error: TSFF box must lie inside the pool with one cell to spare\n");
        return 1;
    }
    if (!(f->focused_thetamax > 0 && f->focused_thetamax <= M_PI / 2)) {
        /* the amplitude takes sqrt(cos theta) */
        fprintf(stderr, "Error: TSFF aperture angle must lie in (0, pi/2]\n");
        return 1;
    }
    if (set->sb.bx0 == SV_BOUNDARY_PERIODIC || set->sb.by0 == SV_BOUNDARY_PERIODIC) {
        fprintf(stderr, "Error: TSFF in a periodic pool is not supported\n");
        return 1;
    }
    memset(&base, 0, sizeof(base));
    base.box_i0 = t->box_i0; base.box_j0 = t->box_j0; base.box_k0 = t->box_k0;
    base.box_in = t->box_in; base.box_jn = t->box_jn; base.box_kn = t->box_kn;
    base.skip_i0 = t->box_boundary_skipi0; base.skip_j0 = t->box_boundary_skipj0; base.skip_k0 = t->box_boundary_skipk0;
    base.skip_in = t->box_boundary_skipin; base.skip_jn = t->box_boundary_skipjn; base.skip_kn = t->box_boundary_skipkn;
    if (load_plane_waveform(set, t, "TSFF", &base))
        return 1;
    base.jp_res = (int)(sqrt((double)(set->sp.xres * set->sp.xres + set->sp.yres * set->sp.yres
                                      + set->sp.zres * set->sp.zres)) + 250);
    base.jp_epsilon = base.jp_mu = 1.0;
    ff = (SvSourceTSFF *)calloc(1, sizeof(SvSourceTSFF));
    mp->src->tsff = ff;
    ff->nw = nip * mip;
    ff->w = (SvSourceTSF *)calloc(ff->nw, sizeof(SvSourceTSF));
    ff->acorr = (double *)calloc(ff->nw, sizeof(double));
    ff->xshift = f->xshift; ff->yshift = f->yshift; ff->zshift = f->zshift;
    xs = (double *)malloc(nip * sizeof(double));
    ws = (double *)malloc(nip * sizeof(double));
    gauss_legendre(nip, xs, ws);
    for (a = 0; a < nip; a++)
        for (b = 0; b < mip; b++) {
            SvSourceTSF *w = ff->w + a * mip + b;
            const double an = ws[a] * f->focused_thetamax / 2, bm = 2.0 * M_PI / mip;
            *w = base;
            w->ia_theta = (xs[a] + 1) * f->focused_thetamax / 2.0;
            w->ia_phi = (b + 0.5) * 2 * M_PI / mip;
            w->ia_psi = w->ia_phi + f->focused_pol;
            w->corr = 1.0 / sqrt(pow(sin(w->ia_theta), 4) * (pow(cos(w->ia_phi), 4) + pow(sin(w->ia_phi), 4))
                                 + pow(cos(w->ia_theta), 4));
            w->jp_dx = set->sp.dx / w->corr;
            /* fs_amplitude, src3d/source.c:167-171, with freal = fdist dx */
            ff->acorr[a * mip + b] = 1e15 * an * bm * (f->focused_fdist * set->sp.dx / (2 * M_PI * SV_LIGHT_SPEED))
                                     * sqrt(cos(w->ia_theta)) * sin(w->ia_theta);
        }
    free(xs);
    free(ws);
    return 0;
}

int
sv_ext_read_planes(const SvSet *set, int family, int step, float *out)
{
    /* src3d/source.c:1340-1399 (H planes for the E step), 2206-2266 (E planes) */
    const int n = set->ss.ext.extxres * set->ss.ext.extyres;
    char name[SV_NAME_LEN + 16];
    int c, q;
    for (c = 0; c < 3; c++) {
        FILE *f;
        snprintf(name, sizeof(name), "%s_%.4d", set->ss.ext.filebase[3 * family + c], step - set->ss.ext.shift);
        f = fopen(name, "r");
        if (!f) {
            fprintf(stderr, "Error: cannot open external source files (e.g. %s, shifted step %d), skipping.\n", name,
                    step - set->ss.ext.shift);
            return 1;
        }
        for (q = 0; q < n; q++)
            if (fscanf(f, "%f", out + (size_t)c * n + q) != 1) {
                fprintf(stderr, "Error: external source file %s holds fewer than %d values, skipping.\n", name, n);
                fclose(f);
                return 1;
            }
        fclose(f);
    }
    return 0;
}

/* SOURCE_EXT geometry: the reference's bounds tests on the target cells can
 * never fail (src3d/source.c:1382-1385), so a region outside the pool is
 * refused here */
static int
check_ext(const SvSet *set)
{
    const int ni = set->ss.ext.iextto - set->ss.ext.iextfrom, nj = set->ss.ext.jextto - set->ss.ext.jextfrom;
    if (set->ss.ext.i != -1 || set->ss.ext.j != -1) {
        fprintf(stderr, "Error: SOURCE_EXT supports the z-plane mode only (i = j = -1); the reference applies no other\n");
        return 1;
    }
    if (set->ss.ext.extxres < 1 || set->ss.ext.extyres < 1 || ni < 1 || nj < 1 || set->ss.ext.iextfrom < 0
        || set->ss.ext.jextfrom < 0 || set->ss.ext.iextto > set->ss.ext.extxres || set->ss.ext.jextto > set->ss.ext.extyres
        || set->ss.ext.ijstart < 0 || set->ss.ext.jkstart < 0 || set->ss.ext.ijstart + ni > set->sp.xres
        || set->ss.ext.jkstart + nj > set->sp.yres || set->ss.ext.k < 1 || set->ss.ext.k >= set->sp.zres) {
        fprintf(stderr, "Error: SOURCE_EXT region must lie inside its planes and inside the pool (1 <= k < zres)\n");
        return 1;
    }
    return 0;
}

static int
load_sources(SvPool *mp, SvSet *set)
{
    int i, q;
    SvSource *src = mp->src;

    src->npnts = set->ss.npnts;
    if (set->ss.npnts)
        src->sp = (SvSourcePoint *)calloc(set->ss.npnts, sizeof(SvSourcePoint));
    for (i = 0; i < set->ss.npnts; i++) {
        SvSetSourcePoint *p = set->ss.pnts + i;
        SvSourcePoint *sp = src->sp + i;
        sp->i = p->point_origin_position_i;
        sp->j = p->point_origin_position_j;
        sp->k = p->point_origin_position_k;
        if (sp->i < 0 || sp->j < 0 || sp->k < 0 || sp->i >= set->sp.xres || sp->j >= set->sp.yres || sp->k >= set->sp.zres) {
            fprintf(stderr, "Error: point source %d at (%d %d %d) is outside the pool\n", i, sp->i, sp->j, sp->k);
            return 1;
        }
        if (p->source_mode == 0) {
            if (read_point_source_file(p->source_filename, sp, set->sc.nsteps))
                return 1;
        } else {
            /* generated source written as SRCF_ARB, src3d/settings.c:525-531 */
            int n = set->sc.nsteps;
            double *v = (double *)malloc(n * sizeof(double));
            double ax = sin(p->point_origin_theta) * cos(p->point_origin_phi);
            double ay = sin(p->point_origin_theta) * sin(p->point_origin_phi);
            double az = cos(p->point_origin_theta);
            if (sv_generate_waveform(p->source_mode, n, p->source_wl, p->source_pulsewidth, set->plan.dt,
                                     p->source_amplitude, v)) {
                free(v);
                return 1;
            }
            sp->ndata = n;
            sp->layered_zpos = (int *)malloc(n * sizeof(int));
            sp->ex = (double *)calloc(6 * (size_t)n, sizeof(double));
            sp->ey = sp->ex + n;
            sp->ez = sp->ey + n;
            sp->hx = sp->ez + n;
            sp->hy = sp->hx + n;
            sp->hz = sp->hy + n;
            for (q = 0; q < n; q++) {
                sp->layered_zpos[q] = q;
                sp->ex[q] = sv_round_g6(ax * v[q]);
                sp->ey[q] = sv_round_g6(ay * v[q]);
                sp->ez[q] = sv_round_g6(az * v[q]);
            }
            free(v);
        }
    }

    if (set->ss.tsf.valid) {
        SvSetSourceTSF *t = &set->ss.tsf;
        SvSourceTSF *ts = (SvSourceTSF *)calloc(1, sizeof(SvSourceTSF));
        src->tsf = ts;
        ts->box_i0 = t->box_i0; ts->box_j0 = t->box_j0; ts->box_k0 = t->box_k0;
        ts->box_in = t->box_in; ts->box_jn = t->box_jn; ts->box_kn = t->box_kn;
        ts->skip_i0 = t->box_boundary_skipi0; ts->skip_j0 = t->box_boundary_skipj0; ts->skip_k0 = t->box_boundary_skipk0;
        ts->skip_in = t->box_boundary_skipin; ts->skip_jn = t->box_boundary_skipjn; ts->skip_kn = t->box_boundary_skipkn;
        ts->ia_theta = t->ia_theta;
        ts->ia_phi = t->ia_phi;
        ts->ia_psi = t->ia_psi;
        /* a domain-periodic axis (no reference counterpart): the box spans the
         * whole axis, its faces normal to it are off, and the wave must not
         * travel along it (the incident field has to be periodic) */
        if (set->sb.bx0 == SV_BOUNDARY_PERIODIC || set->sb.by0 == SV_BOUNDARY_PERIODIC) {
            double ax = sin(t->ia_theta) * cos(t->ia_phi), ay = sin(t->ia_theta) * sin(t->ia_phi);
            if ((set->sb.bx0 == SV_BOUNDARY_PERIODIC && fabs(ax) > 1e-12)
                || (set->sb.by0 == SV_BOUNDARY_PERIODIC && fabs(ay) > 1e-12)) {
                fprintf(stderr, "Error: a TSF wave in a periodic pool must travel normal to the periodic axes\n");
                return 1;
            }
            if (set->sb.bx0 == SV_BOUNDARY_PERIODIC) {
                ts->box_i0 = 0; ts->box_in = set->sp.xres; ts->skip_i0 = ts->skip_in = 1;
            }
            if (set->sb.by0 == SV_BOUNDARY_PERIODIC) {
                ts->box_j0 = 0; ts->box_jn = set->sp.yres; ts->skip_j0 = ts->skip_jn = 1;
            }
        }
        if ((set->sb.bx0 != SV_BOUNDARY_PERIODIC && (ts->box_i0 < 1 || ts->box_in >= set->sp.xres))
            || (set->sb.by0 != SV_BOUNDARY_PERIODIC && (ts->box_j0 < 1 || ts->box_jn >= set->sp.yres))
            || ts->box_k0 < 1 || ts->box_kn >= set->sp.zres) {
            fprintf(stderr, "Error: TSF box must lie inside the pool with one cell to spare\n");
            return 1;
        }
        if (load_plane_waveform(set, t, "TSF", ts))
            return 1;
        /* src3d/plan.c:1584-1595 */
        ts->corr = 1.0 / sqrt(pow(sin(ts->ia_theta), 4) * (pow(cos(ts->ia_phi), 4) + pow(sin(ts->ia_phi), 4))
                              + pow(cos(ts->ia_theta), 4));
        ts->jp_res = (int)(sqrt((double)(set->sp.xres * set->sp.xres + set->sp.yres * set->sp.yres
                                         + set->sp.zres * set->sp.zres)) + 250);
        ts->jp_dx = set->sp.dx / ts->corr;
        ts->jp_epsilon = t->layered_epsilon;
        ts->jp_mu = t->layered_mu;
        /* The line's update (Mur condition on H at the source end, then the Yee
         * update of that H, src3d/1dpool.c:83-162 in the order of
         * src3d/source.c:1469-1473) grows without bound when its Courant number
         * exceeds 0.75 (e.g. theta = 45 deg at the default time step: 0.816).
         * The results stay the reference's; the warning names the remedy. */
        {
            const double s = SV_LIGHT_SPEED * set->plan.dt / ts->jp_dx;
            if (s > 0.75)
                fprintf(stderr, "Warning: the TSF auxiliary line runs at Courant number %.4f; its update diverges "
                                "above 0.75 (DT_MULT <= %.4g keeps it stable)\n", s, set->sc.dtmult * 0.75 / s);
        }
    }
    if (set->ss.tsff.w.valid && load_tsff(mp, set))
        return 1;
    if (set->ss.ext.valid && check_ext(set))
        return 1;
    if (set->ss.sf.valid) {
        /* src3d/plan.c:1985-2040: the waveform and a 1-D line of
         * (int) sqrt(xres^2 + yres^2 + zres^2 + 10) cells of dx, hard source at
         * cell 0, material of the layered keywords; the line is read without
         * the TSF's angular correction */
        SvSourceTSF *sf = (SvSourceTSF *)calloc(1, sizeof(SvSourceTSF));
        src->sf = sf;
        sf->ia_theta = set->ss.sf.ia_theta;
        sf->ia_phi = set->ss.sf.ia_phi;
        sf->ia_psi = set->ss.sf.ia_psi;
        if (load_plane_waveform(set, &set->ss.sf, "SF", sf))
            return 1;
        sf->corr = 1.0;
        sf->jp_res = (int)sqrt((double)(set->sp.xres * set->sp.xres + set->sp.yres * set->sp.yres
                                        + set->sp.zres * set->sp.zres) + 10);
        sf->jp_dx = set->sp.dx;
        sf->jp_epsilon = set->ss.sf.layered_epsilon;
        sf->jp_mu = set->ss.sf.layered_mu;
    }
    if (set->ss.ltsf.valid) {
        /* src3d/plan.c:1600-1685 */
        const SvSetSourceLTSF *t = &set->ss.ltsf;
        SvSourceLTSF *ls = (SvSourceLTSF *)calloc(1, sizeof(SvSourceLTSF));
        const int zres = set->sp.zres;
        FILE *fr;
        int n, q, k;
        mp->src->ltsf = ls;
        ls->box_i0 = t->box_i0; ls->box_j0 = t->box_j0; ls->box_k0 = t->box_k0;
        ls->box_in = t->box_in; ls->box_jn = t->box_jn; ls->box_kn = t->box_kn;
        ls->ia_theta = t->ia_theta; ls->ia_phi = t->ia_phi; ls->ia_psi = t->ia_psi;
        if (ls->box_i0 < 1 || ls->box_j0 < 1 || ls->box_k0 < 1 || ls->box_in >= set->sp.xres || ls->box_jn >= set->sp.yres
            || ls->box_kn >= zres || ls->box_in <= ls->box_i0 || ls->box_jn <= ls->box_j0 || ls->box_kn <= ls->box_k0) {
            fprintf(stderr, "Error: LTSF box must lie inside the pool with one cell to spare\n");
            return 1;
        }
        if (zres < 12) {
            fprintf(stderr, "Error: LTSF needs at least 12 cells along z (its 1-D model injects at cell 10)\n");
            return 1;
        }
        if (!(t->ia_theta >= 0 && t->ia_theta <= M_PI / 2)) {
            fprintf(stderr, "Error: unsupported LTSF angle (theta must lie in 0 .. pi/2)\n");
            return 1;
        }
        if (t->source_mode == 1 || t->source_mode == 2) {
            /* generated at dt / 3 and read back through "%g" as the reference's
             * tmp_source_ltsf file (src3d/settings.c:4420-4428, 445-512) */
            n = 3 * set->sc.nsteps;
            ls->ndata = n;
            ls->e = (double *)calloc(n, sizeof(double));
            if (sv_generate_waveform_ltsf(t->source_mode, set->sc.nsteps, t->source_wl, t->source_pulsewidth,
                                          set->plan.dt, t->source_amplitude, ls->e))
                return 1;
            for (q = 0; q < n; q++)
                ls->e[q] = sv_round_g6(ls->e[q]);
        } else {
            fr = fopen(t->source_filename, "r");
            if (!fr) {
                fprintf(stderr, "Cannot load source file %s\n", t->source_filename);
                return 1;
            }
            if (fscanf(fr, "%d", &n) != 1 || n <= 0) {
                fprintf(stderr, "Error: cannot read number of values from file %s\n", t->source_filename);
                fclose(fr);
                return 1;
            }
            if (n < 3 * set->sc.nsteps)
                fprintf(stderr, "Warning: less data (%d) in \"%s\" than scheduled computation steps (%d) in LTSF source "
                                "(need to be three times more)\n", n, t->source_filename, set->sc.nsteps);
            ls->ndata = n;
            ls->e = (double *)calloc(n, sizeof(double));
            for (q = 0; q < n; q++) {
                int pos;
                if (fscanf(fr, "%d %lf", &pos, &ls->e[q]) != 2) {
                    fprintf(stderr, "Error: cannot read LTSF sample %d from %s\n", q, t->source_filename);
                    fclose(fr);
                    return 1;
                }
            }
            fclose(fr);
        }
        ls->corr = 1.0 / sqrt(pow(sin(ls->ia_theta), 4) * (pow(cos(ls->ia_phi), 4) + pow(sin(ls->ia_phi), 4))
                              + pow(cos(ls->ia_theta), 4));
        ls->zdx = set->sp.dx / ls->corr;
        ls->zdt = set->plan.dt / 3;
        /* a face cell's delay is 3 (step - r dx sin(theta) / (c dt)) substeps with
         * r >= -(|cos phi| + |sin phi|) >= -sqrt(2) */
        ls->nrows = 3 * set->sc.nsteps + (int)ceil(3 * 1.5 * set->sp.dx / (SV_LIGHT_SPEED * set->plan.dt)) + 3;
        /* src3d/zpool.c:118-131: each layer from its z position to the end */
        ls->zeps = (double *)malloc(zres * sizeof(double));
        ls->zmu = (double *)malloc(zres * sizeof(double));
        for (k = 0; k < zres; k++)
            ls->zeps[k] = ls->zmu[k] = 1;
        for (q = 0; q < t->layered_count; q++) {
            double eps = t->layered_epsilon[q], mu = t->layered_mu[q];
            if (t->layered_zpos[q] < 0) {
                fprintf(stderr, "Error: LTSF layer %d starts below z = 0\n", q);
                return 1;
            }
            if (t->layered_material[q][0]) {
                /* a database layer: the reference reads its name and leaves the
                 * layer's eps at 0 (src3d/settings.c:4372-4376); here it is
                 * resolved at the centre wavelength like a vector-file entry */
                double sg, sa;
                if (sv_database_linear(set, t->layered_material[q], &eps, &mu, &sg, &sa))
                    return 1;
            }
            for (k = t->layered_zpos[q]; k < zres; k++) {
                ls->zeps[k] = eps;
                ls->zmu[k] = mu;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* materials (src3d/plan.c:2243-2543, 3112-3300)                       */
/* ------------------------------------------------------------------ */

static void
calculate_ade(SvMatProp *mat, double epsilon, double sigma, double omegap, double nu, double dt)
{
    int i;
    /* src3d/plan.c:2736-2762 */
    mat->ade_a0 = (nu * dt - 2) / (nu * dt + 2);
    mat->ade_a1 = 4 / (nu * dt + 2);
    mat->ade_a2 = dt * dt * EPS0 * omegap * omegap / (2 * (nu * dt + 2));
    /* a third pole (type 7, no reference counterpart) follows the same
     * formulas; for type 5 its parameters are zero and its terms vanish */
    for (i = 0; i < 3; i++) {
        double a = mat->cp3_a[i], phi = mat->cp3_phi[i], omega = mat->cp3_omega[i], gamma = mat->cp3_gamma[i];
        double den = (omega * omega * dt * dt + (2 + gamma * dt) * (2 + gamma * dt));
        mat->ade_bp0[i] = -(omega * omega * dt * dt + (2 - gamma * dt) * (2 - gamma * dt)) / den;
        mat->ade_bp1[i] = 2 * (4 - (gamma * gamma + omega * omega) * dt * dt) / den;
        mat->ade_bp2[i] = 2 * EPS0 * a * omega * dt * (omega * dt * cos(phi) + (2 - gamma * dt) * sin(phi)) / den;
        mat->ade_bp3[i] = 4 * EPS0 * a * omega * dt * (omega * dt * cos(phi) - gamma * dt * sin(phi)) / den;
        mat->ade_bp4[i] = 2 * EPS0 * a * omega * dt * (omega * dt * cos(phi) - (2 + gamma * dt) * sin(phi)) / den;
    }
    {
        double den = sigma * dt / 2 + mat->ade_a2 + mat->ade_bp4[0] + mat->ade_bp4[1] + mat->ade_bp4[2] + EPS0 * epsilon;
        mat->ade_c0 = dt / den;
        mat->ade_c1 = 1 / den;
        mat->ade_c2 = -(mat->ade_a2 + mat->ade_bp2[0] + mat->ade_bp2[1] + mat->ade_bp2[2]) / den;
        mat->ade_c3 = -(sigma * dt / 2 + 2 * mat->ade_a2 + mat->ade_bp3[0] + mat->ade_bp3[1] + mat->ade_bp3[2]
                        - EPS0 * epsilon) / den;
    }
}

/* Susceptibility integrals of the piecewise-linear recursive convolution
 * (type 6), src3d/plan.c:2765-2824: with chi(t) the material's
 * susceptibility, chi0 = int_0^dt chi, xi0 = (1/dt) int_0^dt t chi, and the
 * increments dchi = chi0 - chi1, dxi = xi0 - xi1 of the next interval; the
 * Drude term chi(t) = (wp^2/nu)(1 - e^(-nu t)) is real, a critical point
 * chi(t) = 2 A W e^(-G t) sin(W t - phi) is carried as a complex exponential
 * (conjugate convention e^((-G - iW) t)) whose real part is the physical one. */
static void
calculate_plrc(SvMatProp *mat, double dt)
{
    const double w2 = mat->drude_omega_p * mat->drude_omega_p / (mat->drude_nu * mat->drude_nu);
    const double nd = mat->drude_nu * dt;
    const double lin = 1 / (1 - exp(nd)) + 1 / nd;
    int q;
    mat->plrc_d_chi = w2 * (exp(-nd) + nd - 1);
    mat->plrc_d_xi = mat->plrc_d_chi * lin - w2 * (1 - nd / 2 * (1 / tanh(nd / 2)));
    mat->plrc_d_dchi = -w2 * (1 - exp(-nd)) * (1 - exp(-nd));
    mat->plrc_d_dxi = mat->plrc_d_dchi * lin;
    for (q = 0; q < 2; q++) {
        const double a = mat->cp3_a[q], phi = mat->cp3_phi[q], om = mat->cp3_omega[q], ga = mat->cp3_gamma[q];
        /* amplitude -2 A W (sin phi - i cos phi), one-interval factor 1 - e^((-G - iW) dt) */
        const double complex amp = -2 * a * om * sin(phi) + I * (2 * a * om * cos(phi));
        const double complex f1 = (1 - exp(-dt * ga) * cos(dt * om)) + I * (exp(-dt * ga) * sin(dt * om));
        /* the linear-interpolation terms divide by 1 - e^((G + iW) dt) and by (G + iW) dt */
        const double complex fe = (1 - exp(dt * ga) * cos(dt * om)) - I * (exp(dt * ga) * sin(dt * om));
        const double complex fl = dt * ga + I * (dt * om);
        const double complex chi0 = amp * f1 / (ga + I * om);
        const double complex dchi = chi0 * f1;
        const double complex dxi = dchi / fe + dchi / fl;
        mat->plrc_p_chi[q] = creal(chi0);
        mat->plrc_p_xi[q] = creal(chi0 / fe + chi0 / fl);
        mat->plrc_p_dchir[q] = creal(dchi);
        mat->plrc_p_dchii[q] = cimag(dchi);
        mat->plrc_p_dxir[q] = creal(dxi);
        mat->plrc_p_dxii[q] = cimag(dxi);
    }
}

static int
scan_floats(FILE *fr, float *v, int n)
{
    int q;
    for (q = 0; q < n; q++)
        if (fscanf(fr, "%f", v + q) != 1)
            return 1;
    return 0;
}

/* src3d/plan.c:3112-3300; returns 1 on a read error */
static int
scan_material(FILE *fr, SvMatProp *mat, double dt)
{
    float v[16];
    int type, q, np;

    memset(mat, 0, sizeof(*mat));
    mat->epsilon = mat->mu = 1;
    mat->pos = -1;
    if (fscanf(fr, "%d", &type) != 1)
        return 1;
    mat->type = type;
    if (type == SV_MAT_LINEAR || type == SV_MAT_LINTAB) {
        if (scan_floats(fr, v, 4))
            return 1;
        mat->epsilon = v[0];
        mat->mu = v[1];
        mat->sigma = v[2];
        mat->sigast = v[3];
    } else if (type == SV_MAT_PEC) {
    } else if (type == SV_MAT_DRUDE) {
        /* epsilon, omega_p, nu (omega_p and nu in the units the reference
         * scales by EV_TO_J after loading, src3d/plan.c:4673-4677) */
        if (scan_floats(fr, v, 3))
            return 1;
        mat->epsilon = v[0];
        mat->drude_omega_p = v[1];
        mat->drude_nu = v[2];
    } else if (type == SV_MAT_CP || type == SV_MAT_PLRC) {
        /* CP: epsilon, omega_p, nu, 2 x (A, phi, Omega, Gamma); PLRC: epsilon,
         * sigma, omega_p, nu, 2 x (A, phi, Omega, Gamma) */
        const int off = type == SV_MAT_PLRC;
        if (scan_floats(fr, v, 11 + off))
            return 1;
        mat->epsilon = v[0];
        mat->sigma = off ? v[1] : 0;
        mat->drude_omega_p = v[1 + off];
        mat->drude_nu = v[2 + off];
        for (q = 0; q < 2; q++) {
            mat->cp3_a[q] = v[3 + off + 4 * q];
            mat->cp3_phi[q] = v[4 + off + 4 * q];
            mat->cp3_omega[q] = v[5 + off + 4 * q];
            mat->cp3_gamma[q] = v[6 + off + 4 * q];
        }
        if (type == SV_MAT_PLRC)
            calculate_plrc(mat, dt);
    } else if (type == SV_MAT_ADE || type == SV_MAT_ADE3) {
        np = type == SV_MAT_ADE3 ? 3 : 2;
        if (scan_floats(fr, v, 4 + 4 * np))
            return 1;
        mat->type = SV_MAT_ADE;
        mat->ade_npoles = np;
        mat->epsilon = v[0];
        mat->sigma = v[1];
        mat->drude_omega_p = v[2];
        mat->drude_nu = v[3];
        for (q = 0; q < np; q++) {
            mat->cp3_a[q] = v[4 + 4 * q];
            mat->cp3_phi[q] = v[5 + 4 * q];
            mat->cp3_omega[q] = v[6 + 4 * q];
            mat->cp3_gamma[q] = v[7 + 4 * q];
        }
        calculate_ade(mat, v[0], v[1], v[2], v[3], dt);
    } else {
        fprintf(stderr, "Error: material type %d is not supported by the MI355X engine "
                "(supported: 0 linear, 1 tabulated linear, 2 recursive-convolution Drude, 4 recursive-convolution "
                "Drude + 2 critical points, 5 ADE Drude + 2 critical points, 6 piecewise-linear recursive "
                "convolution, 7 ADE Drude + 3 critical points, 10 PEC)\n", type);
        return 1;
    }
    return 0;
}

static int
almost_eq(double a, double b)
{
    const double TOL = 1e-3;
    if (fabs(a) >= 1.0 && fabs(b) >= 1.0)
        return fabs(a / b - 1.0) < TOL;
    return fabs(a - b) < TOL;
}

static int
get_matindex(const SvPool *mp, const SvMatProp *mat)
{
    int i;
    for (i = 0; i < mp->nmat; i++) {
        const SvMatProp *t = mp->mats + i;
        if (t->type != mat->type)
            continue;
        if (mat->type == SV_MAT_PEC
            || (almost_eq(t->epsilon, mat->epsilon) && almost_eq(t->mu, mat->mu) && almost_eq(t->sigma, mat->sigma)
                && almost_eq(t->sigast, mat->sigast) && almost_eq(t->drude_omega_p, mat->drude_omega_p)
                && almost_eq(t->drude_nu, mat->drude_nu) && almost_eq(t->cp3_a[0], mat->cp3_a[0])
                && almost_eq(t->cp3_a[1], mat->cp3_a[1]) && almost_eq(t->cp3_phi[2], mat->cp3_a[2])
                && almost_eq(t->cp3_phi[0], mat->cp3_phi[0]) && almost_eq(t->cp3_phi[1], mat->cp3_phi[1])
                && almost_eq(t->cp3_phi[2], mat->cp3_phi[2]) && almost_eq(t->cp3_omega[0], mat->cp3_omega[0])
                && almost_eq(t->cp3_omega[1], mat->cp3_omega[1]) && almost_eq(t->cp3_omega[2], mat->cp3_omega[2])
                && almost_eq(t->cp3_gamma[0], mat->cp3_gamma[0]) && almost_eq(t->cp3_gamma[1], mat->cp3_gamma[1])
                && almost_eq(t->cp3_gamma[2], mat->cp3_gamma[2])
                /* three-pole materials (no reference counterpart) also compare the third amplitude */
                && (t->ade_npoles < 3 && mat->ade_npoles < 3 ? 1 : almost_eq(t->cp3_a[2], mat->cp3_a[2]))))
            return i;
    }
    return -1;
}

static void
add_material(SvPool *mp, SvMatProp *mat)
{
    if (mp->nmat == 0) {
        mp->mats = (SvMatProp *)calloc(1, sizeof(SvMatProp));
        mp->mats[0].type = SV_MAT_LINTAB;
        mp->mats[0].epsilon = 1;
        mp->mats[0].mu = 1;
        mp->mats[0].pos = 0;
        mp->nmat = 1;
    }
    mp->mats = (SvMatProp *)realloc(mp->mats, (mp->nmat + 1) * sizeof(SvMatProp));
    mp->mats[mp->nmat] = *mat;
    mp->mats[mp->nmat].pos = mp->nmat;
    mat->pos = mp->nmat;
    if (mp->set->sc.verbose > 1)
        printf("material %d of type %d added\n", mp->nmat, mat->type);
    mp->nmat++;
}

/* matmode probes, src3d/plan.c:139-200 (voxel file) and :250-360, :573-711 (vector file) */
static int
test_material_file(const char *fname)
{
    FILE *fr = fopen(fname, "rb");
    int res[3], is_e = 0, is_m = 0;
    size_t n, q;
    float v;
    if (!fr)
        return SV_MATMODE_NONE;
    if (fread(res, sizeof(int), 3, fr) != 3) {
        fclose(fr);
        return SV_MATMODE_NONE;
    }
    n = (size_t)res[0] * res[1] * res[2];
    for (q = 0; q < n; q++) {
        if (fread(&v, sizeof(float), 1, fr) != 1) { fclose(fr); return SV_MATMODE_NONE; }
        if (v != 1) { is_e = 1; break; }
    }
    if (!is_e) {
        for (q = 0; q < n; q++) {
            if (fread(&v, sizeof(float), 1, fr) != 1) { fclose(fr); return SV_MATMODE_NONE; }
            if (v != 0) { is_e = 1; break; }
        }
    }
    /* the magnetic blocks start after both electric blocks */
    fseek(fr, 12 + 2 * n * sizeof(float), SEEK_SET);
    for (q = 0; q < n; q++) {
        if (fread(&v, sizeof(float), 1, fr) != 1) { fclose(fr); return is_e ? SV_MATMODE_ELECTRIC : SV_MATMODE_NONE; }
        if (v != 1) { is_m = 1; break; }
    }
    if (!is_m) {
        fseek(fr, 12 + 3 * n * sizeof(float), SEEK_SET);
        for (q = 0; q < n; q++) {
            if (fread(&v, sizeof(float), 1, fr) != 1) break;
            if (v != 0) { is_m = 1; break; }
        }
    }
    fclose(fr);
    if (is_e && is_m) return SV_MATMODE_FULL;
    if (is_e) return SV_MATMODE_ELECTRIC;
    if (is_m) return SV_MATMODE_MAGNETIC;
    return SV_MATMODE_NONE;
}

/* numbers of an entity line before its material (TetGen: after the file
 * name and the two attribute integers), -1 for an unsupported type */
static int
entity_nfloats(int type)
{
    switch (type) {
    case SV_ENTITY_SPHERE: return 4;
    case SV_ENTITY_CYLINDER: return 7;
    case SV_ENTITY_CONE: return 7;
    case SV_ENTITY_RCONE: return 8;
    case SV_ENTITY_VOXEL: return 6;
    case SV_ENTITY_TETRAHEDRON: return 12;
    case SV_ENTITY_TETGEN: return 6;
    default: return -1;
    }
}

static int
test_script_file(const SvSet *set, const char *fname)
{
    FILE *fr = sv_open_vector_file(set, fname);
    int type, mtype, is_e = 0, is_m = 0;
    float p[16];
    if (!fr) {
        fprintf(stderr, "Error: cannot open file %s for reading\n", fname);
        return SV_MATMODE_ERROR;
    }
    while (fscanf(fr, "%d", &type) == 1) {
        /* src3d/plan.c:590-692: the entity's numbers, then its material */
        int npts = entity_nfloats(type);
        if (npts < 0) {
            fprintf(stderr, "Error: entity type %d is not supported by the MI355X engine (4 sphere, 7 cylinder, "
                            "8 voxel, 12 tetrahedron, 21 TetGen mesh, 107 cone, 108 rounded cone)\n", type);
            fclose(fr);
            return SV_MATMODE_ERROR;
        }
        if (type == SV_ENTITY_TETGEN) {
            char base[SV_NAME_LEN];
            int a, b;
            if (fscanf(fr, "%511s %d %d", base, &a, &b) != 3)
                goto err;
        }
        if (scan_floats(fr, p, npts) || fscanf(fr, "%d", &mtype) != 1)
            goto err;
        if (mtype == SV_MAT_LINEAR || mtype == SV_MAT_LINTAB) {
            if (scan_floats(fr, p, 4))
                goto err;
            if (mtype == SV_MAT_LINEAR) {
                if (p[0] != 1.0 || p[2] > 0) is_e = 1;
                if (p[1] != 1.0 || p[3] > 0) is_m = 1;
            }
        } else if (mtype == SV_MAT_ADE || mtype == SV_MAT_ADE3 || mtype == SV_MAT_DRUDE || mtype == SV_MAT_CP
                   || mtype == SV_MAT_PLRC) {
            const int nv = mtype == SV_MAT_ADE3 ? 16 : mtype == SV_MAT_DRUDE ? 3 : mtype == SV_MAT_CP ? 11 : 12;
            if (scan_floats(fr, p, nv))
                goto err;
            is_e = 1;
        } else if (mtype == SV_MAT_PEC) {
        } else {
            fprintf(stderr, "Error: material type %d is not supported by the MI355X engine (supported: 0, 1, 2, 4 "
                            "and 6 recursive convolution, 5 and 7 ADE Drude + critical points, 10 PEC)\n", mtype);
            goto err;
        }
    }
    fclose(fr);
    if ((is_e + is_m) == 0) return SV_MATMODE_NONE;
    if (is_e == 0 && is_m != 0) return SV_MATMODE_MAGNETIC;
    if (is_e != 0 && is_m == 0) return SV_MATMODE_ELECTRIC;
    return SV_MATMODE_NONE;
err:
    fclose(fr);
    return SV_MATMODE_ERROR;
}

static void
set_data(SvPool *mp, const SvMatProp *mat, int i, int j, int k)
{
    /* src3d/plan.c:2243-2258 */
    SvYeeData *d = mp->d;
    int mm = mp->set->plan.matmode;
    if (mat->type == SV_MAT_LINEAR) {
        if (mm == SV_MATMODE_ELECTRIC || mm == SV_MATMODE_FULL) {
            d->epsilon->data[i][j][k] = (float)(mat->epsilon * EPS0);
            d->sigma->data[i][j][k] = (float)mat->sigma;
        }
        if (mm == SV_MATMODE_MAGNETIC || mm == SV_MATMODE_FULL) {
            d->mu->data[i][j][k] = (float)(mat->mu * MU0);
            d->sigast->data[i][j][k] = (float)mat->sigast;
        }
        if (d->mat)
            d->mat->data[i][j][k] = 0;
    } else
        d->mat->data[i][j][k] = mat->pos;
}

typedef struct {
    int type;
    float p[12];
    SvMatProp mat;
} SvEntity;

/* ------------------------------------------------------------------ */
/* entity geometry, src3d/plan.c:2062-2240: inside tests at voxel      */
/* centres in float arithmetic (the reference's GLfloat), evaluated in */
/* the reference's operation order (the C build has contraction off)   */
/* ------------------------------------------------------------------ */

static float
geo_dist(const float *a, const float *b)
{
    return (float)sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

/* (b - a) . (c - b) */
static float
geo_dot(const float *a, const float *b, const float *c)
{
    return (b[0] - a[0]) * (c[0] - b[0]) + (b[1] - a[1]) * (c[1] - b[1]) + (b[2] - a[2]) * (c[2] - b[2]);
}

static float
geo_norm(float u, float v, float w)
{
    return (float)sqrt(u * u + v * v + w * w);
}

/* distance of point a from the segment b-c, 1e6 beyond either end's
 * perpendicular plane */
static float
geo_segdist(const float *a, const float *b, const float *c)
{
    float u1, u2, u3, v1, v2, v3;
    if (geo_dot(b, c, a) > 0 || geo_dot(c, b, a) > 0)
        return 1e6f;
    u1 = a[0] - b[0]; u2 = a[1] - b[1]; u3 = a[2] - b[2];
    v1 = a[0] - c[0]; v2 = a[1] - c[1]; v3 = a[2] - c[2];
    return (float)fabs(geo_norm(u2 * v3 - u3 * v2, u3 * v1 - u1 * v3, u1 * v2 - u2 * v1)
                       / geo_norm(c[0] - b[0], c[1] - b[1], c[2] - b[2]));
}

static float
geo_det3(float a11, float a12, float a13, float a21, float a22, float a23, float a31, float a32, float a33)
{
    return a11 * a22 * a33 - a11 * a23 * a32 + a12 * a23 * a31 - a12 * a21 * a33 + a13 * a21 * a32 - a13 * a22 * a31;
}

/* 4 x 4 determinant of rows a, b, c, d by expansion along a */
static float
geo_det4(const float *a, const float *b, const float *c, const float *d)
{
    const float m0 = geo_det3(b[1], b[2], b[3], c[1], c[2], c[3], d[1], d[2], d[3]);
    const float m1 = geo_det3(b[0], b[2], b[3], c[0], c[2], c[3], d[0], d[2], d[3]);
    const float m2 = geo_det3(b[0], b[1], b[3], c[0], c[1], c[3], d[0], d[1], d[3]);
    const float m3 = geo_det3(b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]);
    return a[0] * m0 - a[1] * m1 + a[2] * m2 - a[3] * m3;
}

static int
geo_in_cylinder(const float *x, const float *p1, const float *p2, float r)
{
    return geo_segdist(x, p1, p2) < r;
}

/* radius grows from 0 at p1 to r at p2, measured by the distance from p1 */
static int
geo_in_cone(const float *x, const float *p1, const float *p2, float r)
{
    const float length = geo_dist(p1, p2), pos = geo_dist(x, p1);
    return geo_segdist(x, p1, p2) < pos * r / length;
}

/* truncated cone, radius r1 at p1 and r2 at p2 along the axis */
static int
geo_in_rcone(const float *x, const float *p1, const float *p2, float r1, float r2)
{
    const float length = geo_dist(p1, p2);
    float n[3], q[3], t;
    n[0] = (p2[0] - p1[0]) / length;
    n[1] = (p2[1] - p1[1]) / length;
    n[2] = (p2[2] - p1[2]) / length;
    t = (x[0] - p1[0]) * n[0] + (x[1] - p1[1]) * n[1] + (x[2] - p1[2]) * n[2];
    q[0] = p1[0] + t * n[0];
    q[1] = p1[1] + t * n[1];
    q[2] = p1[2] + t * n[2];
    return t >= 0 && t <= length && geo_dist(x, q) <= r1 + t * (r2 - r1) / length;
}

/* the point and each vertex-replaced determinant have one sign (zero counts
 * as either), src3d/plan.c:2213-2240 */
static int
geo_in_tetrahedron(const float *x, const float *p)
{
    float v[4][4], xp[4], d[5];
    int q;
    for (q = 0; q < 4; q++) {
        v[q][0] = p[3 * q];
        v[q][1] = p[3 * q + 1];
        v[q][2] = p[3 * q + 2];
        v[q][3] = 1;
    }
    xp[0] = x[0]; xp[1] = x[1]; xp[2] = x[2]; xp[3] = 1;
    d[0] = geo_det4(v[0], v[1], v[2], v[3]);
    d[1] = geo_det4(xp, v[1], v[2], v[3]);
    d[2] = geo_det4(v[0], xp, v[2], v[3]);
    d[3] = geo_det4(v[0], v[1], xp, v[3]);
    d[4] = geo_det4(v[0], v[1], v[2], xp);
    return (d[0] >= 0 && d[1] >= 0 && d[2] >= 0 && d[3] >= 0 && d[4] >= 0)
           || (d[0] <= 0 && d[1] <= 0 && d[2] <= 0 && d[3] <= 0 && d[4] <= 0);
}

static float
fmin2(float a, float b)
{
    return a < b ? a : b;
}

static float
fmax2(float a, float b)
{
    return a > b ? a : b;
}

/* TetGen mesh (base.node, base.ele) as tetrahedron entities appended to
 * *ents, src3d/plan.c:3410-3626: node coordinates are read as float, scaled
 * and shifted in double and rounded to float; comment and blank lines are
 * skipped only before the two headers; nodes may be numbered from 0 or 1;
 * node attributes are read as integers and boundary markers are not read.
 * An element is kept when the files have no attributes or the filter
 * (position, value) is off (-1); the reference then compares an attribute
 * past the ones it read, which keeps nothing here. */
static void
tetgen_skip_comments(FILE *fr)
{
    char line[256];
    int c;
    while ((c = getc(fr)) != EOF) {
        ungetc(c, fr);
        if (c != '#' && c != '\n')
            break;
        if (!fgets(line, sizeof(line), fr))
            break;
    }
}

static int
load_tetgen(const char *base, int att_pos, int att_val, const float *shift, const float *mult, const SvMatProp *mat,
            SvEntity **ents, int *nent)
{
    char fname[SV_NAME_LEN + 8];
    FILE *fr;
    int nn, dim, natt, nbm, i, j, nt, npe, nattt, buf, corr = 0, kept = 0;
    float *xyz;
    snprintf(fname, sizeof(fname), "%s.node", base);
    if (!(fr = fopen(fname, "r"))) {
        fprintf(stderr, "Error: node file %s cannot be opened for reading\n", fname);
        return 1;
    }
    tetgen_skip_comments(fr);
    if (fscanf(fr, "%d %d %d %d", &nn, &dim, &natt, &nbm) != 4 || nn <= 0) {
        fclose(fr);
        fprintf(stderr, "Error: cannot read the header of %s\n", fname);
        return 1;
    }
    tetgen_skip_comments(fr);
    xyz = (float *)malloc(3 * (size_t)nn * sizeof(float));
    for (i = 0; i < nn; i++) {
        float v[3];
        if (fscanf(fr, "%d %f %f %f", &buf, v, v + 1, v + 2) != 4) {
            fclose(fr);
            free(xyz);
            fprintf(stderr, "Error: cannot read node %d of %s\n", i, fname);
            return 1;
        }
        if (i == 0 && buf == 1)
            corr = 1;
        for (j = 0; j < 3; j++)
            xyz[3 * i + j] = (float)((double)v[j] * (double)mult[j] + (double)shift[j]);
        for (j = 0; j < natt; j++)
            if (fscanf(fr, "%d", &buf) != 1)
                break;
    }
    fclose(fr);

    snprintf(fname, sizeof(fname), "%s.ele", base);
    if (!(fr = fopen(fname, "r"))) {
        free(xyz);
        fprintf(stderr, "Error: element file %s cannot be opened for reading\n", fname);
        return 1;
    }
    tetgen_skip_comments(fr);
    if (fscanf(fr, "%d %d %d", &nt, &npe, &nattt) != 3 || nt < 0 || (npe != 4 && npe != 10)) {
        fclose(fr);
        free(xyz);
        fprintf(stderr, "Error: cannot read the header of %s\n", fname);
        return 1;
    }
    tetgen_skip_comments(fr);
    for (i = 0; i < nt; i++) {
        int tp[4];
        if (fscanf(fr, "%d %d %d %d %d", &buf, tp, tp + 1, tp + 2, tp + 3) != 5)
            goto bad;
        for (j = 4; j < npe; j++)
            if (fscanf(fr, "%d", &buf) != 1)
                goto bad;
        for (j = 0; j < nattt; j++)
            if (fscanf(fr, "%d", &buf) != 1)
                goto bad;
        if (nattt == 0 || att_pos == -1 || att_val == -1) {
            SvEntity e;
            memset(&e, 0, sizeof(e));
            e.type = SV_ENTITY_TETRAHEDRON;
            for (j = 0; j < 4; j++) {
                const int node = tp[j] - corr;
                if (node < 0 || node >= nn)
                    goto bad;
                e.p[3 * j] = xyz[3 * node];
                e.p[3 * j + 1] = xyz[3 * node + 1];
                e.p[3 * j + 2] = xyz[3 * node + 2];
            }
            e.mat = *mat;
            *ents = (SvEntity *)realloc(*ents, (*nent + 1) * sizeof(SvEntity));
            (*ents)[(*nent)++] = e;
            kept++;
        }
    }
    fclose(fr);
    free(xyz);
    if (!kept && nt > 0)
        fprintf(stderr, "Warning: no element of TetGen mesh %s passes its attribute filter (reference behaviour)\n", base);
    return 0;
bad:
    fclose(fr);
    free(xyz);
    fprintf(stderr, "Error: cannot read element %d of %s\n", i, fname);
    return 1;
}

/* fast reader for the ASCII STRUCTURED_POINTS integer map, src3d/plan.c:3877-3960 */
static int
load_vtk(const char *fname, SvPool *mp)
{
    FILE *fr = fopen(fname, "r");
    char line[256], word[64];
    int xres, yres, zres, nexp, i, j, k;
    SvSet *set = mp->set;
    char *buf, *s, *end;
    long pos, len;

    if (!fr) {
        fprintf(stderr, "Error: can not open file %s\n", fname);
        return 1;
    }
    if (!fgets(line, sizeof(line), fr) || !fgets(line, sizeof(line), fr))
        goto bad;
    if (fscanf(fr, "%63s", word) != 1 || strcmp(word, "ASCII")) {
        fprintf(stderr, "Error in VTK file, it should be ascii\n");
        goto bad;
    }
    if (fscanf(fr, "%63s %63s", word, word) != 2 || strcmp(word, "STRUCTURED_POINTS")) {
        fprintf(stderr, "Error in VTK file, it should be structured points dataset\n");
        goto bad;
    }
    if (fscanf(fr, "%63s", word) != 1 || strcmp(word, "DIMENSIONS")
        || fscanf(fr, "%d %d %d", &xres, &yres, &zres) != 3) {
        fprintf(stderr, "Error in VTK file, DIMENSIONS keyword is expected here\n");
        goto bad;
    }
    if (xres != set->sp.xres || yres != set->sp.yres || zres != set->sp.zres) {
        fprintf(stderr, "Error in VTK file, Dimensions do not match expected values\n");
        goto bad;
    }
    for (i = 0; i < 3; i++)
        if (!fgets(line, sizeof(line), fr))
            goto bad;
    if (fscanf(fr, "%63s %d", word, &nexp) != 2 || strcmp(word, "POINT_DATA") || nexp != xres * yres * zres) {
        fprintf(stderr, "Error in VTK file, POINT_DATA does not match dimensions\n");
        goto bad;
    }
    for (i = 0; i < 3; i++)
        if (!fgets(line, sizeof(line), fr))
            goto bad;
    pos = ftell(fr);
    fseek(fr, 0, SEEK_END);
    len = ftell(fr) - pos;
    fseek(fr, pos, SEEK_SET);
    buf = (char *)malloc(len + 1);
    if (!buf || fread(buf, 1, len, fr) != (size_t)len) {
        free(buf);
        goto bad;
    }
    buf[len] = 0;
    s = buf;
    for (k = 0; k < zres; k++)
        for (j = 0; j < yres; j++)
            for (i = 0; i < xres; i++) {
                long v = strtol(s, &end, 10);
                if (end == s) {
                    fprintf(stderr, "Error in VTK file, Error while reading datapoints\n");
                    free(buf);
                    goto bad;
                }
                mp->d->mat->data[i][j][k] = (int)v;
                s = end;
            }
    free(buf);
    fclose(fr);
    return 0;
bad:
    fclose(fr);
    return 1;
}

/* MEDIUM_CROP, src3d/plan.c:5244-5285: voxels on or outside the crop box
 * faces take the crop medium, as float per-voxel data without a material
 * table, else as one tabulated material (the reference gives that one
 * sigma = sigma* = 0 whatever the keyword says).  The tabulated material is
 * looked up with its relative constants, as the reference does, but stored
 * with absolute ones: the table was converted to absolute units before the
 * crop runs (sv_pool_add_material's multme, src3d/plan.c:2439-2442). */
static void
apply_crop(SvPool *mp, SvSet *set)
{
    SvYeeData *d = mp->d;
    const int xres = set->sp.xres, yres = set->sp.yres, zres = set->sp.zres;
    int i, j, k;
    if (set->sm.crop) {
        const SvSetMat *m = &set->sm;
        SvMatProp cm;
        int idx = -1;
        if (mp->nmat > 0) {
            memset(&cm, 0, sizeof(cm));
            cm.type = SV_MAT_LINTAB;
            cm.epsilon = m->crop_epsilon;
            cm.mu = m->crop_mu;
            idx = get_matindex(mp, &cm);
            if (idx == -1) {
                cm.epsilon *= EPS0;
                cm.mu *= MU0;
                add_material(mp, &cm);
                idx = cm.pos;
            }
        }
        for (i = 0; i < xres; i++)
            for (j = 0; j < yres; j++)
                for (k = 0; k < zres; k++) {
                    if (!(i <= m->crop_i0 || i >= m->crop_in || j <= m->crop_j0 || j >= m->crop_jn || k <= m->crop_k0
                          || k >= m->crop_kn))
                        continue;
                    if (mp->nmat == 0) {
                        if (set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_ELECTRIC) {
                            d->epsilon->data[i][j][k] = (float)(m->crop_epsilon * EPS0);
                            d->sigma->data[i][j][k] = (float)m->crop_sigma;
                        }
                        if (set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_MAGNETIC) {
                            d->mu->data[i][j][k] = (float)(m->crop_mu * MU0);
                            d->sigast->data[i][j][k] = (float)m->crop_sigast;
                        }
                    } else
                        d->mat->data[i][j][k] = idx;
                }
    }
}

static int
load_data(SvPool *mp, SvSet *set)
{
    SvYeeData *d = mp->d;
    int xres = set->sp.xres, yres = set->sp.yres, zres = set->sp.zres;
    int i, j, k, n;

    if (set->sm.in_voxel_filename[0]) {
        FILE *fr = fopen(set->sm.in_voxel_filename, "rb");
        if (!fr)
            fprintf(stderr, "Error: cannot open material file %s for reading\n", set->sm.in_voxel_filename);
        if (set->plan.matmode != SV_MATMODE_NONE && fr) {
            int res[3];
            size_t plane = (size_t)zres;
            if (fread(res, sizeof(int), 3, fr) != 3 || res[0] != xres || res[1] != yres || res[2] != zres) {
                fprintf(stderr, "Error: material dimensions differ from pool settings\n");
                fclose(fr);
                return 1;
            }
            if (set->plan.matmode == SV_MATMODE_MAGNETIC) {
                /* src3d/plan.c:4148: absolute seek that ignores the 12-byte header */
                fseek(fr, 2 * (long)xres * yres * zres * sizeof(float), SEEK_SET);
            } else {
                for (i = 0; i < xres; i++)
                    for (j = 0; j < yres; j++)
                        if (fread(d->epsilon->data[i][j], sizeof(float), plane, fr) != plane) goto short_read;
                for (n = 0; n < xres * yres * zres; n++)
                    d->epsilon->block[n] *= (float)EPS0;
                for (i = 0; i < xres; i++)
                    for (j = 0; j < yres; j++)
                        if (fread(d->sigma->data[i][j], sizeof(float), plane, fr) != plane) goto short_read;
            }
            if (set->plan.matmode == SV_MATMODE_ELECTRIC) {
                /* src3d/plan.c:4162-4165: returns before the vector file, the crop
                 * (src3d/plan.c:5244, same function) and the boundary map */
                fclose(fr);
                return 0;
            }
            for (i = 0; i < xres; i++)
                for (j = 0; j < yres; j++)
                    if (fread(d->mu->data[i][j], sizeof(float), plane, fr) != plane) goto short_read;
            for (n = 0; n < xres * yres * zres; n++)
                d->mu->block[n] *= (float)MU0;
            for (i = 0; i < xres; i++)
                for (j = 0; j < yres; j++)
                    if (fread(d->sigast->data[i][j], sizeof(float), plane, fr) != plane) goto short_read;
            fclose(fr);
            fr = NULL;
        }
        if (fr)
            fclose(fr);
    }

    if (set->sm.in_vector_filename[0]) {
        FILE *fr = sv_open_vector_file(set, set->sm.in_vector_filename);
        SvEntity *ents = NULL;
        int nent = 0, type;
        if (!fr) {
            fprintf(stderr, "Error: cannot open material vector file for loading.\n");
            return 1;
        }
        while (fscanf(fr, "%d", &type) == 1) {
            SvEntity e;
            memset(&e, 0, sizeof(e));
            e.type = type;
            if (type == SV_ENTITY_TETGEN) {
                char base[SV_NAME_LEN];
                int att_pos, att_val;
                if (fscanf(fr, "%511s %d %d", base, &att_pos, &att_val) != 3 || scan_floats(fr, e.p, 6)
                    || scan_material(fr, &e.mat, set->plan.dt))
                    goto vec_err;
                if (e.mat.type != SV_MAT_LINEAR) {
                    int mi = get_matindex(mp, &e.mat);
                    if (mi == -1)
                        add_material(mp, &e.mat);
                    else
                        e.mat.pos = mi;
                }
                if (load_tetgen(base, att_pos, att_val, e.p, e.p + 3, &e.mat, &ents, &nent))
                    goto vec_err;
                continue;
            }
            if (entity_nfloats(type) < 0) {
                fprintf(stderr, "Error: entity type %d is not supported by the MI355X engine\n", type);
                goto vec_err;
            }
            if (scan_floats(fr, e.p, entity_nfloats(type)))
                goto vec_err;
            if (scan_material(fr, &e.mat, set->plan.dt))
                goto vec_err;
            if (e.mat.type != SV_MAT_LINEAR) {
                int mi = get_matindex(mp, &e.mat);
                if (mi == -1)
                    add_material(mp, &e.mat);
                else
                    e.mat.pos = mi;
            }
            ents = (SvEntity *)realloc(ents, (nent + 1) * sizeof(SvEntity));
            ents[nent++] = e;
        }
        fclose(fr);

        if (mp->nmat != 0)
            d->mat = sv_icube_new(xres, yres, zres);
        if (set->sm.in_vtk_data_filename[0]) {
            if (!d->mat) {
                fprintf(stderr, "Error: MEDIUM_VTK_DATA needs at least one tabulated material in MEDIUM_VECTOR\n");
                free(ents);
                return 1;
            }
            if (load_vtk(set->sm.in_vtk_data_filename, mp)) {
                free(ents);
                return 1;
            }
            for (n = 0; n < xres * yres * zres; n++)
                if (d->mat->block[n] < 0 || d->mat->block[n] >= mp->nmat) {
                    fprintf(stderr, "Error: MEDIUM_VTK_DATA holds material index %d, but only %d materials are defined\n",
                            d->mat->block[n], mp->nmat);
                    free(ents);
                    return 1;
                }
        }

        for (n = 0; n < nent; n++) {
            SvEntity *e = ents + n;
            if (e->type == SV_ENTITY_SPHERE) {
                /* src3d/plan.c:4378-4391, float arithmetic as GLfloat */
                float cx = e->p[0], cy = e->p[1], cz = e->p[2], r = e->p[3];
                for (i = (int)(0 > cx - r - 1 ? 0 : cx - r - 1); i < (xres < cx + r + 1 ? xres : cx + r + 1); i++)
                    for (j = (int)(0 > cy - r - 1 ? 0 : cy - r - 1); j < (yres < cy + r + 1 ? yres : cy + r + 1); j++)
                        for (k = (int)(0 > cz - r - 1 ? 0 : cz - r - 1); k < (zres < cz + r + 1 ? zres : cz + r + 1); k++) {
                            float x = i + 0.5f, y = j + 0.5f, z = k + 0.5f;
                            float ddx = x - cx, ddy = y - cy, ddz = z - cz;
                            volatile float s = ddx * ddx + ddy * ddy + ddz * ddz;
                            volatile float rr = r * r;
                            if (s <= rr)
                                set_data(mp, &e->mat, i, j, k);
                        }
            } else if (e->type == SV_ENTITY_VOXEL) {
                /* src3d/plan.c:4467-4479 */
                for (i = (int)(0 > e->p[0] ? 0 : e->p[0]); i < (xres < e->p[3] ? xres : e->p[3]); i++)
                    for (j = (int)(0 > e->p[1] ? 0 : e->p[1]); j < (yres < e->p[4] ? yres : e->p[4]); j++)
                        for (k = (int)(0 > e->p[2] ? 0 : e->p[2]); k < (zres < e->p[5] ? zres : e->p[5]); k++)
                            set_data(mp, &e->mat, i, j, k);
            } else if (e->type == SV_ENTITY_TETRAHEDRON) {
                /* src3d/plan.c:4418-4446: bounding box truncated to integers and clamped, inclusive */
                int lo[3], hi[3], a;
                for (a = 0; a < 3; a++) {
                    const int res = a == 0 ? xres : a == 1 ? yres : zres;
                    lo[a] = (int)fmin2(fmin2(e->p[a], e->p[3 + a]), fmin2(e->p[6 + a], e->p[9 + a]));
                    hi[a] = (int)fmax2(fmax2(e->p[a], e->p[3 + a]), fmax2(e->p[6 + a], e->p[9 + a]));
                    lo[a] = lo[a] < 0 ? 0 : lo[a] > res - 1 ? res - 1 : lo[a];
                    hi[a] = hi[a] < 0 ? 0 : hi[a] > res - 1 ? res - 1 : hi[a];
                }
                if (geo_det4((const float[4]){ e->p[0], e->p[1], e->p[2], 1 }, (const float[4]){ e->p[3], e->p[4], e->p[5], 1 },
                             (const float[4]){ e->p[6], e->p[7], e->p[8], 1 }, (const float[4]){ e->p[9], e->p[10], e->p[11], 1 }) == 0)
                    fprintf(stderr, "Warning: degenerated tetrahedron\n");
                for (i = lo[0]; i <= hi[0]; i++)
                    for (j = lo[1]; j <= hi[1]; j++)
                        for (k = lo[2]; k <= hi[2]; k++) {
                            const float x[3] = { i + 0.5f, j + 0.5f, k + 0.5f };
                            if (geo_in_tetrahedron(x, e->p))
                                set_data(mp, &e->mat, i, j, k);
                        }
            } else {
                /* cylinder, cone, rounded cone, src3d/plan.c:4516-4587: the box of both end points
                 * widened by the (largest) radius + 1, float bounds as in the sphere loop */
                const float *p1 = e->p, *p2 = e->p + 3;
                const float r = e->type == SV_ENTITY_RCONE ? fmax2(e->p[6], e->p[7]) : e->p[6];
                const float x0 = fmin2(p1[0], p2[0]) - r - 1, x1 = fmax2(p1[0], p2[0]) + r + 1;
                const float y0 = fmin2(p1[1], p2[1]) - r - 1, y1 = fmax2(p1[1], p2[1]) + r + 1;
                const float z0 = fmin2(p1[2], p2[2]) - r - 1, z1 = fmax2(p1[2], p2[2]) + r + 1;
                for (i = (int)(0 > x0 ? 0 : x0); i < (xres < x1 ? xres : x1); i++)
                    for (j = (int)(0 > y0 ? 0 : y0); j < (yres < y1 ? yres : y1); j++)
                        for (k = (int)(0 > z0 ? 0 : z0); k < (zres < z1 ? zres : z1); k++) {
                            const float x[3] = { i + 0.5f, j + 0.5f, k + 0.5f };
                            int in;
                            if (e->type == SV_ENTITY_CYLINDER)
                                in = geo_in_cylinder(x, p1, p2, e->p[6]);
                            else if (e->type == SV_ENTITY_CONE)
                                in = geo_in_cone(x, p1, p2, e->p[6]);
                            else
                                in = geo_in_rcone(x, p1, p2, e->p[6], e->p[7]);
                            if (in)
                                set_data(mp, &e->mat, i, j, k);
                        }
            }
        }
        free(ents);

        /* src3d/plan.c:4640-4699: relative -> absolute material constants;
         * the Drude type's omega_p and nu are multiplied by EV_TO_J there */
        for (n = 0; n < mp->nmat; n++) {
            if (sv_material_is_tabulated(mp->mats[n].type)) {
                mp->mats[n].epsilon *= EPS0;
                mp->mats[n].mu *= MU0;
            }
            if (mp->mats[n].type == SV_MAT_DRUDE) {
                mp->mats[n].drude_omega_p *= SV_EV_TO_J;
                mp->mats[n].drude_nu *= SV_EV_TO_J;
            }
        }
        goto after_vector;
vec_err:
        fprintf(stderr, "Error: cannot parse material vector file %s\n", set->sm.in_vector_filename);
        fclose(fr);
        free(ents);
        return 1;
    }
after_vector:

    apply_crop(mp, set);

    /* TSF and layered-TSF skip maps, src3d/plan.c:5297-5357: cells near the
     * box faces with non-vacuum material within the skip depth; the layered
     * TSF's map replaces the TSF's when both ask for one, as there */
    for (n = 0; n < 2; n++) {
        const int sd = n ? set->ss.ltsf.box_boundary_skipdepth : set->ss.tsf.box_boundary_skipdepth;
        int b[6];
        if (sd < 0 || (n ? !mp->src->ltsf : !mp->src->tsf))
            continue;
        if (n) {
            const SvSourceLTSF *t = mp->src->ltsf;
            b[0] = t->box_i0; b[1] = t->box_j0; b[2] = t->box_k0; b[3] = t->box_in; b[4] = t->box_jn; b[5] = t->box_kn;
        } else {
            const SvSourceTSF *t = mp->src->tsf;
            b[0] = t->box_i0; b[1] = t->box_j0; b[2] = t->box_k0; b[3] = t->box_in; b[4] = t->box_jn; b[5] = t->box_kn;
        }
        sv_icube_free(d->sourceskip);
        d->sourceskip = sv_icube_new(xres, yres, zres);
        #pragma omp parallel for private(j, k) schedule(dynamic, 1)
        for (i = 0; i < xres; i++)
            for (j = 0; j < yres; j++)
                for (k = 0; k < zres; k++) {
                    int ii, jj, kk, hit = 0;
                    if (i > (b[0] + sd) && j > (b[1] + sd) && k > (b[2] + sd)
                        && i < (b[3] - sd) && j < (b[4] - sd) && k < (b[5] - sd))
                        continue;
                    for (ii = (i - sd > 0 ? i - sd : 0); ii < (i + sd < xres ? i + sd : xres) && !hit; ii++)
                        for (jj = (j - sd > 0 ? j - sd : 0); jj < (j + sd < yres ? j + sd : yres) && !hit; jj++)
                            for (kk = (k - sd > 0 ? k - sd : 0); kk < (k + sd < zres ? k + sd : zres) && !hit; kk++) {
                                /* sv_yee_data_get_tsf_mattype, src3d/yee.c:1684-1705 */
                                int m = (mp->nmat > 0) ? d->mat->data[ii][jj][kk] : -1, vac = 0;
                                if (m == 0 || m == -1) {
                                    if ((set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_ELECTRIC)
                                        && fabs(1.0 - d->epsilon->data[ii][jj][kk] / EPS0) < 0.001)
                                        vac = 1;
                                    if (set->plan.matmode == SV_MATMODE_NONE)
                                        vac = 1;
                                }
                                if (m > 0 && mp->mats[m].type == SV_MAT_LINTAB && fabs(1.0 - mp->mats[m].epsilon / EPS0) < 0.001)
                                    vac = 1;
                                if (!vac)
                                    hit = 1;
                            }
                    if (hit)
                        d->sourceskip->data[i][j][k] = 1;
                }
    }

    /* internal boundaries, src3d/plan.c:5381-5417 */
    sv_icube_fill(d->bnds, 0);
    #pragma omp parallel for private(j, k) schedule(static)
    for (i = 0; i < xres; i++)
        for (j = 0; j < yres; j++)
            for (k = 0; k < zres; k++) {
                int ii, jj, kk, m, nm;
                double e, s, ne, ns;
                if (i == 0 || j == 0 || k == 0 || i == xres - 1 || j == yres - 1 || k == zres - 1) {
                    d->bnds->data[i][j][k] = 1;
                    continue;
                }
                m = sv_pool_epsilon_sigma(mp, i, j, k, &e, &s);
                for (ii = -1; ii <= 1; ii++)
                    for (jj = -1; jj <= 1; jj++)
                        for (kk = -1; kk <= 1; kk++) {
                            if (d->bnds->data[i][j][k] == 1)
                                continue;
                            nm = sv_pool_epsilon_sigma(mp, i + ii, j + jj, k + kk, &ne, &ns);
                            if (nm != m || ne != e || ns != s)
                                d->bnds->data[i][j][k] = 1;
                        }
            }
    return 0;

short_read:
    fprintf(stderr, "Error: material file %s is shorter than the pool\n", set->sm.in_voxel_filename);
    return 1;
}

/* ------------------------------------------------------------------ */
/* CPML grading (src3d/boundary.c:33-99, 101-280)                      */
/* ------------------------------------------------------------------ */

static void
becebhch(int near, int i, int imax, double sigma, double kappa, double a, int m, double dt,
         double *be, double *ce, double *bh, double *ch, double *vkappa, double *hkappa)
{
    /* near faces use the H offset -0.5 (becebhch0), far faces +0.5 (becebhch) */
    double off = near ? -0.5 : 0.5;
    double powdepthe = pow((imax - i) / (double)(imax - 0), m);
    double powdepthh = pow((imax - i + off) / (double)(imax - 0), m);
    double sve = sigma * powdepthe;
    double kve = 1 + (kappa - 1) * powdepthe;
    double ave = a * pow(((i - 0) / ((double)imax - 0)), 1.0);
    double svh = sigma * powdepthh;
    double kvh = 1 + (kappa - 1) * powdepthh;
    /* alpha grows into the layer: (i - 0.5) on near faces, (i + 0.5) on far faces */
    double avh = a * pow(((i + off) / ((double)imax - 0)), 1.0);

    *be = exp(-(sve / kve + ave) * dt / EPS0);
    *ce = (sve == 0) ? 0 : (sve / (sve + kve * ave) / kve) * ((*be) - 1.0);
    *bh = exp(-(svh / kvh + avh) * dt / EPS0);
    *ch = (svh == 0) ? 0 : (svh / (svh + kvh * avh) / kvh) * ((*bh) - 1.0);
    *vkappa = kve;
    *hkappa = kvh;
}

static void
cpml_face(SvCpmlFace *f, int near, int depth, double sigma, double kappa, double a, int m, double dt)
{
    int i;
    f->depth = depth;
    f->be = (double *)calloc(6 * (size_t)depth, sizeof(double));
    f->ce = f->be + depth;
    f->bh = f->ce + depth;
    f->ch = f->bh + depth;
    f->kappae = f->ch + depth;
    f->kappah = f->kappae + depth;
    for (i = 0; i < depth; i++) {
        int arg = near ? i : depth - i;
        becebhch(near, arg, depth, sigma, kappa, a, m, dt, f->be + i, f->ce + i, f->bh + i, f->ch + i,
                 f->kappae + i, f->kappah + i);
    }
}

static SvBoundary *
boundary_new(SvSet *set)
{
    SvBoundary *b = (SvBoundary *)calloc(1, sizeof(SvBoundary));
    SvSetBoundary *s = &set->sb;
    double dt = set->plan.dt;
    if (s->bx0 == SV_BOUNDARY_CPML) cpml_face(&b->x0, 1, s->depth_bx0, s->sigma_bx0, s->kappa_bx0, s->a_bx0, s->m_bx0, dt);
    if (s->bxn == SV_BOUNDARY_CPML) cpml_face(&b->xn, 0, s->depth_bxn, s->sigma_bxn, s->kappa_bxn, s->a_bxn, s->m_bxn, dt);
    if (s->by0 == SV_BOUNDARY_CPML) cpml_face(&b->y0, 1, s->depth_by0, s->sigma_by0, s->kappa_by0, s->a_by0, s->m_by0, dt);
    if (s->byn == SV_BOUNDARY_CPML) cpml_face(&b->yn, 0, s->depth_byn, s->sigma_byn, s->kappa_byn, s->a_byn, s->m_byn, dt);
    if (s->bz0 == SV_BOUNDARY_CPML) cpml_face(&b->z0, 1, s->depth_bz0, s->sigma_bz0, s->kappa_bz0, s->a_bz0, s->m_bz0, dt);
    if (s->bzn == SV_BOUNDARY_CPML) cpml_face(&b->zn, 0, s->depth_bzn, s->sigma_bzn, s->kappa_bzn, s->a_bzn, s->m_bzn, dt);
    return b;
}

/* ------------------------------------------------------------------ */
/* far field (src3d/farfield.c:56-172)                                  */
/* ------------------------------------------------------------------ */

static SvFarfield *
farfield_new(SvSet *set)
{
    SvFarfield *ff = (SvFarfield *)calloc(1, sizeof(SvFarfield));
    double div = SV_LIGHT_SPEED * set->plan.dt / set->sp.dx;
    int n, set_now = 0, set_istart = 0;

    ff->ndata = (int)(2000 + 2 * set->sc.nsteps
                      + sqrt((double)(set->sp.xres * set->sp.xres + set->sp.yres * set->sp.yres + set->sp.zres * set->sp.zres)) / div);
    ff->nrs = set->sf.nrs;
    if (ff->nrs)
        ff->rpoints = (SvRPoint *)calloc(ff->nrs, sizeof(SvRPoint));
    for (n = 0; n < ff->nrs; n++) {
        double ri = set->sf.ri[n] / div, rj = set->sf.rj[n] / div, rk = set->sf.rk[n] / div;
        double v = sqrt(ri * ri + rj * rj + rk * rk) - 500;
        const int ind = set->sf.individual ? set->sf.individual[n] : 0;
        ff->rpoints[n].istart = (int)(v > 0 ? v : 0);
        /* the points of an area set share the first point's istart
         * (src3d/farfield.c:81-91) */
        if (ind > 0) {
            if (ind != set_now) {
                set_now = ind;
                set_istart = ff->rpoints[n].istart;
            }
            ff->rpoints[n].istart = set_istart;
        }
        ff->rpoints[n].ex = (double *)calloc(3 * (size_t)ff->ndata, sizeof(double));
        ff->rpoints[n].ey = ff->rpoints[n].ex + ff->ndata;
        ff->rpoints[n].ez = ff->rpoints[n].ey + ff->ndata;
    }
    ff->ndft = set->sf.ndft;
    if (ff->ndft) {
        int q;
        long ix = set->sf.box_in - set->sf.box_i0, iy = set->sf.box_jn - set->sf.box_j0, iz = set->sf.box_kn - set->sf.box_k0;
        ff->dft_omega = (double *)malloc(ff->ndft * sizeof(double));
        for (q = 0; q < ff->ndft; q++)
            ff->dft_omega[q] = 2 * M_PI * SV_LIGHT_SPEED / set->sf.dft_lambda[q];
        ff->dft_ncells = 2 * (iy * iz + ix * iz + ix * iy);
        ff->dft_data = (double *)calloc((size_t)ff->dft_ncells * 4 * ff->ndft * 2, sizeof(double));
    }
    ff->nflux = set->sx.nplanes ? set->sx.nw : 0;
    if (ff->nflux) {
        int q;
        ff->flux_omega = (double *)malloc(ff->nflux * sizeof(double));
        for (q = 0; q < ff->nflux; q++)
            ff->flux_omega[q] = 2 * M_PI * SV_LIGHT_SPEED / set->sx.lambda[q];
        ff->flux_ncells = (long)set->sx.nplanes * (set->sx.i1 - set->sx.i0) * (set->sx.j1 - set->sx.j0);
        ff->flux_data = (double *)calloc((size_t)ff->flux_ncells * ff->nflux * 8, sizeof(double));
    }
    return ff;
}

/* periodic far field, src3d/pfarfield.c:36-62 (the postprocess plane store
 * is not allocated here: the engine integrates every step from ppstart on) */
static SvPFarfield *
pfarfield_new(SvSet *set)
{
    SvPFarfield *ff = (SvPFarfield *)calloc(1, sizeof(SvPFarfield));
    double div = SV_LIGHT_SPEED * set->plan.dt / set->sp.dx;
    int n;

    ff->ndata = 1000 + 2 * set->sc.nsteps
                + (int)(sqrt((double)(set->sp.xres * set->sp.xres + set->sp.yres * set->sp.yres + set->sp.zres * set->sp.zres))
                        / div);
    ff->nrs = set->spf.nrs;
    if (ff->nrs)
        ff->prpoints = (SvRPoint *)calloc(ff->nrs, sizeof(SvRPoint));
    for (n = 0; n < ff->nrs; n++) {
        double ri = set->spf.ri[n] / div, rj = set->spf.rj[n] / div, rk = set->spf.rk[n] / div;
        double v = sqrt(ri * ri + rj * rj + rk * rk) - 500;
        ff->prpoints[n].istart = (int)(v > 0 ? v : 0);
        ff->prpoints[n].ex = (double *)calloc(3 * (size_t)ff->ndata, sizeof(double));
        ff->prpoints[n].ey = ff->prpoints[n].ex + ff->ndata;
        ff->prpoints[n].ez = ff->prpoints[n].ey + ff->ndata;
    }
    return ff;
}

/* ------------------------------------------------------------------ */
/* planner entry point                                                  */
/* ------------------------------------------------------------------ */

SvPool *
init_and_make_plan(SvSet *set)
{
    SvPool *mp = (SvPool *)calloc(1, sizeof(SvPool));
    SvYeeData *d;
    int xres = set->sp.xres, yres = set->sp.yres, zres = set->sp.zres;
    int matmode_mat, matmode_script, n;

    set->plan.gpumode = SV_GPUMODE_NONE;
    set->plan.dt = 1.0 / SV_LIGHT_SPEED / sqrt(1.0 / set->sp.dx / set->sp.dx + 1.0 / set->sp.dy / set->sp.dy
                                               + 1.0 / set->sp.dz / set->sp.dz) * set->sc.dtmult;
    mp->set = set;
    mp->d = d = (SvYeeData *)calloc(1, sizeof(SvYeeData));
    d->dt = set->plan.dt;
    mp->src = (SvSource *)calloc(1, sizeof(SvSource));
    mp->out = (SvOutput *)calloc(1, sizeof(SvOutput));

    if (xres < 3 || yres < 3 || zres < 3) {
        fprintf(stderr, "Error: pool must be at least 3 cells along each axis\n");
        return NULL;
    }
    /* domain-periodic axes (no reference counterpart): x and y only, both faces,
     * not combined with the interior periodic planes of MBOUNDARY */
    if ((set->sb.bx0 == SV_BOUNDARY_PERIODIC) != (set->sb.bxn == SV_BOUNDARY_PERIODIC)
        || (set->sb.by0 == SV_BOUNDARY_PERIODIC) != (set->sb.byn == SV_BOUNDARY_PERIODIC)
        || set->sb.bz0 == SV_BOUNDARY_PERIODIC || set->sb.bzn == SV_BOUNDARY_PERIODIC) {
        fprintf(stderr, "Error: periodic boundaries are supported on both x faces and/or both y faces only\n");
        return NULL;
    }
    if ((set->sb.bx0 == SV_BOUNDARY_PERIODIC || set->sb.by0 == SV_BOUNDARY_PERIODIC)
        && (set->smb.bx0 || set->smb.bxn || set->smb.by0 || set->smb.byn || set->smb.bz0 || set->smb.bzn)) {
        fprintf(stderr, "Error: periodic pool boundaries cannot be combined with MBOUNDARY planes\n");
        return NULL;
    }
    if (load_sources(mp, set))
        return NULL;

    /* OUT_SUMTAB filters: the database material at the centre wavelength, as
     * absolute permittivity and permeability (src3d/plan.c:5475-5486) */
    for (n = 0; n < set->so.nsums; n++) {
        SvOutputSum *o = set->so.sums + n;
        if (!o->stringbased)
            continue;
        if (sv_database_linear(set, o->string, &o->layered_epsilon, &o->layered_mu, &o->layered_sigma, &o->layered_sigast))
            return NULL;
        o->layered_epsilon *= EPS0;
        o->layered_mu *= MU0;
        o->stringbased = 0;
    }

    /* matmode, src3d/plan.c:5489-5518 */
    if (set->sm.matmode_check) {
        matmode_mat = SV_MATMODE_NONE;
        matmode_script = SV_MATMODE_NONE;
        if (set->sm.in_voxel_filename[0])
            matmode_mat = test_material_file(set->sm.in_voxel_filename);
        if (set->sm.in_vector_filename[0])
            matmode_script = test_script_file(set, set->sm.in_vector_filename);
        if (matmode_script == SV_MATMODE_ERROR)
            return NULL;
        if (matmode_mat == SV_MATMODE_NONE && matmode_script == SV_MATMODE_NONE)
            set->plan.matmode = SV_MATMODE_NONE;
        else if (matmode_mat == SV_MATMODE_ELECTRIC && (matmode_script == SV_MATMODE_NONE || matmode_script == SV_MATMODE_ELECTRIC))
            set->plan.matmode = SV_MATMODE_ELECTRIC;
        else if (matmode_mat == SV_MATMODE_MAGNETIC && (matmode_script == SV_MATMODE_NONE || matmode_script == SV_MATMODE_MAGNETIC))
            set->plan.matmode = SV_MATMODE_MAGNETIC;
        else if (matmode_script == SV_MATMODE_ELECTRIC && (matmode_mat == SV_MATMODE_NONE || matmode_mat == SV_MATMODE_ELECTRIC))
            set->plan.matmode = SV_MATMODE_ELECTRIC;
        else if (matmode_script == SV_MATMODE_MAGNETIC && (matmode_mat == SV_MATMODE_NONE || matmode_mat == SV_MATMODE_MAGNETIC))
            set->plan.matmode = SV_MATMODE_MAGNETIC;
        else
            set->plan.matmode = SV_MATMODE_FULL;
    } else {
        if (set->sm.in_vector_filename[0] && test_script_file(set, set->sm.in_vector_filename) == SV_MATMODE_ERROR)
            return NULL;
        set->plan.matmode = SV_MATMODE_FULL;
    }

    /* field storage, src3d/yee.c:101-141 */
    d->xres = xres; d->yres = yres; d->zres = zres;
    d->dx = set->sp.dx; d->dy = set->sp.dy; d->dz = set->sp.dz;
    d->ex = sv_dcube_new(xres, yres, zres);
    d->ey = sv_dcube_new(xres, yres, zres);
    d->ez = sv_dcube_new(xres, yres, zres);
    d->hx = sv_dcube_new(xres, yres, zres);
    d->hy = sv_dcube_new(xres, yres, zres);
    d->hz = sv_dcube_new(xres, yres, zres);
    d->bnds = sv_icube_new(xres, yres, zres);
    if (!d->ex || !d->ey || !d->ez || !d->hx || !d->hy || !d->hz || !d->bnds) {
        fprintf(stderr, "Error: cannot allocate host field storage\n");
        return NULL;
    }
    if (set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_ELECTRIC) {
        d->epsilon = sv_fcube_new(xres, yres, zres);
        d->sigma = sv_fcube_new(xres, yres, zres);
        sv_fcube_fill(d->epsilon, (float)EPS0);
    }
    if (set->plan.matmode == SV_MATMODE_FULL || set->plan.matmode == SV_MATMODE_MAGNETIC) {
        d->mu = sv_fcube_new(xres, yres, zres);
        d->sigast = sv_fcube_new(xres, yres, zres);
        sv_fcube_fill(d->mu, (float)MU0);
    }

    mp->bnd = boundary_new(set);
    mp->farfield = farfield_new(set);
    mp->pfarfield = pfarfield_new(set);

    if (set->sm.in_voxel_filename[0] || set->sm.in_vector_filename[0]) {
        if (load_data(mp, set))
            return NULL;
    }

    /* outputs, src3d/output.c (sv_pool_allocate_output) */
    for (n = 0; n < set->so.nplns; n++) {
        const SvOutputRegion *o = set->so.plns + n;
        const int fixed = (o->i >= 0) + (o->j >= 0) + (o->k >= 0);
        if (fixed != 1 || o->i >= xres || o->j >= yres || o->k >= zres) {
            fprintf(stderr, "Error: output plane %d needs one index inside the pool and -1 for the other two\n", n);
            return NULL;
        }
    }
    mp->out->nabs = 0;
    for (n = 0; n < set->so.ncubs; n++)
        if (set->so.cubs[n].component >= SV_OVOLUME_ABS)
            mp->out->nabs++;
    if (mp->out->nabs) {
        mp->out->outabsdata = (SvDCube **)calloc(mp->out->nabs, sizeof(SvDCube *));
        for (n = 0; n < mp->out->nabs; n++)
            mp->out->outabsdata[n] = sv_dcube_new(xres, yres, zres);
    }
    if (set->so.npnts) {
        mp->out->outpointdata = (double **)calloc(set->so.npnts, sizeof(double *));
        for (n = 0; n < set->so.npnts; n++) {
            SvOutputPar *o = set->so.pnts + n;
            if (o->i < 0 || o->j < 0 || o->k < 0 || o->i >= xres || o->j >= yres || o->k >= zres) {
                fprintf(stderr, "Error: output point %d at (%d %d %d) is outside the pool\n", n, o->i, o->j, o->k);
                return NULL;
            }
            mp->out->outpointdata[n] = (double *)calloc(6 * (size_t)(set->sc.nsteps + 1), sizeof(double));
        }
    }

    if (set->so.nsums) {
        mp->out->outsumdata = (double **)calloc(set->so.nsums, sizeof(double *));
        for (n = 0; n < set->so.nsums; n++) {
            SvOutputSum *o = set->so.sums + n;
            if (o->box_i0 < 0 || o->box_j0 < 0 || o->box_k0 < 0 || o->box_in > xres || o->box_jn > yres
                || o->box_kn > zres) {
                fprintf(stderr, "Error: sum box %d (%d %d %d - %d %d %d) is outside the pool\n", n, o->box_i0,
                        o->box_j0, o->box_k0, o->box_in, o->box_jn, o->box_kn);
                return NULL;
            }
            mp->out->outsumdata[n] = (double *)calloc((size_t)set->sc.nsteps + 1, sizeof(double));
        }
    }

    if (set->so.nforces) {
        mp->out->outforcedata = (double **)calloc(set->so.nforces, sizeof(double *));
        for (n = 0; n < set->so.nforces; n++) {
            SvOutputForce *o = set->so.forces + n;
            if (o->box_i0 < 0 || o->box_j0 < 0 || o->box_k0 < 0 || o->box_in >= xres || o->box_jn >= yres
                || o->box_kn >= zres || o->box_in <= o->box_i0 || o->box_jn <= o->box_j0 || o->box_kn <= o->box_k0) {
                fprintf(stderr, "Error: force box %d (%d %d %d - %d %d %d) must lie inside the pool\n", n, o->box_i0,
                        o->box_j0, o->box_k0, o->box_in, o->box_jn, o->box_kn);
                return NULL;
            }
            mp->out->outforcedata[n] = (double *)calloc(3 * (size_t)(set->sc.nsteps + 1), sizeof(double));
        }
    }

    if (set->sf.nrs || set->sf.ndft) {
        SvSetFarfield *f = &set->sf;
        if (f->box_i0 < 1 || f->box_j0 < 1 || f->box_k0 < 1 || f->box_in >= xres - 1 || f->box_jn >= yres - 1
            || f->box_kn >= zres - 1 || f->box_in <= f->box_i0 || f->box_jn <= f->box_j0 || f->box_kn <= f->box_k0) {
            fprintf(stderr, "Error: NFFF box must lie inside the pool with one cell to spare on each side\n");
            return NULL;
        }
    }

    if (set->spf.nrs) {
        /* the plane integrals read E at k0 +- 1 and kn +- 1 */
        SvSetPFarfield *f = &set->spf;
        if (f->box_i0 < 0 || f->box_j0 < 0 || f->box_k0 < 1 || f->box_in > xres || f->box_jn > yres
            || f->box_kn > zres - 2 || f->box_in <= f->box_i0 || f->box_jn <= f->box_j0 || f->box_kn <= f->box_k0) {
            fprintf(stderr, "Error: PERIODIC_NFFF box must lie inside the pool, its z planes one cell inside\n");
            return NULL;
        }
        if (f->postprocess && (f->ppstart < 0 || f->ppstart >= set->sc.nsteps)) {
            fprintf(stderr, "Error: PERIODIC_NFFF_POSTPROCESS start step must lie in 0 .. nsteps-1\n");
            return NULL;
        }
    }

    if (set->sx.nplanes) {
        const SvSetFlux *x = &set->sx;
        int q;
        if (x->i0 < 0 || x->j0 < 0 || x->i1 > xres || x->j1 > yres || x->i1 <= x->i0 || x->j1 <= x->j0) {
            fprintf(stderr, "Error: FLUX_DFT rectangle must lie inside the pool\n");
            return NULL;
        }
        for (q = 0; q < x->nplanes; q++)
            if (x->k[q] < 1 || x->k[q] > zres - 1) {
                fprintf(stderr, "Error: FLUX_DFT plane %d outside 1 .. %d\n", x->k[q], zres - 1);
                return NULL;
            }
    }

    /* gradient extension: the design voxels carry per-voxel permittivity
     * (MEDIUM_LINEAR data) and the coefficients they feed must stay off the
     * zeroed boundary planes */
    if (set->sa.design) {
        SvSetAdjoint *a = &set->sa;
        int i, j, k;
        if (a->design_i0 < 1 || a->design_j0 < 1 || a->design_k0 < 1 || a->design_in > xres - 2
            || a->design_jn > yres - 2 || a->design_kn > zres - 2 || a->design_in <= a->design_i0
            || a->design_jn <= a->design_j0 || a->design_kn <= a->design_k0) {
            fprintf(stderr, "Error: DESIGN_BOX must lie inside the pool with two cells to spare on each side\n");
            return NULL;
        }
        if (!mp->d->epsilon) {
            fprintf(stderr, "Error: DESIGN_BOX needs per-voxel permittivity (MEDIUM_LINEAR)\n");
            return NULL;
        }
        for (i = a->design_i0; i < a->design_in; i++)
            for (j = a->design_j0; j < a->design_jn; j++)
                for (k = a->design_k0; k < a->design_kn; k++)
                    if (mp->nmat > 0 && mp->d->mat->data[i][j][k] > 0) {
                        fprintf(stderr, "Error: DESIGN_BOX voxel (%d %d %d) holds a tabulated material\n", i, j, k);
                        return NULL;
                    }
    }
    if (set->sa.objective) {
        SvSetAdjoint *a = &set->sa;
        if (a->obj_i0 < 0 || a->obj_j0 < 0 || a->obj_k0 < 0 || a->obj_in > xres || a->obj_jn > yres || a->obj_kn > zres
            || a->obj_in <= a->obj_i0 || a->obj_jn <= a->obj_j0 || a->obj_kn <= a->obj_k0) {
            fprintf(stderr, "Error: OBJECTIVE_DFT box must lie inside the pool\n");
            return NULL;
        }
    }
    return mp;
}

void
sv_pool_free(SvPool *mp, SvSet *set)
{
    int n;
    SvYeeData *d;
    (void)set;
    if (!mp)
        return;
    d = mp->d;
    if (d) {
        sv_dcube_free(d->ex); sv_dcube_free(d->ey); sv_dcube_free(d->ez);
        sv_dcube_free(d->hx); sv_dcube_free(d->hy); sv_dcube_free(d->hz);
        sv_fcube_free(d->epsilon); sv_fcube_free(d->sigma);
        sv_fcube_free(d->mu); sv_fcube_free(d->sigast);
        sv_icube_free(d->mat); sv_icube_free(d->bnds); sv_icube_free(d->sourceskip);
        free(d);
    }
    if (mp->src) {
        for (n = 0; n < mp->src->npnts; n++) {
            free(mp->src->sp[n].layered_zpos);
            free(mp->src->sp[n].ex);
        }
        free(mp->src->sp);
        if (mp->src->ltsf) {
            free(mp->src->ltsf->e);
            free(mp->src->ltsf->zeps);
            free(mp->src->ltsf->zmu);
            free(mp->src->ltsf);
        }
        if (mp->src->tsf) {
            free(mp->src->tsf->e);
            free(mp->src->tsf);
        }
        if (mp->src->tsff) {
            if (mp->src->tsff->nw)
                free(mp->src->tsff->w[0].e);
            free(mp->src->tsff->w);
            free(mp->src->tsff->acorr);
            free(mp->src->tsff);
        }
        if (mp->src->sf) {
            free(mp->src->sf->e);
            free(mp->src->sf);
        }
        free(mp->src);
    }
    if (mp->out) {
        if (mp->out->outpointdata) {
            for (n = 0; n < mp->set->so.npnts; n++)
                free(mp->out->outpointdata[n]);
            free(mp->out->outpointdata);
        }
        if (mp->out->outforcedata) {
            for (n = 0; n < mp->set->so.nforces; n++)
                free(mp->out->outforcedata[n]);
            free(mp->out->outforcedata);
        }
        if (mp->out->outsumdata) {
            for (n = 0; n < mp->set->so.nsums; n++)
                free(mp->out->outsumdata[n]);
            free(mp->out->outsumdata);
        }
        for (n = 0; n < mp->out->nabs; n++)
            sv_dcube_free(mp->out->outabsdata[n]);
        free(mp->out->outabsdata);
        free(mp->out);
    }
    if (mp->bnd) {
        free(mp->bnd->x0.be); free(mp->bnd->xn.be); free(mp->bnd->y0.be);
        free(mp->bnd->yn.be); free(mp->bnd->z0.be); free(mp->bnd->zn.be);
        free(mp->bnd);
    }
    if (mp->farfield) {
        for (n = 0; n < mp->farfield->nrs; n++)
            free(mp->farfield->rpoints[n].ex);
        free(mp->farfield->rpoints);
        free(mp->farfield->dft_omega);
        free(mp->farfield->dft_data);
        free(mp->farfield->flux_omega);
        free(mp->farfield->flux_data);
        free(mp->farfield);
    }
    if (mp->pfarfield) {
        for (n = 0; n < mp->pfarfield->nrs; n++)
            free(mp->pfarfield->prpoints[n].ex);
        free(mp->pfarfield->prpoints);
        free(mp->pfarfield);
    }
    free(mp->mats);
    free(mp);
}
