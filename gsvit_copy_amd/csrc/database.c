/*
 * database.c -- optical-database materials (material type 99) in MEDIUM_VECTOR
 * files.
 *
 * The reference resolves them before planning (src3d/plan.c:5463-5468): when
 * a vector file names any database material, the whole file is rewritten,
 * entity by entity, into a temporary material file, every number passing
 * through float and "%g", and each database entry replaced by the material it
 * stands for at the source's centre wavelength (override_matfile_from_database
 * and adjust_material, src3d/plan.c:719-1312).  The planner then reads the
 * rewritten file as an ordinary vector file.  Here the rewritten text goes to
 * an anonymous temporary file (tmpfile) instead of tmp_matfile_<suffix>; its
 * content is the reference's.
 *
 * Database entry file (found as given, relative to the working directory, or
 * under the spectra directory, see open_database_entry):
 *   line 1: <type> <from> <to>    wavelength range in micrometres
 *   line 2: description
 *   then the parameters of the type:
 *     drude      eps_inf omega_p nu                       -> type 2
 *     cp         eps_inf omega_p nu (a phi omega gamma) x2 -> type 4
 *     ade, plrc  eps_inf sigma omega_p nu (a phi omega gamma) x2 -> type 5, 6
 *     cp3        eps_inf sigma (a phi omega gamma) x3      -> type 3
 *     sellmeier  B0 B1 C1 B2 C2 B3 C3:  n^2 = B0 + sum Bi w^2 / (w^2 - Ci)
 *     cauchy     A B C:                 n = A + B / w^2 + C / w^4
 *     kasarova   A0..A5:  n^2 = A0 + A1 w^2 + A2 / w^2 + A3 / w^4 + A4 / w^6 + A5 / w^8
 *                (w the centre wavelength in micrometres; these three give a
 *                tabulated linear material eps = n^2, mu = 1)
 *     nk         count, then rows "wavelength n k" (micrometres or nanometres,
 *                ascending or descending, told from the first row matching
 *                the header range); n and k interpolated linearly at the centre
 *                wavelength give eps = n^2 - k^2, sigma = 4 pi n k c eps0 / lambda
 *                (src3d/plan.c:1028-1125).
 * Material types 2, 3, 4 and 6 are rewritten as the reference does (a PLRC
 * vector-file line comes back as type 5) and then planned like the same
 * types written directly: types 2 and 4 as recursive convolution, type 3
 * refused.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "gsvit_internal.h"

#define DB_TYPE 99

/* the database entry's file: as named (relative to the working directory),
 * else <dir>/<name> for dir in $GSVIT_SPECTRA_DIR and the reference's install
 * locations of its spectra (src3d/global.c:92-104, <prefix>/share/gsvit/data/spectra) */
static FILE *
open_database_entry(const char *name, char *path, size_t npath)
{
    static const char *const dirs[] = { "/usr/local/share/gsvit/data/spectra", "/usr/share/gsvit/data/spectra" };
    const char *env = getenv("GSVIT_SPECTRA_DIR");
    FILE *fs = fopen(name, "r");
    size_t q;
    snprintf(path, npath, "%s", name);
    if (fs)
        return fs;
    if (env && env[0]) {
        snprintf(path, npath, "%s/%s", env, name);
        if ((fs = fopen(path, "r")))
            return fs;
    }
    for (q = 0; q < sizeof(dirs) / sizeof(dirs[0]); q++) {
        snprintf(path, npath, "%s/%s", dirs[q], name);
        if ((fs = fopen(path, "r")))
            return fs;
    }
    return NULL;
}

/* copies n floats from fr to fw as "%g " (the last followed by "\n" when nl) */
static int
copy_floats(FILE *fr, FILE *fw, int n, int nl)
{
    float v;
    int q;
    for (q = 0; q < n; q++) {
        if (fscanf(fr, "%f", &v) != 1)
            return 1;
        fprintf(fw, (q == n - 1 && nl) ? "%g\n" : "%g ", v);
    }
    return 0;
}

/* nk table at the centre wavelength wc (metres), src3d/plan.c:1028-1125;
 * returns 1 on an error (message printed) */
static int
nk_at(FILE *fs, float swfrom, float swto, double wc, double *n, double *k)
{
    int ndata, i, dir = 0, found = 0;
    double mult = 1e-6, wprev = 0, wnext = 0, nprev = 0, nnext = 0, kprev = 0, knext = 0;
    double *w, *nv, *kv;
    if (fscanf(fs, "%d", &ndata) != 1 || ndata < 2) {
        fprintf(stderr, "Error: database nk table needs at least two rows\n");
        return 1;
    }
    w = (double *)malloc(3 * (size_t)ndata * sizeof(double));
    nv = w + ndata;
    kv = nv + ndata;
    for (i = 0; i < ndata; i++) {
        float wf, nf, kf;
        if (fscanf(fs, "%g %g %g", &wf, &nf, &kf) != 3) {
            fprintf(stderr, "Error: database nk table: cannot read row %d\n", i);
            free(w);
            return 1;
        }
        if (i == 0) {
            /* the first row is the header's lower end (ascending table) or upper end
             * (descending), in micrometres or nanometres */
            if (fabs(wf * 1e-6 - swfrom) < 1e-10) {
                dir = 1;
                mult = 1e-6;
            } else if (fabs(wf * 1e-6 - swto) < 1e-10) {
                dir = 0;
                mult = 1e-6;
            } else if (fabs(wf * 1e-9 - swfrom) < 1e-10) {
                dir = 1;
                mult = 1e-9;
            } else if (fabs(wf * 1e-9 - swto) < 1e-10) {
                dir = 0;
                mult = 1e-9;
            } else {
                fprintf(stderr, "Database entry error: wavelength range (%g %g) does not match the spectrum value %g\n",
                        swfrom, swto, wf);
                free(w);
                return 1;
            }
        }
        w[i] = wf * mult;
        nv[i] = nf;
        kv[i] = kf;
    }
    if (dir == 1 && wc < w[0]) {
        wprev = w[0]; nprev = nv[0]; kprev = kv[0];
        wnext = w[1]; nnext = nv[1]; knext = kv[1];
        found = 1;
        fprintf(stderr, "Warning: Spectral data extrapolated at lower side of spectrum\n");
    } else if (dir == 0 && wc < w[ndata - 1]) {
        wprev = w[ndata - 1]; nprev = nv[ndata - 1]; kprev = kv[ndata - 1];
        wnext = w[ndata - 2]; nnext = nv[ndata - 2]; knext = kv[ndata - 2];
        found = 1;
        fprintf(stderr, "Warning: Spectral data extrapolated at lower side of spectrum\n");
    } else {
        for (i = 0; i < ndata - 1 && !found; i++) {
            if (dir == 1 && w[i] <= wc && w[i + 1] >= wc) {
                wprev = w[i]; nprev = nv[i]; kprev = kv[i];
                wnext = w[i + 1]; nnext = nv[i + 1]; knext = kv[i + 1];
                found = 1;
            } else if (dir == 0 && w[i] >= wc && w[i + 1] <= wc) {
                wprev = w[i + 1]; nprev = nv[i + 1]; kprev = kv[i + 1];
                wnext = w[i]; nnext = nv[i]; knext = kv[i];
                found = 1;
            }
        }
    }
    free(w);
    if (!found) {
        /* the reference interpolates between zeros here (0/0) */
        fprintf(stderr, "Error: centre wavelength %g m lies above the database nk table\n", wc);
        return 1;
    }
    *n = nprev + (nnext - nprev) * (wc - wprev) / (wnext - wprev);
    *k = kprev + (knext - kprev) * (wc - wprev) / (wnext - wprev);
    return 0;
}

/* the database entry `name` written as a material (adjust_material's type-99
 * branch, src3d/plan.c:883-1133) */
static int
write_database_material(const SvSet *set, const char *name, FILE *fw)
{
    char path[2 * SV_NAME_LEN], stype[100], desc[256];
    float swfrom, swto, p[7];
    double wav, n, k, eps;
    const double wfrom = set->ss.lambda_min, wc = set->ss.lambda_center, wto = set->ss.lambda_max;
    int rc = 1;
    FILE *fs;
    if (!(wc > 0)) {
        fprintf(stderr, "Error: database material %s needs the source wavelength (SOURCE_WAVELENGTH min centre max); "
                        "the automatic estimate from the source spectrum is not built\n", name);
        return 1;
    }
    if (!(fs = open_database_entry(name, path, sizeof(path)))) {
        printf("no database file for this material (%s)!\n", name);
        fprintf(stderr, "Error: cannot open data from optical database: %s\n", name);
        return 1;
    }
    if (fscanf(fs, "%99s %f %f", stype, &swfrom, &swto) != 3 || !fgets(desc, sizeof(desc), fs)
        || !fgets(desc, sizeof(desc), fs)) {
        fprintf(stderr, "Error: cannot read the header of database entry %s\n", path);
        goto out;
    }
    desc[strcspn(desc, "\r\n")] = 0;
    swfrom *= (float)1.0e-6;            /* the header range is in micrometres */
    swto *= (float)1.0e-6;
    if (swfrom > wfrom || swto < wto)
        fprintf(stderr, "Warning: your source wavelength (%g to %g, center at %g) does not match database spectrum "
                        "range (%g, %g)\n", wfrom, wc, wto, swfrom, swto);
    wav = wc * 1e6;                     /* micrometres */
    if (!strcmp(stype, "drude")) {
        fprintf(fw, "%d ", SV_MAT_DRUDE);
        rc = copy_floats(fs, fw, 3, 1);
    } else if (!strcmp(stype, "cp")) {
        fprintf(fw, "%d ", SV_MAT_CP);
        rc = copy_floats(fs, fw, 11, 1);
    } else if (!strcmp(stype, "ade")) {
        fprintf(fw, "%d ", SV_MAT_ADE);
        rc = copy_floats(fs, fw, 12, 1);
    } else if (!strcmp(stype, "plrc")) {
        fprintf(fw, "%d ", SV_MAT_PLRC);
        rc = copy_floats(fs, fw, 12, 1);
    } else if (!strcmp(stype, "cp3")) {
        fprintf(fw, "%d ", SV_MAT_CP3);
        rc = copy_floats(fs, fw, 14, 1);
    } else if (!strcmp(stype, "sellmeier") || !strcmp(stype, "cauchy") || !strcmp(stype, "kasarova")) {
        const int np = !strcmp(stype, "sellmeier") ? 7 : !strcmp(stype, "cauchy") ? 3 : 6;
        int q;
        for (q = 0; q < np; q++)
            if (fscanf(fs, "%g", p + q) != 1) {
                fprintf(stderr, "Error: database entry %s: cannot read %s parameter %d\n", path, stype, q + 1);
                goto out;
            }
        if (np == 7)
            n = sqrt((double)(p[0] + p[1] * wav * wav / (wav * wav - p[2]) + p[3] * wav * wav / (wav * wav - p[4])
                              + p[5] * wav * wav / (wav * wav - p[6])));
        else if (np == 3)
            n = p[0] + p[1] / wav / wav + p[2] / wav / wav / wav / wav;
        else
            n = sqrt((double)(p[0] + p[1] * wav * wav + p[2] / wav / wav + p[3] / wav / wav / wav / wav
                              + p[4] / wav / wav / wav / wav / wav / wav
                              + p[5] / wav / wav / wav / wav / wav / wav / wav / wav));
        fprintf(fw, "%d %g 1 0 0\n", SV_MAT_LINTAB, n * n);
        rc = 0;
    } else if (!strcmp(stype, "nk")) {
        if (nk_at(fs, swfrom, swto, wc, &n, &k))
            goto out;
        eps = n * n - k * k;
        fprintf(fw, "%d %g 1 %g 0\n", SV_MAT_LINTAB, eps, 4 * M_PI * n * k * SV_LIGHT_SPEED / wc * SV_EPSILON_0);
        rc = 0;
    } else {
        fprintf(stderr, "Error: wrong spectrum type (%s)\n", stype);
        goto out;
    }
    if (rc)
        fprintf(stderr, "Error: database entry %s: cannot read its %s parameters\n", path, stype);
out:
    fclose(fs);
    return rc;
}

/* one material of the vector file copied to fw as the reference rewrites it
 * (adjust_material, src3d/plan.c:719-1140); *db set for a database entry */
static int
rewrite_material(const SvSet *set, FILE *fr, FILE *fw, int *db)
{
    int type;
    char name[SV_NAME_LEN];
    if (fscanf(fr, "%d", &type) != 1)
        return 1;
    switch (type) {
    case SV_MAT_LINEAR:
    case SV_MAT_LINTAB:
        fprintf(fw, "%d ", type);
        return copy_floats(fr, fw, 4, 1);
    case SV_MAT_DRUDE:
        fprintf(fw, "%d ", type);
        return copy_floats(fr, fw, 3, 1);
    case SV_MAT_CP:
        fprintf(fw, "%d ", type);
        return copy_floats(fr, fw, 11, 1);
    case SV_MAT_ADE:
    case SV_MAT_PLRC:
        fprintf(fw, "%d ", SV_MAT_ADE);   /* the reference writes a PLRC entry back as type 5 */
        return copy_floats(fr, fw, 12, 1);
    case SV_MAT_ADE3:                     /* no reference counterpart: kept, rounded like type 5 */
        fprintf(fw, "%d ", type);
        return copy_floats(fr, fw, 16, 1);
    case SV_MAT_PEC:
        fprintf(fw, "%d\n", type);
        return 0;
    case SV_MAT_CP3:
        fprintf(fw, "%d ", type);
        return copy_floats(fr, fw, 14, 1);
    case DB_TYPE:
        if (fscanf(fr, "%511s", name) != 1)
            return 1;
        *db = 1;
        return write_database_material(set, name, fw);
    default:
        fprintf(stderr, "Error: adj: Unknown material type?\n");
        return 1;
    }
}

/* the entity lines of the vector file rewritten into fw (override_matfile_from_database,
 * src3d/plan.c:1142-1312); *db set when some material is a database entry */
static int
rewrite_vector_file(const SvSet *set, FILE *fr, FILE *fw, int *db)
{
    int type, q;
    char base[SV_NAME_LEN];
    float v;
    while (fscanf(fr, "%d", &type) == 1) {
        int npts;
        fprintf(fw, "%d ", type);
        switch (type) {
        case SV_ENTITY_SPHERE: npts = 4; break;
        case SV_ENTITY_CYLINDER: case SV_ENTITY_CONE: npts = 7; break;
        case SV_ENTITY_RCONE: npts = 8; break;
        case SV_ENTITY_VOXEL: npts = 6; break;
        case SV_ENTITY_TETRAHEDRON: npts = 12; break;
        case SV_ENTITY_TETGEN:
            /* file base, two integers read as floats and written truncated, two points */
            if (fscanf(fr, "%511s", base) != 1)
                return 1;
            fprintf(fw, "%s ", base);
            for (q = 0; q < 2; q++) {
                if (fscanf(fr, "%f", &v) != 1)
                    return 1;
                fprintf(fw, "%d ", (int)v);
            }
            npts = 6;
            break;
        default:
            return 1;    /* an entity the planner does not take: it reports the original file */
        }
        if (copy_floats(fr, fw, npts, 0) || rewrite_material(set, fr, fw, db))
            return 1;
    }
    return ferror(fr) != 0;
}

FILE *
sv_open_vector_file(const SvSet *set, const char *fname)
{
    FILE *fr = fopen(fname, "r"), *fw;
    int db = 0, rc;
    if (!fr)
        return NULL;
    if (!(fw = tmpfile())) {
        fprintf(stderr, "Error: cannot create a temporary material file\n");
        fclose(fr);
        return NULL;
    }
    rc = rewrite_vector_file(set, fr, fw, &db);
    fclose(fr);
    if (db && rc) {
        fprintf(stderr, "Error: material vector file %s: cannot resolve its database materials\n", fname);
        fclose(fw);
        return NULL;
    }
    if (!db) {
        /* no database entries: the reference reads the file itself */
        fclose(fw);
        return fopen(fname, "r");
    }
    rewind(fw);
    return fw;
}

int
sv_database_linear(const SvSet *set, const char *name, double *eps, double *mu, double *sigma, double *sigast)
{
    FILE *fw = tmpfile();
    int type = -1, rc = 1;
    if (!fw) {
        fprintf(stderr, "Error: cannot create a temporary material file\n");
        return 1;
    }
    if (!write_database_material(set, name, fw)) {
        rewind(fw);
        if (fscanf(fw, "%d", &type) == 1 && type == SV_MAT_LINTAB && fscanf(fw, "%lf %lf %lf %lf", eps, mu, sigma, sigast) == 4)
            rc = 0;
        else
            fprintf(stderr, "Error: database entry %s is not a linear (non-dispersive) material\n", name);
    }
    fclose(fw);
    return rc;
}
