/*
 * containers.c -- 3-D host arrays with the reference's data[i][j][k] accessor
 * (src3d/dcube.c:31-53, src3d/fcube.c, src3d/icube.c:28-48), backed by one
 * contiguous, zero-initialised block (k fastest) instead of one malloc per row.
 */
#include <stdlib.h>
#include <string.h>
#include "gsvit_pool.h"

#define SV_CUBE_NEW(TYPE, ELEM)                                                     \
    TYPE *c = (TYPE *)calloc(1, sizeof(TYPE));                                      \
    long i, j;                                                                      \
    if (!c)                                                                         \
        return NULL;                                                                \
    c->xres = xres;                                                                 \
    c->yres = yres;                                                                 \
    c->zres = zres;                                                                 \
    c->block = (ELEM *)calloc((size_t)xres * yres * zres, sizeof(ELEM));            \
    c->data = (ELEM ***)malloc((size_t)xres * sizeof(ELEM **));                     \
    if (!c->block || !c->data) {                                                    \
        free(c->block);                                                             \
        free(c->data);                                                              \
        free(c);                                                                    \
        return NULL;                                                                \
    }                                                                               \
    c->data[0] = (ELEM **)malloc((size_t)xres * yres * sizeof(ELEM *));             \
    if (!c->data[0]) {                                                              \
        free(c->block);                                                             \
        free(c->data);                                                              \
        free(c);                                                                    \
        return NULL;                                                                \
    }                                                                               \
    for (i = 0; i < xres; i++) {                                                    \
        c->data[i] = c->data[0] + i * yres;                                         \
        for (j = 0; j < yres; j++)                                                  \
            c->data[i][j] = c->block + ((size_t)i * yres + j) * zres;               \
    }                                                                               \
    return c;

SvDCube *sv_dcube_new(int xres, int yres, int zres) { SV_CUBE_NEW(SvDCube, double) }
SvFCube *sv_fcube_new(int xres, int yres, int zres) { SV_CUBE_NEW(SvFCube, float) }
SvICube *sv_icube_new(int xres, int yres, int zres) { SV_CUBE_NEW(SvICube, int) }

#define SV_CUBE_FREE                 \
    if (!c)                          \
        return;                      \
    if (c->data)                     \
        free(c->data[0]);            \
    free(c->data);                   \
    free(c->block);                  \
    free(c);

void sv_dcube_free(SvDCube *c) { SV_CUBE_FREE }
void sv_fcube_free(SvFCube *c) { SV_CUBE_FREE }
void sv_icube_free(SvICube *c) { SV_CUBE_FREE }

void
sv_fcube_fill(SvFCube *c, float v)
{
    size_t n = (size_t)c->xres * c->yres * c->zres, q;
    for (q = 0; q < n; q++)
        c->block[q] = v;
}

void
sv_icube_fill(SvICube *c, int v)
{
    size_t n = (size_t)c->xres * c->yres * c->zres, q;
    if (v == 0) {
        memset(c->block, 0, n * sizeof(int));
        return;
    }
    for (q = 0; q < n; q++)
        c->block[q] = v;
}
