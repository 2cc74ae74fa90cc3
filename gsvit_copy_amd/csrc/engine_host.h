/*
 * engine_host.h -- host-side state of one MI355X engine (one pool on one
 * GPU), shared by the drop-in entry points (gpu_engine.hip) and the adjoint
 * driver (adjoint.hip).  Not part of the C-ABI.
 */
#ifndef GSV_ENGINE_HOST_H
#define GSV_ENGINE_HOST_H

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <vector>
#include "gsvit_internal.h"
#include "engine.h"
#include "kernels.h"

#define CK(x)                                                                                 \
    do {                                                                                      \
        hipError_t e_ = (x);                                                                  \
        if (e_ != hipSuccess) {                                                               \
            fprintf(stderr, "HIP error: %s (%s) at %s:%d\n", hipGetErrorString(e_), #x, __FILE__, __LINE__); \
            return 1;                                                                         \
        }                                                                                     \
    } while (0)


struct GsvAdjoint;   /* adjoint.hip */
struct Engine;
int gsv_build_pctab(Engine *E);
void gsv_adjoint_free(Engine *E);
int gsv_adjoint_pre_e(Engine *E, int step);
int gsv_adjoint_post_e(Engine *E, int step);
int gsv_adjoint_after_step(Engine *E, int step);
int gsv_engine_dft_flush(Engine *E);   /* gpu_engine.hip: fold pending running-DFT samples */
extern "C" int gsv_engine_farfield_dft(void *eng, int ndirs, const double *dirs, double *out);   /* nfff.hip */
extern "C" void *gsv_session_engine(void *session);   /* session.c: the engine of a session */

struct Engine {
    SvPool *mp;
    SvSet *set;
    int dev;
    hipStream_t st;
    int ready;
    GsvGrid g;
    int idbytes_e, idbytes_h;
    int bx_h, by_h, kchunk_h, bx_e, by_e, kchunk_e;
    std::vector<GsvECoef> etab;
    std::vector<GsvHCoef> htab;
    std::vector<uint8_t> ecls_host, hcls_host;   /* device layout, idbytes per cell */
    std::vector<GsvECoef> pctab_host;            /* per-cell coefficients of the design region */
    GsvAdjoint *adj;                             /* gradient state, adjoint.hip (NULL until used) */
    std::vector<float> src_vals_host;            /* point-source table [nsteps][n][6] */
    int ade_slot_of_mat[256];
    GsvAde adem_host[GSV_MAX_ADE];
    GsvRc rcm_host[GSV_MAX_ADE];
    std::vector<void *> allocs;
    float *fbase[6];
    size_t field_floats;
    /* sources */
    int has_tsf;
    GsvTsf tsf;
    int tsf_face_on[6];       /* face has at least one nonzero correction coefficient */
    int has_src;
    GsvPointSrc src;
    /* scattered-field source: its 1-D line and the PEC-rule cells it drives
     * (i, j, k, component mask 1 Ex / 2 Ey / 4 Ez) */
    int has_sf;
    GsvTsf sfl;
    int4 *sf_cells;
    int n_sf;
    /* traces */
    int npts;
    long long *d_ppos;
    float *d_rec;
    /* boundaries */
    int has_liao;
    GsvLiao liao;
    int rezero[3];
    int has_per;
    GsvPeriodic per;
    /* far field */
    int has_ff;
    GsvFarfieldTD ff;
    int has_ext;                  /* external plane source: window buffers, host staging */
    GsvExt ext;
    float *ext_px, *ext_py;
    std::vector<float> ext_planes, ext_win;
    int has_tsff;                 /* focused TSF: per-wave lines, wave table, E face tables */
    GsvTsff tsff;
    int has_ltsf;                 /* layered TSF: 1-D model history and face tables */
    GsvLtsf ltsf;
    int has_pff;                  /* periodic far field */
    GsvPFarfield pff;
    int has_dft;
    GsvDft dft;
    int dft_first, dft_pending;   /* pending ring samples: steps dft_first .. + dft_pending - 1 */
    int has_flux;                 /* FLUX_DFT planes: the same ring scheme, face code 6 */
    GsvDft flx;
    int flx_first, flx_pending;
    /* box sums */
    std::vector<GsvBoxSum> sums;
    std::vector<GsvForce> forces;
    /* timing of the two volume kernels */
    int timing;
    hipEvent_t ev[4];
    double t_h_ms, t_e_ms;
    long n_h, n_e;
};

template <typename T>
static int
dmalloc(Engine *E, T **p, size_t count)
{
    void *q = NULL;
    if (count == 0)
        count = 1;
    CK(hipMalloc(&q, count * sizeof(T)));
    E->allocs.push_back(q);
    *p = (T *)q;
    return 0;
}

template <typename T>
static int
dupload(Engine *E, T **p, const T *h, size_t count)
{
    if (dmalloc(E, p, count))
        return 1;
    /* synchronous: the host staging vectors of the callers are temporaries */
    CK(hipMemcpyAsync(*p, h, count * sizeof(T), hipMemcpyHostToDevice, E->st));
    CK(hipStreamSynchronize(E->st));
    return 0;
}

#endif
