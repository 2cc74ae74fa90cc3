/*
 * batch.hip -- batched stepping of independent runs on one GPU.
 *
 * A batch packs the volume and TSF work of many engines (runs of uneven grid
 * sizes, each set up by sv_gpu_new as usual and keeping its own device
 * memory) into one launch per kernel type and half step:
 *   k_hstep_batch  over the H tile list of every grid
 *   k_tsf_line_batch, k_tsf_h_batch   (grids with a TSF source)
 *   k_estep_batch  over the E tile list
 *   k_tsf_e_batch  (one launch per face axis)
 * which is the order of the single-run step (src3d/svit.c:239-257 as run by
 * gpu_engine.hip) restricted to the stages a batched run may have.  Every
 * cell's arithmetic is that of the single-run kernels (the same column and
 * face-cell functions), so a batched run's fields equal its single-run fields
 * bit for bit.  The TSF step index of every grid is kept in device memory and
 * advanced by the line kernel, so a captured sequence of steps (a HIP graph)
 * replays with advancing steps; no memset or host copy is captured.
 *
 * A batched run has CPML or no absorbing boundary, at most a TSF source, no
 * Liao or periodic boundaries, point sources, point traces, box sums, far
 * field, DFT, design region or adjoint; its class ids use one width across
 * the batch and its field arrays stay below 4 GiB (32-bit offsets).
 */
#include <string.h>
#include <algorithm>
#include "engine_host.h"

struct GsvBatch {
    std::vector<Engine *> eng;
    hipStream_t st;
    int idb_h, idb_e, ade;
    int per;                  /* some run has a periodic axis */
    int shape[8];              /* bx_h by_h kchunk_h bx_e by_e kchunk_e blocks_h blocks_e */
    GsvGrid *d_grids;
    GsvTsf *d_tsfs;
    int *d_steps;              /* TSF step index per run */
    int *d_truns;              /* runs with a TSF source */
    int ntsf;
    GsvFaceTile *d_faces;      /* face tiles ordered by axis */
    int nfaces;
    int face_off[4];           /* first face tile of each axis, face_off[3] = nfaces */
    GsvTile *d_th, *d_te;
    int nth, nte;
    std::vector<int> h_steps;
    int graph_steps;           /* steps per captured graph, 0: direct launches */
    hipGraphExec_t gexec;
    int have_graph;
    int timing;
    hipEvent_t ev[4];
    double t_h_ms, t_e_ms;
    long n_h, n_e;
};

static void
tile_list(const GsvBatch *b, int bx, int by, int kc, std::vector<GsvTile> &out)
{
    out.clear();
    for (size_t r = 0; r < b->eng.size(); r++) {
        const GsvGrid &g = b->eng[r]->g;
        const int ntx = (g.nx + bx - 1) / bx, nty = (g.ny + by - 1) / by, ntz = (g.nz + kc - 1) / kc;
        for (int kz = 0; kz < ntz; kz++)
            for (int ty = 0; ty < nty; ty++)
                for (int tx = 0; tx < ntx; tx++)
                    out.push_back(GsvTile{ (uint16_t)r, (uint16_t)tx, (uint16_t)ty, (uint16_t)kz });
    }
}

static int
upload_tiles(GsvBatch *b, GsvTile **d, int *n, const std::vector<GsvTile> &h)
{
    if (*d)
        CK(hipFree(*d));
    *d = NULL;
    *n = (int)h.size();
    CK(hipMalloc(d, std::max<size_t>(h.size(), 1) * sizeof(GsvTile)));
    CK(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(GsvTile), hipMemcpyHostToDevice, b->st));
    CK(hipStreamSynchronize(b->st));
    return 0;
}

static int
drop_graph(GsvBatch *b)
{
    if (b->have_graph)
        CK(hipGraphExecDestroy(b->gexec));
    b->have_graph = 0;
    return 0;
}

static int
batch_apply_shape(GsvBatch *b, const int *shape)
{
    for (int q = 0; q < 2; q++) {
        const int bx = shape[3 * q], by = shape[3 * q + 1], kc = shape[3 * q + 2];
        if (bx <= 0 || by <= 0 || bx * by > 512 || bx % 64 || kc <= 0 || shape[6 + q] < 0)
            return 1;
        for (Engine *E : b->eng)
            if ((E->g.nx + bx - 1) / bx > 65535 || (E->g.ny + by - 1) / by > 65535 || (E->g.nz + kc - 1) / kc > 65535)
                return 1;
    }
    std::vector<GsvTile> h;
    if (shape[0] != b->shape[0] || shape[1] != b->shape[1] || shape[2] != b->shape[2] || !b->d_th) {
        tile_list(b, shape[0], shape[1], shape[2], h);
        if (upload_tiles(b, &b->d_th, &b->nth, h))
            return 1;
    }
    if (shape[3] != b->shape[3] || shape[4] != b->shape[4] || shape[5] != b->shape[5] || !b->d_te) {
        tile_list(b, shape[3], shape[4], shape[5], h);
        if (upload_tiles(b, &b->d_te, &b->nte, h))
            return 1;
    }
    memcpy(b->shape, shape, sizeof(b->shape));
    return drop_graph(b);
}

/* why engine E cannot take part in a batch, or NULL */
static const char *
batch_reject(const Engine *E)
{
    if (!E || !E->ready) return "engine not initialised";
    if (E->has_src) return "point sources";
    if (E->has_liao) return "Liao boundaries";
    if (E->has_per) return "periodic boundaries";
    if (E->has_ff) return "time-domain far field";
    if (E->has_pff) return "periodic far field";
    if (E->has_ltsf) return "layered TSF";
    if (E->has_tsff) return "focused TSF";
    if (E->has_ext) return "external plane source";
    if (E->has_sf) return "scattered-field source";
    if (E->has_dft) return "running DFT";
    if (E->has_flux) return "flux spectra";
    if (!E->sums.empty()) return "box sums";
    if (!E->forces.empty()) return "force boxes";
    if (E->npts) return "point outputs";
    if (E->adj) return "adjoint state";
    if (E->g.pc_nx > 0) return "a design region";
    if (E->g.rc) return "recursive-convolution materials";
    if ((unsigned long long)E->g.plane * (E->g.nz + 2) * 4ull >= (1ull << 32)) return "field arrays of 4 GiB or more";
    return NULL;
}

extern "C" void gsv_engine_batch_free(void *bp);

extern "C" void *
gsv_engine_batch_new(void *const *engines, int n)
{
    if (n <= 0 || n > 65535) {
        fprintf(stderr, "Error: a batch holds 1 to 65535 runs (got %d)\n", n);
        return NULL;
    }
    GsvBatch *b = new GsvBatch();
    b->eng.assign((Engine *const *)engines, (Engine *const *)engines + n);
    for (int r = 0; r < n; r++) {
        const char *why = batch_reject(b->eng[r]);
        if (!why && (b->eng[r]->idbytes_h != b->eng[0]->idbytes_h || b->eng[r]->idbytes_e != b->eng[0]->idbytes_e))
            why = "class-id width differing from run 0";
        if (!why && b->eng[r]->dev != b->eng[0]->dev)
            why = "a different GPU than run 0";
        for (int q = 0; !why && q < r; q++)   /* the pole state layout follows the pole count */
            if (b->eng[r]->g.ade_poles && b->eng[q]->g.ade_poles && b->eng[r]->g.ade_poles != b->eng[q]->g.ade_poles)
                why = "a critical-point pole count differing from another run's";
        if (why) {
            fprintf(stderr, "Error: run %d cannot be batched: %s\n", r, why);
            delete b;
            return NULL;
        }
    }
    b->idb_h = b->eng[0]->idbytes_h;
    b->idb_e = b->eng[0]->idbytes_e;
    /* the widest dispersive branch covers the others: a grid without ADE
     * cells (or without critical-point terms) never takes it */
    b->ade = 0;
    b->per = 0;
    for (Engine *E : b->eng) {
        b->ade = std::max(b->ade, E->g.ade ? (E->g.ade_poles ? E->g.ade_poles : 1) : 0);
        b->per |= E->g.perx || E->g.pery;
    }

    bool ok = hipSetDevice(b->eng[0]->dev) == hipSuccess && hipStreamCreate(&b->st) == hipSuccess;
    std::vector<GsvGrid> grids(n);
    std::vector<GsvTsf> tsfs(n);
    std::vector<int> truns;
    std::vector<GsvFaceTile> faces;
    for (int r = 0; r < n; r++) {
        const Engine *E = b->eng[r];
        grids[r] = E->g;
        memset(&tsfs[r], 0, sizeof(GsvTsf));
        if (!E->has_tsf)
            continue;
        tsfs[r] = E->tsf;
        truns.push_back(r);
    }
    /* E corrections of faces of different axes meet on the box edges (both
     * correct the same component there), so the E faces run one axis per
     * launch, in the single-run order; the H corrections of different faces
     * touch different cells or components and run in one launch */
    for (int axis = 0; axis < 3; axis++) {
        b->face_off[axis] = (int)faces.size();
        for (int r : truns) {
            const Engine *E = b->eng[r];
            const GsvTsf &T = E->tsf;
            int na, nb;
            tsf_face_extent(T, E->g.nx, E->g.ny, axis, &na, &nb);
            for (int far = 0; far < 2; far++) {
                if (!E->tsf_face_on[2 * axis + far])
                    continue;
                for (int tb = 0; tb < (nb + 3) / 4; tb++)
                    for (int ta = 0; ta < (na + 63) / 64; ta++)
                        faces.push_back(GsvFaceTile{ (uint16_t)r, (uint16_t)(2 * axis + far), (uint16_t)ta, (uint16_t)tb });
            }
        }
    }
    b->face_off[3] = (int)faces.size();
    b->ntsf = (int)truns.size();
    b->nfaces = (int)faces.size();
    b->h_steps.assign(n, 0);
    ok = ok && hipMalloc(&b->d_grids, n * sizeof(GsvGrid)) == hipSuccess
         && hipMalloc(&b->d_tsfs, n * sizeof(GsvTsf)) == hipSuccess
         && hipMalloc(&b->d_steps, n * sizeof(int)) == hipSuccess
         && hipMalloc(&b->d_truns, std::max(b->ntsf, 1) * sizeof(int)) == hipSuccess
         && hipMalloc(&b->d_faces, std::max(b->nfaces, 1) * sizeof(GsvFaceTile)) == hipSuccess
         && hipMemcpy(b->d_grids, grids.data(), n * sizeof(GsvGrid), hipMemcpyHostToDevice) == hipSuccess
         && hipMemcpy(b->d_tsfs, tsfs.data(), n * sizeof(GsvTsf), hipMemcpyHostToDevice) == hipSuccess
         && hipMemcpy(b->d_truns, truns.data(), b->ntsf * sizeof(int), hipMemcpyHostToDevice) == hipSuccess
         && hipMemcpy(b->d_faces, faces.data(), b->nfaces * sizeof(GsvFaceTile), hipMemcpyHostToDevice) == hipSuccess;
    for (int q = 0; ok && q < 4; q++)
        ok = hipEventCreate(&b->ev[q]) == hipSuccess;
    /* the single-run default shape, one workgroup per tile, direct launches */
    const int shape[8] = { 64, 4, 16, 64, 4, 16, 0, 0 };
    if (!ok || batch_apply_shape(b, shape)) {
        fprintf(stderr, "Error: batch set-up failed (%s)\n", hipGetErrorString(hipGetLastError()));
        gsv_engine_batch_free(b);
        return NULL;
    }
    return b;
}

static int
enqueue_step(GsvBatch *b)
{
    const int *s = b->shape;
    if (b->timing)
        CK(hipEventRecord(b->ev[0], b->st));
    CK(gsv_launch_hstep_batch(b->d_grids, b->d_th, b->nth, b->idb_h, b->per, s[0], s[1], s[2], s[6], b->st));
    if (b->timing)
        CK(hipEventRecord(b->ev[1], b->st));
    if (b->ntsf) {
        CK(gsv_launch_tsf_line_batch(b->d_tsfs, b->d_truns, b->ntsf, b->d_steps, b->st));
        CK(gsv_launch_tsf_faces_batch(b->d_grids, b->d_tsfs, b->d_faces, b->nfaces, 1, b->st));
    }
    if (b->timing)
        CK(hipEventRecord(b->ev[2], b->st));
    CK(gsv_launch_estep_batch(b->d_grids, b->d_te, b->nte, b->idb_e, b->ade, b->per, s[3], s[4], s[5], s[7], b->st));
    if (b->timing)
        CK(hipEventRecord(b->ev[3], b->st));
    for (int axis = 0; b->ntsf && axis < 3; axis++)
        CK(gsv_launch_tsf_faces_batch(b->d_grids, b->d_tsfs, b->d_faces + b->face_off[axis],
                                      b->face_off[axis + 1] - b->face_off[axis], 0, b->st));
    if (b->timing) {
        float ms;
        CK(hipEventSynchronize(b->ev[3]));
        CK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        b->t_h_ms += ms;
        b->n_h++;
        CK(hipEventElapsedTime(&ms, b->ev[2], b->ev[3]));
        b->t_e_ms += ms;
        b->n_e++;
    }
    return 0;
}

static int
capture(GsvBatch *b)
{
    hipGraph_t graph;
    CK(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
    int rc = 0;
    for (int q = 0; q < b->graph_steps && !rc; q++)
        rc = enqueue_step(b);
    hipError_t e = hipStreamEndCapture(b->st, &graph);
    if (rc)
        return 1;
    CK(e);
    e = hipGraphInstantiate(&b->gexec, graph, NULL, NULL, 0);
    CK(hipGraphDestroy(graph));
    CK(e);
    b->have_graph = 1;
    return 0;
}

/* nsteps steps of every run; steps[r]: run r's current step (its TSF index) */
extern "C" int
gsv_engine_batch_run(void *bp, int nsteps, const int *steps)
{
    GsvBatch *b = (GsvBatch *)bp;
    CK(hipSetDevice(b->eng[0]->dev));
    for (Engine *E : b->eng)
        CK(hipStreamSynchronize(E->st));
    for (size_t r = 0; r < b->eng.size(); r++)
        b->h_steps[r] = steps[r];
    CK(hipMemcpyAsync(b->d_steps, b->h_steps.data(), b->eng.size() * sizeof(int), hipMemcpyHostToDevice, b->st));
    int done = 0;
    if (b->graph_steps > 0 && !b->timing) {
        if (!b->have_graph && nsteps >= b->graph_steps && capture(b))
            return 1;
        for (; done + b->graph_steps <= nsteps; done += b->graph_steps)
            CK(hipGraphLaunch(b->gexec, b->st));
    }
    for (; done < nsteps; done++)
        if (enqueue_step(b))
            return 1;
    CK(hipStreamSynchronize(b->st));
    return 0;
}

extern "C" int
gsv_engine_batch_set_launch(void *bp, const int *shape)
{
    GsvBatch *b = (GsvBatch *)bp;
    if (!memcmp(shape, b->shape, sizeof(b->shape)))
        return 0;
    return batch_apply_shape(b, shape);
}

extern "C" void
gsv_engine_batch_get_launch(void *bp, int *shape)
{
    memcpy(shape, ((GsvBatch *)bp)->shape, sizeof(((GsvBatch *)bp)->shape));
}

extern "C" int
gsv_engine_batch_set_graph(void *bp, int steps_per_graph)
{
    GsvBatch *b = (GsvBatch *)bp;
    if (steps_per_graph < 0)
        return 1;
    if (steps_per_graph != b->graph_steps && drop_graph(b))
        return 1;
    b->graph_steps = steps_per_graph;
    return 0;
}

extern "C" void
gsv_engine_batch_set_timing(void *bp, int on)
{
    GsvBatch *b = (GsvBatch *)bp;
    b->timing = on;
    b->t_h_ms = b->t_e_ms = 0;
    b->n_h = b->n_e = 0;
}

/* out: mean H ms, #H launches, mean E ms, #E launches, H tiles, E tiles, face tiles, TSF runs */
extern "C" void
gsv_engine_batch_info(void *bp, double *out)
{
    GsvBatch *b = (GsvBatch *)bp;
    out[0] = b->n_h ? b->t_h_ms / b->n_h : 0;
    out[1] = (double)b->n_h;
    out[2] = b->n_e ? b->t_e_ms / b->n_e : 0;
    out[3] = (double)b->n_e;
    out[4] = b->nth;
    out[5] = b->nte;
    out[6] = b->nfaces;
    out[7] = b->ntsf;
}

extern "C" void
gsv_engine_batch_free(void *bp)
{
    GsvBatch *b = (GsvBatch *)bp;
    if (!b)
        return;
    if (b->st) {
        hipStreamSynchronize(b->st);
        drop_graph(b);
    }
    hipFree(b->d_grids);
    hipFree(b->d_tsfs);
    hipFree(b->d_steps);
    hipFree(b->d_truns);
    hipFree(b->d_faces);
    hipFree(b->d_th);
    hipFree(b->d_te);
    for (int q = 0; q < 4; q++)
        if (b->ev[q])
            hipEventDestroy(b->ev[q]);
    if (b->st)
        hipStreamDestroy(b->st);
    delete b;
}
