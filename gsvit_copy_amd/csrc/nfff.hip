/*
 * nfff.hip -- far field of the running DFT (NFFF_DFT accumulators on the NFFF
 * box faces).  No reference counterpart: it is the frequency-domain
 * counterpart of the time-domain Kirchhoff transform of
 * src3d/farfield.c:174-407, evaluated once after the run.
 *
 * Surface equivalence on the box: J = n x H and M = E x n (n the outward
 * normal) from the accumulated tangential E and face-plane H (the sampling
 * kernel averages H over the two planes either side of a face), each
 * component at its own Yee position on the face:
 *   x faces: Ey, Hz at (p, j+1/2, k);  Ez, Hy at (p, j, k+1/2)
 *   y faces: Ex, Hz at (i+1/2, p, k);  Ez, Hx at (i, p, k+1/2)
 *   z faces: Ex, Hy at (i+1/2, j, p);  Ey, Hx at (i, j+1/2, p)
 * weighted by the cell area and dt (the DFT sums approximate Fourier
 * integrals).  Radiation vectors N = sum J exp(+i k r.r') dA and
 * L = sum M exp(+i k r.r') dA, r' from the grid origin; the far-field
 * pattern (E = F exp(-i k r) / r, fields ~ exp(+i w t)) is
 *   F_theta = -i k / (4 pi) (L_phi + eta0 N_theta)
 *   F_phi   = +i k / (4 pi) (L_theta - eta0 N_phi).
 * Face cells inside an NFFF_SKIP rectangle are left out, as in the
 * time-domain transform.
 *
 * Kernels: k_nfff_partial -- one thread per direction, a block per
 * (256 directions, wavelength, chunk of face cells); the chunk's cells are
 * staged through LDS 128 at a time (positions and signed J, M of both pairs),
 * every thread reads the same cell (LDS broadcast) and accumulates N and L in
 * double; the chunk's sums go to a partial buffer.  k_nfff_reduce -- one
 * thread per (direction, wavelength): sums the chunks and projects.  FP64
 * throughout: the work is two sincos and 48 FMA per cell, direction and
 * wavelength, run once per job.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <mutex>
#include "engine_host.h"

#define NFFF_THREADS 256   /* directions per block */
#define NFFF_TILE 128      /* face cells staged per pass */

/* the NFFF box and its skip rectangles (per face: flag, min a, min b, max a,
 * max b over the face's two in-plane indices: x faces j, k; y faces i, k;
 * z faces i, j) */
struct NfffBox {
    int i0, j0, k0, i1, j1, k1, ix, iy, iz;
    int skip[6][5];
    double dx, dy, dz, dt;
};

/* one staged face cell: the positions (m) of its two component pairs and the
 * area- and dt-weighted J, M vectors at each (x, y, z as re, im) */
struct NfffCell {
    double pa[3], pb[3];
    double ja[6], ma[6], jb[6], mb[6];
};

__device__ static void
nfff_stage(const NfffBox &b, const GsvDft &d, long long q, int w, NfffCell &c)
{
    for (int m = 0; m < 6; m++)
        c.ja[m] = c.ma[m] = c.jb[m] = c.mb[m] = 0.0;
    const long long r0 = q / 2, nyz = (long long)b.iy * b.iz, nxz = (long long)b.ix * b.iz;
    const int far = (int)(q % 2);
    const double s = far ? 1.0 : -1.0;
    int i, j, k, face, ua, ub;
    double area;
    if (r0 < nyz) {
        j = b.j0 + (int)(r0 / b.iz); k = b.k0 + (int)(r0 % b.iz); i = far ? b.i1 : b.i0;
        face = far; ua = j; ub = k; area = b.dy * b.dz;
    } else if (r0 < nyz + nxz) {
        const long long r = r0 - nyz;
        i = b.i0 + (int)(r / b.iz); k = b.k0 + (int)(r % b.iz); j = far ? b.j1 : b.j0;
        face = 2 + far; ua = i; ub = k; area = b.dx * b.dz;
    } else {
        const long long r = r0 - nyz - nxz;
        i = b.i0 + (int)(r / b.iy); j = b.j0 + (int)(r % b.iy); k = far ? b.k1 : b.k0;
        face = 4 + far; ua = i; ub = j; area = b.dx * b.dy;
    }
    const int *sk = b.skip[face];
    c.pa[0] = c.pb[0] = i * b.dx;
    c.pa[1] = c.pb[1] = j * b.dy;
    c.pa[2] = c.pb[2] = k * b.dz;
    if (sk[0] && ua >= sk[1] && ub >= sk[2] && ua <= sk[3] && ub <= sk[4])
        return;
    const double *v = d.acc + ((size_t)q * d.nw + w) * 8;   /* v0, v1 (E), v2, v3 (face-plane H), re / im */
    const double f = s * area * b.dt;
    if (face < 2) {
        c.pa[1] += 0.5 * b.dy;  /* Ey, Hz */
        c.pb[2] += 0.5 * b.dz;  /* Ez, Hy */
        c.ja[2] = -f * v[6]; c.ja[3] = -f * v[7];   /* J_y = -s Hz */
        c.ma[4] = -f * v[0]; c.ma[5] = -f * v[1];   /* M_z = -s Ey */
        c.jb[4] = f * v[4]; c.jb[5] = f * v[5];     /* J_z = s Hy */
        c.mb[2] = f * v[2]; c.mb[3] = f * v[3];     /* M_y = s Ez */
    } else if (face < 4) {
        c.pa[0] += 0.5 * b.dx;  /* Ex, Hz */
        c.pb[2] += 0.5 * b.dz;  /* Ez, Hx */
        c.ja[0] = f * v[6]; c.ja[1] = f * v[7];     /* J_x = s Hz */
        c.ma[4] = f * v[0]; c.ma[5] = f * v[1];     /* M_z = s Ex */
        c.jb[4] = -f * v[4]; c.jb[5] = -f * v[5];   /* J_z = -s Hx */
        c.mb[0] = -f * v[2]; c.mb[1] = -f * v[3];   /* M_x = -s Ez */
    } else {
        c.pa[0] += 0.5 * b.dx;  /* Ex, Hy */
        c.pb[1] += 0.5 * b.dy;  /* Ey, Hx */
        c.ja[0] = -f * v[6]; c.ja[1] = -f * v[7];   /* J_x = -s Hy */
        c.ma[2] = -f * v[0]; c.ma[3] = -f * v[1];   /* M_y = -s Ex */
        c.jb[2] = f * v[4]; c.jb[3] = f * v[5];     /* J_y = s Hx */
        c.mb[0] = f * v[2]; c.mb[1] = f * v[3];     /* M_x = s Ey */
    }
}

/* partial[chunk][w][dir][12] = N (x, y, z as re, im), L (same) of the chunk's cells */
__global__ __launch_bounds__(NFFF_THREADS) void
k_nfff_partial(NfffBox b, GsvDft d, const double *dirs, int ndirs, long long per_chunk, double *partial)
{
    __shared__ NfffCell tile[NFFF_TILE];
    const int w = blockIdx.y, chunk = blockIdx.z;
    const int dir = blockIdx.x * NFFF_THREADS + threadIdx.x;
    const bool on = dir < ndirs;
    const double kk = d.omega[w] / SV_LIGHT_SPEED;
    double rx = 0.0, ry = 0.0, rz = 0.0;
    if (on) {
        double st, ct, sp, cp;
        sincos(dirs[2 * dir], &st, &ct);
        sincos(dirs[2 * dir + 1], &sp, &cp);
        rx = kk * st * cp;
        ry = kk * st * sp;
        rz = kk * ct;
    }
    double N[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 }, L[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    const long long q0 = (long long)chunk * per_chunk;
    const long long q1 = min(q0 + per_chunk, d.ncells);
    for (long long t0 = q0; t0 < q1; t0 += NFFF_TILE) {
        const int n = (int)min((long long)NFFF_TILE, q1 - t0);
        __syncthreads();
        for (int t = threadIdx.x; t < n; t += NFFF_THREADS)
            nfff_stage(b, d, t0 + t, w, tile[t]);
        __syncthreads();
        if (on) {
            for (int t = 0; t < n; t++) {
                const NfffCell &c = tile[t];
                double sa, ca, sb, cb;
                sincos(rx * c.pa[0] + ry * c.pa[1] + rz * c.pa[2], &sa, &ca);
                sincos(rx * c.pb[0] + ry * c.pb[1] + rz * c.pb[2], &sb, &cb);
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    N[2 * m] += c.ja[2 * m] * ca - c.ja[2 * m + 1] * sa + c.jb[2 * m] * cb - c.jb[2 * m + 1] * sb;
                    N[2 * m + 1] += c.ja[2 * m] * sa + c.ja[2 * m + 1] * ca + c.jb[2 * m] * sb + c.jb[2 * m + 1] * cb;
                    L[2 * m] += c.ma[2 * m] * ca - c.ma[2 * m + 1] * sa + c.mb[2 * m] * cb - c.mb[2 * m + 1] * sb;
                    L[2 * m + 1] += c.ma[2 * m] * sa + c.ma[2 * m + 1] * ca + c.mb[2 * m] * sb + c.mb[2 * m + 1] * cb;
                }
            }
        }
    }
    if (on) {
        double *p = partial + (((size_t)chunk * d.nw + w) * ndirs + dir) * 12;
        for (int m = 0; m < 6; m++) {
            p[m] = N[m];
            p[6 + m] = L[m];
        }
    }
}

/* out[dir][w][4] = re / im of F_theta and F_phi from the chunks' partial sums */
__global__ __launch_bounds__(256) void
k_nfff_reduce(const double *partial, int nchunks, int nw, int ndirs, const double *dirs, const double *omega,
              double *out)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ndirs * nw)
        return;
    const int dir = (int)(t / nw), w = (int)(t % nw);
    double v[12];
    for (int m = 0; m < 12; m++)
        v[m] = 0.0;
    for (int c = 0; c < nchunks; c++) {
        const double *p = partial + (((size_t)c * nw + w) * ndirs + dir) * 12;
        for (int m = 0; m < 12; m++)
            v[m] += p[m];
    }
    double st, ct, sp, cp;
    sincos(dirs[2 * dir], &st, &ct);
    sincos(dirs[2 * dir + 1], &sp, &cp);
    const double th[3] = { ct * cp, ct * sp, -st }, ph[3] = { -sp, cp, 0.0 };
    const double eta0 = SV_MU_0 * SV_LIGHT_SPEED;
    const double f = omega[w] / SV_LIGHT_SPEED / (4.0 * M_PI);
    double At[2], Ap[2];
    for (int c = 0; c < 2; c++) {
        const double Nt = v[c] * th[0] + v[2 + c] * th[1] + v[4 + c] * th[2];
        const double Np = v[c] * ph[0] + v[2 + c] * ph[1] + v[4 + c] * ph[2];
        const double Lt = v[6 + c] * th[0] + v[8 + c] * th[1] + v[10 + c] * th[2];
        const double Lp = v[6 + c] * ph[0] + v[8 + c] * ph[1] + v[10 + c] * ph[2];
        At[c] = Lp + eta0 * Nt;   /* F_theta = -i f At */
        Ap[c] = Lt - eta0 * Np;   /* F_phi   = +i f Ap */
    }
    double *o = out + (size_t)t * 4;
    o[0] = f * At[1];
    o[1] = -f * At[0];
    o[2] = -f * Ap[1];
    o[3] = f * Ap[0];
}

static void
nfff_box(const Engine *E, NfffBox *b)
{
    const SvSetFarfield *sf = &E->set->sf;
    memset(b, 0, sizeof(*b));
    b->i0 = sf->box_i0; b->j0 = sf->box_j0; b->k0 = sf->box_k0;
    b->i1 = sf->box_in; b->j1 = sf->box_jn; b->k1 = sf->box_kn;
    b->ix = b->i1 - b->i0; b->iy = b->j1 - b->j0; b->iz = b->k1 - b->k0;
    const int sk[6][5] = {
        { sf->box_boundary_skipi0, sf->skipi0_jmin, sf->skipi0_kmin, sf->skipi0_jmax, sf->skipi0_kmax },
        { sf->box_boundary_skipin, sf->skipin_jmin, sf->skipin_kmin, sf->skipin_jmax, sf->skipin_kmax },
        { sf->box_boundary_skipj0, sf->skipj0_imin, sf->skipj0_kmin, sf->skipj0_imax, sf->skipj0_kmax },
        { sf->box_boundary_skipjn, sf->skipjn_imin, sf->skipjn_kmin, sf->skipjn_imax, sf->skipjn_kmax },
        { sf->box_boundary_skipk0, sf->skipk0_imin, sf->skipk0_jmin, sf->skipk0_imax, sf->skipk0_jmax },
        { sf->box_boundary_skipkn, sf->skipkn_imin, sf->skipkn_jmin, sf->skipkn_imax, sf->skipkn_jmax } };
    memcpy(b->skip, sk, sizeof(sk));
    b->dx = E->set->sp.dx; b->dy = E->set->sp.dy; b->dz = E->set->sp.dz;
    b->dt = E->set->plan.dt;
}

/* chunks of face cells per (direction tile, wavelength): enough blocks to
 * fill the chip (about 2048), at least one tile of cells per chunk, and a
 * partial buffer of at most max_doubles; 0 when one chunk does not fit */
static long long
nfff_chunks(long long ncells, int ndirs, int nw, size_t max_doubles)
{
    const long long per_chunk_doubles = 12LL * nw * ndirs;
    const long long bxy = (long long)((ndirs + NFFF_THREADS - 1) / NFFF_THREADS) * nw;
    long long c = (2048 + bxy - 1) / bxy;
    c = std::min(c, (ncells + NFFF_TILE - 1) / NFFF_TILE);
    c = std::min(c, 256LL);
    c = std::max(c, 1LL);
    c = std::min(c, (long long)(max_doubles / (size_t)per_chunk_doubles));
    return c;
}

/* the two kernels on the engine's stream; dirs, partial and out are device memory */
static int
nfff_launch(Engine *E, const double *dirs, int ndirs, double *partial, long long nchunks, double *out)
{
    NfffBox b;
    nfff_box(E, &b);
    const GsvDft &d = E->dft;
    const long long per_chunk = (d.ncells + nchunks - 1) / nchunks;
    const dim3 grd((ndirs + NFFF_THREADS - 1) / NFFF_THREADS, d.nw, (unsigned)nchunks);
    hipLaunchKernelGGL(k_nfff_partial, grd, dim3(NFFF_THREADS), 0, E->st, b, d, dirs, ndirs, per_chunk, partial);
    CK(hipGetLastError());
    const long long nt = (long long)ndirs * d.nw;
    hipLaunchKernelGGL(k_nfff_reduce, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, E->st, partial, (int)nchunks,
                       d.nw, ndirs, dirs, d.omega, out);
    CK(hipGetLastError());
    return 0;
}

/* Far-field pattern of the running DFT at ndirs directions: dirs [ndirs][2]
 * (theta, phi; host), out [ndirs][nw][4] (host).  Device buffers for the
 * directions, the per-chunk partial sums and the result are allocated for
 * the call.  Also the transform attached to the far-field data
 * (NFFF_DFT_PATTERN files).  Returns 0 on success. */
extern "C" int
gsv_engine_farfield_dft(void *eng, int ndirs, const double *dirs, double *out)
{
    Engine *E = (Engine *)eng;
    if (!E || !E->ready || !E->has_dft) {
        fprintf(stderr, "Error: far-field DFT: the pool has no NFFF_DFT accumulators\n");
        return 1;
    }
    if (ndirs <= 0)
        return ndirs < 0;
    CK(hipSetDevice(E->dev));
    if (gsv_engine_dft_flush(E))
        return 1;
    const int nw = E->dft.nw;
    const long long nchunks = nfff_chunks(E->dft.ncells, ndirs, nw, (size_t)1 << 28);
    if (nchunks < 1) {
        fprintf(stderr, "Error: far-field DFT: %d directions x %d wavelengths is too many\n", ndirs, nw);
        return 1;
    }
    double *d_dirs = NULL, *d_part = NULL, *d_out = NULL;
    int rc = 1;
    const size_t nout = (size_t)ndirs * nw * 4;
    if (hipMalloc(&d_dirs, (size_t)ndirs * 2 * 8) == hipSuccess
        && hipMalloc(&d_part, (size_t)nchunks * 12 * nw * ndirs * 8) == hipSuccess
        && hipMalloc(&d_out, nout * 8) == hipSuccess
        && hipMemcpyAsync(d_dirs, dirs, (size_t)ndirs * 2 * 8, hipMemcpyHostToDevice, E->st) == hipSuccess
        && nfff_launch(E, d_dirs, ndirs, d_part, nchunks, d_out) == 0
        && hipMemcpyAsync(out, d_out, nout * 8, hipMemcpyDeviceToHost, E->st) == hipSuccess
        && hipStreamSynchronize(E->st) == hipSuccess)
        rc = 0;
    else
        fprintf(stderr, "Error: far-field DFT: %s\n", hipGetErrorString(hipGetLastError()));
    hipFree(d_dirs);
    hipFree(d_part);
    hipFree(d_out);
    return rc;
}

/* ------------------------------------------------------------------ */
/* the same transform without allocation                              */
/* ------------------------------------------------------------------ */

/* Partial sums of gsv_farfield_dft_noalloc: a fixed 16 MiB per GPU, held
 * until the process ends.  Its host-side name is not a device address; the
 * address comes from hipGetSymbolAddress once per device. */
#define NFFF_SCRATCH_DOUBLES (2u << 20)
static __device__ __align__(256) double scratch[NFFF_SCRATCH_DOUBLES];

#define NFFF_MAX_DEVICES 64
static double *scratch_ptr[NFFF_MAX_DEVICES];
static std::mutex scratch_mutex;

/* the device address of this device's scratch array (the device current) */
static double *
scratch_on(int dev)
{
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (dev < 0 || dev >= NFFF_MAX_DEVICES)
        return NULL;
    if (!scratch_ptr[dev]) {
        void *p = NULL;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            fprintf(stderr, "Error: far-field DFT: no scratch address on device %d (%s)\n", dev, hipGetErrorString(e));
            return NULL;
        }
        scratch_ptr[dev] = (double *)p;
    }
    return scratch_ptr[dev];
}

/* As gsv_farfield_dft for callers that cannot allocate: dirs [ndirs][2] and
 * out [ndirs][nw][4] are device memory; the partial sums go to the fixed
 * per-GPU scratch array, so ndirs x wavelengths may not exceed
 * NFFF_SCRATCH_DOUBLES / 12.  Calls on one GPU share that array: calls that
 * overlap in time on one GPU race.  Runs on the session's stream and
 * synchronises it before returning. */
extern "C" int
gsv_farfield_dft_noalloc(void *session, int ndirs, const double *dirs, double *out)
{
    Engine *E = (Engine *)gsv_session_engine(session);
    if (!E || !E->ready || !E->has_dft) {
        fprintf(stderr, "Error: far-field DFT: the pool has no NFFF_DFT accumulators\n");
        return 1;
    }
    if (ndirs <= 0)
        return ndirs < 0;
    const int nw = E->dft.nw;
    const long long nchunks = nfff_chunks(E->dft.ncells, ndirs, nw, NFFF_SCRATCH_DOUBLES);
    if (nchunks < 1) {
        fprintf(stderr, "Error: far-field DFT without allocation: %d directions x %d wavelengths exceed the "
                        "scratch limit of %u (directions x wavelengths)\n", ndirs, nw, NFFF_SCRATCH_DOUBLES / 12);
        return 1;
    }
    CK(hipSetDevice(E->dev));
    double *part = scratch_on(E->dev);
    if (!part)
        return 1;
    if (gsv_engine_dft_flush(E) || nfff_launch(E, dirs, ndirs, part, nchunks, out))
        return 1;
    CK(hipStreamSynchronize(E->st));
    return 0;
}
