/*
 * bench_probe.hip -- measurement tooling for bench.py, not part of the
 * engine: a plain streaming copy b[i] = a[i] of a known size on the current
 * device, timed with HIP events.  It gives the attainable HBM bandwidth that
 * the engine's kernels are read against, and, run under rocprofv3 --pmc
 * FETCH_SIZE / WRITE_SIZE, the counter readings of a known traffic that
 * scale the engine's kernel counters to bytes.  Built by
 * gsvit_copy_amd/build.py as gsvit_copy_amd/lib/libgsv_probe.so.
 *
 * Variants: 0 -- one float4 per thread, plain loads and stores; 1 -- the
 * same with nontemporal loads and stores; 2 -- four float4 per thread
 * (strided by the grid), nontemporal.
 */
#include <hip/hip_runtime.h>

typedef float v4f __attribute__((ext_vector_type(4)));   /* 16 bytes; what the nontemporal builtins take */

__global__ __launch_bounds__(256) void
k_copy_plain(const v4f *__restrict__ a, v4f *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a[i];
}

__global__ __launch_bounds__(256) void
k_copy_nt(const v4f *__restrict__ a, v4f *__restrict__ b, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const v4f v = __builtin_nontemporal_load(a + i);
        __builtin_nontemporal_store(v, b + i);
    }
}

__global__ __launch_bounds__(256) void
k_copy_nt4(const v4f *__restrict__ a, v4f *__restrict__ b, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    v4f v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (i + u * stride < n)
            v[u] = __builtin_nontemporal_load(a + i + u * stride);
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (i + u * stride < n)
            __builtin_nontemporal_store(v[u], b + i + u * stride);
}

static void
launch(int variant, const v4f *a, v4f *b, long long n, hipStream_t st)
{
    if (variant == 2) {
        const long long blocks = (n + 4 * 256 - 1) / (4 * 256);
        hipLaunchKernelGGL(k_copy_nt4, dim3((unsigned)blocks), dim3(256), 0, st, a, b, n);
    } else {
        const long long blocks = (n + 255) / 256;
        if (variant == 1)
            hipLaunchKernelGGL(k_copy_nt, dim3((unsigned)blocks), dim3(256), 0, st, a, b, n);
        else
            hipLaunchKernelGGL(k_copy_plain, dim3((unsigned)blocks), dim3(256), 0, st, a, b, n);
    }
}

/* copies nbytes (a multiple of 16; read nbytes + write nbytes) reps times
 * after one untimed copy; *ms_per_copy = mean time of one copy.  0 on success. */
extern "C" int
gsv_probe_copy(long long nbytes, int reps, int variant, double *ms_per_copy)
{
    v4f *a = NULL, *b = NULL;
    hipEvent_t e0, e1;
    hipStream_t st;
    float ms = 0.f;
    const long long n = nbytes / 16;
    if (n <= 0 || reps <= 0 || variant < 0 || variant > 2 || (n + 255) / 256 > 0x7fffffffLL)
        return 1;
    if (hipMalloc(&a, n * 16) != hipSuccess)
        return 1;
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        return 1;
    }
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(a);
        (void)hipFree(b);
        return 1;
    }
    (void)hipMemsetAsync(a, 0, n * 16, st);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch(variant, a, b, n, st);
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++)
        launch(variant, a, b, n, st);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return hipGetLastError() != hipSuccess;
}
