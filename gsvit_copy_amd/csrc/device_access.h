/*
 * device_access.h -- load/store helpers of the volume kernels (forward and
 * adjoint): every per-cell array is addressed by a byte offset from its
 * allocation start, so that one offset register serves all arrays.
 */
#ifndef GSV_DEVICE_ACCESS_H
#define GSV_DEVICE_ACCESS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Global-memory view of a pointer.  Pointers that reach a kernel through its
 * arguments are known to be global; pointers read from device memory (the
 * per-run parameter blocks of a batch) are not, and without this view their
 * accesses would take the generic (flat) path. */
#define GSV_GLOBAL __attribute__((address_space(1)))

template <typename T>
__device__ __forceinline__ const GSV_GLOBAL T *
gp(const T *p)
{
    return (const GSV_GLOBAL T *)p;
}

template <typename T>
__device__ __forceinline__ GSV_GLOBAL T *
gp(T *p)
{
    return (GSV_GLOBAL T *)p;
}

/* float2 / float4 through a global view (HIP's vector classes do not copy
 * from another address space; clang's vector types do) */
typedef float gsv_v2f __attribute__((ext_vector_type(2)));
typedef float gsv_v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float2
ldg2(const float2 *p)
{
    const gsv_v2f v = *(const GSV_GLOBAL gsv_v2f *)(const void *)p;
    return make_float2(v.x, v.y);
}

__device__ __forceinline__ float4
ldg4(const float4 *p)
{
    const gsv_v4f v = *(const GSV_GLOBAL gsv_v4f *)(const void *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}

/* a plain-data struct whose size is a multiple of 16 bytes (the coefficient
 * table entries) through a global view, as 16-byte vector loads */
template <typename T>
__device__ __forceinline__ T
ldgs(const T *p)
{
    static_assert(sizeof(T) % 16 == 0, "ldgs loads whole 16-byte chunks");
    T out;
    gsv_v4f *o = (gsv_v4f *)(void *)&out;
    const GSV_GLOBAL gsv_v4f *q = (const GSV_GLOBAL gsv_v4f *)(const void *)p;
#pragma unroll
    for (unsigned c = 0; c < sizeof(T) / 16; c++)
        o[c] = q[c];
    return out;
}

/* an E class entry without its padding word: 16 + 12 bytes, so that no
 * register receives a value nobody reads (the compiler would otherwise wait
 * for that load before reusing the register) */
template <typename T>
__device__ __forceinline__ T
ldec(const T *p)
{
    static_assert(sizeof(T) == 32, "an E class entry is 32 bytes");
    typedef float v3f __attribute__((ext_vector_type(3)));
    T out;
    const gsv_v4f a = *(const GSV_GLOBAL gsv_v4f *)(const void *)p;
    const GSV_GLOBAL float *q = (const GSV_GLOBAL float *)(const void *)p;
    float *o = (float *)(void *)&out;
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = q[4]; o[5] = q[5]; o[6] = q[6]; o[7] = 0.f;
    return out;
}

/* an H class entry without its padding: 16 + 4 bytes */
template <typename T>
__device__ __forceinline__ T
ldhc(const T *p)
{
    static_assert(sizeof(T) == 32, "an H class entry is 32 bytes");
    T out;
    const gsv_v4f a = *(const GSV_GLOBAL gsv_v4f *)(const void *)p;
    float *o = (float *)(void *)&out;
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = ((const GSV_GLOBAL float *)(const void *)p)[4];
    o[5] = o[6] = o[7] = 0.f;
    return out;
}

/* Every per-cell array is addressed by a byte offset from its allocation
 * start (the guard plane), shared by all arrays: with a 32-bit offset the
 * loads take the uniform-base + 32-bit-offset form, so the z-march carries
 * one offset register instead of a 64-bit pointer per array. */
template <typename OFF>
__device__ __forceinline__ float
ldf(const float *base, OFF off)
{
    return *gp((const float *)((const char *)base + off));
}

template <typename OFF>
__device__ __forceinline__ void
stf(float *base, OFF off, float v)
{
    *gp((float *)((char *)base + off)) = v;
}


/* Streams a kernel reads or writes exactly once (the updated field family,
 * the ADE state) bypass the caches' retention with nontemporal accesses;
 * the other family, whose neighbours are re-read by adjacent rows, stays
 * cached. */
template <typename OFF>
__device__ __forceinline__ float
ldn(const float *base, OFF off)
{
    return __builtin_nontemporal_load(gp((const float *)((const char *)base + off)));
}

template <typename OFF>
__device__ __forceinline__ void
stn(float *base, OFF off, float v)
{
    __builtin_nontemporal_store(v, gp((float *)((char *)base + off)));
}

template <typename OFF>
__device__ __forceinline__ float
ldsn(const float *base, OFF off, int slot_bytes)
{
    return __builtin_nontemporal_load(gp((const float *)((const char *)base + (size_t)off + slot_bytes)));
}

template <typename OFF>
__device__ __forceinline__ void
stsn(float *base, OFF off, int slot_bytes, float v)
{
    __builtin_nontemporal_store(v, gp((float *)((char *)base + (size_t)off + slot_bytes)));
}

/* state slot of a blocked ADE cell: the 32-bit lane offset is widened before
 * the constant slot offset is added, so the slot becomes an immediate */
template <typename OFF>
__device__ __forceinline__ float
lds(const float *base, OFF off, int slot_bytes)
{
    return *gp((const float *)((const char *)base + (size_t)off + slot_bytes));
}

template <typename OFF>
__device__ __forceinline__ void
sts(float *base, OFF off, int slot_bytes, float v)
{
    *gp((float *)((char *)base + (size_t)off + slot_bytes)) = v;
}

template <typename OFF>
__device__ __forceinline__ float2
ld2(const float2 *base, OFF off)
{
    return ldg2((const float2 *)((const char *)base + off));
}

template <typename OFF>
__device__ __forceinline__ void
st2(float2 *base, OFF off, float a, float b)
{
    gsv_v2f v;
    v.x = a;
    v.y = b;
    *(GSV_GLOBAL gsv_v2f *)(void *)((char *)base + off) = v;
}

template <typename ID, typename OFF>
__device__ __forceinline__ unsigned
ldid(const void *base, OFF off)
{
    return *gp((const ID *)((const char *)base + off));
}

#endif
