/*
 * device_access.h -- load/store helpers of the volume kernels (forward and
 * adjoint): every per-cell array is addressed by a byte offset from its
 * allocation start, so that one offset register serves all arrays.
 */
#ifndef GSV_DEVICE_ACCESS_H
#define GSV_DEVICE_ACCESS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* Every per-cell array is addressed by a byte offset from its allocation
 * start (the guard plane), shared by all arrays: with a 32-bit offset the
 * loads take the uniform-base + 32-bit-offset form, so the z-march carries
 * one offset register instead of a 64-bit pointer per array. */
template <typename OFF>
__device__ __forceinline__ float
ldf(const float *base, OFF off)
{
    return *(const float *)((const char *)base + off);
}

template <typename OFF>
__device__ __forceinline__ void
stf(float *base, OFF off, float v)
{
    *(float *)((char *)base + off) = v;
}


/* Streams a kernel reads or writes exactly once (the updated field family,
 * the ADE state) bypass the caches' retention with nontemporal accesses;
 * the other family, whose neighbours are re-read by adjacent rows, stays
 * cached. */
template <typename OFF>
__device__ __forceinline__ float
ldn(const float *base, OFF off)
{
    return __builtin_nontemporal_load((const float *)((const char *)base + off));
}

template <typename OFF>
__device__ __forceinline__ void
stn(float *base, OFF off, float v)
{
    __builtin_nontemporal_store(v, (float *)((char *)base + off));
}

template <typename OFF>
__device__ __forceinline__ float
ldsn(const float *base, OFF off, int slot_bytes)
{
    return __builtin_nontemporal_load((const float *)((const char *)base + (size_t)off + slot_bytes));
}

template <typename OFF>
__device__ __forceinline__ void
stsn(float *base, OFF off, int slot_bytes, float v)
{
    __builtin_nontemporal_store(v, (float *)((char *)base + (size_t)off + slot_bytes));
}

/* state slot of a blocked ADE cell: the 32-bit lane offset is widened before
 * the constant slot offset is added, so the slot becomes an immediate */
template <typename OFF>
__device__ __forceinline__ float
lds(const float *base, OFF off, int slot_bytes)
{
    return *(const float *)((const char *)base + (size_t)off + slot_bytes);
}

template <typename OFF>
__device__ __forceinline__ void
sts(float *base, OFF off, int slot_bytes, float v)
{
    *(float *)((char *)base + (size_t)off + slot_bytes) = v;
}

template <typename OFF>
__device__ __forceinline__ float2
ld2(const float2 *base, OFF off)
{
    return *(const float2 *)((const char *)base + off);
}

template <typename OFF>
__device__ __forceinline__ void
st2(float2 *base, OFF off, float a, float b)
{
    *(float2 *)((char *)base + off) = make_float2(a, b);
}

template <typename ID, typename OFF>
__device__ __forceinline__ unsigned
ldid(const void *base, OFF off)
{
    return *(const ID *)((const char *)base + off);
}

#endif
