// march.hpp -- device helpers shared by the marching stencil kernels (FD step,
// separable SEM step, batched FD shots).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace phc {

// Workgroup barrier that orders LDS only: global loads issued before it stay
// in flight (a plain __syncthreads() also waits for them).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders one wavefront's LDS accesses across its lanes (a row one lane wrote
// and another reads): compiler fences at wavefront scope around a
// wave-level scheduling barrier; no workgroup barrier.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

__device__ __forceinline__ float mul_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a * b;
}

__device__ __forceinline__ float add_nc(float a, float b)
{
#pragma clang fp contract(off)
    return a + b;
}

// Load / store through a kernel-argument base pointer.  With 32-bit offsets
// the offset is in bytes, so the access is scalar base + 32-bit vector offset.
template <typename IT, typename T = float>
__device__ __forceinline__ T ldo(const T *base, IT off)
{
    if constexpr (sizeof(IT) == 4) return *(const T *)((const char *)base + off);
    else return base[off];
}
template <typename IT, typename T = float>
__device__ __forceinline__ T ldo_nt(const T *base, IT off)
{
    if constexpr (sizeof(IT) == 4) return __builtin_nontemporal_load((const T *)((const char *)base + off));
    else return __builtin_nontemporal_load(base + off);
}
template <typename IT, typename T = float>
__device__ __forceinline__ void sto_nt(T *base, IT off, T v)
{
    if constexpr (sizeof(IT) == 4) __builtin_nontemporal_store(v, (T *)((char *)base + off));
    else __builtin_nontemporal_store(v, base + off);
}

// XCD-aware block order: hardware deals consecutive block ids round-robin
// over the 8 XCDs; give each XCD a contiguous range of tiles so halo re-reads
// hit its own L2.  XCD x receives the hardware ids x, x+8, ...: nb/8 + 1 of
// them for x < nb%8, nb/8 otherwise; the map is a bijection of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, l = b >> 3;
    return (x < r) ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

} // namespace phc
