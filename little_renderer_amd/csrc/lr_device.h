// lr_device.h - device helpers shared by the HIP kernels (gfx950, wave64).
//
// Floating point: the library is compiled with -ffp-contract=off, so every expression rounds
// as written; min / max follow WGSL's rule (e2 if e2 < e1, else e1), the rule the reference's
// builder shaders use, so that boxes are the same to the bit whatever the order of the joins.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define LR_LEAF 0x80000000u
#define LR_NONE 0xffffffffu

// loads through the global address space, so that they are global loads, not flat ones (which
// loads through the walks' pointers, read from job records, would otherwise be)
typedef float lr_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ldg4(const float4 *p, size_t i)
{
    const lr_f4v v = ((const __attribute__((address_space(1))) lr_f4v *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint32_t ldg(const uint32_t *p, size_t i)
{
    return ((const __attribute__((address_space(1))) uint32_t *)p)[i];
}

// one BVH node as the reference lays it out (trace.wgsl:2): 32 bytes
struct alignas(16) LrNode {
    float lox, loy, loz;
    uint32_t a;
    float hix, hiy, hiz;
    uint32_t b;
};

__device__ __forceinline__ float wmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float wmax(float a, float b) { return b > a ? b : a; }
// r_fmin / r_fmax of lrmath.h:56-57 (the other number when b is a NaN)
__device__ __forceinline__ float rfmin(float a, float b) { return b != b ? a : (a < b ? a : b); }
__device__ __forceinline__ float rfmax(float a, float b) { return b != b ? a : (a > b ? a : b); }

// device-scope atomics on one word: flag and payload travel together, so relaxed order is
// enough and no fence is needed around them
__device__ __forceinline__ uint32_t ld_dev(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_dev64(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev64(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }
