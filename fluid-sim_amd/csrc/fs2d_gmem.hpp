// fs2d_gmem.hpp -- global-memory pointers, hand-off atomics and bit casts
// shared by the sweep engine (fs2d_sweep.hpp) and the kernels.
//
// Hand-off words between the strips of a sweep are 64-bit {payload, epoch}
// pairs written and read with agent-scope relaxed atomics, so a consumer that
// sees the current epoch also sees the payload; no fences are needed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fs2d {

constexpr int kSpinLimit = 1 << 16;  // polls of one memory round trip each: well under 1 s

// Pointers typed as global memory.  Accesses through them compile to GLOBAL
// instructions, counted by vmcnt alone.  A pointer read from memory (the
// batched kernels' descriptors) is otherwise generic, and FLAT accesses also
// count in lgkmcnt, so every LDS or scalar-load wait would wait for them too.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__host__ __device__ inline gptr<T> G(T* p) {
    return (gptr<T>)p;
}
// int4 / float4 layouts as builtin vectors, which load and store through gptr
// (the HIP vector classes' copy operators expect generic pointers)
typedef int i4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned long long atomic_ld64(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_ld64(gptr<const unsigned long long> p) {
    return __hip_atomic_load(const_cast<gptr<unsigned long long>>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_st64(gptr<unsigned long long> p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned readlane_u32(unsigned v, int lane) {
    return (unsigned)__builtin_amdgcn_readlane((int)v, lane);
}

// Bit casts for carried state.
__device__ __forceinline__ uint32_t f2u(float x) { return __float_as_uint(x); }
__device__ __forceinline__ float u2f(uint32_t x) { return __uint_as_float(x); }
__device__ __forceinline__ void d2u(double x, uint32_t* w) {
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    w[0] = (uint32_t)b;
    w[1] = (uint32_t)(b >> 32);
}
__device__ __forceinline__ double u2d(const uint32_t* w) {
    return __longlong_as_double((long long)(((unsigned long long)w[1] << 32) | w[0]));
}

}  // namespace fs2d
