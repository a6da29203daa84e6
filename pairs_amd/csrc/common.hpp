// Shared device helpers for the gfx950 particle kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PAIRS_WAVE 64

// Blocks of one launch are handed to the 8 XCDs round-robin (block b runs on
// XCD b % 8).  Remapping so that XCD x processes one contiguous range of
// logical blocks keeps each XCD's L2 working set to its own slab of particles.
__device__ __forceinline__ int xcd_block(int bid, int nblocks) {
    const int q = nblocks >> 3, r = nblocks & 7;
    const int x = bid & 7, k = bid >> 3;
    return x * q + (x < r ? x : r) + k;
}

// number of set bits of a wave mask below this lane
__device__ __forceinline__ int lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned) (mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned) mask, 0));
}

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

// type id kept in the 4th lane of a position (bit pattern, never used in arithmetic)
__device__ __host__ __forceinline__ int pos_type(const double4 &p) {
#ifdef __HIP_DEVICE_COMPILE__
    return (int) __double_as_longlong(p.w);
#else
    union { double d; long long l; } u; u.d = p.w; return (int) u.l;
#endif
}
__device__ __host__ __forceinline__ double type_bits(int t) {
#ifdef __HIP_DEVICE_COMPILE__
    return __longlong_as_double((long long) t);
#else
    union { double d; long long l; } u; u.l = (long long) t; return u.d;
#endif
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void atomic_add_f64(double *p, double v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
