// Shared device helpers for the gfx950 particle kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>


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

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
