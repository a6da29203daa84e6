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

// 1 / x rounded to nearest, for x in [2^-700, 2^700]: the hardware
// reciprocal estimate, two Newton-Raphson refinements e = 1 - x r,
// r <- r + r e, and Markstein's final correction r + r (1 - x r) (P. Markstein,
// "IA-64 and Elementary Functions", 2000, ch. 8).  These are the steps of the
// IEEE double division that hipcc emits for 1.0 / x; its v_div_scale /
// v_div_fmas / v_div_fixup wrappers only rescale operands near the ends of the
// exponent range or patch special values, and are the identity in this range,
// so the result is the same correctly rounded quotient bit for bit in four
// fewer instructions.  Outside the range (x = 0 included) the result is
// meaningless: callers use it only where x is a squared distance below a
// cutoff, or discard it.
__device__ __forceinline__ double recip_rn(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(e, r, r);
}
