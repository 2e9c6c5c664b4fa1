// Device arithmetic shared by the dense and sparse kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rbg_types.hpp"

namespace rbg {

// Montgomery product with 32-bit halves (v_mul_lo_u32 / v_mul_hi_u32); the
// carry of hi + mh + (lo != 0) is kept because the sum reaches 2p > 2^32 for
// p > 2^31.
__device__ __forceinline__ u32 mont_mul(u32 a, u32 b, u32 p, u32 ninv) {
    const u32 lo = a * b;
    const u32 hi = __umulhi(a, b);
    const u32 m = lo * ninv;
    const u32 mh = __umulhi(m, p);
    const u64 s = (u64)hi + mh + (lo != 0u);
    return (u32)(s >= p ? s - p : s);
}

__device__ __forceinline__ u32 mont_pow(u32 base, u32 e, u32 p, u32 ninv, u32 one) {
    if (e == 0) return one;
    u32 y = base;
    const int top = 31 - __builtin_clz(e);
    for (int bit = top - 1; bit >= 0; --bit) {
        y = mont_mul(y, y, p, ninv);
        if ((e >> bit) & 1u) y = mont_mul(y, base, p, ninv);
    }
    return y;
}

__device__ __forceinline__ u32 add_mod(u32 a, u32 b, u32 p) {
    const u64 s = (u64)a + b;
    return (u32)(s >= p ? s - p : s);
}

// Index of y in the sorted table vals[0..m) (binary search, fixed trip count
// for a given m so that a wave stays converged).
__device__ __forceinline__ u32 table_index(const u32 *vals, u32 m, u32 y) {
    u32 lo = 0, n = m;
    while (n > 1) {
        const u32 half = n >> 1;
        lo = (vals[lo + half] <= y) ? lo + half : lo;
        n -= half;
    }
    return lo;
}

}  // namespace rbg
