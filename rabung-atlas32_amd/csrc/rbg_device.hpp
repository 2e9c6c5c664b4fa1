// Device arithmetic shared by the dense and sparse kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rbg_types.hpp"

namespace rbg {

// Montgomery product with 32-bit halves (v_mul_lo_u32 / v_mul_hi_u32).  The
// sum u = hi + mh + (lo != 0) reaches 2p > 2^32 for p > 2^31: it is formed
// with one add-with-carry (carry c1 is bit 32 of u), and u >= p iff c1 or
// s - p does not borrow, so the reduction is one subtract-with-borrow and one
// select (8 VALU instructions, 4 of them multiplies).
__device__ __forceinline__ u32 mont_mul(u32 a, u32 b, u32 p, u32 ninv) {
    const u32 lo = a * b;
    const u32 hi = __umulhi(a, b);
    const u32 m = lo * ninv;
    const u32 mh = __umulhi(m, p);
    unsigned int c1, borrow;
    const u32 s = __builtin_addc(hi, mh, lo != 0u ? 1u : 0u, &c1);
    const u32 t = __builtin_subc(s, p, 0u, &borrow);
    return (c1 | (borrow ^ 1u)) ? t : s;
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
