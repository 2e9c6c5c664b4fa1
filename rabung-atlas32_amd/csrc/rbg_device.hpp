// Device arithmetic shared by the dense and sparse kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "rbg_types.hpp"

namespace rbg {

// Montgomery product with 32-bit halves, subtractive form: with
// m = lo(a b) p^{-1} mod 2^32 the low words of a b and m p are equal, so
// REDC(a b) = (a b - m p) / 2^32 = hi(a b) - hi(m p) exactly, a value in
// (-p, p) for a b < p 2^32; a borrow adds p back.  One v_mad_u64_u32, one
// v_mul_lo_u32, one v_mul_hi_u32, a subtract whose borrow lands in a lane
// mask, a select and an add: 6 VALU instructions.  `pinv` is p^{-1} mod
// 2^32 (pos_inverse of the descriptors' -p^{-1}).
__device__ __forceinline__ u32 mont_mul(u32 a, u32 b, u32 p, u32 pinv) {
    const u64 t = (u64)a * b;  // one v_mad_u64_u32
    const u32 lo = (u32)t, hi = (u32)(t >> 32);
    const u32 m = lo * pinv;
    const u32 mh = __umulhi(m, p);
    u32 r;
    const bool borrow = __builtin_sub_overflow(hi, mh, &r);
    return r + (borrow ? p : 0u);
}

// p^{-1} mod 2^32 from the descriptors' -p^{-1}, for a wave-uniform p.  The
// negation goes through readfirstlane so that the compiler cannot see it:
// otherwise it rewrites lo * (0 - ninv) as 0 - lo * ninv, one more
// instruction in every product.
__device__ __forceinline__ u32 pos_inverse(u32 ninv) {
    return __builtin_amdgcn_readfirstlane(0u - ninv);
}

__device__ __forceinline__ u32 mont_pow(u32 base, u32 e, u32 p, u32 pinv, u32 one) {
    if (e == 0) return one;
    u32 y = base;
    const int top = 31 - __builtin_clz(e);
    for (int bit = top - 1; bit >= 0; --bit) {
        y = mont_mul(y, y, p, pinv);
        if ((e >> bit) & 1u) y = mont_mul(y, base, p, pinv);
    }
    return y;
}

__device__ __forceinline__ u32 add_mod(u32 a, u32 b, u32 p) {
    const u64 s = (u64)a + b;
    return (u32)(s >= p ? s - p : s);
}

// Orders the LDS writes of a wave before its following LDS reads of other
// lanes' slots (the wave's LDS instructions execute in order; the fences
// keep the compiler from moving accesses across).
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Index of y in the sorted table vals[0..m) (binary search, fixed trip count
// for a given m so that a wave stays converged).
template <class V>
__device__ __forceinline__ u32 table_index(const V *vals, u32 m, V y) {
    u32 lo = 0, n = m;
    while (n > 1) {
        const u32 half = n >> 1;
        lo = (vals[lo + half] <= y) ? lo + half : lo;
        n -= half;
    }
    return lo;
}

// Montgomery product for primes 2^32 < p < 2^33, still with R = 2^32 (now
// R < p).  Residues a, b < p are a low word and a high bit:
// a b = al bl + (ah bl + bh al) 2^32 + ah bh 2^64.  With m = lo(al bl) n' and
// p = 2^32 + pl, the low word of a b + m p is lo(al bl) + lo(m pl) = 0 mod
// 2^32 (carry [lo(al bl) != 0]) and the exact quotient (a b + m p) / 2^32 is
//     hi(al bl) + hi(m pl) + carry + ah bl + bh al + ah bh 2^32 + m,
// which is below p (p / 2^32 + 1) < 3p: two conditional subtractions
// reduce it (4 multiplies).
__device__ __forceinline__ u64 mont_mul33(u64 a, u64 b, u64 p, u32 ninv) {
    const u32 al = (u32)a, bl = (u32)b, pl = (u32)p;
    const u32 ah = (u32)(a >> 32), bh = (u32)(b >> 32);  // 0 or 1
    const u32 lo = al * bl, hi = __umulhi(al, bl);
    const u32 m = lo * ninv;
    const u32 mh = __umulhi(m, pl);
    u64 u = (u64)hi + mh + m + (lo != 0u ? 1u : 0u);
    u += (u64)(bl & (0u - ah)) + (al & (0u - bh)) + ((u64)(ah & bh) << 32);
    u = u >= p ? u - p : u;
    return u >= p ? u - p : u;
}

__device__ __forceinline__ u64 mont_pow33(u64 base, u64 e, u64 p, u32 ninv, u64 one) {
    if (e == 0) return one;
    u64 y = base;
    const int top = 63 - __builtin_clzll(e);
    for (int bit = top - 1; bit >= 0; --bit) {
        y = mont_mul33(y, y, p, ninv);
        if ((e >> bit) & 1ull) y = mont_mul33(y, base, p, ninv);
    }
    return y;
}

// Residue arithmetic of one prime, as a template argument of the kernels
// that work for both domains: Dom32 (p < 2^32, u32 residues) and Dom33
// (2^32 < p < 2^33, u64 residues).  Both use R = 2^32.
struct Dom32 {
    typedef u32 V;
    static constexpr bool wide = false;
    u32 p, pinv;  // pinv = p^{-1} mod 2^32 (pos_inverse)
    __device__ __forceinline__ V mul(V a, V b) const { return mont_mul(a, b, p, pinv); }
    __device__ __forceinline__ V add(V a, V b) const { return add_mod(a, b, p); }
    __device__ __forceinline__ V pow(V base, u64 e, V one) const {
        return mont_pow(base, (u32)e, p, pinv, one);
    }
};

struct Dom33 {
    typedef u64 V;
    static constexpr bool wide = true;
    u64 p;
    u32 ninv;
    __device__ __forceinline__ V mul(V a, V b) const { return mont_mul33(a, b, p, ninv); }
    __device__ __forceinline__ V add(V a, V b) const {
        const u64 s = a + b;
        return s >= p ? s - p : s;
    }
    __device__ __forceinline__ V pow(V base, u64 e, V one) const {
        return mont_pow33(base, e, p, ninv, one);
    }
};

}  // namespace rbg
