// Exact 32-bit modular arithmetic shared by host and device code.
//
// Montgomery multiplication with R = 2^32 and n' = -p^{-1} mod 2^32 for odd
// p < 2^32 (reference: src/census/engine32/src/mont32.hpp:9-10 and
// src/census/engine32/build/original_scan_gpu.cu:60-77).  For p > 2^31 the
// sum t + m*p can exceed 2^64, so the carry out of the high word is kept and
// folded into the final conditional subtraction.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define RBG_HD __host__ __device__ __forceinline__
#else
#define RBG_HD inline
#endif

namespace rbg {

typedef uint32_t u32;
typedef uint64_t u64;

// -p^{-1} mod 2^32 by Newton iteration (p odd).
RBG_HD u32 neg_inverse32(u32 p) {
    u32 x = p;  // correct to 3 bits for odd p
    for (int i = 0; i < 5; ++i) x *= 2u - p * x;
    return 0u - x;
}

// REDC(a*b) = a*b*2^-32 mod p for a*b < p*2^32; result in [0, p).
RBG_HD u32 montmul(u32 a, u32 b, u32 p, u32 ninv) {
    u64 t = (u64)a * b;
    u32 lo = (u32)t;
    u32 m = lo * ninv;
    u64 mp = (u64)m * p;
    // t + mp has zero low word; its carry into the high word is (lo != 0).
    u64 hi = (t >> 32) + (mp >> 32) + (lo != 0u);  // < 2p < 2^33
    return (u32)(hi >= p ? hi - p : hi);
}

RBG_HD u32 addmod(u32 a, u32 b, u32 p) {
    u64 s = (u64)a + b;
    return (u32)(s >= p ? s - p : s);
}

// Left-to-right binary power in the Montgomery domain; base and result are
// Montgomery forms, `one` is R mod p.
RBG_HD u32 montpow(u32 base, u32 e, u32 p, u32 ninv, u32 one) {
    if (e == 0) return one;
    u32 y = base;
    int top = 31 - __builtin_clz(e);
    for (int bit = top - 1; bit >= 0; --bit) {
        y = montmul(y, y, p, ninv);
        if ((e >> bit) & 1u) y = montmul(y, base, p, ninv);
    }
    return y;
}

// Plain (non-Montgomery) helpers for the host: p < 2^32 so products fit u64.
RBG_HD u32 mulmod64(u32 a, u32 b, u32 p) { return (u32)((u64)a * b % p); }

RBG_HD u32 powmod64(u32 a, u64 e, u32 p) {
    u64 v = 1 % p, x = a % p;
    while (e) {
        if (e & 1) v = v * x % p;
        x = x * x % p;
        e >>= 1;
    }
    return (u32)v;
}

// Montgomery product count of one left-to-right power: bitlen(e)+popcount(e)-2
// (original_scan_gpu.cu:383).
RBG_HD u32 montpow_products(u32 e) {
    if (e < 2) return 0;
    return (32 - __builtin_clz(e)) + __builtin_popcount(e) - 2;
}

}  // namespace rbg
