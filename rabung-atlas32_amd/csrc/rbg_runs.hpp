// Wave-level run bookkeeping on 64-position change masks (device code).
//
// A wave holds 64 consecutive positions; one ballot gives the change mask C
// (bit i set when position i starts a new run; bit 0 of the first chunk of a
// stretch is forced).  The scalar state below follows the open run across
// chunks and records the prefix run, the longest closed run with its least
// start, and (at the end) the suffix run.  Every value is wave-uniform, so the
// compiler keeps it in SGPRs and the work runs on the scalar unit.
#pragma once
#include <hip/hip_runtime.h>

#include "rbg_types.hpp"

namespace rbg {

__device__ __forceinline__ void take(u32 len, u32 start, u32 &best, u32 &bstart) {
    if (len > best || (len == best && start < bstart)) {
        best = len;
        bstart = start;
    }
}

__device__ __forceinline__ u64 ballot64(bool v) { return __builtin_amdgcn_ballot_w64(v); }

// Lane i receives v of lane i-1; lane 0 receives `fill` (DPP wave_shr:1).
__device__ __forceinline__ u32 shift_up_one(u32 v, u32 fill) {
    return (u32)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xF, 0xF, false);
}

// Positions i of X with ones at i..i+k-1 (k >= 1): doubling up to s, the
// largest power of two <= k, as straight-line code on wave-uniform
// branches, then one shift by k - s.
__device__ __forceinline__ u64 runs_of_ones(u64 X, u32 k) {
    if (k >= 2) {
        X &= X >> 1;
        if (k >= 4) {
            X &= X >> 2;
            if (k >= 8) {
                X &= X >> 4;
                if (k >= 16) {
                    X &= X >> 8;
                    if (k >= 32) X &= X >> 16;
                }
            }
        }
    }
    if (X == 0) return 0;
    return X & (X >> (k - (1u << (31 - __builtin_clz(k)))));
}

// Prefix maximum over the wave's lanes (lane i gets the maximum of lanes
// 0..i): four in-row shifts and two row broadcasts, one V_MAX_U32 with DPP
// each.
template <int CTRL, int ROWS>
__device__ __forceinline__ u32 max_dpp(u32 v) {
    return max(v, (u32)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, true));
}
__device__ __forceinline__ u32 prefix_max64(u32 v) {
    v = max_dpp<0x111, 0xF>(v);  // row_shr:1
    v = max_dpp<0x112, 0xF>(v);  // row_shr:2
    v = max_dpp<0x114, 0xF>(v);  // row_shr:4
    v = max_dpp<0x118, 0xF>(v);  // row_shr:8
    v = max_dpp<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = max_dpp<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Whether the same-as-left mask N of the wave holds k >= 1 consecutive ones,
// computed on the vector unit: zp is lane + 1 on the lanes whose bit of N is
// clear and 0 elsewhere, so after the prefix maximum lane i holds one past
// the last clear bit at or below i, and i + 1 - zp ones end at lane i.  The
// same answer as runs_of_ones(N, k) != 0, for a kernel whose scalar unit is
// the busier one.
__device__ __forceinline__ bool has_ones_run(u32 zp, u32 k) {
    const u32 lane1 = (threadIdx.x & 63u) + 1u;
    return __builtin_amdgcn_ballot_w64(prefix_max64(zp) + k <= lane1) != 0ull;
}

// Longest interior run of one 64-position chunk.  N: same-as-left mask of the
// valid positions; f < h are the first and last change bits.  Interior runs
// start at the change bits in [f, h) and close before h; they start after
// every run already seen, so only a run strictly longer than `best` matters.
// `cand` false: the chunk holds no `kq` consecutive ones of N for some
// 1 <= kq <= best, found beforehand (no run can beat best).
__device__ __forceinline__ void interior_best(u64 N, u32 f, u32 h, u32 base, u32 &best,
                                              u32 &bstart, bool cand = true) {
    if (best > 62) return;
    if (!cand) return;
    // quick reject on the whole chunk (the interior mask only narrows N):
    // once best is large, most chunks hold no best consecutive equal pairs
    if (best != 0 && runs_of_ones(N, best) == 0) return;
    const u64 between = ((1ull << h) - 1ull) & ~((2ull << f) - 1ull);
    const u64 T = N & between;
    u64 X;
    u32 l;
    if (best == 0) {
        if (T == 0) {  // every interior run has length 1; the earliest starts at f
            take(1, base + f, best, bstart);
            return;
        }
        X = T;
        l = 1;
    } else {
        // X: positions i with ones at i..i+best-1 (a run of length >= best+1)
        X = runs_of_ones(T, best);
        if (X == 0) return;
        l = best;
    }
    for (;;) {
        const u64 Y = X & (T >> l);
        if (Y == 0) break;
        X = Y;
        ++l;
    }
    // ones at c..c+l-1 with c = ctz(X): the colour run starts at c-1, length l+1
    const u32 c = (u32)__builtin_ctzll(X);
    take(l + 1, base + c - 1, best, bstart);
}

// prefix == 0 while the run that began at relative position 0 is still
// open: a closed head run has length >= 1, so no separate flag is kept (one
// scalar register less per order).
struct OrderState {
    u32 prefix, cur, best, bstart;
};

// floor >= 1: runs shorter than `floor` are not recorded in best (it starts
// at floor - 1, a placeholder that any run of length >= floor replaces).  A
// summary whose final best stays below the floor is then not exact, and the
// caller must rescan with floor 1.
__device__ __forceinline__ void order_init(OrderState &s, u32 floor = 1) {
    s.prefix = 0;
    s.cur = 0;
    s.best = floor ? floor - 1u : 0u;
    s.bstart = 0;
}

// Advance the state over one chunk at relative offset `base` with change mask
// C and same mask N (both restricted to the valid positions).
// cand: as for interior_best (true when the caller has not tested the chunk).
__device__ __forceinline__ void order_update(OrderState &s, u64 C, u64 N, u32 base,
                                             bool cand = true) {
    if (C == 0) return;
    const u32 f = (u32)__builtin_ctzll(C);
    const u32 closed = base + f - s.cur;
    if (closed) {
        if (s.prefix == 0) s.prefix = closed;
        take(closed, s.cur, s.best, s.bstart);
    }
    s.cur = base + f;
    const u64 rest = C & (C - 1ull);
    if (rest) {
        const u32 h = 63u - (u32)__builtin_clzll(rest);
        if (s.prefix == 0)  // the head run closes at the second change (first chunk)
            s.prefix = base + (u32)__builtin_ctzll(rest) - s.cur;
        interior_best(N, f, h, base, s.best, s.bstart, cand);
        s.cur = base + h;
    }
}

// Close the stretch of `len` positions: the suffix run (partial) counts for
// `best`; a head run never closed makes the whole stretch the prefix.
__device__ __forceinline__ u32 order_finish(OrderState &s, u32 len) {
    const u32 suffix = len - s.cur;
    if (s.prefix == 0) s.prefix = len;
    if (len) take(suffix, s.cur, s.best, s.bstart);
    return len ? suffix : 0u;
}

__device__ __forceinline__ u32 f4(u32 len) { return len > 3 ? len - 3 : 0; }

// Associative run monoid of a contiguous stretch (the reference's SummOp,
// original_scan_gpu.cu:116-128, with the least-start witness of
// dense_oracle.cu:32-44 and the A4 window count): merge(A, B) for A
// immediately left of B.  `planes` selects the colour bits of the order.
// Positions are absolute.
struct Agg {
    u32 first_x, len, fcode, lcode;
    u32 prefix, suffix, best, bstart, a4;
};

__device__ __forceinline__ Agg merge(const Agg &A, const Agg &B, u32 planes) {
    if (A.len == 0) return B;
    if (B.len == 0) return A;
    Agg C;
    C.first_x = A.first_x;
    C.len = A.len + B.len;
    C.fcode = A.fcode;
    C.lcode = B.lcode;
    C.best = A.best;
    C.bstart = A.bstart;
    take(B.best, B.bstart, C.best, C.bstart);
    const bool eq = ((A.lcode ^ B.fcode) & planes) == 0u;
    if (eq) {
        const u32 cross = A.suffix + B.prefix;
        const u32 cstart = A.first_x + A.len - A.suffix;
        take(cross, cstart, C.best, C.bstart);
        C.prefix = (A.prefix == A.len) ? A.len + B.prefix : A.prefix;
        C.suffix = (B.suffix == B.len) ? B.len + A.suffix : B.suffix;
        C.a4 = A.a4 + B.a4 + f4(cross) - f4(A.suffix) - f4(B.prefix);
    } else {
        C.prefix = A.prefix;
        C.suffix = B.suffix;
        C.a4 = A.a4 + B.a4;
    }
    return C;
}

}  // namespace rbg
