// Certified rational-anchor sparse search for gfx950.
//
// Reference: paper/GPU_METHOD_EN.tex:246-568 (integer cover certificate,
// reflection-compatible anchor schedule, conservative rejection, exact
// completion) and its CUDA realisation src/census/engine32/src/*.  This is an
// independent design for wave64:
//
//   * cover: B, D certified with exact integers (4BD - HD^2 > 4p, HD <= 2B);
//     every window of Q positions has numerators a + jd, |a + jd| <= B, d <= D
//     (reference clipped_cover.hpp:7-21, GPU_METHOD_EN.tex:254-300);
//   * table: labels j(y) of y = 1..B (one Montgomery power each) and, per
//     order r, colour bit planes over the signed range [-B, B] with
//     cl(-y) = cl(y) + cl(-1);
//   * seed: one wave scans 1..B in order with the ballot run machinery; runs
//     there are real runs and give the first threshold;
//   * filter: one wave per (order, denominator d) task.  The scheduled anchors
//     b (b = n d + c, n = phi or t-2-phi mod s by the side of c, folded to
//     b >= (t-1)d/2) form bands of consecutive b; each lane tests 64
//     consecutive b of a band at once: for every plane, the 64-bit windows at
//     numerators b - j d (j = 1..t-1) are compared with the window at b, and a
//     lane stops as soon as no candidate is left;
//   * completion: a surviving anchor is mapped to x = -b / d mod p and
//     extended left and right through the table (numerators stay in [-B, B]
//     for the short runs that occur) and, past the table, by direct powers;
//     the run (length, least reflected start) is folded into a packed 64-bit
//     key with one atomicMax;
//   * thresholds are read per task from the running best: any q with
//     3 <= q <= min(best, Q) preserves every run of maximal length, so no
//     epoch barrier is needed.
#include <hip/hip_runtime.h>

#include "rabung.h"
#include "rbg_device.hpp"
#include "rbg_runs.hpp"
#include "rbg_sparse.hpp"
#include "rbg_types.hpp"

namespace rbg {

__device__ __forceinline__ u32 find_prime(const SparsePrime *P, u32 n, u32 blk, int which) {
    u32 lo = 0, hi = n;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        const u32 b = which == 0 ? P[mid].lblk : which == 1 ? P[mid].iblk : P[mid].pblk;
        if (b <= blk) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr u32 kLabelsPerThread = 8;

// labels[lab_off + y] = j(y) for y in [1, B]; slot 0 holds 0.
__global__ __launch_bounds__(256) void sp_labels(const SparsePrime *__restrict__ P, u32 nprimes,
                                                 const u32 *__restrict__ tab_val,
                                                 const u16 *__restrict__ tab_lab,
                                                 u16 *__restrict__ labels, u32 *__restrict__ err) {
    __shared__ u32 s_val[kMaxM];
    __shared__ u16 s_lab[kMaxM];
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 0));
    const SparsePrime d = P[pi];
    for (u32 j = threadIdx.x; j < d.m; j += blockDim.x) {
        s_val[j] = tab_val[d.tab_off + j];
        s_lab[j] = tab_lab[d.tab_off + j];
    }
    __syncthreads();
    const u32 y0 = 1u + (blockIdx.x - d.lblk) * (256u * kLabelsPerThread);
    if (blockIdx.x == d.lblk && threadIdx.x == 0) labels[d.lab_off] = 0;
    u32 bad = 0;
    for (u32 k = 0; k < kLabelsPerThread; ++k) {
        const u32 y = y0 + k * 256u + threadIdx.x;
        if (y > d.B) break;
        const u32 v = mont_pow(mont_mul(y, d.r2, d.p, d.ninv), d.e, d.p, d.ninv, d.one);
        const u32 idx = table_index(s_val, d.m, v);
        bad |= s_val[idx] != v;
        labels[(u64)d.lab_off + y] = s_lab[idx];
    }
    if (bad) atomicOr(err, 1u);
}

// inv[inv_off + d] = Montgomery form d^{-1} R mod p for d in [1, D] (d^(p-2)).
__global__ __launch_bounds__(256) void sp_inverses(const SparsePrime *__restrict__ P, u32 nprimes,
                                                   u32 *__restrict__ inv) {
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 1));
    const SparsePrime s = P[pi];
    const u32 d = (blockIdx.x - s.iblk) * 256u + threadIdx.x;
    if (d == 0 || d > s.D) return;
    const u32 dm = mont_mul(d, s.r2, s.p, s.ninv);
    const u32 im = mont_pow(dm, s.p - 2, s.p, s.ninv, s.one);
    inv[(u64)s.inv_off + d] = im;
}

constexpr u32 kPlanePairsPerWave = 16;

// Colour bit planes of one order over the signed table: bit i of plane k is
// bit k of cl(i - B).  One wave builds 64 bits (two words) per step.
__global__ __launch_bounds__(256) void sp_planes(const SparsePrime *__restrict__ P, u32 nprimes,
                                                 const SparseProf *__restrict__ prof, u32 nprof,
                                                 const u16 *__restrict__ labels,
                                                 u32 *__restrict__ planes) {
    // block -> profile through the per-profile block offsets kept in pblk of
    // the profile records (searched linearly from the prime's first profile)
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 2));
    const SparsePrime s = P[pi];
    u32 k = s.prof_begin;
    while (k + 1 < s.prof_begin + s.nprof && prof[k + 1].pblk <= blockIdx.x) ++k;
    const SparseProf f = prof[k];
    const u32 lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const u32 pairs = (f.wpp + 1) / 2;
    const u32 first = ((blockIdx.x - f.pblk) * 4u + wave) * kPlanePairsPerWave;
    for (u32 q = 0; q < kPlanePairsPerWave; ++q) {
        const u32 pair = first + q;
        if (pair >= pairs) break;
        const int64_t y = (int64_t)pair * 64 + lane - (int64_t)s.B;
        u32 c = 0;
        if (y > 0 && y <= (int64_t)s.B) {
            c = labels[(u64)s.lab_off + (u64)y] % f.r;
        } else if (y < 0 && -y <= (int64_t)s.B) {
            c = ((u32)labels[(u64)s.lab_off + (u64)(-y)] + s.jneg) % s.m % f.r;
        }
        for (u32 b = 0; b < f.np; ++b) {
            const u64 bits = ballot64((c >> b) & 1u);
            if (lane == 0) {
                u32 *pl = planes + f.plane_off + (u64)b * f.wpp;
                pl[2 * pair] = (u32)bits;
                if (2 * pair + 1 < f.wpp) pl[2 * pair + 1] = (u32)(bits >> 32);
            }
        }
    }
}

constexpr u32 kSeedSeg = 4096;  // positions of 1..B per seed wave

// Seed runs of 1..B (real runs of the sequence): one wave per segment of
// kSeedSeg positions writes the segment's run summary (Agg over colour values).
__global__ __launch_bounds__(256) void sp_seed_scan(const SparsePrime *__restrict__ P,
                                                    const SparseProf *__restrict__ prof,
                                                    u32 nprof, const u16 *__restrict__ labels,
                                                    Agg *__restrict__ segs, u32 segs_per_prof) {
    const u32 wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const u32 k = __builtin_amdgcn_readfirstlane(wave / segs_per_prof);
    const u32 sg = __builtin_amdgcn_readfirstlane(wave % segs_per_prof);
    if (k >= nprof) return;
    const SparseProf f = prof[k];
    const SparsePrime s = P[f.prime];
    const u32 lane = threadIdx.x & 63u;
    const u32 first = 1u + sg * kSeedSeg;
    const u32 len = first > s.B ? 0u : min(kSeedSeg, s.B - first + 1u);
    OrderState st;
    order_init(st);
    u32 carry = 0, fcol = 0, lcol = 0;
    const u32 nchunks = (len + 63u) >> 6;
    for (u32 c = 0; c < nchunks; ++c) {
        const u32 base = c << 6;
        const u32 valid = min(64u, len - base);
        const u64 vmask = valid == 64u ? ~0ull : ((1ull << valid) - 1ull);
        const u32 y = first + base + lane;
        const u32 col = y <= s.B ? labels[(u64)s.lab_off + y] % f.r : 0u;
        const u32 left = shift_up_one(col, carry);
        carry = __builtin_amdgcn_readlane(col, 63);
        if (c == 0) fcol = __builtin_amdgcn_readfirstlane(col);
        lcol = __builtin_amdgcn_readlane(col, valid - 1);
        u64 C = ballot64(col != left);
        if (c == 0) C |= 1ull;
        C &= vmask;
        order_update(st, C, ~C & vmask, base);
    }
    const u32 suffix = order_finish(st, len);
    if (lane == 0) {
        Agg a;
        a.first_x = first;
        a.len = len;
        a.fcode = fcol;
        a.lcode = lcol;
        a.prefix = st.prefix;
        a.suffix = suffix;
        a.best = st.best;
        a.bstart = st.bstart + first;
        a.a4 = 0;
        segs[(u64)k * segs_per_prof + sg] = a;
    }
}

// One thread per profile: ordered merge of its segment summaries; the key of
// the longest seed run (partial end run included) and the prefix L.
__global__ void sp_seed_merge(const SparsePrime *__restrict__ P, SparseProf *__restrict__ prof,
                              u32 nprof, const Agg *__restrict__ segs, u32 segs_per_prof,
                              unsigned long long *__restrict__ best, u32 *__restrict__ failed) {
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nprof) return;
    const SparseProf f = prof[k];
    const SparsePrime s = P[f.prime];
    Agg acc;
    acc.len = 0;
    for (u32 i = 0; i < segs_per_prof; ++i) acc = merge(acc, segs[(u64)k * segs_per_prof + i], ~0u);
    // runs inside 1..B < p/2 have reflected start = start
    best[k] = ((unsigned long long)acc.best << 32) | (0xFFFFFFFFu - acc.bstart);
    const bool whole = acc.prefix == s.B;  // the identity run may continue past B
    prof[k].prefix = acc.prefix;
    if (whole || acc.best < 3u || acc.len != s.B || acc.first_x != 1u) atomicOr(&failed[f.prime], 1u);
}

__device__ __forceinline__ u32 colour_at(const u32 *pl, u32 wpp, u32 np, int64_t a, u32 B) {
    const u32 i = (u32)(a + (int64_t)B);
    u32 c = 0;
    for (u32 b = 0; b < np; ++b) c |= ((pl[(u64)b * wpp + (i >> 5)] >> (i & 31u)) & 1u) << b;
    return c;
}

// Absolute colour of position z by a direct power (past the table).
__device__ u32 colour_direct(const SparsePrime &s, u32 r, const u32 *tab_val,
                             const u16 *tab_lab, u32 z) {
    const u32 v = mont_pow(mont_mul(z, s.r2, s.p, s.ninv), s.e, s.p, s.ninv, s.one);
    u32 lo = 0, n = s.m;
    const u32 *vals = tab_val + s.tab_off;
    while (n > 1) {
        const u32 half = n >> 1;
        lo = (vals[lo + half] <= v) ? lo + half : lo;
        n -= half;
    }
    return tab_lab[s.tab_off + lo] % r;
}

constexpr u32 kExtendLimit = 1u << 22;

// Extend the monochromatic window [x, x+t-1] (numerators -b + i d) to its run.
__device__ void complete_anchor(const SparsePrime &s, const SparseProf &f, const u32 *pl, u32 b,
                                u32 d, u32 dinv, u32 t, const u32 *tab_val, const u16 *tab_lab,
                                unsigned long long *best_slot, u32 *failed) {
    const u32 p = s.p;
    // dinv is the Montgomery form d^{-1} R, so REDC(b * dinv) = b / d mod p
    const u32 bd = mont_mul(b, dinv, p, s.ninv);
    const u32 x = p - bd;  // bd != 0 because 0 < b <= B < p
    const int64_t B = s.B;
    const u32 c0 = colour_at(pl, f.wpp, f.np, -(int64_t)b, s.B);
    u32 abs0 = 0xFFFFFFFFu;
    u32 steps = 0;
    // left
    u32 lo = x;
    int64_t a = -(int64_t)b - d;
    bool table = true;
    while (lo > 1) {
        bool same;
        if (table && a >= -B) {
            same = colour_at(pl, f.wpp, f.np, a, s.B) == c0;
        } else {
            table = false;
            if (abs0 == 0xFFFFFFFFu) abs0 = colour_direct(s, f.r, tab_val, tab_lab, x);
            same = colour_direct(s, f.r, tab_val, tab_lab, lo - 1) == abs0;
        }
        if (!same || ++steps > kExtendLimit) break;
        --lo;
        a -= d;
    }
    // right
    u32 hi = x + t - 1u;
    a = -(int64_t)b + (int64_t)t * d;
    table = true;
    while (hi < p - 1u) {
        bool same;
        if (table && a <= B) {
            same = colour_at(pl, f.wpp, f.np, a, s.B) == c0;
        } else {
            table = false;
            if (abs0 == 0xFFFFFFFFu) abs0 = colour_direct(s, f.r, tab_val, tab_lab, x);
            same = colour_direct(s, f.r, tab_val, tab_lab, hi + 1u) == abs0;
        }
        if (!same || ++steps > kExtendLimit) break;
        ++hi;
        a += d;
    }
    if (steps > kExtendLimit) {
        atomicOr(&failed[f.prime], 2u);
        return;
    }
    const u32 len = hi - lo + 1u;
    const u32 w = lo < p - hi ? lo : p - hi;
    atomicMax(best_slot, ((unsigned long long)len << 32) | (0xFFFFFFFFu - w));
}

// 64-bit window of plane `pl` starting at signed table index i (i >= 0).
__device__ __forceinline__ u64 window64(const u32 *pl, u32 i) {
    const u32 w = i >> 5, sh = i & 31u;
    const u32 a = pl[w], b = pl[w + 1], c = pl[w + 2];
    const u32 lo = __builtin_amdgcn_alignbit(b, a, sh);
    const u32 hi = __builtin_amdgcn_alignbit(c, b, sh);
    return ((u64)hi << 32) | lo;
}

constexpr u64 kThirds = 0x9249249249249249ull;  // bits i = 0 mod 3
constexpr u32 kXcds = 8;  // XCDs of an MI355X

// Bilateral screen (every placement of a q-window containing the anchor;
// reference lemma GPU_METHOD_EN.tex:457-480, vdw_cut::screen): with
// h = q - t extra positions needed, a candidate anchor is kept iff it has
// L matching positions to its left and R to its right with L + R >= h.  In
// numerator space (negated, so equality is unchanged) position x-k reads
// b + k d and position x+t-1+k reads b - (t-1+k) d; numerators outside
// [-B, B] are unknown and never reject.  R is kept as a bit-sliced count
// (5 planes, h <= 23) so that every candidate of the word is screened at once.
__device__ __forceinline__ u64 match_left(const u32 *pl, u32 wpp, u32 np, const u64 *base,
                                          u32 b0, u32 B, u32 d, u32 k, u64 &n_windows) {
    const int64_t lo = (int64_t)b0 + (int64_t)k * d, Bs = B;
    if (lo > Bs) return ~0ull;
    u64 mm = 0;
    const u32 idx = (u32)(lo + Bs);
    for (u32 pb = 0; pb < np; ++pb) mm |= base[pb] ^ window64(pl + (u64)pb * wpp, idx);
    n_windows += np;
    const int64_t known = Bs - lo + 1;  // bits 0..known-1 lie in the table
    if (known < 64) mm &= (1ull << known) - 1ull;
    return ~mm;
}

__device__ __forceinline__ u64 match_right(const u32 *pl, u32 wpp, u32 np, const u64 *base,
                                           u32 b0, u32 B, u32 d, u32 t, u32 k, u64 &n_windows) {
    const int64_t lo = (int64_t)b0 - (int64_t)(t - 1 + k) * d, Bs = B;
    if (lo + 63 < -Bs) return ~0ull;
    u64 mm = 0;
    if (lo >= -Bs) {
        const u32 idx = (u32)(lo + Bs);
        for (u32 pb = 0; pb < np; ++pb) mm |= base[pb] ^ window64(pl + (u64)pb * wpp, idx);
    } else {
        const u32 shift = (u32)(-Bs - lo);  // 1..63 low bits are below -B: unknown
        for (u32 pb = 0; pb < np; ++pb) mm |= base[pb] ^ (window64(pl + (u64)pb * wpp, 0) << shift);
        mm &= ~((1ull << shift) - 1ull);
    }
    n_windows += np;
    return ~mm;
}

__device__ __forceinline__ u64 screen(const u32 *pl, u32 wpp, u32 np, const u64 *base, u32 b0,
                                      u32 B, u32 d, u32 t, u32 h, u64 &n_windows) {
    // R: consecutive right matches per candidate, bit-sliced (r0 = LSB)
    u64 r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0, cr = ~0ull;
    for (u32 k = 1; k <= h; ++k) {
        cr &= match_right(pl, wpp, np, base, b0, B, d, t, k, n_windows);
        if (!cr) break;
        u64 c = cr, x;
        x = r0 & c; r0 ^= c; c = x;
        x = r1 & c; r1 ^= c; c = x;
        x = r2 & c; r2 ^= c; c = x;
        x = r3 & c; r3 ^= c; c = x;
        r4 ^= c;
    }
    u64 pass = 0, cl = ~0ull;
    for (u32 k = 0; k <= h; ++k) {
        if (k) {
            cl &= match_left(pl, wpp, np, base, b0, B, d, k, n_windows);
            if (!cl) break;
        }
        // candidates with R >= h - k (compare the sliced count with a constant)
        const u32 need = h - k;
        const u64 rb[5] = {r0, r1, r2, r3, r4};
        u64 gt = 0, eq = ~0ull;
        for (int bit = 4; bit >= 0; --bit) {
            if ((need >> bit) & 1u) {
                eq &= rb[bit];
            } else {
                gt |= eq & rb[bit];
                eq &= ~rb[bit];
            }
        }
        pass |= cl & (gt | eq);
    }
    return pass;
}

// Anchor test of one 64-candidate word: for j = 1..t-1 the windows at
// numerators b - j d must equal the base windows.  Stages are issued in
// batches (8, 4 or 2 stages for 1, 2 or more planes; all loads of a batch
// before any wait, the base windows with the first batch), so a word costs
// one memory round trip per batch instead of one per stage; the batch
// stops early once no candidate is left.  Stages past t-1 re-read stage t-1
// (harmless, keeps every address inside the table).
template <int NP>
__device__ __forceinline__ u64 anchor_test(const u32 *pl, u32 wpp, u64 *base, u32 ib, u32 d, u32 t,
                                           u64 valid, u64 &n_windows) {
    constexpr int G = NP == 1 ? 8 : NP == 2 ? 4 : 2;  // stages per batch (register budget)
    u64 w[G][NP];
    // the base windows travel with the first batch: one round trip
#pragma unroll
    for (int pb = 0; pb < NP; ++pb) base[pb] = window64(pl + (u64)pb * wpp, ib);
    u64 mism = 0;
    for (u32 j0 = 1; j0 < t; j0 += G) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const u32 j = min(j0 + (u32)g, t - 1u);
            const u32 io = ib - j * d;
#pragma unroll
            for (int pb = 0; pb < NP; ++pb) w[g][pb] = window64(pl + (u64)pb * wpp, io);
        }
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int pb = 0; pb < NP; ++pb) mism |= base[pb] ^ w[g][pb];
        n_windows += (u64)NP * min((u32)G, t - j0);
        if ((valid & ~mism) == 0) break;
    }
    return valid & ~mism;
}

struct FilterCtx {
    const SparsePrime *s;
    const SparseProf *f;
    const u32 *pl;
    const u32 *tab_val;
    const u16 *tab_lab;
    unsigned long long *best;
    u32 *failed;
    u32 d, dinv, t, q, B, bmin;
};

// All scheduled words of one (profile, denominator) task for NP planes.
template <int NP>
__device__ void filter_item(const FilterCtx &c, u32 sp, u32 phi, u64 &n_windows, u64 &n_words,
                            u64 &n_done) {
    const u32 lane = threadIdx.x & 63u;
    const u32 d = c.d, t = c.t, B = c.B, bmin = c.bmin, wpp = c.f->wpp;
    // band layout: kind 0 = c in [1, ca] (2c < d), kind 1 = c in [cb, d-1]
    // (2c > d); for d <= 2 one unscheduled band covers [max(bmin, .), B].
    for (u32 kind = 0; kind < 2; ++kind) {
        u32 c_lo, c_hi, n0, nb, step;
        if (d <= 2) {
            if (kind) break;
            const u32 lo_b = max(bmin, d == 1 ? t : t - 1u);
            if (lo_b > B) break;
            c_lo = lo_b;
            c_hi = B;
            n0 = 0;
            nb = 1;
            step = 0;
        } else {
            const u32 ca = (d - 1u) / 2u, cb = d / 2u + 1u;
            c_lo = kind ? cb : 1u;
            c_hi = kind ? d - 1u : ca;
            if (c_lo > c_hi) continue;
            const u32 res = kind ? (t - 2u - phi) % sp : phi;
            const u32 n_lo = bmin / d, n_hi = B / d;
            n0 = n_lo + (res + sp - n_lo % sp) % sp;
            nb = n0 > n_hi ? 0u : (n_hi - n0) / sp + 1u;
            step = sp;
        }
        const u32 width = c_hi - c_lo + 1u;
        const u32 W = (width + 63u) >> 6;
        const u32 total = nb * W;  // < 2^32: at most B/64 + nb words
        // idx -> (band, word) with a float reciprocal and an exact fix-up
        const float rW = 1.0f / (float)W;
        for (u32 base_idx = 0; base_idx < total; base_idx += 64u) {
            const u32 idx = base_idx + lane;
            if (idx >= total) continue;
            u32 band = (u32)((float)idx * rW);
            while (band * W > idx) --band;
            while ((band + 1u) * W <= idx) ++band;
            const u32 w = idx - band * W;
            const u32 nd = (n0 + band * step) * d;  // <= B + d < 2^32
            const u32 b0 = nd + c_lo + 64u * w;
            const u32 bend = min(nd + c_hi, B);
            if (b0 > bend) continue;
            const u32 nbits = min(64u, bend - b0 + 1u);
            u64 valid = nbits == 64u ? ~0ull : ((1ull << nbits) - 1ull);
            if (b0 < bmin) {
                const u32 below = bmin - b0;
                if (below >= nbits) continue;
                valid &= ~0ull << below;
            }
            if ((d & 1u) == 0) valid &= (b0 & 1u) ? 0x5555555555555555ull : 0xAAAAAAAAAAAAAAAAull;
            if (d % 3u == 0) valid &= ~(kThirds << ((3u - b0 % 3u) % 3u));
            if (!valid) continue;
            const u32 ib = b0 + B;
            ++n_words;
            u64 base[NP];
            u64 surv = anchor_test<NP>(c.pl, wpp, base, ib, d, t, valid, n_windows);
            if (surv && c.q > t)
                surv &= screen(c.pl, wpp, NP, base, b0, B, d, t, c.q - t, n_windows);
            while (surv) {
                const u32 bit = (u32)__builtin_ctzll(surv);
                surv &= surv - 1ull;
                complete_anchor(*c.s, *c.f, c.pl, b0 + bit, d, c.dinv, t, c.tab_val, c.tab_lab,
                                c.best, c.failed);
                ++n_done;
            }
        }
    }
}

__global__ __launch_bounds__(256) void sp_filter(const SparsePrime *__restrict__ P,
                                                 const SparseProf *__restrict__ prof, u32 nprof,
                                                 u32 dmax, const u32 *__restrict__ planes,
                                                 const u32 *__restrict__ inv,
                                                 const u32 *__restrict__ tab_val,
                                                 const u16 *__restrict__ tab_lab,
                                                 unsigned long long *__restrict__ best,
                                                 u32 *__restrict__ failed,
                                                 unsigned long long *__restrict__ stats,
                                                 u32 kGroup) {
    // per-lane work counts: 64-bit window comparisons, candidate words,
    // completed anchors (summed per block, three atomics per block)
    __shared__ unsigned long long s_stats[3];
    if (threadIdx.x < 3) s_stats[threadIdx.x] = 0;
    __syncthreads();
    u64 n_windows = 0, n_words = 0, n_done = 0;
    // XCD-aware schedule: workgroups are dispatched round-robin over the 8
    // XCDs, so block b runs on XCD b % 8.  XCD x owns the profiles
    // k = x, x+8, x+16, ... and walks them in groups of kGroup profiles (a
    // tuning knob, 8 by default),
    // denominator-major inside a group: the waves of one XCD then touch the
    // planes of at most kGroup profiles at a time (they stay in its L2),
    // while each profile's denominators are still spread over the group.
    const u32 xcd = blockIdx.x % kXcds;
    const u32 wpb = blockDim.x >> 6;
    const u32 wl = (blockIdx.x / kXcds) * wpb + (threadIdx.x >> 6);
    const u32 waves_x = (gridDim.x / kXcds) * wpb;  // grid is a multiple of 8
    const u32 nprof_x = nprof > xcd ? (nprof - xcd + kXcds - 1) / kXcds : 0u;
    const u32 ngroups = (nprof_x + kGroup - 1) / kGroup;
    const u32 items = ngroups * kGroup * dmax;  // checked below 2^32 by the host
    for (u32 item = wl; item < items; item += waves_x) {
        const u32 g = item / (kGroup * dmax);
        const u32 rem = item - g * (kGroup * dmax);
        const u32 dd = __builtin_amdgcn_readfirstlane(1u + rem / kGroup);
        const u32 local = g * kGroup + rem % kGroup;
        if (local >= nprof_x) continue;
        const u32 kk = __builtin_amdgcn_readfirstlane(xcd + local * kXcds);
        const SparseProf f = prof[kk];
        const SparsePrime s = P[f.prime];
        if (dd > s.D) continue;
        const u64 key = __hip_atomic_load(&best[kk], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const u32 q = min((u32)(key >> 32), s.Q);
        if (q < 3u) continue;
        u32 t = max(3u, (q + 1u) / 2u);
        if (t > q) t = q;
        const u32 sp = q - t + 1u;  // schedule period
        const u32 phi = ((t - 1u) / 2u) % sp;
        FilterCtx c;
        c.s = &s;
        c.f = &f;
        c.pl = planes + f.plane_off;
        c.tab_val = tab_val;
        c.tab_lab = tab_lab;
        c.best = &best[kk];
        c.failed = failed;
        c.d = dd;
        c.dinv = inv[(u64)s.inv_off + dd];
        c.t = t;
        c.q = q;
        c.B = s.B;
        c.bmin = ((t - 1u) * dd + 1u) / 2u;
        switch (f.np) {
        case 1: filter_item<1>(c, sp, phi, n_windows, n_words, n_done); break;
        case 2: filter_item<2>(c, sp, phi, n_windows, n_words, n_done); break;
        case 3: filter_item<3>(c, sp, phi, n_windows, n_words, n_done); break;
        default: filter_item<4>(c, sp, phi, n_windows, n_words, n_done); break;
        }
    }
    atomicAdd(&s_stats[0], (unsigned long long)n_windows);
    atomicAdd(&s_stats[1], (unsigned long long)n_words);
    atomicAdd(&s_stats[2], (unsigned long long)n_done);
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd(&stats[threadIdx.x], s_stats[threadIdx.x]);
}

// One thread per prime: the output contract from the reduced keys.
__global__ void sp_finalize(const SparsePrime *__restrict__ P, u32 nprimes,
                            const SparseProf *__restrict__ prof,
                            const unsigned long long *__restrict__ best,
                            const u32 *__restrict__ a4, rbg_record *__restrict__ out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprimes) return;
    const SparsePrime s = P[i];
    rbg_record *rec = &out[s.index];
    rec->p = s.p;
    rec->mask = s.mask;
    for (int k = 0; k < 8; ++k) {
        rec->M[k] = 0;
        rec->L[k] = 0;
        rec->W[k] = 0;
        rec->K[k] = 0;
    }
    for (u32 k = s.prof_begin; k < s.prof_begin + s.nprof; ++k) {
        const SparseProf f = prof[k];
        const u64 key = best[k];
        const u32 M = (u32)(key >> 32);
        const u32 W = 0xFFFFFFFFu - (u32)key;
        const u32 L = f.prefix;
        const u64 bound = f.sigma0 ? 2ull * L + 2ull : (u64)L + 2ull;
        rec->M[f.r - 2] = M;
        rec->W[f.r - 2] = W;
        rec->L[f.r - 2] = L;
        rec->K[f.r - 2] = max((u64)M + 1ull, bound);
    }
    rec->A4 = a4[i];
    rec->status = 0;
}

}  // namespace rbg

using rbg::SparsePrime;
using rbg::SparseProf;

extern "C" hipError_t rbg_sparse_prepare(const SparsePrime *P, uint32_t nprimes,
                                         const SparseProf *prof, uint32_t nprof,
                                         uint32_t label_blocks, uint32_t inverse_blocks,
                                         uint32_t plane_blocks, const uint32_t *tab_val,
                                         const uint16_t *tab_lab, uint16_t *labels,
                                         uint32_t *inv, uint32_t *planes, uint32_t *err,
                                         hipStream_t stream) {
    hipLaunchKernelGGL(rbg::sp_labels, dim3(label_blocks), dim3(256), 0, stream, P, nprimes,
                       tab_val, tab_lab, labels, err);
    hipLaunchKernelGGL(rbg::sp_inverses, dim3(inverse_blocks), dim3(256), 0, stream, P, nprimes,
                       inv);
    hipLaunchKernelGGL(rbg::sp_planes, dim3(plane_blocks), dim3(256), 0, stream, P, nprimes, prof,
                       nprof, labels, planes);
    return hipGetLastError();
}

extern "C" uint32_t rbg_seed_segments(uint32_t bmax) {
    return (bmax + rbg::kSeedSeg - 1) / rbg::kSeedSeg;
}

extern "C" size_t rbg_seed_agg_bytes(void) { return sizeof(rbg::Agg); }

extern "C" hipError_t rbg_sparse_seed(const SparsePrime *P, SparseProf *prof, uint32_t nprof,
                                      const uint16_t *labels, void *segs, uint32_t segs_per_prof,
                                      unsigned long long *best, uint32_t *failed,
                                      hipStream_t stream) {
    const uint32_t waves = nprof * segs_per_prof;
    hipLaunchKernelGGL(rbg::sp_seed_scan, dim3((waves + 3) / 4), dim3(256), 0, stream, P, prof,
                       nprof, labels, (rbg::Agg *)segs, segs_per_prof);
    hipLaunchKernelGGL(rbg::sp_seed_merge, dim3((nprof + 63) / 64), dim3(64), 0, stream, P, prof,
                       nprof, (const rbg::Agg *)segs, segs_per_prof, best, failed);
    return hipGetLastError();
}

extern "C" hipError_t rbg_sparse_filter(const SparsePrime *P, const SparseProf *prof,
                                        uint32_t nprof, uint32_t dmax, const uint32_t *planes,
                                        const uint32_t *inv, const uint32_t *tab_val,
                                        const uint16_t *tab_lab, unsigned long long *best,
                                        uint32_t *failed, unsigned long long *stats,
                                        uint32_t grid, uint32_t group, hipStream_t stream) {
    hipLaunchKernelGGL(rbg::sp_filter, dim3(grid), dim3(256), 0, stream, P, prof, nprof, dmax,
                       planes, inv, tab_val, tab_lab, best, failed, stats, group);
    return hipGetLastError();
}

extern "C" hipError_t rbg_sparse_finalize(const SparsePrime *P, uint32_t nprimes,
                                          const SparseProf *prof, const unsigned long long *best,
                                          const uint32_t *a4, rbg_record *out,
                                          hipStream_t stream) {
    hipLaunchKernelGGL(rbg::sp_finalize, dim3((nprimes + 255) / 256), dim3(256), 0, stream, P,
                       nprimes, prof, best, a4, out);
    return hipGetLastError();
}
