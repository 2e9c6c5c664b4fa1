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
//     b >= (t-1)d/2) form spans of consecutive b; each lane tests 96
//     consecutive b of a span at once: for every plane, the 96-bit windows at
//     numerators b - j d (j = 1..t-1) are compared with the window at b, and a
//     lane takes a new word as soon as no candidate is left;
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

// Residue domain and constants of a sparse prime (the 33rd bits of the wide
// primes are kept in s.hi).
template <class D>
struct SpConsts {
    D dom;
    typename D::V one, r2;
};

__device__ __forceinline__ SpConsts<Dom32> sp_consts(const SparsePrime &s, Dom32) {
    return {{s.p, pos_inverse(s.ninv)}, s.one, s.r2};
}

__device__ __forceinline__ SpConsts<Dom33> sp_consts(const SparsePrime &s, Dom33) {
    const auto w = [&](u32 lo, int bit) { return (u64)lo | ((u64)((s.hi >> bit) & 1u) << 32); };
    return {{w(s.p, 0), s.ninv}, w(s.one, 1), w(s.r2, 2)};
}

// Labels are completely multiplicative: (a b)^e = a^e b^e, so
// j(a b) = j(a) + j(b) mod m.  The first pass powers only y = 2, 3 and the y
// coprime to 6 (a third of 1..B), listed densely so that every lane works:
// index i -> 2, 3, then 1, 5, 7, 11, 13, ... (3(i-2) + 1 + ((i-2) & 1)); the
// second pass (sp_labels_rest) fills y = 2^a 3^b x from j(2), j(3), j(x).
__device__ __forceinline__ u32 coprime6_at(u32 i) {
    return i < 2u ? i + 2u : 3u * (i - 2u) + 1u + ((i - 2u) & 1u);
}

template <class D>
__device__ __forceinline__ void labels_of(const SparsePrime &d, const typename D::V *tab_val,
                                          const u16 *tab_lab, u16 *labels, u32 *err,
                                          typename D::V *s_val, u16 *s_lab) {
    typedef typename D::V V;
    const u32 i0 = (blockIdx.x - d.lblk) * (256u * kLabelsPerThread);
    if (coprime6_at(i0) > d.B && i0 >= 2u) return;  // the block's indices lie past B
    const SpConsts<D> k = sp_consts(d, D{});
    const u32 voff = D::wide ? d.tab64 : d.tab_off;
    for (u32 j = threadIdx.x; j < d.m; j += blockDim.x) {
        s_val[j] = tab_val[voff + j];
        s_lab[j] = tab_lab[d.tab_off + j];
    }
    __syncthreads();
    if (blockIdx.x == d.lblk && threadIdx.x == 0) labels[d.lab_off] = 0;
    u32 bad = 0;
    for (u32 i = 0; i < kLabelsPerThread; ++i) {
        const u32 y = coprime6_at(i0 + i * 256u + threadIdx.x);
        if (y > d.B) break;
        const V v = k.dom.pow(k.dom.mul((V)y, k.r2), d.e, k.one);
        const u32 idx = table_index(s_val, d.m, v);
        bad |= s_val[idx] != v;
        labels[(u64)d.lab_off + y] = s_lab[idx];
    }
    if (bad) atomicOr(err, 1u);
}

// labels[lab_off + y] = j(y) for y in [1, B]; slot 0 holds 0.  tab_val holds
// the root tables of the primes below 2^32, tab_val64 those of the wider
// primes.
__global__ __launch_bounds__(256) void sp_labels(const SparsePrime *__restrict__ P, u32 nprimes,
                                                 const u32 *__restrict__ tab_val,
                                                 const u64 *__restrict__ tab_val64,
                                                 const u16 *__restrict__ tab_lab,
                                                 u16 *__restrict__ labels, u32 *__restrict__ err) {
    __shared__ u64 s_val[kMaxM];
    __shared__ u16 s_lab[kMaxM];
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 0));
    const SparsePrime d = P[pi];
    if (d.hi & 1u)
        labels_of<Dom33>(d, tab_val64, tab_lab, labels, err, s_val, s_lab);
    else
        labels_of<Dom32>(d, tab_val, tab_lab, labels, err, (u32 *)s_val, s_lab);
}

// x mod r for x < 2^20 and 2 <= r < 2^12 (labels, sums of a few labels,
// colours) by one high multiply: with g = floor((2^32-1)/r) + 1 the
// quotient floor(x g / 2^32) is exact because x g / 2^32 exceeds x / r by
// less than x / 2^32 < 1 / r.
struct SmallMod {
    u32 r, g;
    __device__ __forceinline__ explicit SmallMod(u32 r_) : r(r_), g(0xFFFFFFFFu / r_ + 1u) {}
    __device__ __forceinline__ u32 operator()(u32 x) const { return x - r * __umulhi(x, g); }
};

// Second pass: y = 2^a 3^b x with x coprime to 6 and y >= 4 not coprime to
// 6 gets j(y) = (a j(2) + b j(3) + j(x)) mod m (j(1) = 0).
__global__ __launch_bounds__(256) void sp_labels_rest(const SparsePrime *__restrict__ P,
                                                      u32 nprimes, u16 *__restrict__ labels) {
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 0));
    const SparsePrime d = P[pi];
    const u16 *lab = labels + d.lab_off;
    const u32 j2 = lab[2], j3 = lab[3];
    const SmallMod mod_m(d.m);
    const u32 y0 = 1u + (blockIdx.x - d.lblk) * (256u * kLabelsPerThread);
    for (u32 i = 0; i < kLabelsPerThread; ++i) {
        const u32 y = y0 + i * 256u + threadIdx.x;
        if (y > d.B) break;
        if (y < 4u || ((y & 1u) && y % 3u != 0u)) continue;
        const u32 a = (u32)__builtin_ctz(y);
        u32 x = y >> a, b = 0;
        while (x % 3u == 0u) {
            x /= 3u;
            ++b;
        }
        // a <= 20, b <= 12: the sum stays below 33 m < 2^17
        labels[(u64)d.lab_off + y] = (u16)mod_m(a * j2 + b * j3 + lab[x]);
    }
}

// inv[inv_off + d] = Montgomery form d^{-1} R mod p for d in [1, D] (d^(p-2));
// the wider primes write inv64.
__global__ __launch_bounds__(256) void sp_inverses(const SparsePrime *__restrict__ P, u32 nprimes,
                                                   u32 *__restrict__ inv, u64 *__restrict__ inv64) {
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 1));
    const SparsePrime s = P[pi];
    const u32 d = (blockIdx.x - s.iblk) * 256u + threadIdx.x;
    if (d == 0 || d > s.D) return;
    if (s.hi & 1u) {
        const SpConsts<Dom33> k = sp_consts(s, Dom33{});
        inv64[(u64)s.inv_off + d] = k.dom.pow(k.dom.mul(d, k.r2), k.dom.p - 2u, k.one);
    } else {
        const u32 pinv = pos_inverse(s.ninv);
        const u32 dm = mont_mul(d, s.r2, s.p, pinv);
        inv[(u64)s.inv_off + d] = mont_pow(dm, s.p - 2, s.p, pinv, s.one);
    }
}

constexpr u32 kPlanePairsPerWave = kPlanePairsPerBlock / 4;  // 64: one per lane

// Full colour row (row 0) of one order: bit k of plane b is bit b of the
// colour of the numerator y = k - B, cl(-y) = cl(y) + cl(-1), and 0 for
// numerators outside [-B, B] and for y = 0.  One wave builds 64 bits (two
// words) of each of the NP planes per step; lane q keeps the words of step
// q, and after 64 steps the lanes store them together (one 512-byte store
// per plane word).  |y| <= B + 64 < 2^21, so positions are 32-bit.
template <int NP>
__device__ __forceinline__ void row0_planes(const SparsePrime &s, const SparseProf &f, u32 first,
                                            u32 nq, u32 pairs, u32 wpr, const u16 *lab,
                                            u32 *base) {
    const u32 lane = threadIdx.x & 63u;
    u64 mine[NP];  // this lane's step: words of planes 0..NP-1
#pragma unroll
    for (int b = 0; b < NP; ++b) mine[b] = 0;
    const SmallMod mod_r(f.r), mod_m(s.m);
    const int B = (int)s.B;
    // eight steps at a time: the eight label loads first (every lane loads an
    // in-range label, position 1 for the positions outside the table), then
    // the colours and ballots of the steps that exist
    for (u32 q0 = 0; q0 < nq; q0 += 8u) {
        u32 jv[8];
        int yv[8];
#pragma unroll
        for (u32 u = 0; u < 8u; ++u) {
            yv[u] = (int)((first + q0 + u) * 64u + lane) - B;
            const int a = yv[u] < 0 ? -yv[u] : yv[u];
            jv[u] = lab[(yv[u] != 0 && a <= B) ? a : 1];
        }
#pragma unroll
        for (u32 u = 0; u < 8u; ++u) {
            if (q0 + u >= nq) break;
            // the step's 64 numerators: all positive and in the table, all
            // negative and in it (wave-uniform tests), or neither
            const int ylo = (int)((first + q0 + u) * 64u) - B;
            u32 c;
            if (ylo > 0 && ylo + 63 <= B) {
                c = mod_r(jv[u]);
            } else if (ylo + 63 < 0 && ylo >= -B) {
                c = mod_r(mod_m(jv[u] + s.jneg));
            } else {
                const int y = yv[u];
                const int a = y < 0 ? -y : y;
                const bool in = y != 0 && a <= B;
                c = !in ? 0u : y > 0 ? mod_r(jv[u]) : mod_r(mod_m(jv[u] + s.jneg));
            }
#pragma unroll
            for (int b = 0; b < NP; ++b) {
                const u64 bits = ballot64((c >> b) & 1u);
                if (lane == q0 + u) mine[b] = bits;
            }
        }
    }
    const u32 pair = first + lane;
    if (pair < pairs) {
#pragma unroll
        for (int b = 0; b < NP; ++b) {
            u32 *pl = base + (u64)b * wpr;
            pl[2 * pair] = (u32)mine[b];
            pl[2 * pair + 1] = (u32)(mine[b] >> 32);
        }
    }
}

// Row 0 of every profile.  The kernel's grid covers every row's blocks (the
// plane-block map of the profile records), but the blocks of the decimated
// rows return at once: sp_rows copies their bits out of row 0.
__global__ __launch_bounds__(256) void sp_planes(const SparsePrime *__restrict__ P, u32 nprimes,
                                                 const SparseProf *__restrict__ prof, u32 nprof,
                                                 const u16 *__restrict__ labels,
                                                 u32 *__restrict__ planes) {
    // block -> profile through the per-profile block offsets kept in pblk of
    // the profile records (searched linearly from the prime's first
    // profile), then -> row by the rows' block counts
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 2));
    const SparsePrime s = P[pi];
    u32 k = s.prof_begin;
    while (k + 1 < s.prof_begin + s.nprof && prof[k + 1].pblk <= blockIdx.x) ++k;
    const SparseProf f = prof[k];
    const u32 local = blockIdx.x - f.pblk;
    if (local >= row_blocks(0, s.B)) return;  // a block of a decimated row
    const u32 wpr = row_words(1, s.B);
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 pairs = wpr / 2;
    const u32 first = (local * 4u + wave) * kPlanePairsPerWave;
    const u32 nq = first < pairs ? min(kPlanePairsPerWave, pairs - first) : 0u;
    const u16 *lab = labels + s.lab_off;
    u32 *base = planes + (u64)f.plane_off * 16u;
    switch (f.np) {
    case 1: row0_planes<1>(s, f, first, nq, pairs, wpr, lab, base); break;
    case 2: row0_planes<2>(s, f, first, nq, pairs, wpr, lab, base); break;
    case 3: row0_planes<3>(s, f, first, nq, pairs, wpr, lab, base); break;
    default: row0_planes<4>(s, f, first, nq, pairs, wpr, lab, base); break;
    }
}

// 32 bits of a decimated row from the full row: bit b of the result is bit
// g b of the stream s[0], s[1], ... (the row-0 bits from rho on, already
// aligned by alignbit).
template <u32 G>
__device__ __forceinline__ u32 every_gth_bit(const u32 *s) {
    u32 out = 0;
#pragma unroll
    for (u32 b = 0; b < 32; ++b) {
        constexpr u32 kStride = G;
        const u32 j = kStride * b;
        out |= ((s[j >> 5] >> (j & 31u)) & 1u) << b;
    }
    return out;
}

template <u32 G>
__device__ __forceinline__ u32 decimated_word(const u32 *full, u32 wfull, u32 rho, u32 w) {
    // row bit 32 w + b is row-0 bit rho + G (32 w + b) = 32 G w + rho + G b
    u32 in[G + 1];
#pragma unroll
    for (u32 k = 0; k <= G; ++k) in[k] = G * w + k < wfull ? full[G * w + k] : 0u;
    u32 s[G];
#pragma unroll
    for (u32 k = 0; k < G; ++k) s[k] = __builtin_amdgcn_alignbit(in[k + 1], in[k], rho);
    return every_gth_bit<G>(s);
}

// The five decimated rows of every profile (rbg_sparse.hpp) from its full
// row: row (g, c) bit k is row-0 bit rho + g k.  Same grid and block map as
// sp_planes; blocks of row 0 return at once, the others write 512 words of
// each plane (two per thread).
__global__ __launch_bounds__(256) void sp_rows(const SparsePrime *__restrict__ P, u32 nprimes,
                                               const SparseProf *__restrict__ prof,
                                               u32 *__restrict__ planes) {
    const u32 pi = __builtin_amdgcn_readfirstlane(find_prime(P, nprimes, blockIdx.x, 2));
    const SparsePrime s = P[pi];
    u32 k = s.prof_begin;
    while (k + 1 < s.prof_begin + s.nprof && prof[k + 1].pblk <= blockIdx.x) ++k;
    const SparseProf f = prof[k];
    u32 local = blockIdx.x - f.pblk, row = 0;
    while (row + 1 < kRows && local >= row_blocks(row, s.B)) local -= row_blocks(row++, s.B);
    row = __builtin_amdgcn_readfirstlane(row);
    if (row == 0) return;
    const u32 g = row_g(row), rho = row_rho(row, s.B);
    const u32 wpr = row_words(g, s.B), wfull = row_words(1, s.B);
    u32 *prof_rows = planes + (u64)f.plane_off * 16u;
    for (u32 half = 0; half < 2; ++half) {
        const u32 w = local * 2u * kPlanePairsPerBlock + half * 256u + threadIdx.x;
        if (w >= wpr) break;
        for (u32 b = 0; b < f.np; ++b) {
            const u32 *full = prof_rows + (u64)b * wfull;
            u32 v;
            if (g == 2) v = decimated_word<2>(full, wfull, rho, w);
            else if (g == 3) v = decimated_word<3>(full, wfull, rho, w);
            else v = decimated_word<6>(full, wfull, rho, w);
            prof_rows[row_offset(row, s.B, f.np) + (u64)b * wpr + w] = v;
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
    const SmallMod mod_r(f.r);
    const u32 lane = threadIdx.x & 63u;
    const u32 first = 1u + sg * kSeedSeg;
    const u32 len = first > s.B ? 0u : min(kSeedSeg, s.B - first + 1u);
    OrderState st;
    order_init(st);
    u32 carry = 0, fcol = 0, lcol = 0;
    const u32 nchunks = (len + 63u) >> 6;
    const u16 *lab = labels + s.lab_off;
    for (u32 c0 = 0; c0 < nchunks; c0 += 4u) {
        // the labels of four chunks in flight before the first is used
        u32 jv[4];
#pragma unroll
        for (u32 i = 0; i < 4u; ++i) {
            const u32 y = first + ((c0 + i) << 6) + lane;
            jv[i] = lab[y <= s.B ? y : 0u];
        }
#pragma unroll
        for (u32 i = 0; i < 4u; ++i) {
            const u32 c = c0 + i;
            if (c >= nchunks) break;
            const u32 base = c << 6;
            const u32 valid = min(64u, len - base);
            const u64 vmask = valid == 64u ? ~0ull : ((1ull << valid) - 1ull);
            const u32 y = first + base + lane;
            const u32 col = y <= s.B ? mod_r(jv[i]) : 0u;
            const u32 left = shift_up_one(col, carry);
            carry = __builtin_amdgcn_readlane(col, 63);
            if (c == 0) fcol = __builtin_amdgcn_readfirstlane(col);
            lcol = __builtin_amdgcn_readlane(col, valid - 1);
            const bool chg = col != left;
            u64 C = ballot64(chg);
            if (c == 0) C |= 1ull;
            C &= vmask;
            // the interior-run search's rejection test on the vector unit
            // (the scalar unit carries the rest of the run bookkeeping)
            bool cand = true;
            if (st.best != 0u && st.best <= 62u)
                cand = has_ones_run(chg || lane >= valid || (c == 0 && lane == 0) ? lane + 1u : 0u,
                                    st.best);
            order_update(st, C, ~C & vmask, base, cand);
        }
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

// One wave per profile: ordered merge of its segment summaries (each lane
// merges a contiguous range of segments, then the lanes' results are merged
// pairwise in lane order, six rounds); the key of the longest seed run
// (partial end run included) and the prefix L.
__device__ __forceinline__ Agg shfl_down_agg(const Agg &a, u32 off) {
    Agg o;
    o.first_x = (u32)__shfl_down((int)a.first_x, off);
    o.len = (u32)__shfl_down((int)a.len, off);
    o.fcode = (u32)__shfl_down((int)a.fcode, off);
    o.lcode = (u32)__shfl_down((int)a.lcode, off);
    o.prefix = (u32)__shfl_down((int)a.prefix, off);
    o.suffix = (u32)__shfl_down((int)a.suffix, off);
    o.best = (u32)__shfl_down((int)a.best, off);
    o.bstart = (u32)__shfl_down((int)a.bstart, off);
    o.a4 = (u32)__shfl_down((int)a.a4, off);
    return o;
}

__global__ __launch_bounds__(256) void sp_seed_merge(const SparsePrime *__restrict__ P,
                                                     SparseProf *__restrict__ prof, u32 nprof,
                                                     const Agg *__restrict__ segs,
                                                     u32 segs_per_prof,
                                                     unsigned long long *__restrict__ best,
                                                     u32 *__restrict__ failed) {
    const u32 k = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6));
    if (k >= nprof) return;
    const u32 lane = threadIdx.x & 63u;
    const u32 per = (segs_per_prof + 63u) / 64u;
    const u32 i0 = min(lane * per, segs_per_prof), i1 = min(i0 + per, segs_per_prof);
    Agg acc;
    acc.len = 0;
    for (u32 i = i0; i < i1; ++i) acc = merge(acc, segs[(u64)k * segs_per_prof + i], ~0u);
#pragma unroll
    for (u32 off = 1; off < 64u; off <<= 1) {
        const Agg right = shfl_down_agg(acc, off);
        if ((lane & (2u * off - 1u)) == 0) acc = merge(acc, right, ~0u);
    }
    if (lane != 0) return;
    const SparseProf f = prof[k];
    const SparsePrime s = P[f.prime];
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
template <class D>
__device__ u32 colour_direct(const SparsePrime &s, const SpConsts<D> &k, u32 r,
                             const typename D::V *tab_val, const u16 *tab_lab, typename D::V z) {
    const typename D::V v = k.dom.pow(k.dom.mul(z, k.r2), s.e, k.one);
    const u32 lo = table_index(tab_val + (D::wide ? s.tab64 : s.tab_off), s.m, v);
    return tab_lab[s.tab_off + lo] % r;
}

// Extend the monochromatic window [x, x+t-1] (numerators -b + i d) to its run.
// Rare (a few per task), so kept out of line: the filter loop then holds only
// the two record pointers, not the records.  A run that needs more than
// `limit` extension steps refuses the prime (failed bit 2): the host then
// recomputes it with the dense scan.
template <class D>
__device__ void complete_anchor(
    const SparsePrime *sg, const SparseProf *fg, const u32 *pl, u32 b, u32 d, typename D::V dinv,
    u32 t, const typename D::V *tab_val, const u16 *tab_lab, unsigned long long *best_slot,
    u32 *failed, u32 limit) {
    typedef typename D::V V;
    const SparsePrime s = *sg;
    const SparseProf f = *fg;
    const SpConsts<D> k = sp_consts(s, D{});
    const V p = k.dom.p;
    // dinv is the Montgomery form d^{-1} R, so REDC(b * dinv) = b / d mod p
    const V bd = k.dom.mul((V)b, dinv);
    const V x = p - bd;  // bd != 0 because 0 < b <= B < p
    const int64_t B = s.B;
    const u32 c0 = colour_at(pl, f.wpp, f.np, -(int64_t)b, s.B);
    u32 abs0 = 0xFFFFFFFFu;
    u32 steps = 0;
    // left
    V lo = x;
    int64_t a = -(int64_t)b - d;
    bool table = true;
    while (lo > 1) {
        bool same;
        if (table && a >= -B) {
            same = colour_at(pl, f.wpp, f.np, a, s.B) == c0;
        } else {
            table = false;
            if (abs0 == 0xFFFFFFFFu) abs0 = colour_direct<D>(s, k, f.r, tab_val, tab_lab, x);
            same = colour_direct<D>(s, k, f.r, tab_val, tab_lab, lo - 1u) == abs0;
        }
        if (!same || ++steps > limit) break;
        --lo;
        a -= d;
    }
    // right
    V hi = x + t - 1u;
    a = -(int64_t)b + (int64_t)t * d;
    table = true;
    while (hi < p - 1u) {
        bool same;
        if (table && a <= B) {
            same = colour_at(pl, f.wpp, f.np, a, s.B) == c0;
        } else {
            table = false;
            if (abs0 == 0xFFFFFFFFu) abs0 = colour_direct<D>(s, k, f.r, tab_val, tab_lab, x);
            same = colour_direct<D>(s, k, f.r, tab_val, tab_lab, hi + 1u) == abs0;
        }
        if (!same || ++steps > limit) break;
        ++hi;
        a += d;
    }
    if (steps > limit) {
        atomicOr(&failed[f.prime], 2u);
        return;
    }
    // the run lies in [1, p-1] and its reflected start is at most (p-1)/2
    const u32 len = (u32)(hi - lo + 1u);
    const u32 w = (u32)(lo < p - hi ? lo : p - hi);
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

constexpr u32 kXcds = 8;  // XCDs of an MI355X

// A task's rows (wave-uniform).  Denominator d tests only anchors b coprime
// to g = gcd(d, 6) (the others share a factor with d and are tested with a
// smaller denominator), so its anchors fall into the classes c mod g with
// gcd(c, g) = 1: one class for g = 1 (every b, the full row) and g = 2
// (odd b), two for g = 3 (b = 1, 2 mod 3) and g = 6 (b = 1, 5 mod 6).  An
// anchor b of class ci is row bit k = (b + B - rho_ci) / g, and b - j d is
// row bit k - j d / g.  Plane pb of the first class's row starts at
// plb[pb]; the second class's row lies `cstride` bits further in every
// plane, so row bit k of class ci is bit k + ci cstride from plb[pb] and
// every window of the task is read through the same plane pointers.
struct TaskRows {
    u32 g, dg, nc;      // decimation, d / g, classes
    u32 cstride;        // bits from the first class's row to the second's
    u32 rho[2];         // (c + B) mod g of each class
    u32 nbits[2];       // row bits that carry colours: numerators in [-B, B]
};

// Bilateral screen (every placement of a q-window containing the anchor;
// reference lemma GPU_METHOD_EN.tex:457-480, vdw_cut::screen): with
// h = q - t extra positions needed, a candidate anchor is kept iff it has
// L matching positions to its left and R to its right with L + R >= h.  In
// numerator space (negated, so equality is unchanged) position x-k reads
// b + k d and position x+t-1+k reads b - (t-1+k) d; in row space these are
// row bits kk + k dg and kk - (t-1+k) dg.  Bits outside the class's
// colours [start, lim) (numerators outside [-B, B]) are unknown and never
// reject.  R is kept as a bit-sliced count (5 planes, h <= 31; the cover
// cap Q <= 63 and t = ceil(q/2) keep h there) so that every candidate of
// the word is screened at once.
template <int NP>
__device__ __forceinline__ u64 match_left(const u32 *const *plb, const u64 *base, u32 kk, u32 lim,
                                          u32 dg, u32 k, u32 &n_windows) {
    const u32 lo = kk + k * dg;
    if (lo >= lim) return ~0ull;
    u64 mm = 0;
#pragma unroll
    for (int pb = 0; pb < NP; ++pb) mm |= base[pb] ^ window64(plb[pb], lo);
    n_windows += 2u * NP;  // counted in 32-anchor plane dwords
    const u32 known = lim - lo;  // bits 0..known-1 carry colours
    if (known < 64u) mm &= (1ull << known) - 1ull;
    return ~mm;
}

template <int NP>
__device__ __forceinline__ u64 match_right(const u32 *const *plb, const u64 *base, u32 kk,
                                           u32 start, u32 dg, u32 t, u32 k, u32 &n_windows) {
    const int64_t lo = (int64_t)kk - (int64_t)(t - 1 + k) * dg;
    if (lo + 63 < (int64_t)start) return ~0ull;
    u64 mm = 0;
    if (lo >= (int64_t)start) {
#pragma unroll
        for (int pb = 0; pb < NP; ++pb) mm |= base[pb] ^ window64(plb[pb], (u32)lo);
    } else {
        const u32 shift = (u32)((int64_t)start - lo);  // 1..63 low bits lie before the row
#pragma unroll
        for (int pb = 0; pb < NP; ++pb) mm |= base[pb] ^ (window64(plb[pb], start) << shift);
        mm &= ~((1ull << shift) - 1ull);
    }
    n_windows += 2u * NP;
    return ~mm;
}

// Screen of the candidates `cand` of the word at row bit kk (the anchors
// that passed every stage); candidates outside `cand` are never reported,
// so the counts follow `cand` only and the loops end as soon as those runs
// end.
template <int NP>
__device__ __forceinline__ u64 screen(const u32 *const *plb, const u64 *base, u64 cand, u32 kk,
                                      u32 start, u32 lim, u32 dg, u32 t, u32 h, u32 &n_windows) {
    // R: consecutive right matches per candidate, bit-sliced (r0 = LSB)
    u64 r0 = 0, r1 = 0, r2 = 0, r3 = 0, r4 = 0, cr = cand;
    for (u32 k = 1; k <= h; ++k) {
        cr &= match_right<NP>(plb, base, kk, start, dg, t, k, n_windows);
        if (!cr) break;
        u64 c = cr, x;
        x = r0 & c; r0 ^= c; c = x;
        x = r1 & c; r1 ^= c; c = x;
        x = r2 & c; r2 ^= c; c = x;
        x = r3 & c; r3 ^= c; c = x;
        r4 ^= c;
    }
    u64 pass = 0, cl = cand;
    for (u32 k = 0; k <= h; ++k) {
        if (k) {
            cl &= match_left<NP>(plb, base, kk, lim, dg, k, n_windows);
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

__device__ __forceinline__ u32 lanes_below(u64 m) {
    return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// Per-task context (wave-uniform).
template <class D>
struct FilterCtx {
    const SparsePrime *s;
    const SparseProf *f;
    const u32 *pl;    // the full row, plane 0 (completion reads colours there)
    const u32 *rows;  // plane 0 of the task's first class row
    const typename D::V *tab_val;
    const u16 *tab_lab;
    unsigned long long *best;
    u32 *failed;
    typename D::V dinv;
    u32 d, t, q, B, bmin, wpr, ext_limit;
    TaskRows R;
};

// Scheduled anchors of one task as spans.  The schedule keeps the anchors
// b = n d + c (b in [bmin, B]) with c in [1, (d-1)/2] and n = phi mod s, or
// c in [d/2 + 1, d-1] and n = t-2-phi mod s (reference schedule
// GPU_METHOD_EN.tex:380-455).  Since t-2-phi = phi (t even) or phi - 1 (t
// odd) mod s, the two sides pair up into contiguous spans
// S(m) = [m d + s_lo, m d + s_lo + width - 1] for m = rm mod s:
//   t even: S(m) = [m d + 1, m d + d - 1]; c = d/2 (d even) is not scheduled;
//   t odd:  S(m) = [m d + d/2 + 1, (m+1) d + (d-1)/2]; (m+1) d is not.
// At stage j every span reads the span j d below it, so consecutive words of
// a span are consecutive in the colour rows at every stage.  For d <= 2
// one span [max(bmin, t or t-1), B] holds every anchor.
struct Spans {
    u32 s_lo, width, excl;  // span offset and width; excluded offset (~0u: none)
    u32 m_first, nsb;       // first span index, spans (step s)
};

__device__ __forceinline__ Spans spans_of(u32 d, u32 t, u32 sp, u32 phi, u32 bmin, u32 B) {
    Spans S;
    S.excl = ~0u;
    S.m_first = 0;
    S.nsb = 0;
    if (d <= 2) {
        const u32 lo_b = max(bmin, d == 1 ? t : t - 1u);
        S.s_lo = lo_b;
        S.width = lo_b <= B ? B - lo_b + 1u : 0u;
        S.nsb = lo_b <= B ? 1u : 0u;
    } else {
        const bool odd_t = (t & 1u) != 0;
        const u32 rm = odd_t ? (phi + sp - 1u) % sp : phi;
        S.s_lo = odd_t ? d / 2u + 1u : 1u;
        S.width = odd_t ? d + (d - 1u) / 2u - d / 2u : d - 1u;
        if (odd_t) S.excl = d - S.s_lo;
        else if ((d & 1u) == 0) S.excl = d / 2u - 1u;
        // spans meeting [bmin, B]: m d + s_lo + width - 1 >= bmin and
        // m d + s_lo <= B
        const int x = (int)bmin + 1 - (int)S.s_lo - (int)S.width;
        const u32 m_lo = x <= 0 ? 0u : ((u32)x + d - 1u) / d;
        S.m_first = m_lo + (rm + sp - m_lo % sp) % sp;
        const u32 m_hi = (B - S.s_lo) / d;
        S.nsb = S.m_first > m_hi ? 0u : (m_hi - S.m_first) / sp + 1u;
    }
    return S;
}

// Words are 96 consecutive bits of one class row (96 anchors b, b + g, ...,
// b + 95 g of one span): a stage reads one 96-bit window per plane, at any
// bit offset, with one four-dword load.
constexpr u32 kWordBits = 96;
constexpr u32 kQueue = 128;  // entries of a wave's queue of passing words (two per word at most)

// Screen and complete the queued words (one lane per entry: a row bit kk and
// the 64 anchors from it that passed every stage).
template <class D, int NP>
__device__ __forceinline__ void drain_queue(const FilterCtx<D> &c, const u32 *const *plb, u32 qn,
                                            const u32 *qb, const u64 *qs, u32 lane,
                                            u32 &n_windows, u32 &n_done) {
    const TaskRows &R = c.R;
    for (u32 e = lane; e < qn; e += 64u) {
        const u32 kk = qb[e];
        u64 surv = qs[e];
        if (!surv) continue;
        const u32 ci = kk >= R.cstride && R.nc > 1 ? 1u : 0u;
        const u32 start = ci ? R.cstride : 0u;
        if (c.q > c.t) {
            u64 base[NP];
#pragma unroll
            for (int pb = 0; pb < NP; ++pb) base[pb] = window64(plb[pb], kk);
            surv &= screen<NP>(plb, base, surv, kk, start, start + R.nbits[ci], R.dg, c.t,
                               c.q - c.t, n_windows);
        }
        // anchor of bit 0: b0 = g (kk - start) + rho - B
        const u32 b0 = R.g * (kk - start) + R.rho[ci] - c.B;
        while (surv) {
            const u32 bit = (u32)__builtin_ctzll(surv);
            surv &= surv - 1ull;
            complete_anchor<D>(c.s, c.f, c.pl, b0 + R.g * bit, c.d, c.dinv, c.t, c.tab_val,
                               c.tab_lab, c.best, c.failed, c.ext_limit);
            ++n_done;
        }
    }
}
// r & ~(a ^ b) in one V_BITOP3_B32 (truth table 0x90 over the inputs r, a, b,
// r the most significant index bit).
__device__ __forceinline__ u32 and_xnor(u32 r, u32 a, u32 b) {
    return __builtin_amdgcn_bitop3_b32(r, a, b, 0x90);
}


// Per class of a task: where its words lie and which of their bits are
// anchors (wave-uniform).  Every span holds the same anchors of the class
// relative to its start (g | d), so a word is named by (span, w) and
//   k0 = klo + span * spk + 96 w     (row bit of its first anchor),
// with `n` anchors of the class per span in W words; bits at row bits
// below kmin (anchors below bmin) or above kmax (above B) are not anchors,
// nor is bit exb of word exw (the span's unscheduled point).
struct ClassWords {
    u32 n, W, klo, kmin, kmax, exw, exb, start;
};

__device__ __forceinline__ ClassWords class_words(const Spans &S, const TaskRows &R, u32 ci,
                                                  u32 d, u32 B, u32 bmin) {
    const u32 g = R.g;
    const u32 cls = g == 1 ? 0u : (g == 2 ? 1u : (ci ? g - 1u : 1u));  // 1, or 1 and g-1
    const u32 delta = (cls + g - S.s_lo % g) % g;  // first anchor of the class in a span
    ClassWords w;
    w.n = S.width > delta ? (S.width - 1u - delta) / g + 1u : 0u;
    w.W = (w.n + kWordBits - 1u) / kWordBits;
    w.start = ci * R.cstride;
    const u32 rho = R.rho[ci];
    w.klo = w.start + (S.m_first * d + S.s_lo + delta + B - rho) / g;
    w.kmin = w.start + (bmin + B - rho + g - 1u) / g;
    w.kmax = w.start + (2u * B - rho) / g;
    w.exw = ~0u;
    w.exb = 0;
    if (S.excl != ~0u && S.excl >= delta && (S.excl - delta) % g == 0) {
        const u32 i = (S.excl - delta) / g;
        w.exw = i / kWordBits;
        w.exb = i % kWordBits;
    }
    return w;
}

// Every scheduled word of one (profile, denominator) task for NP planes.
//
// Each lane holds one word (96 anchors of one class row) and tests it G
// stages per step: for stages j the 96-bit windows at row bits k - j dg must
// equal the base window at k.  The base words of a new word are loaded when
// the lane takes it and aligned only after the step's stage loads are
// issued, so taking a word costs no extra memory round trip.  A lane whose
// word dies or passes takes the next word at once (words are handed out in
// order with a ballot), so the lanes stay busy however differently the
// words end.  Passing words (anchors that survived every stage) go to a
// per-wave LDS queue, one entry per half (anchors 0..63, 64..95) that has
// survivors; the queue is screened and completed with all lanes at once
// outside the word loop, at the end of the task or when it is full (the
// lanes whose words did not fit keep them and the word loop stops handing
// out words until the queue has been drained).  A step tests min(G, t-1)
// distinct stages from min(j, t - min(G, t-1)), so a word's last step
// re-tests stages it has already passed rather than read past stage t-1: every
// address stays inside the row and the stage offsets of a step are
// wave-uniform.  The work count n_windows takes, per word, the stages through
// the one that killed its last anchor (t - 1 for a passing word), not the
// stages its last step loaded after that.
constexpr u32 kRefill = 48;  // smallest cohort of idle lanes that takes new words

template <class D, int NP>
__device__ void filter_task(const FilterCtx<D> &c, u32 sp, u32 phi, u32 *qb, u64 *qs, u32 *cw,
                            u32 &n_windows, u32 &n_words, u32 &n_done, u32 &n_loads) {
    constexpr u32 WB = kWordBits;
    constexpr u32 NW = 4;  // dwords per window load
    constexpr int G = NP == 1 ? 6 : NP == 2 ? 3 : NP == 3 ? 2 : 1;  // stages per step (register budget)
    const u32 lane = threadIdx.x & 63u;
    const TaskRows &R = c.R;
    const u32 t = c.t, dg = R.dg;
    const Spans S = spans_of(c.d, t, sp, phi, c.bmin, c.B);
    // both classes' word parameters go to the wave's LDS table (a lane reads
    // its class's eight words when it takes a word), only W0 stays in a
    // register
    u32 W0, Wsum;
    {
        const ClassWords C0 = class_words(S, R, 0, c.d, c.B, c.bmin);
        ClassWords C1 = C0;
        C1.W = 0;
        if (R.nc > 1) C1 = class_words(S, R, 1, c.d, c.B, c.bmin);
        W0 = C0.W;
        Wsum = C0.W + C1.W;  // words per span
        if (lane < 2u) {
            const ClassWords &C = lane ? C1 : C0;
            u32 *row = cw + 8u * lane;
            row[0] = C.n;
            row[1] = C.klo;
            row[2] = C.kmin;
            row[3] = C.kmax;
            row[4] = C.exw;
            row[5] = C.exb;
            row[6] = C.start;
            row[7] = 0;
        }
        wave_lds_fence();
    }
    const u32 total = S.nsb * Wsum;  // < 2^23: at most 2B/96 + 2 nsb words
    const float rW = 1.0f / (float)max(Wsum, 1u);
    const u32 spk = sp * dg;  // row bits from one span to the next
    const u32 *plb[NP];
#pragma unroll
    for (int pb = 0; pb < NP; ++pb) plb[pb] = c.rows + (u32)pb * c.wpr;
    // stage offsets of a step: Gt = min(G, t-1) distinct stages, the rest
    // repeat the last one (wave-uniform)
    const u32 Gt = min((u32)G, t - 1u);
    u32 ug[G];
#pragma unroll
    for (int g = 0; g < G; ++g) ug[g] = min((u32)g, Gt - 1u) * dg;
    u32 next = 0, qn = 0;
    u32 used = 0;  // stages the lane's last word needed (counted when it takes the next)
    bool pend = false;  // a passing word that did not fit the queue
    u32 k0 = 0;
    u64 live = 0;    // surviving anchors 0 .. 63 of the word
    u32 live_hi = 0; // surviving anchors 64 .. 95
    for (;;) {
        bool stall = false;
        bool have = false;
        u32 j = 1, sh = 0;
        // row bit of stage j, and of stage t - Gt: a step tests Gt stages
        // from min(j, t - Gt), so a word's last step re-tests a few stages
        // instead of reading past stage t-1
        int io = 0, io_min = 0;
        u32 braw[NP][NW];
#pragma unroll
        for (int pb = 0; pb < NP; ++pb)
#pragma unroll
            for (u32 k = 0; k < NW; ++k) braw[pb][k] = 0;
        for (;;) {
            // idle lanes take words in cohorts of at least kRefill (or all
            // 64 at once): a cohort starts together on consecutive words, so
            // its loads stay on few cache lines
            const u64 idle = ballot64(!have && !pend);
            const u32 nidle = (u32)__builtin_popcountll(idle);
            if (!stall && idle && next < total && (nidle >= kRefill || nidle == 64u)) {
                bool took = false;
                if (!have && !pend) {
                    // the stages this lane's last word took (t - 1 when it
                    // passed) are counted now, once per word
                    n_windows += (u32)NP * (WB / 32u) * used;
                    used = 0;
                    j = 1;
                    const u32 my = next + lanes_below(idle);
                    if (my < total) {
                        // exact quotient: the float estimate is at most one
                        // low (|error| < 1/Wsum because total < 2^23)
                        u32 span = (u32)((float)my * rW);
                        if ((span + 1u) * Wsum <= my) ++span;
                        const u32 rem = my - span * Wsum;
                        const bool second = rem >= W0;
                        const uint4 *cp = (const uint4 *)(cw + (second ? 8u : 0u));
                        const uint4 ca = cp[0], cb = cp[1];
                        ClassWords C;
                        C.n = ca.x;
                        C.klo = ca.y;
                        C.kmin = ca.z;
                        C.kmax = ca.w;
                        C.exw = cb.x;
                        C.exb = cb.y;
                        C.start = cb.z;
                        const u32 w = second ? rem - W0 : rem;
                        k0 = C.klo + span * spk + WB * w;
                        // anchors of the word: row bits [lo_k, hi_k], the
                        // class's anchors of the span clipped to [kmin, kmax]
                        // (compared as row bits first, so the offsets below
                        // are only formed for a word that holds anchors)
                        const u32 nvalid = min(WB, C.n - WB * w);
                        const u32 lo_k = max(k0, C.kmin);
                        const u32 hi_k = min(k0 + nvalid - 1u, C.kmax);
                        u64 valid = 0;
                        u32 valid_hi = 0;
                        if (lo_k <= hi_k) {
                            const u32 first = lo_k - k0, last = hi_k - k0;
                            if (first < 64u)
                                valid = (~0ull << first) & (~0ull >> (63u - min(last, 63u)));
                            if (last >= 64u)
                                valid_hi = (~0u << (max(first, 64u) - 64u)) & (~0u >> (95u - last));
                            if (w == C.exw) {
                                if (C.exb < 64u) valid &= ~(1ull << C.exb);
                                else valid_hi &= ~(1u << (C.exb - 64u));
                            }
                        }
                        if (valid | valid_hi) {
                            have = true;
                            live = valid;
                            live_hi = valid_hi;
                            j = 1;
                            io = (int)k0 - (int)dg;
                            // anchors are >= bmin, so stage t-1 stays in the
                            // row; bits below bmin only need an address in it
                            io_min = max((int)k0 - (int)((t - Gt) * dg),
                                         (int)(C.start + (Gt - 1u) * dg));
                            sh = k0 & 31u;
                            const u32 wo = k0 >> 5;
#pragma unroll
                            for (int pb = 0; pb < NP; ++pb)
#pragma unroll
                                for (u32 k = 0; k < NW; ++k) braw[pb][k] = plb[pb][wo + k];
                            took = true;
                        }
                    }
                }
                next += (u32)__builtin_popcountll(idle);
                n_words += (u32)__builtin_popcountll(ballot64(took));  // wave-uniform
                // the new words' base windows (counted for the 32-bit filter
                // only: the wide one has no registers to spare)
                if constexpr (!D::wide) n_loads += (u32)NP;
            }
            const u64 busy = ballot64(have);
            if (!busy) {
                if (stall || next >= total) break;
                continue;
            }
            if constexpr (!D::wide) n_loads += (u32)(G * NP);  // one per stage and plane
            bool fin = false;
            if (have) {
                // every load of the step is issued before the first use
                const int ia = max(io, io_min);
                u32 raw[G][NP][NW], shg[G];
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    const u32 ig = (u32)(ia - (int)ug[g]);
                    shg[g] = ig;  // alignbit takes the low 5 bits
#pragma unroll
                    for (int pb = 0; pb < NP; ++pb) {
                        const u32 *q = plb[pb] + (ig >> 5);
#pragma unroll
                        for (u32 k = 0; k < NW; ++k) raw[g][pb][k] = q[k];
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                // live anchors 0-31, 32-63 and 64-95 cleared where a window
                // dword differs from the base, one 3-input bit operation per
                // window dword, stage by stage; `alive` counts the stages
                // after which the word still had a live anchor, so a word that
                // dies in this step is known to have needed exactly the
                // stages through the one that killed it
                u32 bw[NP][3];
#pragma unroll
                for (int pb = 0; pb < NP; ++pb) {
                    bw[pb][0] = __builtin_amdgcn_alignbit(braw[pb][1], braw[pb][0], sh);
                    bw[pb][1] = __builtin_amdgcn_alignbit(braw[pb][2], braw[pb][1], sh);
                    bw[pb][2] = __builtin_amdgcn_alignbit(braw[pb][3], braw[pb][2], sh);
                }
                u32 r0 = (u32)live, r1 = (u32)(live >> 32), r2 = live_hi, alive = 0;
#pragma unroll
                for (int g = 0; g < G; ++g) {
#pragma unroll
                    for (int pb = 0; pb < NP; ++pb) {
                        const u32 *w = raw[g][pb];
                        r0 = and_xnor(r0, bw[pb][0], __builtin_amdgcn_alignbit(w[1], w[0], shg[g]));
                        r1 = and_xnor(r1, bw[pb][1], __builtin_amdgcn_alignbit(w[2], w[1], shg[g]));
                        r2 = and_xnor(r2, bw[pb][2], __builtin_amdgcn_alignbit(w[3], w[2], shg[g]));
                    }
                    if (g + 1 < G) alive += (r0 | r1 | r2) != 0u ? 1u : 0u;
                }
                live = ((u64)r1 << 32) | r0;
                live_hi = r2;
                // first stage of this step: j, or t - Gt on a word's last
                // (clamped) step, whose earlier stages repeat passed ones
                const u32 s0 = io < io_min ? t - Gt : j;
                j += G;
                io -= G * (int)dg;
                if (!(live | live_hi)) {
                    have = false;
                    used = s0 + alive;  // stages 1 .. s0 + alive
                } else if (j >= t) {
                    have = false;
                    fin = true;
                    used = t - 1u;
                }
            }
            // a passing word takes one queue entry per half with survivors:
            // anchors 0 .. 63 (row bit k0) and 64 .. 95 (row bit k0 + 64)
            if (ballot64(fin)) {
                const u64 fin_lo = ballot64(fin && live != 0);
                const u64 fin_hi = ballot64(fin && live_hi != 0);
                if (fin) {
                    u32 slot = qn + lanes_below(fin_lo) + lanes_below(fin_hi);
                    const u32 need = (live != 0 ? 1u : 0u) + (live_hi != 0 ? 1u : 0u);
                    if (slot + need <= kQueue) {
                        if (live) {
                            qb[slot] = k0;
                            qs[slot++] = live;
                        }
                        if (live_hi) {
                            qb[slot] = k0 + 64u;
                            qs[slot] = live_hi;
                        }
                    } else {
                        // a slot this word cannot fill is left empty
                        if (slot < kQueue) qs[slot] = 0;
                        pend = true;
                    }
                }
                const u32 n = (u32)__builtin_popcountll(fin_lo) + (u32)__builtin_popcountll(fin_hi);
                if (qn + n > kQueue) stall = true;
                qn = min(qn + n, kQueue);
            }
        }
        // no word in flight: count the last words' stages, screen and
        // complete the queue, then requeue the lanes that were held back
        n_windows += (u32)NP * (WB / 32u) * used;
        used = 0;
        wave_lds_fence();
        if (qn) drain_queue<D, NP>(c, plb, qn, qb, qs, lane, n_windows, n_done);
        wave_lds_fence();
        const u64 pm_lo = ballot64(pend && live != 0);
        const u64 pm_hi = ballot64(pend && live_hi != 0);
        qn = 0;
        if (pm_lo | pm_hi) {
            if (pend) {
                u32 slot = lanes_below(pm_lo) + lanes_below(pm_hi);
                if (live) {
                    qb[slot] = k0;
                    qs[slot++] = live;
                }
                if (live_hi) {
                    qb[slot] = k0 + 64u;
                    qs[slot] = live_hi;
                }
                pend = false;
            }
            qn = (u32)__builtin_popcountll(pm_lo) + (u32)__builtin_popcountll(pm_hi);
            wave_lds_fence();
            continue;
        }
        if (next >= total) break;
    }
}

constexpr u32 kFilterWaves = 4;  // waves per filter block (blockDim 256)

// Next task of the wave: the rest of its chunk of `csz` tasks, then the
// chunk of `nsz` tasks claimed during the chunk's last task (`claim` is
// valid in lane 0).
__device__ __forceinline__ u32 next_task(u32 &k, u32 &first, u32 claim, u32 &csz, u32 nsz) {
    if (++k == csz) {
        first = __builtin_amdgcn_readfirstlane(claim);
        csz = nsz;
        k = 0;
    }
    return first + k;
}

// Tasks per queue claim: `chunk` when set, else by the profile's cover
// radius B (about the words a task holds, so the time a claim's round trip
// has to hide behind): long tasks near 2^32 share their XCD's queue best in
// small claims (4: more waves on the same few profiles' rows), short ones
// in larger claims that amortise the atomic (16).  Any size gives the same
// tasks and records.
__device__ __forceinline__ u32 claim_size(u32 chunk, u32 B) {
    return chunk ? chunk : B >= 400000u ? 4u : B >= 100000u ? 8u : 16u;
}

// Context of one (profile, denominator) task; false when the task has
// nothing to do (d beyond the prime's cover, or a threshold below 3).
template <class D>
__device__ __forceinline__ bool task_setup(const SparsePrime *P, const SparseProf *fg, u32 dd,
                                           const u32 *planes, const typename D::V *inv,
                                           const typename D::V *tab_val, const u16 *tab_lab,
                                           unsigned long long *best_slot, u32 *failed,
                                           u32 ext_limit, FilterCtx<D> &c, u32 &sp, u32 &phi) {
    const SparsePrime *sg = P + fg->prime;
    if (dd > fg->D) return false;
    const u64 key = __hip_atomic_load(best_slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the running best, or the speculative threshold when it is higher (a
    // profile whose filter finds no run reaching it is recomputed)
    const u32 q = min(max((u32)(key >> 32), fg->qspec), fg->Q);
    if (q < 3u) return false;
    u32 t = max(3u, (q + 1u) / 2u);
    if (t > q) t = q;
    sp = q - t + 1u;  // schedule period
    phi = ((t - 1u) / 2u) % sp;
    c.s = sg;
    c.f = fg;
    // rows of g = gcd(d, 6): the full row for g = 1, else the class rows
    const u32 B = fg->B, np = fg->np;
    const u32 wg = ((dd & 1u) ? 1u : 2u) * (dd % 3u ? 1u : 3u);
    const u32 row0 = first_row_of(wg);
    const u32 *prof_rows = planes + (u64)fg->plane_off * 16u;
    c.pl = prof_rows;
    c.rows = prof_rows + row_offset(row0, B, np);
    c.wpr = row_words(wg, B);
    c.R.g = wg;
    c.R.dg = dd / wg;
    c.R.nc = wg >= 3 ? 2u : 1u;
    c.R.cstride = np * c.wpr * 32u;
    for (u32 ci = 0; ci < 2; ++ci) {
        const u32 row = row0 + (ci < c.R.nc ? ci : 0u);
        c.R.rho[ci] = row_rho(row, B);
        c.R.nbits[ci] = row_bits(row, B);
    }
    c.tab_val = tab_val;
    c.tab_lab = tab_lab;
    c.best = best_slot;
    c.failed = failed;
    c.d = dd;
    c.dinv = inv[(u64)fg->inv_off + dd];
    c.t = t;
    c.q = q;
    c.B = B;
    c.bmin = ((t - 1u) * dd + 1u) / 2u;
    c.ext_limit = ext_limit;
    return true;
}

// D = Dom32 for the profiles of primes below 2^32, Dom33 for the wider
// primes (one launch per domain, each over its own range of profiles).
// Waves per SIMD the filter is compiled for: 6 (at most 80 VGPRs) for the
// 32-bit domain; the 33-bit one keeps 64-bit residues and its completion
// beside the word loop and would spill at 80 (a scratch reload per step and
// a wait on the base loads at refill), so it takes 5 waves and 88 VGPRs
// (filter of the 256 largest primes below 2^33: 80.2 -> 70.6 ms).
template <class D> struct FilterWaves { static constexpr int value = 6; };
template <> struct FilterWaves<Dom33> { static constexpr int value = 5; };

template <class D>
__global__ __launch_bounds__(256, FilterWaves<D>::value) void sp_filter(const SparsePrime *__restrict__ P,
                                                 const SparseProf *__restrict__ prof, u32 nprof,
                                                 u32 dmax, const u32 *__restrict__ tpre,
                                                 const u32 *__restrict__ planes,
                                                 const typename D::V *__restrict__ inv,
                                                 const typename D::V *__restrict__ tab_val,
                                                 const u16 *__restrict__ tab_lab,
                                                 unsigned long long *__restrict__ best,
                                                 u32 *__restrict__ failed,
                                                 unsigned long long *__restrict__ stats,
                                                 u32 *__restrict__ queue, u32 kGroup,
                                                 u32 chunk, u32 ext_limit) {
    // work counts: window comparisons in 32-anchor plane dwords, candidate
    // words, completed anchors (per lane) and the word loop's wave-level
    // load instructions (per wave), summed per block (four atomics per
    // block)
    __shared__ unsigned long long s_stats[4];
    // per-wave queue of passing words: start b0 and surviving anchors
    __shared__ u32 s_qb[kFilterWaves][kQueue];
    __shared__ u64 s_qs[kFilterWaves][kQueue];
    // per-wave word parameters of the current task's classes (16 words)
    __shared__ __attribute__((aligned(16))) u32 s_cw[kFilterWaves][16];
    if (threadIdx.x < 4) s_stats[threadIdx.x] = 0;
    __syncthreads();
    // u32 counts, added to the block's totals before they can overflow: a
    // task adds less than 2^28 to a lane's n_windows and n_done and less
    // than 2^23 to its wave's n_words
    u32 n_windows = 0, n_done = 0;  // per lane
    u32 n_words = 0;  // words taken by the wave (wave-uniform)
    u32 n_loads = 0;  // wave-level load instructions of the word loop (wave-uniform)
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    u32 *qb = s_qb[wave];
    u64 *qs = s_qs[wave];
    u32 *cw = s_cw[wave];
    // XCD-aware schedule: workgroups are dispatched round-robin over the 8
    // XCDs, so block b runs on XCD b % 8.  XCD x owns the profiles
    // k = x, x+8, x+16, ... and walks them in groups of kGroup profiles (a
    // tuning knob), denominator-major inside a group: the waves of one XCD
    // then touch the planes of at most kGroup profiles at a time (they stay
    // in its L2), while each profile's denominators are still spread over
    // the group.  Tasks are taken from a per-XCD counter (a device work
    // queue, zeroed before the launch) in chunks of `chunk` consecutive
    // tasks, so the tasks in flight on an XCD stay within a short window
    // however many waves are resident or queued; the next chunk is claimed
    // while the last task of the current one runs.  A wave whose XCD has no
    // unclaimed task left takes chunks from the other XCDs' queues in turn
    // (their planes then come through its own L2), so no XCD idles while
    // another still has work.
    const u32 lane = threadIdx.x & 63u;
    u32 lastB = 0;  // cover radius of the wave's last task (sizes its next claim)
    for (u32 v = 0; v < kXcds; ++v) {
    const u32 xcd = (blockIdx.x + v) % kXcds;
    const u32 nprof_x = nprof > xcd ? (nprof - xcd + kXcds - 1) / kXcds : 0u;
    // groups of one profile (the usual shape): the XCD's tasks are its
    // profiles' denominators 1..D back to back, tp[l] the first task of its
    // profile l and tp[nprof_x] their number, so no claimed task is empty;
    // larger groups interleave the denominators of kGroup profiles up to
    // the launch's largest D (tasks past a profile's own D return at once)
    u32 off = xcd;  // the lists before this XCD's, one entry more than profiles each
    for (u32 x = 0; x < xcd; ++x) off += nprof > x ? (nprof - x + kXcds - 1) / kXcds : 0u;
    const u32 *tp = tpre + off;
    const u32 ngroups = (nprof_x + kGroup - 1) / kGroup;
    // checked below 2^32 by the host
    const u32 items = kGroup == 1 ? (nprof_x ? tp[nprof_x] : 0u) : ngroups * kGroup * dmax;
    u32 cur = 0, lo = 0, hi = 0;  // profile of the list being walked and its tasks [lo, hi)
    u32 claim = 0;
    u32 csz = claim_size(chunk, lastB), nsz = csz;
    if (lane == 0) claim = atomicAdd(&queue[xcd], csz);
    u32 first = __builtin_amdgcn_readfirstlane(claim), k = 0;
    for (u32 item = first; item < items; item = next_task(k, first, claim, csz, nsz)) {
        if (k == csz - 1u) {
            nsz = claim_size(chunk, lastB);
            if (lane == 0) claim = atomicAdd(&queue[xcd], nsz);
        }
        u32 dd, local;
        if (kGroup == 1) {
            if (item >= hi) {
                // the profile holding `item`: tasks come in increasing order,
                // so it is at or after the current one (binary search)
                u32 a = cur, b = nprof_x - 1u;
                while (a < b) {
                    const u32 m = (a + b) >> 1;
                    if (tp[m + 1] > item) b = m;
                    else a = m + 1u;
                }
                cur = a;
                lo = tp[a];
                hi = tp[a + 1];
            }
            dd = __builtin_amdgcn_readfirstlane(1u + item - lo);
            local = cur;
        } else {
            const u32 g = item / (kGroup * dmax);
            const u32 rem = item - g * (kGroup * dmax);
            dd = __builtin_amdgcn_readfirstlane(1u + rem / kGroup);
            local = g * kGroup + rem % kGroup;
            if (local >= nprof_x) continue;
        }
        const u32 kk = __builtin_amdgcn_readfirstlane(xcd + local * kXcds);
        FilterCtx<D> c;
        u32 sp = 0, phi = 0;
        if (!task_setup<D>(P, prof + kk, dd, planes, inv, tab_val, tab_lab, &best[kk], failed,
                           ext_limit, c, sp, phi))
            continue;
        lastB = c.B;
        const SparseProf *fg = c.f;
        switch (fg->np) {
        case 1: filter_task<D, 1>(c, sp, phi, qb, qs, cw, n_windows, n_words, n_done,
                                  n_loads); break;
        case 2: filter_task<D, 2>(c, sp, phi, qb, qs, cw, n_windows, n_words, n_done,
                                  n_loads); break;
        case 3: filter_task<D, 3>(c, sp, phi, qb, qs, cw, n_windows, n_words, n_done,
                                  n_loads); break;
        default: filter_task<D, 4>(c, sp, phi, qb, qs, cw, n_windows, n_words, n_done,
                                  n_loads); break;
        }
        if ((n_windows | n_done | n_words) >= 0x80000000u) {
            atomicAdd(&s_stats[0], (unsigned long long)n_windows);
            atomicAdd(&s_stats[2], (unsigned long long)n_done);
            if (lane == 0) {
                atomicAdd(&s_stats[1], (unsigned long long)n_words);
                atomicAdd(&s_stats[3], (unsigned long long)n_loads);
            }
            n_windows = n_words = n_done = n_loads = 0;
        }
    }
    }
    atomicAdd(&s_stats[0], (unsigned long long)n_windows);
    atomicAdd(&s_stats[2], (unsigned long long)n_done);
    if (lane == 0) {
        atomicAdd(&s_stats[1], (unsigned long long)n_words);
        atomicAdd(&s_stats[3], (unsigned long long)n_loads);
    }
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&stats[threadIdx.x], s_stats[threadIdx.x]);
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
                                         const uint64_t *tab_val64, const uint16_t *tab_lab,
                                         uint16_t *labels, uint32_t *inv, uint64_t *inv64,
                                         uint32_t *planes, uint32_t *err, hipStream_t stream) {
    hipLaunchKernelGGL(rbg::sp_labels, dim3(label_blocks), dim3(256), 0, stream, P, nprimes,
                       tab_val, tab_val64, tab_lab, labels, err);
    hipLaunchKernelGGL(rbg::sp_labels_rest, dim3(label_blocks), dim3(256), 0, stream, P, nprimes,
                       labels);
    hipLaunchKernelGGL(rbg::sp_inverses, dim3(inverse_blocks), dim3(256), 0, stream, P, nprimes,
                       inv, inv64);
    hipLaunchKernelGGL(rbg::sp_planes, dim3(plane_blocks), dim3(256), 0, stream, P, nprimes, prof,
                       nprof, labels, planes);
    hipLaunchKernelGGL(rbg::sp_rows, dim3(plane_blocks), dim3(256), 0, stream, P, nprimes, prof,
                       planes);
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
    hipLaunchKernelGGL(rbg::sp_seed_merge, dim3((nprof + 3) / 4), dim3(256), 0, stream, P, prof,
                       nprof, (const rbg::Agg *)segs, segs_per_prof, best, failed);
    return hipGetLastError();
}

// prof / best: the first profile of the launch and its key; wide = 0 for
// profiles of primes below 2^32 (inv and tab_val hold u32 values), 1 for the
// wider primes (u64 values).
extern "C" hipError_t rbg_sparse_filter(const SparsePrime *P, const SparseProf *prof,
                                        uint32_t nprof, uint32_t dmax, const uint32_t *tpre,
                                        const uint32_t *planes,
                                        const void *inv, const void *tab_val,
                                        const uint16_t *tab_lab, unsigned long long *best,
                                        uint32_t *failed, unsigned long long *stats,
                                        uint32_t *queue, int wide, uint32_t grid, uint32_t group,
                                        uint32_t chunk, uint32_t extend_limit,
                                        hipStream_t stream) {
    if (wide)
        hipLaunchKernelGGL(rbg::sp_filter<rbg::Dom33>, dim3(grid), dim3(256), 0, stream, P, prof,
                           nprof, dmax, tpre, planes, (const uint64_t *)inv, (const uint64_t *)tab_val,
                           tab_lab, best, failed, stats, queue, group, chunk, extend_limit);
    else
        hipLaunchKernelGGL(rbg::sp_filter<rbg::Dom32>, dim3(grid), dim3(256), 0, stream, P, prof,
                           nprof, dmax, tpre, planes, (const uint32_t *)inv, (const uint32_t *)tab_val,
                           tab_lab, best, failed, stats, queue, group, chunk, extend_limit);
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
