// Dense census scan for gfx950 (the "plain method").
//
// For each prime p of a block and each admissible order r, the colour of x is
// cl_r(x) = ind(x) mod r.  With m = lcm of the admissible orders, y = x^((p-1)/m)
// is an m-th root of unity z^j and j mod r is the colour (any z of exact order
// m gives a unit relabelling, which preserves colour equality; reference
// paper/GPU_METHOD_EN.tex:229-244, frozen_entry.cu:15-24).  Only the half
// interval 1..(p-1)/2 is scanned; reflection cl(p-x) = cl(x) + sigma recovers
// the full sequence (GPU_METHOD_EN.tex:175-177; original_scan_gpu.cu:380-382).
//
// Work decomposition (MI355X-first, not the reference's thread-per-segment):
//   * one wave = one segment of `chunks` x 64 consecutive positions; lane i
//     owns position base+i of each 64-position chunk;
//   * per chunk every lane evaluates its colour code (Montgomery power + LDS
//     root-table search), one DPP wave shift gives the left neighbour's code,
//     and one 64-lane ballot per admissible order yields the uint64 change
//     mask of that order;
//   * run bookkeeping (prefix, suffix, longest run with least start, A4
//     windows) is done on those masks with scalar (SALU) bit operations, so
//     it is wave-uniform and overlaps the vector work of other waves;
//   * the waves of a block hold consecutive segments and are merged in LDS
//     into one summary per (tile, order); tiles of all primes of the block
//     form one flattened space, consumed either by a persistent grid through
//     a device work counter (batched method) or by one grid per prime
//     (plain method);
//   * merge_primes reduces the tile summaries of each prime in order (the
//     associative run monoid) and finalises M, L, witness, K and A4.
#include <hip/hip_runtime.h>

#include "rabung.h"
#include "rbg_device.hpp"
#include "rbg_runs.hpp"
#include "rbg_types.hpp"

namespace rbg {

// Residue constants of a prime in the kernel's domain (the 33rd bits of the
// wide primes are kept in d.hi).
template <class D>
struct PrimeConsts {
    D dom;
    typename D::V one, r2, step;
};

__device__ __forceinline__ PrimeConsts<Dom32> consts_of(const PrimeDesc &d, Dom32) {
    return {{d.p, pos_inverse(d.ninv)}, d.one, d.r2, d.step};
}

__device__ __forceinline__ PrimeConsts<Dom33> consts_of(const PrimeDesc &d, Dom33) {
    const auto w = [&](u32 lo, int bit) { return (u64)lo | ((u64)((d.hi >> bit) & 1u) << 32); };
    return {{w(d.p, 0), d.ninv}, w(d.one, 1), w(d.r2, 2), w(d.step, 3)};
}

constexpr int kPowChains = 4;  // chunks whose powers a lane computes together

// y[i] = x[i]^e (Montgomery forms) for K independent bases and one exponent:
// the chains share the square-and-multiply schedule, so each multiply of
// one chain hides the latency of the others.
template <class D, int K>
__device__ __forceinline__ void pow_many(const D &dm, const typename D::V *x, u32 e,
                                         typename D::V one, typename D::V *y) {
#pragma unroll
    for (int i = 0; i < K; ++i) y[i] = e ? x[i] : one;
    if (e < 2) return;
    const int top = 31 - __builtin_clz(e);
    for (int bit = top - 1; bit >= 0; --bit) {
#pragma unroll
        for (int i = 0; i < K; ++i) y[i] = dm.mul(y[i], y[i]);
        if ((e >> bit) & 1u) {
#pragma unroll
            for (int i = 0; i < K; ++i) y[i] = dm.mul(y[i], x[i]);
        }
    }
}

// Run bookkeeping of one wave over its segment: per admissible order the
// run state, fed one 64-position chunk of colour codes at a time.
struct WaveRuns {
    OrderState st[8];
    u32 a4;
    u64 prevN;
    u32 a4v;         // per-lane A4 counts (the vector-unit form of runs_chunk)
    u32 pn62, pn63;  // r = 2 same-as-left bits of the previous chunk's lanes 62, 63
    u32 carry_code, first_code, last_code;
};

__device__ __forceinline__ void runs_init(WaveRuns &w, const PrimeDesc &d) {
#pragma unroll
    for (int k = 0; k < 8; ++k) order_init(w.st[k], run_floor(d, (u32)k));
    w.a4 = 0;
    w.prevN = 0;
    w.a4v = 0;
    w.pn62 = w.pn63 = 0;
    w.carry_code = w.first_code = w.last_code = 0;
}

// Chunk c (positions base .. base+valid-1 of the segment) with this lane's
// colour code: one DPP shift for the left neighbour, one ballot per order.
// kVecReject: the test that lets most chunks skip the interior-run search
// (no `best` consecutive equal neighbours) runs on the vector unit (a prefix
// maximum over the lanes, nine VALU instructions) instead of the scalar
// unit's doubling (about 30 scalar instructions), and so does the A4 count
// (three consecutive equal neighbours of r = 2, per lane, summed at the end
// of the segment); for the label-reuse scan, whose scalar unit is the busier
// one.
template <bool kVecReject = false>
__device__ __forceinline__ void runs_chunk(WaveRuns &w, u32 code, u32 c, u32 base, u32 valid,
                                           u32 mask) {
    const u64 vmask = valid == 64u ? ~0ull : ((1ull << valid) - 1ull);
    const u32 left = shift_up_one(code, w.carry_code);
    const u32 diff = code ^ left;
    const u32 lane = threadIdx.x & 63u;
    // lanes whose bit of N is clear whatever the colours: past the valid
    // positions, and position 0 of the segment (a forced change)
    const u32 zfix = (lane >= valid || (c == 0 && lane == 0)) ? lane + 1u : 0u;
    if (c == 0) w.first_code = __builtin_amdgcn_readfirstlane(code);
    w.carry_code = __builtin_amdgcn_readlane(code, 63);
    w.last_code = __builtin_amdgcn_readlane(code, valid - 1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const u32 r = (u32)k + 2u;
        if (!(mask & (1u << k))) continue;
        const bool chg = (diff & plane_bits(r)) != 0u;
        u64 C = ballot64(chg);
        if (c == 0) C |= 1ull;
        C &= vmask;
        const u64 N = ~C & vmask;
        bool cand = true;
        if constexpr (kVecReject) {
            const u32 b = w.st[k].best;
            if (b != 0u && b <= 62u) cand = has_ones_run(chg ? lane + 1u : zfix, b);
        }
        if (r == 2) {
            if constexpr (kVecReject) {
                // lane i: same-as-left at i, i-1 and i-2 (the previous
                // chunk's lanes 63 and 62 before lane 0)
                const u32 nb = (chg || zfix) ? 0u : 1u;
                const u32 s1 = shift_up_one(nb, w.pn63);
                const u32 s2 = shift_up_one(s1, w.pn62);
                w.a4v += nb & s1 & s2;
                w.pn63 = __builtin_amdgcn_readlane(nb, 63);
                w.pn62 = __builtin_amdgcn_readlane(nb, 62);
            } else {
                u64 W = N & ((N << 1) | (w.prevN >> 63)) & ((N << 2) | (w.prevN >> 62));
                w.a4 += (u32)__builtin_popcountll(W);
                w.prevN = N;
            }
        }
        order_update(w.st[k], C, N, base, cand);
    }
}

// Per-order summaries of the segment (positions relative to its start)
// into out[r-2]; codes_out = first code | last code << 16.
template <bool kVecReject = false>
__device__ __forceinline__ void runs_finish(WaveRuns &w, u32 seg_len, u32 mask, RunSum *out,
                                            u32 &codes_out) {
    const u32 lane = threadIdx.x & 63u;
    if constexpr (kVecReject) {
        u32 a4 = w.a4v;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) a4 += (u32)__shfl_xor((int)a4, o);
        w.a4 = a4;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!(mask & (1u << k))) continue;
        OrderState &s = w.st[k];
        const u32 suffix = order_finish(s, seg_len);
        if (lane == 0) {
            RunSum o;
            o.prefix = s.prefix;
            o.suffix = suffix;
            o.best = s.best;
            o.bstart = s.bstart;
            o.a4 = (k == 0) ? w.a4 : 0;
            out[k] = o;
        }
    }
    codes_out = w.first_code | (w.last_code << 16);
}

// Scan one segment [seg_start, seg_start+seg_len) with the calling wave.
// Writes per-order summaries (positions relative to seg_start) into out[r-2].
template <class D>
__device__ void wave_scan(const PrimeDesc &d, const PrimeConsts<D> &k, u32 seg_start, u32 seg_len,
                          const typename D::V *s_val, const u16 *s_code, RunSum *out,
                          u32 &codes_out, u32 *err) {
    typedef typename D::V V;
    const u32 lane = threadIdx.x & 63u;
    const u32 e = d.e, m = d.m, mask = d.mask;
    WaveRuns w;
    runs_init(w, d);
    // Montgomery form of this lane's first position
    V xm = k.dom.mul((V)(seg_start + lane), k.r2);
    const u32 nchunks = (seg_len + 63u) >> 6;
    u32 bad = 0;
    for (u32 c0 = 0; c0 < nchunks; c0 += kPowChains) {
        // colour codes of x: y = x^e for kPowChains chunks at once, then
        // each y's index in the sorted root table, chunk by chunk
        V xs[kPowChains], ys[kPowChains];
        xs[0] = xm;
#pragma unroll
        for (int i = 1; i < kPowChains; ++i) xs[i] = k.dom.add(xs[i - 1], k.step);
        pow_many<D, kPowChains>(k.dom, xs, e, k.one, ys);
        xm = k.dom.add(xs[kPowChains - 1], k.step);
#pragma unroll
        for (int ci = 0; ci < kPowChains; ++ci) {
            const u32 c = c0 + (u32)ci;
            if (c >= nchunks) break;
            const u32 base = c << 6;
            const u32 valid = min(64u, seg_len - base);
            const V y = ys[ci];
            const u32 lo = table_index(s_val, m, y);
            bad |= (s_val[lo] != y) & (lane < valid);
            runs_chunk(w, s_code[lo], c, base, valid, mask);
        }
    }
    if (__builtin_amdgcn_ballot_w64(bad != 0)) atomicOr(err, 1u);
    runs_finish(w, seg_len, mask, out, codes_out);
}

// ---- label reuse (the plain method for primes below 2^32) ----------------
//
// The colour of x depends on its discrete-log label j(x) (x^e = z^j) only,
// and labels are additive: j(q u) = j(q) + j(u) mod m.  So only positions
// coprime to 6 need a Montgomery power: an even x takes j(x/2) + j(2) and
// an odd multiple of 3 takes j(x/3) + j(3) from a label array in HBM (one
// u16 per position up to nh/2), written by the tiles of lower positions.
// The reference's V2b variant computes powers for odd positions only
// (src/census/engine32/build/original_scan_gpu.cu:386-390); here the wheel
// is 6, a third of the positions.
//
// Tiles run in phases of doubling position ranges (launches on one
// stream): a phase reads only labels of earlier phases (x/2 <= end/2 <=
// the previous phase's end) and of a seed range computed by direct powers.
// In a wave, lane l computes the power of the l-th position coprime to 6 of
// a 192-position group (x = g + 6 (l >> 1) + 4 (l & 1), g = 1 mod 6), four
// groups at once (768 positions, 12 chunks); the labels of the other
// positions are gathered from HBM before the powers start, so their latency
// hides behind the products.  Segments start at 1 mod 6 (their length is a
// multiple of 192).

constexpr u32 kGroup = 192;                       // positions per 64 coprime ones
constexpr u32 kReuseChunks = kPowChains * 3;      // chunks per pass (4 groups)

__device__ __forceinline__ void wave_scan_reuse(const PrimeDesc &d, const PrimeConsts<Dom32> &k,
                                                u32 seg_start, u32 seg_len, const u32 *s_val,
                                                const u16 *s_lab, const u16 *s_code, u16 *s_jp,
                                                u16 *lab, u32 lab_from, u32 lab_to, u32 j2, u32 j3,
                                                RunSum *out, u32 &codes_out, u32 *err) {
    const u32 lane = threadIdx.x & 63u;
    const u32 e = d.e, m = d.m, mask = d.mask;
    const Dom32 &dm = k.dom;
    WaveRuns w;
    runs_init(w, d);
    // per chunk phase s = c mod 3: offset o = 64 s + lane in its group and
    // the kind of x = g + o (0: coprime to 6, 1: even, 2: odd multiple of 3)
    u32 kind[3];
#pragma unroll
    for (u32 s = 0; s < 3; ++s) {
        const u32 rr = (64u * s + lane) % 6u;
        kind[s] = (rr == 0u || rr == 4u) ? 0u : (rr & 1u) ? 1u : 2u;
    }
    const u32 step = dm.mul(kGroup, k.r2);  // Montgomery form of 192
    // this lane's coprime position in each group: g + 6 (lane >> 1) + 4 (lane & 1)
    const u32 coff0 = 6u * (lane >> 1) + 4u * (lane & 1u);
    u32 xm = dm.mul(seg_start + coff0, k.r2);
    const u32 nchunks = (seg_len + 63u) >> 6;
    u32 bad = 0;
    for (u32 c0 = 0; c0 < nchunks; c0 += kReuseChunks) {
        // labels of the non-coprime positions (loads in flight during the powers)
        u32 lv[kReuseChunks];
#pragma unroll
        for (u32 c = 0; c < kReuseChunks; ++c) {
            const u32 off = ((c0 + c) << 6) + lane;
            const u32 x = seg_start + off;
            const u32 kd = kind[c % 3u];
            lv[c] = 0;
            if (kd != 0u && off < seg_len) lv[c] = lab[kd == 1u ? x >> 1 : x / 3u];
        }
        u32 xs[kPowChains], ys[kPowChains];
        xs[0] = xm;
#pragma unroll
        for (int i = 1; i < kPowChains; ++i) xs[i] = dm.add(xs[i - 1], step);
        pow_many<Dom32, kPowChains>(dm, xs, e, k.one, ys);
        xm = dm.add(xs[kPowChains - 1], step);
        // every label of the pass into s_jp (pass-relative positions)
#pragma unroll
        for (int i = 0; i < kPowChains; ++i) {
            const u32 lo = table_index(s_val, m, ys[i]);
            const u32 coff = (u32)i * kGroup + coff0;
            bad |= (s_val[lo] != ys[i]) & ((c0 << 6) + coff < seg_len);
            s_jp[coff] = s_lab[lo];
        }
#pragma unroll
        for (u32 c = 0; c < kReuseChunks; ++c) {
            const u32 kd = kind[c % 3u];
            u32 jr = lv[c] + (kd == 1u ? j2 : j3);
            jr = jr >= m ? jr - m : jr;
            bad |= jr >= m;  // a label never written (a phase out of order)
            if (kd != 0u) s_jp[(c << 6) + lane] = (u16)min(jr, m - 1u);
        }
        wave_lds_fence();
        const u32 cend = min(kReuseChunks, nchunks - c0);
        for (u32 c = 0; c < cend; ++c) {
            const u32 cc = c0 + c;
            const u32 base = cc << 6;
            const u32 valid = min(64u, seg_len - base);
            const u32 j = s_jp[(c << 6) + lane];
            const u32 x = seg_start + base + lane;
            if (lane < valid && x > lab_from && x <= lab_to) lab[x] = (u16)j;
            runs_chunk<true>(w, s_code[j], cc, base, valid, mask);
        }
        wave_lds_fence();  // the next pass rewrites s_jp
    }
    if (__builtin_amdgcn_ballot_w64(bad != 0)) atomicOr(err, 1u);
    runs_finish<true>(w, seg_len, mask, out, codes_out);
}

constexpr int kWavesPerBlock = 4;

// Wave summaries of one tile, merged into the tile's summary.
struct TileWaves {
    RunSum run[kWavesPerBlock][8];
    u32 codes[kWavesPerBlock];
    u32 first[kWavesPerBlock];
    u32 len[kWavesPerBlock];
};

// Every wave of the block calls this after its segment: the per-order
// merge of the waves' summaries in position order, written as tile lt of
// prime d (global tile number `tile`).  Ends with a block barrier.
__device__ __forceinline__ void write_tile(TileWaves &tw, u32 wave, const PrimeDesc &d, u32 lt,
                                           u32 tile, u32 seg_start, u32 seg_len, u32 codes,
                                           TileHdr *__restrict__ hdr, RunSum *__restrict__ runs) {
    if ((threadIdx.x & 63u) == 0) {
        tw.codes[wave] = codes;
        tw.first[wave] = seg_start;
        tw.len[wave] = seg_len;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const u32 k = threadIdx.x, r = k + 2;
        if (d.mask & (1u << k)) {
            const u32 planes = plane_bits(r);
            Agg acc;
            acc.len = 0;
            for (int w = 0; w < kWavesPerBlock; ++w) {
                Agg b;
                b.first_x = tw.first[w];
                b.len = tw.len[w];
                b.fcode = tw.codes[w] & 0xFFFFu;
                b.lcode = tw.codes[w] >> 16;
                b.prefix = tw.run[w][k].prefix;
                b.suffix = tw.run[w][k].suffix;
                b.best = tw.run[w][k].best;
                b.bstart = tw.run[w][k].bstart + tw.first[w];
                b.a4 = tw.run[w][k].a4;
                acc = merge(acc, b, planes);
            }
            const u32 kk = __builtin_popcount(d.mask & ((1u << k) - 1u));
            RunSum o;
            o.prefix = acc.prefix;
            o.suffix = acc.suffix;
            o.best = acc.best;
            o.bstart = acc.bstart;
            o.a4 = acc.a4;
            runs[(u64)d.sum_off + (u64)lt * d.nr + kk] = o;
            if (k == 0) {
                TileHdr h;
                h.first_x = acc.first_x;
                h.len = acc.len;
                h.codes = acc.fcode | (acc.lcode << 16);
                h.pad = 0;
                hdr[tile] = h;
            }
        }
    }
    __syncthreads();
}

// D = Dom32 for primes below 2^32, Dom33 for 2^32 < p < 2^33 (the tiles of
// one launch belong to primes of one domain).
template <class D>
__global__ __launch_bounds__(kWavesPerBlock * 64) void scan_tiles(
    const PrimeDesc *__restrict__ descs, u32 nprimes, u32 tile_first, u32 tile_end, u32 chunks,
    const typename D::V *__restrict__ tab_val, const u16 *__restrict__ tab_code,
    TileHdr *__restrict__ hdr, RunSum *__restrict__ runs, u32 *__restrict__ counter,
    u32 *__restrict__ err, int persistent) {
    __shared__ typename D::V s_val[kMaxM];
    __shared__ u16 s_code[kMaxM];
    __shared__ TileWaves s_tw;
    __shared__ u32 s_tile, s_pi;
    // wave-uniform: everything derived from it (segment bounds, chunk masks,
    // the run state of every order) then lives in scalar registers and runs
    // on the scalar unit beside the vector powers
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const u32 seg_positions = chunks * 64u;
    u32 cur_pi = 0xFFFFFFFFu;
    for (u32 iter = 0;; ++iter) {
        if (threadIdx.x == 0) {
            u32 t;
            if (persistent)
                t = tile_first + atomicAdd(counter, 1u);
            else
                t = iter == 0 ? tile_first + blockIdx.x : tile_end;
            u32 pi = cur_pi;
            if (t < tile_end) {
                if (pi == 0xFFFFFFFFu || descs[pi].tile_begin > t) {
                    u32 lo = 0, hi = nprimes;  // last prime with tile_begin <= t
                    while (hi - lo > 1) {
                        u32 mid = (lo + hi) >> 1;
                        if (descs[mid].tile_begin <= t) lo = mid; else hi = mid;
                    }
                    pi = lo;
                }
                while (t >= descs[pi].tile_begin + descs[pi].ntiles) ++pi;
            }
            s_tile = t;
            s_pi = pi;
        }
        __syncthreads();
        const u32 tile = __builtin_amdgcn_readfirstlane(s_tile);
        if (tile >= tile_end) break;
        const u32 pi = __builtin_amdgcn_readfirstlane(s_pi);
        const PrimeDesc d = descs[pi];
        if (pi != cur_pi) {
            const u32 voff = D::wide ? d.tab64 : d.tab_off;
            for (u32 j = threadIdx.x; j < d.m; j += blockDim.x) {
                s_val[j] = tab_val[voff + j];
                s_code[j] = tab_code[d.tab_off + j];
            }
            cur_pi = pi;
            __syncthreads();
        }
        const u32 lt = tile - d.tile_begin;
        const u64 seg = (u64)lt * kWavesPerBlock + wave;
        const u64 seg_start = 1ull + seg * seg_positions;
        const u32 seg_len =
            seg_start > d.nh ? 0u : (u32)min((u64)seg_positions, (u64)d.nh - seg_start + 1ull);
        u32 codes = 0;
        wave_scan<D>(d, consts_of(d, D{}), (u32)seg_start, seg_len, s_val, s_code, s_tw.run[wave],
                     codes, err);
        write_tile(s_tw, wave, d, lt, tile, (u32)seg_start, seg_len, codes, hdr, runs);
    }
}

// The piece of a batched launch that block b belongs to (pieces sorted by
// block0, the first at 0).
__device__ __forceinline__ u32 piece_of(const ReusePiece *pieces, u32 npieces, u32 b) {
    u32 lo = 0, hi = npieces;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (pieces[mid].block0 <= b) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr u32 kSeedPerThread = 16;  // seed positions per thread

// Label seed of the reuse phases, for several primes at once: lab[x] = j(x)
// for 1 <= x <= n by direct powers (the positions below the first phase's
// reads), kSeedPerThread positions per thread.
__global__ __launch_bounds__(256) void label_seed(const PrimeDesc *__restrict__ descs,
                                                  const ReusePiece *__restrict__ pieces,
                                                  u32 npieces, const u32 *__restrict__ tab_val,
                                                  const u16 *__restrict__ tab_lab,
                                                  u16 *__restrict__ lab_base, u32 n,
                                                  u32 *__restrict__ err) {
    __shared__ u32 s_val[kMaxM];
    __shared__ u16 s_lab[kMaxM];
    const ReusePiece pc = pieces[__builtin_amdgcn_readfirstlane(piece_of(pieces, npieces, blockIdx.x))];
    const PrimeDesc d = descs[pc.pi];
    u16 *lab = lab_base + pc.lab_off;
    for (u32 j = threadIdx.x; j < d.m; j += blockDim.x) {
        s_val[j] = tab_val[d.tab_off + j];
        s_lab[j] = tab_lab[d.tab_off + j];
    }
    __syncthreads();
    const PrimeConsts<Dom32> k = consts_of(d, Dom32{});
    u32 bad = 0;
    const u32 x0 = 1u + (blockIdx.x - pc.block0) * (256u * kSeedPerThread) + threadIdx.x;
    for (u32 i = 0; i < kSeedPerThread; ++i) {
        const u32 x = x0 + i * 256u;
        if (x > n) break;
        const u32 y = k.dom.pow(k.dom.mul(x, k.r2), d.e, k.one);
        const u32 lo = table_index(s_val, d.m, y);
        bad |= s_val[lo] != y;
        lab[x] = s_lab[lo];
    }
    if (bad) atomicOr(err, 1u);
}

// One phase of the label-reuse scan for several primes at once: block b of
// piece (prime pi) scans its prime's tile tile_lo + b - block0 (one tile per
// workgroup).  The prime's labels hold positions 1 .. lab_from (seed) and
// every position of the earlier phases up to lab_to = nh / 2; this phase
// adds its own up to lab_to.
__global__ __launch_bounds__(kWavesPerBlock * 64) void scan_reuse(
    const PrimeDesc *__restrict__ descs, const ReusePiece *__restrict__ pieces, u32 npieces,
    u32 chunks, const u32 *__restrict__ tab_val, const u16 *__restrict__ tab_lab, u16 *lab_base,
    u32 lab_from, TileHdr *__restrict__ hdr, RunSum *__restrict__ runs, u32 *__restrict__ err) {
    __shared__ u32 s_val[kMaxM];
    __shared__ u16 s_lab[kMaxM];   // sorted root-table index -> label j
    __shared__ u16 s_code[kMaxM];  // label j -> colour code
    __shared__ u16 s_jp[kWavesPerBlock][kReuseChunks * 64];
    __shared__ TileWaves s_tw;
    const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const ReusePiece pc = pieces[__builtin_amdgcn_readfirstlane(piece_of(pieces, npieces, blockIdx.x))];
    const PrimeDesc d = descs[pc.pi];
    for (u32 j = threadIdx.x; j < d.m; j += blockDim.x) {
        s_val[j] = tab_val[d.tab_off + j];
        s_lab[j] = tab_lab[d.tab_off + j];
        s_code[j] = (u16)colour_code(j, d.m);
    }
    __syncthreads();
    const u32 lt = pc.tile_lo + (blockIdx.x - pc.block0);
    const u32 seg_positions = chunks * 64u;
    const u64 seg = (u64)lt * kWavesPerBlock + wave;
    const u64 seg_start = 1ull + seg * seg_positions;
    const u32 seg_len =
        seg_start > d.nh ? 0u : (u32)min((u64)seg_positions, (u64)d.nh - seg_start + 1ull);
    u32 codes = 0;
    wave_scan_reuse(d, consts_of(d, Dom32{}), (u32)seg_start, seg_len, s_val, s_lab, s_code,
                    s_jp[wave], lab_base + pc.lab_off, lab_from, d.nh / 2u, pc.j2, pc.j3,
                    s_tw.run[wave], codes, err);
    write_tile(s_tw, wave, d, lt, d.tile_begin + lt, (u32)seg_start, seg_len, codes, hdr, runs);
}

constexpr int kMergeThreads = 256;

// One block per prime: ordered reduction of its tile summaries per order,
// then the reflection junction and the output contract.
__global__ __launch_bounds__(kMergeThreads) void merge_primes(const PrimeDesc *__restrict__ descs,
                                                              const TileHdr *__restrict__ hdr,
                                                              const RunSum *__restrict__ runs,
                                                              rbg_record *__restrict__ out,
                                                              u32 *__restrict__ err) {
    __shared__ Agg s_agg[kMergeThreads];
    const PrimeDesc d = descs[blockIdx.x];
    const u32 t = threadIdx.x;
    const u32 t0 = (u32)((u64)d.ntiles * t / kMergeThreads);
    const u32 t1 = (u32)((u64)d.ntiles * (t + 1) / kMergeThreads);
    rbg_record *rec = &out[d.index];
    if (t == 0) {
        rec->p = d.p;
        rec->mask = d.mask;
        for (int k = 0; k < 8; ++k) {
            rec->M[k] = 0;
            rec->L[k] = 0;
            rec->W[k] = 0;
            rec->K[k] = 0;
        }
        rec->A4 = 0;
        rec->status = 0;
    }
    u32 kk = 0;
    for (int k = 0; k < 8; ++k) {
        if (!(d.mask & (1u << k))) continue;
        const u32 r = (u32)k + 2u;
        const u32 planes = plane_bits(r);
        Agg acc;
        acc.len = 0;
        for (u32 i = t0; i < t1; ++i) {
            const TileHdr h = hdr[d.tile_begin + i];
            const RunSum s = runs[(u64)d.sum_off + (u64)i * d.nr + kk];
            Agg b;
            b.first_x = h.first_x;
            b.len = h.len;
            b.fcode = h.codes & 0xFFFFu;
            b.lcode = h.codes >> 16;
            b.prefix = s.prefix;
            b.suffix = s.suffix;
            b.best = s.best;
            b.bstart = s.bstart;
            b.a4 = s.a4;
            acc = merge(acc, b, planes);
        }
        s_agg[t] = acc;
        __syncthreads();
        for (u32 stride = 1; stride < kMergeThreads; stride <<= 1) {
            if ((t % (2 * stride)) == 0) s_agg[t] = merge(s_agg[t], s_agg[t + stride], planes);
            __syncthreads();
        }
        if (t == 0) {
            const Agg S = s_agg[0];
            if (S.len != d.nh || S.first_x != 1u) atomicOr(err, 2u);
            // runs below the floor were not tracked: exact only if a run
            // reached it
            if (S.best < run_floor(d, (u32)k)) atomicOr(err, 4u);
            const u64 p = (u64)d.p | ((u64)(d.hi & 1u) << 32);
            const bool sigma0 = (((p - 1u) / 2u) % r) == 0u;
            u32 M = S.best, W = S.bstart;
            if (sigma0) take(2u * S.suffix, d.nh - S.suffix + 1u, M, W);  // central junction
            u32 L = S.prefix;
            if (sigma0 && S.prefix == d.nh) L = (u32)(p - 1u);
            const u64 bound = sigma0 ? 2ull * L + 2ull : (u64)L + 2ull;
            rec->M[k] = M;
            rec->L[k] = L;
            rec->W[k] = W;
            rec->K[k] = max((u64)M + 1ull, bound);
            if (r == 2)
                rec->A4 = sigma0 ? 2u * (S.a4 - f4(S.suffix)) + f4(2u * S.suffix) : 2u * S.a4;
        }
        __syncthreads();
        ++kk;
    }
}

// Fills n words with v (used instead of hipMemsetAsync so that a chain can
// be recorded into a graph).
__global__ void fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = v;
}

}  // namespace rbg

// ---- launch wrappers used by the host orchestration -----------------------

extern "C" hipError_t rbg_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t stream) {
    if (!n) return hipSuccess;
    const uint32_t blocks = (n + 255u) / 256u < 1024u ? (n + 255u) / 256u : 1024u;
    hipLaunchKernelGGL(rbg::fill_u32, dim3(blocks), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

// wide = 0: tiles of primes below 2^32 (tab_val: u32 table values);
// wide = 1: tiles of primes between 2^32 and 2^33 (u64 table values).
extern "C" hipError_t rbg_launch_scan(const rbg::PrimeDesc *descs, uint32_t nprimes,
                                      uint32_t tile_first, uint32_t tile_end, uint32_t chunks,
                                      const void *tab_val, const uint16_t *tab_code,
                                      rbg::TileHdr *hdr, rbg::RunSum *runs, uint32_t *counter,
                                      uint32_t *err, int persistent, int wide, uint32_t grid,
                                      hipStream_t stream) {
    if (wide)
        hipLaunchKernelGGL(rbg::scan_tiles<rbg::Dom33>, dim3(grid), dim3(rbg::kWavesPerBlock * 64),
                           0, stream, descs, nprimes, tile_first, tile_end, chunks,
                           (const uint64_t *)tab_val, tab_code, hdr, runs, counter, err,
                           persistent);
    else
        hipLaunchKernelGGL(rbg::scan_tiles<rbg::Dom32>, dim3(grid), dim3(rbg::kWavesPerBlock * 64),
                           0, stream, descs, nprimes, tile_first, tile_end, chunks,
                           (const uint32_t *)tab_val, tab_code, hdr, runs, counter, err,
                           persistent);
    return hipGetLastError();
}

// Label-reuse scan of descriptor pi (a prime below 2^32): the seed of
// labels 1 .. lab_from, then phases of doubling tile ranges, each phase
// reading only labels of positions below its first tile.  first_phase = the
// tiles of the first phase (2 lab_from >= first_phase * tile positions).
// Batched label-reuse launches (plain method): the seed of every piece's
// prime (pieces[0 .. npieces), nblocks blocks of kSeedPerThread x 256
// positions), or one phase (one block per tile).
extern "C" hipError_t rbg_launch_label_seed(const rbg::PrimeDesc *descs,
                                            const rbg::ReusePiece *pieces, uint32_t npieces,
                                            uint32_t nblocks, const uint32_t *tab_val,
                                            const uint16_t *tab_lab, uint16_t *lab,
                                            uint32_t lab_from, uint32_t *err,
                                            hipStream_t stream) {
    hipLaunchKernelGGL(rbg::label_seed, dim3(nblocks), dim3(256), 0, stream, descs, pieces,
                       npieces, tab_val, tab_lab, lab, lab_from, err);
    return hipGetLastError();
}

extern "C" hipError_t rbg_launch_scan_reuse(const rbg::PrimeDesc *descs,
                                            const rbg::ReusePiece *pieces, uint32_t npieces,
                                            uint32_t nblocks, uint32_t chunks,
                                            const uint32_t *tab_val, const uint16_t *tab_lab,
                                            uint16_t *lab, uint32_t lab_from, rbg::TileHdr *hdr,
                                            rbg::RunSum *runs, uint32_t *err,
                                            hipStream_t stream) {
    if (chunks % rbg::kReuseChunks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(rbg::scan_reuse, dim3(nblocks), dim3(rbg::kWavesPerBlock * 64), 0, stream,
                       descs, pieces, npieces, chunks, tab_val, tab_lab, lab, lab_from, hdr, runs,
                       err);
    return hipGetLastError();
}

extern "C" uint32_t rbg_seed_per_block(void) { return 256u * rbg::kSeedPerThread; }

extern "C" hipError_t rbg_launch_merge(const rbg::PrimeDesc *descs, uint32_t nprimes,
                                       const rbg::TileHdr *hdr, const rbg::RunSum *runs,
                                       rbg_record *out, uint32_t *err, hipStream_t stream) {
    hipLaunchKernelGGL(rbg::merge_primes, dim3(nprimes), dim3(rbg::kMergeThreads), 0, stream,
                       descs, hdr, runs, out, err);
    return hipGetLastError();
}

extern "C" int rbg_waves_per_block(void) { return rbg::kWavesPerBlock; }
