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

// Scan one segment [seg_start, seg_start+seg_len) with the calling wave.
// Writes per-order summaries (positions relative to seg_start) into out[r-2].
template <class D>
__device__ void wave_scan(const PrimeDesc &d, const PrimeConsts<D> &k, u32 seg_start, u32 seg_len,
                          const typename D::V *s_val, const u16 *s_code, RunSum *out,
                          u32 &codes_out, u32 *err) {
    typedef typename D::V V;
    const u32 lane = threadIdx.x & 63u;
    const u32 e = d.e, m = d.m, mask = d.mask;
    OrderState st[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) order_init(st[k]);
    u32 a4 = 0;
    u64 prevN = 0;
    u32 carry_code = 0, first_code = 0, last_code = 0;
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
        const u64 vmask = valid == 64u ? ~0ull : ((1ull << valid) - 1ull);
        const V y = ys[ci];
        const u32 lo = table_index(s_val, m, y);
        bad |= (s_val[lo] != y) & (lane < valid);
        u32 code = s_code[lo];
        u32 left = shift_up_one(code, carry_code);
        u32 diff = code ^ left;
        if (c == 0) first_code = __builtin_amdgcn_readfirstlane(code);
        carry_code = __builtin_amdgcn_readlane(code, 63);
        last_code = __builtin_amdgcn_readlane(code, valid - 1);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const u32 r = (u32)k + 2u;
            if (!(mask & (1u << k))) continue;
            u64 C = ballot64((diff & plane_bits(r)) != 0u);
            if (c == 0) C |= 1ull;
            C &= vmask;
            const u64 N = ~C & vmask;
            if (r == 2) {
                u64 W = N & ((N << 1) | (prevN >> 63)) & ((N << 2) | (prevN >> 62));
                a4 += (u32)__builtin_popcountll(W);
                prevN = N;
            }
            order_update(st[k], C, N, base);
        }
    }
    }
    if (__builtin_amdgcn_ballot_w64(bad != 0)) atomicOr(err, 1u);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (!(mask & (1u << k))) continue;
        OrderState &s = st[k];
        const u32 suffix = order_finish(s, seg_len);
        if (lane == 0) {
            RunSum o;
            o.prefix = s.prefix;
            o.suffix = suffix;
            o.best = s.best;
            o.bstart = s.bstart;
            o.a4 = (k == 0) ? a4 : 0;
            out[k] = o;
        }
    }
    codes_out = first_code | (last_code << 16);
}

constexpr int kWavesPerBlock = 4;

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
    __shared__ RunSum s_run[kWavesPerBlock][8];
    __shared__ u32 s_codes[kWavesPerBlock];
    __shared__ u32 s_first[kWavesPerBlock];
    __shared__ u32 s_len[kWavesPerBlock];
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
        wave_scan<D>(d, consts_of(d, D{}), (u32)seg_start, seg_len, s_val, s_code, s_run[wave],
                     codes, err);
        if ((threadIdx.x & 63u) == 0) {
            s_codes[wave] = codes;
            s_first[wave] = (u32)seg_start;
            s_len[wave] = seg_len;
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
                    b.first_x = s_first[w];
                    b.len = s_len[w];
                    b.fcode = s_codes[w] & 0xFFFFu;
                    b.lcode = s_codes[w] >> 16;
                    b.prefix = s_run[w][k].prefix;
                    b.suffix = s_run[w][k].suffix;
                    b.best = s_run[w][k].best;
                    b.bstart = s_run[w][k].bstart + s_first[w];
                    b.a4 = s_run[w][k].a4;
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

extern "C" hipError_t rbg_launch_merge(const rbg::PrimeDesc *descs, uint32_t nprimes,
                                       const rbg::TileHdr *hdr, const rbg::RunSum *runs,
                                       rbg_record *out, uint32_t *err, hipStream_t stream) {
    hipLaunchKernelGGL(rbg::merge_primes, dim3(nprimes), dim3(rbg::kMergeThreads), 0, stream,
                       descs, hdr, runs, out, err);
    return hipGetLastError();
}

extern "C" int rbg_waves_per_block(void) { return rbg::kWavesPerBlock; }
