// Device-resident prime table: every prime below 2^32, for interval queries
// resolved on the GPU.
//
// Layout (HBM): a bitmap of the odd numbers, bit b of 64-bit word w standing
// for n = 2 (64 w + b) + 1 (1 if n is prime), 2^25 words = 256 MiB; the
// popcount of every block of 64 words (4096 odd numbers) and the exclusive
// prefix sums of those counts, 2^19 u32 each.  The prime 2 is implicit:
// rank 0.  Ranks count primes from 0 (p_0 = 2, p_1 = 3, ...).
//
// Built by a segmented sieve (one workgroup per 2^19 numbers, the segment's
// bitmap in LDS, the odd primes below 2^16 from the host), then one
// workgroup scans the block counts.  Queries: rank of x (primes below x) and
// the primes of a range of ranks (one thread per rank: binary search over
// the block prefixes, then a word scan and a bit select inside the block).
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace rbg_primes {

typedef uint32_t u32;
typedef uint64_t u64;

constexpr u32 kWords = 1u << 25;         // 64-bit bitmap words (odd n < 2^32)
constexpr u32 kBlockWords = 64;          // words per counted block
constexpr u32 kBlocks = kWords / kBlockWords;
constexpr u32 kSegWords = 4096;          // bitmap words per sieve workgroup (32 KiB of LDS)
constexpr u32 kSegments = kWords / kSegWords;
constexpr u32 kCoopLimit = 8192;         // smaller primes are marked by the whole workgroup

// One segment of the bitmap: odd numbers n = 2 i + 1 for i in
// [seg * 2^18, (seg + 1) * 2^18).  Clears the odd multiples q m >= q^2 of
// every odd prime q < 2^16 (and n = 1).
__global__ __launch_bounds__(256) void pt_sieve(const u32 *__restrict__ small, u32 nsmall,
                                                u64 *__restrict__ bits, u32 *__restrict__ counts) {
    __shared__ u32 seg[kSegWords * 2];
    const u32 tid = threadIdx.x, nt = blockDim.x;
    for (u32 i = tid; i < kSegWords * 2; i += nt) seg[i] = ~0u;
    __syncthreads();
    const u64 i_lo = (u64)blockIdx.x * kSegWords * 64;  // first odd index of the segment
    const u64 i_hi = i_lo + (u64)kSegWords * 64;
    // first index i >= i_lo with 2 i + 1 an odd multiple of q, at least q^2
    auto first_index = [&](u64 q) -> u64 {
        const u64 n_lo = 2 * i_lo + 1;
        u64 n0 = q * q;
        if (n0 < n_lo) {
            n0 = (n_lo + q - 1) / q * q;
            if (!(n0 & 1)) n0 += q;
        }
        return (n0 - 1) / 2;  // consecutive odd multiples are q apart in index space
    };
    // small primes: every thread takes every nt-th multiple
    for (u32 k = 0; k < nsmall; ++k) {
        const u32 q = small[k];
        if (q >= kCoopLimit) break;
        for (u64 i = first_index(q) + (u64)tid * q; i < i_hi; i += (u64)nt * q) {
            const u32 b = (u32)(i - i_lo);
            atomicAnd(&seg[b >> 5], ~(1u << (b & 31u)));
        }
    }
    // larger primes (at most 32 multiples per segment): one thread each
    for (u32 k = tid; k < nsmall; k += nt) {
        const u32 q = small[k];
        if (q < kCoopLimit) continue;
        for (u64 i = first_index(q); i < i_hi; i += q) {
            const u32 b = (u32)(i - i_lo);
            atomicAnd(&seg[b >> 5], ~(1u << (b & 31u)));
        }
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) seg[0] &= ~1u;  // n = 1
    __syncthreads();
    u64 *out = bits + (u64)blockIdx.x * kSegWords;
    for (u32 w = tid; w < kSegWords; w += nt) out[w] = ((u64)seg[2 * w + 1] << 32) | seg[2 * w];
    // popcounts of the segment's blocks
    for (u32 b = tid; b < kSegWords / kBlockWords; b += nt) {
        u32 c = 0;
        for (u32 w = 0; w < kBlockWords; ++w) {
            c += (u32)__builtin_popcount(seg[2 * (b * kBlockWords + w)]);
            c += (u32)__builtin_popcount(seg[2 * (b * kBlockWords + w) + 1]);
        }
        counts[blockIdx.x * (kSegWords / kBlockWords) + b] = c;
    }
}

// Exclusive prefix sums of the block counts (one workgroup of 1024 threads,
// 512 blocks each) and the total number of odd primes.
__global__ __launch_bounds__(1024) void pt_scan(const u32 *__restrict__ counts,
                                                u32 *__restrict__ prefix, u32 *__restrict__ total) {
    __shared__ u32 part[1024];
    const u32 tid = threadIdx.x;
    constexpr u32 kPer = kBlocks / 1024;
    u32 s = 0;
    for (u32 k = 0; k < kPer; ++k) s += counts[tid * kPer + k];
    part[tid] = s;
    __syncthreads();
    // inclusive Hillis-Steele scan of the 1024 partial sums
    for (u32 off = 1; off < 1024; off <<= 1) {
        const u32 v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    u32 run = part[tid] - s;  // exclusive start of this thread's chunk
    for (u32 k = 0; k < kPer; ++k) {
        prefix[tid * kPer + k] = run;
        run += counts[tid * kPer + k];
    }
    if (tid == 1023) *total = part[1023];
}

// Position (0..63) of the k-th set bit of v (k < popcount(v)).
__device__ __forceinline__ u32 select_bit(u64 v, u32 k) {
    for (u32 i = 0; i < k; ++i) v &= v - 1;
    return (u32)__builtin_ctzll(v);
}

// Primes of ranks first .. first + n - 1 (rank 0 is 2; ranks past the last
// prime below 2^32 give 0).
__global__ __launch_bounds__(256) void pt_select(const u64 *__restrict__ bits,
                                                 const u32 *__restrict__ counts,
                                                 const u32 *__restrict__ prefix, u32 nodd, u64 first,
                                                 u32 n, u32 *__restrict__ out) {
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const u64 rank = first + k;
    if (rank == 0) {
        out[k] = 2;
        return;
    }
    if (rank > nodd) {
        out[k] = 0;
        return;
    }
    const u32 r = (u32)(rank - 1);  // rank among the odd primes
    // last block whose prefix is <= r
    u32 lo = 0, hi = kBlocks;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (prefix[mid] <= r) lo = mid; else hi = mid;
    }
    u32 left = r - prefix[lo];
    const u64 *blk = bits + (u64)lo * kBlockWords;
    u32 w = 0;
    for (; w < kBlockWords - 1u; ++w) {  // the count of block lo bounds the scan
        const u32 c = (u32)__builtin_popcountll(blk[w]);
        if (left < c) break;
        left -= c;
    }
    const u64 i = ((u64)lo * kBlockWords + w) * 64 + select_bit(blk[w], left);
    out[k] = (u32)(2 * i + 1);
}

// Number of primes below x (x <= 2^32).
__global__ void pt_rank(const u64 *__restrict__ bits, const u32 *__restrict__ prefix, u64 x,
                        u64 *__restrict__ out) {
    if (threadIdx.x != 0) return;
    if (x <= 2) {
        *out = 0;
        return;
    }
    const u64 I = x / 2;  // odd numbers 2 i + 1 < x: i < x / 2
    const u64 blk = I / (kBlockWords * 64);
    u64 c = blk < kBlocks ? prefix[blk] : 0;
    if (blk >= kBlocks) {  // x = 2^32: every odd number below it
        c = prefix[kBlocks - 1];
        for (u32 w = 0; w < kBlockWords; ++w)
            c += (u32)__builtin_popcountll(bits[(u64)(kBlocks - 1) * kBlockWords + w]);
    } else {
        const u64 w0 = blk * kBlockWords, wend = I / 64;
        for (u64 w = w0; w < wend; ++w) c += (u32)__builtin_popcountll(bits[w]);
        const u32 rem = (u32)(I % 64);
        if (rem) c += (u32)__builtin_popcountll(bits[wend] & ((1ull << rem) - 1ull));
    }
    *out = c + 1;  // and the prime 2
}

}  // namespace rbg_primes

using namespace rbg_primes;

extern "C" size_t rbg_pt_bytes(void) {
    return (size_t)kWords * 8 + (size_t)kBlocks * 4 * 2 + 64;
}

// Builds the table into `bits` (kWords u64), `counts` and `prefix` (kBlocks
// u32 each) and writes the number of odd primes to *d_total (device).
extern "C" hipError_t rbg_pt_build(const uint32_t *d_small, uint32_t nsmall, uint64_t *bits,
                                   uint32_t *counts, uint32_t *prefix, uint32_t *d_total,
                                   hipStream_t stream) {
    hipLaunchKernelGGL(pt_sieve, dim3(kSegments), dim3(256), 0, stream, d_small, nsmall, bits,
                       counts);
    hipLaunchKernelGGL(pt_scan, dim3(1), dim3(1024), 0, stream, counts, prefix, d_total);
    return hipGetLastError();
}

extern "C" hipError_t rbg_pt_select(const uint64_t *bits, const uint32_t *counts,
                                    const uint32_t *prefix, uint32_t nodd, uint64_t first,
                                    uint32_t n, uint32_t *d_out, hipStream_t stream) {
    if (n) hipLaunchKernelGGL(pt_select, dim3((n + 255) / 256), dim3(256), 0, stream, bits, counts,
                              prefix, nodd, first, n, d_out);
    return hipGetLastError();
}

extern "C" hipError_t rbg_pt_rank(const uint64_t *bits, const uint32_t *prefix, uint64_t x,
                                  uint64_t *d_out, hipStream_t stream) {
    hipLaunchKernelGGL(pt_rank, dim3(1), dim3(64), 0, stream, bits, prefix, x, d_out);
    return hipGetLastError();
}
