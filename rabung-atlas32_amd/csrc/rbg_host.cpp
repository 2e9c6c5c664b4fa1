// Host orchestration and C ABI of the census library.
//
// Replaces the reference's per-prime host chain (src/census/worker32.cu:20-41;
// src/census/engine32/src/frozen_entry.cu:149-242 prepare_gpu_table /
// gpu_prepared_group).  Per-prime constants, certified covers and the sorted
// root tables are prepared on the host (character_root, frozen_entry.cu:18-24,
// 158-167; original_scan_gpu.cu:369-376) and uploaded in one copy each.  The
// batched method sends every eligible prime of the block through the sparse
// engine in one chain of launches (labels, inverses, planes, seeds, one
// persistent filter, finalize) while the host counts A4 by elliptic curves;
// primes below the sparse domain, or refused by it, go through the dense
// scan (one persistent launch plus one ordered merge).  The plain method is
// the dense scan with one launch per prime.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "rabung.h"
#include "rbg_sparse.hpp"
#include "rbg_types.hpp"

using rbg::PrimeDesc;
using rbg::RunSum;
using rbg::SparsePrime;
using rbg::SparseProf;
using rbg::TileHdr;
using rbg::kPlanePairsPerBlock;
using rbg::u16;
using rbg::u32;
using rbg::u64;

extern "C" hipError_t rbg_launch_label_seed(const PrimeDesc *descs, const rbg::ReusePiece *pieces,
                                            uint32_t npieces, uint32_t nblocks,
                                            const uint32_t *tab_val, const uint16_t *tab_lab,
                                            uint16_t *lab, uint32_t lab_from, uint32_t *err,
                                            hipStream_t stream);
extern "C" hipError_t rbg_launch_scan_reuse(const PrimeDesc *descs, const rbg::ReusePiece *pieces,
                                            uint32_t npieces, uint32_t nblocks, uint32_t chunks,
                                            const uint32_t *tab_val, const uint16_t *tab_lab,
                                            uint16_t *lab, uint32_t lab_from, TileHdr *hdr,
                                            RunSum *runs, uint32_t *err, hipStream_t stream);
extern "C" uint32_t rbg_seed_per_block(void);
extern "C" hipError_t rbg_launch_scan(const PrimeDesc *descs, uint32_t nprimes, uint32_t tile_first,
                                      uint32_t tile_end, uint32_t chunks, const void *tab_val,
                                      const uint16_t *tab_code, TileHdr *hdr, RunSum *runs,
                                      uint32_t *counter, uint32_t *err, int persistent, int wide,
                                      uint32_t grid, hipStream_t stream);
extern "C" hipError_t rbg_launch_merge(const PrimeDesc *descs, uint32_t nprimes,
                                       const TileHdr *hdr, const RunSum *runs, rbg_record *out,
                                       uint32_t *err, hipStream_t stream);
extern "C" int rbg_waves_per_block(void);
extern "C" hipError_t rbg_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t stream);
extern "C" hipError_t rbg_sparse_prepare(const SparsePrime *P, uint32_t nprimes,
                                         const SparseProf *prof, uint32_t nprof,
                                         uint32_t label_blocks, uint32_t inverse_blocks,
                                         uint32_t plane_blocks, const uint32_t *tab_val,
                                         const uint64_t *tab_val64, const uint16_t *tab_lab,
                                         uint16_t *labels, uint32_t *inv, uint64_t *inv64,
                                         uint32_t *planes, uint32_t *err, hipStream_t stream);
extern "C" uint32_t rbg_seed_segments(uint32_t bmax);
extern "C" size_t rbg_seed_agg_bytes(void);
extern "C" hipError_t rbg_sparse_seed(const SparsePrime *P, SparseProf *prof, uint32_t nprof,
                                      const uint16_t *labels, void *segs, uint32_t segs_per_prof,
                                      unsigned long long *best, uint32_t *failed,
                                      hipStream_t stream);
extern "C" hipError_t rbg_sparse_filter(const SparsePrime *P, const SparseProf *prof,
                                        uint32_t nprof, uint32_t dmax, const uint32_t *tpre,
                                        const uint32_t *planes,
                                        const void *inv, const void *tab_val,
                                        const uint16_t *tab_lab, unsigned long long *best,
                                        uint32_t *failed, unsigned long long *stats,
                                        uint32_t *queue, int wide, uint32_t grid, uint32_t group,
                                        uint32_t chunk, uint32_t extend_limit,
                                        hipStream_t stream);
extern "C" size_t rbg_pt_bytes(void);
extern "C" hipError_t rbg_pt_build(const uint32_t *d_small, uint32_t nsmall, uint64_t *bits,
                                   uint32_t *counts, uint32_t *prefix, uint32_t *d_total,
                                   hipStream_t stream);
extern "C" hipError_t rbg_pt_select(const uint64_t *bits, const uint32_t *counts,
                                    const uint32_t *prefix, uint32_t nodd, uint64_t first,
                                    uint32_t n, uint32_t *d_out, hipStream_t stream);
extern "C" hipError_t rbg_pt_rank(const uint64_t *bits, const uint32_t *prefix, uint64_t x,
                                  uint64_t *d_out, hipStream_t stream);
extern "C" size_t rbg_scratch_capacity(void);
extern "C" hipError_t rbg_scratch_setup(int device, void **address);
extern "C" hipError_t rbg_sparse_finalize(const SparsePrime *P, uint32_t nprimes,
                                          const SparseProf *prof, const unsigned long long *best,
                                          const uint32_t *a4, rbg_record *out, hipStream_t stream);

namespace {

// Deterministic Miller-Rabin for n < 2^32 with bases 2, 7, 61 (Jaeschke 1993:
// no strong pseudoprime to all three below 4,759,123,141).
bool prime_u32(u32 n) {
    if (n < 2) return false;
    for (u32 q : {2u, 3u, 5u, 7u, 11u, 13u, 61u})
        if (n % q == 0) return n == q;
    u32 d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (u32 a : {2u, 7u, 61u}) {
        u64 x = rbg::powmod64(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool ok = false;
        for (int i = 1; i < s; ++i) {
            x = x * x % n;
            if (x == n - 1) {
                ok = true;
                break;
            }
        }
        if (!ok) return false;
    }
    return true;
}

// The wider domain 2^32 < p < 2^33 (reference worker33 and engine33,
// reproducibility/gpu_method/source/worker33.cu:30;
// engine33/src/math33.hpp:9-72): exact host arithmetic through 128-bit
// products.
constexpr u64 kWideLimit = 1ull << 33;

inline u64 mulmod_w(u64 a, u64 b, u64 m) { return (u64)((unsigned __int128)a * b % m); }

u64 powmod_w(u64 a, u64 e, u64 m) {
    u64 v = 1 % m, x = a % m;
    while (e) {
        if (e & 1) v = mulmod_w(v, x, m);
        x = mulmod_w(x, x, m);
        e >>= 1;
    }
    return v;
}

// Deterministic Miller-Rabin with the first twelve prime bases, exact for
// n < 3.18e23 (Sorenson and Webster, Math. Comp. 86 (2017), Theorem 1.1; the
// reference's exact_prime64, src/census/theorie/bounded_prime.hpp:5-17, rests
// on the same theorem).
bool prime_u64(u64 n) {
    if (n < (1ull << 32)) return prime_u32((u32)n);
    static const u64 bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (u64 q : bases)
        if (n % q == 0) return false;
    u64 d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (u64 a : bases) {
        u64 x = powmod_w(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool ok = false;
        for (int i = 1; i < s; ++i) {
            x = mulmod_w(x, x, n);
            if (x == n - 1) {
                ok = true;
                break;
            }
        }
        if (!ok) return false;
    }
    return true;
}

u32 mask_of(u64 p) {
    if (p < 3) return 0;
    u32 m = 0;
    for (u32 r = 2; r <= 9; ++r)
        if ((p - 1) % r == 0) m |= 1u << (r - 2);
    return m;
}

u32 lcm_of(u32 mask) {
    u32 m = 1;
    for (u32 r = 2; r <= 9; ++r)
        if (mask & (1u << (r - 2))) m = m / std::gcd(m, r) * r;
    return m;
}

// An element of exact multiplicative order m: z = a^((p-1)/m) for the least
// a >= 2 with z^(m/q) != 1 for each prime q | m (GPU_METHOD_EN.tex:229-244).
u64 root_of_order(u64 p, u32 m) {
    const u64 e = (p - 1) / m;
    for (u64 a = 2; a < p; ++a) {
        const u64 z = powmod_w(a, e, p);
        bool ok = true;
        for (u32 q : {2u, 3u, 5u, 7u})
            if (m % q == 0 && powmod_w(z, m / q, p) == 1) {
                ok = false;
                break;
            }
        if (ok) return z;
    }
    return 0;
}

// Sorted Montgomery images z^j R mod p (R = 2^32) with their label j (and
// colour code); V = u32 below 2^32, u64 for the wider primes.
template <class V>
bool root_table(u64 p, u32 m, std::vector<V> &vals, std::vector<u16> *codes,
                std::vector<u16> *labs) {
    const u64 z = root_of_order(p, m);
    if (!z) return false;
    std::vector<std::pair<u64, u32>> tab(m);
    u64 cur = 1;
    for (u32 j = 0; j < m; ++j) {
        tab[j] = {(u64)(((unsigned __int128)cur << 32) % p), j};
        cur = mulmod_w(cur, z, p);
    }
    std::sort(tab.begin(), tab.end());
    for (auto &t : tab) {
        vals.push_back((V)t.first);
        if (codes) codes->push_back((u16)rbg::colour_code(t.second, m));
        if (labs) labs->push_back((u16)t.second);
    }
    return true;
}

// Montgomery constants of p for R = 2^32: one = R mod p, r2 = R^2 mod p and
// the form of 64; each as a low word and bit 32 (set only for p > 2^32).
struct MontConsts {
    u32 ninv, one, r2, step, hi;
};

MontConsts mont_consts(u64 p) {
    const u64 one = (1ull << 32) % p;
    const u64 r2 = mulmod_w(one, one, p), step = mulmod_w(64, one, p);
    MontConsts k;
    k.ninv = rbg::neg_inverse32((u32)p);
    k.one = (u32)one;
    k.r2 = (u32)r2;
    k.step = (u32)step;
    k.hi = (u32)(p >> 32) | (u32)(one >> 32) << 1 | (u32)(r2 >> 32) << 2 | (u32)(step >> 32) << 3;
    return k;
}

u64 isqrt64(u64 n) {
    u64 x = (u64)std::sqrt((double)n);
    while (x * x > n) --x;
    while ((x + 1) * (x + 1) <= n) ++x;
    return x;
}

// Certified clipped cover for threshold cap Q (GPU_METHOD_EN.tex:254-326;
// reference vdw_clipped::from_denominator): D ~ sqrt(p/H), least B with
// 4BD - HD^2 > 4p, all checked in integers.
// Least certified radius for a given D (GPU_METHOD_EN.tex eq. Bchoice), or 0
// when the branch condition H D^2 <= 4p fails.
u64 cover_radius(u64 p, u64 H, u64 d) {
    if (d < 1 || H * d * d > 4ull * p) return 0;
    return (4ull * p + H * d * d) / (4ull * d) + 1;
}

// Cover shape n (paper/GPU_METHOD_EN.tex:322-336): D = floor(2 sqrt(p/H) / n),
// so n = 2 is the reference's D = floor(sqrt(p/H)) and B D ~ p (1 + 1/n^2).
// Where the least radius for that D would pass the table limit, D grows
// to the least value whose radius fits (never past the n = 2 value).  The
// returned pair is certified with exact integers.
bool choose_cover(u64 p, u32 Q, u32 div, u32 &B, u32 &D) {
    const u64 H = Q - 1;
    const u64 n = div ? div : 2;
    u64 dmax = isqrt64(p / H);  // n = 2
    while (dmax > 1 && H * dmax * dmax > 4ull * p) --dmax;
    u64 d = n <= 2 ? dmax : std::max<u64>(1, isqrt64(4ull * p / (H * n * n)));
    if (d > dmax) d = dmax;
    // table radius limit: 4096 primes of at most 2^20 - 1 label entries each
    // keep a block's label offsets below 2^32
    constexpr u64 kMaxB = (1u << 20) - 2;
    if (d < dmax && cover_radius(p, H, d) > kMaxB) {
        // radius(d) ~ p/d + H d/4 falls as d grows below 2 sqrt(p/H): start
        // from the real root of H d^2 - 4 kMaxB d + 4p = 0 and step to the
        // least integer that fits
        const double bm = (double)kMaxB - 1.0;
        const double disc = bm * bm - (double)H * (double)p;
        u64 d0 = disc > 0 ? (u64)((2.0 * bm - 2.0 * std::sqrt(disc)) / (double)H) : dmax;
        d0 = std::min(std::max<u64>(d0, d), dmax);
        while (d0 > d && cover_radius(p, H, d0 - 1) && cover_radius(p, H, d0 - 1) <= kMaxB) --d0;
        while (d0 < dmax && (!cover_radius(p, H, d0) || cover_radius(p, H, d0) > kMaxB)) ++d0;
        d = d0;
    }
    if (d < 1) return false;
    const u64 b = cover_radius(p, H, d);
    if (!b || b >= p || d >= p || b > kMaxB || H * d > 2 * b) return false;
    if (4ull * b * d - H * d * d <= 4ull * p) return false;
    if (2 * b + 2 >= (u64)p - 1) return false;  // table stays inside the first half
    B = (u32)b;
    D = (u32)d;
    return true;
}

// Device memory carved from a caller-provided region instead of hipMalloc
// (the allocation-free entry point): a bump pointer over [base, base + cap).
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
};

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    Arena *arena = nullptr;  // set: carve from the arena, never allocate or free
    hipError_t ensure(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (arena) {
            const size_t at = (arena->used + 255) & ~(size_t)255;
            if (at > arena->cap || want > arena->cap - at) return hipErrorOutOfMemory;
            ptr = arena->base + at;
            bytes = want;
            arena->used = at + want;
            return hipSuccess;
        }
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        size_t cap = std::max(want, (size_t)4096);
        hipError_t e = hipMalloc(&ptr, cap);
        if (e == hipSuccess) bytes = cap;
        return e;
    }
    void release() {
        if (ptr && !arena) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

constexpr u32 kSparseMin = 1u << 16;  // smallest prime sent to the sparse engine
constexpr u32 kCoverQ = 48;           // default threshold cap of the cover
constexpr u32 kMaxCoverQ = 63;        // the screen counts up to q - t = 31 (t = ceil(q/2))
constexpr u32 kExtendLimit = 1u << 22;  // default extension steps of one passing anchor
constexpr u32 kBatchChunk = 16384;    // primes per launch chain of a batch
// filter statistics (4 x u64) and the filter's per-XCD task counters (8 x u32
// per launch: primes below 2^32, wider primes)
constexpr size_t kQueueOffset = 32, kStatsBytes = kQueueOffset + 16 * 4;

enum Ev { kAll0, kAll1, kDense0, kDense1, kPrep0, kPrep1, kSeed0, kSeed1, kFilt0, kFilt1, kNumEv };

}  // namespace

struct rbg_ctx {
    // graph mode of the current call and the cached executable graphs of the
    // dense chain [0] and the sparse chain [1]
    bool graph = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int device = 0;
    int cus = 0;
    int clock_khz = 0;
    size_t hbm_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[kNumEv] = {};
    rbg_config cfg{};
    std::string error;
    // dense path
    DevBuf descs, tab_val, tab_val64, tab_code, hdr, runs, recs, flags;
    std::vector<PrimeDesc> h_descs;
    std::vector<u32> h_val;
    std::vector<u64> h_val64;  // root tables of primes above 2^32
    std::vector<u16> h_code;
    std::vector<u16> h_lab;  // sorted root-table index -> label j (primes below 2^32)
    // per dense prime: tile chunks, and for the label-reuse scan the labels
    // of 2 and 3 (reuse = 0: every position by its power)
    struct DenseLaunch {
        u32 chunks, reuse, j2, j3;
    };
    std::vector<DenseLaunch> h_launch;
    // batched label-reuse launches of the plain method: pieces (one per
    // prime and launch) and the launches over them
    struct ReuseLaunch {
        u32 first, count, blocks, seed;  // pieces [first, first+count); seed launch or phase
    };
    std::vector<rbg::ReusePiece> h_pieces;
    std::vector<ReuseLaunch> h_reuse;
    bool exact_runs = false;  // dense scans with run floor 1 (the rescan)
    DevBuf tab_lab, dense_lab, dense_pieces;
    // sparse path
    DevBuf sp_primes, sp_profs, sp_val, sp_val64, sp_lab, sp_labels, sp_inv, sp_inv64, sp_planes,
        sp_best, sp_failed, sp_a4, sp_stats, sp_segs, sp_tpre;
    std::vector<SparsePrime> h_sp;
    std::vector<SparseProf> h_prof;
    std::vector<u32> h_sval;
    std::vector<u64> h_sval64;
    std::vector<u16> h_slab;
    std::vector<u32> h_tpre;  // filter task lists (first task of each profile per XCD)
    u32 nprof32 = 0;  // profiles of primes below 2^32 (they come first)
    // timing of the last call
    rbg_timing timing{};
    // the allocation-free entry point's scratch region while one of its calls
    // runs (null otherwise)
    Arena *arena = nullptr;
    // primes selected from the device prime table (rank queries)
    DevBuf sel;
    // this device's copy of the library's static scratch array (device
    // address, resolved when the context is opened) or the reason it is not
    // available
    void *scratch = nullptr;
    std::string scratch_error;

    // every device buffer of the context
    std::vector<DevBuf *> buffers() {
        return {&descs, &tab_val, &tab_val64, &tab_code, &tab_lab, &dense_lab, &hdr, &runs,
                &recs, &flags, &dense_pieces, &sp_primes, &sp_profs, &sp_val, &sp_val64, &sp_lab,
                &sp_labels, &sp_inv,
                &sp_inv64, &sp_planes, &sp_best, &sp_failed, &sp_a4, &sp_stats, &sp_segs, &sp_tpre,
                &sel};
    }

    int fail(int code, const std::string &msg) {
        error = msg;
        return code;
    }
};

namespace {

int hip_fail(rbg_ctx *c, hipError_t e, const char *what) {
    if (c->arena && e == hipErrorOutOfMemory) {
        char msg[160];
        snprintf(msg, sizeof msg,
                 "the call needs more device workspace than the %zu-byte scratch of "
                 "rbg_profile_block_noalloc",
                 c->arena->cap);
        return c->fail(RBG_ECAPACITY, msg);
    }
    return c->fail(RBG_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCK(ctx, expr)                                          \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return hip_fail((ctx), _e, #expr); \
    } while (0)

u32 tile_chunks(const rbg_ctx *c) { return c->cfg.tile_chunks ? c->cfg.tile_chunks : 64u; }

// Enqueues a launch chain on the context's stream.  In graph mode the chain
// is recorded by stream capture and launched as one graph; the cached
// executable graph of the slot is updated in place when only launch
// parameters changed and instantiated again otherwise.  The chain must not
// contain memsets or pageable copies (fills are kernels).
template <class F>
int launch_chain(rbg_ctx *c, int slot, F &&enqueue) {
    if (!c->graph) return enqueue();
    hipStream_t s = c->stream;
    HIPCK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &g);
    struct GraphGuard {
        hipGraph_t g;
        ~GraphGuard() {
            if (g) (void)hipGraphDestroy(g);
        }
    } guard{g};
    if (rc) return rc;
    HIPCK(c, end);
    hipGraphExec_t &x = c->gexec[slot];
    if (x) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(x, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
    }
    if (!x) HIPCK(c, hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    HIPCK(c, hipGraphLaunch(x, s));
    ++c->timing.launches;
    return RBG_OK;
}

u32 blocks_per_cu(const rbg_ctx *c) { return c->cfg.blocks_per_cu ? c->cfg.blocks_per_cu : 8u; }

u32 xcd_group(const rbg_ctx *c) { return c->cfg.xcd_group ? c->cfg.xcd_group : 1u; }

u64 budget(const rbg_ctx *c) { return c->cfg.hbm_budget ? c->cfg.hbm_budget : (16ull << 30); }

u32 cover_q(const rbg_ctx *c) { return c->cfg.cover_q ? c->cfg.cover_q : kCoverQ; }

u32 cover_div(const rbg_ctx *c) { return c->cfg.cover_div ? c->cfg.cover_div : 3u; }

// 0: chosen per launch (sparse_run)
u32 filter_chunk(const rbg_ctx *c) { return c->cfg.filter_chunk; }

// Speculative threshold of order r for p: the largest L whose expected
// number of runs of length >= L over the half interval of a random
// r-colouring is at least 14, (p-1)/2 (1 - 1/r) r^-(L-1) >= 14 (so a
// longest run below L has probability about e^-14; 0 of the 35,676
// sparse-domain profiles of the committed vectors falls below it).  0 when
// no L >= 2 qualifies.
u32 spec_threshold(u64 p, u32 r) {
    double v = (double)((p - 1) / 2) * (1.0 - 1.0 / r) / 14.0;
    u32 L = 1;
    for (; v >= r; v /= r) ++L;
    return L >= 2 ? L : 0u;
}

u32 extend_limit(const rbg_ctx *c) {
    return c->cfg.extend_limit ? c->cfg.extend_limit : kExtendLimit;
}

double elapsed(rbg_ctx *c, int a, int b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0;
    return ms;
}

// Host verification of one profile by direct powers (reference
// certify_witness, VBB/src/advance.cpp:7-13, and the prefix loop of
// src/census/worker32.cu:37-38).
// POW(x) = x^e mod p: 64-bit products below 2^32, 128-bit above.
template <class POW>
bool certify_with(u64 p, u32 M, u32 L, u32 W, POW pw) {
    if (M == 0 || W == 0 || (u64)W + M > p || L == 0) return false;
    const u64 colour = pw(W);
    for (u32 j = 1; j < M; ++j)
        if (pw((u64)W + j) != colour) return false;
    if (W > 1 && pw(W - 1) == colour) return false;
    if ((u64)W + M < p && pw((u64)W + M) == colour) return false;
    for (u32 x = 1; x <= L; ++x)
        if (pw(x) != 1) return false;
    if ((u64)L + 1 < p && pw((u64)L + 1) == 1) return false;
    return true;
}

bool certify(u64 p, u32 r, u32 M, u32 L, u32 W) {
    const u64 e = (p - 1) / r;
    if (p < (1ull << 32))
        return certify_with(p, M, L, W, [&](u64 x) { return (u64)rbg::powmod64((u32)x, e, (u32)p); });
    return certify_with(p, M, L, W, [&](u64 x) { return powmod_w(x, e, p); });
}

// Index of the first record that fails its direct verification (the
// admissibility mask, K from M and L, and for every admissible order the
// witness run, its neighbours and the prefix by direct powers), or -1.
// primes[i] is the full prime of record i (records keep its low word).
long long first_uncertified(const rbg_record *recs, const u64 *primes, u32 n, unsigned threads) {
    const unsigned nt = std::max(1u, std::min(threads, n));
    std::vector<long long> bad(nt, -1);
    auto check = [&](u32 i) {
        const rbg_record &rec = recs[i];
        const u64 p = primes[i];
        if (rec.mask != mask_of(p)) return false;
        for (u32 r = 2; r <= 9; ++r) {
            const u32 k = r - 2;
            if (!(rec.mask & (1u << k))) continue;
            const u64 bound = ((p - 1) / 2) % r == 0 ? 2ull * rec.L[k] + 2 : (u64)rec.L[k] + 2;
            if (rec.K[k] != std::max<u64>((u64)rec.M[k] + 1, bound)) return false;
            if (rec.L[k] > rec.M[k] || rec.M[k] > (p - 1) / r || rec.W[k] > (p - 1) / 2)
                return false;
            if (!certify(p, r, rec.M[k], rec.L[k], rec.W[k])) return false;
        }
        return true;
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (u32 i = t; i < n; i += nt)
                if (!check(i)) {
                    bad[t] = i;
                    return;
                }
        });
    for (auto &t : th) t.join();
    long long first = -1;
    for (long long b : bad)
        if (b >= 0 && (first < 0 || b < first)) first = b;
    return first;
}

unsigned host_threads() { return std::max(1u, std::thread::hardware_concurrency()); }

int certify_all(rbg_ctx *c, const rbg_record *out, const u64 *primes, u32 n) {
    const long long b = first_uncertified(out, primes, n, std::min(16u, host_threads()));
    if (b >= 0) {
        char msg[128];
        snprintf(msg, sizeof msg, "direct verification failed for p=%llu",
                 (unsigned long long)primes[b]);
        return c->fail(RBG_ECERTIFY, msg);
    }
    return RBG_OK;
}

// ---- dense scan ----------------------------------------------------------

// Primes below 2^32 first, then the wider ones (each domain is one
// contiguous range of descriptors, tiles and launches).
std::vector<u32> by_domain(const u64 *primes, const std::vector<u32> &idx) {
    std::vector<u32> out;
    for (u32 i : idx)
        if (primes[i] < (1ull << 32)) out.push_back(i);
    for (u32 i : idx)
        if (primes[i] >= (1ull << 32)) out.push_back(i);
    return out;
}

// Run floor of order r for a half interval of nh positions: 1 + the
// largest t with r^t <= nh / 4096, so that runs of that length are expected
// about 4096 (1 - 1/r) times; segments then skip the bookkeeping of shorter
// runs (most chunks hold none that long).  A prime whose longest run stays
// below its floor is rescanned with floor 1, so results never depend on it.
u32 run_floor(u64 nh, u32 r) {
    u32 t = 1;
    for (u64 v = nh / 4096; v >= r && t < 62; v /= r) ++t;
    return t;
}

// Label-reuse scan (plain method, primes below 2^32): tiles of the first
// phase, whose positions read the labels seeded by direct powers.  Batched
// phases fill the chip from small windows on, so the seed can be short
// (block 2^24: 0.287 s with 256 first-phase tiles, 0.264 / 0.253 / 0.247 s
// with 128 / 64 / 32; the top primes do not change).
constexpr u32 kReuseFirstPhase = 32;
constexpr u32 kReuseChunkUnit = 12;  // segments start at 1 mod 6, passes of 12 chunks

// Label j of v (x^e = z^j) from a prime's sorted root table.
u32 label_of(u64 p, u32 e, u32 v, const u32 *vals, const u16 *labs, u32 m) {
    const u32 y = (u32)(((unsigned __int128)rbg::powmod64((u32)v, e, (u32)p) << 32) % p);
    const u32 *at = std::lower_bound(vals, vals + m, y);
    return (at != vals + m && *at == y) ? labs[at - vals] : 0xFFFFFFFFu;
}

int dense_run(rbg_ctx *c, const u64 *primes, const std::vector<u32> &idx_in, bool plain,
              rbg_record *recs, u32 *err) {
    const u32 waves = (u32)rbg_waves_per_block();
    const u32 reuse_chunks = (tile_chunks(c) + kReuseChunkUnit - 1) / kReuseChunkUnit *
                             kReuseChunkUnit;
    const std::vector<u32> idx = by_domain(primes, idx_in);
    c->h_descs.clear();
    c->h_val.clear();
    c->h_val64.clear();
    c->h_code.clear();
    c->h_lab.clear();
    c->h_launch.clear();
    u64 tiles = 0, runs = 0, tiles32 = 0, lab_bytes = 0;
    for (u32 i : idx) {
        const u64 p = primes[i];
        const bool wide = p >= (1ull << 32);
        // the plain method scans the large primes below 2^32 with label reuse
        const u64 nh = (p - 1) / 2;
        const bool reuse = plain && !wide &&
                           nh > (u64)kReuseFirstPhase * waves * reuse_chunks * 64u;
        const u32 chunks = reuse ? reuse_chunks : tile_chunks(c);
        const u32 tile_pos = waves * chunks * 64u;
        const MontConsts k = mont_consts(p);
        PrimeDesc d{};
        d.p = (u32)p;
        d.nh = (u32)((p - 1) / 2);
        d.mask = mask_of(p);
        d.m = lcm_of(d.mask);
        d.e = (u32)((p - 1) / d.m);
        d.ninv = k.ninv;
        d.one = k.one;
        d.r2 = k.r2;
        d.step = k.step;
        d.hi = k.hi;
        d.tab_off = (u32)c->h_code.size();
        d.tab64 = (u32)c->h_val64.size();
        d.planes = 0;
        for (u32 r = 2; r <= 9; ++r)
            if (d.mask & (1u << (r - 2))) d.planes |= rbg::plane_bits(r);
        d.ntiles = (u32)(((u64)d.nh + tile_pos - 1) / tile_pos);  // nh reaches 2^32 - 5 below 2^33
        d.tile_begin = (u32)tiles;
        d.nr = (u32)__builtin_popcount(d.mask);
        d.sum_off = (u32)runs;
        d.index = i;
        d.floors[0] = d.floors[1] = 0x01010101u;
        if (!c->exact_runs)
            for (u32 r = 2; r <= 9; ++r)
                if (d.mask & (1u << (r - 2))) {
                    const u32 f = std::min(62u, run_floor(d.nh, r) + c->cfg.floor_boost);
                    d.floors[(r - 2) >> 2] += (f - 1u) << (8u * ((r - 2) & 3u));
                }
        if (tiles + d.ntiles > 0xFFFFFFFFull || runs + (u64)d.ntiles * d.nr > 0xFFFFFFFFull)
            return c->fail(RBG_ECAPACITY, "block too large for one call");
        tiles += d.ntiles;
        if (!wide) tiles32 = tiles;
        runs += (u64)d.ntiles * d.nr;
        c->timing.positions += d.nh;
        const bool ok = wide ? root_table(p, d.m, c->h_val64, &c->h_code, nullptr)
                             : root_table(p, d.m, c->h_val, &c->h_code, &c->h_lab);
        if (!ok) return c->fail(RBG_EINVAL, "no root of the required order");
        if (wide) c->h_lab.resize(c->h_code.size(), 0);
        rbg_ctx::DenseLaunch L{chunks, 0, 0, 0};
        if (reuse) {
            const u32 *vals = c->h_val.data() + d.tab_off;
            const u16 *labs = c->h_lab.data() + d.tab_off;
            L.j2 = label_of(p, d.e, 2, vals, labs, d.m);
            L.j3 = label_of(p, d.e, 3, vals, labs, d.m);
            if (L.j2 >= d.m || L.j3 >= d.m) return c->fail(RBG_EINVAL, "label of 2 or 3");
            L.reuse = 1;
            lab_bytes = std::max<u64>(lab_bytes, (u64)(d.nh / 2 + 1) * 2);
            // powers: the seed, then the positions coprime to 6
            const u64 seed = (u64)kReuseFirstPhase * tile_pos / 2;
            c->timing.products += (seed + (d.nh + 5) / 6 + (d.nh + 1) / 6) *
                                  rbg::montpow_products(d.e);
        } else {
            c->timing.products += (u64)d.nh * rbg::montpow_products(d.e);
        }
        c->h_launch.push_back(L);
        c->h_descs.push_back(d);
    }
    const u32 np = (u32)c->h_descs.size();
    c->timing.dense_primes = np;
    if (!np) return RBG_OK;
    u64 bytes = np * sizeof(PrimeDesc) + c->h_val.size() * 4 + c->h_val64.size() * 8 +
                c->h_code.size() * 4 + tiles * sizeof(TileHdr) + runs * sizeof(RunSum);
    if (bytes > budget(c)) return c->fail(RBG_ECAPACITY, "dense workspace exceeds the HBM budget");
    // no room for labels (the HBM budget, or the static scratch of the
    // allocation-free entry point, far below a label array's 2 GiB):
    // every position by its power
    if (bytes + lab_bytes > budget(c) || c->arena) {
        for (rbg_ctx::DenseLaunch &L : c->h_launch) L.reuse = 0;
        lab_bytes = 0;
    }
    // the label-reuse primes run in groups whose label arrays fit side by
    // side in the budget the other workspaces leave (and below 2^32
    // entries): per group one seed launch and one launch per phase over all
    // its primes, so that a phase fills the chip however few tiles each
    // prime has in it.  Phase windows of tiles [0, f), [f, 2f), [2f, 4f),
    // ... (f = kReuseFirstPhase): a position x of a phase reads x/2 and x/3,
    // which lie in the seed or an earlier phase of its own prime.
    c->h_pieces.clear();
    c->h_reuse.clear();
    const u32 reuse_pos = waves * reuse_chunks * 64u;  // positions per tile
    const u32 lab_from = kReuseFirstPhase * reuse_pos / 2u;
    u64 lab_entries_max = 0;
    if (lab_bytes) {
        const u64 cap = std::min<u64>((budget(c) - bytes) / 2, 0xFFFFFFFFull);
        const u32 seed_blocks = (lab_from + rbg_seed_per_block() - 1) / rbg_seed_per_block();
        std::vector<u32> group;
        std::vector<u64> offs;
        u64 entries = 0;
        auto flush = [&]() {
            if (group.empty()) return;
            rbg_ctx::ReuseLaunch S{(u32)c->h_pieces.size(), 0, 0, 1};
            for (size_t g = 0; g < group.size(); ++g) {
                const rbg_ctx::DenseLaunch &L = c->h_launch[group[g]];
                c->h_pieces.push_back({S.blocks, group[g], 0, L.j2, L.j3, 0, offs[g]});
                S.blocks += seed_blocks;
                ++S.count;
            }
            c->h_reuse.push_back(S);
            for (u32 lo = 0, hi = kReuseFirstPhase;; lo = hi, hi *= 2u) {
                rbg_ctx::ReuseLaunch P{(u32)c->h_pieces.size(), 0, 0, 0};
                for (size_t g = 0; g < group.size(); ++g) {
                    const u32 nt = c->h_descs[group[g]].ntiles;
                    if (nt <= lo) continue;
                    const rbg_ctx::DenseLaunch &L = c->h_launch[group[g]];
                    c->h_pieces.push_back({P.blocks, group[g], lo, L.j2, L.j3, 0, offs[g]});
                    P.blocks += std::min(hi, nt) - lo;
                    ++P.count;
                }
                if (!P.count) break;
                c->h_reuse.push_back(P);
            }
            lab_entries_max = std::max(lab_entries_max, entries);
            group.clear();
            offs.clear();
            entries = 0;
        };
        for (u32 i = 0; i < np; ++i) {
            if (!c->h_launch[i].reuse) continue;
            const u64 e = (u64)c->h_descs[i].nh / 2 + 1;
            if (!group.empty() && entries + e > cap) flush();
            group.push_back(i);
            offs.push_back(entries);
            entries += e;
        }
        flush();
        lab_bytes = lab_entries_max * 2;
    }
    HIPCK(c, c->descs.ensure(np * sizeof(PrimeDesc)));
    HIPCK(c, c->tab_val.ensure(c->h_val.size() * 4));
    HIPCK(c, c->tab_val64.ensure(c->h_val64.size() * 8));
    HIPCK(c, c->tab_code.ensure(c->h_code.size() * 2));
    HIPCK(c, c->tab_lab.ensure(c->h_lab.size() * 2));
    if (lab_bytes) {
        HIPCK(c, c->dense_lab.ensure(lab_bytes));
        HIPCK(c, c->dense_pieces.ensure(c->h_pieces.size() * sizeof(rbg::ReusePiece)));
    }
    HIPCK(c, c->hdr.ensure(tiles * sizeof(TileHdr)));
    HIPCK(c, c->runs.ensure(runs * sizeof(RunSum)));
    hipStream_t s = c->stream;
    HIPCK(c, hipMemcpyAsync(c->descs.ptr, c->h_descs.data(), np * sizeof(PrimeDesc),
                            hipMemcpyHostToDevice, s));
    if (!c->h_val.empty())
        HIPCK(c, hipMemcpyAsync(c->tab_val.ptr, c->h_val.data(), c->h_val.size() * 4,
                                hipMemcpyHostToDevice, s));
    if (!c->h_val64.empty())
        HIPCK(c, hipMemcpyAsync(c->tab_val64.ptr, c->h_val64.data(), c->h_val64.size() * 8,
                                hipMemcpyHostToDevice, s));
    HIPCK(c, hipMemcpyAsync(c->tab_code.ptr, c->h_code.data(), c->h_code.size() * 2,
                            hipMemcpyHostToDevice, s));
    HIPCK(c, hipMemcpyAsync(c->tab_lab.ptr, c->h_lab.data(), c->h_lab.size() * 2,
                            hipMemcpyHostToDevice, s));
    if (!c->h_pieces.empty())
        HIPCK(c, hipMemcpyAsync(c->dense_pieces.ptr, c->h_pieces.data(),
                                c->h_pieces.size() * sizeof(rbg::ReusePiece),
                                hipMemcpyHostToDevice, s));
    // tile counters of the two persistent launches (flags[0], flags[2])
    u32 *counter = (u32 *)c->flags.ptr, *counter64 = (u32 *)c->flags.ptr + 2;
    const PrimeDesc *dd = (const PrimeDesc *)c->descs.ptr;
    HIPCK(c, hipEventRecord(c->ev[kDense0], s));
    // one chain: clear the tile counter, scan (one launch per prime for the
    // plain method, one persistent launch otherwise), merge per prime
    const int rc = launch_chain(c, 0, [&]() -> int {
        HIPCK(c, rbg_fill_u32(counter, 1, 0u, s));
        HIPCK(c, rbg_fill_u32(counter64, 1, 0u, s));
        if (plain) {
            const rbg::ReusePiece *pieces = (const rbg::ReusePiece *)c->dense_pieces.ptr;
            for (const rbg_ctx::ReuseLaunch &L : c->h_reuse) {
                if (L.seed)
                    HIPCK(c, rbg_launch_label_seed(dd, pieces + L.first, L.count, L.blocks,
                                                   (const u32 *)c->tab_val.ptr,
                                                   (const u16 *)c->tab_lab.ptr,
                                                   (u16 *)c->dense_lab.ptr, lab_from, err, s));
                else
                    HIPCK(c, rbg_launch_scan_reuse(dd, pieces + L.first, L.count, L.blocks,
                                                   reuse_chunks, (const u32 *)c->tab_val.ptr,
                                                   (const u16 *)c->tab_lab.ptr,
                                                   (u16 *)c->dense_lab.ptr, lab_from,
                                                   (TileHdr *)c->hdr.ptr, (RunSum *)c->runs.ptr,
                                                   err, s));
                ++c->timing.launches;
            }
            for (u32 i = 0; i < np; ++i) {
                const PrimeDesc &d = c->h_descs[i];
                const rbg_ctx::DenseLaunch &L = c->h_launch[i];
                const int wide = (int)(d.hi & 1u);
                if (L.reuse) continue;  // scanned by the batched phases above
                HIPCK(c, rbg_launch_scan(dd, np, d.tile_begin, d.tile_begin + d.ntiles, L.chunks,
                                         wide ? c->tab_val64.ptr : c->tab_val.ptr,
                                         (const u16 *)c->tab_code.ptr, (TileHdr *)c->hdr.ptr,
                                         (RunSum *)c->runs.ptr, counter, err, 0, wide, d.ntiles,
                                         s));
                ++c->timing.launches;
            }
        } else {
            // one persistent launch per domain over its contiguous tile range
            const u64 ranges[2][2] = {{0, tiles32}, {tiles32, tiles}};
            for (int wide = 0; wide < 2; ++wide) {
                const u64 t0 = ranges[wide][0], t1 = ranges[wide][1];
                if (t1 == t0) continue;
                const u32 grid = (u32)std::min<u64>((u64)c->cus * blocks_per_cu(c), t1 - t0);
                HIPCK(c, rbg_launch_scan(dd, np, (u32)t0, (u32)t1, tile_chunks(c),
                                         wide ? c->tab_val64.ptr : c->tab_val.ptr,
                                         (const u16 *)c->tab_code.ptr, (TileHdr *)c->hdr.ptr,
                                         (RunSum *)c->runs.ptr, wide ? counter64 : counter, err, 1,
                                         wide, std::max(1u, grid), s));
                ++c->timing.launches;
            }
        }
        // the scan time excludes the merge when the chain runs eagerly
        if (!c->graph) HIPCK(c, hipEventRecord(c->ev[kDense1], s));
        HIPCK(c, rbg_launch_merge(dd, np, (const TileHdr *)c->hdr.ptr, (const RunSum *)c->runs.ptr,
                                  recs, err, s));
        return RBG_OK;
    });
    if (rc) return rc;
    if (c->graph) HIPCK(c, hipEventRecord(c->ev[kDense1], s));
    return RBG_OK;
}

// ---- sparse engine -------------------------------------------------------

// Plans the sparse primes; indices that the sparse engine cannot take are
// appended to `rest`.  Per prime the certified cover, the Montgomery
// constants and the sorted root table are built on up to 16 host threads
// (contiguous ranges of primes), then one ordered pass assigns every offset
// and lays out the profiles.
int sparse_plan(rbg_ctx *c, const u64 *primes, const std::vector<u32> &idx_in,
                std::vector<u32> &rest) {
    const std::vector<u32> idx = by_domain(primes, idx_in);
    c->h_sp.clear();
    c->h_prof.clear();
    c->h_sval.clear();
    c->h_sval64.clear();
    c->h_slab.clear();
    c->nprof32 = 0;
    const u32 n = (u32)idx.size();
    struct Item {
        SparsePrime s;
        u32 part, val_at, lab_at;  // root table: thread part and its offsets
        bool sparse, ok;
    };
    struct Part {
        std::vector<u32> val;
        std::vector<u64> val64;
        std::vector<u16> lab;
    };
    std::vector<Item> items(n);
    // a prime's plan takes a few microseconds (its root of order m), while
    // starting and joining a thread costs tens: threads only for large blocks
    const u32 nt = std::max(1u, std::min({16u, std::thread::hardware_concurrency(), n / 256u}));
    std::vector<Part> parts(nt);
    auto build = [&](u32 t) {
        Part &pt = parts[t];
        for (u32 k = (u32)((u64)n * t / nt); k < (u32)((u64)n * (t + 1) / nt); ++k) {
            Item &it = items[k];
            const u64 p = primes[idx[k]];
            const bool wide = p >= (1ull << 32);
            u32 B = 0, D = 0;
            it.sparse = p >= kSparseMin && choose_cover(p, cover_q(c), cover_div(c), B, D);
            it.ok = true;
            if (!it.sparse) continue;
            const MontConsts mc = mont_consts(p);
            SparsePrime &sp = it.s;
            sp = SparsePrime{};
            sp.p = (u32)p;
            sp.mask = mask_of(p);
            sp.m = lcm_of(sp.mask);
            sp.e = (u32)((p - 1) / sp.m);
            sp.ninv = mc.ninv;
            sp.one = mc.one;
            sp.r2 = mc.r2;
            sp.hi = mc.hi & 7u;
            sp.index = idx[k];
            sp.B = B;
            sp.D = D;
            sp.Q = cover_q(c);
            sp.jneg = (sp.e & 1u) ? sp.m / 2 : 0;  // (-1)^e = z^(m/2) when e is odd
            it.part = t;
            it.val_at = (u32)(wide ? pt.val64.size() : pt.val.size());
            it.lab_at = (u32)pt.lab.size();
            it.ok = wide ? root_table(p, sp.m, pt.val64, nullptr, &pt.lab)
                         : root_table(p, sp.m, pt.val, nullptr, &pt.lab);
        }
    };
    if (nt == 1) {
        build(0);
    } else {
        std::vector<std::thread> th;
        for (u32 t = 0; t < nt; ++t) th.emplace_back(build, t);
        for (auto &t : th) t.join();
    }
    u64 lab = 0, inv = 0, inv64 = 0, words = 0, lblk = 0, iblk = 0, pblk = 0;
    for (u32 k = 0; k < n; ++k) {
        const Item &it = items[k];
        if (!it.sparse) {
            rest.push_back(idx[k]);
            continue;
        }
        if (!it.ok) return c->fail(RBG_EINVAL, "no root of the required order");
        const u64 p = primes[idx[k]];
        const bool wide = p >= (1ull << 32);
        SparsePrime s = it.s;
        const u32 B = s.B, D = s.D;
        s.tab_off = (u32)c->h_slab.size();
        s.tab64 = (u32)c->h_sval64.size();
        s.lab_off = (u32)lab;
        s.inv_off = (u32)(wide ? inv64 : inv);
        s.prof_begin = (u32)c->h_prof.size();
        s.lblk = (u32)lblk;
        s.iblk = (u32)iblk;
        s.pblk = (u32)pblk;
        lab += B + 1ull;
        (wide ? inv64 : inv) += D + 1ull;
        lblk += (B + 2047ull) / 2048ull;
        iblk += (D + 1ull + 255ull) / 256ull;
        const Part &pt = parts[it.part];
        c->h_slab.insert(c->h_slab.end(), pt.lab.begin() + it.lab_at,
                         pt.lab.begin() + it.lab_at + s.m);
        if (wide)
            c->h_sval64.insert(c->h_sval64.end(), pt.val64.begin() + it.val_at,
                               pt.val64.begin() + it.val_at + s.m);
        else
            c->h_sval.insert(c->h_sval.end(), pt.val.begin() + it.val_at,
                             pt.val.begin() + it.val_at + s.m);
        for (u32 r = 2; r <= 9; ++r)
            if (s.mask & (1u << (r - 2))) {
                SparseProf f{};
                f.prime = (u32)c->h_sp.size();
                f.r = r;
                f.np = 32u - (u32)__builtin_clz(r - 1);
                f.wpp = rbg::row_words(1, B);
                f.plane_off = (u32)words;  // units of 16 words
                f.sigma0 = (((p - 1) / 2) % r) == 0;
                f.pblk = (u32)pblk;
                f.B = s.B;
                f.D = s.D;
                f.Q = s.Q;
                f.inv_off = s.inv_off;
                if (!c->cfg.spec_off) {
                    const u32 q0 = spec_threshold(p, r);
                    f.qspec = q0 ? std::min(s.Q, q0 + c->cfg.spec_boost) : 0u;
                }
                words += rbg::row_offset(rbg::kRows, B, f.np) / 16u;
                for (u32 row = 0; row < rbg::kRows; ++row) pblk += rbg::row_blocks(row, B);
                c->h_prof.push_back(f);
                ++s.nprof;
                if (!wide) ++c->nprof32;
            }
        c->h_sp.push_back(s);
        if (lab > 0xFFFFFFFFull || words > 0xFFFFFFFFull || pblk > 0x7FFFFFFFull)
            return c->fail(RBG_ECAPACITY, "sparse block too large for one call");
    }
    c->timing.sparse_primes = c->h_sp.size();
    c->timing.sparse_profiles = c->h_prof.size();
    return RBG_OK;
}

int sparse_run(rbg_ctx *c, rbg_record *recs, u32 *err, std::vector<u32> &failed_idx) {
    const u32 np = (u32)c->h_sp.size(), nf = (u32)c->h_prof.size();
    if (!np) return RBG_OK;
    u64 lab = 0, inv = 0, inv64 = 0, words = 0, lblk = 0, iblk = 0, pblk = 0, bmax = 0;
    u64 dmax[2] = {0, 0};  // largest denominator per domain
    for (const SparsePrime &s : c->h_sp) {
        const bool wide = s.hi & 1u;
        lab += s.B + 1ull;
        (wide ? inv64 : inv) += s.D + 1ull;
        lblk += (s.B + 2047ull) / 2048ull;
        iblk += (s.D + 1ull + 255ull) / 256ull;
        dmax[wide] = std::max<u64>(dmax[wide], s.D);
        bmax = std::max<u64>(bmax, s.B);
    }
    const u32 segs_per_prof = rbg_seed_segments((u32)bmax);
    for (const SparseProf &f : c->h_prof) {
        const u32 B = c->h_sp[f.prime].B;
        words += rbg::row_offset(rbg::kRows, B, f.np);  // 32-bit words
        for (u32 row = 0; row < rbg::kRows; ++row) pblk += rbg::row_blocks(row, B);
    }
    const u64 bytes = np * sizeof(SparsePrime) + nf * sizeof(SparseProf) + c->h_sval.size() * 4 +
                      c->h_sval64.size() * 8 + c->h_slab.size() * 2 + lab * 2 + inv * 4 +
                      inv64 * 8 + words * 4 + nf * 8 + np * 8;
    if (bytes > budget(c)) return c->fail(RBG_ECAPACITY, "sparse workspace exceeds the HBM budget");
    HIPCK(c, c->sp_primes.ensure(np * sizeof(SparsePrime)));
    HIPCK(c, c->sp_profs.ensure(nf * sizeof(SparseProf)));
    HIPCK(c, c->sp_val.ensure(c->h_sval.size() * 4));
    HIPCK(c, c->sp_val64.ensure(c->h_sval64.size() * 8));
    HIPCK(c, c->sp_lab.ensure(c->h_slab.size() * 2));
    HIPCK(c, c->sp_labels.ensure(lab * 2));
    HIPCK(c, c->sp_inv.ensure(inv * 4));
    HIPCK(c, c->sp_inv64.ensure(inv64 * 8));
    HIPCK(c, c->sp_planes.ensure(words * 4));
    HIPCK(c, c->sp_best.ensure(nf * 8));
    HIPCK(c, c->sp_failed.ensure(np * 4));
    HIPCK(c, c->sp_a4.ensure(np * 4));
    HIPCK(c, c->sp_stats.ensure(kStatsBytes));
    HIPCK(c, c->sp_segs.ensure((u64)nf * segs_per_prof * rbg_seed_agg_bytes()));
    hipStream_t s = c->stream;
    // A4 by elliptic point counts on the host while the device works
    std::vector<u64> sp_p(np);
    std::vector<u32> a4(np, 0);
    for (u32 i = 0; i < np; ++i) sp_p[i] = c->h_sp[i].p | ((u64)(c->h_sp[i].hi & 1u) << 32);
    int a4_rc = RBG_OK;
    double a4_ms = 0.0;
    std::thread a4_thread([&] {
        const auto t0 = std::chrono::steady_clock::now();
        a4_rc = rbg_a4_host_wide(sp_p.data(), np, a4.data(), 0);
        a4_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
    });
    struct Joiner {
        std::thread &t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } joiner{a4_thread};
    HIPCK(c, hipMemcpyAsync(c->sp_primes.ptr, c->h_sp.data(), np * sizeof(SparsePrime),
                            hipMemcpyHostToDevice, s));
    HIPCK(c, hipMemcpyAsync(c->sp_profs.ptr, c->h_prof.data(), nf * sizeof(SparseProf),
                            hipMemcpyHostToDevice, s));
    if (!c->h_sval.empty())
        HIPCK(c, hipMemcpyAsync(c->sp_val.ptr, c->h_sval.data(), c->h_sval.size() * 4,
                                hipMemcpyHostToDevice, s));
    if (!c->h_sval64.empty())
        HIPCK(c, hipMemcpyAsync(c->sp_val64.ptr, c->h_sval64.data(), c->h_sval64.size() * 8,
                                hipMemcpyHostToDevice, s));
    HIPCK(c, hipMemcpyAsync(c->sp_lab.ptr, c->h_slab.data(), c->h_slab.size() * 2,
                            hipMemcpyHostToDevice, s));
    const SparsePrime *dP = (const SparsePrime *)c->sp_primes.ptr;
    SparseProf *dF = (SparseProf *)c->sp_profs.ptr;
    unsigned long long *best = (unsigned long long *)c->sp_best.ptr;
    u32 *failed = (u32 *)c->sp_failed.ptr;
    // the filter's XCD-aware schedule needs a grid that is a multiple of 8;
    // the 33-bit filter is compiled for 5 resident blocks per CU (its
    // register budget), so its grid stops there (blocks past the resident
    // ones would only start once the queues are empty)
    const u32 grid = std::max(8u, ((u32)c->cus * blocks_per_cu(c) + 7u) / 8u * 8u);
    const u32 grid_wide =
        std::max(8u, ((u32)c->cus * std::min(blocks_per_cu(c), 5u) + 7u) / 8u * 8u);
    // task lists of the XCDs, per filter launch (domain): XCD x owns the
    // launch's profiles x, x+8, ...; its list holds their denominators 1..D
    // back to back, and tpre holds, per XCD, the first task of each of its
    // profiles followed by the list's length (nf + 16 entries in all)
    const u32 nfd[2] = {c->nprof32, nf - c->nprof32};
    std::vector<u32> &tpre = c->h_tpre;
    tpre.clear();
    tpre.reserve(nf + 16);
    u64 most = 0, tasks = 0;
    for (int wide = 0; wide < 2; ++wide) {
        const u32 f0 = wide ? c->nprof32 : 0u;
        for (u32 x = 0; x < 8; ++x) {
            u64 at = 0;
            for (u32 k = x; k < nfd[wide]; k += 8) {
                tpre.push_back((u32)at);
                at += c->h_sp[c->h_prof[f0 + k].prime].D;
            }
            tpre.push_back((u32)at);
            most = std::max(most, at);
            tasks += at;
        }
    }
    // queue claims: the configured size, else 16 tasks for launches with more
    // than 10^4 tasks per wave (the end of the launch is then a small share
    // and large claims amortise the atomic: batch64 3.59 s of filter against
    // 3.73 with claims sized per profile) and sized by each profile's cover
    // radius in the kernel below that (top-256 40.7 -> 39.6 ms, 128-prime
    // batches at 0.257e9 3.95 -> 3.57 ms and at 4.237e9 23.3 -> 20.8 ms)
    const u32 chunk = filter_chunk(c) ? filter_chunk(c)
                      : tasks > 10000ull * grid * 4u ? 16u : 0u;
    // a per-XCD counter ends at most one claim per wave past its items
    if (std::max<u64>(std::max(dmax[0], dmax[1]) * (nf + 8ull * xcd_group(c)), most) +
            4ull * grid * (chunk ? chunk : 16u) > 0xFFFFFFFFull)
        return c->fail(RBG_ECAPACITY, "too many filter tasks");
    HIPCK(c, c->sp_tpre.ensure(tpre.size() * 4));
    HIPCK(c, hipMemcpyAsync(c->sp_tpre.ptr, tpre.data(), tpre.size() * 4, hipMemcpyHostToDevice,
                            s));
    // one chain: clear flags and counters, tables and planes, seed, filter
    // (eagerly with an event per stage; in graph mode the whole chain is
    // timed as the filter)
    HIPCK(c, hipEventRecord(c->ev[kPrep0], s));
    if (c->graph)
        for (int e : {kPrep1, kSeed0, kSeed1, kFilt0}) HIPCK(c, hipEventRecord(c->ev[e], s));
    const int chain_rc = launch_chain(c, 1, [&]() -> int {
        HIPCK(c, rbg_fill_u32(failed, np, 0u, s));
        HIPCK(c, rbg_fill_u32((u32 *)c->sp_stats.ptr, (u32)(kStatsBytes / 4), 0u, s));
        HIPCK(c, rbg_sparse_prepare(dP, np, dF, nf, (u32)lblk, (u32)iblk, (u32)pblk,
                                    (const u32 *)c->sp_val.ptr, (const u64 *)c->sp_val64.ptr,
                                    (const u16 *)c->sp_lab.ptr, (u16 *)c->sp_labels.ptr,
                                    (u32 *)c->sp_inv.ptr, (u64 *)c->sp_inv64.ptr,
                                    (u32 *)c->sp_planes.ptr, err, s));
        if (!c->graph) HIPCK(c, hipEventRecord(c->ev[kPrep1], s));
        if (!c->graph) HIPCK(c, hipEventRecord(c->ev[kSeed0], s));
        HIPCK(c, rbg_sparse_seed(dP, dF, nf, (const u16 *)c->sp_labels.ptr, c->sp_segs.ptr,
                                 segs_per_prof, best, failed, s));
        if (!c->graph) HIPCK(c, hipEventRecord(c->ev[kSeed1], s));
        if (!c->graph) HIPCK(c, hipEventRecord(c->ev[kFilt0], s));
        // one filter launch per domain, each over its own profiles (the
        // profiles of the primes below 2^32 come first)
        for (int wide = 0; wide < 2; ++wide) {
            if (!nfd[wide]) continue;
            const u32 f0 = wide ? c->nprof32 : 0u;
            // the 32-bit launch's lists take nprof32 + 8 entries of tpre
            const u32 *tp = (const u32 *)c->sp_tpre.ptr + (wide ? c->nprof32 + 8u : 0u);
            HIPCK(c, rbg_sparse_filter(dP, dF + f0, nfd[wide], (u32)dmax[wide], tp,
                                       (const u32 *)c->sp_planes.ptr,
                                       wide ? c->sp_inv64.ptr : c->sp_inv.ptr,
                                       wide ? c->sp_val64.ptr : c->sp_val.ptr,
                                       (const u16 *)c->sp_lab.ptr, best + f0, failed,
                                       (unsigned long long *)c->sp_stats.ptr,
                                       (u32 *)((char *)c->sp_stats.ptr + kQueueOffset) + 8 * wide,
                                       wide, wide ? grid_wide : grid, xcd_group(c),
                                       chunk,
                                       extend_limit(c), s));
        }
        return RBG_OK;
    });
    if (chain_rc) return chain_rc;
    HIPCK(c, hipEventRecord(c->ev[kFilt1], s));
    c->timing.launches += 6;
    std::vector<u32> h_failed(np, 0);
    u64 h_stats[4] = {0, 0, 0, 0};
    std::vector<u64> h_best(nf, 0);
    HIPCK(c, hipMemcpyAsync(h_failed.data(), failed, np * 4, hipMemcpyDeviceToHost, s));
    HIPCK(c, hipMemcpyAsync(h_stats, c->sp_stats.ptr, sizeof h_stats, hipMemcpyDeviceToHost, s));
    HIPCK(c, hipMemcpyAsync(h_best.data(), best, nf * 8, hipMemcpyDeviceToHost, s));
    HIPCK(c, hipStreamSynchronize(s));
    // a profile whose longest run found stays below its speculative
    // threshold may hold longer runs below it that the filter skipped: its
    // prime is recomputed like a refused one
    for (u32 k = 0; k < nf; ++k) {
        const SparseProf &f = c->h_prof[k];
        if (f.qspec && (u32)(h_best[k] >> 32) < f.qspec) {
            h_failed[f.prime] |= 8u;
            ++c->timing.spec_misses;
        }
    }
    const auto wait0 = std::chrono::steady_clock::now();
    a4_thread.join();
    c->timing.a4_wait_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wait0).count();
    c->timing.a4_host_ms += a4_ms;
    if (a4_rc) return c->fail(a4_rc, "elliptic A4 count failed");
    HIPCK(c, hipMemcpyAsync(c->sp_a4.ptr, a4.data(), np * 4, hipMemcpyHostToDevice, s));
    HIPCK(c, rbg_sparse_finalize(dP, np, dF, best, (const u32 *)c->sp_a4.ptr, recs, s));
    HIPCK(c, hipStreamSynchronize(s));  // a4 lives on this stack frame
    for (u32 i = 0; i < np; ++i)
        if (h_failed[i]) failed_idx.push_back(c->h_sp[i].index);
    c->timing.fallbacks += failed_idx.size();
    c->timing.sparse_prep_ms = elapsed(c, kPrep0, kPrep1);
    c->timing.sparse_seed_ms = elapsed(c, kSeed0, kSeed1);
    c->timing.sparse_filter_ms = elapsed(c, kFilt0, kFilt1);
    c->timing.filter_windows = h_stats[0];
    c->timing.filter_words = h_stats[1];
    c->timing.filter_completions = h_stats[2];
    c->timing.filter_loads = h_stats[3];
    return RBG_OK;
}

// Device workspace one prime needs under `flags` (the planning rules of
// sparse_plan and dense_run).
u64 prime_bytes(u64 p, u32 flags, u32 tile_pos, u32 q, u32 div) {
    if (p < 3) return sizeof(rbg_record);
    const u32 mask = mask_of(p), m = lcm_of(mask), nr = (u32)__builtin_popcount(mask);
    const u64 vbytes = p >= (1ull << 32) ? 8 : 4;  // root-table and inverse values
    u64 bytes = sizeof(rbg_record) + (u64)m * (vbytes + 4) + 64;
    u32 B = 0, D = 0;
    if ((flags & RBG_METHOD_MASK) == RBG_METHOD_BATCHED && p >= kSparseMin &&
        choose_cover(p, q, div, B, D)) {
        bytes += 2ull * (B + 1) + vbytes * (D + 1) + 88;
        for (u32 r = 2; r <= 9; ++r)
            if (mask & (1u << (r - 2)))
                bytes += 4ull * rbg::row_offset(rbg::kRows, B, 32u - (u32)__builtin_clz(r - 1)) + 48;
        return bytes;
    }
    const u64 tiles = ((p - 1) / 2 + tile_pos - 1) / tile_pos;
    return bytes + tiles * (sizeof(TileHdr) + nr * sizeof(RunSum));
}

// Offsets one sparse prime adds to a call's label array (u16 entries), plane
// array (units of 16 words) and plane-block grid; each has a 32-bit (blocks:
// 31-bit) range per call.  Zero for a prime the dense scan takes.
struct SparseExtent {
    u64 labels = 0, words16 = 0, blocks = 0;
};

SparseExtent sparse_extent(u64 p, u32 flags, u32 q, u32 div) {
    SparseExtent e;
    u32 B = 0, D = 0;
    if ((flags & RBG_METHOD_MASK) != RBG_METHOD_BATCHED || p < kSparseMin ||
        !choose_cover(p, q, div, B, D))
        return e;
    const u32 mask = mask_of(p);
    e.labels = B + 1ull;
    for (u32 r = 2; r <= 9; ++r)
        if (mask & (1u << (r - 2))) {
            e.words16 += rbg::row_offset(rbg::kRows, B, 32u - (u32)__builtin_clz(r - 1)) / 16u;
            for (u32 row = 0; row < rbg::kRows; ++row) e.blocks += rbg::row_blocks(row, B);
        }
    return e;
}

// Profiles a block of primes below `pmax` (2^32 for the 32-bit entry points,
// 2^33 for the wide ones).  Records keep the low word of p; the wide entry
// points widen them.
int run_block(rbg_ctx *c, const u64 *h_primes, u32 n, u32 flags, rbg_record *d_out,
              rbg_record *h_out, u32 limit = RBG_MAX_BLOCK, u64 pmax = 1ull << 32,
              bool trusted = false) {
    if (n > limit) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    c->error.clear();
    c->timing = rbg_timing{};
    c->graph = (flags & RBG_GRAPH) != 0;
    const u32 method = flags & RBG_METHOD_MASK;
    if (method > RBG_METHOD_DENSE) return c->fail(RBG_EINVAL, "unknown method");
    if (cover_q(c) < 3 || cover_q(c) > kMaxCoverQ)
        return c->fail(RBG_EINVAL, "cover_q outside 3..63");
    if (c->cfg.cover_div == 1 || c->cfg.cover_div > 8)
        return c->fail(RBG_EINVAL, "cover_div outside 2..8");
    if (c->cfg.filter_chunk > 256) return c->fail(RBG_EINVAL, "filter_chunk outside 1..256");
    std::vector<u32> odd, dense, sparse, failed;
    for (u32 i = 0; i < n; ++i) {
        const u64 p = h_primes[i];
        if (p >= pmax || (!trusted && !prime_u64(p))) {
            char msg[128];
            snprintf(msg, sizeof msg, "not a prime below 2^%d: %llu (index %u)",
                     pmax > (1ull << 32) ? 33 : 32, (unsigned long long)p, i);
            return c->fail(RBG_EINVAL, msg);
        }
        if (p != 2) odd.push_back(i);
    }
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, c->flags.ensure(3 * sizeof(u32)));  // tile counter, error flags, tile counter
    rbg_record *recs = d_out;
    if (!recs) {
        HIPCK(c, c->recs.ensure((size_t)std::max(1u, n) * sizeof(rbg_record)));
        recs = (rbg_record *)c->recs.ptr;
    }
    hipStream_t s = c->stream;
    u32 *err = (u32 *)c->flags.ptr + 1;
    HIPCK(c, hipEventRecord(c->ev[kAll0], s));
    HIPCK(c, hipMemsetAsync(err, 0, sizeof(u32), s));
    if (n) {  // p = 2 (mask 0, A4 placeholder 0) and zeroed slots for the rest
        HIPCK(c, hipMemsetAsync(recs, 0, (size_t)n * sizeof(rbg_record), s));
        for (u32 i = 0; i < n; ++i)
            if (h_primes[i] == 2) {
                static const u32 two = 2;
                HIPCK(c, hipMemcpyAsync(&recs[i].p, &two, sizeof two, hipMemcpyHostToDevice, s));
            }
    }
    if (method == RBG_METHOD_BATCHED) {
        int rc = sparse_plan(c, h_primes, odd, dense);
        if (rc) return rc;
        rc = sparse_run(c, recs, err, failed);
        if (rc) return rc;
        dense.insert(dense.end(), failed.begin(), failed.end());
    } else {
        dense = odd;
    }
    int rc = dense_run(c, h_primes, dense, method == RBG_METHOD_PLAIN, recs, err);
    if (rc) return rc;
    u32 h_err = 0;
    HIPCK(c, hipMemcpyAsync(&h_err, err, sizeof(u32), hipMemcpyDeviceToHost, s));
    if (!dense.empty()) {
        HIPCK(c, hipStreamSynchronize(s));
        if (h_err == 4u) {  // a longest run below its floor: rescan the dense primes exactly
            c->exact_runs = true;
            HIPCK(c, hipMemsetAsync(err, 0, sizeof(u32), s));
            rc = dense_run(c, h_primes, dense, method == RBG_METHOD_PLAIN, recs, err);
            c->exact_runs = false;
            if (rc) return rc;
            ++c->timing.rescans;
            HIPCK(c, hipMemcpyAsync(&h_err, err, sizeof(u32), hipMemcpyDeviceToHost, s));
        }
    }
    if (h_out && n)
        HIPCK(c, hipMemcpyAsync(h_out, recs, (size_t)n * sizeof(rbg_record), hipMemcpyDeviceToHost,
                                s));
    HIPCK(c, hipEventRecord(c->ev[kAll1], s));
    HIPCK(c, hipStreamSynchronize(s));
    c->timing.total_ms = elapsed(c, kAll0, kAll1);
    if (c->timing.dense_primes) c->timing.dense_scan_ms = elapsed(c, kDense0, kDense1);
    if (h_err & 1u) return c->fail(RBG_EDEVICE, "character value missing from the root table");
    if (h_err & 2u) return c->fail(RBG_EDEVICE, "incomplete coverage of the half interval");
    if (h_err & 4u) return c->fail(RBG_EDEVICE, "longest run below its floor in an exact scan");
    if ((flags & RBG_CERTIFY) && h_out) return certify_all(c, h_out, h_primes, n);
    return RBG_OK;
}

std::vector<u64> widen(const uint32_t *primes, u32 n) { return std::vector<u64>(primes, primes + n); }

// rbg_profile_batch over 64-bit primes: greedy sub-batches under the HBM
// budget and the chunk limit, records into out (low word of p).
int run_batch(rbg_ctx *c, const u64 *primes, u32 n, u32 flags, rbg_record *out, u64 pmax) {
    const u32 tile_pos = (u32)rbg_waves_per_block() * tile_chunks(c) * 64u;
    const u64 cap = budget(c) / 10 * 9;
    rbg_timing sum{};
    u32 first = 0;
    while (first < n) {
        u32 last = first;
        u64 bytes = 0;
        SparseExtent used;
        while (last < n && last - first < kBatchChunk) {
            const u64 b = prime_bytes(primes[last], flags, tile_pos, cover_q(c), cover_div(c));
            const SparseExtent e = sparse_extent(primes[last], flags, cover_q(c), cover_div(c));
            // a sub-batch also keeps the sparse engine's per-call offsets in range
            const bool full = used.labels + e.labels > 0xFFFFFFFFull ||
                              used.words16 + e.words16 > 0xFFFFFFFFull ||
                              used.blocks + e.blocks > 0x7FFFFFFFull;
            if (last > first && (bytes + b > cap || full)) break;
            bytes += b;
            used.labels += e.labels;
            used.words16 += e.words16;
            used.blocks += e.blocks;
            ++last;
        }
        int rc = run_block(c, primes + first, last - first, flags, nullptr, out + first,
                           kBatchChunk, pmax);
        if (rc) return rc;
        for (u32 i = first; i < last; ++i) out[i].status = 0;
        const rbg_timing &t = c->timing;
        sum.total_ms += t.total_ms;
        sum.dense_scan_ms += t.dense_scan_ms;
        sum.sparse_prep_ms += t.sparse_prep_ms;
        sum.sparse_seed_ms += t.sparse_seed_ms;
        sum.sparse_filter_ms += t.sparse_filter_ms;
        sum.dense_primes += t.dense_primes;
        sum.sparse_primes += t.sparse_primes;
        sum.sparse_profiles += t.sparse_profiles;
        sum.fallbacks += t.fallbacks;
        sum.positions += t.positions;
        sum.products += t.products;
        sum.launches += t.launches;
        sum.filter_windows += t.filter_windows;
        sum.filter_words += t.filter_words;
        sum.filter_completions += t.filter_completions;
        sum.filter_loads += t.filter_loads;
        sum.spec_misses += t.spec_misses;
        sum.a4_host_ms += t.a4_host_ms;
        sum.a4_wait_ms += t.a4_wait_ms;
        sum.rescans += t.rescans;
        first = last;
    }
    c->timing = sum;
    return RBG_OK;
}

void widen_records(const u64 *primes, const rbg_record *in, u32 n, rbg_record_wide *out) {
    for (u32 i = 0; i < n; ++i) {
        rbg_record_wide &w = out[i];
        w.p = primes[i];
        w.mask = in[i].mask;
        std::memcpy(w.M, in[i].M, sizeof w.M);
        std::memcpy(w.L, in[i].L, sizeof w.L);
        std::memcpy(w.W, in[i].W, sizeof w.W);
        std::memcpy(w.K, in[i].K, sizeof w.K);
        w.A4 = in[i].A4;
        w.status = in[i].status;
    }
}

}  // namespace

extern "C" {

rbg_ctx *rbg_open(const rbg_config *cfg) {
    rbg_ctx *c = new rbg_ctx();
    if (cfg) c->cfg = *cfg;
    c->device = c->cfg.device;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= c->device || c->device < 0) {
        delete c;
        return nullptr;
    }
    hipDeviceProp_t prop;
    if (hipSetDevice(c->device) != hipSuccess ||
        hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
        delete c;
        return nullptr;
    }
    c->cus = prop.multiProcessorCount;
    c->clock_khz = prop.clockRate;
    c->hbm_bytes = prop.totalGlobalMem;
    const hipError_t se = rbg_scratch_setup(c->device, &c->scratch);
    if (se != hipSuccess) c->scratch_error = hipGetErrorString(se);
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
    for (int i = 0; i < kNumEv && ok; ++i) ok = hipEventCreate(&c->ev[i]) == hipSuccess;
    if (!ok) {
        rbg_close(c);
        return nullptr;
    }
    return c;
}

int rbg_profile_block(rbg_ctx *c, const uint32_t *primes, uint32_t n, uint32_t flags,
                      rbg_record *out) {
    if (!c || (!primes && n) || (!out && n)) return RBG_EINVAL;
    const std::vector<u64> wp = widen(primes, n);
    int rc = run_block(c, wp.data(), n, flags, nullptr, out);
    if (rc && out) memset(out, 0, sizeof(rbg_record) * n);
    return rc;
}

int rbg_profile_block_wide(rbg_ctx *c, const uint64_t *primes, uint32_t n, uint32_t flags,
                           rbg_record_wide *out) {
    if (!c || (!primes && n) || (!out && n)) return RBG_EINVAL;
    if (n > RBG_MAX_BLOCK) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    std::vector<rbg_record> recs(n);
    int rc = run_block(c, primes, n, flags, nullptr, recs.data(), RBG_MAX_BLOCK, kWideLimit);
    if (rc) {
        if (out) memset(out, 0, sizeof(rbg_record_wide) * n);
        return rc;
    }
    widen_records(primes, recs.data(), n, out);
    return RBG_OK;
}

int rbg_profile_batch_wide(rbg_ctx *c, const uint64_t *primes, uint32_t n, uint32_t flags,
                           rbg_record_wide *out) {
    if (!c || (!primes && n) || (!out && n)) return RBG_EINVAL;
    if (n > RBG_MAX_BATCH) return c->fail(RBG_EINVAL, "batch larger than RBG_MAX_BATCH");
    std::vector<rbg_record> recs(n);
    int rc = run_batch(c, primes, n, flags, recs.data(), kWideLimit);
    if (rc) {
        if (out) memset(out, 0, sizeof(rbg_record_wide) * n);
        return rc;
    }
    widen_records(primes, recs.data(), n, out);
    return RBG_OK;
}

int rbg_estimate_bytes_wide(const rbg_ctx *c, const uint64_t *primes, uint32_t n, uint32_t flags,
                            uint64_t *bytes) {
    if (!c || (!primes && n) || !bytes) return RBG_EINVAL;
    const u32 tile_pos = (u32)rbg_waves_per_block() * tile_chunks(c) * 64u;
    u64 total = 0;
    for (u32 i = 0; i < n; ++i) total += prime_bytes(primes[i], flags, tile_pos, cover_q(c), cover_div(c));
    *bytes = total;
    return RBG_OK;
}

int rbg_estimate_bytes(const rbg_ctx *c, const uint32_t *primes, uint32_t n, uint32_t flags,
                       uint64_t *bytes) {
    if (!c || (!primes && n) || !bytes) return RBG_EINVAL;
    const u32 tile_pos = (u32)rbg_waves_per_block() * tile_chunks(c) * 64u;
    u64 total = 0;
    for (u32 i = 0; i < n; ++i) total += prime_bytes(primes[i], flags, tile_pos, cover_q(c), cover_div(c));
    *bytes = total;
    return RBG_OK;
}

int rbg_profile_batch(rbg_ctx *c, const uint32_t *primes, uint32_t n, uint32_t flags,
                      rbg_record *out) {
    if (!c || (!primes && n) || (!out && n)) return RBG_EINVAL;
    if (n > RBG_MAX_BATCH) return c->fail(RBG_EINVAL, "batch larger than RBG_MAX_BATCH");
    const std::vector<u64> wp = widen(primes, n);
    int rc = run_batch(c, wp.data(), n, flags, out, 1ull << 32);
    if (rc) memset(out, 0, sizeof(rbg_record) * n);
    return rc;
}

int rbg_autotune(rbg_ctx *c, const uint32_t *primes, uint32_t n, uint32_t flags,
                 rbg_config *chosen) {
    if (!c || !primes || !n || n > RBG_MAX_BLOCK) return RBG_EINVAL;
    const u32 method = flags & RBG_METHOD_MASK;
    struct Cand {
        u32 chunks, bpc, group;
    };
    std::vector<Cand> cands;
    if (method == RBG_METHOD_BATCHED) {
        // the filter runs 6 waves per SIMD (6 blocks per CU resident): fewer
        // blocks leave SIMDs idle, more only queue behind the resident ones
        for (u32 bpc : {6u, 8u, 12u})
            for (u32 g : {1u, 2u, 4u}) cands.push_back({c->cfg.tile_chunks, bpc, g});
    } else {
        for (u32 ch : {16u, 32u, 64u, 128u, 256u})
            for (u32 bpc : {4u, 8u}) cands.push_back({ch, bpc, c->cfg.xcd_group});
    }
    const rbg_config saved = c->cfg;
    const std::vector<u64> wp = widen(primes, n);
    std::vector<rbg_record> ref(n), got(n);
    double best_ms = 1e300;
    Cand best = cands[0];
    bool have_ref = false;
    for (const Cand &k : cands) {
        c->cfg.tile_chunks = k.chunks;
        c->cfg.blocks_per_cu = k.bpc;
        c->cfg.xcd_group = k.group;
        double ms = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
            int rc = run_block(c, wp.data(), n, method, nullptr, got.data());
            if (rc) {
                c->cfg = saved;
                return rc;
            }
            const rbg_timing &t = c->timing;
            const double stage = method == RBG_METHOD_BATCHED
                                     ? t.sparse_filter_ms + t.sparse_seed_ms + t.dense_scan_ms
                                     : t.dense_scan_ms;
            ms = std::min(ms, stage);
        }
        // every configuration must give the same records
        if (!have_ref) {
            ref = got;
            have_ref = true;
        } else if (memcmp(ref.data(), got.data(), n * sizeof(rbg_record)) != 0) {
            c->cfg = saved;
            return c->fail(RBG_ECERTIFY, "autotune: configurations disagree");
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = k;
        }
    }
    c->cfg.tile_chunks = best.chunks;
    c->cfg.blocks_per_cu = best.bpc;
    c->cfg.xcd_group = best.group;
    if (chosen) *chosen = c->cfg;
    return RBG_OK;
}

int rbg_profile_block_device(rbg_ctx *c, const uint32_t *d_primes, uint32_t n, uint32_t flags,
                             rbg_record *d_out) {
    if (!c || (!d_primes && n) || (!d_out && n)) return RBG_EINVAL;
    if (n > RBG_MAX_BLOCK) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    std::vector<u32> h(n);
    if (n) HIPCK(c, hipMemcpy(h.data(), d_primes, n * 4, hipMemcpyDeviceToHost));
    const std::vector<u64> wp = widen(h.data(), n);
    return run_block(c, wp.data(), n, flags & ~RBG_CERTIFY, d_out, nullptr);
}

// ---- device-resident prime table (one per device, kept until exit) --------

namespace {

struct PrimeTable {
    std::mutex lock;
    bool built = false;
    u64 *bits = nullptr;
    u32 *counts = nullptr, *prefix = nullptr, *total = nullptr;
    u32 nodd = 0;  // odd primes below 2^32
    u64 *rank_out = nullptr;
};

PrimeTable g_tables[64];

// Odd primes 3 <= q < 2^16 (every odd composite below 2^32 has one of them
// as a factor q <= its square root).
std::vector<u32> small_odd_primes() {
    std::vector<unsigned char> comp(1u << 16, 0);
    std::vector<u32> out;
    for (u32 q = 3; q < (1u << 16); q += 2) {
        if (comp[q]) continue;
        out.push_back(q);
        for (u32 m = q * q; m < (1u << 16); m += 2 * q) comp[m] = 1;
    }
    return out;
}

// The table of ctx's device, built on first use (segmented sieve on the
// device, about 258 MiB of HBM outside the context's workspace budget).
int prime_table(rbg_ctx *c, PrimeTable **out) {
    if (c->device < 0 || c->device >= 64) return c->fail(RBG_EDEVICE, "device ordinal");
    PrimeTable &t = g_tables[c->device];
    std::lock_guard<std::mutex> lock(t.lock);
    if (!t.built) {
        HIPCK(c, hipSetDevice(c->device));
        const std::vector<u32> small = small_odd_primes();
        const size_t words = (size_t)1 << 25, blocks = words / 64;
        const size_t small_at = (rbg_pt_bytes() + 64 + 255) & ~(size_t)255;
        char *mem = nullptr;
        HIPCK(c, hipMalloc((void **)&mem, small_at + small.size() * 4));
        u32 *d_small = (u32 *)(mem + small_at);
        hipError_t e = hipMemcpy(d_small, small.data(), small.size() * 4, hipMemcpyHostToDevice);
        u64 *bits = (u64 *)mem;
        u32 *counts = (u32 *)(mem + words * 8), *prefix = counts + blocks,
            *total = prefix + blocks;
        if (e == hipSuccess)
            e = rbg_pt_build(d_small, (u32)small.size(), bits, counts, prefix, total, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(&t.nodd, total, 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            (void)hipFree(mem);
            return hip_fail(c, e, "building the prime table");
        }
        t.bits = bits;
        t.counts = counts;
        t.prefix = prefix;
        t.total = total;
        t.rank_out = (u64 *)(mem + rbg_pt_bytes());
        t.built = true;
    }
    *out = &t;
    return RBG_OK;
}

}  // namespace

int rbg_prime_table(rbg_ctx *c, uint64_t *count) {
    if (!c) return RBG_EINVAL;
    PrimeTable *t;
    if (int rc = prime_table(c, &t)) return rc;
    if (count) *count = (u64)t->nodd + 1;
    return RBG_OK;
}

int rbg_prime_rank(rbg_ctx *c, uint64_t x, uint64_t *rank) {
    if (!c || !rank) return RBG_EINVAL;
    if (x > (1ull << 32)) return c->fail(RBG_EINVAL, "rank beyond 2^32");
    PrimeTable *t;
    if (int rc = prime_table(c, &t)) return rc;
    std::lock_guard<std::mutex> lock(t->lock);  // one rank slot per table
    HIPCK(c, rbg_pt_rank(t->bits, t->prefix, x, t->rank_out, c->stream));
    HIPCK(c, hipMemcpyAsync(rank, t->rank_out, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return RBG_OK;
}

namespace {

// Primes of ranks first .. first+n-1 into the context's device buffer and
// into h_out; *found = how many exist below 2^32.
int select_ranks(rbg_ctx *c, u64 first, u32 n, u32 *h_out, u32 *found) {
    PrimeTable *t;
    if (int rc = prime_table(c, &t)) return rc;
    const u64 total = (u64)t->nodd + 1;
    const u32 m = first >= total ? 0u : (u32)std::min<u64>(n, total - first);
    if (found) *found = m;
    if (!m) return RBG_OK;
    HIPCK(c, c->sel.ensure((size_t)m * 4));
    HIPCK(c, rbg_pt_select(t->bits, t->counts, t->prefix, t->nodd, first, m, (u32 *)c->sel.ptr,
                           c->stream));
    HIPCK(c, hipMemcpyAsync(h_out, c->sel.ptr, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    return RBG_OK;
}

}  // namespace

int rbg_primes_by_rank(rbg_ctx *c, uint64_t first, uint32_t n, uint32_t *out, uint32_t *found) {
    if (!c || (!out && n) || !found) return RBG_EINVAL;
    if (n > RBG_MAX_BATCH) return c->fail(RBG_EINVAL, "more than RBG_MAX_BATCH ranks");
    return select_ranks(c, first, n, out, found);
}

int rbg_profile_ranked(rbg_ctx *c, uint64_t first, uint32_t n, uint32_t flags, rbg_record *out,
                       uint32_t *found) {
    if (!c || (!out && n) || !found) return RBG_EINVAL;
    if (n > RBG_MAX_BLOCK) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    std::vector<u32> primes(n);
    if (int rc = select_ranks(c, first, n, primes.data(), found)) return rc;
    const std::vector<u64> wp(primes.begin(), primes.begin() + *found);
    // primes of the sieve need no primality test
    const int rc = run_block(c, wp.data(), *found, flags, nullptr, out, RBG_MAX_BLOCK,
                             1ull << 32, true);
    if (rc && out) memset(out, 0, sizeof(rbg_record) * n);
    return rc;
}

// One lock per device: calls of the allocation-free entry point share that
// device's scratch array and run one at a time.
static std::mutex g_scratch_users[64];

int rbg_profile_block_noalloc(rbg_ctx *c, const uint32_t *d_primes, uint32_t n, uint32_t flags,
                              rbg_record *d_out) {
    if (!c || (!d_primes && n) || (!d_out && n)) return RBG_EINVAL;
    if (n > RBG_MAX_BLOCK) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    if (flags & RBG_GRAPH)
        return c->fail(RBG_EINVAL, "graph capture is not available without allocation");
    if (!c->scratch || c->device < 0 || c->device >= 64)
        return c->fail(RBG_EDEVICE, "no scratch array on this device: " + c->scratch_error);
    HIPCK(c, hipSetDevice(c->device));
    std::vector<u32> h(n);
    if (n) HIPCK(c, hipMemcpy(h.data(), d_primes, n * 4, hipMemcpyDeviceToHost));
    const std::vector<u64> wp = widen(h.data(), n);
    const size_t cap = rbg_scratch_capacity();
    // refuse before any work when the planned workspace cannot fit (primes
    // that the sparse engine refuses later need the dense workspace too,
    // which the carving checks again)
    const u32 tile_pos = (u32)rbg_waves_per_block() * tile_chunks(c) * 64u;
    u64 need = 0;
    for (u32 i = 0; i < n; ++i)
        need += prime_bytes(wp[i], flags & RBG_METHOD_MASK, tile_pos, cover_q(c), cover_div(c));
    if (need > cap) {
        char msg[160];
        snprintf(msg, sizeof msg, "block needs about %llu bytes of device workspace; "
                 "rbg_profile_block_noalloc has %zu", (unsigned long long)need, cap);
        return c->fail(RBG_ECAPACITY, msg);
    }
    std::lock_guard<std::mutex> lock(g_scratch_users[c->device]);
    Arena arena;
    arena.base = (char *)c->scratch;
    arena.cap = cap;
    std::vector<DevBuf *> bufs = c->buffers();
    std::vector<DevBuf> kept;
    for (DevBuf *b : bufs) {
        kept.push_back(*b);
        *b = DevBuf{};
        b->arena = &arena;
    }
    c->arena = &arena;
    int rc = run_block(c, wp.data(), n, flags & ~RBG_CERTIFY, d_out, nullptr);
    // nothing of this call may still run on the scratch when the next one
    // takes it
    if (rc) (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < bufs.size(); ++i) *bufs[i] = kept[i];
    c->arena = nullptr;
    return rc;
}

const char *rbg_last_error(const rbg_ctx *c) { return c ? c->error.c_str() : "no context"; }

int rbg_last_timing(const rbg_ctx *c, rbg_timing *out) {
    if (!c || !out) return RBG_EINVAL;
    *out = c->timing;
    return RBG_OK;
}

int rbg_device_info(const rbg_ctx *c, uint32_t *cus, uint32_t *clock_khz, uint64_t *hbm_bytes) {
    if (!c) return RBG_EINVAL;
    if (cus) *cus = (uint32_t)c->cus;
    if (clock_khz) *clock_khz = (uint32_t)c->clock_khz;
    if (hbm_bytes) *hbm_bytes = c->hbm_bytes;
    return RBG_OK;
}

void rbg_close(rbg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    for (DevBuf *b : c->buffers()) b->release();
    for (int i = 0; i < kNumEv; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    for (hipGraphExec_t &x : c->gexec)
        if (x) (void)hipGraphExecDestroy(x);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int rbg_certify_records(const rbg_record *recs, uint32_t n, uint32_t threads,
                        uint32_t *first_bad) {
    if (!recs && n) return RBG_EINVAL;
    std::vector<u64> primes(n);
    for (u32 i = 0; i < n; ++i) {
        primes[i] = recs[i].p;
        if (!prime_u32(recs[i].p)) return RBG_EINVAL;
    }
    const long long b = first_uncertified(recs, primes.data(), n,
                                          threads ? threads : host_threads());
    if (first_bad) *first_bad = b < 0 ? n : (uint32_t)b;
    return b < 0 ? RBG_OK : RBG_ECERTIFY;
}

uint32_t rbg_mask(uint32_t p) { return mask_of(p); }

uint32_t rbg_primes_from(uint64_t start, uint32_t count, uint32_t *out) {
    if (!out) return 0;
    u32 k = 0;
    for (u64 x = start < 2 ? 2 : start; k < count && x < (1ull << 32); ++x)
        if (prime_u32((u32)x)) out[k++] = (u32)x;
    return k;
}

int rbg_is_prime(uint32_t p) { return prime_u32(p) ? 1 : 0; }

int rbg_is_prime_wide(uint64_t p) { return p < kWideLimit && prime_u64(p) ? 1 : 0; }

uint32_t rbg_mask_wide(uint64_t p) { return mask_of(p); }

uint32_t rbg_primes_from_wide(uint64_t start, uint32_t count, uint64_t *out) {
    if (!out) return 0;
    u32 k = 0;
    for (u64 x = start < 2 ? 2 : start; k < count && x < kWideLimit; ++x)
        if (prime_u64(x)) out[k++] = x;
    return k;
}

const char *rbg_version(void) { return "rabung-mi355x 0.2 (gfx950 dense scan + sparse anchor filter)"; }

// ---- reference bridge --------------------------------------------------

static std::mutex g_bridge_mutex;
static rbg_ctx *g_bridge = nullptr;

void baseline_init(unsigned cap) {
    (void)cap;  // advisory: the context's workspaces grow on demand
    std::lock_guard<std::mutex> lock(g_bridge_mutex);
    if (!g_bridge) g_bridge = rbg_open(nullptr);
}

// Return codes: 0; 2 for a mask the call cannot serve (without `full`: empty,
// or naming an order that does not divide p - 1); -1 for any other failure
// (not a prime below 2^32, no device, a device error), which the reference
// would have ended the process for.
int baseline_profile(unsigned p, unsigned mask, unsigned *out, unsigned *prefix, int full) {
    std::lock_guard<std::mutex> lock(g_bridge_mutex);
    const u32 adm = mask_of(p);
    if (!full && (mask == 0 || (mask & ~adm))) return 2;
    if (!out || !prefix) return -1;
    if (!g_bridge) g_bridge = rbg_open(nullptr);
    if (!g_bridge) return -1;
    rbg_record rec;
    if (rbg_profile_block(g_bridge, &p, 1, RBG_METHOD_PLAIN, &rec) != RBG_OK) return -1;
    // only the caller's orders are written; with `full` an order of the mask
    // that does not divide p - 1 has no profile and gets 0
    for (u32 r = 2; r <= 9; ++r)
        if (mask & (1u << (r - 2))) {
            const bool have = (adm >> (r - 2)) & 1u;
            out[r] = have ? rec.M[r - 2] : 0u;
            prefix[r] = have ? rec.L[r - 2] : 0u;
        }
    return 0;
}

}  // extern "C"
