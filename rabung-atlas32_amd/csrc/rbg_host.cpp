// Host orchestration and C ABI of the census library.
//
// Replaces the reference's per-prime host chain (src/census/worker32.cu:20-41;
// src/census/engine32/src/frozen_entry.cu:149-242 prepare_gpu_table /
// gpu_prepared_group) for the dense path: per-prime constants and the sorted
// root table are prepared on the host (character_root, frozen_entry.cu:18-24,
// 158-167; original_scan_gpu.cu:369-376), uploaded in one copy, and the whole
// block is scanned by one persistent launch (or one launch per prime for the
// plain method), followed by one ordered merge launch and one copy back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "rabung.h"
#include "rbg_types.hpp"

using rbg::PrimeDesc;
using rbg::RunSum;
using rbg::TileHdr;
using rbg::u16;
using rbg::u32;
using rbg::u64;

extern "C" hipError_t rbg_launch_scan(const PrimeDesc *descs, uint32_t nprimes, uint32_t tile_first,
                                      uint32_t tile_end, uint32_t chunks, const uint32_t *tab_val,
                                      const uint16_t *tab_code, TileHdr *hdr, RunSum *runs,
                                      uint32_t *counter, uint32_t *err, int persistent,
                                      uint32_t grid, hipStream_t stream);
extern "C" hipError_t rbg_launch_merge(const PrimeDesc *descs, uint32_t nprimes,
                                       const TileHdr *hdr, const RunSum *runs, rbg_record *out,
                                       uint32_t *err, hipStream_t stream);
extern "C" int rbg_waves_per_block(void);

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

u32 mask_of(u32 p) {
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
u32 root_of_order(u32 p, u32 m) {
    const u32 e = (p - 1) / m;
    for (u32 a = 2; a < p; ++a) {
        u32 z = rbg::powmod64(a, e, p);
        bool ok = true;
        for (u32 q : {2u, 3u, 5u, 7u})
            if (m % q == 0 && rbg::powmod64(z, m / q, p) == 1) {
                ok = false;
                break;
            }
        if (ok) return z;
    }
    return 0;
}

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t want) {
        if (want <= bytes) return hipSuccess;
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        size_t cap = std::max(want, (size_t)4096);
        hipError_t e = hipMalloc(&ptr, cap);
        if (e == hipSuccess) bytes = cap;
        return e;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

}  // namespace

struct rbg_ctx {
    int device = 0;
    int cus = 0;
    int clock_khz = 0;
    size_t hbm_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_scan0 = nullptr, ev_scan1 = nullptr, ev_all0 = nullptr, ev_all1 = nullptr;
    rbg_config cfg{};
    std::string error;
    DevBuf descs, tab_val, tab_code, hdr, runs, recs, flags, primes;
    std::vector<PrimeDesc> h_descs;
    std::vector<u32> h_val;
    std::vector<u16> h_code;
    double scan_ms = 0, total_ms = 0;
    u64 launches = 0, positions = 0, products = 0;

    int fail(int code, const std::string &msg) {
        error = msg;
        return code;
    }
};

namespace {

int hip_fail(rbg_ctx *c, hipError_t e, const char *what) {
    return c->fail(RBG_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCK(ctx, expr)                                          \
    do {                                                          \
        hipError_t _e = (expr);                                   \
        if (_e != hipSuccess) return hip_fail((ctx), _e, #expr); \
    } while (0)

u32 tile_chunks(const rbg_ctx *c) { return c->cfg.tile_chunks ? c->cfg.tile_chunks : 64u; }

u32 blocks_per_cu(const rbg_ctx *c) { return c->cfg.blocks_per_cu ? c->cfg.blocks_per_cu : 8u; }

// Host verification of one profile by direct powers (reference
// certify_witness, VBB/src/advance.cpp:7-13, and the prefix loop of
// src/census/worker32.cu:37-38).
bool certify(u32 p, u32 r, u32 M, u32 L, u32 W) {
    const u64 e = (p - 1) / r;
    if (M == 0 || W == 0 || (u64)W + M > p || L == 0) return false;
    const u32 colour = rbg::powmod64(W, e, p);
    for (u32 j = 1; j < M; ++j)
        if (rbg::powmod64(W + j, e, p) != colour) return false;
    if (W > 1 && rbg::powmod64(W - 1, e, p) == colour) return false;
    if ((u64)W + M < p && rbg::powmod64(W + M, e, p) == colour) return false;
    for (u32 x = 1; x <= L; ++x)
        if (rbg::powmod64(x, e, p) != 1) return false;
    if ((u64)L + 1 < p && rbg::powmod64(L + 1, e, p) == 1) return false;
    return true;
}

int certify_all(rbg_ctx *c, const rbg_record *out, u32 n) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<int> bad(nt, -1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (u32 i = t; i < n; i += nt) {
                const rbg_record &rec = out[i];
                for (u32 r = 2; r <= 9; ++r)
                    if (rec.mask & (1u << (r - 2)))
                        if (!certify(rec.p, r, rec.M[r - 2], rec.L[r - 2], rec.W[r - 2])) {
                            bad[t] = (int)i;
                            return;
                        }
            }
        });
    for (auto &t : th) t.join();
    for (int b : bad)
        if (b >= 0) {
            char msg[128];
            snprintf(msg, sizeof msg, "direct verification failed for p=%u", out[b].p);
            return c->fail(RBG_ECERTIFY, msg);
        }
    return RBG_OK;
}

// Prepare the device descriptors for the odd primes of the block; p = 2 and
// the layout of records are handled by the caller.
int prepare(rbg_ctx *c, const u32 *primes, u32 n, u64 *total_tiles, u64 *total_runs) {
    const u32 tile_pos = (u32)rbg_waves_per_block() * tile_chunks(c) * 64u;
    c->h_descs.clear();
    c->h_val.clear();
    c->h_code.clear();
    u64 tiles = 0, runs = 0, products = 0, positions = 0;
    for (u32 i = 0; i < n; ++i) {
        const u32 p = primes[i];
        if (!prime_u32(p)) {
            char msg[96];
            snprintf(msg, sizeof msg, "not a prime below 2^32: %u (index %u)", p, i);
            return c->fail(RBG_EINVAL, msg);
        }
        if (p == 2) continue;
        PrimeDesc d{};
        d.p = p;
        d.nh = (p - 1) / 2;
        d.mask = mask_of(p);
        d.m = lcm_of(d.mask);
        d.e = (p - 1) / d.m;
        d.ninv = rbg::neg_inverse32(p);
        d.one = (u32)((1ull << 32) % p);
        d.r2 = (u32)((u64)d.one * d.one % p);
        d.step = (u32)((u64)64 * d.one % p);
        d.tab_off = (u32)c->h_val.size();
        d.planes = 0;
        for (u32 r = 2; r <= 9; ++r)
            if (d.mask & (1u << (r - 2))) d.planes |= rbg::plane_bits(r);
        d.ntiles = (d.nh + tile_pos - 1) / tile_pos;
        d.tile_begin = (u32)tiles;
        d.nr = (u32)__builtin_popcount(d.mask);
        d.sum_off = (u32)runs;
        d.index = i;
        if (tiles + d.ntiles > 0xFFFFFFFFull || runs + (u64)d.ntiles * d.nr > 0xFFFFFFFFull)
            return c->fail(RBG_ECAPACITY, "block too large for one call");
        tiles += d.ntiles;
        runs += (u64)d.ntiles * d.nr;
        positions += d.nh;
        products += (u64)d.nh * (rbg::montpow_products(d.e) + 1);
        const u32 z = root_of_order(p, d.m);
        if (!z) return c->fail(RBG_EINVAL, "no root of the required order");
        std::vector<std::pair<u32, u32>> tab(d.m);
        u32 cur = 1;
        for (u32 j = 0; j < d.m; ++j) {
            tab[j] = {(u32)(((u64)cur << 32) % p), rbg::colour_code(j, d.m)};
            cur = rbg::mulmod64(cur, z, p);
        }
        std::sort(tab.begin(), tab.end());
        for (auto &t : tab) {
            c->h_val.push_back(t.first);
            c->h_code.push_back((u16)t.second);
        }
        c->h_descs.push_back(d);
    }
    c->positions = positions;
    c->products = products;
    *total_tiles = tiles;
    *total_runs = runs;
    return RBG_OK;
}

int run_block(rbg_ctx *c, const u32 *h_primes, u32 n, u32 flags, rbg_record *d_out,
              rbg_record *h_out) {
    if (n > RBG_MAX_BLOCK) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    c->error.clear();
    c->scan_ms = c->total_ms = 0;
    c->launches = 0;
    u64 tiles = 0, runs = 0;
    int rc = prepare(c, h_primes, n, &tiles, &runs);
    if (rc) return rc;
    const u32 np = (u32)c->h_descs.size();
    const size_t bytes = np * sizeof(PrimeDesc) + c->h_val.size() * 6 + tiles * sizeof(TileHdr) +
                         runs * sizeof(RunSum) + (size_t)n * sizeof(rbg_record);
    const u64 budget = c->cfg.hbm_budget ? c->cfg.hbm_budget : (16ull << 30);
    if (bytes > budget) return c->fail(RBG_ECAPACITY, "workspace exceeds the HBM budget");
    HIPCK(c, hipSetDevice(c->device));
    HIPCK(c, c->descs.ensure(std::max<size_t>(1, np) * sizeof(PrimeDesc)));
    HIPCK(c, c->tab_val.ensure(std::max<size_t>(1, c->h_val.size()) * 4));
    HIPCK(c, c->tab_code.ensure(std::max<size_t>(1, c->h_code.size()) * 2));
    HIPCK(c, c->hdr.ensure(std::max<u64>(1, tiles) * sizeof(TileHdr)));
    HIPCK(c, c->runs.ensure(std::max<u64>(1, runs) * sizeof(RunSum)));
    HIPCK(c, c->flags.ensure(2 * sizeof(u32)));
    rbg_record *recs = d_out;
    if (!recs) {
        HIPCK(c, c->recs.ensure((size_t)std::max(1u, n) * sizeof(rbg_record)));
        recs = (rbg_record *)c->recs.ptr;
    }
    hipStream_t s = c->stream;
    HIPCK(c, hipEventRecord(c->ev_all0, s));
    if (np) {
        HIPCK(c, hipMemcpyAsync(c->descs.ptr, c->h_descs.data(), np * sizeof(PrimeDesc),
                                hipMemcpyHostToDevice, s));
        HIPCK(c, hipMemcpyAsync(c->tab_val.ptr, c->h_val.data(), c->h_val.size() * 4,
                                hipMemcpyHostToDevice, s));
        HIPCK(c, hipMemcpyAsync(c->tab_code.ptr, c->h_code.data(), c->h_code.size() * 2,
                                hipMemcpyHostToDevice, s));
    }
    HIPCK(c, hipMemsetAsync(c->flags.ptr, 0, 2 * sizeof(u32), s));
    // p = 2 records (mask 0, A4 placeholder 0) and the rest come from the merge
    {
        std::vector<rbg_record> twos;
        for (u32 i = 0; i < n; ++i)
            if (h_primes[i] == 2) {
                rbg_record r{};
                r.p = 2;
                HIPCK(c, hipMemcpyAsync(recs + i, &r, sizeof r, hipMemcpyHostToDevice, s));
                HIPCK(c, hipStreamSynchronize(s));
            }
    }
    u32 *counter = (u32 *)c->flags.ptr;
    u32 *err = counter + 1;
    const PrimeDesc *dd = (const PrimeDesc *)c->descs.ptr;
    HIPCK(c, hipEventRecord(c->ev_scan0, s));
    if (np) {
        if (flags & RBG_METHOD_PLAIN) {
            for (u32 i = 0; i < np; ++i) {
                const PrimeDesc &d = c->h_descs[i];
                HIPCK(c, rbg_launch_scan(dd, np, d.tile_begin, d.tile_begin + d.ntiles,
                                         tile_chunks(c), (const u32 *)c->tab_val.ptr,
                                         (const u16 *)c->tab_code.ptr, (TileHdr *)c->hdr.ptr,
                                         (RunSum *)c->runs.ptr, counter, err, 0, d.ntiles, s));
                ++c->launches;
            }
        } else {
            const u32 grid = (u32)std::min<u64>((u64)c->cus * blocks_per_cu(c), tiles);
            HIPCK(c, rbg_launch_scan(dd, np, 0, (u32)tiles, tile_chunks(c),
                                     (const u32 *)c->tab_val.ptr, (const u16 *)c->tab_code.ptr,
                                     (TileHdr *)c->hdr.ptr, (RunSum *)c->runs.ptr, counter, err, 1,
                                     std::max(1u, grid), s));
            ++c->launches;
        }
    }
    HIPCK(c, hipEventRecord(c->ev_scan1, s));
    if (np)
        HIPCK(c, rbg_launch_merge(dd, np, (const TileHdr *)c->hdr.ptr, (const RunSum *)c->runs.ptr,
                                  recs, err, s));
    u32 h_err = 0;
    HIPCK(c, hipMemcpyAsync(&h_err, err, sizeof(u32), hipMemcpyDeviceToHost, s));
    if (h_out && n)
        HIPCK(c, hipMemcpyAsync(h_out, recs, (size_t)n * sizeof(rbg_record), hipMemcpyDeviceToHost,
                                s));
    HIPCK(c, hipEventRecord(c->ev_all1, s));
    HIPCK(c, hipStreamSynchronize(s));
    float ms = 0;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev_scan0, c->ev_scan1));
    c->scan_ms = ms;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev_all0, c->ev_all1));
    c->total_ms = ms;
    if (h_err & 1u) return c->fail(RBG_EDEVICE, "character value missing from the root table");
    if (h_err & 2u) return c->fail(RBG_EDEVICE, "incomplete coverage of the half interval");
    if ((flags & RBG_CERTIFY) && h_out) return certify_all(c, h_out, n);
    return RBG_OK;
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
    if (hipSetDevice(c->device) != hipSuccess) {
        delete c;
        return nullptr;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) {
        delete c;
        return nullptr;
    }
    c->cus = prop.multiProcessorCount;
    c->clock_khz = prop.clockRate;
    c->hbm_bytes = prop.totalGlobalMem;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_scan0) != hipSuccess || hipEventCreate(&c->ev_scan1) != hipSuccess ||
        hipEventCreate(&c->ev_all0) != hipSuccess || hipEventCreate(&c->ev_all1) != hipSuccess) {
        delete c;
        return nullptr;
    }
    return c;
}

int rbg_profile_block(rbg_ctx *c, const uint32_t *primes, uint32_t n, uint32_t flags,
                      rbg_record *out) {
    if (!c || (!primes && n) || (!out && n)) return RBG_EINVAL;
    int rc = run_block(c, primes, n, flags, nullptr, out);
    if (rc && out) memset(out, 0, sizeof(rbg_record) * n);
    return rc;
}

int rbg_profile_block_device(rbg_ctx *c, const uint32_t *d_primes, uint32_t n, uint32_t flags,
                             rbg_record *d_out) {
    if (!c || (!d_primes && n) || (!d_out && n)) return RBG_EINVAL;
    if (n > RBG_MAX_BLOCK) return c->fail(RBG_EINVAL, "more than 4096 primes in one block");
    std::vector<u32> h(n);
    if (n) HIPCK(c, hipMemcpy(h.data(), d_primes, n * 4, hipMemcpyDeviceToHost));
    return run_block(c, h.data(), n, flags & ~RBG_CERTIFY, d_out, nullptr);
}

const char *rbg_last_error(const rbg_ctx *c) { return c ? c->error.c_str() : "no context"; }

int rbg_last_timing(const rbg_ctx *c, double *scan_ms, double *total_ms, uint64_t *launches,
                    uint64_t *positions, uint64_t *products) {
    if (!c) return RBG_EINVAL;
    if (scan_ms) *scan_ms = c->scan_ms;
    if (total_ms) *total_ms = c->total_ms;
    if (launches) *launches = c->launches;
    if (positions) *positions = c->positions;
    if (products) *products = c->products;
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
    for (DevBuf *b : {&c->descs, &c->tab_val, &c->tab_code, &c->hdr, &c->runs, &c->recs, &c->flags,
                      &c->primes})
        b->release();
    if (c->ev_scan0) (void)hipEventDestroy(c->ev_scan0);
    if (c->ev_scan1) (void)hipEventDestroy(c->ev_scan1);
    if (c->ev_all0) (void)hipEventDestroy(c->ev_all0);
    if (c->ev_all1) (void)hipEventDestroy(c->ev_all1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

uint32_t rbg_mask(uint32_t p) { return mask_of(p); }

int rbg_is_prime(uint32_t p) { return prime_u32(p) ? 1 : 0; }

const char *rbg_version(void) { return "rabung-mi355x 0.1 (gfx950 dense scan)"; }

// ---- reference bridge --------------------------------------------------

static std::mutex g_bridge_mutex;
static rbg_ctx *g_bridge = nullptr;

void baseline_init(unsigned cap) {
    (void)cap;  // workspaces grow on demand
    std::lock_guard<std::mutex> lock(g_bridge_mutex);
    if (!g_bridge) g_bridge = rbg_open(nullptr);
}

int baseline_profile(unsigned p, unsigned mask, unsigned *out, unsigned *prefix, int full) {
    std::lock_guard<std::mutex> lock(g_bridge_mutex);
    if (!g_bridge) g_bridge = rbg_open(nullptr);
    if (!g_bridge) return RBG_EDEVICE;
    const u32 adm = mask_of(p);
    if (full) mask = adm;
    if (mask & ~adm) return 2;
    rbg_record rec;
    int rc = rbg_profile_block(g_bridge, &p, 1, RBG_METHOD_PLAIN, &rec);
    if (rc) return rc;
    for (u32 r = 2; r <= 9; ++r)
        if (mask & (1u << (r - 2))) {
            out[r] = rec.M[r - 2];
            prefix[r] = rec.L[r - 2];
        }
    return 0;
}

}  // extern "C"
