// Measurement probes for the roofline of the census kernels (gfx950).
//
//   * VALU issue rates of the instructions the kernels are made of: 32-bit
//     add, v_mul_lo_u32, v_mul_hi_u32, v_alignbit_b32, and whole Montgomery
//     products (the dense scan's unit of work).  Each thread runs 8
//     independent dependency chains so that latency is hidden and the rate is
//     the issue rate.
//   * The throughput cost of one wave-wide vector load: 1 to 4 dwords per
//     lane, lanes 4 bytes apart (consecutive dwords), 8 bytes apart (one
//     contiguous 512-byte span) or 4160 bytes apart (one line per lane),
//     from a buffer that stays in L2; each wave keeps 8 independent loads in
//     flight, so the figure is the issue cost, not the latency.
//
// No reference counterpart: the reference quotes its dense rate in source
// comments only (original_scan_gpu.cu:12-14).
#include <hip/hip_runtime.h>

#include "rabung.h"
#include "rbg_device.hpp"

namespace rbg {

constexpr int kChains = 8;

template <int OP>
__global__ __launch_bounds__(256) void valu_rate(u32 seed, u32 iters, u32 p, u32 ninv, u32 *out) {
    u32 v[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) v[c] = seed + threadIdx.x * 977u + c * 131u;
    for (u32 it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if (OP == 0) v[c] = v[c] + (v[(c + 1) % kChains] | 1u);
            if (OP == 1) v[c] = v[c] * (v[(c + 1) % kChains] | 1u);
            if (OP == 2) v[c] = __umulhi(v[c], v[(c + 1) % kChains] | 0x80000001u);
            if (OP == 3) v[c] = __builtin_amdgcn_alignbit(v[c], v[(c + 1) % kChains], v[c]);
            if (OP == 4) v[c] = mont_mul(v[c], v[c], p, pos_inverse(ninv));
            if (OP == 5) v[c] = v[c] + v[(c + 1) % kChains];
            if (OP == 6) v[c] = v[c] * v[(c + 1) % kChains];
        }
    }
    u32 x = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) x ^= v[c];
    if (x == 0x9E3779B9u) out[0] = x;  // keeps the chains alive
}

constexpr u32 kLoadsInFlight = 8;

// iters must be a multiple of kLoadsInFlight.
template <int NW>
__global__ __launch_bounds__(256) void load_rate(const u32 *__restrict__ buf, u32 mask, u32 stride,
                                                 u32 iters, u32 *out) {
    const u32 lane = threadIdx.x & 63u;
    const u32 wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    u32 acc = 0, off = (wave * 1024u) & mask;
    for (u32 it = 0; it < iters; it += kLoadsInFlight) {
        u32 v[kLoadsInFlight][NW];
#pragma unroll
        for (u32 u = 0; u < kLoadsInFlight; ++u) {
            // the buffer has 4 spare dwords past mask
            const u32 a = ((off + lane * stride + u * 12480u) & mask) >> 2;
#pragma unroll
            for (int w = 0; w < NW; ++w) v[u][w] = buf[a + w];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (u32 u = 0; u < kLoadsInFlight; ++u)
#pragma unroll
            for (int w = 0; w < NW; ++w) acc ^= v[u][w];
        off = (off + kLoadsInFlight * 12480u + 4160u) & mask;
    }
    if (acc == 0x9E3779B9u) out[0] = acc;
}

// Wide residue arithmetic on given operands (2^32 < p < 2^33): kind 0
// out[i] = a[i] b[i] 2^-32 mod p (one Montgomery product), kind 1
// out[i] = Montgomery power of a[i] to the exponent b[i].
__global__ void wide_check(const u64 *a, const u64 *b, u32 n, u64 p, u32 ninv, u64 one, int kind,
                           u64 *out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = kind == 0 ? mont_mul33(a[i], b[i], p, ninv) : mont_pow33(a[i], b[i], p, ninv, one);
}

// 32-bit residue arithmetic of the dense scan on given operands (odd p):
// kind 0 out[i] = a[i] b[i] 2^-32 mod p (one product, Dom32::mul);
// kind 1 out[i] = a[i]^b[i] mod p through the Montgomery domain (into the
// domain by r2 = 2^64 mod p, Dom32::pow, out by a product with 1);
// kind 2 out[i] = Montgomery form of a[i] + b[i] reached as the scan walks
// its positions: the form of a[i], then b[i] additions of step = 2^32 mod p.
__global__ void mont32_check(const u32 *a, const u32 *b, u32 n, u32 p, u32 ninv, u32 r2, u32 one,
                             int kind, u32 *out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Dom32 dm{p, pos_inverse(ninv)};
    if (kind == 0) {
        out[i] = dm.mul(a[i], b[i]);
    } else if (kind == 1) {
        out[i] = dm.mul(dm.pow(dm.mul(a[i], r2), b[i], one), 1u);
    } else {
        u32 x = dm.mul(a[i], r2);
        for (u32 k = 0; k < b[i]; ++k) x = dm.add(x, one);
        out[i] = x;
    }
}

}  // namespace rbg

namespace {

template <class L>
double time_ms(L launch) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1.0;
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        (void)hipEventRecord(a, 0);
        launch();
        (void)hipEventRecord(b, 0);
        if (hipEventSynchronize(b) != hipSuccess) return -1.0;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        if (ms < best) best = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return hipGetLastError() == hipSuccess ? best : -1.0;
}

}  // namespace

// kind 0..6: issue rate, in lane operations per second, of add(x, y|1) /
// mul_lo(x, y|1) / mul_hi(x, y|0x80000001) / alignbit / Montgomery product /
// plain add / plain mul_lo (the first three also issue one v_or_b32 per
// operation).
// kind 10 + w (w = 1..4): cycles of one CU per wave-wide load of w dwords per
// lane; param = lane stride in bytes (8: contiguous span, 4160: one line per
// lane).  *out receives the figure.
extern "C" int rbg_microbench(int device, int kind, uint32_t param, double *out) {
    if (!out || hipSetDevice(device) != hipSuccess) return RBG_EDEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return RBG_EDEVICE;
    const double cus = prop.multiProcessorCount, clock_hz = prop.clockRate * 1e3;
    uint32_t *sink = nullptr;
    if (hipMalloc(&sink, 64) != hipSuccess) return RBG_EDEVICE;
    int rc = RBG_OK;
    const uint32_t blocks = (uint32_t)cus * 8u;  // 8 waves per SIMD
    if (kind >= 0 && kind <= 6) {
        const uint32_t iters = 4096, p = 4294967291u;
        uint32_t ninv = 1;
        for (int i = 0; i < 5; ++i) ninv *= 2u - p * ninv;
        ninv = 0u - ninv;
        double ms = -1.0;
        switch (kind) {
#define RBG_VALU_CASE(K)                                                                         \
    case K:                                                                                      \
        ms = time_ms([&] {                                                                       \
            hipLaunchKernelGGL(rbg::valu_rate<K>, dim3(blocks), dim3(256), 0, 0, 12345u, iters, \
                               p, ninv, sink);                                                   \
        });                                                                                      \
        break;
            RBG_VALU_CASE(0)
            RBG_VALU_CASE(1)
            RBG_VALU_CASE(2)
            RBG_VALU_CASE(3)
            RBG_VALU_CASE(4)
            RBG_VALU_CASE(5)
            RBG_VALU_CASE(6)
#undef RBG_VALU_CASE
        }
        if (ms <= 0) rc = RBG_EDEVICE;
        else *out = (double)blocks * 256.0 * iters * rbg::kChains / (ms * 1e-3);
    } else if (kind >= 11 && kind <= 14 &&
               (param == 4u || param == 8u || param == 12u || param == 4160u)) {
        const uint32_t bytes = 1u << 20, iters = 2000;
        uint32_t *buf = nullptr;
        if (hipMalloc(&buf, bytes + 64) != hipSuccess || hipMemset(buf, 1, bytes + 64) != hipSuccess) {
            rc = RBG_EDEVICE;
        } else {
            const int nw = kind - 10;
            double ms = -1.0;
            switch (nw) {
#define RBG_LOAD_CASE(W)                                                                         \
    case W:                                                                                      \
        ms = time_ms([&] {                                                                       \
            hipLaunchKernelGGL(rbg::load_rate<W>, dim3(blocks), dim3(256), 0, 0, buf, bytes - 1, \
                               param, iters, sink);                                              \
        });                                                                                      \
        break;
                RBG_LOAD_CASE(1)
                RBG_LOAD_CASE(2)
                RBG_LOAD_CASE(3)
                RBG_LOAD_CASE(4)
#undef RBG_LOAD_CASE
            }
            if (ms <= 0) rc = RBG_EDEVICE;
            else {
                const double wave_loads_per_cu = (double)blocks * 4.0 * iters / cus;
                *out = ms * 1e-3 * clock_hz / wave_loads_per_cu;
            }
        }
        if (buf) (void)hipFree(buf);
    } else {
        rc = RBG_EINVAL;
    }
    (void)hipFree(sink);
    return rc;
}

// Device check of the 32-bit residue arithmetic (a test hook: the dense
// scan's Montgomery product, power and position recurrence on caller
// operands, host arrays).
extern "C" int rbg_check_mont32(int device, const uint32_t *a, const uint32_t *b, uint32_t n,
                                uint32_t p, int kind, uint32_t *out) {
    if (!a || !b || !out || !n || p < 3 || !(p & 1) || kind < 0 || kind > 2) return RBG_EINVAL;
    for (uint32_t i = 0; i < n; ++i)
        if (a[i] >= p || (kind == 0 && b[i] >= p) || (kind == 2 && b[i] > 4096)) return RBG_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return RBG_EDEVICE;
    const uint32_t ninv = rbg::neg_inverse32(p);
    const uint32_t one = (uint32_t)((1ull << 32) % p);
    const uint32_t r2 = (uint32_t)(((unsigned __int128)1 << 64) % p);
    uint32_t *d = nullptr;
    if (hipMalloc(&d, 3ull * n * 4) != hipSuccess) return RBG_EDEVICE;
    int rc = RBG_OK;
    if (hipMemcpy(d, a, n * 4ull, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, b, n * 4ull, hipMemcpyHostToDevice) != hipSuccess) {
        rc = RBG_EDEVICE;
    } else {
        hipLaunchKernelGGL(rbg::mont32_check, dim3((n + 255) / 256), dim3(256), 0, 0, d, d + n, n,
                           p, ninv, r2, one, kind, d + 2ull * n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, d + 2ull * n, n * 4ull, hipMemcpyDeviceToHost) != hipSuccess)
            rc = RBG_EDEVICE;
    }
    (void)hipFree(d);
    return rc;
}

// Device check of the wide residue arithmetic (a test hook: the kernels'
// 33-bit Montgomery product and power on caller operands, host arrays).
extern "C" int rbg_check_wide(int device, const uint64_t *a, const uint64_t *b, uint32_t n,
                              uint64_t p, int kind, uint64_t *out) {
    if (!a || !b || !out || !n || p <= (1ull << 32) || p >= (1ull << 33) || !(p & 1) ||
        (kind != 0 && kind != 1))
        return RBG_EINVAL;
    for (uint32_t i = 0; i < n; ++i)
        if (a[i] >= p || (kind == 0 && b[i] >= p)) return RBG_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return RBG_EDEVICE;
    uint32_t ninv = (uint32_t)p;  // -p^{-1} mod 2^32 (Newton)
    for (int i = 0; i < 5; ++i) ninv *= 2u - (uint32_t)p * ninv;
    ninv = 0u - ninv;
    uint64_t *d = nullptr;
    if (hipMalloc(&d, 3ull * n * 8) != hipSuccess) return RBG_EDEVICE;
    int rc = RBG_OK;
    if (hipMemcpy(d, a, n * 8ull, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, b, n * 8ull, hipMemcpyHostToDevice) != hipSuccess) {
        rc = RBG_EDEVICE;
    } else {
        hipLaunchKernelGGL(rbg::wide_check, dim3((n + 255) / 256), dim3(256), 0, 0, d, d + n, n, p,
                           ninv, 1ull << 32, kind, d + 2ull * n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, d + 2ull * n, n * 8ull, hipMemcpyDeviceToHost) != hipSuccess)
            rc = RBG_EDEVICE;
    }
    (void)hipFree(d);
    return rc;
}
