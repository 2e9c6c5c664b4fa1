// Census worker front-end over librabung.so, speaking the reference's two
// process protocols so external drivers see the same worker.
//
//   rbg_worker profile [--method batched|plain|dense] [--max-batch N]
//                      [--no-certify] [--device D] [--domain-bits 32|33]
//                      [--floor F] [--with-a4] [--hbm-budget BYTES]
//     stdin : "<id> p1 ... pk" lines (ids consecutive from 0, 1 <= k <= N,
//             default N = 128, primes F <= p < 2^32 with F = 1000 as in
//             src/census/worker32.cu:24 -- `--floor 2` accepts the small
//             primes that the CPU worker src/census/small_cpu.cpp takes;
//             with --domain-bits 33 primes 1000 <= p < 2^33, the
//             reference's wider worker,
//             reproducibility/gpu_method/source/worker33.cu:24-41), then
//             "QUIT";
//     stdout: JSON lines {"type":"ready"}, one {"type":"measurement"} per
//             prime in input order (with --with-a4 an "A4" field is added),
//             {"type":"batch_done"} per batch, {"type":"bye"}.  Reference:
//             src/census/worker32.cu:10-45 and src/census/small_cpu.cpp:
//             34-40 (docs/REPRODUCTION.md:29-39).
//     A batch is profiled by one library call (all-or-nothing).  A capacity
//     refusal replays the same whole batch in halves, recursively, within a
//     30-second deadline (the reference's replay rule, GPU_CONFIG.json:
//     164-170, with the batch split in place of the epoch) and notes each
//     split on stderr; --hbm-budget lowers the library's workspace budget,
//     which makes large batches take that path.  Any other failure, or a
//     replay past the deadline, ends the process with
//     "VDW_CAMPAIGN_FAILURE: <reason>" on stderr and exit status 2, with no
//     output for the batch.
//
//   rbg_worker a4
//     stdin : "<id> <audit> <n> p1 ... pn" (n <= 4096 strictly increasing
//             primes < 2^32; audit = -1 or the index of the prime whose
//             certificate is wanted), then "QUIT";
//     stdout: JSON ready line, then per batch the binary record
//             '<8sQIIQQIIQQ' (b"A4BATCH1", id, n, certificate length,
//             nanoseconds, sum of A4, max points drawn, count of p < 1000,
//             first prime, last prime), n x u32 LE A4, certificate JSON.
//             Reference: src/census/moment_worker.cpp:19-52.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "rabung.h"

namespace {

using Clock = std::chrono::steady_clock;

[[noreturn]] void fail(const char *tag, const std::string &msg) {
    std::cerr << tag << ": " << msg << '\n';
    std::exit(2);
}

double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

bool parse_u64(const std::string &tok, uint64_t &v) {
    if (tok.empty() || tok.size() > 20 || tok.find_first_not_of("0123456789") != std::string::npos)
        return false;
    v = std::strtoull(tok.c_str(), nullptr, 10);
    return true;
}

struct ProfileOptions {
    uint32_t method = RBG_METHOD_BATCHED;
    uint32_t max_batch = 128;
    bool certify = true;
    int device = 0;
    int domain_bits = 32;
    uint64_t floor = 1000;
    bool with_a4 = false;
    uint64_t hbm_budget = 0;
};

// Profiles `primes` into `out`, replaying halves on a capacity refusal.
// (The wide entry point takes both domains; primes below 2^32 run the
// 32-bit kernels.)
int profile_replay(rbg_ctx *ctx, const uint64_t *primes, uint32_t n, uint32_t flags,
                   rbg_record_wide *out, Clock::time_point deadline, std::string &why) {
    const int rc = rbg_profile_block_wide(ctx, primes, n, flags, out);
    if (rc == RBG_OK) return RBG_OK;
    why = rbg_last_error(ctx);
    if (rc != RBG_ECAPACITY || n == 1 || Clock::now() > deadline) return rc;
    const uint32_t half = n / 2;
    std::cerr << "rbg_worker: capacity refusal of " << n << " primes, replayed as " << half
              << " + " << n - half << '\n';
    const int a = profile_replay(ctx, primes, half, flags, out, deadline, why);
    if (a) return a;
    return profile_replay(ctx, primes + half, n - half, flags, out + half, deadline, why);
}

int run_profile(const ProfileOptions &o) {
    const char *tag = "VDW_CAMPAIGN_FAILURE";
    rbg_config cfg{};
    cfg.device = o.device;
    cfg.hbm_budget = o.hbm_budget;
    rbg_ctx *ctx = rbg_open(&cfg);
    if (!ctx) fail(tag, "device context");
    std::cout << "{\"type\":\"ready\",\"domain_bits\":" << o.domain_bits << ",\"engine\":\""
              << rbg_version() << "\"}\n"
              << std::flush;
    const bool wide = o.domain_bits == 33;
    std::string line;
    uint64_t expected = 0;
    std::vector<rbg_record_wide> recs;
    while (std::getline(std::cin, line)) {
        if (line == "QUIT") {
            std::cout << "{\"type\":\"bye\"}\n" << std::flush;
            rbg_close(ctx);
            return 0;
        }
        std::istringstream in(line);
        std::string tok;
        uint64_t id = 0, v = 0;
        if (!(in >> tok) || !parse_u64(tok, id) || id != expected) fail(tag, "batch sequence");
        std::vector<uint64_t> primes;
        while (in >> tok) {
            if (!parse_u64(tok, v)) fail(tag, "prime token");
            if (wide ? (v < 1000 || v >= (1ull << 33)) : (v < o.floor || v >= (1ull << 32)))
                fail(tag, "worker prime range");
            if (!rbg_is_prime_wide(v)) fail(tag, "worker primality");
            primes.push_back(v);
        }
        if (primes.empty() || primes.size() > o.max_batch) fail(tag, "worker batch size");
        const auto start = Clock::now();
        recs.assign(primes.size(), rbg_record_wide{});
        std::string why;
        const uint32_t flags = o.method | (o.certify ? RBG_CERTIFY : 0u);
        const int rc = profile_replay(ctx, primes.data(), (uint32_t)primes.size(), flags,
                                      recs.data(), start + std::chrono::seconds(30), why);
        if (rc) fail(tag, "status " + std::to_string(rc) + ": " + why);
        const double wall = ms_since(start);
        std::ostringstream outs;
        outs.precision(12);
        for (const rbg_record_wide &r : recs) {
            outs << "{\"type\":\"measurement\",\"status\":\"PASS\",\"p\":" << r.p
                 << ",\"mask\":" << r.mask << ",\"total_ms\":" << wall / recs.size();
            if (o.with_a4) outs << ",\"A4\":" << r.A4;
            outs << ",\"profiles\":[";
            bool comma = false;
            for (uint32_t k = 0; k < 8; ++k) {
                if (!(r.mask & (1u << k))) continue;
                if (comma) outs << ',';
                comma = true;
                outs << "{\"r\":" << k + 2 << ",\"max\":" << r.M[k] << ",\"prefix\":" << r.L[k]
                     << ",\"K\":" << r.K[k] << ",\"witness\":" << r.W[k] << '}';
            }
            outs << "]}\n";
        }
        outs << "{\"type\":\"batch_done\",\"id\":" << id << ",\"count\":" << primes.size()
             << ",\"wall_ms\":" << wall << "}\n";
        std::cout << outs.str() << std::flush;
        ++expected;
    }
    rbg_close(ctx);
    return 0;
}

void put_le(std::string &out, uint64_t v, int bytes) {
    for (int i = 0; i < bytes; ++i) out.push_back((char)(v >> (8 * i)));
}

int run_a4() {
    const char *tag = "VDW_A4_FAILURE";
    std::ios::sync_with_stdio(false);
    std::cout << "{\"type\":\"ready\",\"algorithm\":\"mestre_bsgs_A4\",\"maximum_batch\":4096}\n"
              << std::flush;
    std::string line;
    while (std::getline(std::cin, line)) {
        if (line == "QUIT") return 0;
        std::istringstream in(line);
        std::string tok;
        uint64_t id = 0, n = 0, v = 0;
        long long audit = 0;
        if (!(in >> tok) || !parse_u64(tok, id) || !(in >> audit) || !(in >> tok) ||
            !parse_u64(tok, n) || n > RBG_MAX_BLOCK || audit < -1 || audit >= (long long)n)
            fail(tag, "batch header");
        std::vector<uint32_t> primes(n);
        uint64_t previous = 0;
        for (auto &p : primes) {
            if (!(in >> tok) || !parse_u64(tok, v) || v <= previous || v >= (1ull << 32) ||
                !rbg_is_prime((uint32_t)v))
                fail(tag, "ordered prime batch");
            p = (uint32_t)v;
            previous = v;
        }
        if (in >> tok) fail(tag, "request suffix");
        const auto begin = Clock::now();
        // one point count per prime on the host threads; each also reports
        // the points it drew (the header's maximum) and its certificate
        std::vector<uint32_t> values(n, 0);
        std::vector<uint64_t> attempts(n, 0);
        std::vector<int> bad(n, 0);
        std::string certificate;
        const unsigned threads =
            std::max(1u, std::min({16u, std::thread::hardware_concurrency(), (unsigned)n}));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < threads; ++t)
            pool.emplace_back([&, t] {
                for (uint64_t i = t; i < n; i += threads) {
                    if ((long long)i != audit) {
                        bad[i] = rbg_a4_certificate(primes[i], &values[i], &attempts[i],
                                                    nullptr, nullptr);
                        continue;
                    }
                    // the audited prime: ask for the length, then the text
                    size_t len = 0;
                    bad[i] = rbg_a4_certificate(primes[i], &values[i], &attempts[i], nullptr, &len);
                    if (bad[i] != RBG_ECAPACITY) continue;
                    std::string text(len, '\0');
                    bad[i] = rbg_a4_certificate(primes[i], &values[i], &attempts[i], &text[0], &len);
                    if (!bad[i]) certificate = text;
                }
            });
        for (auto &th : pool) th.join();
        uint64_t sum = 0, small = 0, max_attempts = 0;
        for (uint64_t i = 0; i < n; ++i) {
            if (bad[i]) fail(tag, "A4 point count");
            sum += values[i];
            small += primes[i] < 1000;
            max_attempts = std::max(max_attempts, attempts[i]);
        }
        const uint64_t ns = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                Clock::now() - begin)
                                .count();
        std::string out = "A4BATCH1";
        put_le(out, id, 8);
        put_le(out, n, 4);
        put_le(out, certificate.size(), 4);
        put_le(out, ns, 8);
        put_le(out, sum, 8);
        put_le(out, max_attempts, 4);
        put_le(out, small, 4);
        put_le(out, n ? primes.front() : 0, 8);
        put_le(out, n ? primes.back() : 0, 8);
        for (uint32_t a : values) put_le(out, a, 4);
        out += certificate;
        std::cout.write(out.data(), (std::streamsize)out.size());
        std::cout.flush();
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: rbg_worker profile [--method batched|plain|dense] [--max-batch N] "
                     "[--no-certify] [--device D] [--domain-bits 32|33] [--floor F] "
                     "[--with-a4] [--hbm-budget BYTES] | rbg_worker a4\n";
        return 2;
    }
    const std::string mode = argv[1];
    if (mode == "a4") return run_a4();
    if (mode != "profile") {
        std::cerr << "rbg_worker: unknown mode " << mode << '\n';
        return 2;
    }
    ProfileOptions o;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--method" && i + 1 < argc) {
            const std::string m = argv[++i];
            if (m == "batched") o.method = RBG_METHOD_BATCHED;
            else if (m == "plain") o.method = RBG_METHOD_PLAIN;
            else if (m == "dense") o.method = RBG_METHOD_DENSE;
            else fail("VDW_CAMPAIGN_FAILURE", "unknown method " + m);
        } else if (a == "--max-batch" && i + 1 < argc) {
            o.max_batch = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
            if (o.max_batch < 1 || o.max_batch > RBG_MAX_BLOCK)
                fail("VDW_CAMPAIGN_FAILURE", "max batch outside 1..4096");
        } else if (a == "--floor" && i + 1 < argc) {
            if (!parse_u64(argv[++i], o.floor) || o.floor < 2)
                fail("VDW_CAMPAIGN_FAILURE", "floor must be an integer >= 2");
        } else if (a == "--with-a4") {
            o.with_a4 = true;
        } else if (a == "--hbm-budget" && i + 1 < argc) {
            if (!parse_u64(argv[++i], o.hbm_budget) || o.hbm_budget == 0)
                fail("VDW_CAMPAIGN_FAILURE", "hbm budget must be a positive integer");
        } else if (a == "--no-certify") {
            o.certify = false;
        } else if (a == "--device" && i + 1 < argc) {
            o.device = std::atoi(argv[++i]);
        } else if (a == "--domain-bits" && i + 1 < argc) {
            o.domain_bits = std::atoi(argv[++i]);
            if (o.domain_bits != 32 && o.domain_bits != 33)
                fail("VDW_CAMPAIGN_FAILURE", "domain bits must be 32 or 33");
        } else {
            fail("VDW_CAMPAIGN_FAILURE", "unknown option " + a);
        }
    }
    return run_profile(o);
}
