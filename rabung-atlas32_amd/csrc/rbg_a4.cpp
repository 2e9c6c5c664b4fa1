// A4 four-window counts through an elliptic point count (host side).
//
// For odd p, A4 = #{x in 1..p-4 : x..x+3 share the quadratic character}.
// Character-sum identity used by the reference (src/census/theorie/
// elliptic4_mestre.cpp:64-67, quadruples):
//     A4 = (p - 10 - Bd + S) / 8,   S = N - p - 2,
//     Bd = 4 chi(2) + 2 chi(6) + 2 chi(3) + 2 chi(-1) + 2 chi(-2),
// where N = #E(F_p) for E : y^2 = x (x - 1)(x - 9).  N is found by Mestre's
// method: random points on E and on its quadratic twist E' (y^2 = x(x-s)(x-9s)
// with s a non-residue, #E + #E' = 2p + 2) give point orders by baby-step
// giant-step over the Hasse interval; the lcms of those orders are refined
// until exactly one N in the interval is consistent with both curves.
// For p < 1000 the count is dense.  The dense GPU scan computes A4 directly
// from its order-2 run summaries; this path serves the sparse engine, which
// never visits every position.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rabung.h"

namespace {

typedef uint64_t u64;
typedef uint32_t u32;

struct Field {
    u64 p;
    u64 mul(u64 a, u64 b) const { return (u64)((unsigned __int128)a * b % p); }  // p < 2^33
    u64 add(u64 a, u64 b) const { u64 s = a + b; return s >= p ? s - p : s; }
    u64 sub(u64 a, u64 b) const { return a >= b ? a - b : a + p - b; }
    u64 pow(u64 a, u64 e) const {
        u64 v = 1 % p;
        a %= p;
        while (e) {
            if (e & 1) v = mul(v, a);
            a = mul(a, a);
            e >>= 1;
        }
        return v;
    }
    u64 inv(u64 a) const {  // extended Euclid, a != 0
        int64_t t = 0, nt = 1;
        int64_t r = (int64_t)p, nr = (int64_t)a;
        while (nr) {
            int64_t q = r / nr, tmp;
            tmp = t - q * nt; t = nt; nt = tmp;
            tmp = r - q * nr; r = nr; nr = tmp;
        }
        if (t < 0) t += (int64_t)p;
        return (u64)t;
    }
    int chi(u64 a) const {  // Legendre symbol
        a %= p;
        if (!a) return 0;
        return pow(a, (p - 1) / 2) == 1 ? 1 : -1;
    }
    bool sqrt(u64 a, u64 &root) const {  // Tonelli-Shanks
        a %= p;
        if (!a) { root = 0; return true; }
        if (chi(a) != 1) return false;
        if (p % 4 == 3) { root = pow(a, (p + 1) / 4); return true; }
        u64 q = p - 1, s = 0;
        while (!(q & 1)) { q >>= 1; ++s; }
        u64 z = 2;
        while (chi(z) != -1) ++z;
        u64 c = pow(z, q), x = pow(a, (q + 1) / 2), t = pow(a, q), m = s;
        while (t != 1) {
            u64 i = 1, t2 = mul(t, t);
            while (t2 != 1) { t2 = mul(t2, t2); ++i; }
            u64 b = pow(c, 1ull << (m - i - 1));
            x = mul(x, b);
            c = mul(b, b);
            t = mul(t, c);
            m = i;
        }
        root = x;
        return true;
    }
};

struct Pt {
    u64 x, y;
    bool inf;
    bool operator==(const Pt &o) const { return inf == o.inf && (inf || (x == o.x && y == o.y)); }
};

// y^2 = x^3 + a2 x^2 + a4 x with a2 = -10 s, a4 = 9 s^2  (= x (x - s)(x - 9 s))
struct Curve {
    Field F;
    u64 a2, a4;
    Curve(u64 p, u64 s) : F{p} {
        a2 = F.sub(0, F.mul(10 % p, s));
        a4 = F.mul(9 % p, F.mul(s, s));
    }
    u64 rhs(u64 x) const { return F.mul(x, F.add(F.mul(x, F.add(x, a2)), a4)); }
    Pt neg(Pt P) const {
        if (!P.inf) P.y = F.sub(0, P.y);
        return P;
    }
    Pt add(const Pt &P, const Pt &Q) const {
        if (P.inf) return Q;
        if (Q.inf) return P;
        u64 lam;
        if (P.x == Q.x) {
            if (F.add(P.y, Q.y) == 0) return Pt{0, 0, true};
            u64 num = F.add(F.add(F.mul(3, F.mul(P.x, P.x)), F.mul(F.add(a2, a2), P.x)), a4);
            lam = F.mul(num, F.inv(F.add(P.y, P.y)));
        } else {
            lam = F.mul(F.sub(Q.y, P.y), F.inv(F.sub(Q.x, P.x)));
        }
        u64 x3 = F.sub(F.sub(F.sub(F.mul(lam, lam), a2), P.x), Q.x);
        u64 y3 = F.sub(F.mul(lam, F.sub(P.x, x3)), P.y);
        return Pt{x3, y3, false};
    }
    Pt mul(Pt P, u64 n) const {
        Pt R{0, 0, true};
        while (n) {
            if (n & 1) R = add(R, P);
            P = add(P, P);
            n >>= 1;
        }
        return R;
    }
};

struct PtHash {
    size_t operator()(const Pt &a) const {
        return a.inf ? 0 : (size_t)(a.x * 0x9E3779B97F4A7C15ull ^ (a.y + 0x632BE59BD9B4E019ull));
    }
};

u64 isqrt64(u64 n) {
    u64 x = (u64)std::sqrt((long double)n);
    while (x * x > n) --x;
    while ((x + 1) * (x + 1) <= n) ++x;
    return x;
}

std::vector<u64> prime_factors(u64 n) {
    std::vector<u64> f;
    for (u64 q = 2; q * q <= n; ++q)
        if (n % q == 0) {
            f.push_back(q);
            while (n % q == 0) n /= q;
        }
    if (n > 1) f.push_back(n);
    return f;
}

u64 splitmix(u64 &s) {
    u64 z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

u64 gcd64(u64 a, u64 b) {
    while (b) { u64 t = a % b; a = b; b = t; }
    return a;
}

// A point drawn on one curve and its exact order.
struct Witness {
    u64 x, y, order;
};

// Evidence of one point count: the Hasse window [lower, upper], the lcm of the
// orders of the random points found on the curve and on its quadratic twist
// (parameter `twist`, a non-residue), and the number of points drawn.  In
// certificate mode also the curve whose own lcm fixes the count (`cert_s`:
// 1 for E, the twist parameter for E') and that curve's points.
struct CountInfo {
    u64 lower = 0, upper = 0, lcm_curve = 1, lcm_twist = 1, attempts = 0, twist = 0;
    u64 cert_s = 0, cert_lcm = 0;
    std::vector<Witness> cert_points;
};

// The one multiple of l in [L, H], or 0 when there are none or several.
u64 unique_multiple(u64 l, u64 L, u64 H) {
    const u64 first = (L + l - 1) / l * l;
    return first <= H && first + l > H ? first : 0;
}

// #E(F_p) for E: y^2 = x(x-1)(x-9); returns 0 on failure.  Points are drawn
// alternately on E and E'.  The count is settled as soon as the two lcms
// leave one N in the Hasse window with lcm(E) | N and lcm(E') | 2p + 2 - N.
// With `certify` the draw goes on until one curve's lcm alone has a single
// multiple in the window (Mestre: for p > 229 one of the two curves has such
// points), so that the points of that curve certify N by themselves; that
// N must agree with the other curve's lcm too.
u64 curve_order(u64 p, CountInfo *info = nullptr, bool certify = false) {
    const u64 R = isqrt64(4 * p), L = p + 1 - R, H = p + 1 + R, W = H - L;
    u64 m = isqrt64(W + 1);
    if (m * m < W + 1) ++m;
    Field F{p};
    u64 rng = p * 0x2545F4914F6CDD1Dull + 20261203ull;
    u64 s = 2;
    while (F.chi(s) != -1) ++s;
    Curve curves[2] = {Curve(p, 1), Curve(p, s)};
    u64 order_lcm[2] = {1, 1};
    std::vector<Witness> points[2];
    for (int attempt = 0; attempt < 400; ++attempt) {
        const int which = attempt & 1;
        const Curve &E = curves[which];
        Pt P{0, 0, true};
        for (int tries = 0; tries < 1000 && P.inf; ++tries) {
            u64 x = splitmix(rng) % p, y;
            u64 f = E.rhs(x);
            if (f == 0 || !F.sqrt(f, y)) continue;
            P = Pt{x, (splitmix(rng) & 1) ? F.sub(0, y) : y, false};
        }
        if (P.inf) return 0;
        // baby steps j*P, j in [0, m)
        std::unordered_map<Pt, u64, PtHash> baby;
        baby.reserve(2 * m);
        Pt J{0, 0, true};
        for (u64 j = 0; j < m; ++j) {
            baby.emplace(J, j);
            J = E.add(J, P);
        }
        // giant steps: find n = L + i m + j with n P = O  <=>  j P = -(L + i m) P
        const Pt G = E.neg(E.mul(P, m));
        Pt Qi = E.neg(E.mul(P, L));
        u64 n = 0;
        for (u64 i = 0; i <= W / m + 1; ++i) {
            auto it = baby.find(Qi);
            if (it != baby.end() && i * m + it->second <= W) {
                n = L + i * m + it->second;
                break;
            }
            Qi = E.add(Qi, G);
        }
        if (!n || !E.mul(P, n).inf) return 0;
        u64 ord = n;  // exact order of P
        for (u64 q : prime_factors(n))
            while (ord % q == 0 && E.mul(P, ord / q).inf) ord /= q;
        order_lcm[which] = order_lcm[which] / gcd64(order_lcm[which], ord) * ord;
        if (certify) points[which].push_back({P.x, P.y, ord});
        const u64 l0 = order_lcm[0], l1 = order_lcm[1];
        if (certify) {
            const u64 own = unique_multiple(order_lcm[which], L, H);
            if (!own) continue;
            const u64 N = which ? 2 * p + 2 - own : own;
            if (N % l0 || (2 * p + 2 - N) % l1) return 0;  // the curves disagree
            if (info) {
                info->lower = L;
                info->upper = H;
                info->lcm_curve = l0;
                info->lcm_twist = l1;
                info->attempts = (u64)attempt + 1;
                info->twist = s;
                info->cert_s = which ? s : 1;
                info->cert_lcm = order_lcm[which];
                info->cert_points = points[which];
            }
            return N;
        }
        // candidates N in [L, H]: lcm0 | N and lcm1 | (2p + 2 - N)
        u64 found = 0, count = 0;
        for (u64 N = (L + l0 - 1) / l0 * l0; N <= H && count < 2; N += l0)
            if ((2 * p + 2 - N) % l1 == 0) {
                found = N;
                ++count;
            }
        if (count == 1) {
            if (info) {
                info->lower = L;
                info->upper = H;
                info->lcm_curve = l0;
                info->lcm_twist = l1;
                info->attempts = (u64)attempt + 1;
                info->twist = s;
            }
            return found;
        }
    }
    return 0;
}

u32 dense_a4(u32 p) {
    if (p <= 3) return 0;
    std::vector<unsigned char> sq(p, 0);
    for (u64 x = 1; x < p; ++x) sq[x * x % p] = 1;
    u32 n = 0;
    for (u32 x = 1; x + 3 < p; ++x)
        n += (sq[x] == sq[x + 1] && sq[x] == sq[x + 2] && sq[x] == sq[x + 3]);
    return n;
}

// 0 on success; a4 = UINT32_MAX on failure.
int a4_one(u64 p, u32 *a4, u64 *points = nullptr, CountInfo *info = nullptr,
           bool certify = false) {
    if (p == 2) { *a4 = 0; return 0; }
    if (p < 1000) { *a4 = dense_a4((u32)p); return 0; }
    const u64 N = curve_order(p, info, certify);
    if (points) *points = N;
    if (!N) return 1;
    Field F{p};
    const int64_t bd = 4 * F.chi(2) + 2 * F.chi(6) + 2 * F.chi(3) + 2 * F.chi(p - 1) +
                       2 * F.chi(p - 2);
    const int64_t S = (int64_t)N - (int64_t)p - 2;
    const int64_t num = (int64_t)p - 10 - bd + S;
    if (num < 0 || num % 8) return 1;
    const int64_t v = num / 8;
    if (v > (int64_t)(p - 3) / 2) return 1;
    *a4 = (u32)v;
    return 0;
}

template <class P>
int a4_many(const P *primes, uint32_t n, uint32_t *out, uint32_t threads) {
    if (!threads) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    threads = std::min<u32>(threads, std::max(1u, n));
    std::vector<int> rc(threads, 0);
    std::vector<std::thread> pool;
    for (u32 t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            for (u32 i = t; i < n; i += threads)
                if (a4_one(primes[i], &out[i])) rc[t] = RBG_ECERTIFY;
        });
    for (auto &th : pool) th.join();
    for (int r : rc)
        if (r) return r;
    return RBG_OK;
}

}  // namespace

extern "C" int rbg_a4_host(const uint32_t *primes, uint32_t n, uint32_t *out, uint32_t threads) {
    if ((!primes || !out) && n) return RBG_EINVAL;
    for (u32 i = 0; i < n; ++i)
        if (!rbg_is_prime(primes[i])) return RBG_EINVAL;
    return a4_many(primes, n, out, threads);
}

extern "C" int rbg_a4_host_wide(const uint64_t *primes, uint32_t n, uint32_t *out,
                                uint32_t threads) {
    if ((!primes || !out) && n) return RBG_EINVAL;
    for (u32 i = 0; i < n; ++i)
        if (!rbg_is_prime_wide(primes[i])) return RBG_EINVAL;
    return a4_many(primes, n, out, threads);
}

extern "C" int rbg_a4_certificate(uint32_t p, uint32_t *a4, uint64_t *attempts, char *out,
                                  size_t *out_len) {
    if (!a4 || !rbg_is_prime(p)) return RBG_EINVAL;
    u64 N = 0;
    CountInfo info;
    const bool certify = out_len != nullptr;
    if (a4_one(p, a4, &N, &info, certify)) return RBG_ECERTIFY;
    if (attempts) *attempts = info.attempts;
    if (!certify) return RBG_OK;
    std::string text;
    char buf[512];
    if (p < 1000) {
        std::snprintf(buf, sizeof buf,
                      "{\"status\":\"PASS\",\"p\":%u,\"r\":2,\"quadruples\":%u,"
                      "\"certificate\":\"small_dense\"}",
                      p, *a4);
        text = buf;
    } else {
        // field names and meaning of the reference's certificate
        // (elliptic4_mestre.cpp:68-73): the points of the curve y^2 =
        // x (x - s)(x - 9 s) whose orders' lcm has one multiple in the
        // Hasse window, each with its order and the order's prime factors
        std::snprintf(buf, sizeof buf,
                      "{\"status\":\"PASS\",\"p\":%u,\"r\":2,\"quadruples\":%u,"
                      "\"elliptic_points\":%llu,\"order_lcm\":%llu,\"hasse_lower\":%llu,"
                      "\"hasse_upper\":%llu,\"attempts\":%llu,\"twist_parameter\":%llu,"
                      "\"certificate_points\":[",
                      p, *a4, (unsigned long long)N, (unsigned long long)info.cert_lcm,
                      (unsigned long long)info.lower, (unsigned long long)info.upper,
                      (unsigned long long)info.attempts, (unsigned long long)info.cert_s);
        text = buf;
        for (size_t i = 0; i < info.cert_points.size(); ++i) {
            const Witness &w = info.cert_points[i];
            std::snprintf(buf, sizeof buf, "%s{\"x\":%llu,\"y\":%llu,\"order\":%llu,\"prime_factors\":[",
                          i ? "," : "", (unsigned long long)w.x, (unsigned long long)w.y,
                          (unsigned long long)w.order);
            text += buf;
            const std::vector<u64> f = prime_factors(w.order);
            for (size_t k = 0; k < f.size(); ++k)
                text += (k ? "," : "") + std::to_string(f[k]);
            text += "]}";
        }
        text += "],\"certificate\":\"unique_lcm_multiple_in_Hasse\"}";
    }
    const size_t cap = *out_len;
    *out_len = text.size();
    if (!out || cap < text.size()) return RBG_ECAPACITY;
    std::memcpy(out, text.data(), text.size());
    return RBG_OK;
}
