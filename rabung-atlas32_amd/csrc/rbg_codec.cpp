// Pack-record frames of the census (host code): scan frames (M, L, witness per
// admissible order) and A4 frames, byte-identical to the reference's writer.
//
// Layout from the reference's docs/DATA_FORMAT.md:9-47 (decoder
// scripts/atlas_codec.py:20-60):
//   scan frame  = header '<8sQQQQII32s' (b"VDWSM1\0\0", compressed length,
//                 raw length, lo, hi, count, witness flag, sha256(raw)) +
//                 zlib(raw); raw = per prime varint(p - previous) (previous
//                 starts at lo-1), the mask byte, then for r = 2..9 in the
//                 mask the varints M, L[, witness];
//   A4 frame    = header '<8sQQIIII32s32s32s' (b"VDWA4V1\0", lo, hi, n,
//                 compressed length, raw length, flags 0, prime hash, profile
//                 science hash, sha256(raw)) + zlib(n x u32 LE);
//   prime hash  = sha256("vdw-primes-u64le" || p as u64 LE ...);
//   science hash= sha256("vdw-p-profile-v1-u64le" || per prime (p u64, mask
//                 u8, then per admissible r: M, L, K as u64)).
// The same guards as the decoder are applied before writing (ascending primes
// in [lo, hi), mask = admissible mask, 1 <= L <= M <= (p-1)/r, 1 <= witness
// <= (p-1)/2, K consistent); a record that fails them refuses the whole frame.
#include <openssl/evp.h>
#include <zlib.h>

#include <cstring>
#include <vector>

#include "rabung.h"

namespace {

using u8 = uint8_t;
using u32 = uint32_t;
using u64 = uint64_t;

void put_varint(std::vector<u8> &out, u64 v) {
    while (v >= 0x80) {
        out.push_back((u8)(v | 0x80));
        v >>= 7;
    }
    out.push_back((u8)v);
}

void put_le(std::vector<u8> &out, u64 v, int bytes) {
    for (int i = 0; i < bytes; ++i) out.push_back((u8)(v >> (8 * i)));
}

class Sha256 {
  public:
    Sha256() : ctx_(EVP_MD_CTX_new()) { ok_ = ctx_ && EVP_DigestInit_ex(ctx_, EVP_sha256(), nullptr); }
    ~Sha256() { EVP_MD_CTX_free(ctx_); }
    void update(const void *p, size_t n) { ok_ = ok_ && EVP_DigestUpdate(ctx_, p, n); }
    bool final(u8 out[32]) {
        unsigned len = 0;
        return ok_ && EVP_DigestFinal_ex(ctx_, out, &len) && len == 32;
    }

  private:
    EVP_MD_CTX *ctx_;
    bool ok_;
};

u32 admissible(u32 p) {
    if (p < 3) return 0;
    u32 m = 0;
    for (u32 r = 2; r <= 9; ++r)
        if ((p - 1) % r == 0) m |= 1u << (r - 2);
    return m;
}

u64 threshold(u32 p, u32 r, u64 m, u64 l) {
    const u64 bound = (((p - 1) / 2) % r == 0) ? 2 * l + 2 : l + 2;
    return m + 1 > bound ? m + 1 : bound;
}

// The decoder's guards for one record (witness checked only when written).
bool record_ok(const rbg_record &rec, bool witnesses) {
    if (rec.mask != admissible(rec.p)) return false;
    for (u32 r = 2; r <= 9; ++r) {
        if (!(rec.mask & (1u << (r - 2)))) continue;
        const u32 k = r - 2;
        if (!(1 <= rec.L[k] && rec.L[k] <= rec.M[k] && rec.M[k] <= (rec.p - 1) / r)) return false;
        if (rec.K[k] != threshold(rec.p, r, rec.M[k], rec.L[k])) return false;
        if (witnesses && !(1 <= rec.W[k] && rec.W[k] <= (rec.p - 1) / 2)) return false;
    }
    return true;
}

bool ordered(const rbg_record *recs, u32 n, u64 lo, u64 hi) {
    u64 prev = lo - 1;
    for (u32 i = 0; i < n; ++i) {
        if (recs[i].p <= prev || recs[i].p < lo || recs[i].p >= hi) return false;
        prev = recs[i].p;
    }
    return true;
}

bool deflate_into(const std::vector<u8> &raw, int level, std::vector<u8> &comp) {
    uLongf cap = compressBound((uLong)raw.size());
    comp.resize(cap);
    if (compress2(comp.data(), &cap, raw.data(), (uLong)raw.size(), level) != Z_OK) return false;
    comp.resize(cap);
    return true;
}

int emit(const std::vector<u8> &frame, u8 *out, size_t *out_len) {
    const size_t cap = *out_len;
    *out_len = frame.size();
    if (!out || cap < frame.size()) return RBG_ECAPACITY;
    std::memcpy(out, frame.data(), frame.size());
    return RBG_OK;
}

bool science_digest(const rbg_record *recs, u32 n, u8 out[32]) {
    Sha256 h;
    static const char tag[] = "vdw-p-profile-v1-u64le";
    h.update(tag, sizeof tag - 1);
    std::vector<u8> buf;
    for (u32 i = 0; i < n; ++i) {
        buf.clear();
        put_le(buf, recs[i].p, 8);
        buf.push_back((u8)recs[i].mask);
        for (u32 r = 2; r <= 9; ++r) {
            if (!(recs[i].mask & (1u << (r - 2)))) continue;
            put_le(buf, recs[i].M[r - 2], 8);
            put_le(buf, recs[i].L[r - 2], 8);
            put_le(buf, recs[i].K[r - 2], 8);
        }
        h.update(buf.data(), buf.size());
    }
    return h.final(out);
}

bool prime_digest(const rbg_record *recs, u32 n, u8 out[32]) {
    Sha256 h;
    static const char tag[] = "vdw-primes-u64le";
    h.update(tag, sizeof tag - 1);
    for (u32 i = 0; i < n; ++i) {
        u8 le[8];
        for (int b = 0; b < 8; ++b) le[b] = (u8)((u64)recs[i].p >> (8 * b));
        h.update(le, 8);
    }
    return h.final(out);
}

}  // namespace

extern "C" int rbg_encode_scan_frame(const rbg_record *recs, uint32_t n, uint64_t lo, uint64_t hi,
                                     int witnesses, int level, uint8_t *out, size_t *out_len) {
    if (!out_len || (n && !recs) || n > RBG_MAX_BLOCK || lo == 0 || hi <= lo || level < 0 ||
        level > 9 || !ordered(recs, n, lo, hi))
        return RBG_EINVAL;
    std::vector<u8> raw;
    u64 prev = lo - 1;
    for (u32 i = 0; i < n; ++i) {
        const rbg_record &rec = recs[i];
        if (!record_ok(rec, witnesses != 0)) return RBG_EINVAL;
        put_varint(raw, rec.p - prev);
        prev = rec.p;
        raw.push_back((u8)rec.mask);
        for (u32 r = 2; r <= 9; ++r) {
            if (!(rec.mask & (1u << (r - 2)))) continue;
            put_varint(raw, rec.M[r - 2]);
            put_varint(raw, rec.L[r - 2]);
            if (witnesses) put_varint(raw, rec.W[r - 2]);
        }
    }
    std::vector<u8> comp;
    u8 digest[32];
    Sha256 h;
    h.update(raw.data(), raw.size());
    if (!deflate_into(raw, level, comp) || !h.final(digest)) return RBG_EINVAL;
    std::vector<u8> frame;
    frame.reserve(80 + comp.size());
    static const u8 magic[8] = {'V', 'D', 'W', 'S', 'M', '1', 0, 0};
    frame.insert(frame.end(), magic, magic + 8);
    put_le(frame, comp.size(), 8);
    put_le(frame, raw.size(), 8);
    put_le(frame, lo, 8);
    put_le(frame, hi, 8);
    put_le(frame, n, 4);
    put_le(frame, witnesses ? 1 : 0, 4);
    frame.insert(frame.end(), digest, digest + 32);
    frame.insert(frame.end(), comp.begin(), comp.end());
    return emit(frame, out, out_len);
}

extern "C" int rbg_encode_a4_frame(const rbg_record *recs, uint32_t n, uint64_t lo, uint64_t hi,
                                   int level, uint8_t *out, size_t *out_len) {
    if (!out_len || (n && !recs) || n > RBG_MAX_BLOCK || lo == 0 || hi <= lo || level < 0 ||
        level > 9 || !ordered(recs, n, lo, hi))
        return RBG_EINVAL;
    for (u32 i = 0; i < n; ++i)
        if (!record_ok(recs[i], false) || recs[i].A4 > (recs[i].p < 3 ? 0u : (recs[i].p - 3) / 2))
            return RBG_EINVAL;
    std::vector<u8> raw;
    for (u32 i = 0; i < n; ++i) put_le(raw, recs[i].A4, 4);
    std::vector<u8> comp;
    u8 ph[32], sh[32], vh[32];
    Sha256 h;
    h.update(raw.data(), raw.size());
    if (!deflate_into(raw, level, comp) || !h.final(vh) || !prime_digest(recs, n, ph) ||
        !science_digest(recs, n, sh))
        return RBG_EINVAL;
    std::vector<u8> frame;
    frame.reserve(136 + comp.size());
    static const u8 magic[8] = {'V', 'D', 'W', 'A', '4', 'V', '1', 0};
    frame.insert(frame.end(), magic, magic + 8);
    put_le(frame, lo, 8);
    put_le(frame, hi, 8);
    put_le(frame, n, 4);
    put_le(frame, comp.size(), 4);
    put_le(frame, raw.size(), 4);
    put_le(frame, 0, 4);
    frame.insert(frame.end(), ph, ph + 32);
    frame.insert(frame.end(), sh, sh + 32);
    frame.insert(frame.end(), vh, vh + 32);
    frame.insert(frame.end(), comp.begin(), comp.end());
    return emit(frame, out, out_len);
}

extern "C" int rbg_science_hash(const rbg_record *recs, uint32_t n, uint8_t out[32]) {
    if ((n && !recs) || !out) return RBG_EINVAL;
    return science_digest(recs, n, out) ? RBG_OK : RBG_EINVAL;
}

extern "C" int rbg_prime_hash(const rbg_record *recs, uint32_t n, uint8_t out[32]) {
    if ((n && !recs) || !out) return RBG_EINVAL;
    return prime_digest(recs, n, out) ? RBG_OK : RBG_EINVAL;
}

extern "C" int rbg_values_hash(const rbg_record *recs, uint32_t n, uint8_t out[32]) {
    if ((n && !recs) || !out) return RBG_EINVAL;
    std::vector<u8> raw;
    raw.reserve(4ull * n);
    for (u32 i = 0; i < n; ++i) put_le(raw, recs[i].A4, 4);
    Sha256 h;
    h.update(raw.data(), raw.size());
    return h.final(out) ? RBG_OK : RBG_EINVAL;
}

namespace {
// SHA-256 over the given 32-byte digests followed by u64 lo, u64 hi (LE).
int chain_step(const u8 *const *parts, int nparts, u64 lo, u64 hi, u8 out[32]) {
    std::vector<u8> msg;
    for (int i = 0; i < nparts; ++i) {
        if (!parts[i]) return RBG_EINVAL;
        msg.insert(msg.end(), parts[i], parts[i] + 32);
    }
    put_le(msg, lo, 8);
    put_le(msg, hi, 8);
    Sha256 h;
    h.update(msg.data(), msg.size());
    return out && h.final(out) ? RBG_OK : RBG_EINVAL;
}
}  // namespace

extern "C" int rbg_chain_scan(const uint8_t prev[32], const uint8_t prime_hash[32],
                              const uint8_t profile_hash[32], uint64_t lo, uint64_t hi,
                              uint8_t out[32]) {
    const u8 *parts[3] = {prev, prime_hash, profile_hash};
    return chain_step(parts, 3, lo, hi, out);
}

extern "C" int rbg_chain_a4(const uint8_t prev[32], const uint8_t prime_hash[32],
                            const uint8_t profile_hash[32], const uint8_t values_hash[32],
                            uint64_t lo, uint64_t hi, uint8_t out[32]) {
    const u8 *parts[4] = {prev, prime_hash, profile_hash, values_hash};
    return chain_step(parts, 4, lo, hi, out);
}
