// Device/host shared layouts for the census dense path.
#pragma once
#include <stdint.h>

#include "rbg_arith.hpp"

namespace rbg {

typedef uint16_t u16;

// Largest lcm of admissible orders in 2..9.
constexpr u32 kMaxM = 2520;

// Colour code of a label j (mod m): a bijective storage form in which the
// colour of every admissible order r is a fixed set of code bits, so equality
// of colours under r is equality of those bits.
//   bits 0-2  : j mod 2^a  (2^a = 2-part of m: 2, 4 or 8)
//   bits 3-4  : j mod 3                      (when 3 | m)
//   bits 5-6  : (j mod 9) div 3              (when 9 | m)
//   bits 7-9  : j mod 5                      (when 5 | m)
//   bits 10-12: j mod 7                      (when 7 | m)
// r=2 -> bit 0, r=4 -> bits 0-1, r=8 -> bits 0-2, r=3 -> bits 3-4,
// r=9 -> bits 3-6, r=6 -> bits 0,3,4, r=5 -> bits 7-9, r=7 -> bits 10-12.
RBG_HD u32 plane_bits(u32 r) {
    switch (r) {
    case 2: return 0x001u;
    case 3: return 0x018u;
    case 4: return 0x003u;
    case 5: return 0x380u;
    case 6: return 0x019u;
    case 7: return 0x1C00u;
    case 8: return 0x007u;
    case 9: return 0x078u;
    default: return 0u;
    }
}

RBG_HD u32 colour_code(u32 j, u32 m) {
    u32 two = (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : 1;
    u32 c = j % two;
    if (m % 3 == 0) c |= (j % 3) << 3;
    if (m % 9 == 0) c |= ((j % 9) / 3) << 5;
    if (m % 5 == 0) c |= (j % 5) << 7;
    if (m % 7 == 0) c |= (j % 7) << 10;
    return c;
}

// Per-prime constants, one record per prime of the block.
// One prime's part of a batched label-reuse launch (plain method): blocks
// [block0, next piece's block0) of the launch work for prime pi, the scan
// tiles from tile_lo (seed launches: positions from 1), its labels at
// lab + lab_off, labels j2 = j(2), j3 = j(3).
struct ReusePiece {
    u32 block0, pi, tile_lo, j2, j3, pad;
    u64 lab_off;
};

struct PrimeDesc {
    u32 p;          // the prime
    u32 nh;         // (p-1)/2, length of the half interval scanned
    u32 e;          // (p-1)/m
    u32 ninv;       // -p^{-1} mod 2^32
    u32 one;        // R mod p (Montgomery form of 1)
    u32 r2;         // R^2 mod p
    u32 step;       // Montgomery form of the per-chunk stride (64)
    u32 mask;       // admissible orders, bit r-2
    u32 m;          // lcm of admissible orders
    u32 tab_off;    // offset of the root table in the table arrays
    u32 planes;     // union of plane bits used by the admissible orders
    u32 tile_begin; // first tile of this prime in the flattened tile space
    u32 ntiles;     // tiles covering [1, nh]
    u32 sum_off;    // first RunSum slot of this prime
    u32 nr;         // number of admissible orders
    u32 index;      // position in the caller's block
    u32 hi;         // 2^32 < p < 2^33: bit 32 of p, one, r2, step (bits 0-3);
                    // p, one, r2 and step above hold the low words
    u32 tab64;      // 2^32 < p: offset of the root-table values in the u64
                    // value array (codes stay at tab_off)
    u32 floors[2];  // run floor of order r in byte r-2 (bytes 0-3 of [0],
                    // 4-7 of [1]): runs shorter than the floor are not
                    // tracked inside a segment (1: every run)
};

// Run floor of order k + 2.
RBG_HD u32 run_floor(const PrimeDesc &d, u32 k) { return (d.floors[k >> 2] >> (8u * (k & 3u))) & 0xFFu; }

// Per (tile, admissible order) run summary of a contiguous stretch of the
// half interval, the associative run monoid of the reference's dense scans
// (original_scan_gpu.cu:116-128 SummOp; dense_oracle.cu:7,32-44 Summary with
// least-start witness).  Positions are absolute (1-based).
struct RunSum {
    u32 prefix;  // length of the run starting at the first position
    u32 suffix;  // length of the run ending at the last position
    u32 best;    // longest run inside (partial end runs included)
    u32 bstart;  // least start among runs of length `best`
    u32 a4;      // windows of 4 equal colours inside (order 2 only)
};

struct TileHdr {
    u32 first_x;  // first position
    u32 len;      // positions covered (0 = empty)
    u32 codes;    // first colour code | last colour code << 16
    u32 pad;
};

}  // namespace rbg
