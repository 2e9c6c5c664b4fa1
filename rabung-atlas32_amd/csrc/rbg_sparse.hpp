// Layouts of the sparse engine (host and device).
#pragma once
#include "rbg_types.hpp"

namespace rbg {

// Per prime handled by the sparse engine.
struct SparsePrime {
    u32 p, e, ninv, one;     // e = (p-1)/m; Montgomery constants (R = 2^32)
    u32 r2, m, mask, index;  // R^2 mod p; lcm of admissible orders; mask; caller index
    u32 B, D, Q, jneg;       // certified cover; label of -1 (0 or m/2)
    u32 tab_off, lab_off, inv_off, prof_begin;
    u32 nprof, lblk, iblk, pblk;  // profiles; first label / inverse / plane block
    // 2^32 < p < 2^33: bit 32 of p, one, r2 (bits 0-2; the fields above hold
    // the low words), the offset of the root-table values in the u64 value
    // array (labels stay at tab_off), and inv_off counts in the u64 inverse
    // array
    u32 hi, tab64;
};

// Colour-plane construction: a block of 4 waves builds 64-bit steps
// (two plane words) for this many consecutive steps of one profile.
constexpr u32 kPlanePairsPerBlock = 256;

// Per (prime, admissible order).
struct SparseProf {
    u32 prime;      // index into the SparsePrime array
    u32 r;          // order
    u32 np;         // colour bit planes (bits of r-1)
    u32 wpp;        // 32-bit words per plane of the full row (row 0)
    u32 plane_off;  // first word of the profile's rows, in units of 16 words
    u32 prefix;     // initial run of the identity colour (seed kernel)
    u32 sigma0;     // r | (p-1)/2
    u32 pblk;       // first plane block
    u32 qspec;      // speculative threshold (0: none), at most Q
    // copies of the prime's cover, threshold cap and inverse offset, so that
    // a filter task's setup needs no second dependent load
    u32 B, D, Q, inv_off;
};

// Colour rows of one profile.  Row 0 is the full signed table: bit i of
// plane b is bit b of the colour of the numerator i - B, i in [0, 2B].
// Rows 1-5 decimate it by the wheel: row (g, c) holds the numerators
// a = c (mod g), c coprime to g, in increasing order -- bit k is the full
// row's bit rho + g k with rho = (c + B) mod g.  A denominator d with
// gcd(d, 6) = g > 1 tests only anchors b coprime to g (the others share a
// factor with d), and b - j d stays in b's class, so a word of 96
// consecutive bits of one row holds 96 anchors, all live, and every stage
// window is 96 consecutive bits of the same row.  Each row is padded by 256
// bits (a four-dword load at its last bit stays inside) and every row
// starts on a 16-word boundary; the two rows of one g have equal sizes, so
// the second class lies a fixed number of bits after the first.
constexpr u32 kRows = 6;

RBG_HD u32 row_g(u32 row) { return row == 0 ? 1u : row == 1 ? 2u : row <= 3 ? 3u : 6u; }

RBG_HD u32 row_class(u32 row) { return row == 3 ? 2u : row == 5 ? 5u : row == 0 ? 0u : 1u; }

RBG_HD u32 row_rho(u32 row, u32 B) { return (row_class(row) + B) % row_g(row); }

// class numerators in [-B, B]: row bits that carry colours
RBG_HD u32 row_bits(u32 row, u32 B) { return (2u * B - row_rho(row, B)) / row_g(row) + 1u; }

// 32-bit words of one plane of a row of decimation g (a multiple of 16)
RBG_HD u32 row_words(u32 g, u32 B) { return ((2u * B / g + 1u + 256u + 31u) / 32u + 15u) & ~15u; }

// words from a profile's first word to plane 0 of `row`
RBG_HD u32 row_offset(u32 row, u32 B, u32 np) {
    u32 off = 0;
    for (u32 r = 0; r < row; ++r) off += np * row_words(row_g(r), B);
    return off;
}

// blocks of the plane kernel for one row (kPlanePairsPerBlock 64-bit steps each)
RBG_HD u32 row_blocks(u32 row, u32 B) {
    return (row_words(row_g(row), B) / 2u + kPlanePairsPerBlock - 1u) / kPlanePairsPerBlock;
}

// first row of decimation g and the number of its classes
RBG_HD u32 first_row_of(u32 g) { return g == 1 ? 0u : g == 2 ? 1u : g == 3 ? 2u : 4u; }

}  // namespace rbg
