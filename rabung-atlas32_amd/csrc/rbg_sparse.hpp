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
    u32 wpp;        // 32-bit words per plane
    u32 plane_off;  // first word of plane 0
    u32 prefix;     // initial run of the identity colour (seed kernel)
    u32 sigma0;     // r | (p-1)/2
    u32 pblk;       // first plane block
};

}  // namespace rbg
