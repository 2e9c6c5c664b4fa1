/*
 * rabung.h -- C ABI of the MI355X census-worker library (librabung.so).
 *
 * One call profiles a block of primes p < 2^32: for every admissible order
 * r in 2..9 with r | p-1 it returns the longest monochromatic run M, the
 * initial run L of the identity colour, the canonical reflected witness W and
 * the threshold K, plus the quadratic four-window count A4 (docs/METHODS.md
 * of the reference).  Plain pointers and sizes only; no exit() inside the
 * library; every failure is a nonzero status with nothing valid written.
 *
 * Reference interfaces replaced:
 *   - rbg_profile_block: the per-prime loop of the persistent GPU worker,
 *     src/census/worker32.cu:20-41 (prepare_gpu_table + gpu_prepared_group,
 *     src/census/engine32/src/frozen_entry.cu:149-242), and the dense CPU
 *     worker's emit(), src/census/small_cpu.cpp:15-33, extended to p < 2^32.
 *     The A4 value replaces the moment worker's per-prime loop,
 *     src/census/moment_worker.cpp:35-41.
 *   - baseline_init / baseline_profile: the reference's only in-process C ABI,
 *     src/census/engine32/src/baseline_bridge.cu:7,21-27 (same names,
 *     argument meaning and return codes; errors return nonzero instead of
 *     terminating the process).
 */
#ifndef RABUNG_H
#define RABUNG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBG_MAX_BLOCK 4096u /* primes per call: the pack block limit */

/* flags for rbg_profile_block */
#define RBG_METHOD_BATCHED 0x0u /* sparse anchor filter over the whole block in one
                                   persistent launch; dense scan for primes below
                                   2^16 and for any prime the sparse engine refuses */
#define RBG_METHOD_PLAIN 0x1u   /* dense scan, one grid launch per prime */
#define RBG_METHOD_DENSE 0x2u   /* dense scan, one persistent launch over all tiles */
#define RBG_METHOD_MASK 0xFu
#define RBG_CERTIFY 0x10u       /* direct powmod check of witness and prefix */
#define RBG_GRAPH 0x20u         /* record each launch chain (fills, tables, seed,
                                   filter; dense scans and merge) into a hipGraph
                                   by stream capture and launch the graph, updating
                                   the context's cached executable graph in place
                                   when only launch parameters change; timing then
                                   reports the sparse chain as the filter stage.  No
                                   reference counterpart: the reference's optional
                                   capture, frozen_entry.cu:232-233 */

/* status codes */
#define RBG_OK 0
#define RBG_EINVAL 1    /* bad argument: not a prime below 2^32, n too large */
#define RBG_EDEVICE 2   /* HIP runtime error */
#define RBG_ECERTIFY 3  /* a result failed its direct verification */
#define RBG_ECAPACITY 4 /* workspace beyond the configured HBM budget */

/* One prime's profile; arrays are indexed by r-2 and only entries whose bit
 * is set in `mask` are meaningful (others are zero). */
typedef struct rbg_record {
    uint32_t p;
    uint32_t mask; /* bit r-2 set iff r | p-1 (p=2: 0) */
    uint32_t M[8]; /* longest monochromatic run */
    uint32_t L[8]; /* initial run of the identity colour from x=1 */
    uint32_t W[8]; /* least reflected start min(s, p-e) among runs of length M */
    uint64_t K[8]; /* max(M+1, 2L+2) if r | (p-1)/2 else max(M+1, L+2) */
    uint32_t A4;   /* #{x <= p-4 : x..x+3 share the quadratic character}; p=2: 0 */
    uint32_t status;
} rbg_record;

/* The same record for the wider domain p < 2^33 (M, L, W, A4 still fit 32
 * bits: W <= (p-1)/2 and A4 <= (p-3)/2). */
typedef struct rbg_record_wide {
    uint64_t p;
    uint32_t mask;
    uint32_t M[8];
    uint32_t L[8];
    uint32_t W[8];
    uint64_t K[8];
    uint32_t A4;
    uint32_t status;
} rbg_record_wide;

typedef struct rbg_config {
    int32_t device;          /* HIP device ordinal */
    uint32_t tile_chunks;    /* dense scan: 64-position chunks per wave segment; 0 = 64
                                (the label-reuse scan rounds it up to a multiple of 12) */
    uint32_t blocks_per_cu;  /* persistent blocks (4 waves) per CU; 0 = 8 */
    uint64_t hbm_budget;     /* bytes of device workspace allowed; 0 = 16 GiB */
    uint32_t xcd_group;      /* sparse filter: profiles one XCD works on at a time; 0 = 1 */
    uint32_t cover_q;        /* sparse engine: threshold cap Q of the certified cover,
                                3..63; 0 = 48 (the reference's cover, frozen_entry.cu:
                                132-148, takes 3 <= Q <= 256; the bit-sliced screen here
                                counts up to q - t = 31) */
    uint32_t extend_limit;   /* sparse engine: most steps one passing anchor's run may be
                                extended by before the prime is refused and recomputed by
                                the dense scan; 0 = 2^22.  A small value forces the
                                refusal path (the reference's queue-overflow refusal,
                                gpu_pipeline.cuh:335,342,363) */
    uint32_t floor_boost;    /* dense scan: added to every run floor (runs shorter than
                                the floor are not tracked inside segments; a prime whose
                                longest run stays below it is rescanned exactly); 0 for
                                normal use.  A large value forces the rescan path */
    uint32_t cover_div;      /* sparse engine: cover shape n, D = floor(2 sqrt(p/H) / n)
                                (H = Q - 1), B the least radius certified for that D;
                                larger n trades a wider table for fewer denominators
                                (paper/GPU_METHOD_EN.tex:322-336; n = 2 is the
                                reference's D = floor(sqrt(p/H))).  2..8; 0 = 3.  Where B
                                would pass 2^20 - 2, D grows back toward the n = 2 value */
    uint32_t filter_chunk;   /* sparse filter: tasks one wave claims from its XCD's
                                queue at a time, 1..256; 0 = 16 for launches of more
                                than 10^4 tasks per wave, else by each profile's
                                cover radius B (4 for B >= 400000, 8 for B >= 100000,
                                else 16) */
    uint32_t spec_off;       /* sparse filter: nonzero starts every profile from its
                                seed threshold only; 0 also applies the speculative
                                threshold (the largest L with at least 14 runs of
                                length >= L expected over the half interval of a random
                                r-colouring), and a profile whose filter finds no run
                                reaching it is recomputed by the dense scan */
    uint32_t spec_boost;     /* added to every speculative threshold; 0 for normal use.
                                A large value makes every profile miss and forces the
                                dense recomputation */
} rbg_config;

typedef struct rbg_ctx rbg_ctx;

/* Create a context bound to cfg->device (cfg may be NULL for defaults).
 * Returns NULL on failure.  One context per host thread. */
rbg_ctx *rbg_open(const rbg_config *cfg);

/* Profile n <= RBG_MAX_BLOCK primes.  primes: host array; out: host array of n
 * records (caller-owned).  All-or-nothing: on a nonzero status no record is
 * valid.  Blocking. */
int rbg_profile_block(rbg_ctx *ctx, const uint32_t *primes, uint32_t n,
                      uint32_t flags, rbg_record *out);

/* Batching across blocks: n <= RBG_MAX_BATCH primes (e.g. many pack blocks
 * concatenated) profiled in as few launch chains as the HBM budget allows
 * (greedy sub-batches of at most 16384 primes, one persistent filter launch
 * each).  No reference counterpart: the reference processed 128-prime
 * batches per worker process (reproducibility/gpu_method/source/worker32.cu:
 * 37-49).  All-or-nothing like rbg_profile_block. */
#define RBG_MAX_BATCH (1u << 20)
int rbg_profile_batch(rbg_ctx *ctx, const uint32_t *primes, uint32_t n,
                      uint32_t flags, rbg_record *out);

/* Autotuner: times the method of `flags` on the sample primes (n <= 4096)
 * over launch shapes (batched: blocks per CU x XCD profile group; dense:
 * segment length x blocks per CU), checks that every shape gives identical
 * records, keeps the fastest in the context and copies it to *chosen.  No
 * reference counterpart (the reference hand-picked its launch constants,
 * src/census/GPU_CONFIG.json:38-59). */
int rbg_autotune(rbg_ctx *ctx, const uint32_t *primes, uint32_t n, uint32_t flags,
                 rbg_config *chosen);

/* Device workspace (bytes) that profiling these primes with `flags` needs;
 * serving loops use it for admission by HBM budget. */
int rbg_estimate_bytes(const rbg_ctx *ctx, const uint32_t *primes, uint32_t n,
                       uint32_t flags, uint64_t *bytes);

/* Same with device-resident input and output (d_primes: n uint32 in device
 * memory, d_out: n records in device memory).  Enqueues on the context's
 * stream and waits for completion.  Host-side preparation reads the primes
 * back once (4 bytes each). */
int rbg_profile_block_device(rbg_ctx *ctx, const uint32_t *d_primes, uint32_t n,
                             uint32_t flags, rbg_record *d_out);

/* rbg_profile_block_device without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * inputs, outputs and results; every device buffer beyond d_primes and d_out
 * is carved from a static array of 128 MiB in the library's code object,
 * which each GPU holds from the first context opened on it until the process
 * ends, and which keeps the last call's contents (the call clears what must
 * start at zero).  Calls on one GPU share that array: they serialise on a
 * per-device lock inside the process, so calls that overlap on one GPU wait
 * for each other (the call returns only when its work is complete).  Returns
 * RBG_ECAPACITY when the block needs more workspace than the array holds
 * (about 30 primes near 2^32 with the batched method), RBG_EINVAL for the
 * RBG_GRAPH flag; RBG_CERTIFY is ignored as in rbg_profile_block_device. */
int rbg_profile_block_noalloc(rbg_ctx *ctx, const uint32_t *d_primes, uint32_t n,
                              uint32_t flags, rbg_record *d_out);

/* ---- device-resident prime table (interval queries) ---------------------
 *
 * Every prime below 2^32 as a bitmap of the odd numbers in HBM (256 MiB,
 * plus 4 MiB of block counts and prefixes), built once per device by a
 * segmented sieve on the device the first time a context needs it and kept
 * until the process ends (outside rbg_config.hbm_budget).  Ranks count
 * primes from 0: p_0 = 2, p_1 = 3, ...  No reference counterpart: the
 * reference's supervisor hands its workers prime lists from a host sieve
 * (src/census/prime_sieve.cpp:8-22); these calls let a serving loop resolve
 * interval queries on the device and stream them in pieces by rank. */

/* Builds the table if needed; *count = number of primes below 2^32. */
int rbg_prime_table(rbg_ctx *ctx, uint64_t *count);

/* *rank = number of primes below x (x <= 2^32). */
int rbg_prime_rank(rbg_ctx *ctx, uint64_t x, uint64_t *rank);

/* Primes of ranks first .. first+n-1 into out (host); *found = how many of
 * them lie below 2^32 (n <= RBG_MAX_BATCH). */
int rbg_primes_by_rank(rbg_ctx *ctx, uint64_t first, uint32_t n, uint32_t *out, uint32_t *found);

/* Profiles of the primes of ranks first .. first+n-1 (n <= RBG_MAX_BLOCK),
 * selected on the device from the table (no host primality test); out gets
 * *found records.  Same flags and all-or-nothing contract as
 * rbg_profile_block. */
int rbg_profile_ranked(rbg_ctx *ctx, uint64_t first, uint32_t n, uint32_t flags,
                       rbg_record *out, uint32_t *found);

/* The wider arithmetic domain: primes 1000 <= p < 2^33 (and any smaller
 * prime).  Replaces the reference's 33-bit worker and engine
 * (reproducibility/gpu_method/source/worker33.cu:19-41 calculate_prime /
 * main, engine33/src/frozen_entry.cu, engine33/src/math33.hpp:9-72): the
 * same census with residues of 33 bits (Montgomery products with R = 2^32,
 * two conditional subtractions).  Primes below 2^32 of the same call take
 * the 32-bit kernels.  Same flags, limits and all-or-nothing contract as
 * rbg_profile_block / rbg_profile_batch / rbg_estimate_bytes. */
int rbg_profile_block_wide(rbg_ctx *ctx, const uint64_t *primes, uint32_t n,
                           uint32_t flags, rbg_record_wide *out);
int rbg_profile_batch_wide(rbg_ctx *ctx, const uint64_t *primes, uint32_t n,
                           uint32_t flags, rbg_record_wide *out);
int rbg_estimate_bytes_wide(const rbg_ctx *ctx, const uint64_t *primes, uint32_t n,
                            uint32_t flags, uint64_t *bytes);

/* Message for the last nonzero status of this context ("" if none). */
const char *rbg_last_error(const rbg_ctx *ctx);

/* Device-time breakdown of the last call (HIP events on the context's
 * stream) and its work counts. */
typedef struct rbg_timing {
    double total_ms;         /* whole call, first copy to last copy */
    double dense_scan_ms;    /* dense scan launch(es) */
    double sparse_prep_ms;   /* labels, inverses, colour planes */
    double sparse_seed_ms;   /* seed runs of 1..B */
    double sparse_filter_ms; /* anchor filter + completion */
    uint64_t dense_primes, sparse_primes, sparse_profiles, fallbacks;
    uint64_t positions;      /* half-interval positions of the dense scan */
    uint64_t products;       /* Montgomery products of the dense scan */
    uint64_t launches;       /* kernel launches of the dominant stages */
    uint64_t filter_windows; /* plane windows the filter needed, in 32-anchor dwords: per
                                word the stages through the one that killed its last
                                anchor (all t - 1 for a passing word), plus the screen's */
    uint64_t filter_words;   /* 64- or 96-anchor words the filter examined */
    uint64_t filter_completions; /* surviving anchors extended to their runs */
    double a4_host_ms;       /* elliptic A4 counts on host threads (beside the device) */
    double a4_wait_ms;       /* of that, time the call waited after the device was done */
    uint64_t rescans;        /* dense scans repeated with every run tracked (a longest
                                run below its floor; expected never above small primes) */
    uint64_t filter_loads;   /* wave-level 4-dword load instructions of the filter's word
                                loop for the primes below 2^32 (stage and base windows;
                                each moves 64 x 16 bytes through a CU's vector-memory
                                path whatever the active lanes); 0 for wider primes */
    uint64_t spec_misses;    /* profiles whose longest run stayed below their speculative
                                threshold (their primes are among the fallbacks) */
} rbg_timing;

int rbg_last_timing(const rbg_ctx *ctx, rbg_timing *out);

/* Compute units, peak engine clock (kHz) and HBM size of the context's device. */
int rbg_device_info(const rbg_ctx *ctx, uint32_t *cus, uint32_t *clock_khz, uint64_t *hbm_bytes);

void rbg_close(rbg_ctx *ctx);

/* A4 values of n primes on the host through an elliptic point count
 * (p >= 1000; dense count below; p = 2 gives 0).  Replaces the moment
 * worker's per-prime loop, src/census/moment_worker.cpp:35-41 (cardinality +
 * quadruples, src/census/theorie/elliptic4_mestre.cpp:42-67).  threads = 0
 * picks the host's core count (at most 16).  Returns RBG_EINVAL for a
 * non-prime and RBG_ECERTIFY if a count fails its congruence checks. */
int rbg_a4_host(const uint32_t *primes, uint32_t n, uint32_t *out, uint32_t threads);

/* Same for primes below 2^33 (the reference's point count accepts
 * 1000 <= p < 2^33, src/census/theorie/elliptic4_mestre.cpp:75). */
int rbg_a4_host_wide(const uint64_t *primes, uint32_t n, uint32_t *out, uint32_t threads);

/* A4 of one prime and the number of points its count drew (*attempts).  With
 * out_len non-NULL also its audit certificate as JSON in the caller's buffer
 * (same length convention as the frame writers): the field names and meaning
 * of the reference's certificate (src/census/theorie/elliptic4_mestre.cpp:
 * 68-73, read by src/census/theorie/certify_elliptic.py:13-47) -- point count
 * N, Hasse window, the curve y^2 = x (x - s)(x - 9 s) (twist_parameter s,
 * 1 for the curve itself) and its certificate_points (x, y, exact order and
 * the order's prime factors) whose orders' lcm has exactly one multiple in
 * the Hasse window.  p < 1000 gives the small_dense certificate.  Serves the
 * audit entry of the moment worker's protocol, moment_worker.cpp:43-45. */
int rbg_a4_certificate(uint32_t p, uint32_t *a4, uint64_t *attempts, char *out, size_t *out_len);

/* Measurement probes for the roofline (no reference counterpart).  kind 0-6:
 * issue rate, in lane operations per second, of add(x, y|1), mul_lo(x, y|1),
 * mul_hi(x, y|0x80000001) (each with one v_or_b32), v_alignbit_b32, whole
 * Montgomery products, plain add and plain v_mul_lo_u32 (8 independent
 * chains per thread, 8 waves per SIMD); kind 11-14: CU cycles per wave-wide
 * load of 1-4 dwords per lane with lanes `param` = 4 (consecutive dwords),
 * 8 or 12 (one contiguous span; 12 is the anchor filter's word stride) or
 * 4160 (one line each) bytes apart, from an
 * L2-resident buffer, 8 independent loads in flight per wave (issue cost,
 * not latency). */
int rbg_microbench(int device, int kind, uint32_t param, double *out);

/* Test hook for the 32-bit residue arithmetic of the dense scan, on
 * `device` for odd p >= 3 (prime or not): kind 0 writes out[i] =
 * a[i] b[i] 2^-32 mod p (one Montgomery product; a[i], b[i] < p), kind 1
 * out[i] = a[i]^b[i] mod p computed in the Montgomery domain (a[i] < p),
 * kind 2 the Montgomery form (a[i] + b[i]) 2^32 mod p reached by the scan's
 * position recurrence (b[i] additions of 2^32 mod p, b[i] <= 4096).
 * Replaces the reference's device Montgomery self-check
 * (src/census/engine32/build/original_scan_gpu.cu:550-574). */
int rbg_check_mont32(int device, const uint32_t *a, const uint32_t *b, uint32_t n, uint32_t p,
                     int kind, uint32_t *out);

/* Test hook for the wide residue arithmetic: on `device`, kind 0 writes
 * out[i] = a[i] b[i] 2^-32 mod p (the kernels' 33-bit Montgomery product),
 * kind 1 the Montgomery power of a[i] to b[i] (R = 2^32 mod p = 2^32), for
 * odd 2^32 < p < 2^33 and a[i] (and for kind 0 b[i]) < p.  No reference
 * counterpart (the reference checks its wide multiply by the reciprocal
 * lemma, paper/GPU_METHOD_EN.tex:1001-1025). */
int rbg_check_wide(int device, const uint64_t *a, const uint64_t *b, uint32_t n, uint64_t p,
                   int kind, uint64_t *out);

/* Direct verification of n records on the host (no device needed): the
 * admissibility mask, K from M and L, and for every admissible order the
 * witness run (monochromatic over W..W+M-1 with differing neighbours) and
 * the prefix (x = 1..L in the identity class, L+1 not) by direct powers.
 * Replaces the reference's certify_witness and prefix check
 * (src/census/engine32/vendor/research/baseline/baseline/src/advance.cpp:
 * 7-13, src/census/worker32.cu:37-38), the check that RBG_CERTIFY runs on
 * every result.  Returns RBG_OK, RBG_ECERTIFY with the index of the first
 * failing record in *first_bad (n when all pass), or RBG_EINVAL for a
 * record whose p is not prime.  threads = 0: one per host core. */
int rbg_certify_records(const rbg_record *recs, uint32_t n, uint32_t threads,
                        uint32_t *first_bad);

/* Admissibility mask of p (bit r-2 iff r | p-1); 0 for p < 3. */
uint32_t rbg_mask(uint32_t p);

/* Up to `count` consecutive primes >= start (and < 2^32) into out; returns how
 * many were written.  Builds prime blocks (the reference's segmented sieve,
 * src/census/prime_sieve.cpp:8-22, serves the same purpose). */
uint32_t rbg_primes_from(uint64_t start, uint32_t count, uint32_t *out);

/* Deterministic primality for 32-bit inputs. */
int rbg_is_prime(uint32_t p);

/* Wider domain helpers (p < 2^33; the reference's wide33::prime and
 * prime_sieve bounds, engine33/src/math33.hpp:72, src/census/prime_sieve.cpp:11):
 * primality (0 for p >= 2^33), admissibility mask, consecutive primes. */
int rbg_is_prime_wide(uint64_t p);
uint32_t rbg_mask_wide(uint64_t p);
uint32_t rbg_primes_from_wide(uint64_t start, uint32_t count, uint64_t *out);

/* Library version string. */
const char *rbg_version(void);

/* ---- pack records (host only; no device needed) ------------------------- */

/* Scan frame of n records, ascending primes in [lo, hi) (lo >= 1): the
 * reference's pack writer, docs/DATA_FORMAT.md:9-31 (decoder
 * scripts/atlas_codec.py:20-44), byte-identical for the same zlib level.
 * witnesses: write the witness varints (flag 1).  out: caller buffer,
 * *out_len in = capacity, out = frame length; RBG_ECAPACITY (with the needed
 * length in *out_len) if the buffer is short, RBG_EINVAL if a record fails
 * the decoder's guards (mask, 1 <= L <= M <= (p-1)/r, 1 <= W <= (p-1)/2, K). */
int rbg_encode_scan_frame(const rbg_record *recs, uint32_t n, uint64_t lo, uint64_t hi,
                          int witnesses, int level, uint8_t *out, size_t *out_len);

/* A4 frame of the same records (DATA_FORMAT.md:35-47, decoder
 * atlas_codec.py:54-60): A4 values as u32 LE, with the block's prime hash and
 * profile science hash in the header.  Same buffer convention. */
int rbg_encode_a4_frame(const rbg_record *recs, uint32_t n, uint64_t lo, uint64_t hi, int level,
                        uint8_t *out, size_t *out_len);

/* Commitments of a block (atlas_codec.py:46-51 and the prime hash of
 * DATA_FORMAT.md): 32-byte SHA-256 digests. */
int rbg_science_hash(const rbg_record *recs, uint32_t n, uint8_t out[32]);
int rbg_prime_hash(const rbg_record *recs, uint32_t n, uint8_t out[32]);
/* SHA-256 of the block's A4 values as u32 LE (the values hash of an A4 frame,
 * docs/DATA_FORMAT.md:43-47). */
int rbg_values_hash(const rbg_record *recs, uint32_t n, uint8_t out[32]);

/* One step of the atlas hash chains (docs/DATA_FORMAT.md:63-69, walked by
 * scripts/verify_atlas.py:79-81): scan = SHA256(prev || prime_hash ||
 * profile_hash || u64 lo || u64 hi), A4 the same with values_hash before
 * the interval.  out may alias prev.  Chains start from 32 zero bytes. */
int rbg_chain_scan(const uint8_t prev[32], const uint8_t prime_hash[32],
                   const uint8_t profile_hash[32], uint64_t lo, uint64_t hi, uint8_t out[32]);
int rbg_chain_a4(const uint8_t prev[32], const uint8_t prime_hash[32],
                 const uint8_t profile_hash[32], const uint8_t values_hash[32], uint64_t lo,
                 uint64_t hi, uint8_t out[32]);

/* ---- reference bridge (baseline_bridge.cu:7,21-27) ---------------------- */

/* Prepares the process-wide context of the bridge.  `cap` (the largest
 * prime the reference sized its buffers for) is advisory here: workspaces
 * grow on demand. */
void baseline_init(unsigned cap);

/* M by order into out[r] and prefixes into prefix[r] (arrays of 10, indexed
 * by r); only the orders r whose bit r-2 is set in `mask` are written, as in
 * the reference.  `full` nonzero computes every admissible order of p (an
 * order of `mask` that does not divide p-1 then gets 0); zero requires
 * `mask` to be a nonempty subset of p's admissible orders (the reference
 * serves five fixed masks there, baseline_bridge.cu:24).  Returns 0; 2 for a
 * mask that cannot be served; -1 for any other failure (not a prime below
 * 2^32, no device, a device error), where the reference ends the process.
 * Serialised by a lock (the reference is not thread-safe). */
int baseline_profile(unsigned p, unsigned mask, unsigned *out, unsigned *prefix, int full);

#ifdef __cplusplus
}
#endif

#endif /* RABUNG_H */
