/*
 * rbg_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference census worker's dense profile, used as the
 * checker for the HIP library.  Only tests/, the smoke check and the bench's
 * cpu_baseline leg load it; the product path never does.
 *
 * Method (deliberately independent of the GPU path): a multiplicative walk
 * x = g^i over a primitive root g writes the discrete-log label i mod m for
 * every residue, then each order r is scanned over the FULL interval 1..p-1
 * in ordinary integer order, without reflection folding and without
 * Montgomery arithmetic.  This is the scheme of the reference CPU workers:
 *   - src/census/small_cpu.cpp:15-33            (emit: walk, run scan, witness,
 *                                                 prefix, K)
 *   - reproducibility/bounds_update/independent_profile.cpp:57-81
 *                                               (walk closure, colour
 *                                                 populations, reflection)
 *   - src/census/moment_worker.cpp:11-18        (dense A4)
 *   - docs/METHODS.md:1-29                      (contract: M, L, witness, K, A4)
 *
 * Pinning (see tests/test_oracle.py): the reference sample frames for block 0
 * (all 4096 primes, with A4) and a seeded subset of block 49900 (primes just
 * below 2^32), the reference holdout/night JSON profiles, the worked example
 * p=101 r=2 -> (7,1,8,19), and literal pow(x,(p-1)/r,p) brute force for small p.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;

/* Per-prime result; index k = r-2 for r in 2..9. */
typedef struct {
    u32 p;
    u32 mask;
    u32 M[8];
    u32 L[8];
    u32 W[8];
    u64 K[8];
    u32 A4;
    u32 status; /* 0 ok, nonzero = refused */
} orc_record;

static u64 mulmod(u64 a, u64 b, u64 p) { return a * b % p; } /* p < 2^32 */

static u64 powmod(u64 a, u64 e, u64 p) {
    u64 v = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) v = mulmod(v, a, p);
        a = mulmod(a, a, p);
        e >>= 1;
    }
    return v;
}

/* Trial division, small_cpu.cpp:14 (prime). */
int orc_is_prime(u64 n) {
    if (n < 2) return 0;
    if (n % 2 == 0) return n == 2;
    for (u64 q = 3; q * q <= n; q += 2)
        if (n % q == 0) return 0;
    return 1;
}

/* Admissibility mask, scripts/atlas_codec.py:6 (expected_mask). */
u32 orc_mask(u64 p) {
    u32 mask = 0;
    if (p < 3) return 0;
    for (u32 r = 2; r <= 9; ++r)
        if ((p - 1) % r == 0) mask |= 1u << (r - 2);
    return mask;
}

static u64 gcd64(u64 a, u64 b) {
    while (b) { u64 t = a % b; a = b; b = t; }
    return a;
}

/* Least primitive root, small_cpu.cpp:19. */
u64 orc_primitive_root(u64 p) {
    u64 f[64];
    int nf = 0;
    u64 n = p - 1;
    for (u64 q = 2; q * q <= n; ++q)
        if (n % q == 0) {
            f[nf++] = q;
            while (n % q == 0) n /= q;
        }
    if (n > 1) f[nf++] = n;
    for (u64 g = 2; g < p; ++g) {
        int ok = 1;
        for (int i = 0; i < nf; ++i)
            if (powmod(g, (p - 1) / f[i], p) == 1) { ok = 0; break; }
        if (ok) return g;
    }
    return p == 2 ? 1 : 0;
}

/* ---- the walk -------------------------------------------------------- */

typedef struct {
    uint16_t *lab;
    u64 p, g, m, i0, i1;
    u64 last_x;
    int dup;
} walk_job;

/* x*g mod p with Shoup's precomputed quotient gq = floor(g*2^64/p)
 * (V. Shoup, NTL MulModPrecon): for x < p the estimate q = hi(x*gq) is the
 * true quotient or one less, so one conditional subtraction finishes. */
static inline u64 mulmod_shoup(u64 x, u64 g, u64 gq, u64 p) {
    u64 q = (u64)(((unsigned __int128)x * gq) >> 64);
    u64 r = x * g - q * p;
    return r >= p ? r - p : r;
}

static void *walk_thread(void *arg) {
    walk_job *j = (walk_job *)arg;
    const u64 p = j->p, g = j->g, m = j->m;
    const u64 gq = (u64)(((unsigned __int128)g << 64) / p);
    u64 x = powmod(g, j->i0, p), lab = j->i0 % m;
    /* p-1 writes over p-1 residues: a repeated visit would leave some
     * residue at the 0xFFFF sentinel, which the scans reject */
    for (u64 i = j->i0; i < j->i1; ++i) {
        j->lab[x] = (uint16_t)lab;
        if (++lab == m) lab = 0;
        x = mulmod_shoup(x, g, gq, p);
    }
    j->last_x = x;
    j->dup = 0;
    return 0;
}

typedef struct {
    uint16_t *lab;
    u64 a, b;
} fill_job;

static void *fill_thread(void *arg) {
    fill_job *f = (fill_job *)arg;
    memset(f->lab + f->a, 0xFF, sizeof(uint16_t) * (size_t)(f->b - f->a));
    return 0;
}

/* ---- per-order scan --------------------------------------------------- */

/* Summary of the colour sequence over one chunk [x0, x1] of 1..p-1 for one
 * order r: the run touching x0 (head), the run touching x1 (tail), and the
 * complete runs strictly inside (best length, least reflected start among
 * them, and their A4 window count).  Runs are taken in plain integer order;
 * the merge below joins a tail with the next head when the colours agree,
 * which is the only way a run can cross a chunk boundary. */
typedef struct {
    u64 x0, x1;
    u32 head_c, tail_c;
    u64 head_len, tail_len;
    u64 best, best_w, a4;
    u64 cnt[9];
    int bad;
} chunk_sum;

typedef struct {
    const uint16_t *lab;
    u64 p;
    u32 r;
    chunk_sum *s;
} chunk_job;

/* reflected start of the run [s, e] (docs/METHODS.md:7-9) */
static inline u64 refl_start(u64 s, u64 e, u64 p) { return s < p - e ? s : p - e; }

static inline void take(u64 len, u64 w, u64 *best, u64 *best_w) {
    if (len > *best || (len == *best && w < *best_w)) { *best = len; *best_w = w; }
}

#define CHUNK_BODY(R)                                                         \
    do {                                                                      \
        u64 x = s->x0;                                                        \
        u32 first = lab[x];                                                   \
        bad |= first == 0xFFFFu;                                              \
        u32 c0 = first % (R);                                                 \
        s->head_c = c0;                                                       \
        u64 start = x, last = c0;                                             \
        int in_head = 1;                                                      \
        cnt[c0]++;                                                            \
        for (x = s->x0 + 1; x <= s->x1; ++x) {                                \
            u32 v = lab[x];                                                   \
            bad |= v == 0xFFFFu;                                              \
            u32 c = v % (R);                                                  \
            cnt[c]++;                                                         \
            if (c != last) {                                                  \
                u64 len = x - start;                                          \
                if (in_head) { s->head_len = len; in_head = 0; }              \
                else {                                                        \
                    take(len, refl_start(start, x - 1, p), &best, &best_w);   \
                    if (len > 3) a4 += len - 3;                               \
                }                                                             \
                start = x; last = c;                                          \
            }                                                                 \
        }                                                                     \
        if (in_head) s->head_len = s->x1 - s->x0 + 1;                         \
        s->tail_c = (u32)last;                                                \
        s->tail_len = s->x1 - start + 1;                                      \
    } while (0)

static void *chunk_thread(void *arg) {
    chunk_job *j = (chunk_job *)arg;
    const uint16_t *lab = j->lab;
    const u64 p = j->p;
    chunk_sum *s = j->s;
    u64 best = 0, best_w = p, a4 = 0, cnt[9] = {0};
    int bad = 0;
    switch (j->r) {
    case 2: CHUNK_BODY(2); break;
    case 3: CHUNK_BODY(3); break;
    case 4: CHUNK_BODY(4); break;
    case 5: CHUNK_BODY(5); break;
    case 6: CHUNK_BODY(6); break;
    case 7: CHUNK_BODY(7); break;
    case 8: CHUNK_BODY(8); break;
    case 9: CHUNK_BODY(9); break;
    default: bad = 1;
    }
    s->best = best; s->best_w = best_w; s->a4 = a4; s->bad = bad;
    memcpy(s->cnt, cnt, sizeof(cnt));
    return 0;
}

/* Walk the chunk summaries in order; an open run is carried across chunk
 * boundaries while the colour agrees.  Produces M, reflected witness, A4. */
static void merge_chunks(const chunk_sum *s, int n, u64 p, u64 *M, u64 *W, u64 *A4) {
    u64 best = 0, best_w = p, a4 = 0;
    u64 open_start = 0, open_len = 0;
    u32 open_c = 0xFFFFFFFFu;
    for (int i = 0; i < n; ++i) {
        const chunk_sum *c = &s[i];
        int whole = c->head_len == c->x1 - c->x0 + 1; /* chunk is one run */
        if (open_len && c->head_c == open_c) {
            open_len += c->head_len;
        } else {
            if (open_len) {
                take(open_len, refl_start(open_start, open_start + open_len - 1, p), &best, &best_w);
                if (open_len > 3) a4 += open_len - 3;
            }
            open_start = c->x0;
            open_len = c->head_len;
            open_c = c->head_c;
        }
        if (!whole) {
            /* the head run closes inside this chunk */
            take(open_len, refl_start(open_start, open_start + open_len - 1, p), &best, &best_w);
            if (open_len > 3) a4 += open_len - 3;
            take(c->best, c->best_w, &best, &best_w);
            a4 += c->a4;
            open_start = c->x1 - c->tail_len + 1;
            open_len = c->tail_len;
            open_c = c->tail_c;
        }
    }
    if (open_len) {
        take(open_len, refl_start(open_start, open_start + open_len - 1, p), &best, &best_w);
        if (open_len > 3) a4 += open_len - 3;
    }
    *M = best; *W = best_w; *A4 = a4;
}

/*
 * Dense profile of one prime p < 2^32.  threads = worker threads for the walk
 * (the scan uses one thread per admissible order).  Returns 0 on success;
 * nonzero when p is not a prime below 2^32, memory is short, or an internal
 * consistency check (walk closure, duplicate visit, colour population) fails.
 * Memory: 2 bytes per residue (8.6 GB just below 2^32).
 */
int orc_profile(u64 p, int threads, orc_record *out) {
    memset(out, 0, sizeof(*out));
    out->p = (u32)p;
    if (p >= (1ull << 32) || !orc_is_prime(p)) { out->status = 1; return 1; }
    if (p == 2) return 0; /* mask 0, A4 placeholder 0 */
    u32 mask = orc_mask(p);
    out->mask = mask;
    u64 m = 1;
    for (u32 r = 2; r <= 9; ++r)
        if (mask & (1u << (r - 2))) m = m / gcd64(m, r) * r;
    u64 g = orc_primitive_root(p);
    uint16_t *lab = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)p);
    if (!lab) { out->status = 2; return 2; }
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    {
        fill_job fj[64];
        pthread_t ft[64];
        for (int t = 0; t < threads; ++t) {
            fj[t].lab = lab;
            fj[t].a = p * (u64)t / (u64)threads;
            fj[t].b = p * (u64)(t + 1) / (u64)threads;
            pthread_create(&ft[t], 0, fill_thread, &fj[t]);
        }
        for (int t = 0; t < threads; ++t) pthread_join(ft[t], 0);
    }
    u64 steps = p - 1;
    if ((u64)threads * 64 > steps) threads = 1;
    walk_job jobs[64];
    pthread_t tid[64];
    for (int t = 0; t < threads; ++t) {
        jobs[t].lab = lab; jobs[t].p = p; jobs[t].g = g; jobs[t].m = m;
        jobs[t].i0 = steps * (u64)t / (u64)threads;
        jobs[t].i1 = steps * (u64)(t + 1) / (u64)threads;
        jobs[t].dup = 0;
        pthread_create(&tid[t], 0, walk_thread, &jobs[t]);
    }
    int bad = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], 0);
        bad |= jobs[t].dup;
    }
    if (jobs[threads - 1].last_x != 1) bad = 1; /* walk closure g^(p-1)=1 */
    /* unvisited residues are caught by the scans (sentinel 0xFFFF) */
    if (bad) { free(lab); out->status = 3; return 3; }

    /* chunked scans: every admissible r is cut into `threads` chunks */
    int nr = 0;
    u32 rs[8];
    for (u32 r = 2; r <= 9; ++r)
        if (mask & (1u << (r - 2))) rs[nr++] = r;
    int nchunk = threads;
    if ((u64)nchunk * 16 > p - 1) nchunk = 1;
    chunk_sum *cs = (chunk_sum *)calloc((size_t)nr * nchunk, sizeof(chunk_sum));
    chunk_job *cj = (chunk_job *)calloc((size_t)nr * nchunk, sizeof(chunk_job));
    pthread_t *ct = (pthread_t *)calloc((size_t)nr * nchunk, sizeof(pthread_t));
    if (!cs || !cj || !ct) { free(cs); free(cj); free(ct); free(lab); out->status = 2; return 2; }
    for (int i = 0; i < nr; ++i)
        for (int k = 0; k < nchunk; ++k) {
            chunk_sum *c = &cs[i * nchunk + k];
            c->x0 = 1 + (p - 1) * (u64)k / (u64)nchunk;
            c->x1 = (p - 1) * (u64)(k + 1) / (u64)nchunk;
            cj[i * nchunk + k].lab = lab; cj[i * nchunk + k].p = p;
            cj[i * nchunk + k].r = rs[i]; cj[i * nchunk + k].s = c;
        }
    /* run at most `threads` chunk jobs at a time */
    int total = nr * nchunk;
    for (int base = 0; base < total; base += threads) {
        int top = base + threads < total ? base + threads : total;
        for (int t = base; t < top; ++t) pthread_create(&ct[t], 0, chunk_thread, &cj[t]);
        for (int t = base; t < top; ++t) pthread_join(ct[t], 0);
    }
    for (int i = 0; i < nr; ++i) {
        u32 r = rs[i], k = r - 2;
        u64 cnt[9] = {0};
        for (int c = 0; c < nchunk; ++c) {
            if (cs[i * nchunk + c].bad) bad = 1;
            for (u32 q = 0; q < r; ++q) cnt[q] += cs[i * nchunk + c].cnt[q];
        }
        for (u32 q = 0; q < r; ++q)
            if (cnt[q] != (p - 1) / r) bad = 1; /* colour populations */
        u64 M, W, A4;
        merge_chunks(&cs[i * nchunk], nchunk, p, &M, &W, &A4);
        u64 pre = 0; /* prefix: small_cpu.cpp:27 */
        while (pre + 1 < p && lab[pre + 1] % r == 0) ++pre;
        u64 b = (((p - 1) / 2) % r == 0) ? 2 * pre + 2 : pre + 2;
        out->M[k] = (u32)M; out->W[k] = (u32)W; out->L[k] = (u32)pre;
        out->K[k] = M + 1 > b ? M + 1 : b;
        if (r == 2) out->A4 = (u32)A4;
    }
    free(cs); free(cj); free(ct);
    free(lab);
    if (bad) { out->status = 4; return 4; }
    return 0;
}

/* Profiles of n primes (each single-threaded walk), primes spread over
 * `threads` worker threads.  Intended for many small primes. */
typedef struct {
    const u64 *ps;
    orc_record *out;
    int n, stride, first;
    int rc;
} many_job;

static void *many_thread(void *arg) {
    many_job *j = (many_job *)arg;
    for (int i = j->first; i < j->n; i += j->stride) {
        int rc = orc_profile(j->ps[i], 1, &j->out[i]);
        if (rc) j->rc = rc;
    }
    return 0;
}

int orc_profile_many(const u64 *ps, int n, int threads, orc_record *out) {
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    many_job jobs[64];
    pthread_t tid[64];
    for (int t = 0; t < threads; ++t) {
        jobs[t].ps = ps; jobs[t].out = out; jobs[t].n = n; jobs[t].stride = threads;
        jobs[t].first = t; jobs[t].rc = 0;
        pthread_create(&tid[t], 0, many_thread, &jobs[t]);
    }
    int rc = 0;
    for (int t = 0; t < threads; ++t) {
        pthread_join(tid[t], 0);
        if (jobs[t].rc) rc = jobs[t].rc;
    }
    return rc;
}

/* n consecutive primes >= start (trial division); returns count written. */
int orc_next_primes(u64 start, int n, u64 *out) {
    int k = 0;
    for (u64 x = start; k < n && x < (1ull << 32); ++x)
        if (orc_is_prime(x)) out[k++] = x;
    return k;
}

/* Literal reflection check used by tests: cl(p-y) = cl(y) + sigma. */
u64 orc_powmod(u64 a, u64 e, u64 p) { return powmod(a, e, p); }
