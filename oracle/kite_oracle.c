/*
 * kite_oracle.c -- CPU restatement of jiehua1995/Kite v0.7.5's k-mer counting
 * path: FASTQ chunking and parsing, rolling k-mers, Fibonacci partitioning,
 * sort + run-length counting, DVLE, the .kite writer/reader, dump, histo,
 * merge and the union rows that feed analyze.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker and the timed CPU
 * baseline.  The shipped engine never links, loads or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 */
#define _GNU_SOURCE
#include "kite_oracle.h"

#include <errno.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

typedef unsigned __int128 u128;

static __thread char g_err[512];

static int fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

const char *ok_last_error(void) { return g_err; }

/* ---- base encoding: count/kmer.rs:2-13 --------------------------------- */
static uint8_t ENC[256];
static void enc_init(void) {
    static int done = 0;
    if (done) return;
    memset(ENC, 255, sizeof ENC);
    ENC['A'] = ENC['a'] = 0;
    ENC['C'] = ENC['c'] = 1;
    ENC['G'] = ENC['g'] = 2;
    ENC['T'] = ENC['t'] = 3;
    done = 1;
}

int ok_key_words(int k) { return k <= 31 ? 1 : (k <= 64 ? 2 : 4); }

/* ---- rolling extraction ------------------------------------------------- */
typedef void (*key_sink)(void *ctx, const uint64_t *key);

/* KmerIter (count/kmer.rs:203-268): u64, k <= 32. */
static size_t extract_w1(const uint8_t *s, size_t len, int k, int canon, key_sink f, void *ctx) {
    uint64_t mask = (k == 32) ? ~0ull : ((1ull << (2 * k)) - 1);
    unsigned shift = 2u * (unsigned)(k - 1);
    uint64_t fwd = 0, rev = 0;
    uint32_t valid = 0;
    size_t n = 0;
    for (size_t i = 0; i < len; i++) {
        uint8_t e = ENC[s[i]];
        if (e == 255) { fwd = rev = 0; valid = 0; continue; }
        fwd = ((fwd << 2) | e) & mask;
        rev = (rev >> 2) | ((uint64_t)(e ^ 3) << shift);
        if (++valid >= (uint32_t)k) {
            uint64_t key = (canon && rev < fwd) ? rev : fwd;
            f(ctx, &key);
            n++;
        }
    }
    return n;
}

/* KmerIter128 (count/kmer.rs:273-338): u128, k <= 64. */
static size_t extract_w2(const uint8_t *s, size_t len, int k, int canon, key_sink f, void *ctx) {
    u128 mask = (k == 64) ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
    unsigned shift = 2u * (unsigned)(k - 1);
    u128 fwd = 0, rev = 0;
    uint32_t valid = 0;
    size_t n = 0;
    for (size_t i = 0; i < len; i++) {
        uint8_t e = ENC[s[i]];
        if (e == 255) { fwd = rev = 0; valid = 0; continue; }
        fwd = ((fwd << 2) | e) & mask;
        rev = (rev >> 2) | ((u128)(e ^ 3) << shift);
        if (++valid >= (uint32_t)k) {
            u128 v = (canon && rev < fwd) ? rev : fwd;
            uint64_t key[2] = {(uint64_t)v, (uint64_t)(v >> 64)};
            f(ctx, key);
            n++;
        }
    }
    return n;
}

/* KmerIterU256 (count/kmer.rs:348-432): (hi, lo) u128 pair, 65 <= k <= 128. */
static size_t extract_w4(const uint8_t *s, size_t len, int k, int canon, key_sink f, void *ctx) {
    unsigned hi_bits = 2u * (unsigned)k - 128u;
    u128 mask_hi = (hi_bits == 128) ? ~(u128)0 : (((u128)1 << hi_bits) - 1);
    unsigned rev_top = 2u * (unsigned)(k - 1) - 128u;
    u128 flo = 0, fhi = 0, rlo = 0, rhi = 0;
    uint32_t valid = 0;
    size_t n = 0;
    for (size_t i = 0; i < len; i++) {
        uint8_t e = ENC[s[i]];
        if (e == 255) { flo = fhi = rlo = rhi = 0; valid = 0; continue; }
        u128 carry_f = flo >> 126;
        flo = (flo << 2) | e;
        fhi = ((fhi << 2) | carry_f) & mask_hi;
        u128 carry_r = rhi << 126;
        rhi = (rhi >> 2) | ((u128)(e ^ 3) << rev_top);
        rlo = (rlo >> 2) | carry_r;
        if (++valid >= (uint32_t)k) {
            int use_rev = canon && (rhi < fhi || (rhi == fhi && rlo < flo));
            u128 lo = use_rev ? rlo : flo, hi = use_rev ? rhi : fhi;
            uint64_t key[4] = {(uint64_t)lo, (uint64_t)(lo >> 64), (uint64_t)hi, (uint64_t)(hi >> 64)};
            f(ctx, key);
            n++;
        }
    }
    return n;
}

static size_t extract_any(const uint8_t *s, size_t len, int k, int canon, int words, key_sink f,
                          void *ctx) {
    enc_init();
    if (words == 1) return extract_w1(s, len, k, canon, f, ctx);
    if (words == 2) return extract_w2(s, len, k, canon, f, ctx);
    return extract_w4(s, len, k, canon, f, ctx);
}

typedef struct { uint64_t *out; size_t cap, n; int words; } vec_sink;
static void vec_push(void *ctx, const uint64_t *key) {
    vec_sink *v = (vec_sink *)ctx;
    if (v->n < v->cap) memcpy(v->out + v->n * v->words, key, 8 * (size_t)v->words);
    v->n++;
}

size_t ok_extract(const uint8_t *seq, size_t len, int k, int canonical, int words, uint64_t *out,
                  size_t cap) {
    vec_sink v = {out, cap, 0, words};
    return extract_any(seq, len, k, canonical, words, vec_push, &v);
}

/* ---- partition hash: count/kmer.rs:67-71, 111-116, 156-161 -------------- */
uint32_t ok_bucket(const uint64_t *key, int words, int bits) {
    uint64_t f = key[0];
    if (words >= 2) f ^= key[1];
    if (words == 4) f ^= key[2] ^ key[3];
    if (bits == 0) return 0; /* the reference shifts by 64 here; P=1 is not usable there */
    return (uint32_t)((f * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

/* ---- FASTQ parse: count/parser.rs:19-65 --------------------------------- */
typedef void (*seq_sink)(void *ctx, const uint8_t *seq, size_t len);

static int parse_seqs(const uint8_t *d, size_t len, seq_sink f, void *ctx) {
    if (len == 0) return 0;
    if (d[0] != '@') return -1;
    int state = 0; /* 0 header, 1 seq, 2 plus, 3 qual */
    size_t seq_start = 0;
    for (size_t i = 0; i < len; i++) {
        if (d[i] != '\n') continue;
        if (state == 0) { state = 1; seq_start = i + 1; }
        else if (state == 1) {
            if (i > seq_start) f(ctx, d + seq_start, i - seq_start);
            state = 2;
        } else if (state == 2) state = 3;
        else state = 0;
    }
    if (state == 1 && seq_start < len) f(ctx, d + seq_start, len - seq_start);
    return 0;
}

typedef struct { uint64_t *spans; size_t cap, n; } span_sink;
static void span_push(void *ctx, const uint8_t *seq, size_t len) {
    span_sink *s = (span_sink *)ctx;
    (void)seq;
    if (s->n < s->cap) { s->spans[2 * s->n] = 0; s->spans[2 * s->n + 1] = len; }
    s->n++;
}

long ok_parse_sequences(const uint8_t *data, size_t len, uint64_t *spans, size_t cap) {
    span_sink s = {spans, cap, 0};
    /* record the start offset relative to data */
    if (len && data[0] != '@') return -1;
    int state = 0;
    size_t seq_start = 0;
    for (size_t i = 0; i < len; i++) {
        if (data[i] != '\n') continue;
        if (state == 0) { state = 1; seq_start = i + 1; }
        else if (state == 1) {
            if (i > seq_start) {
                if (s.n < cap) { spans[2 * s.n] = seq_start; spans[2 * s.n + 1] = i - seq_start; }
                s.n++;
            }
            state = 2;
        } else if (state == 2) state = 3;
        else state = 0;
    }
    if (state == 1 && seq_start < len) {
        if (s.n < cap) { spans[2 * s.n] = seq_start; spans[2 * s.n + 1] = len - seq_start; }
        s.n++;
    }
    (void)span_push;
    return (long)s.n;
}

/* ---- record-aligned chunking: count/bgzf.rs:109-121, decompress.rs:85-138 */
size_t ok_fastq_split_4lines(const uint8_t *data, size_t len) {
    size_t nl = 0, last = 0;
    for (size_t i = 0; i < len; i++)
        if (data[i] == '\n' && (++nl % 4) == 0) last = i + 1;
    return last;
}

size_t ok_plain_chunks(const uint8_t *data, size_t len, size_t chunk_size, uint64_t *chunks,
                       size_t cap) {
    /* The reader keeps a carry; each chunk = carry + next chunk_size bytes.
     * A short read ends the stream with the whole remainder as one chunk. */
    size_t pos = 0;      /* file bytes consumed by reads */
    size_t cstart = 0;   /* start of the current chunk (carry begins here) */
    size_t n = 0;
    for (;;) {
        size_t take = len - pos < chunk_size ? len - pos : chunk_size;
        size_t cend = pos + take;
        if (cend == cstart) break; /* empty chunk -> end of stream */
        pos = cend;
        size_t clen = cend - cstart;
        if (take < chunk_size) {
            if (n < cap) { chunks[2 * n] = cstart; chunks[2 * n + 1] = clen; }
            n++;
            cstart = cend;
            if (pos >= len) {
                /* next read returns 0 bytes with an empty carry -> None */
                break;
            }
            continue;
        }
        size_t split = ok_fastq_split_4lines(data + cstart, clen);
        if (split == 0 || split >= clen) {
            if (n < cap) { chunks[2 * n] = cstart; chunks[2 * n + 1] = clen; }
            n++;
            cstart = cend;
        } else {
            if (n < cap) { chunks[2 * n] = cstart; chunks[2 * n + 1] = split; }
            n++;
            cstart = cstart + split;
        }
    }
    return n;
}

/* ---- sort + RLE: count/sorter.rs:67-93, 388-502, 514-530 ---------------- */
static int key_cmp(const uint64_t *a, const uint64_t *b, int words) {
    for (int i = words - 1; i >= 0; i--) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

/* LSD byte radix over the 2k significant bits; the order is the unsigned
 * integer order of the (up to 256-bit) key, which for KmerU256 is (hi, lo). */
static void radix_sort_keys(uint64_t *keys, size_t n, int words, int k) {
    if (n < 2) return;
    size_t nbytes = (size_t)(2 * k + 7) / 8;
    uint64_t *tmp = (uint64_t *)malloc(n * 8 * (size_t)words);
    uint64_t *src = keys, *dst = tmp;
    size_t cnt[256];
    for (size_t b = 0; b < nbytes; b++) {
        int w = (int)(b / 8);
        unsigned sh = (unsigned)(b % 8) * 8u;
        memset(cnt, 0, sizeof cnt);
        for (size_t i = 0; i < n; i++) cnt[(src[i * words + w] >> sh) & 255]++;
        int single = 0;
        for (int d = 0; d < 256; d++) if (cnt[d] == n) single = 1;
        if (single) continue;
        size_t sum = 0;
        for (int d = 0; d < 256; d++) { size_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (size_t i = 0; i < n; i++) {
            size_t o = cnt[(src[i * words + w] >> sh) & 255]++;
            memcpy(dst + o * words, src + i * words, 8 * (size_t)words);
        }
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != keys) memcpy(keys, src, n * 8 * (size_t)words);
    free(tmp);
}

/* Sorts, then writes unique keys in place and u64 run lengths; returns u. */
static size_t sort_rle_u64(uint64_t *keys, size_t n, int words, int k, uint64_t *counts) {
    radix_sort_keys(keys, n, words, k);
    if (n == 0) return 0;
    size_t u = 0;
    uint64_t c = 1;
    for (size_t i = 1; i <= n; i++) {
        if (i < n && key_cmp(keys + i * words, keys + (i - 1) * words, words) == 0) { c++; continue; }
        if (u != i - 1) memcpy(keys + u * words, keys + (i - 1) * words, 8 * (size_t)words);
        counts[u++] = c;
        c = 1;
    }
    return u;
}

size_t ok_sort_rle(uint64_t *keys, size_t n, int words, int k, uint32_t *counts) {
    uint64_t *c64 = (uint64_t *)malloc((n ? n : 1) * 8);
    size_t u = sort_rle_u64(keys, n, words, k, c64);
    for (size_t i = 0; i < u; i++) counts[i] = c64[i] > 0xffffffffull ? 0xffffffffu : (uint32_t)c64[i];
    free(c64);
    return u;
}

/* ---- DVLE: format/run_block.rs:9-95 ------------------------------------- */
static size_t put_varint(uint8_t *o, uint64_t lo, uint64_t hi) {
    size_t n = 0;
    for (;;) {
        uint8_t b = (uint8_t)(lo & 0x7f);
        lo = (lo >> 7) | (hi << 57);
        hi >>= 7;
        if (lo == 0 && hi == 0) { o[n++] = b; return n; }
        o[n++] = b | 0x80;
    }
}

size_t ok_dvle_encode(const uint64_t *keys, const uint32_t *counts, size_t n, int words,
                      uint8_t *out) {
    size_t p = 0;
    if (n == 0) return 0;
    if (words == 4) { /* KmerU256 is written raw: lo16 | hi16 | count4 */
        for (size_t i = 0; i < n; i++) {
            memcpy(out + p, keys + i * 4, 32);
            memcpy(out + p + 32, counts + i, 4);
            p += 36;
        }
        return p;
    }
    memcpy(out, keys, 8 * (size_t)words);
    memcpy(out + 8 * words, counts, 4);
    p = 8 * (size_t)words + 4;
    for (size_t i = 1; i < n; i++) {
        const uint64_t *a = keys + (i - 1) * words, *b = keys + i * words;
        uint64_t dlo, dhi = 0;
        if (words == 1) dlo = b[0] - a[0];
        else {
            u128 av = (u128)a[0] | ((u128)a[1] << 64), bv = (u128)b[0] | ((u128)b[1] << 64);
            u128 d = bv - av;
            dlo = (uint64_t)d;
            dhi = (uint64_t)(d >> 64);
        }
        p += put_varint(out + p, dlo, dhi);
        p += put_varint(out + p, counts[i], 0);
    }
    return p;
}

/* LEB128 reads with the reference's 10 / 19 byte limits. */
static size_t get_varint(const uint8_t *d, size_t avail, size_t maxb, uint64_t *lo, uint64_t *hi) {
    u128 r = 0;
    unsigned sh = 0;
    size_t i;
    for (i = 0; i < avail && i < maxb; i++) {
        if (sh < 128) r |= (u128)(d[i] & 0x7f) << sh;
        if (!(d[i] & 0x80)) { *lo = (uint64_t)r; *hi = (uint64_t)(r >> 64); return i + 1; }
        sh += 7;
    }
    *lo = (uint64_t)r;
    *hi = (uint64_t)(r >> 64);
    return i;
}

/* ---- count: count/mod.rs:512-540, worker.rs:196-271, 296-449 ------------ */
typedef struct { uint64_t *v; size_t n, cap; } kvec;

typedef struct {
    const uint8_t *data;
    const uint64_t *chunks;
    size_t n_chunks;
    int k, canon, words, bits;
    uint32_t partitions;
    int tid, nthreads;
    kvec *parts;
    uint64_t kmers, reads;
} count_job;

static void kvec_push(kvec *v, const uint64_t *key, int words) {
    if (v->n + (size_t)words > v->cap) {
        size_t nc = v->cap ? v->cap * 2 : 1024;
        while (nc < v->n + (size_t)words) nc *= 2;
        v->v = (uint64_t *)realloc(v->v, nc * 8);
        v->cap = nc;
    }
    memcpy(v->v + v->n, key, 8 * (size_t)words);
    v->n += (size_t)words;
}

static void job_key(void *ctx, const uint64_t *key) {
    count_job *j = (count_job *)ctx;
    uint32_t b = ok_bucket(key, j->words, j->bits);
    kvec_push(&j->parts[b], key, j->words);
    j->kmers++;
}

static void job_seq(void *ctx, const uint8_t *seq, size_t len) {
    count_job *j = (count_job *)ctx;
    j->reads++;
    extract_any(seq, len, j->k, j->canon, j->words, job_key, j);
}

static void *count_worker(void *arg) {
    count_job *j = (count_job *)arg;
    for (size_t c = (size_t)j->tid; c < j->n_chunks; c += (size_t)j->nthreads) {
        /* a chunk that does not start with '@' is dropped (worker.rs:217) */
        (void)parse_seqs(j->data + j->chunks[2 * c], j->chunks[2 * c + 1], job_seq, j);
    }
    return NULL;
}

typedef struct {
    count_job *jobs;
    int nthreads, words, k;
    uint32_t partitions;
    uint64_t *sizes; /* unique per partition */
    uint64_t **pkeys, **pcounts;
    int tid;
} reduce_job;

static void *reduce_worker(void *arg) {
    reduce_job *r = (reduce_job *)arg;
    for (uint32_t p = (uint32_t)r->tid; p < r->partitions; p += (uint32_t)r->nthreads) {
        size_t tot = 0;
        for (int t = 0; t < r->nthreads; t++) tot += r->jobs[t].parts[p].n;
        uint64_t *keys = (uint64_t *)malloc((tot ? tot : 1) * 8);
        size_t o = 0;
        for (int t = 0; t < r->nthreads; t++) {
            kvec *v = &r->jobs[t].parts[p];
            if (v->n) memcpy(keys + o, v->v, v->n * 8);
            o += v->n;
            free(v->v);
            v->v = NULL;
        }
        size_t n = tot / (size_t)r->words;
        uint64_t *counts = (uint64_t *)malloc((n ? n : 1) * 8);
        size_t u = sort_rle_u64(keys, n, r->words, r->k, counts);
        r->pkeys[p] = keys;
        r->pcounts[p] = counts;
        r->sizes[p] = u;
    }
    return NULL;
}

int ok_count_fastq(const uint8_t *data, size_t len, int k, int canonical, uint32_t partitions,
                   size_t chunk_size, int threads, ok_table *out) {
    enc_init();
    memset(out, 0, sizeof *out);
    if (k < 1 || k > 128) return fail("invalid k-mer length: --k must be between 1 and 128");
    if (partitions == 0 || (partitions & (partitions - 1)))
        return fail("invalid partition count: must be a power of 2");
    if (threads < 1) threads = 1;
    int words = ok_key_words(k);
    int bits = __builtin_ctz(partitions);
    size_t cap = len / (chunk_size ? chunk_size : 1) + 4;
    uint64_t *chunks = (uint64_t *)malloc(cap * 16);
    size_t nch = ok_plain_chunks(data, len, chunk_size, chunks, cap);
    if (nch > cap) { free(chunks); return fail("chunk table overflow"); }
    count_job *jobs = (count_job *)calloc((size_t)threads, sizeof *jobs);
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof *th);
    for (int t = 0; t < threads; t++) {
        count_job *j = &jobs[t];
        j->data = data; j->chunks = chunks; j->n_chunks = nch;
        j->k = k; j->canon = canonical; j->words = words; j->bits = bits;
        j->partitions = partitions; j->tid = t; j->nthreads = threads;
        j->parts = (kvec *)calloc(partitions, sizeof(kvec));
        pthread_create(&th[t], NULL, count_worker, j);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    uint64_t kmers = 0, reads = 0;
    for (int t = 0; t < threads; t++) { kmers += jobs[t].kmers; reads += jobs[t].reads; }

    reduce_job *rj = (reduce_job *)calloc((size_t)threads, sizeof *rj);
    uint64_t *sizes = (uint64_t *)calloc(partitions, 8);
    uint64_t **pkeys = (uint64_t **)calloc(partitions, sizeof(uint64_t *));
    uint64_t **pcounts = (uint64_t **)calloc(partitions, sizeof(uint64_t *));
    for (int t = 0; t < threads; t++) {
        rj[t] = (reduce_job){jobs, threads, words, k, partitions, sizes, pkeys, pcounts, t};
        pthread_create(&th[t], NULL, reduce_worker, &rj[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);

    uint64_t total = 0;
    for (uint32_t p = 0; p < partitions; p++) total += sizes[p];
    out->k = k; out->canonical = canonical; out->words = words; out->partitions = partitions;
    out->part_off = (uint64_t *)malloc(((size_t)partitions + 1) * 8);
    out->keys = (uint64_t *)malloc((total ? total : 1) * 8 * (size_t)words);
    out->counts = (uint64_t *)malloc((total ? total : 1) * 8);
    uint64_t o = 0;
    for (uint32_t p = 0; p < partitions; p++) {
        out->part_off[p] = o;
        memcpy(out->keys + o * words, pkeys[p], sizes[p] * 8 * (size_t)words);
        memcpy(out->counts + o, pcounts[p], sizes[p] * 8);
        o += sizes[p];
        free(pkeys[p]);
        free(pcounts[p]);
    }
    out->part_off[partitions] = o;
    out->n_unique = total;
    out->n_kmers = kmers;
    out->n_reads = reads;
    for (int t = 0; t < threads; t++) free(jobs[t].parts);
    free(jobs); free(th); free(rj); free(sizes); free(pkeys); free(pcounts); free(chunks);
    return 0;
}

/* ---- the reference's count hot path, for the CPU baseline ---------------
 * ok_count_fastq above is the checker: per-thread partition vectors, then
 * one sort of each whole partition.  ok_count_ref restates the reference's
 * own count loop for one-word keys (k <= 31) so that the CPU baseline times
 * the reference's algorithm:
 *   - workers take record-aligned chunks in turn (the bounded channel of
 *     count/mod.rs:512-540), parse them (count/parser.rs:19-65) and roll
 *     k-mers inline (count/kmer.rs:203-268);
 *   - each k-mer goes to a 64-entry write-combining slot of its partition,
 *     which is appended to the partition's local vector when full
 *     (worker.rs:15, 56-175);
 *   - at the end each worker drains its slots and flushes every partition
 *     as one run (worker.rs:273-449): radix_sort_u64_kmer (sorter.rs:67-93:
 *     MSD byte radix from byte ceil(2k/8) - 1 down for n >= 131072 with a
 *     comparison-sort leaf at <= 1024, else LSD over ceil(2k/8) bytes, a
 *     comparison sort below 1024), RLE with u32 saturation (sorter.rs:514-530),
 *     DVLE encode (run_block.rs:9-35).
 * The reference's AVX-512 histogram and its zstd step (not used in fast
 * mode) are left out.  Runs stay per worker, as the reference writes them;
 * with `out` the runs are merged into a table for the tests. */
#define REF_WCOM 64
#define REF_MSD_THRESHOLD 131072
#define REF_MSD_LEAF 1024

static void ref_isort(uint64_t *d, size_t n) {
    for (size_t i = 1; i < n; i++) {
        uint64_t v = d[i];
        size_t j = i;
        while (j > 0 && d[j - 1] > v) { d[j] = d[j - 1]; j--; }
        d[j] = v;
    }
}

/* comparison sort of a small array: median-of-three quicksort, insertion
 * sort below 24 elements */
static void ref_qsort(uint64_t *d, size_t n) {
    while (n > 24) {
        uint64_t a = d[0], b = d[n / 2], c = d[n - 1];
        uint64_t piv = a < b ? (b < c ? b : (a < c ? c : a)) : (a < c ? a : (b < c ? c : b));
        size_t i = 0, j = n - 1;
        for (;;) {
            while (d[i] < piv) i++;
            while (d[j] > piv) j--;
            if (i >= j) break;
            uint64_t t = d[i]; d[i] = d[j]; d[j] = t;
            i++; j--;
        }
        /* recurse into the smaller side */
        if (j + 1 < n - j - 1) { ref_qsort(d, j + 1); d += j + 1; n -= j + 1; }
        else { ref_qsort(d + j + 1, n - j - 1); n = j + 1; }
    }
    ref_isort(d, n);
}

static void ref_msd(uint64_t *d, size_t n, int byte_hi, int byte_lo, uint64_t *tmp) {
    if (n <= REF_MSD_LEAF) { ref_qsort(d, n); return; }
    const unsigned sh = (unsigned)byte_hi * 8u;
    size_t hist[256] = {0}, pre[256];
    for (size_t i = 0; i < n; i++) hist[(d[i] >> sh) & 255]++;
    size_t acc = 0;
    for (int b = 0; b < 256; b++) { pre[b] = acc; acc += hist[b]; }
    for (size_t i = 0; i < n; i++) tmp[pre[(d[i] >> sh) & 255]++] = d[i];
    memcpy(d, tmp, n * 8);
    if (byte_hi > byte_lo) {
        size_t off = 0;
        for (int b = 0; b < 256; b++) {
            if (hist[b] > 1) ref_msd(d + off, hist[b], byte_hi - 1, byte_lo, tmp);
            off += hist[b];
        }
    }
}

static void ref_lsd(uint64_t *d, size_t n, int passes, uint64_t *tmp) {
    if (n < 2) return;
    if (n < 1024) { ref_qsort(d, n); return; }
    uint64_t *src = d, *dst = tmp;
    for (int p = 0; p < passes; p++) {
        const unsigned sh = (unsigned)p * 8u;
        size_t hist[256] = {0};
        for (size_t i = 0; i < n; i++) hist[(src[i] >> sh) & 255]++;
        size_t acc = 0;
        for (int b = 0; b < 256; b++) { size_t c = hist[b]; hist[b] = acc; acc += c; }
        for (size_t i = 0; i < n; i++) dst[hist[(src[i] >> sh) & 255]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    if (src != d) memcpy(d, src, n * 8);
}

typedef struct {
    const uint8_t *data;
    const uint64_t *chunks;
    size_t n_chunks;
    size_t *next;          /* shared chunk cursor */
    int k, canon, bits, keep;
    uint32_t P;
    uint64_t **bk;         /* per-partition local vectors */
    size_t *bn, *bcap;
    uint64_t *wc;          /* P x REF_WCOM staging slots */
    uint8_t *wl;
    uint64_t kmers, reads, rows, bytes;
    uint64_t **rk, **rc;   /* kept runs (tests) */
    size_t *rn;
} ref_worker;

static void ref_append(ref_worker *w, uint32_t b, const uint64_t *src, size_t cnt) {
    if (w->bn[b] + cnt > w->bcap[b]) {
        size_t nc = w->bcap[b] ? w->bcap[b] * 2 : 512;
        while (nc < w->bn[b] + cnt) nc *= 2;
        w->bk[b] = (uint64_t *)realloc(w->bk[b], nc * 8);
        w->bcap[b] = nc;
    }
    memcpy(w->bk[b] + w->bn[b], src, cnt * 8);
    w->bn[b] += cnt;
}

static void *ref_worker_main(void *arg) {
    ref_worker *w = (ref_worker *)arg;
    const int k = w->k;
    const uint64_t mask = (1ull << (2 * k)) - 1;
    const unsigned rsh = 2u * (unsigned)(k - 1);
    const unsigned bsh = 64u - (unsigned)w->bits;
    for (;;) {
        const size_t c = __atomic_fetch_add(w->next, 1, __ATOMIC_RELAXED);
        if (c >= w->n_chunks) break;
        const uint8_t *d = w->data + w->chunks[2 * c];
        const size_t len = w->chunks[2 * c + 1];
        if (len == 0 || d[0] != '@') continue;  /* worker.rs:217 */
        int state = 0;
        size_t s0 = 0;
        for (size_t i = 0; i <= len; i++) {
            /* the end of the chunk closes an unterminated sequence line */
            if (i < len && d[i] != '\n') continue;
            if (state == 1 && i > s0) {
                w->reads++;
                uint64_t fwd = 0, rev = 0;
                uint32_t valid = 0;
                for (size_t q = s0; q < i; q++) {
                    const uint8_t e = ENC[d[q]];
                    if (e == 255) { fwd = rev = 0; valid = 0; continue; }
                    fwd = ((fwd << 2) | e) & mask;
                    rev = (rev >> 2) | ((uint64_t)(e ^ 3) << rsh);
                    if (++valid >= (uint32_t)k) {
                        const uint64_t key = (w->canon && rev < fwd) ? rev : fwd;
                        const uint32_t b = w->bits ? (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> bsh) : 0u;
                        uint8_t l = w->wl[b];
                        w->wc[(size_t)b * REF_WCOM + l] = key;
                        if (++l == REF_WCOM) { ref_append(w, b, w->wc + (size_t)b * REF_WCOM, REF_WCOM); l = 0; }
                        w->wl[b] = l;
                        w->kmers++;
                    }
                }
            }
            if (i == len) break;
            if (state == 0) { state = 1; s0 = i + 1; }
            else if (state == 1) state = 2;
            else if (state == 2) state = 3;
            else state = 0;
        }
    }
    /* flush_all: drain each slot, sort, RLE, encode */
    const int passes = (2 * k + 7) / 8 < 8 ? (2 * k + 7) / 8 : 8;
    uint64_t *tmp = NULL;
    size_t tcap = 0;
    uint8_t *enc = NULL;
    size_t ecap = 0;
    uint32_t *cnt = NULL;
    size_t ccap = 0;
    for (uint32_t b = 0; b < w->P; b++) {
        if (w->wl[b]) { ref_append(w, b, w->wc + (size_t)b * REF_WCOM, w->wl[b]); w->wl[b] = 0; }
        const size_t n = w->bn[b];
        if (!n) continue;
        uint64_t *keys = w->bk[b];
        if (n > tcap) { free(tmp); tcap = n; tmp = (uint64_t *)malloc(tcap * 8); }
        if (n >= REF_MSD_THRESHOLD) ref_msd(keys, n, passes - 1, 0, tmp);
        else ref_lsd(keys, n, passes, tmp);
        if (n > ccap) { free(cnt); ccap = n; cnt = (uint32_t *)malloc(ccap * 4); }
        size_t u = 0;
        uint32_t run = 1;
        for (size_t i = 1; i <= n; i++) {
            if (i < n && keys[i] == keys[i - 1]) { if (run != 0xffffffffu) run++; continue; }
            keys[u] = keys[i - 1];
            cnt[u++] = run;
            run = 1;
        }
        if (n * 15 + 16 > ecap) { free(enc); ecap = n * 15 + 16; enc = (uint8_t *)malloc(ecap); }
        w->bytes += ok_dvle_encode(keys, cnt, u, 1, enc);
        w->rows += u;
        if (w->keep) {
            w->rk[b] = (uint64_t *)malloc(u * 8);
            w->rc[b] = (uint64_t *)malloc(u * 8);
            memcpy(w->rk[b], keys, u * 8);
            for (size_t i = 0; i < u; i++) w->rc[b][i] = cnt[i];
            w->rn[b] = u;
        }
        free(w->bk[b]);
        w->bk[b] = NULL;
        w->bn[b] = w->bcap[b] = 0;
    }
    free(tmp);
    free(enc);
    free(cnt);
    return NULL;
}

typedef struct { uint64_t key, count; } ref_pair;
static int ref_pair_cmp(const void *a, const void *b) {
    const uint64_t x = ((const ref_pair *)a)->key, y = ((const ref_pair *)b)->key;
    return x < y ? -1 : (x > y);
}

int ok_count_ref(const uint8_t *data, size_t len, int k, int canonical, uint32_t partitions, size_t chunk_size,
                 int threads, ok_table *out, ok_ref_stats *st) {
    enc_init();
    if (st) memset(st, 0, sizeof *st);
    if (out) memset(out, 0, sizeof *out);
    if (k < 1 || k > 31) return fail("ok_count_ref: one-word keys only (k between 1 and 31)");
    if (partitions == 0 || (partitions & (partitions - 1)))
        return fail("invalid partition count: must be a power of 2");
    if (threads < 1) threads = 1;
    const size_t cap = len / (chunk_size ? chunk_size : 1) + 4;
    uint64_t *chunks = (uint64_t *)malloc(cap * 16);
    const size_t nch = ok_plain_chunks(data, len, chunk_size, chunks, cap);
    if (nch > cap) { free(chunks); return fail("chunk table overflow"); }
    size_t next = 0;
    ref_worker *ws = (ref_worker *)calloc((size_t)threads, sizeof *ws);
    pthread_t *th = (pthread_t *)calloc((size_t)threads, sizeof *th);
    for (int t = 0; t < threads; t++) {
        ref_worker *w = &ws[t];
        w->data = data; w->chunks = chunks; w->n_chunks = nch; w->next = &next;
        w->k = k; w->canon = canonical; w->bits = __builtin_ctz(partitions); w->P = partitions;
        w->keep = out != NULL;
        w->bk = (uint64_t **)calloc(partitions, sizeof(uint64_t *));
        w->bn = (size_t *)calloc(partitions, sizeof(size_t));
        w->bcap = (size_t *)calloc(partitions, sizeof(size_t));
        w->wc = (uint64_t *)malloc((size_t)partitions * REF_WCOM * 8);
        w->wl = (uint8_t *)calloc(partitions, 1);
        if (out) {
            w->rk = (uint64_t **)calloc(partitions, sizeof(uint64_t *));
            w->rc = (uint64_t **)calloc(partitions, sizeof(uint64_t *));
            w->rn = (size_t *)calloc(partitions, sizeof(size_t));
        }
        pthread_create(&th[t], NULL, ref_worker_main, w);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    for (int t = 0; t < threads; t++) {
        if (st) {
            st->reads += ws[t].reads;
            st->kmers += ws[t].kmers;
            st->run_rows += ws[t].rows;
            st->encoded_bytes += ws[t].bytes;
        }
    }
    if (out) {
        /* tests: merge the workers' runs of each partition into totals */
        out->k = k; out->canonical = canonical; out->words = 1; out->partitions = partitions;
        out->part_off = (uint64_t *)malloc(((size_t)partitions + 1) * 8);
        size_t total = 0;
        for (int t = 0; t < threads; t++)
            for (uint32_t p = 0; p < partitions; p++) total += ws[t].rn[p];
        out->keys = (uint64_t *)malloc((total ? total : 1) * 8);
        out->counts = (uint64_t *)malloc((total ? total : 1) * 8);
        ref_pair *pairs = (ref_pair *)malloc((total ? total : 1) * sizeof(ref_pair));
        uint64_t o = 0;
        for (uint32_t p = 0; p < partitions; p++) {
            size_t m = 0;
            for (int t = 0; t < threads; t++)
                for (size_t i = 0; i < ws[t].rn[p]; i++) pairs[m++] = (ref_pair){ws[t].rk[p][i], ws[t].rc[p][i]};
            qsort(pairs, m, sizeof *pairs, ref_pair_cmp);
            out->part_off[p] = o;
            for (size_t i = 0; i < m; i++) {
                if (i && pairs[i].key == pairs[i - 1].key) { out->counts[o - 1] += pairs[i].count; continue; }
                out->keys[o] = pairs[i].key;
                out->counts[o++] = pairs[i].count;
            }
        }
        out->part_off[partitions] = o;
        out->n_unique = o;
        free(pairs);
        for (int t = 0; t < threads; t++) { out->n_kmers += ws[t].kmers; out->n_reads += ws[t].reads; }
        for (int t = 0; t < threads; t++) {
            for (uint32_t p = 0; p < partitions; p++) { free(ws[t].rk[p]); free(ws[t].rc[p]); }
            free(ws[t].rk); free(ws[t].rc); free(ws[t].rn);
        }
    }
    for (int t = 0; t < threads; t++) {
        free(ws[t].bk); free(ws[t].bn); free(ws[t].bcap); free(ws[t].wc); free(ws[t].wl);
    }
    free(ws); free(th); free(chunks);
    return 0;
}

void ok_table_free(ok_table *t) {
    free(t->part_off); free(t->keys); free(t->counts);
    memset(t, 0, sizeof *t);
}

/* ---- .kite writer: count/writer.rs:76-495 and src/format ---------------- */
static void le16(uint8_t *p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
static void le32(uint8_t *p, uint32_t v) { for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i)); }
static void le64(uint8_t *p, uint64_t v) { for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i)); }
static uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static uint32_t rd32(const uint8_t *p) { uint32_t v = 0; for (int i = 3; i >= 0; i--) v = (v << 8) | p[i]; return v; }
static uint64_t rd64(const uint8_t *p) { uint64_t v = 0; for (int i = 7; i >= 0; i--) v = (v << 8) | p[i]; return v; }

typedef struct {
    uint64_t rel_or_off, rows, enc, raw;
    uint32_t codec;
    uint64_t kmin[4], kmax[4]; /* (lo128, hi128) as 4 words */
} chunk_meta;

typedef struct {
    uint8_t *buf; size_t len, cap;
    chunk_meta *pend; size_t npend, cappend;
    chunk_meta *comm; size_t ncomm, capcomm;
} pstage;

typedef struct {
    FILE *f;
    int regular;
    uint64_t next_offset;
    size_t S, high, low, staged;
    uint32_t P;
    pstage *st;
} kwriter;

static void meta_push(chunk_meta **a, size_t *n, size_t *cap, const chunk_meta *m) {
    if (*n == *cap) { *cap = *cap ? *cap * 2 : 4; *a = (chunk_meta *)realloc(*a, *cap * sizeof **a); }
    (*a)[(*n)++] = *m;
}

static uint64_t claim(kwriter *w, const uint8_t *data, size_t len) {
    uint64_t off = w->next_offset;
    if (w->regular && len) {
        fseeko(w->f, (off_t)off, SEEK_SET);
        fwrite(data, 1, len, w->f);
    }
    w->next_offset += len;
    return off;
}

static void flush_part(kwriter *w, uint32_t p) {
    pstage *s = &w->st[p];
    if (s->len == 0) return;
    uint64_t off = claim(w, s->buf, s->len);
    w->staged -= s->len;
    s->len = 0;
    for (size_t i = 0; i < s->npend; i++) {
        chunk_meta m = s->pend[i];
        m.rel_or_off += off;
        meta_push(&s->comm, &s->ncomm, &s->capcomm, &m);
    }
    s->npend = 0;
}

static int drain_one(kwriter *w) {
    size_t best = 0;
    int64_t bp = -1;
    for (uint32_t p = 0; p < w->P; p++)
        if (w->st[p].len > best) { best = w->st[p].len; bp = p; }
    if (bp < 0) return 0;
    flush_part(w, (uint32_t)bp);
    return 1;
}

static void writer_append(kwriter *w, uint32_t p, const chunk_meta *m0, const uint8_t *data, size_t len) {
    pstage *s = &w->st[p];
    chunk_meta m = *m0;
    m.enc = len;
    if (s->len && s->len + len > w->S) flush_part(w, p);
    if (len >= w->S) {
        m.rel_or_off = claim(w, data, len);
        meta_push(&s->comm, &s->ncomm, &s->capcomm, &m);
        return;
    }
    if (s->len + len > s->cap) {
        s->cap = (s->len + len) * 2;
        s->buf = (uint8_t *)realloc(s->buf, s->cap);
    }
    m.rel_or_off = s->len;
    memcpy(s->buf + s->len, data, len);
    s->len += len;
    meta_push(&s->pend, &s->npend, &s->cappend, &m);
    w->staged += len;
    if (s->len >= w->S) flush_part(w, p);
    if (w->staged >= w->high) {
        int attempts = 0;
        while (w->staged > w->low && attempts < 4) {
            if (!drain_one(w)) break;
            attempts++;
        }
    }
}

static int cmp_meta_off(const void *a, const void *b) {
    const chunk_meta *x = (const chunk_meta *)a, *y = (const chunk_meta *)b;
    return x->rel_or_off < y->rel_or_off ? -1 : x->rel_or_off > y->rel_or_off;
}

static void header_bytes(uint8_t *h, int k, int canonical, uint32_t P, uint16_t key_width,
                         uint16_t codec_mask, uint64_t cio, uint64_t pto, uint64_t fo, uint64_t tc,
                         uint64_t tr, int has_min, uint32_t min, int has_max, uint32_t max) {
    memset(h, 0, 128);
    memcpy(h, "KX2!", 4);
    le16(h + 4, 1);
    h[6] = (uint8_t)k;
    h[7] = (uint8_t)(canonical ? 1 : 0);
    le32(h + 8, P);
    le16(h + 12, key_width);
    le16(h + 14, codec_mask);
    le32(h + 16, 0);
    le64(h + 20, cio);
    le64(h + 28, pto);
    le64(h + 36, fo);
    le64(h + 44, tc);
    le64(h + 52, tr);
    if (has_min || has_max) {
        uint8_t *r = h + 60;
        memcpy(r, "KMT1", 4);
        le32(r + 4, (has_min ? 1u : 0u) | (has_max ? 2u : 0u));
        le32(r + 8, has_min ? min : 0);
        le32(r + 12, has_max ? max : 0);
    }
}

static int key_width_for_words(int words) { return words == 1 ? 64 : (words == 2 ? 128 : 256); }

int ok_kite_write_table(const char *path, const ok_table *t, uint16_t codec_mask, int has_min,
                        uint32_t min, int has_max, uint32_t max, size_t staging_bytes,
                        int write_lanes) {
    kwriter w;
    memset(&w, 0, sizeof w);
    w.f = fopen(path, "w+b");
    if (!w.f) return fail("cannot create %s: %s", path, strerror(errno));
    struct stat stt;
    w.regular = (fstat(fileno(w.f), &stt) == 0) ? S_ISREG(stt.st_mode) : 1;
    w.next_offset = 128;
    w.S = staging_bytes;
    size_t hw = staging_bytes * (size_t)write_lanes * 4;
    w.high = hw > staging_bytes ? hw : staging_bytes;
    w.low = w.high / 2;
    w.P = t->partitions;
    w.st = (pstage *)calloc(t->partitions, sizeof(pstage));
    uint8_t hdr[128];
    header_bytes(hdr, t->k, t->canonical, t->partitions, (uint16_t)key_width_for_words(t->words),
                 codec_mask, 0, 0, 0, 0, 0, has_min, min, has_max, max);
    if (w.regular) fwrite(hdr, 1, 128, w.f);
    int words = t->words;
    size_t maxn = 0;
    for (uint32_t p = 0; p < t->partitions; p++) {
        size_t n = t->part_off[p + 1] - t->part_off[p];
        if (n > maxn) maxn = n;
    }
    uint8_t *enc = (uint8_t *)malloc(maxn * 48 + 64);
    uint32_t *c32 = (uint32_t *)malloc((maxn + 1) * 4);
    for (uint32_t p = 0; p < t->partitions; p++) {
        uint64_t a = t->part_off[p], n = t->part_off[p + 1] - a;
        if (!n) continue;
        for (uint64_t i = 0; i < n; i++) {
            if (t->counts[a + i] > 0xffffffffull) {
                free(enc); free(c32);
                fclose(w.f);
                return fail("merged count exceeds u32");
            }
            c32[i] = (uint32_t)t->counts[a + i];
        }
        size_t len = ok_dvle_encode(t->keys + a * words, c32, n, words, enc);
        chunk_meta m;
        memset(&m, 0, sizeof m);
        m.rows = n;
        m.raw = len;
        m.codec = words == 4 ? 1u : 2u;
        const uint64_t *f = t->keys + a * words, *l = t->keys + (a + n - 1) * words;
        /* split_key (count/kmer.rs:104-106,149-151,195-197; kiteops 850-857):
         * (lo, hi) = (key, 0) for u64/u128, (lo, hi) for u256 */
        for (int i = 0; i < words; i++) { m.kmin[i] = f[i]; m.kmax[i] = l[i]; }
        writer_append(&w, p, &m, enc, len);
    }
    free(enc); free(c32);
    /* finalize (writer.rs:350-495) */
    uint64_t total_chunks = 0, total_rows = 0;
    for (uint32_t p = 0; p < t->partitions; p++) {
        flush_part(&w, p);
        qsort(w.st[p].comm, w.st[p].ncomm, sizeof(chunk_meta), cmp_meta_off);
        total_chunks += w.st[p].ncomm;
        for (size_t i = 0; i < w.st[p].ncomm; i++) total_rows += w.st[p].comm[i].rows;
    }
    uint64_t data_end = w.next_offset;
    uint64_t cio = data_end, pto = cio + total_chunks * 128, fo = pto + (uint64_t)t->partitions * 40;
    if (w.regular) {
        fseeko(w.f, (off_t)data_end, SEEK_SET);
        uint8_t e[128];
        for (uint32_t p = 0; p < t->partitions; p++)
            for (size_t i = 0; i < w.st[p].ncomm; i++) {
                chunk_meta *m = &w.st[p].comm[i];
                memset(e, 0, 128);
                le32(e, p);
                le32(e + 4, m->codec);
                le64(e + 8, m->rel_or_off);
                le64(e + 16, m->enc);
                le64(e + 24, m->raw);
                le64(e + 32, m->rows);
                le64(e + 40, m->kmin[0]); le64(e + 48, m->kmin[1]);
                le64(e + 56, m->kmin[2]); le64(e + 64, m->kmin[3]);
                le64(e + 72, m->kmax[0]); le64(e + 80, m->kmax[1]);
                le64(e + 88, m->kmax[2]); le64(e + 96, m->kmax[3]);
                fwrite(e, 1, 128, w.f);
            }
        uint64_t first = 0;
        for (uint32_t p = 0; p < t->partitions; p++) {
            uint8_t pe[40];
            uint64_t rows = 0, bytes = 0, dmin = 0;
            for (size_t i = 0; i < w.st[p].ncomm; i++) {
                chunk_meta *m = &w.st[p].comm[i];
                rows += m->rows;
                bytes += m->enc;
                if (i == 0 || m->rel_or_off < dmin) dmin = m->rel_or_off;
            }
            le64(pe, first); le64(pe + 8, w.st[p].ncomm); le64(pe + 16, rows);
            le64(pe + 24, dmin); le64(pe + 32, bytes);
            fwrite(pe, 1, 40, w.f);
            first += w.st[p].ncomm;
        }
        uint8_t ft[24];
        le64(ft, total_chunks); le64(ft + 8, total_rows); memcpy(ft + 16, "KX2FOOT!", 8);
        fwrite(ft, 1, 24, w.f);
        header_bytes(hdr, t->k, t->canonical, t->partitions, (uint16_t)key_width_for_words(words),
                     codec_mask, cio, pto, fo, total_chunks, total_rows, has_min, min, has_max, max);
        fseeko(w.f, 0, SEEK_SET);
        fwrite(hdr, 1, 128, w.f);
        fflush(w.f);
        if (ftruncate(fileno(w.f), (off_t)(fo + 24)) != 0) { /* same length already */ }
    }
    fclose(w.f);
    for (uint32_t p = 0; p < t->partitions; p++) {
        free(w.st[p].buf); free(w.st[p].pend); free(w.st[p].comm);
    }
    free(w.st);
    return 0;
}

/* ---- .kite reader: dump/reader.rs:43-185 -------------------------------- */
typedef struct {
    uint8_t *img;
    size_t len;
    int k, canonical, words;
    uint32_t P;
    uint16_t key_width, codec_mask;
    int has_min, has_max;
    uint32_t min, max;
    uint64_t n_chunks;
    /* per chunk */
    uint32_t *c_part, *c_codec;
    uint64_t *c_off, *c_enc, *c_rows;
    uint64_t *p_first, *p_count;
} kreader;

static void kreader_free(kreader *r) {
    free(r->img); free(r->c_part); free(r->c_codec); free(r->c_off); free(r->c_enc);
    free(r->c_rows); free(r->p_first); free(r->p_count);
    memset(r, 0, sizeof *r);
}

static int kreader_open(const char *path, kreader *r) {
    memset(r, 0, sizeof *r);
    FILE *f = fopen(path, "rb");
    if (!f) return fail("cannot open %s: %s", path, strerror(errno));
    fseeko(f, 0, SEEK_END);
    off_t sz = ftello(f);
    fseeko(f, 0, SEEK_SET);
    r->len = (size_t)sz;
    r->img = (uint8_t *)malloc(r->len ? r->len : 1);
    if (r->len && fread(r->img, 1, r->len, f) != r->len) { fclose(f); kreader_free(r); return fail("short read"); }
    fclose(f);
    const uint8_t *h = r->img;
    if (r->len < 128 + 24) { kreader_free(r); return fail("file too small"); }
    if (memcmp(h, "KX2!", 4)) { kreader_free(r); return fail("invalid magic"); }
    if (rd16(h + 4) != 1) { kreader_free(r); return fail("unsupported version: %u", rd16(h + 4)); }
    r->k = h[6];
    r->canonical = h[7];
    r->P = rd32(h + 8);
    r->key_width = rd16(h + 12);
    r->codec_mask = rd16(h + 14);
    r->words = r->key_width == 128 ? 2 : (r->key_width == 256 ? 4 : 1);
    uint64_t cio = rd64(h + 20), pto = rd64(h + 28), fo = rd64(h + 36);
    if (!memcmp(h + 60, "KMT1", 4)) {
        uint32_t fl = rd32(h + 64);
        r->has_min = (fl & 1) != 0; r->min = rd32(h + 68);
        r->has_max = (fl & 2) != 0; r->max = rd32(h + 72);
    }
    if (fo + 24 > r->len) { kreader_free(r); return fail("footer out of bounds"); }
    if (memcmp(r->img + fo + 16, "KX2FOOT!", 8)) { kreader_free(r); return fail("invalid footer magic"); }
    if (pto + (uint64_t)r->P * 40 > r->len) { kreader_free(r); return fail("partition table out of bounds"); }
    if (pto < cio) { kreader_free(r); return fail("invalid chunk index offsets"); }
    if ((pto - cio) % 128) { kreader_free(r); return fail("chunk index size misaligned"); }
    uint64_t nc = (pto - cio) / 128;
    r->n_chunks = nc;
    r->c_part = (uint32_t *)malloc((nc + 1) * 4);
    r->c_codec = (uint32_t *)malloc((nc + 1) * 4);
    r->c_off = (uint64_t *)malloc((nc + 1) * 8);
    r->c_enc = (uint64_t *)malloc((nc + 1) * 8);
    r->c_rows = (uint64_t *)malloc((nc + 1) * 8);
    for (uint64_t i = 0; i < nc; i++) {
        const uint8_t *e = r->img + cio + i * 128;
        r->c_part[i] = rd32(e);
        r->c_codec[i] = rd32(e + 4);
        r->c_off[i] = rd64(e + 8);
        r->c_enc[i] = rd64(e + 16);
        r->c_rows[i] = rd64(e + 32);
    }
    r->p_first = (uint64_t *)malloc(((size_t)r->P + 1) * 8);
    r->p_count = (uint64_t *)malloc(((size_t)r->P + 1) * 8);
    for (uint32_t p = 0; p < r->P; p++) {
        const uint8_t *e = r->img + pto + (uint64_t)p * 40;
        r->p_first[p] = rd64(e);
        r->p_count[p] = rd64(e + 8);
        if (r->p_first[p] + r->p_count[p] > nc) { kreader_free(r); return fail("partition %u chunk range out of bounds", p); }
        for (uint64_t c = r->p_first[p]; c < r->p_first[p] + r->p_count[p]; c++) {
            if (r->c_enc[c] == 0 || r->c_off[c] + r->c_enc[c] > r->len) { kreader_free(r); return fail("partition %u chunk out of bounds", p); }
            if (r->c_codec[c] & ~7u) { kreader_free(r); return fail("partition %u chunk uses unsupported codec mask %#x", p, r->c_codec[c]); }
            if (r->c_codec[c] & 4u) { kreader_free(r); return fail("zstd chunks are not decoded by this oracle"); }
        }
    }
    return 0;
}

/* Decode all rows of one partition (kiteops/mod.rs:433-585): keys, u64 counts. */
static int load_partition(const kreader *r, uint32_t p, kvec *keys, uint64_t **counts, size_t *n) {
    int words = r->words;
    size_t tot = 0;
    for (uint64_t c = r->p_first[p]; c < r->p_first[p] + r->p_count[p]; c++) tot += r->c_rows[c];
    keys->n = 0;
    if (keys->cap < tot * words + 4) { keys->cap = tot * words + 4; keys->v = (uint64_t *)realloc(keys->v, keys->cap * 8); }
    *counts = (uint64_t *)realloc(*counts, (tot + 1) * 8);
    size_t o = 0;
    for (uint64_t c = r->p_first[p]; c < r->p_first[p] + r->p_count[p]; c++) {
        const uint8_t *d = r->img + r->c_off[c];
        size_t len = r->c_enc[c], rows = r->c_rows[c];
        if (words == 4 || !(r->c_codec[c] & 2u)) {
            size_t es = words == 4 ? 36 : (words == 2 ? 20 : 12);
            for (size_t i = 0; i < rows; i++) {
                if ((i + 1) * es > len) return fail("raw row payload truncated at row %zu", i);
                for (int w = 0; w < words; w++) keys->v[(o + i) * words + w] = rd64(d + i * es + 8 * w);
                (*counts)[o + i] = rd32(d + i * es + 8 * words);
            }
        } else {
            if (rows == 0) continue;
            size_t head = 8 * (size_t)words + 4;
            if (len < head) return fail("DVLE payload too small");
            uint64_t plo = rd64(d), phi = words == 2 ? rd64(d + 8) : 0;
            keys->v[o * words] = plo;
            if (words == 2) keys->v[o * words + 1] = phi;
            (*counts)[o] = rd32(d + 8 * words);
            size_t pos = head;
            for (size_t i = 1; i < rows; i++) {
                uint64_t dlo, dhi, clo, chi;
                size_t u = get_varint(d + pos, len - pos, words == 2 ? 19 : 10, &dlo, &dhi);
                if (!u) return fail("invalid DVLE delta");
                pos += u;
                size_t v = get_varint(d + pos, len - pos, 10, &clo, &chi);
                if (!v) return fail("invalid DVLE count");
                pos += v;
                if (words == 1) plo += dlo;
                else {
                    u128 pv = ((u128)phi << 64) | plo;
                    pv += ((u128)dhi << 64) | dlo;
                    plo = (uint64_t)pv;
                    phi = (uint64_t)(pv >> 64);
                }
                keys->v[(o + i) * words] = plo;
                if (words == 2) keys->v[(o + i) * words + 1] = phi;
                (*counts)[o + i] = clo;
            }
        }
        o += rows;
    }
    /* sort by key and total equal keys with u64 saturation */
    size_t nrows = o;
    if (nrows > 1) {
        /* sort (key, count) pairs: sort index by radix via a packed copy */
        int W = words + 1;
        uint64_t *pk = (uint64_t *)malloc(nrows * 8 * (size_t)W);
        for (size_t i = 0; i < nrows; i++) {
            memcpy(pk + i * W, keys->v + i * words, 8 * (size_t)words);
            pk[i * W + words] = (*counts)[i];
        }
        /* LSD radix over the key bytes only; stable, count rides along */
        size_t nbytes = (size_t)8 * words;
        uint64_t *tmp = (uint64_t *)malloc(nrows * 8 * (size_t)W);
        uint64_t *src = pk, *dst = tmp;
        size_t cnt[256];
        for (size_t b = 0; b < nbytes; b++) {
            int w = (int)(b / 8);
            unsigned sh = (unsigned)(b % 8) * 8u;
            memset(cnt, 0, sizeof cnt);
            for (size_t i = 0; i < nrows; i++) cnt[(src[i * W + w] >> sh) & 255]++;
            int single = 0;
            for (int dd = 0; dd < 256; dd++) if (cnt[dd] == nrows) single = 1;
            if (single) continue;
            size_t sum = 0;
            for (int dd = 0; dd < 256; dd++) { size_t c2 = cnt[dd]; cnt[dd] = sum; sum += c2; }
            for (size_t i = 0; i < nrows; i++) {
                size_t oo = cnt[(src[i * W + w] >> sh) & 255]++;
                memcpy(dst + oo * W, src + i * W, 8 * (size_t)W);
            }
            uint64_t *t = src; src = dst; dst = t;
        }
        size_t u = 0;
        for (size_t i = 0; i < nrows; i++) {
            const uint64_t *e = src + i * W;
            if (u && key_cmp(keys->v + (u - 1) * words, e, words) == 0) {
                uint64_t s = (*counts)[u - 1] + e[words];
                (*counts)[u - 1] = s < e[words] ? ~0ull : s;
            } else {
                memcpy(keys->v + u * words, e, 8 * (size_t)words);
                (*counts)[u] = e[words];
                u++;
            }
        }
        nrows = u;
        free(pk); free(tmp);
    }
    *n = nrows;
    return 0;
}

static int passes(uint64_t c, int has_min, uint32_t min, int has_max, uint32_t max) {
    if (has_min && c < min) return 0;
    if (has_max && c > max) return 0;
    return 1;
}

int ok_kite_load(const char *path, ok_table *out, int *has_min, uint32_t *min, int *has_max,
                 uint32_t *max) {
    kreader r;
    memset(out, 0, sizeof *out);
    if (kreader_open(path, &r)) return -1;
    out->k = r.k; out->canonical = r.canonical; out->words = r.words; out->partitions = r.P;
    out->part_off = (uint64_t *)malloc(((size_t)r.P + 1) * 8);
    kvec kv = {0};
    uint64_t *cnt = NULL;
    size_t cap = 1024, tot = 0;
    out->keys = (uint64_t *)malloc(cap * 8 * (size_t)r.words);
    out->counts = (uint64_t *)malloc(cap * 8);
    for (uint32_t p = 0; p < r.P; p++) {
        size_t n;
        out->part_off[p] = tot;
        if (load_partition(&r, p, &kv, &cnt, &n)) { kreader_free(&r); free(kv.v); free(cnt); return -1; }
        if (tot + n > cap) {
            while (tot + n > cap) cap *= 2;
            out->keys = (uint64_t *)realloc(out->keys, cap * 8 * (size_t)r.words);
            out->counts = (uint64_t *)realloc(out->counts, cap * 8);
        }
        memcpy(out->keys + tot * r.words, kv.v, n * 8 * (size_t)r.words);
        memcpy(out->counts + tot, cnt, n * 8);
        tot += n;
    }
    out->part_off[r.P] = tot;
    out->n_unique = tot;
    if (has_min) *has_min = r.has_min;
    if (min) *min = r.min;
    if (has_max) *has_max = r.has_max;
    if (max) *max = r.max;
    free(kv.v); free(cnt);
    kreader_free(&r);
    return 0;
}

/* ---- dump formatting: dump/decoder.rs:27-117 ----------------------------- */
static const char DNA[4] = {'A', 'C', 'G', 'T'};

void ok_decode_key(const uint64_t *key, int words, int k, char *buf) {
    if (words <= 2) {
        u128 v = (u128)key[0] | (words == 2 ? ((u128)key[1] << 64) : 0);
        for (int i = 0; i < k; i++) buf[i] = DNA[(v >> (2 * (k - 1 - i))) & 3];
        return;
    }
    /* decode_kmer_u256 takes min(k,64) = 64 bases MSB-first from all 128 bits
     * of hi, then the low 2(k-64) bits of lo as the remaining k-64 bases. */
    u128 lo = (u128)key[0] | ((u128)key[1] << 64);
    u128 hi = (u128)key[2] | ((u128)key[3] << 64);
    int hb = k < 64 ? k : 64, lb = k - hb;
    u128 hs = hi << (128 - 2 * hb);
    for (int i = 0; i < hb; i++) buf[i] = DNA[(hs >> (126 - 2 * i)) & 3];
    for (int j = 0; j < lb; j++) buf[hb + j] = DNA[(lo >> (2 * (lb - 1 - j))) & 3];
}

int ok_kite_dump(const char *path, const char *out_path) {
    kreader r;
    if (kreader_open(path, &r)) return -1;
    FILE *o = strcmp(out_path, "-") ? fopen(out_path, "wb") : stdout;
    if (!o) { kreader_free(&r); return fail("cannot create %s", out_path); }
    kvec kv = {0};
    uint64_t *cnt = NULL;
    char line[256];
    for (uint32_t p = 0; p < r.P; p++) {
        size_t n;
        if (load_partition(&r, p, &kv, &cnt, &n)) { if (o != stdout) fclose(o); kreader_free(&r); return -1; }
        for (size_t i = 0; i < n; i++) {
            uint64_t c = cnt[i] > 0xffffffffull ? 0xffffffffull : cnt[i]; /* u32 saturating sum */
            if (!passes(c, r.has_min, r.min, r.has_max, r.max)) continue;
            ok_decode_key(kv.v + i * r.words, r.words, r.k, line);
            int l = r.k;
            l += sprintf(line + l, "\t%llu\n", (unsigned long long)c);
            fwrite(line, 1, (size_t)l, o);
        }
    }
    if (o != stdout) fclose(o); else fflush(o);
    free(kv.v); free(cnt);
    kreader_free(&r);
    return 0;
}

/* ---- labels: kiteops/mod.rs:123-165 -------------------------------------- */
static void infer_label(const char *path, int index, char *out, size_t cap) {
    const char *base = strrchr(path, '/');
    base = base ? base + 1 : path;
    snprintf(out, cap, "%s", base);
    static const char *suf[] = {".fastq.gz", ".fq.gz", ".fastq", ".fq", ".kite", ".tsv", ".txt", ".gz"};
    for (int i = 0; i < 8; i++) {
        size_t L = strlen(out), s = strlen(suf[i]);
        if (L >= s && !strcmp(out + L - s, suf[i])) { out[L - s] = 0; break; }
    }
    if (!out[0]) snprintf(out, cap, "sample%d", index + 1);
}

/* ---- histo: histo/mod.rs:16-176 ------------------------------------------ */
typedef struct { uint64_t count, freq; } bin;
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int ok_kite_histo(const char *const *paths, int n, const char *const *labels, const char *out_path) {
    uint64_t **vals = (uint64_t **)calloc((size_t)n, sizeof(uint64_t *));
    size_t *nv = (size_t *)calloc((size_t)n, sizeof(size_t));
    for (int i = 0; i < n; i++) {
        ok_table t;
        int hmin, hmax;
        uint32_t mn, mx;
        if (ok_kite_load(paths[i], &t, &hmin, &mn, &hmax, &mx)) return -1;
        vals[i] = (uint64_t *)malloc((t.n_unique + 1) * 8);
        size_t m = 0;
        for (uint64_t j = 0; j < t.n_unique; j++)
            if (passes(t.counts[j], hmin, mn, hmax, mx)) vals[i][m++] = t.counts[j];
        qsort(vals[i], m, 8, cmp_u64);
        nv[i] = m;
        ok_table_free(&t);
    }
    FILE *o = strcmp(out_path, "-") ? fopen(out_path, "wb") : stdout;
    if (!o) return fail("cannot create %s", out_path);
    fprintf(o, "count");
    char lab[4096];
    for (int i = 0; i < n; i++) {
        if (labels) fprintf(o, "\t%s", labels[i]);
        else { infer_label(paths[i], i, lab, sizeof lab); fprintf(o, "\t%s", lab); }
    }
    fprintf(o, "\n");
    size_t *pos = (size_t *)calloc((size_t)n, sizeof(size_t));
    for (;;) {
        int have = 0;
        uint64_t c = 0;
        for (int i = 0; i < n; i++)
            if (pos[i] < nv[i] && (!have || vals[i][pos[i]] < c)) { c = vals[i][pos[i]]; have = 1; }
        if (!have) break;
        fprintf(o, "%llu", (unsigned long long)c);
        for (int i = 0; i < n; i++) {
            uint64_t f = 0;
            while (pos[i] < nv[i] && vals[i][pos[i]] == c) { f++; pos[i]++; }
            fprintf(o, "\t%llu", (unsigned long long)f);
        }
        fprintf(o, "\n");
    }
    if (o != stdout) fclose(o); else fflush(o);
    for (int i = 0; i < n; i++) free(vals[i]);
    free(vals); free(nv); free(pos);
    return 0;
}

/* ---- merge: merge/mod.rs:22-214, kiteops/mod.rs:587-724, 804-859 --------- */
static int check_compat(const ok_table *a, const ok_table *b, int idx) {
    if (a->k != b->k) return fail("incompatible .kite inputs: input #%d has k=%d", idx + 1, b->k);
    if (!!a->canonical != !!b->canonical) return fail("incompatible .kite inputs: canonical differs (input #%d)", idx + 1);
    if (a->partitions != b->partitions) return fail("incompatible .kite inputs: partitions differ (input #%d)", idx + 1);
    if (a->words != b->words) return fail("incompatible .kite inputs: key_width differs (input #%d)", idx + 1);
    return 0;
}

/* Filter each input by its own header markers, then sorted-union per
 * partition; returns a table of union keys and either summed counts (sum=1,
 * u64 saturating) or the N x n count matrix (sum=0). */
static int union_tables(const char *const *paths, int n, int sum, ok_table *out, uint64_t **matrix,
                        int *all_min, uint32_t *omin, int *all_max, uint32_t *omax) {
    ok_table *t = (ok_table *)calloc((size_t)n, sizeof(ok_table));
    int *hmin = (int *)calloc((size_t)n, sizeof(int)), *hmax = (int *)calloc((size_t)n, sizeof(int));
    uint32_t *mn = (uint32_t *)calloc((size_t)n, 4), *mx = (uint32_t *)calloc((size_t)n, 4);
    for (int i = 0; i < n; i++) {
        if (ok_kite_load(paths[i], &t[i], &hmin[i], &mn[i], &hmax[i], &mx[i])) return -1;
        if (i && check_compat(&t[0], &t[i], i)) return -1;
        /* apply_reader_filter_to_rows (kiteops/mod.rs:103-108) */
        uint64_t o = 0;
        for (uint32_t p = 0; p < t[i].partitions; p++) {
            uint64_t a = t[i].part_off[p], b = t[i].part_off[p + 1];
            t[i].part_off[p] = o;
            for (uint64_t j = a; j < b; j++) {
                if (!passes(t[i].counts[j], hmin[i], mn[i], hmax[i], mx[i])) continue;
                memmove(t[i].keys + o * t[i].words, t[i].keys + j * t[i].words, 8 * (size_t)t[i].words);
                t[i].counts[o++] = t[i].counts[j];
            }
        }
        t[i].part_off[t[i].partitions] = o;
    }
    if (all_min) {
        *all_min = 1; *all_max = 1;
        for (int i = 1; i < n; i++) {
            if (hmin[i] != hmin[0] || (hmin[0] && mn[i] != mn[0])) *all_min = 0;
            if (hmax[i] != hmax[0] || (hmax[0] && mx[i] != mx[0])) *all_max = 0;
        }
        *all_min = *all_min && hmin[0];
        *all_max = *all_max && hmax[0];
        *omin = mn[0];
        *omax = mx[0];
    }
    int words = t[0].words;
    uint32_t P = t[0].partitions;
    uint64_t cap = 0;
    for (int i = 0; i < n; i++) cap += t[i].part_off[P];
    memset(out, 0, sizeof *out);
    out->k = t[0].k; out->canonical = t[0].canonical; out->words = words; out->partitions = P;
    out->part_off = (uint64_t *)malloc(((size_t)P + 1) * 8);
    out->keys = (uint64_t *)malloc((cap + 1) * 8 * (size_t)words);
    out->counts = (uint64_t *)malloc((cap + 1) * 8);
    if (!sum) *matrix = (uint64_t *)malloc((cap + 1) * 8 * (size_t)n);
    uint64_t *idx = (uint64_t *)calloc((size_t)n, 8);
    uint64_t o = 0;
    for (uint32_t p = 0; p < P; p++) {
        out->part_off[p] = o;
        for (int i = 0; i < n; i++) idx[i] = t[i].part_off[p];
        for (;;) {
            int have = -1;
            for (int i = 0; i < n; i++)
                if (idx[i] < t[i].part_off[p + 1] &&
                    (have < 0 || key_cmp(t[i].keys + idx[i] * words, t[have].keys + idx[have] * words, words) < 0))
                    have = i;
            if (have < 0) break;
            const uint64_t *key = t[have].keys + idx[have] * words;
            memcpy(out->keys + o * words, key, 8 * (size_t)words);
            uint64_t s = 0;
            uint64_t keycopy[4];
            memcpy(keycopy, key, 8 * (size_t)words);
            for (int i = 0; i < n; i++) {
                uint64_t c = 0;
                if (idx[i] < t[i].part_off[p + 1] && key_cmp(t[i].keys + idx[i] * words, keycopy, words) == 0)
                    c = t[i].counts[idx[i]++];
                if (!sum) (*matrix)[o * n + i] = c;
                uint64_t ns = s + c;
                s = ns < s ? ~0ull : ns;
            }
            out->counts[o] = s;
            o++;
        }
    }
    out->part_off[P] = o;
    out->n_unique = o;
    for (int i = 0; i < n; i++) ok_table_free(&t[i]);
    free(t); free(hmin); free(hmax); free(mn); free(mx); free(idx);
    return 0;
}

int ok_kite_merge(const char *const *paths, int n, const char *out_path, int write_lanes) {
    ok_table u;
    int amin, amax;
    uint32_t omin, omax;
    if (union_tables(paths, n, 1, &u, NULL, &amin, &omin, &amax, &omax)) return -1;
    int rc = ok_kite_write_table(out_path, &u, 3, amin, omin, amax, omax, 16u << 20, write_lanes);
    ok_table_free(&u);
    return rc;
}

int ok_kite_union(const char *const *paths, int n, uint64_t *rows_out, uint64_t **keys_out,
                  uint64_t **counts_out, int *words_out) {
    ok_table u;
    uint64_t *m = NULL;
    if (union_tables(paths, n, 0, &u, &m, NULL, NULL, NULL, NULL)) return -1;
    *rows_out = u.n_unique;
    *keys_out = u.keys;
    *counts_out = m;
    *words_out = u.words;
    u.keys = NULL;
    ok_table_free(&u);
    return 0;
}

/* ---- command-line front end (used by tests and the CPU baseline) --------- */
#ifdef KITE_ORACLE_MAIN
#include <fcntl.h>
#include <sys/mman.h>
#include <time.h>

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static int usage(void) {
    fprintf(stderr,
            "kite_oracle count IN.fastq OUT.kite K CANON PARTITIONS CHUNK_BYTES THREADS [MIN|-] [MAX|-]\n"
            "kite_oracle count-ref IN.fastq /dev/null K CANON PARTITIONS CHUNK_BYTES THREADS  (k <= 31)\n"
            "kite_oracle dump IN.kite OUT.tsv|-\n"
            "kite_oracle histo OUT.tsv IN.kite...\n"
            "kite_oracle merge OUT.kite LANES IN.kite...\n");
    return 2;
}

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    if (!strcmp(argv[1], "count")) {
        if (argc < 9) return usage();
        int fd = open(argv[2], O_RDONLY);
        if (fd < 0) { perror("open"); return 1; }
        struct stat st;
        fstat(fd, &st);
        size_t len = (size_t)st.st_size;
        const uint8_t *img = len ? (const uint8_t *)mmap(NULL, len, PROT_READ, MAP_PRIVATE, fd, 0) : (const uint8_t *)"";
        ok_table t;
        double t0 = now_s();
        if (ok_count_fastq(img, len, atoi(argv[4]), atoi(argv[5]), (uint32_t)strtoul(argv[6], 0, 10),
                           (size_t)strtoull(argv[7], 0, 10), atoi(argv[8]), &t)) {
            fprintf(stderr, "error: %s\n", ok_last_error());
            return 1;
        }
        double t1 = now_s();
        int hmin = argc > 9 && strcmp(argv[9], "-"), hmax = argc > 10 && strcmp(argv[10], "-");
        int rc = 0;
        if (strcmp(argv[3], "/dev/null"))
            rc = ok_kite_write_table(argv[3], &t, 7, hmin, hmin ? (uint32_t)strtoul(argv[9], 0, 10) : 0,
                                     hmax, hmax ? (uint32_t)strtoul(argv[10], 0, 10) : 0, 16u << 20, 2);
        double t2 = now_s();
        fprintf(stderr, "oracle count: reads=%llu kmers=%llu unique=%llu count_s=%.3f write_s=%.3f\n",
                (unsigned long long)t.n_reads, (unsigned long long)t.n_kmers,
                (unsigned long long)t.n_unique, t1 - t0, t2 - t1);
        printf("{\"reads\": %llu, \"kmers\": %llu, \"unique\": %llu, \"count_s\": %.6f}\n",
               (unsigned long long)t.n_reads, (unsigned long long)t.n_kmers,
               (unsigned long long)t.n_unique, t1 - t0);
        ok_table_free(&t);
        if (rc) { fprintf(stderr, "error: %s\n", ok_last_error()); return 1; }
        return 0;
    }
    if (!strcmp(argv[1], "count-ref")) {
        if (argc < 9) return usage();
        int fd = open(argv[2], O_RDONLY);
        if (fd < 0) { perror("open"); return 1; }
        struct stat st;
        fstat(fd, &st);
        size_t len = (size_t)st.st_size;
        const uint8_t *img = len ? (const uint8_t *)mmap(NULL, len, PROT_READ, MAP_PRIVATE, fd, 0) : (const uint8_t *)"";
        ok_ref_stats rs;
        double t0 = now_s();
        if (ok_count_ref(img, len, atoi(argv[4]), atoi(argv[5]), (uint32_t)strtoul(argv[6], 0, 10),
                         (size_t)strtoull(argv[7], 0, 10), atoi(argv[8]), NULL, &rs)) {
            fprintf(stderr, "error: %s\n", ok_last_error());
            return 1;
        }
        double t1 = now_s();
        printf("{\"reads\": %llu, \"kmers\": %llu, \"run_rows\": %llu, \"encoded_bytes\": %llu, \"count_s\": %.6f}\n",
               (unsigned long long)rs.reads, (unsigned long long)rs.kmers, (unsigned long long)rs.run_rows,
               (unsigned long long)rs.encoded_bytes, t1 - t0);
        return 0;
    }
    if (!strcmp(argv[1], "dump")) {
        if (argc < 4) return usage();
        if (ok_kite_dump(argv[2], argv[3])) { fprintf(stderr, "error: %s\n", ok_last_error()); return 1; }
        return 0;
    }
    if (!strcmp(argv[1], "histo")) {
        if (argc < 4) return usage();
        if (ok_kite_histo((const char *const *)argv + 3, argc - 3, NULL, argv[2])) {
            fprintf(stderr, "error: %s\n", ok_last_error());
            return 1;
        }
        return 0;
    }
    if (!strcmp(argv[1], "merge")) {
        if (argc < 6) return usage();
        if (ok_kite_merge((const char *const *)argv + 4, argc - 4, argv[2], atoi(argv[3]))) {
            fprintf(stderr, "error: %s\n", ok_last_error());
            return 1;
        }
        return 0;
    }
    return usage();
}
#endif
