/*
 * synth.c -- seeded synthetic genomes and FASTQ reads for the benchmark
 * workloads (uniform random genome, optional diverged repeat copies,
 * optional per-sample SNV variant genome, uniform read starts, i.i.d.
 * substitution errors and Ns, constant 'I' qualities, "@r<index>" headers).
 *
 * Every base is a pure function of (seed, position), so any read range can
 * be produced independently and in parallel: a bounded prefix of a workload
 * is byte-identical to the start of the full file.
 */
#include "synth.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline uint64_t h3(uint64_t seed, uint64_t stream, uint64_t i) {
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ i);
}

/* uniform double in [0,1) from 53 bits */
static inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

static const char ACGT[4] = {'A', 'C', 'G', 'T'};

static inline int code_of(uint8_t b) { return b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : 3; }

enum { S_BASE = 1, S_REP_POS, S_REP_SEG, S_REP_DIV, S_REP_SUB, S_SNV, S_SNV_SUB, S_READ, S_ERR,
       S_ERR_SUB, S_NPOS };

int kite_synth_genome(const kite_synth_spec *s, uint8_t *g) {
    if (!s || !g || s->genome_len == 0) return -1;
    for (uint64_t i = 0; i < s->genome_len; i++) g[i] = (uint8_t)ACGT[h3(s->seed, S_BASE, i) & 3];
    /* repeat copies: copy c of segment (c % segments) overwrites a
     * seed-chosen window, each base diverged with probability repeat_div */
    if (s->repeat_segments && s->repeat_seg_len && s->repeat_copies &&
        s->repeat_seg_len < s->genome_len) {
        uint32_t L = s->repeat_seg_len;
        uint8_t *seg = (uint8_t *)malloc(L);
        for (uint32_t c = 0; c < s->repeat_copies; c++) {
            uint32_t sidx = c % s->repeat_segments;
            for (uint32_t j = 0; j < L; j++)
                seg[j] = (uint8_t)ACGT[h3(s->seed, S_REP_SEG, (uint64_t)sidx * L + j) & 3];
            uint64_t pos = h3(s->seed, S_REP_POS, c) % (s->genome_len - L + 1);
            for (uint32_t j = 0; j < L; j++) {
                uint8_t b = seg[j];
                uint64_t key = (uint64_t)c * L + j;
                if (unit(h3(s->seed, S_REP_DIV, key)) < s->repeat_div) {
                    int code = code_of(b);
                    code = (code + 1 + (int)(h3(s->seed, S_REP_SUB, key) % 3)) & 3;
                    b = (uint8_t)ACGT[code];
                }
                g[pos + j] = b;
            }
        }
        free(seg);
    }
    /* per-sample SNV variant genome */
    if (s->snv_rate > 0.0) {
        for (uint64_t i = 0; i < s->genome_len; i++) {
            if (unit(h3(s->variant_seed, S_SNV, i)) < s->snv_rate) {
                int code = code_of(g[i]);
                code = (code + 1 + (int)(h3(s->variant_seed, S_SNV_SUB, i) % 3)) & 3;
                g[i] = (uint8_t)ACGT[code];
            }
        }
    }
    return 0;
}

static inline uint64_t digits10(uint64_t v) {
    uint64_t d = 1;
    while (v >= 10) { v /= 10; d++; }
    return d;
}

/* bytes of all records with index < r: "@r" + digits + "\n" + L + "\n+\n" + L + "\n" */
uint64_t kite_synth_offset(const kite_synth_spec *s, uint64_t r) {
    uint64_t fixed = 2 + 1 + (uint64_t)s->read_len + 3 + (uint64_t)s->read_len + 1;
    uint64_t total = r * fixed;
    /* indexes in [0,10) have 1 digit, [10,100) have 2, ... */
    uint64_t lo = 0, hi = 10, d = 1;
    while (lo < r) {
        uint64_t end = hi < r ? hi : r;
        total += (end - lo) * d;
        lo = end;
        d++;
        hi *= 10;
    }
    return total;
}

static void write_read(const kite_synth_spec *s, const uint8_t *g, uint64_t r, uint8_t *o) {
    uint32_t L = s->read_len;
    size_t p = 0;
    o[p++] = '@';
    o[p++] = 'r';
    uint64_t nd = digits10(r), v = r;
    for (uint64_t i = 0; i < nd; i++) { o[p + nd - 1 - i] = (uint8_t)('0' + v % 10); v /= 10; }
    p += nd;
    o[p++] = '\n';
    uint64_t hr = h3(s->seed, S_READ, r);
    uint64_t span = s->genome_len >= L ? s->genome_len - L + 1 : 1;
    uint64_t start = hr % span;
    int rc = s->both_strands ? (int)((hr >> 63) & 1) : 0;
    for (uint32_t j = 0; j < L; j++) {
        uint64_t gp = rc ? start + (L - 1 - j) : start + j;
        uint8_t b = gp < s->genome_len ? g[gp] : 'A';
        if (rc) b = (uint8_t)(b == 'A' ? 'T' : b == 'C' ? 'G' : b == 'G' ? 'C' : 'A');
        uint64_t key = r * (uint64_t)L + j;
        if (s->err_rate > 0 && unit(h3(s->seed, S_ERR, key)) < s->err_rate) {
            int code = code_of(b);
            code = (code + 1 + (int)(h3(s->seed, S_ERR_SUB, key) % 3)) & 3;
            b = (uint8_t)ACGT[code];
        }
        if (s->n_rate > 0 && unit(h3(s->seed, S_NPOS, key)) < s->n_rate) b = 'N';
        o[p + j] = b;
    }
    p += L;
    o[p++] = '\n';
    o[p++] = '+';
    o[p++] = '\n';
    memset(o + p, 'I', L);
    p += L;
    o[p] = '\n';
}

typedef struct {
    const kite_synth_spec *s;
    const uint8_t *g;
    uint64_t r0, r1, base;
    uint8_t *out;
} synth_job;

static void *synth_worker(void *arg) {
    synth_job *j = (synth_job *)arg;
    for (uint64_t r = j->r0; r < j->r1; r++)
        write_read(j->s, j->g, r, j->out + (kite_synth_offset(j->s, r) - j->base));
    return NULL;
}

int64_t kite_synth_reads(const kite_synth_spec *s, const uint8_t *genome, uint64_t r0, uint64_t r1,
                         uint8_t *out, uint64_t cap, int threads) {
    if (!s || !genome || r1 < r0) return -1;
    uint64_t base = kite_synth_offset(s, r0);
    uint64_t need = kite_synth_offset(s, r1) - base;
    if (!out) return (int64_t)need;
    if (need > cap) return -1;
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    uint64_t n = r1 - r0;
    if (n < (uint64_t)threads * 1024) threads = 1;
    synth_job jobs[64];
    pthread_t th[64];
    for (int t = 0; t < threads; t++) {
        jobs[t].s = s;
        jobs[t].g = genome;
        jobs[t].r0 = r0 + n * (uint64_t)t / (uint64_t)threads;
        jobs[t].r1 = r0 + n * (uint64_t)(t + 1) / (uint64_t)threads;
        jobs[t].base = base;
        jobs[t].out = out;
        if (threads == 1) synth_worker(&jobs[t]);
        else pthread_create(&th[t], NULL, synth_worker, &jobs[t]);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    return (int64_t)need;
}
