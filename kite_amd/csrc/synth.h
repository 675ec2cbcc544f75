/* synth.h -- seeded synthetic genomes and FASTQ reads (see synth.c). */
#ifndef KITE_SYNTH_H
#define KITE_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kite_synth_spec {
    uint64_t seed;            /* genome + read seed */
    uint64_t genome_len;      /* bases */
    uint32_t read_len;        /* bases per read */
    uint32_t both_strands;    /* 0: forward reads only, 1: strand 50/50 */
    double err_rate;          /* i.i.d. substitution rate per read base */
    double n_rate;            /* i.i.d. rate of 'N' per read base */
    uint32_t repeat_segments; /* distinct repeat segments (0: none) */
    uint32_t repeat_seg_len;  /* bases per segment */
    uint32_t repeat_copies;   /* total diverged copies placed in the genome */
    uint32_t _pad;
    double repeat_div;        /* per-base divergence of each copy */
    uint64_t variant_seed;    /* seed of the per-sample SNV layer */
    double snv_rate;          /* 0: no variant layer */
} kite_synth_spec;

/* Fill genome[0..genome_len) with ACGT.  Returns 0 on success. */
int kite_synth_genome(const kite_synth_spec *spec, uint8_t *genome);

/* Byte offset of read r's record in the full FASTQ image. */
uint64_t kite_synth_offset(const kite_synth_spec *spec, uint64_t r);

/* Write reads [r0, r1) as FASTQ into out (cap bytes).  With out == NULL
 * returns the byte size only.  Returns bytes written, or -1. */
int64_t kite_synth_reads(const kite_synth_spec *spec, const uint8_t *genome, uint64_t r0,
                         uint64_t r1, uint8_t *out, uint64_t cap, int threads);

#ifdef __cplusplus
}
#endif
#endif
