/*
 * kite_oracle.h -- CPU restatement of the exact k-mer counting path of
 * jiehua1995/Kite v0.7.5 (count -> .kite runs -> dump / histo / merge / union).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the parity checker and the timed
 * CPU baseline ("port").  Nothing in the shipped engine (kite_amd/) links,
 * loads or calls it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do.
 *
 * Keys are held as up to four little-endian 64-bit words (w[0] = least
 * significant).  Word count by k follows src/count/mod.rs:232-238:
 *   k <= 31 -> 1 word (u64), k <= 64 -> 2 words (u128), else 4 (KmerU256,
 *   lo = w[0],w[1], hi = w[2],w[3]).
 */
#ifndef KITE_ORACLE_H
#define KITE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of u64 words the reference uses for k (count/mod.rs:232-238). */
int ok_key_words(int k);

/* Rolling k-mer extraction over one sequence line (count/kmer.rs:203-432).
 * `words` selects the iterator (1: KmerIter, 2: KmerIter128, 4: KmerIterU256).
 * Writes up to `cap` keys of `words` u64 each into `out`; returns the number
 * of k-mers the sequence yields (may exceed cap: then only cap are written). */
size_t ok_extract(const uint8_t *seq, size_t len, int k, int canonical, int words,
                  uint64_t *out, size_t cap);

/* Fibonacci partition of a key (count/kmer.rs:67-71, 111-116, 156-161). */
uint32_t ok_bucket(const uint64_t *key, int words, int bucket_bits);

/* FASTQ sequence-line spans of one chunk (count/parser.rs:19-65).
 * Returns the number of sequences, or -1 when the chunk does not start with
 * '@'.  Up to `cap` (start, len) pairs are written to spans. */
long ok_parse_sequences(const uint8_t *data, size_t len, uint64_t *spans, size_t cap);

/* find_fastq_split_4lines (count/bgzf.rs:109-121). */
size_t ok_fastq_split_4lines(const uint8_t *data, size_t len);

/* Chunk boundaries the reference's PlainChunkedReader produces for a file
 * image of `len` bytes read in `chunk_size` pieces (count/decompress.rs:85-138).
 * Writes up to cap (offset, length) pairs; returns the number of chunks. */
size_t ok_plain_chunks(const uint8_t *data, size_t len, size_t chunk_size,
                       uint64_t *chunks, size_t cap);

/* Per-partition exact count table. */
typedef struct ok_table {
    int k, canonical, words;
    uint32_t partitions;
    uint64_t *part_off;   /* partitions + 1 offsets into keys/counts */
    uint64_t *keys;       /* n_unique * words, ascending within a partition */
    uint64_t *counts;     /* u64 totals (never saturated here) */
    uint64_t n_unique;
    uint64_t n_kmers;     /* input k-mers (windows) */
    uint64_t n_reads;
} ok_table;

/* Count a FASTQ image the way `kite count` does: plain chunker with
 * chunk_size, per-chunk '@' check, parse, extract, partition, sort, RLE.
 * threads > 1 spreads chunks over worker threads.  Returns 0 on success. */
int ok_count_fastq(const uint8_t *data, size_t len, int k, int canonical,
                   uint32_t partitions, size_t chunk_size, int threads, ok_table *out);
void ok_table_free(ok_table *t);

/* The reference's count loop for one-word keys (k <= 31), for the CPU
 * baseline: chunks taken in turn by `threads` workers, inline rolling
 * k-mers into 64-entry write-combining partition slots, then per worker and
 * partition one run: MSD/LSD byte radix sort, saturating RLE, DVLE encode
 * (count/mod.rs:512-540, worker.rs:56-449, sorter.rs:67-272, 514-530).
 * `out` (may be NULL) receives the runs merged into totals (tests). */
typedef struct ok_ref_stats {
    uint64_t reads, kmers;
    uint64_t run_rows;       /* rows over every worker's runs */
    uint64_t encoded_bytes;  /* DVLE bytes over every run */
} ok_ref_stats;
int ok_count_ref(const uint8_t *data, size_t len, int k, int canonical, uint32_t partitions,
                 size_t chunk_size, int threads, ok_table *out, ok_ref_stats *st);

/* Radix sort + saturating RLE of raw keys (count/sorter.rs:514-530):
 * sorts `n` keys of `words` words in place, writes unique keys (in place,
 * compacted) and u32 saturating run lengths; returns the unique count. */
size_t ok_sort_rle(uint64_t *keys, size_t n, int words, int k, uint32_t *counts);

/* LEB128 and DVLE (format/run_block.rs:9-95).  Returns bytes written. */
size_t ok_dvle_encode(const uint64_t *keys, const uint32_t *counts, size_t n,
                      int words, uint8_t *out);

/* ---- .kite files ---------------------------------------------------- */

/* Write a count table as a .kite file: one chunk per non-empty partition,
 * appended in partition order through the reference's staging writer
 * (count/writer.rs:76-495) with `staging_bytes` and `write_lanes`.
 * Counts above u32::MAX are an error (kiteops/mod.rs:804-859).
 * codec_mask: 7 for count, 3 for merge.  has_min/has_max select markers. */
int ok_kite_write_table(const char *path, const ok_table *t, uint16_t codec_mask,
                        int has_min, uint32_t min, int has_max, uint32_t max,
                        size_t staging_bytes, int write_lanes);

/* Load one .kite file, totals summed per key (kiteops/mod.rs:433-473),
 * before any header filter.  Also returns the header markers. */
int ok_kite_load(const char *path, ok_table *out, int *has_min, uint32_t *min,
                 int *has_max, uint32_t *max);

/* `kite dump -o <out>` in the deterministic partition order
 * (dump/mod.rs:293-348, dump/merge.rs:282-428). */
int ok_kite_dump(const char *path, const char *out_path);

/* `kite histo` (histo/mod.rs:16-169).  labels may be NULL (inferred). */
int ok_kite_histo(const char *const *paths, int n, const char *const *labels,
                  const char *out_path);

/* `kite merge` (merge/mod.rs:22-214) with the reference writer layout. */
int ok_kite_merge(const char *const *paths, int n, const char *out_path, int write_lanes);

/* Union count matrix of n .kite files after each file's header filter
 * (kiteops/mod.rs:625-802).  *rows_out receives N; counts is N*n u64,
 * row-major, allocated with malloc.  keys (N*words) likewise. */
int ok_kite_union(const char *const *paths, int n, uint64_t *rows_out,
                  uint64_t **keys_out, uint64_t **counts_out, int *words_out);

/* Decode a key to ACGT exactly as the reference's dump does
 * (dump/decoder.rs:27-107, including its k>=65 layout). */
void ok_decode_key(const uint64_t *key, int words, int k, char *buf);

const char *ok_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
