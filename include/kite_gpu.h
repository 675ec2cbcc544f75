/*
 * kite_gpu.h -- C ABI of the MI355X (gfx950) exact k-mer engine.
 *
 * The reference (jiehua1995/Kite v0.7.5) is a monolithic Rust binary with no
 * plugin or FFI interface.  These entry points sit at the seams its own code
 * has between the count driver, the per-thread workers and the .kite writer,
 * and between the .kite readers and the merge/analyze arithmetic.  Each
 * declaration names the reference function(s) it replaces.
 *
 * Conventions: every function returns 0 on success and a negative code on
 * failure; kite_gpu_last_error() then returns a message (thread-local).
 * Keys cross the boundary as little-endian u64 words, word 0 least
 * significant: 1 word for k <= 31, 2 for k <= 64, 4 for k <= 128
 * (src/count/mod.rs:232-238).  A 4-word key is (lo = w0,w1; hi = w2,w3).
 * No torch or HIP types appear in any signature.
 */
#ifndef KITE_GPU_H
#define KITE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KITE_GPU_ABI_VERSION 5

enum {
    KITE_OK = 0,
    KITE_E_INVALID = -1,  /* bad argument (k, partitions, sizes, state) */
    KITE_E_DEVICE = -2,   /* HIP runtime / device failure */
    KITE_E_NOMEM = -3,    /* device budget exceeded */
    KITE_E_CALLBACK = -4, /* the append callback returned non-zero */
    KITE_E_OVERFLOW = -5  /* a merged count exceeds u32 */
};

/* Chunk codecs (src/format/run_block.rs:1-3). */
#define KITE_CODEC_RAW 0x01u
#define KITE_CODEC_DVLE 0x02u

typedef struct kite_gpu_ctx kite_gpu_ctx;

/* Receives one encoded run chunk, borrowed for the call.  Mirrors
 * ConcurrentKiteWriter::append_chunk (src/count/writer.rs:285-348):
 * key_min / key_max are the split keys (lo128, hi128) as 4 u64 words.
 * Return non-zero to abort. */
typedef int (*kite_append_chunk_fn)(void *user, uint32_t partition, uint32_t codec, uint64_t rows,
                                    uint64_t raw_bytes, const uint64_t key_min[4],
                                    const uint64_t key_max[4], const uint8_t *data, size_t len);

typedef struct kite_count_stats {
    uint64_t input_bytes;     /* FASTQ bytes counted (after the '@' chunk check) */
    uint64_t kmers;           /* input k-mers (valid windows) */
    uint64_t unique;          /* sum over batches of unique keys per batch */
    uint64_t buckets;         /* LDS sort buckets of the last batch */
    uint32_t batches;         /* device batches run */
    uint32_t split_levels;    /* MSD split levels of the last batch */
    uint32_t skipped_chunks;  /* fed chunks dropped for not starting with '@' */
    uint32_t words;           /* key words (1, 2 or 4) */
    uint32_t graph;           /* batch: 1 when the sample's count and chunk encode ran as captured HIP graphs
                                 (one launch each) */
    uint32_t d2h_engine;      /* batch: the SDMA engine (hsa_amd_sdma_engine_id_t bit) the library chose for the
                                 sample's payload D2H; 0 when the HIP runtime chose it */
    double device_ms;         /* device time of all batches (events) */
    double extract_ms, split_ms, sort_ms, compact_ms; /* last batch */
} kite_count_stats;

/* ---- context ------------------------------------------------------------ */

/* Opens the engine on HIP device `device`.  hbm_budget_bytes bounds the
 * device memory one batch may use (0: 90% of the device's free memory).
 * Replaces the host memory planner's role for the hot path
 * (src/count/mod.rs:624-864 derive_heuristic_plan, src/count/memory.rs). */
int kite_gpu_open(int device, uint64_t hbm_budget_bytes, kite_gpu_ctx **out);
void kite_gpu_close(kite_gpu_ctx *ctx);
const char *kite_gpu_last_error(void);
int kite_gpu_abi_version(void);

/* ---- count ---------------------------------------------------------------
 * begin + feed* + finish + emit replace the count driver's worker pool:
 * run_count_typed's worker loop (src/count/mod.rs:512-540) over
 * CountWorker::process_chunk (src/count/worker.rs:196-271) and
 * CountWorker::flush_all / flush_bucket (worker.rs:273-449). */

/* Starts a count session (validation as src/count/mod.rs:208-230). */
int kite_gpu_count_begin(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions);

/* Feeds one record-aligned FASTQ chunk from host memory, as the reader
 * thread hands chunks to workers (src/count/mod.rs:424-509).  A chunk that
 * does not start with '@' yields no k-mers, as in the reference
 * (src/count/parser.rs:24-27 with the error ignored at worker.rs:217). */
int kite_gpu_count_feed(kite_gpu_ctx *ctx, const uint8_t *fastq, size_t len);

/* Feeds FASTQ bytes already resident in device memory (one chunk).  The
 * buffer must stay valid until kite_gpu_count_finish returns. */
int kite_gpu_count_feed_device(kite_gpu_ctx *ctx, const void *d_fastq, size_t len);

/* Feeds BGZF-compressed FASTQ: whole BGZF members from data[0, len), the
 * members scanned as the reference's BGZF reader does (is_bgzf_header /
 * bgzf_block_total_size, src/count/bgzf.rs:37-101; non-BGZF data is an
 * error).  A truncated last member is not consumed: *consumed receives the
 * bytes of the whole members, and the caller passes the rest again with
 * more data.  Members are inflated on the device, one per thread, straight
 * into the FASTQ staging buffer (the thread pool of decompress_bgzf_parallel,
 * src/count/bgzf.rs:152-264).  Batches are cut after the last whole record
 * (find_fastq_split_4lines, :109-121); a batch that does not start with '@'
 * yields no k-mers, as a reader chunk.  One session takes either plain
 * chunks or BGZF input. */
int kite_gpu_count_feed_bgzf(kite_gpu_ctx *ctx, const uint8_t *data, size_t len, size_t *consumed);

/* Inflates the whole BGZF members of data[0, len) on the device into
 * out[0, cap) (host memory).  *out_len receives the output size, also when
 * cap is too small (then KITE_E_INVALID); *device_ms (may be NULL) the
 * inflate kernel time. */
int kite_gpu_inflate_bgzf(kite_gpu_ctx *ctx, const uint8_t *data, size_t len, uint8_t *out, size_t cap,
                          size_t *out_len, double *device_ms);

/* Runs the device pipeline on everything fed; results stay on the device. */
int kite_gpu_count_finish(kite_gpu_ctx *ctx, kite_count_stats *stats);

/* Encodes the results as DVLE (k <= 64) or RAW (k >= 65) run chunks, one
 * per non-empty partition per batch, in partition order, and passes each to
 * `fn` -- the job flush_bucket does before append_chunk
 * (src/count/worker.rs:296-449, src/format/run_block.rs:70-95). */
int kite_gpu_count_emit(kite_gpu_ctx *ctx, kite_append_chunk_fn fn, void *user);

/* Size of the reduced result of the last batch. */
int kite_gpu_count_result_size(kite_gpu_ctx *ctx, uint64_t *n_unique, uint32_t *words,
                               uint32_t *partitions);

/* Copies the last batch's result to host memory: keys (n_unique * words
 * u64), counts (n_unique u32, saturated at u32::MAX) and part_off
 * (partitions + 1 offsets).  Any pointer may be NULL. */
int kite_gpu_count_download(kite_gpu_ctx *ctx, uint64_t *keys, uint32_t *counts,
                            uint64_t *part_off);

/* Device pointers of the last batch's result (valid until the next
 * finish/close): keys, counts and the host copy of part_off. */
int kite_gpu_count_device_result(kite_gpu_ctx *ctx, const uint64_t **d_keys,
                                 const uint32_t **d_counts, const uint64_t **part_off);

/* ---- batched count of many samples ---------------------------------------
 * Counts n_samples independent FASTQ buffers (host memory; pinned memory
 * from kite_gpu_host_alloc lets the copies overlap the kernels).  Each
 * sample is one record-aligned chunk and is reduced completely, as one
 * `kite count` run per sample would (run_count_typed, src/count/mod.rs:
 * 512-540, once per input file).  The reference has no multi-sample entry
 * point; this replaces a loop of count runs over the inputs of a cohort.
 *
 * mode KITE_BATCH_PLAIN: one sample at a time -- fresh device allocations,
 * H2D copy, kernels, D2H copy, synchronise, free.
 * mode KITE_BATCH_SERIAL: one sample at a time like PLAIN, but in
 * persistent device buffers (no per-sample allocation).
 * mode KITE_BATCH_PIPELINED: persistent device buffers in a ring of three
 * (KITE_BATCH_RING: 2..4); the H2D copies of the next samples and the D2H
 * copies of earlier ones run on their own streams while sample i is counted,
 * and results are handed to `fn` on a host thread of the engine's own while
 * later samples run.
 *
 * `fn` receives sample i's result, borrowed for the call: keys
 * (stats->unique * words u64), counts (u32, saturated at u32::MAX) and
 * part_off (partitions + 1).  Calls come in sample order.  Return non-zero
 * to abort the batch (KITE_E_CALLBACK). */
#define KITE_BATCH_PLAIN 0
#define KITE_BATCH_PIPELINED 1
#define KITE_BATCH_SERIAL 2

typedef struct kite_batch_sample {
    const uint8_t *fastq;
    size_t len;
} kite_batch_sample;

typedef int (*kite_batch_result_fn)(void *user, uint32_t sample, const kite_count_stats *stats,
                                    const uint64_t *keys, const uint32_t *counts,
                                    const uint64_t *part_off);

int kite_gpu_count_batch(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions,
                         const kite_batch_sample *samples, uint32_t n_samples, int mode,
                         kite_batch_result_fn fn, void *user);

/* The same batch, each sample delivered as .kite chunk payloads encoded on
 * the device (DVLE for k <= 64, RAW above; only these bytes cross PCIe):
 * partition p's payload is data[byte_off[p], byte_off[p + 1]) holding
 * rows_off[p + 1] - rows_off[p] rows, with first and last keys
 * minmax[8p .. 8p+3] and minmax[8p+4 .. 8p+7] (4 words each, unused words 0).
 * A host writes each sample's .kite from them as the count workers hand
 * chunks to ConcurrentKiteWriter::append_chunk (src/count/writer.rs:285-348). */
typedef int (*kite_batch_chunks_fn)(void *user, uint32_t sample, const kite_count_stats *stats, uint32_t codec,
                                    const uint8_t *data, const uint64_t *byte_off, const uint64_t *rows_off,
                                    const uint64_t *minmax);

int kite_gpu_count_batch_chunks(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions,
                                const kite_batch_sample *samples, uint32_t n_samples, int mode,
                                kite_batch_chunks_fn fn, void *user);

/* Autotuner of the count's own settings: times the count of the sample
 * fastq[0, len) (host memory, whole records) under candidate settings --
 * the MSD split target (1, 2, 4, 8: children split until they hold at most
 * the largest LDS sort bucket / target), the maximum radix digit width of
 * a split level (8, 10, 11, 12 bits), how many sort size classes use the
 * hash-then-bin kernel (2, 3, 4), the extract tile (1, 2, 4, 8 sub-tiles
 * per workgroup) and the size above which a node is split again (the
 * largest bucket capacity / 1 or / 2) -- one setting at a time with the others held, the best of
 * three device timings each, a change kept when it wins by 1%.  Every
 * setting must give the same k-mer and unique counts.  The winner is used
 * by every later count on this context with the same key width and
 * partition count and an input of the same size class (below 64 MiB,
 * 512 MiB, 4 GiB, above).  report (may be NULL) receives a JSON summary:
 * the chosen settings, default and tuned device ms, and every trial.
 * Counterpart of the reference's run-time planner and k-aware sort choice
 * (derive_heuristic_plan, src/count/mod.rs:624-864; radix_sort_u64_kmer,
 * src/count/sorter.rs:67-93).  A count session's device result is first
 * kept as encoded runs. */
int kite_gpu_autotune(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions, const uint8_t *fastq,
                      size_t len, char *report, size_t cap);

/* Sets the count settings of one (key width of k, partition count, input
 * size class 0-3) without timing anything, as if kite_gpu_autotune had
 * chosen them: a caller re-applies choices kept from an earlier run.  The
 * values are the autotuner's candidates (split target 1, 2, 4, 8; digit
 * bits 8-12; hash classes 2-4; extract sub-tiles 1, 2, 4, 8; split-above
 * divisor 1 or 2); others are rejected.  Counterpart of passing the
 * reference planner's derived settings back in by hand (its --partitions,
 * -t and --memory flags, src/count/mod.rs:923-1016). */
int kite_gpu_set_tuned(kite_gpu_ctx *ctx, uint32_t k, uint32_t partitions, int size_class, int split_target_div,
                       int max_digit_bits, int hash_classes, int extract_subs, int split_above_div);

/* Host-link copy rates of this device: page-locked buffers of `bytes`,
 * H2D alone, D2H alone, and both at once on two streams (GB/s; one untimed
 * copy each way, then the best of 5).  The bound of the batch above, whose
 * H2D and D2H run at once.  No counterpart in the reference (it has no
 * device link). */
int kite_gpu_copy_peak(kite_gpu_ctx *ctx, uint64_t bytes, double *h2d_gbps, double *d2h_gbps, double *both_gbps);

/* Page-locked host memory for FASTQ buffers and results (hipHostMalloc). */
int kite_gpu_host_alloc(uint64_t bytes, void **out);
void kite_gpu_host_free(void *p);

/* ---- merge ---------------------------------------------------------------
 * Sorted union with count sums of n_inputs tables that share one partition
 * layout: merge_vectors_total (src/kiteops/mod.rs:587-724) as used per
 * partition by build_merge_partition_chunk (src/merge/mod.rs:198-214).
 * Table i: keys[i] (rows_i * words u64, ascending and unique within each
 * partition), counts[i] (rows_i u64), part_off[i] (partitions + 1).
 * out_keys / out_counts must hold sum(rows_i) rows; out_part_off receives
 * partitions + 1 offsets.  Equal keys sum with u64 saturation. */
int kite_gpu_merge_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t partitions, uint32_t n_inputs,
                          const uint64_t *const *keys, const uint64_t *const *counts,
                          const uint64_t *const *part_off, uint64_t *out_keys, uint64_t *out_counts,
                          uint64_t *out_part_off);

/* Union rows of n_inputs tables with one count column per table, zero
 * filled: for_each_union_row (src/kiteops/mod.rs:625-802).  out_keys holds
 * sum(rows_i) rows; out_matrix holds sum(rows_i) * n_inputs u32 (row major,
 * saturated at u32::MAX); out_part_off receives partitions + 1 offsets. */
int kite_gpu_union_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t partitions, uint32_t n_inputs,
                          const uint64_t *const *keys, const uint64_t *const *counts,
                          const uint64_t *const *part_off, uint64_t *out_keys, uint32_t *out_matrix,
                          uint64_t *out_part_off);

/* ---- analyze ---------------------------------------------------------------
 * Correlation of the columns of a union count matrix (rows = union k-mers,
 * columns = samples): Pearson on ln(1 + count) (PearsonAccumulator,
 * src/analyze/mod.rs:194-270, values :780-784) and Spearman as Pearson of
 * 1-based average-tie ranks over all rows (average_ranks + pearson_slice,
 * :406-438, :559-578).  The Gram runs on bf16 MFMA with a hi/lo split (FP32
 * within a K-tile, FP64 across tiles); column means and variances, the
 * reduction and the normalisation are FP64.  Any output pointer may be NULL.
 * Every count is read through its rank among its column's distinct counts
 * (1-, 2- or 4-byte codes, by the most distinct counts in a column), so any
 * count value works; 1..4096 samples (more than 32 as pairs of 16-column
 * blocks through the same 32 x 32 Gram). */
#define KITE_CORR_PEARSON 0x1u
#define KITE_CORR_SPEARMAN 0x2u
#define KITE_CORR_FP64_BASELINE 0x1u /* flags: also run the FP64 MFMA Gram */
#define KITE_CORR_LOLO 0x2u          /* flags: add the lo*lo product */

typedef struct kite_corr_stats {
    double gram_ms_pearson, gram_ms_spearman; /* split-bf16 Gram + FP64 reduction (share of the pass) */
    double fp64_ms_pearson, fp64_ms_spearman; /* FP64 MFMA Gram + reduction (share of the pass) */
    uint64_t rows;
    uint32_t cols, code_bytes;                /* bytes per matrix cell read by the Gram: 1 = exact
                                                 8-bit rank of the count within its column */
    double gram_ms_total, fp64_ms_total;      /* all requested methods: the split-bf16 Gram computes
                                                 them in one pass over the matrix, the FP64 Gram in
                                                 one pass per method (its faster form) */
    uint32_t passes, fp64_passes;             /* passes over the matrix of each Gram */
} kite_corr_stats;

/* counts: N x n row-major u32 (host).  pearson / spearman / cov: n x n
 * (cov of ln(1+count), divisor N-1: covariance_matrix :238-252). */
int kite_gpu_corr_counts(kite_gpu_ctx *ctx, const uint32_t *counts, uint64_t N, uint32_t n, uint32_t methods,
                         uint32_t flags, double *pearson, double *spearman, double *cov, double *pearson_fp64,
                         double *spearman_fp64, kite_corr_stats *stats);

/* The same on the union of n_inputs tables (see kite_gpu_union_tables),
 * built and consumed on the device; out_keys / out_matrix / out_part_off
 * optionally return the union (for exact_count_matrix.tsv; out_matrix cells
 * saturate at u32::MAX).  When an input count exceeds u32::MAX the union
 * the correlation reads holds exact u64 counts. */
int kite_gpu_analyze_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t partitions, uint32_t n_inputs,
                            const uint64_t *const *keys, const uint64_t *const *counts,
                            const uint64_t *const *part_off, uint32_t methods, uint32_t flags,
                            uint64_t *out_keys, uint32_t *out_matrix, uint64_t *out_part_off, double *pearson,
                            double *spearman, double *cov, double *pearson_fp64, double *spearman_fp64,
                            kite_corr_stats *stats);

/* PCA of an n x n covariance (pca_scores, src/analyze/mod.rs:1605-1625):
 * FP64 cyclic Jacobi on the device (one wave for n <= 64, a round-robin
 * parallel Jacobi on one workgroup above), eigenvalues descending; scores
 * (n x dims) = eigenvector * sqrt(max(eigenvalue, 0)).  n <= 1024. */
int kite_gpu_pca(kite_gpu_ctx *ctx, const double *cov, uint32_t n, uint32_t dims, double *scores,
                 double *eigenvalues);

/* kite_gpu_pca without any allocation: no counterpart in the reference.  It
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * Jacobi works in a fixed per-GPU device array of 2 * 64 * 64 + 64 FP64
 * (66,048 bytes, held until the process ends), so n is limited to 1..64
 * (larger n: KITE_E_INVALID).  Same arguments and results as kite_gpu_pca.
 * Calls on one GPU share that array: within a process they take turns (a
 * per-GPU lock), so overlapping calls wait rather than race. */
int kite_gpu_pca_noalloc(kite_gpu_ctx *ctx, const double *cov, uint32_t n, uint32_t dims, double *scores,
                         double *eigenvalues);

/* Turns profiling on or off for later counts: KITE_PROFILE_KERNELS records
 * HIP events around every launch on the engine's stream, plus each kernel's
 * algorithmic HBM bytes; KITE_PROFILE_TIMELINE records per-sample stage
 * events on the three streams of a pipelined batch (kite_gpu_batch_timeline).
 * Counterpart of the reference's --test-mode stage timers
 * (src/stats.rs:8-273), at kernel granularity. */
#define KITE_PROFILE_KERNELS 0x1
#define KITE_PROFILE_TIMELINE 0x2
int kite_gpu_count_profile(kite_gpu_ctx *ctx, int enable);

/* Writes the last batch's kernel profile as JSON into buf (cap bytes):
 * {"<kernel>": {"ms": total, "launches": n, "bytes": algorithmic}, ...}.
 * Returns the JSON length, or a negative code. */
int kite_gpu_count_kernel_stats(kite_gpu_ctx *ctx, char *buf, size_t cap);

/* Stage times of the last pipelined batch run with profiling on: per sample
 * six values in ms since the batch started -- H2D start and end (copy
 * stream), count start and count (+ chunk encode) end (engine stream), D2H
 * start and end (second copy stream).  out receives up to cap_samples * 6
 * doubles; returns the number of samples timed.  After a batch, the kernel
 * profile of kite_gpu_count_kernel_stats is summed over its samples.
 * Counterpart of the reference's per-stage --test-mode timers
 * (src/stats.rs:8-273) for the multi-sample pipeline. */
int kite_gpu_batch_timeline(kite_gpu_ctx *ctx, double *out, uint32_t cap_samples);

/* ---- dump -------------------------------------------------------------------
 * Replaces the partition workers of `kite dump` (stream_partition_slabs,
 * src/dump/merge.rs:282-428, driven by src/dump/mod.rs:293-348): the stored
 * runs of every partition are decoded (DVLE / RAW, src/dump/merge.rs:13-240),
 * merged in key order with u32 saturating sums (src/dump/loser_tree.rs:
 * 141-175; 256-bit keys: sort + RLE, src/dump/merge.rs:311-371, which equals
 * the merge for the sorted unique runs writers produce; other 256-bit runs
 * are refused), filtered by min <= count <= max (src/kiteops/mod.rs:44-56)
 * and formatted as "KMER\tCOUNT\n" (src/dump/decoder.rs:27-117), partitions
 * in order 0..P-1.  A DVLE varint without its end byte within 10 (64-bit) or
 * 19 (128-bit) bytes, or a payload shorter than its rows, is an error. */
typedef struct kite_dump_run {
    uint32_t partition;   /* non-decreasing over the array */
    uint32_t codec;       /* KITE_CODEC_DVLE or KITE_CODEC_RAW; zstd removed by the caller */
    uint64_t rows;
    const uint8_t *data;  /* payload, borrowed for the call */
    uint64_t bytes;
} kite_dump_run;

typedef struct kite_dump_stats {
    uint64_t runs, rows_in, rows, lines, bytes, groups;
    double decode_ms, merge_ms, format_ms;  /* device time per stage */
    double copy_ms;  /* text phase, wall time: D2H pieces overlapped with the writer callback */
} kite_dump_stats;

/* Receives the text in order, in pieces; non-zero aborts the dump. */
typedef int (*kite_dump_write_fn)(void *user, const char *text, uint64_t len);

int kite_gpu_dump(kite_gpu_ctx *ctx, uint32_t k, uint32_t key_width, uint32_t partitions, const kite_dump_run *runs,
                  uint64_t n_runs, int has_min, uint32_t min_count, int has_max, uint32_t max_count,
                  kite_dump_write_fn fn, void *user, kite_dump_stats *stats);

/* ---- histo ------------------------------------------------------------------
 * Replaces the partition workers of `kite histo` for one input
 * (src/histo/mod.rs:76-111 over load_partition_counts, src/kiteops/mod.rs:
 * 433-473): the stored runs (as for kite_gpu_dump) are decoded with u64
 * counts, the runs of each partition summed per key (u64 saturating), the
 * header filter applied to each total, and the frequency of every total
 * counted.  `fn` receives the bins once, values ascending: n_bins values and
 * their frequencies.  The caller forms the union of bins over its inputs
 * (merge_histograms, src/histo/mod.rs:171-176).  stats->lines is the number
 * of keys that passed the filter. */
typedef int (*kite_histo_fn)(void *user, const uint64_t *values, const uint64_t *freqs, uint64_t n_bins);

int kite_gpu_histo(kite_gpu_ctx *ctx, uint32_t key_width, uint32_t partitions, const kite_dump_run *runs,
                   uint64_t n_runs, int has_min, uint32_t min_count, int has_max, uint32_t max_count,
                   kite_histo_fn fn, void *user, kite_dump_stats *stats);

/* ---- merge of stored runs ------------------------------------------------------
 * `kite merge` on the device, partition by partition as
 * build_merge_partition_chunk does (src/merge/mod.rs:198-214): each input's
 * stored runs of the partition (as for kite_gpu_dump) are decoded and summed
 * per key with u64 saturation (load_partition_counts, src/kiteops/mod.rs:
 * 433-473), that input's header filter is applied to each total
 * (apply_reader_filter_to_rows, :103-108), the inputs are merged with u64
 * saturating sums (merge_vectors_total, :587-724), and the partition becomes
 * one chunk (serialize_partition_chunk, :804-859): DVLE for 64- and 128-bit
 * keys, RAW for 256-bit keys.  `fn` receives the chunks of the non-empty
 * partitions in partition order, as the merge hands them to append_chunk
 * (src/merge/mod.rs:134-151).  A total above u32::MAX fails with
 * KITE_E_OVERFLOW ("merged count exceeds u32 for key <key>"), before any
 * chunk of the partition group that holds it is passed on.  Only payload
 * bytes cross the host link: the stored payloads in, the chunks out. */
typedef struct kite_merge_input {
    const kite_dump_run *runs;  /* the input's stored runs, partitions non-decreasing */
    uint64_t n_runs;
    int has_min;                /* the input header's count filter */
    uint32_t min_count;
    int has_max;
    uint32_t max_count;
} kite_merge_input;

typedef struct kite_merge_stats {
    uint64_t runs, rows_in;   /* stored runs and rows read */
    uint64_t rows_kept;       /* rows of all inputs after the header filters */
    uint64_t rows, bytes;     /* merged rows, chunk payload bytes */
    uint64_t groups;          /* partition groups through the device */
    double decode_ms, merge_ms, encode_ms; /* device time per stage */
    double in_ms, out_ms;     /* host wall time: payloads staged and copied in; chunks copied out and handed to fn */
} kite_merge_stats;

int kite_gpu_merge_runs(kite_gpu_ctx *ctx, uint32_t key_width, uint32_t partitions, const kite_merge_input *inputs,
                        uint32_t n_inputs, kite_append_chunk_fn fn, void *user, kite_merge_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
