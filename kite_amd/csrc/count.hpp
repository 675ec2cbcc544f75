// count.hpp -- the gfx950 exact k-mer count pipeline (host-side interface).
//
// One batch of FASTQ bytes resident in HBM becomes per-partition sorted,
// fully reduced (key, count) runs, also resident in HBM:
//
//   k_nl_count      newline count per sub-tile (8064 bytes + 128 halo)
//   scan            line index at each sub-tile start
//   k_extract<HIST> 2-bit pack, rolling fwd/rc k-mers, Fibonacci partition
//                   histogram per (partition, workgroup)
//   scan            partition-major destination of every workgroup's keys
//   k_extract<SCAT> same k-mers, scattered into partition order
//   k_split<HIST/SCAT> MSD split of large partitions on the top key bits
//                   (the bucket list itself is built on the device)
//   k_bucket_hash_sort / k_bucket_sort
//                   per-bucket sort in LDS + run-length counts
//   scan + k_compact  contiguous per-partition (key, count) output
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace kite {

struct EncodeBufs;

struct Bucket {
    uint64_t start;   // first key (index into the key array)
    uint64_t size;    // keys in the bucket
    uint32_t prefix;  // key bits (from the top of the 2k-bit key) shared by all keys
    uint32_t array;   // 0: keys live in A, 1: in B
};

// ---- device-planned count (a launch sequence fixed before the data) -----------
// Every size the classic host orchestration reads back is kept here, in
// device memory, and the kernels read it there; buffers are sized for a
// capacity.  A count whose data would exceed a capacity sets `overflow` and
// writes nothing further; its caller runs the classic count instead.
constexpr int DEV_LEVELS = 3;  // MSD split levels within one planned count

struct DevLevel {
    uint64_t nn;                     // nodes in the list
    uint64_t nact;                   // nodes split at this level
    uint64_t tiles;                  // split workgroup tiles
    unsigned long long maxsz, keys;  // largest split node, keys in split nodes
    uint64_t nchild;                 // nact << d
    int32_t d, shift, prefix, staged;  // digit bits and shift; key bits above; staged scatter
};

struct DevPlan {
    uint64_t len, n_sub, n;       // input bytes, sub-tiles, k-mers
    uint64_t nwg;                 // extract workgroups with input
    DevLevel lv[DEV_LEVELS + 1];  // lv[DEV_LEVELS].nn: the buckets
    uint32_t cls_n[4], cls_cur[4], cls_off[5], fb_n[4];
    uint32_t sort_cur[12];        // bucket cursors of the sort launches (hash, LSD, LSD fallback per class)
    uint64_t n_unique, bytes;     // unique keys, encoded chunk bytes
    uint32_t overflow;            // bit 0 keys, 1 split levels, 2 nodes, 3 children, 4 chunk bytes
};

// Host view of a planned count's result, written to page-locked memory by
// its last kernel: the summary, then part_off (P + 1 row offsets), byte_off
// (P + 1 chunk byte offsets) and minmax (8 P words: first and last key of
// every partition, 4 words each).
struct DevSummary {
    uint64_t n_kmers, n_unique, bytes, buckets, fallback, len;
    uint64_t nwg, split_keys;  // extract workgroups with input; keys through the split passes
    uint32_t overflow, levels;
};

// Capacities of a planned count (all from the input capacity in bytes).
struct DevCaps {
    uint64_t len = 0, n_sub = 0, nwg = 0, n = 0, nodes = 0, children = 0;
    uint32_t subs = 0;
};

struct StageTimes {
    float extract_ms = 0, split_ms = 0, sort_ms = 0, compact_ms = 0, total_ms = 0;
};

// Per-kernel device time (HIP events on the engine stream) and algorithmic
// HBM bytes of the last run, filled when kernel profiling is on.
struct KernelStat {
    std::string name;
    double ms = 0;
    uint64_t launches = 0;
    uint64_t bytes = 0;
};

struct CountTunables {
    int split_target_div = 0;  // split partitions until max child <= largest cap / div (0: by key width)
    int max_digit_bits = 12;
    bool dedup = true;  // small buckets: hash dedup + one bin pass instead of LSD passes
    int hash_classes = -1;  // size classes [0, hash_classes) use the hash kernel (-1: by key width)
    int split_chunk = 1;   // split tiles per workgroup (longer runs per child, fewer atomics)
    int extract_subs = 4;  // sub-tiles per extract workgroup
    bool extract_staged = true;  // group each chunk's keys by partition in LDS before writing
    int split_staged = -1;       // group each split tile's keys by child in LDS (-1: one-word keys)
    int split_above_div = 1;     // split nodes above the largest bucket capacity / this (1 or 2)
};

class CountEngine {
   public:
    explicit CountEngine(hipStream_t s) : stream_(s) {}
    ~CountEngine();
    CountEngine(const CountEngine &) = delete;
    CountEngine &operator=(const CountEngine &) = delete;

    // Count the FASTQ bytes d_fq[0, len).  Results stay on the device until
    // the next run: keys() (n_unique * words u64), counts() (u32), and
    // part_off() (host copy, partitions + 1 entries).
    void run(const uint8_t *d_fq, uint64_t len, int k, int canonical, uint32_t partitions);
    // Sizes the key arrays run() uses for up to max_kmers k-mers of `words`
    // words now, so that a batch of growing inputs does not grow them (and
    // free the smaller ones) sample by sample.
    void reserve(uint64_t max_kmers, int words);

    uint64_t n_kmers() const { return n_kmers_; }
    uint64_t n_unique() const { return n_unique_; }
    int words() const { return words_; }
    const std::vector<uint64_t> &part_off() const { return part_off_h_; }
    const uint64_t *d_part_off() const { return part_u_.as<uint64_t>(); }  // device copy of part_off()
    const uint64_t *keys() const { return out_keys_[slot_].as<uint64_t>(); }
    const uint32_t *counts() const { return out_counts_[slot_].as<uint32_t>(); }
    // Several output slots let a copy stream drain earlier results while the
    // next run writes another slot.
    static constexpr int OUT_SLOTS = 4;
    void set_slot(int s) { slot_ = s % OUT_SLOTS; }
    int slot() const { return slot_; }
    // Frees every device buffer (the next run allocates afresh).
    void release();
    const StageTimes &times() const { return times_; }
    uint64_t buckets() const { return n_buckets_; }
    // buckets the hash sort handed to the LSD sort in the last run
    uint64_t fallback_buckets() const { return n_fallback_; }
    int levels() const { return levels_; }
    void set_timing(bool on) { timing_ = on; }
    void set_kernel_profile(bool on) { kprof_ = on; }
    const std::vector<KernelStat> &kernel_stats() const { return kstats_; }
    // Device-planned counts run eagerly with kernel profiling on: clear the
    // stats before the first, collect them (a sync on the events) after the
    // last; the caller fills in the algorithmic bytes.
    void begin_kernel_stats();
    void collect_kernel_stats();
    std::vector<KernelStat> &kernel_stats_mut() { return kstats_; }
    CountTunables tun;

    // Bytes of device memory the pipeline needs for n k-mers of `words` words
    // and `fastq` input bytes (used by the batch planner).
    static uint64_t bytes_needed(uint64_t n, int words, uint64_t fastq, uint32_t partitions);

    // Settings the next runs start from (an autotuner's choice, or the
    // built-in defaults); KITE_* environment overrides still apply on top.
    void set_base_tunables(const CountTunables &t) { base_ = t; }

    // Device-planned count.  dev_prepare sizes every buffer for inputs up to
    // cap_len bytes.  dev_enqueue_count then puts the whole count of
    // d_fq[0, *len) on the engine stream, with the settings current at that
    // moment, and no host read and no allocation (so that a HIP graph can
    // capture it): `len` is a word the first kernel reads (device or
    // page-locked host memory), the grids are sized for inputs up to
    // max_len (at most cap_len; a longer input is an overflow).
    // dev_enqueue_encode appends the chunk encode of that count's
    // result into `bytes` (bytes_cap bytes, DVLE or RAW per partition, as
    // kite_gpu_count_emit) and its summary into the page-locked `h_summary`
    // (DevSummary, then part_off, byte_off and minmax).  Buffers of the
    // encode live in `enc`.
    void dev_prepare(uint64_t cap_len, int k, uint32_t partitions, EncodeBufs &enc);
    void dev_enqueue_count(const uint8_t *d_fq, const uint64_t *len, uint64_t max_len, int k, int canonical,
                           uint32_t partitions);
    void dev_enqueue_encode(EncodeBufs &enc, uint8_t *bytes, uint64_t bytes_cap, void *h_summary);
    // Stores `len` into the device word `d_len` on the engine stream (one
    // thread; no copy engine involved): the length a later count reads.
    void dev_set_len(uint64_t *d_len, uint64_t len);
    const DevCaps &dev_caps() const { return dcap_; }
    // The settings the next count uses: the base settings with the KITE_*
    // environment overrides applied.
    CountTunables resolved_tunables() const;
    // bytes of page-locked memory a summary of `partitions` takes
    static size_t dev_summary_bytes(uint32_t partitions) {
        return sizeof(DevSummary) + 8ull * (2ull * (partitions + 1) + 8ull * partitions);
    }

   private:
    template <int W>
    void run_w(const uint8_t *d_fq, uint64_t len, int k, int canonical, int pbits);
    template <int W>
    void dev_enqueue_w(const uint8_t *d_fq, const uint64_t *len, uint64_t max_len, int k, int canonical, int pbits);

    // records events around one launch when kernel profiling is on
    template <class F>
    void launch(const char *name, uint64_t bytes, F &&f);

    hipStream_t stream_;
    CountTunables base_;
    bool timing_ = false;
    bool kprof_ = false;
    std::vector<KernelStat> kstats_;
    struct PendingEv {
        int stat;
        hipEvent_t a, b;
    };
    std::vector<PendingEv> pending_;
    std::vector<hipEvent_t> ev_pool_;
    DevBuf nl_, line0_, scan_tmp_, hist_, off_, part_off_d_;
    DevBuf a_, b_, counts_tmp_, out_keys_[OUT_SLOTS], out_counts_[OUT_SLOTS];
    int slot_ = 0;
    DevBuf segs_, tile_pre_, child_hist_, child_cur_, ucount_, uoff_, part_first_, part_u_, fallback_;
    DevBuf order_, wg_hist_;
    // device bucket plan: node lists (ping-pong), per-level scratch, classes
    DevBuf nodes_[2], npart_[2], act_, tiles_, act_rank_, tile_off_, lstats_, seg_node_;
    DevBuf nz_rank_, nz_count_, new_cnt_, new_base_, cls_, cls_count_;
    // device-planned count: its own buffers (a captured graph keeps their
    // addresses, so the classic path never reallocates them) and capacities
    struct DevBufs {
        DevBuf plan, nl, line0, stmp, hist, off, part_off, a, b, ctmp, nodes[2], npart[2], act, tiles, act_rank,
            tile_off, segs, tpre, seg_node, ch, ccur, nzr, nzc, ncnt, nbase, cls, order, fb, ucount, uoff, pfirst,
            pu, keys, counts;
    } dv_;
    DevCaps dcap_;
    uint32_t dP_ = 0;
    int dW_ = 0;
    void *pin_ = nullptr;  // page-locked scratch for small read-backs
    size_t pin_bytes_ = 0;
    uint64_t *pinned_host(size_t bytes);
    std::vector<uint64_t> part_off_h_;
    uint64_t n_kmers_ = 0, n_unique_ = 0, n_buckets_ = 0, n_fallback_ = 0;
    int words_ = 1, levels_ = 0;
    StageTimes times_;
};

}  // namespace kite
