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

struct Bucket {
    uint64_t start;   // first key (index into the key array)
    uint64_t size;    // keys in the bucket
    uint32_t prefix;  // key bits (from the top of the 2k-bit key) shared by all keys
    uint32_t array;   // 0: keys live in A, 1: in B
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
    CountTunables tun;

    // Bytes of device memory the pipeline needs for n k-mers of `words` words
    // and `fastq` input bytes (used by the batch planner).
    static uint64_t bytes_needed(uint64_t n, int words, uint64_t fastq, uint32_t partitions);

    // Settings the next runs start from (an autotuner's choice, or the
    // built-in defaults); KITE_* environment overrides still apply on top.
    void set_base_tunables(const CountTunables &t) { base_ = t; }

   private:
    template <int W>
    void run_w(const uint8_t *d_fq, uint64_t len, int k, int canonical, int pbits);

    // records events around one launch when kernel profiling is on
    template <class F>
    void launch(const char *name, uint64_t bytes, F &&f);
    void collect_kernel_stats();

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
    void *pin_ = nullptr;  // page-locked scratch for small read-backs
    size_t pin_bytes_ = 0;
    uint64_t *pinned_host(size_t bytes);
    std::vector<uint64_t> part_off_h_;
    uint64_t n_kmers_ = 0, n_unique_ = 0, n_buckets_ = 0, n_fallback_ = 0;
    int words_ = 1, levels_ = 0;
    StageTimes times_;
};

}  // namespace kite
