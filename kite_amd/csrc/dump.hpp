// dump.hpp -- `kite dump` on the device: stored runs of a group of
// partitions are decoded (DVLE or RAW), the runs of each partition are
// merged with u32-saturating sums, the count filter is applied and the
// "KMER\tCOUNT\n" lines are formatted, in partition and key order.
#pragma once

#include <cstdint>
#include <map>
#include <vector>

#include "common.hpp"
#include "merge.hpp"

namespace kite {

constexpr uint32_t DUMP_CODEC_RAW = 0x01u;
constexpr uint32_t DUMP_CODEC_DVLE = 0x02u;

// One stored run of a partition, staged in device memory.
struct DumpRun {
    uint64_t src;    // byte offset of the payload in the staged buffer (16-byte aligned)
    uint64_t bytes;  // payload bytes (DVLE or RAW; zstd already removed)
    uint64_t rows;
    uint64_t row0;   // first row of this run in the decoded row space
    uint64_t dst;    // first row of this run in its merge table
    uint32_t codec;  // DUMP_CODEC_DVLE or DUMP_CODEC_RAW
    uint32_t part;   // partition index within the group
    uint32_t tile0;  // first byte tile of the run (DVLE)
    uint32_t table;  // rank of the run among its partition's runs
};

constexpr uint32_t HISTO_SMALL = 8192;  // totals below this are binned in LDS

struct DumpTimes {
    float decode_ms = 0, merge_ms = 0, format_ms = 0;
};

class DumpEngine {
   public:
    DumpEngine(hipStream_t s, MergeEngine *merge) : stream_(s), merge_(merge) {}
    ~DumpEngine() { release(); }  // frees the page-locked error word too
    DumpEngine(const DumpEngine &) = delete;
    DumpEngine &operator=(const DumpEngine &) = delete;

    // Decodes `runs` (payloads at d_payload, which has 16 readable bytes past
    // the last payload), merges the runs of each of the `np` partitions,
    // filters and formats the lines.  The text stays on the device (text(),
    // returned byte count).  rows_out receives the rows after the merge,
    // lines_out the rows that passed the filter.
    uint64_t run(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int k, int W,
                 const DumpFilter &f, uint64_t *rows_out, uint64_t *lines_out, DumpTimes *times);
    const char *text() const { return text_.as<char>(); }

    // Decodes `runs` and sums the runs of each partition per key with u64
    // saturation (load_partition_counts, src/kiteops/mod.rs:433-473): returns
    // the rows, and device pointers to their keys, u64 counts and the np + 1
    // partition row offsets, valid until the next call on this engine.
    uint64_t decode_sum(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int W,
                        const uint64_t **keys, const uint64_t **counts, const uint64_t **off, DumpTimes *times);

    // `kite histo` of one input over its partition groups: histo_begin,
    // then histo per group (decode, merge with u64 sums, filter on the
    // total, frequency of every total), then histo_end with the bins in
    // ascending order.
    void histo_begin();
    void histo(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int W, const DumpFilter &f,
               uint64_t *rows_out, DumpTimes *times);
    void histo_end(std::vector<uint64_t> &values, std::vector<uint64_t> &freqs);
    void release();

   private:
    template <int W>
    uint64_t decode_merge_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, bool trunc32,
                            const uint64_t **keys, const uint64_t **counts, hipEvent_t ev[3]);
    template <int W>
    uint64_t decode_sum_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np,
                          const uint64_t **keys, const uint64_t **counts, const uint64_t **off, DumpTimes *times);
    template <int W>
    void histo_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, const DumpFilter &f,
                 uint64_t *rows_out, DumpTimes *times);
    template <int W>
    uint64_t run_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int k,
                   const DumpFilter &f, uint64_t *rows_out, uint64_t *lines_out, DumpTimes *times);

    hipStream_t stream_;
    MergeEngine *merge_;
    DevBuf runs_d_, tile_run_, tcnt_, tbase_, ends_, vals_, counts_, scan_tmp_, tlen_, toff_, text_, err_,
        tabptr_, lines_, loff_, hsmall_, hbig_d_, ptile_, ptrun_, psum_, poff_;
    std::map<uint64_t, uint64_t> hbig_;  // frequencies of totals >= HISTO_SMALL
    std::vector<DevTable> tabs_;
    DevTable merged_;
    uint32_t *host_err_ = nullptr;
};

}  // namespace kite
