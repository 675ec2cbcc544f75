// merge.hpp -- GPU sorted union of per-partition (key, count) tables.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"

namespace kite {

// A table in host memory: rows * W key words, rows u64 counts, P+1 offsets.
struct HostTable {
    const uint64_t *keys = nullptr;
    const uint64_t *counts = nullptr;
    std::vector<uint64_t> part_off;
    uint64_t rows() const { return part_off.empty() ? 0 : part_off.back(); }
};

// Inclusive count bounds of a table's header (count_passes_filter,
// src/kiteops/mod.rs:44-56), compared with u64 counts.
struct DumpFilter {
    uint32_t has_min = 0, min = 0, has_max = 0, max = 0;
};

struct DevTable {
    DevBuf keys, counts, off;
    uint64_t rows = 0;
};

class MergeEngine {
   public:
    explicit MergeEngine(hipStream_t s) : stream_(s) {}
    void upload_table(const HostTable &t, DevTable &d, int W);
    // Sorted union of all tables with u64-saturating count sums.
    void merge_all(std::vector<DevTable> &tabs, uint32_t P, int W, DevTable &result);
    // Count matrix (union rows x tables) of the union `u`: u32 cells
    // saturated at u32::MAX (cell_bytes 4), or exact u64 cells (8).
    void union_matrix(const std::vector<DevTable> &tabs, const DevTable &u, uint32_t P, int W, DevBuf &matrix,
                      int cell_bytes = 4);
    // Largest count over the tables.
    uint64_t max_count(const std::vector<DevTable> &tabs);
    void download(const DevTable &t, int W, uint32_t P, uint64_t *keys, uint64_t *counts, uint64_t *part_off);
    // The rows of a device table (n rows, u64 counts, P + 1 partition
    // offsets) whose count passes `f`, copied into `out` with their offsets
    // (apply_reader_filter_to_rows, src/kiteops/mod.rs:103-108).
    void filter_into(const uint64_t *keys, const uint64_t *counts, const uint64_t *off, uint64_t n, uint32_t P, int W,
                     const DumpFilter &f, DevTable &out);
    // u64 counts narrowed to u32 into out; returns the first row whose count
    // exceeds u32::MAX, or n when none does.
    uint64_t narrow_u32(const uint64_t *counts, uint64_t n, uint32_t *out);

    template <int W>
    void merge2(const DevTable &A, const DevTable &B, DevTable &out, uint32_t P);

   private:
    hipStream_t stream_;
    DevBuf mk_, mc_, head_, pos_, scan_tmp_, tile_pre_, bad_;
};

}  // namespace kite
