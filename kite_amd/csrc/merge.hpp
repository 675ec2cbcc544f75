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

    template <int W>
    void merge2(const DevTable &A, const DevTable &B, DevTable &out, uint32_t P);

   private:
    hipStream_t stream_;
    DevBuf mk_, mc_, head_, pos_, scan_tmp_, tile_pre_;
};

}  // namespace kite
