// merge_host.hpp -- `kite merge` data path: decode each input's partitions,
// apply its header filter, union + sum on the GPU, serialize in order.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "kitefmt.hpp"

namespace kh {

struct MergeStats {
    std::string engine;
    double merge_s = 0;
};

MergeStats merge_files(const std::vector<std::unique_ptr<Reader>> &rs, Writer &w, unsigned threads, bool gpu);

// Loads partitions [p0, p1) of one input with its header filter applied:
// keys (rows x words), u64 counts and p1 - p0 + 1 offsets.
void load_filtered(const Reader &r, uint32_t p0, uint32_t p1, unsigned threads, std::vector<uint64_t> &keys,
                   std::vector<uint64_t> &counts, std::vector<uint64_t> &off);

}  // namespace kh
