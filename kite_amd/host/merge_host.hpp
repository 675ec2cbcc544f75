// merge_host.hpp -- `kite merge` data path.  The device engine takes every
// input's stored runs and returns the merged partitions as chunk payloads;
// the host engine decodes each input's partitions, applies its header
// filter, merges them on worker threads and serializes them in order.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "kite_gpu.h"
#include "kitefmt.hpp"

namespace kh {

struct MergeStats {
    std::string engine;
    double merge_s = 0;  // device engine: the kite_gpu_merge_runs call; host engine: the merges alone
};

// Every stored run of a file in partition order, as the device engines
// take them; zstd payloads are decoded into `decoded` first.
void collect_runs(const Reader &r, std::vector<kite_dump_run> &runs, std::vector<std::vector<uint8_t>> &decoded);

MergeStats merge_files(const std::vector<std::unique_ptr<Reader>> &rs, Writer &w, unsigned threads, bool gpu);

// Loads partitions [p0, p1) of one input with its header filter applied:
// keys (rows x words), u64 counts and p1 - p0 + 1 offsets.
void load_filtered(const Reader &r, uint32_t p0, uint32_t p1, unsigned threads, std::vector<uint64_t> &keys,
                   std::vector<uint64_t> &counts, std::vector<uint64_t> &off);

// Exact u64 count of `key` in rows [lo, hi) of one input's loaded table (a
// union cell saturated at u32::MAX is re-read from here).  Throws when the
// key is absent.
uint64_t exact_count(const std::vector<uint64_t> &K, const std::vector<uint64_t> &C, uint64_t lo, uint64_t hi,
                     const uint64_t *key, int W);

}  // namespace kh
