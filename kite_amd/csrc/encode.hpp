// encode.hpp -- device encoding of reduced (key, count) runs into .kite
// chunk payloads (DVLE for 1- and 2-word keys, RAW for 4-word keys).
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace kite {

struct EncodeBufs {
    DevBuf head, len, off, tmp, byte_off, minmax, bytes;
};

// One payload per partition, back to back: partition p is
// bytes[byte_off[p], byte_off[p + 1]) and holds rows_off[p + 1] - rows_off[p]
// rows; minmax[8p .. 8p+3] is its first key, minmax[8p+4 .. 8p+7] its last
// (4 words each, unused words 0).
struct EncodedRuns {
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> byte_off, rows_off, minmax;
    uint32_t codec = 0;
    int words = 1;
};

// d_keys / d_counts: n reduced rows on the device, partition-ordered;
// d_part_off: the partitions + 1 row offsets on the device (part_off the same
// on the host).  Downloads only the encoded bytes and the chunk metadata.
void encode_runs_device(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off,
                        const std::vector<uint64_t> &part_off, uint32_t P, int W, hipStream_t s, EncodeBufs &bufs,
                        EncodedRuns &out);

}  // namespace kite
