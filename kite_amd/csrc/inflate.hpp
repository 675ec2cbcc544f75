// inflate.hpp -- BGZF blocks inflated on the device (RFC 1951 DEFLATE, one
// block per wave), the block-parallel decompression of the reference's
// BGZF reader (src/count/bgzf.rs:152-264) without its host threads.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace kite {

// One BGZF block: raw deflate data at src (clen bytes) in the staged
// compressed buffer; isize output bytes go to dst in the output buffer.
struct InflateBlock {
    uint64_t src, dst;
    uint32_t clen, isize;
};

// A parsed BGZF member (src/count/bgzf.rs:37-101): header checks, BSIZE,
// the deflate payload and ISIZE.
struct BgzfMember {
    uint64_t offset;   // of the member in the input
    uint32_t size;     // BSIZE + 1
    uint32_t data;     // offset of the deflate data inside the member
    uint32_t clen;     // deflate bytes
    uint32_t isize;    // uncompressed bytes
};

// Scans whole members from data[0, len): stops at a truncated last member
// (its bytes are not consumed); throws on data that is not BGZF.
// *consumed receives the bytes of the whole members.
std::vector<BgzfMember> scan_bgzf(const uint8_t *data, size_t len, size_t *consumed);

// Inflates `blocks` from d_comp into d_out on stream s (asynchronous);
// d_err (one u32, zeroed by the caller) receives an error bit mask:
// 1 = invalid stream, 2 = output size differs from ISIZE.
void inflate_blocks(const uint8_t *d_comp, const std::vector<InflateBlock> &blocks, uint8_t *d_out,
                    uint32_t *d_err, DevBuf &d_blocks, hipStream_t s);

// Adds the number of '\n' bytes in d[0, n) (d 16-byte aligned) to *d_total.
void count_newlines_device(const uint8_t *d, uint64_t n, unsigned long long *d_total, hipStream_t s);

}  // namespace kite
