// encode.hip -- .kite chunk payloads built on the device: DVLE for 1- and
// 2-word keys (first row: key little-endian + u32 count; then LEB128 of the
// key delta and LEB128 of the count), RAW rows (key lo16 | hi16 | count4) for
// 4-word keys.  The same bytes as the host's serialize_rows (kite_amd/host/
// kitefmt.cpp, used by merge) and as the reference's chunk writer
// (src/format/run_block.rs:9-35, 70-95; src/count/kmer.rs:187-192).
#include <algorithm>

#include "encode.hpp"

namespace kite {

namespace {

__device__ __forceinline__ uint32_t varint_len64(uint64_t v) {
    const uint32_t bits = v ? 64u - (uint32_t)__clzll((long long)v) : 1u;
    return (bits + 6u) / 7u;
}

__device__ __forceinline__ uint32_t varint_len128(uint64_t lo, uint64_t hi) {
    if (!hi) return varint_len64(lo);
    const uint32_t bits = 128u - (uint32_t)__clzll((long long)hi);
    return (bits + 6u) / 7u;
}

__device__ __forceinline__ uint32_t put_varint(uint8_t *o, uint64_t lo, uint64_t hi) {
    uint32_t n = 0;
    for (;;) {
        const uint8_t b = (uint8_t)(lo & 0x7f);
        lo = (lo >> 7) | (hi << 57);
        hi >>= 7;
        if (!lo && !hi) {
            o[n++] = b;
            return n;
        }
        o[n++] = b | 0x80;
    }
}

// Delta of entry i to entry i-1 as 128 bits (W = 1 or 2).
template <int W>
__device__ __forceinline__ void key_delta(const uint64_t *k, uint64_t i, uint64_t &lo, uint64_t &hi) {
    if constexpr (W == 1) {
        lo = k[i] - k[i - 1];
        hi = 0;
    } else {
        const uint64_t a0 = k[2 * i], a1 = k[2 * i + 1], b0 = k[2 * i - 2], b1 = k[2 * i - 1];
        lo = a0 - b0;
        hi = a1 - b1 - (a0 < b0 ? 1ull : 0ull);
    }
}

__global__ void k_zero_u8(uint8_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_mark_heads(const uint64_t *__restrict__ part_off, uint32_t P, uint8_t *__restrict__ head) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P && part_off[p] < part_off[p + 1]) head[part_off[p]] = 1;
}

template <int W>
__device__ __forceinline__ void entry_len_at(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                                             const uint8_t *__restrict__ head, uint64_t i,
                                             uint32_t *__restrict__ len) {
    if constexpr (W == 4) {
        len[i] = 36;
    } else {
        if (head[i]) {
            len[i] = 8 * W + 4;
        } else {
            uint64_t lo, hi;
            key_delta<W>(keys, i, lo, hi);
            len[i] = varint_len128(lo, hi) + varint_len64(counts[i]);
        }
    }
}

template <int W>
__global__ void k_entry_len(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                            const uint8_t *__restrict__ head, uint64_t n, uint32_t *__restrict__ len) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) entry_len_at<W>(keys, counts, head, i, len);
}

template <int W>
__device__ __forceinline__ void entry_write_at(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                                               const uint8_t *__restrict__ head, const uint64_t *__restrict__ off,
                                               uint64_t i, uint8_t *__restrict__ out) {
    uint8_t *o = out + off[i];
    if (W == 4 || head[i]) {
        // raw row: key words little-endian, then the count
#pragma unroll
        for (int r = 0; r < W; r++) {
            const uint64_t v = keys[i * W + r];
#pragma unroll
            for (int b = 0; b < 8; b++) o[8 * r + b] = (uint8_t)(v >> (8 * b));
        }
        const uint32_t c = counts[i];
#pragma unroll
        for (int b = 0; b < 4; b++) o[8 * W + b] = (uint8_t)(c >> (8 * b));
        return;
    }
    if constexpr (W != 4) {
        uint64_t lo, hi;
        key_delta<W>(keys, i, lo, hi);
        const uint32_t m = put_varint(o, lo, hi);
        put_varint(o + m, counts[i], 0);
    }
}

template <int W>
__global__ void k_entry_write(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                              const uint8_t *__restrict__ head, const uint64_t *__restrict__ off, uint64_t n,
                              uint8_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) entry_write_at<W>(keys, counts, head, off, i, out);
}

// ---- the same with the row count read from device memory (fixed grids) ------
__global__ void k_zero_u8_dev(uint8_t *__restrict__ p, const uint64_t *__restrict__ dn,
                              const uint32_t *__restrict__ overflow) {
    if (*overflow) return;
    const uint64_t n = *dn;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

template <int W>
__global__ void k_entry_len_dev(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                                const uint8_t *__restrict__ head, const uint64_t *__restrict__ dn,
                                const uint32_t *__restrict__ overflow, uint32_t *__restrict__ len) {
    if (*overflow) return;
    const uint64_t n = *dn;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        entry_len_at<W>(keys, counts, head, i, len);
}

__global__ void k_mark_heads_dev(const uint64_t *__restrict__ part_off, uint32_t P, uint8_t *__restrict__ head,
                                 const uint32_t *__restrict__ overflow) {
    if (*overflow) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P && part_off[p] < part_off[p + 1]) head[part_off[p]] = 1;
}

// The payload total; beyond the output's capacity the count is void.
__global__ void k_bytes_total(const uint64_t *__restrict__ off, const uint64_t *__restrict__ dn, uint64_t cap,
                              uint64_t *__restrict__ total, uint32_t *__restrict__ overflow) {
    if (*overflow) return;
    const uint64_t t = off[*dn];
    *total = t;
    if (t > cap) *overflow |= 16u;
}

template <int W>
__global__ void k_entry_write_dev(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ counts,
                                  const uint8_t *__restrict__ head, const uint64_t *__restrict__ off,
                                  const uint64_t *__restrict__ dn, const uint32_t *__restrict__ overflow,
                                  uint8_t *__restrict__ out) {
    if (*overflow) return;
    const uint64_t n = *dn;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        entry_write_at<W>(keys, counts, head, off, i, out);
}

// Per partition: byte offset of its chunk and its first and last keys.
template <int W>
__device__ __forceinline__ void chunk_meta_at(const uint64_t *__restrict__ part_off, uint32_t P, uint32_t p,
                                              const uint64_t *__restrict__ off, const uint64_t *__restrict__ keys,
                                              uint64_t *__restrict__ byte_off, uint64_t *__restrict__ minmax) {
    const uint64_t a = part_off[p];
    byte_off[p] = off[a];  // off has n + 1 entries: off[n] is the total
    if (p == P) return;
    const uint64_t b = part_off[p + 1];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        minmax[(uint64_t)p * 8 + r] = (a < b && r < W) ? keys[a * W + r] : 0;
        minmax[(uint64_t)p * 8 + 4 + r] = (a < b && r < W) ? keys[(b - 1) * W + r] : 0;
    }
}

template <int W>
__global__ void k_chunk_meta(const uint64_t *__restrict__ part_off, uint32_t P, const uint64_t *__restrict__ off,
                             const uint64_t *__restrict__ keys, uint64_t *__restrict__ byte_off,
                             uint64_t *__restrict__ minmax) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= P) chunk_meta_at<W>(part_off, P, p, off, keys, byte_off, minmax);
}

template <int W>
__global__ void k_chunk_meta_dev(const uint64_t *__restrict__ part_off, uint32_t P, const uint64_t *__restrict__ off,
                                 const uint64_t *__restrict__ keys, uint64_t *__restrict__ byte_off,
                                 uint64_t *__restrict__ minmax, const uint32_t *__restrict__ overflow) {
    if (*overflow) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= P) chunk_meta_at<W>(part_off, P, p, off, keys, byte_off, minmax);
}

__global__ void k_put_host_u64(const uint64_t *__restrict__ src, uint32_t n, uint64_t *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <int W>
EncodeDevice encode_w(const uint64_t *keys, const uint32_t *counts, const uint64_t *part_off, uint64_t n, uint32_t P,
                      hipStream_t s, EncodeBufs &bufs, uint64_t *host_byte_off) {
    const unsigned T = 256;
    const dim3 gn((unsigned)std::max<uint64_t>(1, ceil_div(n, T))), gp((unsigned)ceil_div((uint64_t)P + 1, T));
    uint8_t *head = static_cast<uint8_t *>(bufs.head.ensure(std::max<uint64_t>(n, 1)));
    uint32_t *len = static_cast<uint32_t *>(bufs.len.ensure(std::max<uint64_t>(n, 1) * 4));
    uint64_t *off = static_cast<uint64_t *>(bufs.off.ensure((n + 1) * 8));
    uint64_t *tmp = static_cast<uint64_t *>(bufs.tmp.ensure(scan_temp_elems(n + 1) * 8));
    uint64_t *boff = static_cast<uint64_t *>(bufs.byte_off.ensure(((uint64_t)P + 1) * 8));
    uint64_t *mm = static_cast<uint64_t *>(bufs.minmax.ensure((uint64_t)P * 64));
    k_zero_u8<<<gn, T, 0, s>>>(head, n);
    KITE_CHECK_LAUNCH();
    k_mark_heads<<<gp, T, 0, s>>>(part_off, P, head);
    KITE_CHECK_LAUNCH();
    if (n) {
        k_entry_len<W><<<gn, T, 0, s>>>(keys, counts, head, n, len);
        KITE_CHECK_LAUNCH();
    }
    scan_to_u64(len, off, n, tmp, s);
    k_chunk_meta<W><<<gp, T, 0, s>>>(part_off, P, off, keys, boff, mm);
    KITE_CHECK_LAUNCH();
    // the offsets go to the host through a kernel write, not a copy engine
    k_put_host_u64<<<gp, T, 0, s>>>(boff, P + 1, host_byte_off);
    KITE_CHECK_LAUNCH();
    KITE_HIP(hipStreamSynchronize(s));
    EncodeDevice ed;
    ed.total = host_byte_off[P];
    uint8_t *bytes = static_cast<uint8_t *>(bufs.bytes.ensure(std::max<uint64_t>(ed.total, 1)));
    if (n) {
        k_entry_write<W><<<gn, T, 0, s>>>(keys, counts, head, off, n, bytes);
        KITE_CHECK_LAUNCH();
    }
    ed.bytes = bytes;
    ed.minmax = mm;
    return ed;
}

template <int W>
void encode_dev_w(const uint64_t *keys, const uint32_t *counts, const uint64_t *part_off, const uint64_t *dn,
                  uint64_t n_cap, uint32_t P, hipStream_t s, EncodeBufs &bufs, uint8_t *bytes, uint64_t bytes_cap,
                  uint64_t *total, uint32_t *overflow) {
    const unsigned T = 256;
    const dim3 g((unsigned)std::min<uint64_t>(std::max<uint64_t>(1, ceil_div(n_cap, T)), 4096));
    const dim3 gp((unsigned)ceil_div((uint64_t)P + 1, T));
    uint8_t *head = bufs.head.as<uint8_t>();
    uint32_t *len = bufs.len.as<uint32_t>();
    uint64_t *off = bufs.off.as<uint64_t>();
    k_zero_u8_dev<<<g, T, 0, s>>>(head, dn, overflow);
    KITE_CHECK_LAUNCH();
    k_mark_heads_dev<<<gp, T, 0, s>>>(part_off, P, head, overflow);
    KITE_CHECK_LAUNCH();
    k_entry_len_dev<W><<<g, T, 0, s>>>(keys, counts, head, dn, overflow, len);
    KITE_CHECK_LAUNCH();
    scan_to_u64_dev(len, off, dn, n_cap, bufs.tmp.as<uint64_t>(), s);
    k_chunk_meta_dev<W><<<gp, T, 0, s>>>(part_off, P, off, keys, bufs.byte_off.as<uint64_t>(),
                                         bufs.minmax.as<uint64_t>(), overflow);
    KITE_CHECK_LAUNCH();
    k_bytes_total<<<1, 1, 0, s>>>(off, dn, bytes_cap, total, overflow);
    KITE_CHECK_LAUNCH();
    k_entry_write_dev<W><<<g, T, 0, s>>>(keys, counts, head, off, dn, overflow, bytes);
    KITE_CHECK_LAUNCH();
}

}  // namespace

void encode_prepare_dev(uint64_t n_cap, uint32_t P, EncodeBufs &bufs) {
    bufs.head.ensure(std::max<uint64_t>(n_cap, 1));
    bufs.len.ensure(std::max<uint64_t>(n_cap, 1) * 4);
    bufs.off.ensure((n_cap + 1) * 8);
    bufs.tmp.ensure(scan_temp_elems(n_cap + 1) * 8);
    bufs.byte_off.ensure(((uint64_t)P + 1) * 8);
    bufs.minmax.ensure((uint64_t)P * 64);
}

void encode_dev(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off, const uint64_t *dn,
                uint64_t n_cap, uint32_t P, int W, hipStream_t s, EncodeBufs &bufs, uint8_t *bytes,
                uint64_t bytes_cap, uint64_t *total, uint32_t *overflow) {
    if (W == 1) encode_dev_w<1>(d_keys, d_counts, d_part_off, dn, n_cap, P, s, bufs, bytes, bytes_cap, total, overflow);
    else if (W == 2)
        encode_dev_w<2>(d_keys, d_counts, d_part_off, dn, n_cap, P, s, bufs, bytes, bytes_cap, total, overflow);
    else encode_dev_w<4>(d_keys, d_counts, d_part_off, dn, n_cap, P, s, bufs, bytes, bytes_cap, total, overflow);
}

EncodeDevice encode_on_device(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off, uint64_t n,
                              uint32_t P, int W, hipStream_t s, EncodeBufs &bufs, uint64_t *host_byte_off) {
    if (W == 1) return encode_w<1>(d_keys, d_counts, d_part_off, n, P, s, bufs, host_byte_off);
    if (W == 2) return encode_w<2>(d_keys, d_counts, d_part_off, n, P, s, bufs, host_byte_off);
    return encode_w<4>(d_keys, d_counts, d_part_off, n, P, s, bufs, host_byte_off);
}

void encode_runs_device(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off,
                        const std::vector<uint64_t> &part_off, uint32_t P, int W, hipStream_t s, EncodeBufs &bufs,
                        EncodedRuns &out) {
    const uint64_t n = part_off.empty() ? 0 : part_off[P];
    out.rows_off = part_off;
    out.codec = W == 4 ? 0x01u : 0x02u;
    out.words = W;
    uint64_t *hb = static_cast<uint64_t *>(bufs.host_byte_off.ensure(((size_t)P + 1) * 8));
    const EncodeDevice ed = encode_on_device(d_keys, d_counts, d_part_off, n, P, W, s, bufs, hb);
    out.byte_off.assign(hb, hb + P + 1);
    out.bytes.resize(ed.total);
    out.minmax.resize((size_t)P * 8);
    if (ed.total) KITE_HIP(hipMemcpyAsync(out.bytes.data(), ed.bytes, ed.total, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipMemcpyAsync(out.minmax.data(), ed.minmax, (size_t)P * 64, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
}

}  // namespace kite
