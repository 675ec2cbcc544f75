// encode.hpp -- device encoding of reduced (key, count) runs into .kite
// chunk payloads (DVLE for 1- and 2-word keys, RAW for 4-word keys).
#pragma once

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "common.hpp"

namespace kite {

// Page-locked host buffer that grows on demand.
struct HostPinned {
    void *ptr = nullptr;
    size_t bytes = 0;
    HostPinned() = default;
    HostPinned(const HostPinned &) = delete;
    HostPinned &operator=(const HostPinned &) = delete;
    ~HostPinned() { release(); }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void *ensure(size_t need) {
        if (need == 0) need = 256;
        if (need > bytes) {
            release();
            const size_t want = need + need / 8;
            KITE_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
            bytes = want;
        }
        return ptr;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

struct EncodeBufs {
    DevBuf head, len, off, tmp, byte_off, minmax, bytes;
    HostPinned host_byte_off;
};

// Device result of an encode: partition p's payload is bytes[byte_off[p],
// byte_off[p + 1]) (byte_off on the host, written by the encoder), the key
// range minmax[8p .. 8p+7] (device).
struct EncodeDevice {
    const uint8_t *bytes = nullptr;
    const uint64_t *minmax = nullptr;
    uint64_t total = 0;
};

// One payload per partition, back to back: partition p is
// bytes[byte_off[p], byte_off[p + 1]) and holds rows_off[p + 1] - rows_off[p]
// rows; minmax[8p .. 8p+3] is its first key, minmax[8p+4 .. 8p+7] its last
// (4 words each, unused words 0).
// Allocator whose resize leaves new elements uninitialised (a payload that
// a device copy fills at once need not be zeroed first).
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};

struct EncodedRuns {
    std::vector<uint8_t, NoInitAlloc<uint8_t>> bytes;
    std::vector<uint64_t> byte_off, rows_off, minmax;
    uint32_t codec = 0;
    int words = 1;
};

// Runs the encode kernels on stream s and returns with the payload bytes and
// key ranges on the device; host_byte_off (partitions + 1 entries, page-locked)
// receives the byte offsets.  The stream is synchronised once, for the total.
EncodeDevice encode_on_device(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off, uint64_t n,
                              uint32_t P, int W, hipStream_t s, EncodeBufs &bufs, uint64_t *host_byte_off);

// The encode of a launch sequence fixed before the data is seen: the row
// count is read from *dn (at most n_cap; encode_prepare_dev sizes the
// scratch for it), the payload goes to bytes (bytes_cap), its size to
// *total, and nothing is written when *overflow is set (a total above
// bytes_cap sets bit 4).  Byte offsets and key ranges stay in
// bufs.byte_off / bufs.minmax on the device.
void encode_prepare_dev(uint64_t n_cap, uint32_t P, EncodeBufs &bufs);
void encode_dev(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off, const uint64_t *dn,
                uint64_t n_cap, uint32_t P, int W, hipStream_t s, EncodeBufs &bufs, uint8_t *bytes,
                uint64_t bytes_cap, uint64_t *total, uint32_t *overflow);

// d_keys / d_counts: n reduced rows on the device, partition-ordered;
// d_part_off: the partitions + 1 row offsets on the device (part_off the same
// on the host).  Downloads only the encoded bytes and the chunk metadata.
void encode_runs_device(const uint64_t *d_keys, const uint32_t *d_counts, const uint64_t *d_part_off,
                        const std::vector<uint64_t> &part_off, uint32_t P, int W, hipStream_t s, EncodeBufs &bufs,
                        EncodedRuns &out);

}  // namespace kite
