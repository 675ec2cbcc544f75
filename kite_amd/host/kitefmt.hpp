// kitefmt.hpp -- the .kite byte contract on the host: header, run chunks,
// chunk index, partition table, footer; the staging writer and the reader.
//
// Layout (all little endian), from src/format/header.rs:62-151,
// src/format/bucket_dir.rs:5-140 and src/count/writer.rs:350-495:
//   [0,128)  header: "KX2!" | u16 version=1 | u8 k | u8 canonical | u32 partitions
//            | u16 key_width | u16 codec_mask | u32 checksum_mode=0
//            | u64 chunk_index_offset | u64 partition_table_offset
//            | u64 footer_offset | u64 total_chunks | u64 total_rows
//            | 68 reserved bytes: "KMT1" | u32 flags (1 min, 2 max) | u32 min | u32 max
//   chunk payloads, then the chunk index (128 B per chunk, grouped by
//   partition, sorted by offset), the partition table (40 B each) and the
//   footer (u64 chunks | u64 rows | "KX2FOOT!").
#pragma once

#include <cstdint>
#include <cstdio>
#include <optional>
#include <stdexcept>
#include <string>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace kh {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

constexpr uint32_t CODEC_RAW = 1, CODEC_DVLE = 2, CODEC_ZSTD = 4;

inline int words_for_width(uint16_t key_width) { return key_width == 128 ? 2 : (key_width == 256 ? 4 : 1); }
inline uint16_t width_for_k(int k) { return k <= 31 ? 64 : (k <= 64 ? 128 : 256); }
inline int entry_size(uint16_t key_width) { return key_width == 128 ? 20 : (key_width == 256 ? 36 : 12); }

// Up to 256-bit key as little-endian words; ordering is the unsigned integer
// order (for 4-word keys: (hi, lo), src/count/kmer.rs:31-35).
struct Key {
    uint64_t w[4] = {0, 0, 0, 0};
    bool operator<(const Key &o) const {
        for (int i = 3; i >= 0; i--)
            if (w[i] != o.w[i]) return w[i] < o.w[i];
        return false;
    }
    bool operator==(const Key &o) const {
        return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3];
    }
    bool operator!=(const Key &o) const { return !(*this == o); }
};

struct Row {
    Key key;
    uint64_t count;
};

struct Header {
    uint8_t k = 0;
    uint8_t canonical = 0;
    uint32_t partitions = 0;
    uint16_t key_width = 64;
    uint16_t codec_mask = 0;
    uint64_t chunk_index_offset = 0, partition_table_offset = 0, footer_offset = 0;
    uint64_t total_chunks = 0, total_rows = 0;
    std::optional<uint32_t> min_count, max_count;

    void write(uint8_t out[128]) const;
    static Header parse(const uint8_t *in);  // throws on bad magic / version
};

struct ChunkMeta {
    uint32_t partition = 0, codec = 0;
    uint64_t offset = 0, encoded = 0, raw = 0, rows = 0;
    uint64_t kmin[4] = {0, 0, 0, 0}, kmax[4] = {0, 0, 0, 0};  // (lo128, hi128)
};

// std::vector allocator whose resize leaves new elements uninitialised, so
// that growing a byte buffer before a copy into it does not zero it first.
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
    void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void *>(p)) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
using Bytes = std::vector<uint8_t, NoInitAlloc<uint8_t>>;

// ConcurrentKiteWriter (src/count/writer.rs:71-495) restated for appends
// from one thread: per-partition staging buffers of `staging_bytes`, chunks
// at least that large written directly, excess staging drained largest
// partition first, and the trailer plus final header written by finalize().
class Writer {
   public:
    Writer(const std::string &path, const Header &h, size_t staging_bytes, int write_lanes);
    ~Writer();
    void append_chunk(uint32_t partition, uint32_t codec, uint64_t rows, uint64_t raw_bytes,
                      const uint64_t kmin[4], const uint64_t kmax[4], const uint8_t *data, size_t len);
    // returns (total_chunks, total_rows)
    std::pair<uint64_t, uint64_t> finalize();
    static int suggest_write_lanes(int workers);

   private:
    struct Pending {
        ChunkMeta m;
    };
    struct Stage {
        Bytes buf;
        std::vector<ChunkMeta> pending, committed;
    };
    uint64_t claim(const uint8_t *data, size_t len);
    void flush_partition(uint32_t p);
    bool drain_one();
    // regular files: flushed staging buffers are written at their claimed
    // offsets by `write_lanes` threads (the file layout does not change)
    struct Job {
        uint64_t off;
        Bytes buf;
    };
    void write_at(const uint8_t *data, size_t len, uint64_t off);
    void worker();
    void wait_writes();
    void stop_writers();

    FILE *f_ = nullptr;
    int fd_ = -1;
    bool regular_ = true;
    bool finalized_ = false;
    Header h_;
    uint64_t next_offset_ = 128;
    size_t S_, high_, low_, staged_ = 0;
    std::vector<Stage> st_;
    std::vector<std::thread> pool_;
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<Job> jobs_;
    size_t busy_ = 0, queued_bytes_ = 0;
    bool stop_ = false;
    std::string werr_;
    // written staging buffers kept for reuse (their pages stay mapped, so a
    // staged chunk is a plain copy rather than a copy plus page faults)
    std::vector<Bytes> spare_;
    size_t spare_cap_ = 0;
};

// Memory-mapped reader with the validation of KiteReader::open
// (src/dump/reader.rs:43-154).
class Reader {
   public:
    explicit Reader(const std::string &path);
    ~Reader();
    Reader(const Reader &) = delete;
    Reader &operator=(const Reader &) = delete;

    const Header &header() const { return h_; }
    int words() const { return words_for_width(h_.key_width); }
    uint32_t partitions() const { return h_.partitions; }
    uint64_t chunk_count() const { return chunks_.size(); }
    const std::vector<ChunkMeta> &chunks() const { return chunks_; }
    // chunks of partition p: [first, first + count)
    std::pair<uint64_t, uint64_t> partition_chunks(uint32_t p) const { return part_[p]; }
    // Payload of one chunk, zstd-decoded when needed.
    std::vector<uint8_t> payload(const ChunkMeta &c) const;
    const uint8_t *raw_payload(const ChunkMeta &c) const { return base_ + c.offset; }

    // Decoded rows of one run, in stored order (counts as stored).
    void decode_run(const ChunkMeta &c, std::vector<Row> &out) const;
    // All runs of partition p, sorted by key with equal keys summed
    // (u64 saturating): load_partition_counts (src/kiteops/mod.rs:433-473).
    std::vector<Row> load_partition(uint32_t p) const;
    uint64_t file_size() const { return size_; }

   private:
    const uint8_t *base_ = nullptr;
    size_t size_ = 0;
    Header h_;
    std::vector<ChunkMeta> chunks_;
    std::vector<std::pair<uint64_t, uint64_t>> part_;
};

// Count filter (src/kiteops/mod.rs:44-56): inclusive bounds.
inline bool passes(uint64_t c, const std::optional<uint32_t> &mn, const std::optional<uint32_t> &mx) {
    if (mn && c < *mn) return false;
    if (mx && c > *mx) return false;
    return true;
}

// LEB128 / DVLE (src/format/run_block.rs:9-95).
size_t put_varint(uint8_t *o, uint64_t lo, uint64_t hi);
// Serialize one sorted, unique partition as a chunk payload (DVLE, or RAW
// for 4-word keys).  Throws if a count exceeds u32 (src/kiteops/mod.rs:804-859).
std::vector<uint8_t> serialize_rows(const std::vector<Row> &rows, int words, uint32_t *codec);

// Key -> ACGT text exactly as the reference's dump prints it
// (src/dump/decoder.rs:27-107, including its layout for k >= 65).
void decode_key(const Key &k, int words, int kk, char *buf);

// File name based label (src/kiteops/mod.rs:123-165).
std::string infer_label(const std::string &path, size_t index);

// zstd through the system libzstd.so.1 (loaded on first use).
bool zstd_available();
std::vector<uint8_t> zstd_decompress(const uint8_t *src, size_t len, size_t raw_size);
std::vector<uint8_t> zstd_compress(const uint8_t *src, size_t len, int level);

}  // namespace kh
