// dvle.hpp -- host encoding of reduced (key, count) runs into .kite chunk
// payloads: DVLE for 1- and 2-word keys, RAW for 4-word keys
// (src/format/run_block.rs:9-95, src/count/kmer.rs:187-192).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace kite {

struct EncodedChunk {
    std::vector<uint8_t> bytes;
    uint64_t rows = 0;
    uint32_t codec = 0;
    uint64_t key_min[4] = {0, 0, 0, 0};
    uint64_t key_max[4] = {0, 0, 0, 0};
};

// LEB128 of a value up to 128 bits (write_varint_u64/u128).
inline size_t put_varint128(uint8_t *o, uint64_t lo, uint64_t hi) {
    size_t n = 0;
    for (;;) {
        uint8_t b = (uint8_t)(lo & 0x7f);
        lo = (lo >> 7) | (hi << 57);
        hi >>= 7;
        if (!lo && !hi) {
            o[n++] = b;
            return n;
        }
        o[n++] = b | 0x80;
    }
}

inline size_t max_encoded_bytes(size_t rows, int W) {
    if (W == 4) return rows * 36;
    return rows ? (8 * (size_t)W + 4) + (rows - 1) * (W == 1 ? 15 : 24) : 0;
}

// Encodes one sorted, unique run.  Returns the payload length.
inline size_t encode_run(const uint64_t *keys, const uint32_t *counts, size_t n, int W, uint8_t *out) {
    if (n == 0) return 0;
    if (W == 4) {
        for (size_t i = 0; i < n; i++) {
            memcpy(out + 36 * i, keys + 4 * i, 32);
            memcpy(out + 36 * i + 32, counts + i, 4);
        }
        return 36 * n;
    }
    size_t p = 8 * (size_t)W + 4;
    memcpy(out, keys, 8 * (size_t)W);
    memcpy(out + 8 * W, counts, 4);
    if (W == 1) {
        uint64_t prev = keys[0];
        for (size_t i = 1; i < n; i++) {
            p += put_varint128(out + p, keys[i] - prev, 0);
            p += put_varint128(out + p, counts[i], 0);
            prev = keys[i];
        }
    } else {
        unsigned __int128 prev = ((unsigned __int128)keys[1] << 64) | keys[0];
        for (size_t i = 1; i < n; i++) {
            unsigned __int128 v = ((unsigned __int128)keys[2 * i + 1] << 64) | keys[2 * i];
            unsigned __int128 d = v - prev;
            p += put_varint128(out + p, (uint64_t)d, (uint64_t)(d >> 64));
            p += put_varint128(out + p, counts[i], 0);
            prev = v;
        }
    }
    return p;
}

// Encodes every partition of a result in parallel.
inline std::vector<EncodedChunk> encode_partitions(const uint64_t *keys, const uint32_t *counts,
                                                   const uint64_t *part_off, uint32_t P, int W,
                                                   unsigned threads = 0) {
    std::vector<EncodedChunk> out(P);
    if (!threads) threads = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    threads = std::min<unsigned>(threads, std::max<uint32_t>(1, P));
    auto work = [&](unsigned t) {
        for (uint32_t p = t; p < P; p += threads) {
            const uint64_t a = part_off[p], n = part_off[p + 1] - a;
            EncodedChunk &c = out[p];
            c.rows = n;
            if (!n) continue;
            c.codec = W == 4 ? 0x01u : 0x02u;
            c.bytes.resize(max_encoded_bytes(n, W));
            c.bytes.resize(encode_run(keys + a * W, counts + a, n, W, c.bytes.data()));
            for (int r = 0; r < W; r++) {
                c.key_min[r] = keys[a * W + r];
                c.key_max[r] = keys[(a + n - 1) * W + r];
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < threads; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    return out;
}

}  // namespace kite
