// inflate.hip -- DEFLATE (RFC 1951) decoding of BGZF blocks on the device.
//
// BGZF (SAM/BAM specification, section 4.1) is gzip made of independent
// members of at most 64 KiB of output, each with its compressed size in the
// BC extra field and its output size (ISIZE) in the trailer.  The reference
// scans the members and inflates them on a thread pool
// (src/count/bgzf.rs:37-101, 152-264); here every member is one device
// workgroup (one wave), writing straight to its offset in the FASTQ staging
// buffer, so the decompressed text never crosses PCIe.
//
// Decoding follows RFC 1951: stored, fixed-Huffman and dynamic-Huffman
// blocks; canonical Huffman codes (section 3.2.2) decoded through a direct
// table of their first 10 (8) bits, longer codes from per-length code counts
// and the symbol list sorted by code; length and distance bases and extra
// bits from section 3.2.5; back references up to 32 KiB read from the
// member's own output.
#include <algorithm>
#include <cstring>

#include "inflate.hpp"

namespace kite {

namespace {

// RFC 1951 3.2.5: the base and extra bits of length code 257 + s (s < 29)
// and of distance code d (d < 30), from the closed forms the tables follow:
// after the first 8 lengths (3..10) each group of four codes doubles the step
// and adds an extra bit, up to code 284; code 285 is 258.  Distances 1..4
// have no extra bits, then each pair of codes does the same.  (Computed, not
// looked up: a table in device memory costs the serial decoder a load round
// trip per match.)
__device__ __forceinline__ uint32_t len_base(uint32_t s, int *extra) {
    if (s < 8) { *extra = 0; return s + 3; }
    if (s >= 28) { *extra = 0; return 258; }
    const int e = (int)(s >> 2) - 1;
    *extra = e;
    return ((4u + (s & 3u)) << e) + 3u;
}
__device__ __forceinline__ uint32_t dist_base(uint32_t d, int *extra) {
    if (d < 4) { *extra = 0; return d + 1; }
    const int e = (int)(d >> 1) - 1;
    *extra = e;
    return ((2u + (d & 1u)) << e) + 1u;
}
// RFC 1951 3.2.7: order of the code length code lengths
__constant__ uint8_t c_clorder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

enum : uint32_t { E_STREAM = 1, E_SIZE = 2 };

// LSB-first bit reader over the `len` bytes at `base`; reads past the end
// supply zero bytes and are counted, so that consuming them is detected.
// (32-bit offsets: their compares are single scalar instructions.)
struct Bits {
    const uint8_t *base;
    uint32_t off, len;
    uint64_t buf;
    int cnt;
    uint32_t pad;
    // n <= 32.  Four bytes at once from the two aligned words holding them
    // while eight or more bytes are left (one dependent load pair per four
    // bytes instead of per byte), then byte by byte.
    __device__ void fill(int n) {
        if (cnt >= n) return;
        if (len - off >= 8) {
            // (the aligned address by pointer arithmetic on base, not from
            // an integer, so that the compiler still knows its memory space;
            // in 64 bits: at off = 0 the aligned word starts before base)
            const uint32_t mis = ((uint32_t)reinterpret_cast<uintptr_t>(base) + off) & 3u;
            const uint8_t *q = base + off;
            const uint32_t *w = reinterpret_cast<const uint32_t *>(q - mis);
            const uint64_t two = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
            buf |= ((two >> (8 * mis)) & 0xFFFFFFFFull) << cnt;  // cnt < 32: no bit lost
            cnt += 32;
            off += 4;
            return;
        }
        while (cnt < n) {
            uint64_t v = 0;
            if (off < len) v = base[off++];
            else pad++;
            buf |= v << cnt;
            cnt += 8;
        }
    }
    // n bits that are already in the buffer (n <= cnt)
    __device__ uint32_t take(int n) {
        const uint32_t r = (uint32_t)(buf & ((1ull << n) - 1));
        buf >>= n;
        cnt -= n;
        return r;
    }
    __device__ uint32_t get(int n) {
        if (n == 0) return 0;
        fill(n);
        const uint32_t r = (uint32_t)(buf & ((1ull << n) - 1));
        buf >>= n;
        cnt -= n;
        return r;
    }
    // bits consumed beyond the real input
    __device__ bool overrun() const { return (int)(8 * pad) > cnt; }
};

// Canonical code of n symbols from their lengths: count[len] codes per
// length, symbol[] ordered by code.  Returns < 0 if over-subscribed, > 0 if
// incomplete, 0 if complete.
__device__ int construct(uint16_t *count, uint16_t *symbol, uint16_t *offs, const uint8_t *length, int n) {
    for (int len = 0; len <= 15; len++) count[len] = 0;
    for (int s = 0; s < n; s++) count[length[s]]++;
    if (count[0] == n) return 0;
    int left = 1;
    for (int len = 1; len <= 15; len++) {
        left <<= 1;
        left -= count[len];
        if (left < 0) return left;
    }
    offs[1] = 0;
    for (int len = 1; len < 15; len++) offs[len + 1] = offs[len] + count[len];
    for (int s = 0; s < n; s++)
        if (length[s]) symbol[offs[length[s]]++] = (uint16_t)s;
    return left;
}

// One symbol: the code's bits arrive most significant first.
__device__ int decode(Bits &b, const uint16_t *count, const uint16_t *symbol) {
    b.fill(15);
    uint64_t bits = b.buf;
    int code = 0, first = 0, index = 0;
    for (int len = 1; len <= 15; len++) {
        code |= (int)(bits & 1u);
        bits >>= 1;
        const int c = count[len];
        if (code - c < first) {
            b.buf >>= len;
            b.cnt -= len;
            return symbol[index + (code - first)];
        }
        index += c;
        first += c;
        first <<= 1;
        code <<= 1;
    }
    return -1;
}

// Per-member state in LDS: canonical codes, the direct tables of the
// first LB / DB code bits, and the token queue.
constexpr int LB = 10;    // literal/length table bits
constexpr int DB = 8;     // distance table bits
// Token queue entries and waves per SIMD (build-time settings): at 256
// tokens a member holds 7.2 KB of LDS, 22 members in the CU's 160 KB; the
// kernel fits six waves per SIMD's 80 VGPRs.
#ifndef KITE_INFLATE_NTOK
#define KITE_INFLATE_NTOK 256
#endif
#ifndef KITE_INFLATE_WAVES
#define KITE_INFLATE_WAVES 6
#endif
constexpr int NTOK = KITE_INFLATE_NTOK;
static_assert(NTOK % 64 == 0 && NTOK >= 64, "the queue is written 64 tokens at a time");
// Tokens: a literal is TOK_LIT | (count << 24) | its 1 or 2 bytes (first in
// bits 0-7); a match is (length << 16) | distance - 1 (bit 31 clear).
constexpr uint32_t TOK_LIT = 0x80000000u;
// Literal/length table entries, bits 26-29 the bits the entry consumes (0:
// the code is longer than LB bits): for one or two literals the token itself
// with those bits added; for a length code its base length in bits 0-8 and
// extra bits in 9-11; ENT_EOB for the end of block, ENT_BAD for symbols 286
// and 287.
constexpr int ENT_LEN = 26;
constexpr uint32_t ENT_TOK = ~(15u << ENT_LEN);
constexpr uint32_t ENT_EOB = 1u << 12, ENT_BAD = 1u << 13;
// Distance table entries: the base distance in bits 0-14, its extra bits in
// 15-18, the code's length in 19-22; DIST_BAD for codes 30 and 31.
constexpr uint32_t DIST_BAD = 1u << 23;

struct Member {
    uint32_t lut_l[1 << LB];
    uint32_t lut_d[1 << DB];  // distance entries (distance_lut), 0 = longer code
    uint16_t lcount[16], lsym[288];
    uint16_t dcount[16], dsym[32];
    uint16_t offs[16];
    uint8_t lengths[320];
    uint32_t tok[NTOK];
};

// Direct table of the codes up to `bits` long: entry [reversed code + k <<
// length] for every k, so that the next `bits` stream bits index it.  The
// wave runs it together: the code walk is the same on every lane, the
// entries of one code are split over the lanes.
template <class E>
__device__ void fill_lut(E *lut, int bits, const uint16_t *count, const uint16_t *symbol, uint32_t lane) {
    for (uint32_t i = lane; i < (1u << bits); i += 64) lut[i] = 0;
    int code = 0, idx = 0;
    for (int len = 1; len <= bits; len++) {
        for (int q = 0; q < count[len]; q++, idx++, code++) {
            const uint32_t r = __builtin_bitreverse32((uint32_t)code) >> (32 - len);
            const E e = (E)((len << 9) | symbol[idx]);
            for (uint32_t k = r + (lane << len); k < (1u << bits); k += 64u << len) lut[k] = e;
        }
        code <<= 1;
    }
}

// The literal/length table from fill_lut's (length << 9) | symbol entries
// to the decoder's: literal entries become tokens, and where an entry's code
// is a literal shorter than LB bits and the remaining bits of the index hold
// a whole second literal code, the entry is the token of both.
__device__ void literal_lut(uint32_t *lut, uint32_t lane) {
    constexpr int PER = (1 << LB) / 64;
    uint32_t v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const uint32_t i = lane + 64u * (uint32_t)q;
        const uint32_t e1 = lut[i];
        const uint32_t l1 = (e1 >> 9) & 15u, s1 = e1 & 511u;
        if (!l1) {
            v[q] = 0;
        } else if (s1 < 256) {
            v[q] = TOK_LIT | (1u << 24) | s1 | (l1 << ENT_LEN);
        } else if (s1 == 256) {
            v[q] = ENT_EOB | (l1 << ENT_LEN);
        } else if (s1 >= 286) {
            v[q] = ENT_BAD | (l1 << ENT_LEN);
        } else {
            int lx;
            const uint32_t lb = len_base(s1 - 257, &lx);
            v[q] = lb | ((uint32_t)lx << 9) | (l1 << ENT_LEN);
        }
        if (l1 && l1 < (uint32_t)LB && s1 < 256) {
            const uint32_t e2 = lut[i >> l1];
            const uint32_t l2 = (e2 >> 9) & 15u, s2 = e2 & 511u;
            if (l2 && l1 + l2 <= (uint32_t)LB && s2 < 256)
                v[q] = TOK_LIT | (2u << 24) | s1 | (s2 << 8) | ((l1 + l2) << ENT_LEN);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) lut[lane + 64u * (uint32_t)q] = v[q];
    __syncthreads();
}

// The distance table from fill_lut's (length << 9) | symbol entries to the
// decoder's: base distance and extra bits resolved per entry.
__device__ void distance_lut(uint32_t *lut, uint32_t lane) {
#pragma unroll
    for (int q = 0; q < (1 << DB) / 64; q++) {
        const uint32_t i = lane + 64u * (uint32_t)q;
        const uint32_t e1 = lut[i];
        const uint32_t l1 = (e1 >> 9) & 15u, d = e1 & 511u;
        if (!l1) continue;
        if (d >= 30) {
            lut[i] = DIST_BAD | (l1 << 19);
        } else {
            int dx;
            const uint32_t db = dist_base(d, &dx);
            lut[i] = db | ((uint32_t)dx << 15) | (l1 << 19);
        }
    }
}

// Block header: the tables of a fixed or dynamic block (run by every lane
// on the same bits: every lane stores the same values, and the direct
// tables are filled by the lanes together).
__device__ uint32_t read_tables(Bits &b, Member &t, uint32_t type, uint32_t lane) {
    if (type == 1) {
        // fixed codes (RFC 1951 3.2.6)
        for (int s = 0; s < 288; s++) t.lengths[s] = s < 144 ? 8 : (s < 256 ? 9 : (s < 280 ? 7 : 8));
        construct(t.lcount, t.lsym, t.offs, t.lengths, 288);
        for (int s = 0; s < 30; s++) t.lengths[s] = 5;
        construct(t.dcount, t.dsym, t.offs, t.lengths, 30);
    } else {
        const int nlen = (int)b.get(5) + 257, ndist = (int)b.get(5) + 1, ncode = (int)b.get(4) + 4;
        if (nlen > 286 || ndist > 30) return E_STREAM;
        for (int q = 0; q < 19; q++) t.lengths[c_clorder[q]] = q < ncode ? (uint8_t)b.get(3) : 0;
        // the code length code must be complete (held in the distance arrays)
        if (construct(t.dcount, t.dsym, t.offs, t.lengths, 19) != 0) return E_STREAM;
        int idx = 0;
        while (idx < nlen + ndist) {
            const int sym = decode(b, t.dcount, t.dsym);
            if (sym < 0) return E_STREAM;
            if (sym < 16) {
                t.lengths[idx++] = (uint8_t)sym;
                continue;
            }
            uint8_t len = 0;
            int rep;
            if (sym == 16) {
                if (idx == 0) return E_STREAM;
                len = t.lengths[idx - 1];
                rep = 3 + (int)b.get(2);
            } else if (sym == 17) {
                rep = 3 + (int)b.get(3);
            } else {
                rep = 11 + (int)b.get(7);
            }
            if (idx + rep > nlen + ndist) return E_STREAM;
            while (rep--) t.lengths[idx++] = len;
        }
        if (t.lengths[256] == 0) return E_STREAM;  // no end-of-block code
        int r = construct(t.lcount, t.lsym, t.offs, t.lengths, nlen);
        if (r < 0 || (r > 0 && nlen - t.lcount[0] != 1)) return E_STREAM;
        r = construct(t.dcount, t.dsym, t.offs, t.lengths + nlen, ndist);
        if (r < 0 || (r > 0 && ndist - t.dcount[0] != 1)) return E_STREAM;
    }
    fill_lut(t.lut_l, LB, t.lcount, t.lsym, lane);
    fill_lut(t.lut_d, DB, t.dcount, t.dsym, lane);
    __syncthreads();
    distance_lut(t.lut_d, lane);  // (each lane its own entries)
    literal_lut(t.lut_l, lane);   // (ends with a barrier)
    return 0;
}

// One workgroup (one wave) per member.  The wave decodes up to NTOK tokens
// (literals and length/distance pairs) into LDS; then it writes the
// literals in parallel and copies each match with all lanes, in order: a
// copy whose source overlaps its destination (distance below the length)
// goes in rounds that double the copied span, each round reading only
// bytes written before it.
__global__ __launch_bounds__(64, KITE_INFLATE_WAVES) void k_inflate(const uint8_t *__restrict__ comp, const InflateBlock *__restrict__ blocks,
                                                uint32_t nb, uint8_t *__restrict__ out, uint32_t *__restrict__ err) {
    __shared__ Member t;
    const uint32_t lane = threadIdx.x;
    const InflateBlock bk = blocks[blockIdx.x];
    uint8_t *o = out + bk.dst;
    const uint32_t osz = bk.isize;
    // The decoder: every lane runs it on the same bits, so its state and
    // branches are wave-uniform (scalar registers and scalar branches, no
    // execution-mask changes per symbol).  Tokens are queued 64 at a time
    // in one vector register (token i in lane i % 64) and stored to the LDS
    // queue by the whole wave.
    Bits b{comp + bk.src, 0, bk.clen, 0, 0, 0};
    uint32_t mode = 0;         // 0: between blocks, 1: stored, 2: Huffman codes
    bool last = false, done = false;
    uint32_t stored_left = 0;  // bytes left in a stored block
    uint32_t produced = 0;     // output bytes of the tokens decoded so far
    uint32_t pos = 0;          // output bytes written
    uint32_t e = 0;
    while (!done) {
        uint32_t n = 0, qv = 0;
        auto put = [&](uint32_t tok) {
            qv = lane == (n & 63u) ? tok : qv;
            if ((n & 63u) == 63u) t.tok[(n & ~63u) + lane] = qv;
            n++;
        };
        while (n < NTOK) {
            if (mode == 2) {
                // the symbols of a Huffman block, until it ends or the
                // queue is full: through the direct table (one or two
                // literals, or a length code), else the canonical walk for
                // a code longer than LB bits
                for (;;) {
                    b.fill(15);
                    const uint32_t ent = t.lut_l[b.buf & ((1u << LB) - 1)];
                    const int tl = (int)(ent >> ENT_LEN) & 15;
                    if (ent & TOK_LIT) {
                        // one or two literals
                        b.buf >>= tl;
                        b.cnt -= tl;
                        put(ent & ENT_TOK);
                        produced += (ent >> 24) & 3u;
                        if (n == NTOK) break;
                        continue;
                    }
                    uint32_t lb;
                    int lx;
                    if (tl) {
                        b.buf >>= tl;
                        b.cnt -= tl;
                        if (ent & (ENT_EOB | ENT_BAD)) {
                            if (ent & ENT_BAD) e = E_STREAM;
                            mode = 0;
                            if (b.overrun()) e = E_STREAM;
                            break;
                        }
                        lb = ent & 511u;
                        lx = (int)((ent >> 9) & 7u);
                    } else {
                        int sym = decode(b, t.lcount, t.lsym);
                        if (sym < 0) {
                            e = E_STREAM;
                            break;
                        }
                        if (sym < 256) {
                            put(TOK_LIT | (1u << 24) | (uint32_t)sym);
                            produced++;
                            if (n == NTOK) break;
                            continue;
                        }
                        if (sym == 256) {
                            mode = 0;
                            if (b.overrun()) e = E_STREAM;
                            break;
                        }
                        sym -= 257;
                        if (sym >= 29) {
                            e = E_STREAM;
                            break;
                        }
                        lb = len_base((uint32_t)sym, &lx);
                    }
                    // the length's extra bits (<= 5) and the distance code
                    // (<= 15) from one refill
                    b.fill(32);
                    const uint32_t len = lb + b.take(lx);
                    uint32_t dist;
                    const uint32_t de = t.lut_d[b.buf & ((1u << DB) - 1)];
                    if (de) {
                        // code <= DB bits: at least 19 bits are left for the
                        // extra bits (<= 13)
                        b.take((int)((de >> 19) & 15u));
                        if (de & DIST_BAD) {
                            e = E_STREAM;
                            break;
                        }
                        dist = (de & 0x7FFFu) + b.take((int)((de >> 15) & 15u));
                    } else {
                        const int ds = decode(b, t.dcount, t.dsym);
                        if (ds < 0 || ds >= 30) {
                            e = E_STREAM;
                            break;
                        }
                        int dx;
                        const uint32_t db = dist_base((uint32_t)ds, &dx);
                        dist = db + b.get(dx);
                    }
                    if (dist > produced) {
                        e = E_STREAM;
                        break;
                    }
                    put((len << 16) | (dist - 1));
                    produced += len;
                    if (n == NTOK) break;
                }
                if (e) break;
                continue;
            }
            if (mode == 0) {
                if (last) {
                    done = true;
                    break;
                }
                last = b.get(1) != 0;
                const uint32_t type = b.get(2);
                if (type == 0) {
                    b.get(b.cnt & 7);
                    const uint32_t len = b.get(16), nlen = b.get(16);
                    if (len != (~nlen & 0xFFFFu) || b.overrun()) {
                        e = E_STREAM;
                        break;
                    }
                    stored_left = len;
                    mode = len ? 1u : 0u;
                } else if (type == 1 || type == 2) {
                    e = read_tables(b, t, type, lane);
                    if (e) break;
                    mode = 2;
                } else {
                    e = E_STREAM;
                    break;
                }
                continue;
            }
            // mode 1: a byte of a stored block
            put(TOK_LIT | (1u << 24) | b.get(8));
            produced++;
            if (--stored_left == 0) {
                mode = 0;
                if (b.overrun()) {
                    e = E_STREAM;
                    break;
                }
            }
        }
        // no token is written past the member's output size (checked here
        // once for the batch's literals and matches)
        if (!e && produced > osz) e = E_SIZE;
        // the partial group (slots from n on are never read)
        if (n & 63u) t.tok[(n & ~63u) + lane] = qv;
        if (e) done = true;
        __syncthreads();
        if (e) break;
        // output offsets of the tokens: NTOK / 64 consecutive tokens per lane
        constexpr int TPL = NTOK / 64;
        uint32_t tks[TPL], lens[TPL], offs[TPL], sum = 0, mbits = 0;
#pragma unroll
        for (int q = 0; q < TPL; q++) {
            const uint32_t i = lane * TPL + q;
            tks[q] = i < n ? t.tok[i] : TOK_LIT;
            lens[q] = i < n ? ((tks[q] & TOK_LIT) ? (tks[q] >> 24) & 3u : (tks[q] >> 16)) : 0u;
            if (i < n && !(tks[q] & TOK_LIT)) mbits |= 1u << q;
            sum += lens[q];
        }
        const uint32_t incl = wave_incl_scan_u32(sum);
        const uint32_t total = __shfl(incl, 63, 64);
        uint32_t off = pos + incl - sum;
        // literals
#pragma unroll
        for (int q = 0; q < TPL; q++) {
            offs[q] = off;
            if (lens[q] && (tks[q] & TOK_LIT)) {
                o[off] = (uint8_t)tks[q];
                if (lens[q] == 2) o[off + 1] = (uint8_t)(tks[q] >> 8);
            }
            off += lens[q];
        }
        __threadfence_block();
        // matches, in order: the lanes holding one in ascending order, each
        // lane's in token order, read from its registers.  A copy reads
        // bytes of earlier copies only after their stores have completed:
        // `dirty` is the lowest output byte stored since the last wait
        // (every later store lies above it, below the current match), and a
        // match whose source ends at or below it (bytes of earlier batches,
        // this batch's literals, copies before the last wait) copies
        // without waiting.
        uint32_t dirty = 0xFFFFFFFFu;
        uint64_t lanes = __ballot(mbits != 0);
        while (lanes) {
            const int L = __builtin_ctzll(lanes);
            lanes &= lanes - 1;
            const uint32_t lm = (uint32_t)__builtin_amdgcn_readlane((int)mbits, L);
#pragma unroll
            for (int q = 0; q < TPL; q++) {
                if (!((lm >> q) & 1u)) continue;
                const uint32_t tk = (uint32_t)__builtin_amdgcn_readlane((int)tks[q], L);
                const uint32_t at = (uint32_t)__builtin_amdgcn_readlane((int)offs[q], L);
                const uint32_t len = tk >> 16, dist = (tk & 0xFFFFu) + 1;
                uint32_t c = dist < len ? dist : len;
                if (at - dist + c > dirty) {
                    __threadfence_block();
                    dirty = 0xFFFFFFFFu;
                }
                for (uint32_t j = lane; j < c; j += 64) o[at + j] = o[at - dist + j];
                if (dirty > at) dirty = at;
                while (c < len) {
                    // the bytes so far repeat with period c (a multiple of dist)
                    __threadfence_block();
                    const uint32_t m = (len - c) < c ? (len - c) : c;
                    for (uint32_t j = lane; j < m; j += 64) o[at + c + j] = o[at + j];
                    dirty = at + c;
                    c += m;
                }
            }
        }
        pos += total;
        __syncthreads();
    }
    if (lane == 0) {
        if (!e && pos != osz) e = E_SIZE;
        if (e) atomicOr(err, e);
    }
}

// Number of '\n' bytes in d[0, n) added to *total.
__global__ __launch_bounds__(256) void k_count_newlines(const uint8_t *__restrict__ d, uint64_t n,
                                                        unsigned long long *__restrict__ total) {
    __shared__ uint32_t ws[256 / 64 + 1];
    uint32_t c = 0;
    const uint64_t nv = n / 16;
    for (uint64_t v = (uint64_t)blockIdx.x * 256 + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * 256) {
        const uint4 w = reinterpret_cast<const uint4 *>(d)[v];
        const uint32_t x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t y = x[q] ^ 0x0A0A0A0Au;
            c += __popc(~(((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y | 0x7F7F7F7Fu));
        }
    }
    if (blockIdx.x == 0)
        for (uint64_t i = nv * 16 + threadIdx.x; i < n; i += 256) c += d[i] == '\n';
    uint32_t sum;
    (void)block_excl_scan<256, uint32_t>(c, ws, &sum);
    if (threadIdx.x == 0 && sum) atomicAdd(total, (unsigned long long)sum);
}

}  // namespace

void count_newlines_device(const uint8_t *d, uint64_t n, unsigned long long *d_total, hipStream_t s) {
    if (!n) return;
    const uint64_t wgs = std::min<uint64_t>(std::max<uint64_t>(1, n / (256 * 16 * 8)), 4096);
    k_count_newlines<<<dim3((unsigned)wgs), dim3(256), 0, s>>>(d, n, d_total);
    KITE_CHECK_LAUNCH();
}

std::vector<BgzfMember> scan_bgzf(const uint8_t *d, size_t len, size_t *consumed) {
    std::vector<BgzfMember> out;
    size_t pos = 0;
    // is_bgzf_header / bgzf_block_total_size (src/count/bgzf.rs:37-101)
    while (pos + 18 <= len) {
        const uint8_t *h = d + pos;
        if (!(h[0] == 0x1f && h[1] == 0x8b && (h[3] & 0x04) && h[12] == 0x42 && h[13] == 0x43 && h[14] == 0x02 &&
              h[15] == 0x00))
            throw Error("Non-BGZF data at offset " + std::to_string(pos) + " (expected gzip+BGZF header)");
        const uint32_t bsize = (uint32_t)(h[16] | (h[17] << 8)) + 1;
        if (pos + bsize > len) break;  // truncated last member: left for the caller
        const uint32_t xlen = (uint32_t)(h[10] | (h[11] << 8));
        if (h[2] != 8 || 12 + xlen + 8 > bsize)
            throw Error("invalid BGZF member at offset " + std::to_string(pos));
        // only FEXTRA may be set besides the always-present fields
        if (h[3] & 0x1a) throw Error("unsupported gzip header flags in BGZF member at offset " + std::to_string(pos));
        BgzfMember m;
        m.offset = pos;
        m.size = bsize;
        m.data = 12 + xlen;
        m.clen = bsize - 8 - m.data;
        const uint8_t *tr = h + bsize - 4;
        m.isize = (uint32_t)(tr[0] | (tr[1] << 8) | (tr[2] << 16) | ((uint32_t)tr[3] << 24));
        if (m.isize > 65536) throw Error("BGZF member above 64 KiB of output at offset " + std::to_string(pos));
        out.push_back(m);
        pos += bsize;
    }
    if (consumed) *consumed = pos;
    return out;
}

void inflate_blocks(const uint8_t *d_comp, const std::vector<InflateBlock> &blocks, uint8_t *d_out, uint32_t *d_err,
                    DevBuf &d_blocks, hipStream_t s) {
    if (blocks.empty()) return;
    InflateBlock *db = static_cast<InflateBlock *>(d_blocks.ensure(blocks.size() * sizeof(InflateBlock)));
    KITE_HIP(hipMemcpyAsync(db, blocks.data(), blocks.size() * sizeof(InflateBlock), hipMemcpyHostToDevice, s));
    const uint32_t nb = (uint32_t)blocks.size();
    k_inflate<<<dim3(nb), dim3(64), 0, s>>>(d_comp, db, nb, d_out, d_err);
    KITE_CHECK_LAUNCH();
}

}  // namespace kite
