// count.hip -- gfx950 kernels and host orchestration of the exact k-mer count.
//
// Semantics follow jiehua1995/Kite v0.7.5:
//   record split: every 4th line starting at line 1 is a sequence line
//     (src/count/parser.rs:19-65); a buffer not starting with '@' yields
//     nothing (parse error ignored, src/count/worker.rs:217);
//   encoding A/a=0 C/c=1 G/g=2 T/t=3, any other byte resets the window
//     (src/count/kmer.rs:2-13, 241-267);
//   key = 2k-bit integer, first base in the high bits; canonical =
//     min(forward, reverse complement) as unsigned integers
//     (src/count/kmer.rs:203-432);
//   partition = (fold(key) * 0x9E3779B97F4A7C15) >> (64 - log2 P), fold =
//     XOR of the 64-bit words (src/count/kmer.rs:67-71, 111-116, 156-161);
//   per-partition ascending unique keys with run lengths saturated at
//     u32::MAX (src/count/sorter.rs:514-530).
//
// Device layout: keys are arrays of structures of W little-endian u64 words
// (W = 1 for k <= 31, 2 for k <= 64, 4 above), w[0] least significant.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <tuple>
#include <map>

#include "count.hpp"
#include "encode.hpp"

namespace kite {

// ---- extract geometry -------------------------------------------------------
constexpr int XT = 256;                 // extract threads
constexpr int SUB = 8064;               // FASTQ bytes per sub-tile (with the halo, 8 KiB)
constexpr int SEGB = 32;                // bytes per segment (one thread)
constexpr int HALO = 128;               // >= k_max - 1 bytes before each sub-tile
constexpr int HSEG = HALO / SEGB;       // 4 halo segments
constexpr int NSEG = SUB / SEGB + HSEG; // 256 segments per sub-tile: one per thread
constexpr int EXT_U4 = (SUB + HALO) / 16;
static_assert(NSEG == XT, "one segment per extract thread");
constexpr uint32_t MAX_PARTITIONS = 8192;

// ---- split / sort geometry --------------------------------------------------
constexpr int ST = 256;                 // split threads
template <int W> struct SplitCfg { static constexpr int IPT = 32 / W; static constexpr int TK = ST * IPT; };
constexpr int BT = 256;                 // bucket-sort threads
constexpr int BNW = BT / 64;
constexpr int SORT_CLASSES = 4;         // items per thread (4 << class) / W
template <int W, int C> struct SortClass { static constexpr int IPT = (4 << C) / W; static constexpr int CAP = BT * IPT; };
constexpr int MAX_DBITS = 12;

template <int W>
struct Key {
    uint64_t w[W];
};

__device__ __forceinline__ uint64_t rev2(uint64_t x) {
    x = __builtin_bitreverse64(x);
    return ((x >> 1) & 0x5555555555555555ull) | ((x & 0x5555555555555555ull) << 1);
}

// 0x80 in every zero byte of v, 0 elsewhere (exact, no borrow across bytes)
__device__ __forceinline__ uint32_t zero_bytes(uint32_t v) {
    return ~(((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v | 0x7F7F7F7Fu);
}

// The 0x80 flags of zero_bytes as four bits, byte 0 in bit 0.
__device__ __forceinline__ uint32_t byte_mask4(uint32_t z) {
    return (((z >> 7) * 0x00204081u) >> 21) & 0xFu;
}

__device__ __forceinline__ uint32_t newlines_in(uint32_t v) {
    uint32_t x = v ^ 0x0A0A0A0Au;  // zero bytes where '\n'
    uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
    return __popc(z);
}

template <int W>
__device__ __forceinline__ bool key_less(const Key<W> &a, const Key<W> &b) {
#pragma unroll
    for (int r = W - 1; r >= 0; r--) {
        if (a.w[r] != b.w[r]) return a.w[r] < b.w[r];
    }
    return false;
}

template <int W>
__device__ __forceinline__ bool key_eq(const Key<W> &a, const Key<W> &b) {
    bool e = true;
#pragma unroll
    for (int r = 0; r < W; r++) e = e && (a.w[r] == b.w[r]);
    return e;
}

template <int W>
__device__ __forceinline__ uint32_t key_partition(const Key<W> &a, int pbits) {
    uint64_t f = a.w[0];
#pragma unroll
    for (int r = 1; r < W; r++) f ^= a.w[r];
    // top pbits bits of the product (0 for pbits = 0), without a branch
    return (uint32_t)(((f * 0x9E3779B97F4A7C15ull) >> 32) >> (32 - pbits));
}

// Bits [shift, shift + nb) of the key, nb <= 32.
template <int W>
__device__ __forceinline__ uint32_t key_bits(const Key<W> &a, int shift, int nb) {
    int wi = shift >> 6, bi = shift & 63;
    uint64_t lo = 0, hi = 0;
#pragma unroll
    for (int r = 0; r < W; r++) {
        if (r == wi) lo = a.w[r];
        if (r == wi + 1) hi = a.w[r];
    }
    uint64_t v = lo >> bi;
    if (bi && bi + nb > 64) v |= hi << (64 - bi);
    return (uint32_t)(v & ((1ull << nb) - 1));
}

// Loads NI keys of a tile, item `it` from position it * STRIDE + tid of
// src[0, n) (n >= 1).  Positions past n read the last key, so that every
// load is unconditional and the NI loads are in flight together; a guarded
// load whose value is used right away would make each wave wait for one
// load at a time.
template <int W, int NI, int STRIDE>
__device__ __forceinline__ void load_tile(const uint64_t *__restrict__ src, uint64_t n, int tid, Key<W> (&key)[NI]) {
#pragma unroll
    for (int it = 0; it < NI; it++) {
        const uint64_t i = (uint64_t)it * STRIDE + (uint64_t)tid;
        const uint64_t *p = src + (i < n ? i : n - 1) * W;
#pragma unroll
        for (int r = 0; r < W; r++) key[it].w[r] = p[r];
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  This bijection on [0, n) gives XCD x a contiguous range of logical
// tiles, so tiles that write neighbouring destination lines share an L2 and
// their partial-line writes merge there instead of being written back twice.
constexpr uint32_t N_XCD = 8;
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t n) {
    const uint64_t x = b % N_XCD, i = b / N_XCD;
    const uint64_t q = n / N_XCD, r = n % N_XCD;
    return x * q + (x < r ? x : r) + i;
}

// ---- newline count per sub-tile -------------------------------------------
__device__ __forceinline__ void nl_count_tile(const uint8_t *__restrict__ fq, uint64_t len, uint64_t tile,
                                              uint32_t *__restrict__ out, uint32_t *ws) {
    uint64_t base = tile * SUB + (uint64_t)threadIdx.x * SEGB;
    uint32_t c = 0;
    if ((int)threadIdx.x * SEGB >= SUB) {
        // past the sub-tile (its last threads)
    } else if (base + SEGB <= len) {
        const uint4 *p = reinterpret_cast<const uint4 *>(fq + base);
        uint4 a = p[0], b = p[1];
        c = newlines_in(a.x) + newlines_in(a.y) + newlines_in(a.z) + newlines_in(a.w) +
            newlines_in(b.x) + newlines_in(b.y) + newlines_in(b.z) + newlines_in(b.w);
    } else {
        for (uint64_t i = base; i < len && i < base + SEGB; i++) c += fq[i] == '\n';
    }
    uint32_t total;
    (void)block_excl_scan<XT, uint32_t>(c, ws, &total);
    if (threadIdx.x == 0) out[tile] = total;
}

__global__ __launch_bounds__(XT) void k_nl_count(const uint8_t *__restrict__ fq, uint64_t len,
                                                  uint32_t *__restrict__ out) {
    __shared__ uint32_t ws[XT / 64 + 1];
    nl_count_tile(fq, len, blockIdx.x, out, ws);
}

struct ExtractArgs {
    const uint8_t *fq;
    uint64_t len;
    const uint64_t *sub_line0;  // lines before each sub-tile
    uint64_t n_sub;
    uint32_t nwg;               // extract workgroups = matrix columns
    uint32_t subs;              // sub-tiles per workgroup
    int k, canonical, pbits;
    uint32_t *hist;             // [P][nwg] counts (histogram pass)
    const uint64_t *off;        // scanned [P][nwg] (scatter pass)
    uint64_t *out;              // partition-ordered keys (scatter pass)
    const DevPlan *pl;          // device-planned count: len and n_sub from here (else the fields above)
};

// Both passes walk identical k-mers; the scatter pass places each key at
// off[p][wg] + (slot within this workgroup's share of partition p).
// MODE 0: histogram.  MODE 1: scatter, each lane storing its own keys (any
// partition count).  MODE 2: scatter in chunks of SC keys grouped by
// partition in LDS first, so that lanes store runs of neighbouring keys:
// one 8-byte store per lane costs the cache one request per key, a run of
// r keys about r/8 (used when a chunk gives runs of two or more keys).
template <int W>
struct StageCfg {
    static constexpr int SC = 1024 / W;    // keys per chunk: SC * 8W bytes alias the sub-tile buffer
    static constexpr int EPT = SC / XT;    // keys per thread per chunk
};

template <int W, int MODE>
__global__ __launch_bounds__(XT) void k_extract(ExtractArgs a) {
    constexpr bool SCATTER = MODE != 0;
    // sub-tile bytes, then (histogram and staged passes) the k-mer end
    // positions; the staged pass reuses both as one chunk's staging once
    // the positions are read
    constexpr int SPOS_U4 = MODE != 1 ? XT * SEGB * 2 / 16 : 0;
    __shared__ uint4 sraw[EXT_U4 + SPOS_U4];
    uint4 *sb4 = sraw;
    __shared__ uint64_t scode[NSEG];
    __shared__ uint64_t srcode[NSEG];  // canonical: each segment reverse-complemented
    __shared__ uint32_t svalid[NSEG];
    __shared__ uint32_t snl[NSEG];
    __shared__ uint32_t spre[NSEG];
    __shared__ uint32_t ws[XT / 64 + 1];
    // k-mer end positions of the sub-tile (histogram and staged passes:
    // every lane builds k-mers; the direct scatter keeps the per-segment loop)
    uint16_t *spos = reinterpret_cast<uint16_t *>(sraw + EXT_U4);
    __shared__ uint16_t spart[MODE == 2 ? StageCfg<W>::SC : 1];  // partition of each staged key
    // P u32 counters, then (scatter) P u64 bases, then (staged) per-chunk
    // counts and offsets, P u32 each
    extern __shared__ uint32_t dyn[];
    const uint32_t P = 1u << a.pbits;
    uint32_t *cnt = dyn;
    uint64_t *pbase = reinterpret_cast<uint64_t *>(dyn + ((P + 1) & ~1u));
    uint32_t *lcnt = dyn + ((P + 1) & ~1u) + 2 * P;
    uint32_t *loff = lcnt + P;
    const int tid = threadIdx.x;
    const uint32_t *sb = reinterpret_cast<const uint32_t *>(sb4);
    // the input size, from the device plan when the launch was fixed before
    // the data was seen (the grid then covers a capacity)
    const uint64_t alen = a.pl ? a.pl->len : a.len;
    const uint64_t an_sub = a.pl ? a.pl->n_sub : a.n_sub;
    if constexpr (SCATTER) {
        if (a.pl && a.pl->overflow) return;  // the keys would not fit: nothing is written
    }
    // the workgroups that have input are spread over the XCDs as one grid of
    // their own number; the rest hold empty matrix columns
    const uint64_t busy = a.pl ? ceil_div(an_sub, a.subs) : gridDim.x;
    if (blockIdx.x >= busy) {
        if constexpr (!SCATTER)
            for (uint32_t p = tid; p < P; p += XT) a.hist[(uint64_t)p * a.nwg + blockIdx.x] = 0;
        return;
    }
    const uint32_t wg = (uint32_t)xcd_tile(blockIdx.x, busy);
    const bool live = alen > 0 && a.fq[0] == '@';

    for (uint32_t p = tid; p < P; p += XT) {
        cnt[p] = 0;
        if constexpr (SCATTER) pbase[p] = a.off[(uint64_t)p * a.nwg + wg];
        if constexpr (MODE == 2) lcnt[p] = 0;
    }

    // key masks and reverse-complement shift (constant per launch)
    uint64_t kmask[W];
#pragma unroll
    for (int r = 0; r < W; r++) {
        int bits = 2 * a.k - 64 * r;
        kmask[r] = bits >= 64 ? ~0ull : (bits <= 0 ? 0ull : ((1ull << bits) - 1));
    }
    const int rs = 64 * W - 2 * a.k;
    const int rq = W == 1 ? 0 : rs >> 6, rb = rs & 63;  // rs < 64 for one word

    for (uint32_t si = 0; si < a.subs; si++) {
        uint64_t st = (uint64_t)wg * a.subs + si;
        if (st >= an_sub) break;
        const int64_t base = (int64_t)(st * SUB);
        __syncthreads();
        if (base >= HALO && base + SUB <= (int64_t)alen) {
            // the whole sub-tile and its halo are inside: unconditional
            // loads, all in flight before the first LDS store
            constexpr int LPT = EXT_U4 / XT;
            static_assert(EXT_U4 % XT == 0, "whole 16-byte words per thread");
            uint4 v[LPT];
            const uint4 *src = reinterpret_cast<const uint4 *>(a.fq + base - HALO);
#pragma unroll
            for (int u = 0; u < LPT; u++) v[u] = src[u * XT + tid];
#pragma unroll
            for (int u = 0; u < LPT; u++) sb4[u * XT + tid] = v[u];
        } else for (int i = tid; i < EXT_U4; i += XT) {
            int64_t gb = base - HALO + 16 * (int64_t)i;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gb >= 0 && gb + 16 <= (int64_t)alen) {
                v = *reinterpret_cast<const uint4 *>(a.fq + gb);
            } else if (gb + 16 > 0 && gb < (int64_t)alen) {
                uint8_t tmp[16];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    int64_t g = gb + q;
                    tmp[q] = (g >= 0 && g < (int64_t)alen) ? a.fq[g] : 0;
                }
                memcpy(&v, tmp, 16);
            }
            sb4[i] = v;
        }
        __syncthreads();
        for (int s = tid; s < NSEG; s += XT) {
            uint32_t m = 0;
#pragma unroll
            for (int q = 0; q < SEGB / 4; q++) m |= byte_mask4(zero_bytes(sb[s * (SEGB / 4) + q] ^ 0x0A0A0A0Au)) << (4 * q);
            snl[s] = m;  // bit j: byte j of the segment is '\n'
        }
        __syncthreads();
        {
            uint32_t total;
            uint32_t ex = block_excl_scan<XT, uint32_t>(__popc(snl[tid]), ws, &total);
            spre[tid] = ex;
            if (tid < NSEG - XT) {
                uint32_t t = total;
                for (int q = 0; q < tid; q++) t += __popc(snl[XT + q]);
                spre[XT + tid] = t;
            }
        }
        __syncthreads();
        // line index (mod 4) at the first halo byte
        const uint32_t l0 = (uint32_t)(a.sub_line0[st] & 3u) - spre[HSEG];
        for (int s = tid; s < NSEG; s += XT) {
            // bytes on sequence lines (line index 1 of the 4-line record)
            uint32_t line = (l0 + spre[s]) & 3u, l1 = 0, start = 0;
            for (uint32_t e = snl[s]; e; e &= e - 1) {
                const uint32_t p = (uint32_t)__ffs(e) - 1;
                if (line == 1u) l1 |= ((1u << p) - 1u) & ~((1u << start) - 1u);
                line = (line + 1u) & 3u;
                start = p + 1;
            }
            if (line == 1u && start < 32u) l1 |= ~0u << start;
            // four bytes at a time: 2-bit codes (A0 C1 G2 T3, case-blind),
            // first byte most significant, and the ACGT/acgt test
            uint64_t code = 0;
            uint32_t acgt = 0;
#pragma unroll
            for (int q = 0; q < SEGB / 4; q++) {
                const uint32_t w = sb[s * (SEGB / 4) + q];
                const uint32_t c = ((w >> 1) ^ (w >> 2)) & 0x03030303u;
                // byte b's code to bits 30 - 10b (no two partial products
                // overlap), then the four codes as one byte, byte 0 on top
                const uint32_t p8 = (c * 0x40100401u) >> 24;
                code |= (uint64_t)p8 << (56 - 8 * q);
                // the lowercase letter each code stands for; a byte is a base
                // exactly when it equals that letter after case folding
                const uint32_t expect = __builtin_amdgcn_perm(0u, 0x74676361u, c);
                acgt |= byte_mask4(zero_bytes((w | 0x20202020u) ^ expect)) << (4 * q);
            }
            scode[s] = code;
            srcode[s] = rev2(~code);
            svalid[s] = live ? (acgt & l1) : 0u;
        }
        __syncthreads();

        // k-mer end positions of this thread's segment (the last HSEG
        // threads own none: their segments are the next sub-tile's halo)
        const int s = tid + HSEG;
        uint32_t emask = 0;
        if (s < NSEG) {
            uint32_t run = 0;  // valid bases just before the segment
#pragma unroll
            for (int q = 1; q <= HSEG; q++) {
                uint32_t v = svalid[s - q];
                if (v == 0xFFFFFFFFu) {
                    run += 32;
                    continue;
                }
                run += __clz(~v);
                break;
            }
            // byte j ends a k-mer when bytes j-L..j (L = k-1) are valid: no
            // invalid byte of the segment in that window (each invalid bit
            // smeared L places up), and j >= L - run where the window
            // reaches before the segment
            const uint32_t L = (uint32_t)a.k - 1u;
            uint32_t sm = ~svalid[s], len = 1;  // sm: invalid bits smeared by [0, len)
#pragma unroll
            for (int t = 0; t < 5; t++)
                if (2 * len <= L + 1) {
                    sm |= sm << len;
                    len *= 2;
                }
            if (L + 1 - len < 32u) sm |= sm << (L + 1 - len);
            const uint32_t lk = run < L ? L - run : 0u;  // low bytes whose window lacks bases
            const uint32_t low = lk >= 32u ? ~0u : ((1u << lk) - 1u);
            emask = ~(sm | low);
        }
        // the canonical key of the k-mer ending at position j of segment ss
        auto kmer_at = [&](int ss, int j) -> Key<W> {
            uint64_t wv[W + 1];
#pragma unroll
            for (int q = 0; q <= W; q++) wv[q] = scode[ss - W + q];
            const int sh = 62 - 2 * j;
            Key<W> f;
#pragma unroll
            for (int r = 0; r < W; r++) {
                // (x << 1) << (63 - sh) is x << (64 - sh), and 0 for sh = 0
                uint64_t lo = wv[W - r] >> sh;
                uint64_t hi = (wv[W - r - 1] << 1) << (63 - sh);
                f.w[r] = (lo | hi) & kmask[r];
            }
            // (computed in both modes and selected, which keeps the code
            // free of branches so that neighbouring k-mers interleave)
            {
                // the complemented bases from the k-mer's last one backwards
                // are the reverse-complemented segments read downwards from
                // bit 2 * (31 - j) of segment ss: top-aligned, most
                // significant word first, then shifted down to 2k bits
                uint64_t top[W];
#pragma unroll
                for (int r = 0; r < W; r++) {
                    const uint64_t x = srcode[ss - r], y = srcode[ss - r - 1];
                    top[r] = (x << sh) | ((y >> 1) >> (63 - sh));
                }
                Key<W> rc;
#pragma unroll
                for (int r = 0; r < W; r++) {
                    uint64_t lo = 0, hi = 0;
#pragma unroll
                    for (int t = 0; t < W; t++) {
                        if (t == r + rq) lo = top[W - 1 - t];
                        if (t == r + rq + 1) hi = top[W - 1 - t];
                    }
                    rc.w[r] = (lo >> rb) | ((hi << 1) << (63 - rb));
                }
                return (a.canonical && key_less<W>(rc, f)) ? rc : f;
            }
        };
        if constexpr (!SCATTER) {
            // one list for the whole sub-tile, so that every lane builds k-mers
            // (segments of header and quality lines have none)
            uint32_t nk;
            uint32_t slot = block_excl_scan<XT, uint32_t>(__popc(emask), ws, &nk);
            for (uint32_t e = emask; e; e &= e - 1) spos[slot++] = (uint16_t)(s * SEGB + __ffs(e) - 1);
            __syncthreads();
            // HU k-mers per thread and step, so that their LDS reads overlap
            constexpr int HU = 4;
            for (uint32_t i0 = tid; i0 < nk; i0 += HU * XT) {
                uint32_t q[HU];
#pragma unroll
                for (int u = 0; u < HU; u++) {
                    const uint32_t i = i0 + (uint32_t)u * XT;
                    const uint32_t pos = spos[i < nk ? i : i0];
                    q[u] = key_partition<W>(kmer_at((int)(pos / SEGB), (int)(pos % SEGB)), a.pbits);
                }
#pragma unroll
                for (int u = 0; u < HU; u++)
                    if (i0 + (uint32_t)u * XT < nk) atomicAdd(&cnt[q[u]], 1u);
            }
        } else if constexpr (MODE == 1) {
            for (uint32_t e = emask; e; e &= e - 1) {
                const Key<W> key = kmer_at(s, __ffs(e) - 1);
                const uint32_t p = key_partition<W>(key, a.pbits);
                uint64_t *dst = a.out + (pbase[p] + atomicAdd(&cnt[p], 1u)) * W;
#pragma unroll
                for (int r = 0; r < W; r++) dst[r] = key.w[r];
            }
        } else {
            constexpr int SC = StageCfg<W>::SC, EPT = StageCfg<W>::EPT;
            const uint32_t PPT = (P + XT - 1) / XT;  // partitions per thread in the offset scan
            uint32_t nk;
            uint32_t slot = block_excl_scan<XT, uint32_t>(__popc(emask), ws, &nk);
            for (uint32_t e = emask; e; e &= e - 1) spos[slot++] = (uint16_t)(s * SEGB + __ffs(e) - 1);
            __syncthreads();
            // one chunk for the whole sub-tile when its keys fit the bytes and
            // positions together (longer runs per partition), staged in
            // place once every position is read; else chunks of SC keys
            // staged in the sub-tile bytes alone
            constexpr int SCB = (EXT_U4 + SPOS_U4) * 16 / (8 * W), EPTB = SCB / XT;
            static_assert(SCB % XT == 0, "whole keys per thread");
            if (nk <= (uint32_t)SCB) {
                Key<W> key[EPTB];
                uint32_t part[EPTB], rank[EPTB];
#pragma unroll
                for (int u = 0; u < EPTB; u++) {
                    const uint32_t i = (uint32_t)u * XT + tid;
                    if (i < nk) {
                        const uint32_t pos = spos[i];
                        key[u] = kmer_at((int)(pos / SEGB), (int)(pos % SEGB));
                        part[u] = key_partition<W>(key[u], a.pbits);
                        rank[u] = atomicAdd(&lcnt[part[u]], 1u);
                    }
                }
                __syncthreads();
                {
                    const uint32_t p0 = tid * PPT;
                    uint32_t sum = 0;
                    for (uint32_t q = 0; q < PPT; q++)
                        if (p0 + q < P) sum += lcnt[p0 + q];
                    uint32_t tot;
                    uint32_t run = block_excl_scan<XT, uint32_t>(sum, ws, &tot);
                    for (uint32_t q = 0; q < PPT; q++)
                        if (p0 + q < P) {
                            loff[p0 + q] = run;
                            run += lcnt[p0 + q];
                        }
                }
                __syncthreads();
                uint64_t *big = reinterpret_cast<uint64_t *>(sraw);
#pragma unroll
                for (int u = 0; u < EPTB; u++) {
                    const uint32_t i = (uint32_t)u * XT + tid;
                    if (i < nk) {
                        const uint32_t d = loff[part[u]] + rank[u];
#pragma unroll
                        for (int r = 0; r < W; r++) big[d * W + r] = key[u].w[r];
                    }
                }
                __syncthreads();
                for (uint32_t i = tid; i < nk; i += XT) {
                    Key<W> kk;
#pragma unroll
                    for (int r = 0; r < W; r++) kk.w[r] = big[i * W + r];
                    const uint32_t p = key_partition<W>(kk, a.pbits);  // recomputed, not stored
                    uint64_t *dst = a.out + (pbase[p] + cnt[p] + (i - loff[p])) * W;
#pragma unroll
                    for (int r = 0; r < W; r++) dst[r] = kk.w[r];
                }
                __syncthreads();
                for (uint32_t q = 0; q < PPT; q++) {
                    const uint32_t p = tid * PPT + q;
                    if (p < P) {
                        cnt[p] += lcnt[p];
                        lcnt[p] = 0;
                    }
                }
                // (the next sub-tile's fill starts with a barrier)
                continue;
            }
            // the sub-tile bytes are no longer needed: they hold the chunk
            uint64_t *stage = reinterpret_cast<uint64_t *>(sb4);
            for (uint32_t c0 = 0; c0 < nk; c0 += SC) {
                const uint32_t m = min((uint32_t)SC, nk - c0);
                Key<W> key[EPT];
                uint32_t part[EPT], rank[EPT];
#pragma unroll
                for (int u = 0; u < EPT; u++) {
                    const uint32_t i = (uint32_t)u * XT + tid;
                    if (i < m) {
                        const uint32_t pos = spos[c0 + i];
                        key[u] = kmer_at((int)(pos / SEGB), (int)(pos % SEGB));
                        part[u] = key_partition<W>(key[u], a.pbits);
                        rank[u] = atomicAdd(&lcnt[part[u]], 1u);
                    }
                }
                __syncthreads();
                {
                    const uint32_t p0 = tid * PPT;
                    uint32_t sum = 0;
                    for (uint32_t q = 0; q < PPT; q++)
                        if (p0 + q < P) sum += lcnt[p0 + q];
                    uint32_t tot;
                    uint32_t run = block_excl_scan<XT, uint32_t>(sum, ws, &tot);
                    for (uint32_t q = 0; q < PPT; q++)
                        if (p0 + q < P) {
                            loff[p0 + q] = run;
                            run += lcnt[p0 + q];
                        }
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < EPT; u++) {
                    const uint32_t i = (uint32_t)u * XT + tid;
                    if (i < m) {
                        const uint32_t d = loff[part[u]] + rank[u];
#pragma unroll
                        for (int r = 0; r < W; r++) stage[d * W + r] = key[u].w[r];
                        spart[d] = (uint16_t)part[u];
                    }
                }
                __syncthreads();
                for (uint32_t i = tid; i < m; i += XT) {
                    const uint32_t p = spart[i];
                    uint64_t *dst = a.out + (pbase[p] + cnt[p] + (i - loff[p])) * W;
#pragma unroll
                    for (int r = 0; r < W; r++) dst[r] = stage[i * W + r];
                }
                __syncthreads();
                for (uint32_t q = 0; q < PPT; q++) {
                    const uint32_t p = tid * PPT + q;
                    if (p < P) {
                        cnt[p] += lcnt[p];
                        lcnt[p] = 0;
                    }
                }
                __syncthreads();
            }
        }
    }
    if constexpr (!SCATTER) {
        __syncthreads();
        for (uint32_t p = tid; p < P; p += XT) a.hist[(uint64_t)p * a.nwg + wg] = cnt[p];
    }
}

__global__ void k_part_gather(const uint64_t *__restrict__ off, uint32_t nwg, uint32_t P,
                              uint64_t *__restrict__ part_off) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= P) part_off[p] = off[(uint64_t)p * nwg];
}

// ---- MSD split of large segments on the next key bits ----------------------
struct Seg {
    uint64_t start, size, child_base;
};

struct SplitArgs {
    const uint64_t *src;
    uint64_t *dst;
    const Seg *segs;
    uint32_t nseg;
    const uint64_t *tile_pre;  // nseg + 1 prefix of workgroup chunks
    int shift, dbits;
    uint32_t *child_hist;      // histogram pass: global child totals
    uint32_t *wg_hist;         // [workgroup][2^dbits] child counts of each chunk
    uint64_t *child_cur;       // scatter pass: running destination per child
    uint32_t chunk;            // tiles of TK keys per workgroup
};

// One workgroup handles a chunk of `chunk` tiles (TK keys each) of one segment,
// so that each child receives a run of about TK*chunk/2^dbits keys from it.
// MODE 0 (histogram) adds every chunk's child counts to the global totals
// (and keeps them per chunk when chunk > 1).  MODE 1 (scatter, one tile)
// holds the tile's keys in registers, recounts them in LDS and reserves one
// range per child with a single atomic; the ranges are kept as 32-bit
// offsets from the segment start (segments below 2^32 keys).  MODE 2
// (scatter, chunked) reserves from the stored per-chunk counts.
template <int W, int MODE>
__device__ __forceinline__ void split_tile(const SplitArgs &a, uint64_t b) {
    constexpr int IPT = SplitCfg<W>::IPT, TK = SplitCfg<W>::TK;
    const uint64_t TKL = (uint64_t)TK * a.chunk;
    __shared__ uint32_t hist[1 << MAX_DBITS];
    __shared__ uint32_t cbase32[MODE == 1 ? (1 << MAX_DBITS) : 1];
    __shared__ uint64_t cbase[MODE == 2 ? (1 << MAX_DBITS) : 1];
    const int tid = threadIdx.x;
    uint32_t lo = 0, hi = a.nseg;  // largest s with tile_pre[s] <= b
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (a.tile_pre[mid] <= b) lo = mid;
        else hi = mid;
    }
    const Seg sg = a.segs[lo];
    const uint64_t t = b - a.tile_pre[lo];
    const uint64_t first = sg.start + t * TKL;
    const uint64_t n = std::min<uint64_t>(TKL, sg.size - t * TKL);
    const uint32_t nd = 1u << a.dbits;
    uint32_t *row = a.chunk > 1 ? a.wg_hist + b * nd : nullptr;  // chunked: per-workgroup counts
    if constexpr (MODE == 0) {
        for (uint32_t g = tid; g < nd; g += ST) hist[g] = 0;
        __syncthreads();
        for (uint64_t c0 = 0; c0 < n; c0 += TK) {
            Key<W> key[IPT];
            load_tile<W, IPT, ST>(a.src + (first + c0) * W, n - c0, tid, key);
#pragma unroll
            for (int it = 0; it < IPT; it++)
                if (c0 + (uint64_t)it * ST + tid < n) atomicAdd(&hist[key_bits<W>(key[it], a.shift, a.dbits)], 1u);
        }
        __syncthreads();
        for (uint32_t g = tid; g < nd; g += ST) {
            const uint32_t c = hist[g];
            if (a.chunk > 1) row[g] = c;  // one-tile chunks recount in the scatter pass
            if (c) atomicAdd(&a.child_hist[sg.child_base + g], c);
        }
    } else if constexpr (MODE == 1) {
        // one tile: its keys stay in registers between the count and the
        // scatter; digits are recomputed rather than held
        Key<W> key[IPT];
        load_tile<W, IPT, ST>(a.src + first * W, n, tid, key);
        for (uint32_t g = tid; g < nd; g += ST) hist[g] = 0;
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; it++)
            if ((uint64_t)it * ST + tid < n) atomicAdd(&hist[key_bits<W>(key[it], a.shift, a.dbits)], 1u);
        __syncthreads();
        for (uint32_t g = tid; g < nd; g += ST) {
            const uint32_t c = hist[g];
            if (c) cbase32[g] = (uint32_t)(atomicAdd((unsigned long long *)&a.child_cur[sg.child_base + g],
                                                     (unsigned long long)c) - sg.start);
            hist[g] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            uint64_t i = (uint64_t)it * ST + tid;
            if (i < n) {
                const uint32_t dg = key_bits<W>(key[it], a.shift, a.dbits);
                const uint64_t slot = sg.start + cbase32[dg] + atomicAdd(&hist[dg], 1u);
                uint64_t *d = a.dst + slot * W;
#pragma unroll
                for (int r = 0; r < W; r++) d[r] = key[it].w[r];
            }
        }
    } else {
        for (uint32_t g = tid; g < nd; g += ST) {
            const uint32_t c = row[g];
            if (c) cbase[g] = atomicAdd((unsigned long long *)&a.child_cur[sg.child_base + g],
                                        (unsigned long long)c);
            hist[g] = 0;
        }
        __syncthreads();
        for (uint64_t c0 = 0; c0 < n; c0 += TK) {
            Key<W> key[IPT];
            uint32_t dig[IPT];
            load_tile<W, IPT, ST>(a.src + (first + c0) * W, n - c0, tid, key);
#pragma unroll
            for (int it = 0; it < IPT; it++) dig[it] = key_bits<W>(key[it], a.shift, a.dbits);
#pragma unroll
            for (int it = 0; it < IPT; it++) {
                uint64_t i = c0 + (uint64_t)it * ST + tid;
                if (i < n) {
                    uint64_t slot = cbase[dig[it]] + atomicAdd(&hist[dig[it]], 1u);
                    uint64_t *d = a.dst + slot * W;
#pragma unroll
                    for (int r = 0; r < W; r++) d[r] = key[it].w[r];
                }
            }
        }
    }
}

template <int W, int MODE>
__global__ __launch_bounds__(ST) void k_split(SplitArgs a) {
    split_tile<W, MODE>(a, xcd_tile(blockIdx.x, gridDim.x));
}

// Scatter pass of one tile through LDS (children up to 2^SPLIT_STAGED_BITS):
// each key's rank among its child's keys comes from the LDS count itself,
// the tile is regrouped by child in LDS, and the grouped tile is stored in
// order, so consecutive lanes write consecutive addresses of one child's
// run (about TK / 2^dbits keys) rather than one scattered store per key.
// Same tiles, reservations and 32-bit segment offsets as MODE 1; 80 KiB of
// LDS, two workgroups per CU.
constexpr int SST = 512;
constexpr int SPLIT_STAGED_BITS = 11;
template <int W>
__device__ __forceinline__ void split_staged_tile(const SplitArgs &a, uint64_t b) {
    constexpr int TK = SplitCfg<W>::TK, IPT = TK / SST;
    constexpr int ND = 1 << SPLIT_STAGED_BITS, CPT = ND / SST;
    static_assert(TK % SST == 0 && ND % SST == 0, "tile and children split evenly over the threads");
    __shared__ uint64_t stage[TK * W];
    __shared__ uint32_t loff[ND];  // child counts, then each child's first LDS position
    __shared__ uint32_t gofs[ND];  // child's destination minus its LDS position (mod 2^32)
    const int tid = threadIdx.x;
    uint32_t lo = 0, hi = a.nseg;  // largest s with tile_pre[s] <= b
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (a.tile_pre[mid] <= b) lo = mid;
        else hi = mid;
    }
    const Seg sg = a.segs[lo];
    const uint64_t t = b - a.tile_pre[lo];
    const uint64_t first = sg.start + t * TK;
    const uint32_t n = (uint32_t)std::min<uint64_t>(TK, sg.size - t * TK);
    const uint32_t nd = 1u << a.dbits;
    Key<W> key[IPT];
    load_tile<W, IPT, SST>(a.src + first * W, n, tid, key);
    for (uint32_t g = tid; g < nd; g += SST) loff[g] = 0;
    __syncthreads();
    uint32_t dig[IPT], rk[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        if ((uint32_t)it * SST + tid < n) {
            dig[it] = key_bits<W>(key[it], a.shift, a.dbits);
            rk[it] = atomicAdd(&loff[dig[it]], 1u);
        }
    }
    __syncthreads();
    {
        // thread tid owns children [tid*cpt, tid*cpt + cpt)
        const uint32_t cpt = (nd + SST - 1) / SST, g0 = (uint32_t)tid * cpt;
        uint32_t c[CPT], sum = 0;
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            c[q] = ((uint32_t)q < cpt && g0 + q < nd) ? loff[g0 + q] : 0u;
            sum += c[q];
        }
        uint32_t tot;
        // the stage is free until the regrouping below
        uint32_t run = block_excl_scan<SST, uint32_t>(sum, reinterpret_cast<uint32_t *>(stage), &tot);
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            if ((uint32_t)q < cpt && g0 + q < nd) {
                const uint32_t g = g0 + q;
                if (c[q])
                    gofs[g] = (uint32_t)(atomicAdd((unsigned long long *)&a.child_cur[sg.child_base + g],
                                                   (unsigned long long)c[q]) - sg.start) - run;
                loff[g] = run;
                run += c[q];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        const uint32_t i = (uint32_t)it * SST + tid;
        if (i < n) {
            const uint32_t d = loff[dig[it]] + rk[it];
#pragma unroll
            for (int r = 0; r < W; r++) stage[d * W + r] = key[it].w[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        const uint32_t i = (uint32_t)it * SST + tid;
        if (i < n) {
            Key<W> kk;
#pragma unroll
            for (int r = 0; r < W; r++) kk.w[r] = stage[i * W + r];
            const uint32_t g = key_bits<W>(kk, a.shift, a.dbits);
            uint64_t *d = a.dst + (sg.start + (uint32_t)(gofs[g] + i)) * W;
#pragma unroll
            for (int r = 0; r < W; r++) d[r] = kk.w[r];
        }
    }
}

template <int W>
__global__ __launch_bounds__(SST) void k_split_staged(SplitArgs a) {
    split_staged_tile<W>(a, xcd_tile(blockIdx.x, gridDim.x));
}

// ---- per-bucket LSD radix sort in LDS + run-length counts ------------------
// One 256-thread workgroup per bucket; IPT (items per thread) sets the size
// class, so small buckets take little LDS and many workgroups share a CU.
// Keys are held warp-striped: wave w owns positions [w*64*ipt, (w+1)*64*ipt),
// item `it` of lane l is position w*64*ipt + it*64 + l.  Ranking inside a
// wave uses ballot-matched digits, processed item by item, so every pass is
// stable in position order; waves are combined by a (digit, wave) scan.
template <int W, int IPT>
__device__ __forceinline__ void lsd_sort_bucket(uint64_t *__restrict__ A, uint64_t *__restrict__ B,
                                                uint32_t *__restrict__ counts_tmp, const Bucket *__restrict__ buckets,
                                                const uint32_t bid, uint32_t *__restrict__ ucount, int k) {
    constexpr int CAP = BT * IPT;
    __shared__ uint64_t skeys[CAP * W];
    __shared__ uint32_t whist[256 * BNW];  // [digit][wave]
    __shared__ uint16_t upos[CAP + 1];
    __shared__ uint32_t ws[BNW + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const Bucket bk = buckets[bid];
    uint64_t *X = bk.array ? B : A;
    const uint64_t start = bk.start;
    const int bits = 2 * k - (int)bk.prefix;
    if (bk.size == 0) {
        if (tid == 0) ucount[bid] = 0;
        return;
    }
    if (bits <= 0 || bk.size == 1) {
        // every key in the bucket is identical: one run
        if (tid == 0) {
            counts_tmp[start] = bk.size > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bk.size;
            ucount[bid] = 1;
        }
        return;
    }
    const uint32_t m = (uint32_t)bk.size;  // <= CAP by construction
    const int ipt = (int)((m + BT - 1) / BT);
    const uint32_t wbase = (uint32_t)wave * 64u * (uint32_t)ipt;

    Key<W> key[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < m) {
            const uint64_t *p = X + (start + pos) * W;
#pragma unroll
            for (int r = 0; r < W; r++) key[it].w[r] = p[r];
        }
    }
    for (int shift = 0; shift < bits; shift += 8) {
        const int nb = std::min(8, bits - shift);
        for (int j = tid; j < 256 * BNW; j += BT) whist[j] = 0;
        __syncthreads();
        uint32_t rank[IPT], dig[IPT];
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            if (it < ipt) {
                uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
                bool valid = pos < m;
                uint32_t d = valid ? key_bits<W>(key[it], shift, nb) : 0u;
                uint64_t peers = __ballot(valid);
                for (int bb = 0; bb < nb; bb++) {
                    uint64_t mb = __ballot(valid && ((d >> bb) & 1u));
                    peers &= ((d >> bb) & 1u) ? mb : ~mb;
                }
                uint32_t below = __popcll(peers & lanemask_lt());
                uint32_t basev = whist[d * BNW + wave];
                rank[it] = basev + below;
                dig[it] = d;
                if (valid && below == 0) whist[d * BNW + wave] = basev + (uint32_t)__popcll(peers);
            }
        }
        __syncthreads();
        {
            // exclusive scan over whist in (digit, wave) order: 4 entries per thread
            uint32_t v0 = whist[tid * 4 + 0], v1 = whist[tid * 4 + 1];
            uint32_t v2 = whist[tid * 4 + 2], v3 = whist[tid * 4 + 3];
            uint32_t total;
            uint32_t ex = block_excl_scan<BT, uint32_t>(v0 + v1 + v2 + v3, ws, &total);
            whist[tid * 4 + 0] = ex;
            whist[tid * 4 + 1] = ex + v0;
            whist[tid * 4 + 2] = ex + v0 + v1;
            whist[tid * 4 + 3] = ex + v0 + v1 + v2;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
            if (it < ipt && pos < m) {
                uint32_t dst = whist[dig[it] * BNW + wave] + rank[it];
#pragma unroll
                for (int r = 0; r < W; r++) skeys[dst * W + r] = key[it].w[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
            if (it < ipt && pos < m) {
#pragma unroll
                for (int r = 0; r < W; r++) key[it].w[r] = skeys[pos * W + r];
            }
        }
        __syncthreads();
    }
    // run-length counts over the sorted keys in skeys[0, m)
    uint32_t heads = 0;
    const uint32_t p0 = (uint32_t)tid * (uint32_t)ipt;
    for (int it = 0; it < ipt; it++) {
        uint32_t p = p0 + it;
        if (p >= m) break;
        bool h = p == 0;
        if (!h) {
#pragma unroll
            for (int r = 0; r < W; r++) h = h || skeys[p * W + r] != skeys[(p - 1) * W + r];
        }
        heads += h;
    }
    uint32_t utotal;
    uint32_t u = block_excl_scan<BT, uint32_t>(heads, ws, &utotal);
    for (int it = 0; it < ipt; it++) {
        uint32_t p = p0 + it;
        if (p >= m) break;
        bool h = p == 0;
        if (!h) {
#pragma unroll
            for (int r = 0; r < W; r++) h = h || skeys[p * W + r] != skeys[(p - 1) * W + r];
        }
        if (h) upos[u++] = (uint16_t)p;
    }
    if (tid == 0) upos[utotal] = (uint16_t)m;
    __syncthreads();
    for (uint32_t q = tid; q < utotal; q += BT) {
        uint32_t p = upos[q];
        uint32_t c = (uint32_t)upos[q + 1] - p;
        uint64_t *d = X + (start + q) * W;
#pragma unroll
        for (int r = 0; r < W; r++) d[r] = skeys[p * W + r];
        counts_tmp[start + q] = c;
    }
    if (tid == 0) ucount[bid] = utotal;
}

template <int W, int IPT>
__global__ __launch_bounds__(BT) void k_bucket_sort(uint64_t *__restrict__ A, uint64_t *__restrict__ B,
                                                     uint32_t *__restrict__ counts_tmp,
                                                     const Bucket *__restrict__ buckets,
                                                     const uint32_t *__restrict__ order,
                                                     uint32_t *__restrict__ ucount, int k) {
    lsd_sort_bucket<W, IPT>(A, B, counts_tmp, buckets, order[blockIdx.x], ucount, k);
}

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

// Waves per SIMD of every hash-sort class: about 18 B of LDS per item (keys,
// 16-bit counts, two table slots) fits 8 waves, and the registers are
// capped to match.
constexpr int hash_sort_waves(int ipt) { return ipt > 4 ? 4 : 8; }

// Hash-then-bin variant: every key is inserted into an LDS open-addressing
// table (a slot holds the first position of its key, claimed by CAS) and
// duplicates add to that position's count.  The unique keys, carrying their
// counts, are then distributed by one counting pass over CAP / 2 bins on
// their top remaining key bits (order inside a bin does not matter: keys
// are distinct), and each key's place inside its bin is its rank among the
// bin's keys (every key compares itself with the others of its bin at once,
// instead of a per-thread insertion sort whose steps wait on each other).
// A bin above LOCAL_MAX keys (skewed key bits) hands the bucket, untouched,
// to k_bucket_sort through fb_list (the host launches it over the
// handed-back buckets), which keeps the LSD code and its registers out of
// this kernel.  Output equals k_bucket_sort's: ascending unique keys with
// their run lengths.
template <int W, int IPT, int NT>
__device__ __forceinline__ void hash_sort_bucket(uint64_t *__restrict__ A, uint64_t *__restrict__ B,
                                                 uint32_t *__restrict__ counts_tmp,
                                                 const Bucket *__restrict__ buckets, const uint32_t bid,
                                                 uint32_t *__restrict__ ucount, int k,
                                                 uint32_t *__restrict__ fb_list, uint32_t *__restrict__ fb_count) {
    constexpr int NW = NT / 64;
    constexpr int CAP = NT * IPT;
    constexpr int HT = 2 * CAP;  // hash slots (load factor <= 1/2), then bins
    constexpr int NB = HT / 4 >= NT ? HT / 4 : NT;  // bins of the counting pass
    constexpr int BPT = NB / NT; // bins per thread
    constexpr int DB = ilog2c(NB);
    constexpr int LOCAL_MAX = 16;
    static_assert(NB <= HT && NB % NT == 0, "bins alias the hash table");
    __shared__ uint64_t skeys[CAP * W];
    __shared__ uint16_t scnt[CAP];  // counts <= CAP < 2^16; pairs updated as one u32
    __shared__ uint32_t ht[HT + 1];  // [NB]: the number of unique keys once scanned
    __shared__ uint32_t ws[NW + 1];
    __shared__ uint32_t s_fallback;
    uint32_t *scnt32 = reinterpret_cast<uint32_t *>(scnt);
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const Bucket bk = buckets[bid];
    uint64_t *X = bk.array ? B : A;
    const uint64_t start = bk.start;
    const int bits = 2 * k - (int)bk.prefix;
    if (bk.size == 0) {
        if (tid == 0) ucount[bid] = 0;
        return;
    }
    if (bits <= 0 || bk.size == 1) {
        if (tid == 0) {
            counts_tmp[start] = bk.size > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bk.size;
            ucount[bid] = 1;
        }
        return;
    }
    const uint32_t m = (uint32_t)bk.size;
    int ipt = (int)((m + NT - 1) / NT);
    uint32_t wbase = (uint32_t)wave * 64u * (uint32_t)ipt;
    Key<W> key[IPT];
    // all loads first (positions past the bucket read its last key), then
    // the table clears while they are in flight, then the LDS copies
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        const uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        const uint64_t *p = X + (start + (it < ipt && pos < m ? pos : m - 1)) * W;
#pragma unroll
        for (int r = 0; r < W; r++) key[it].w[r] = p[r];
    }
    for (int j = tid; j < HT; j += NT) ht[j] = 0;
    for (int j = tid; j < CAP / 2; j += NT) scnt32[j] = 0;
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        const uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < m)
#pragma unroll
            for (int r = 0; r < W; r++) skeys[pos * W + r] = key[it].w[r];
    }
    __syncthreads();
    // 1. hash dedup: the first position of every key collects its count.
    // Probes go in rounds over all of a thread's items, so the independent
    // LDS compare-and-swaps of one round are in flight together.
    uint32_t h[IPT], rep[IPT];
    uint32_t pending = 0;
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < m) {
            uint64_t f = key[it].w[0];
#pragma unroll
            for (int r = 1; r < W; r++) f ^= key[it].w[r] * 0xC2B2AE3D27D4EB4Full;
            h[it] = (uint32_t)((f * 0xD6E8FEB86659FD93ull) >> 40) & (HT - 1);
            pending |= 1u << it;
        }
        rep[it] = 0;
    }
    while (pending) {
        uint32_t old[IPT];
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            const uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
            old[it] = ((pending >> it) & 1u) ? atomicCAS(&ht[h[it]], 0u, pos + 1) : 0u;
        }
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            if (!((pending >> it) & 1u)) continue;
            const uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
            if (old[it] == 0) {
                rep[it] = pos;
                pending &= ~(1u << it);
                continue;
            }
            bool eq = true;
#pragma unroll
            for (int r = 0; r < W; r++) eq = eq && skeys[(old[it] - 1) * W + r] == key[it].w[r];
            if (eq) {
                rep[it] = old[it] - 1;
                pending &= ~(1u << it);
            } else {
                h[it] = (h[it] + 1) & (HT - 1);
            }
        }
    }
    uint32_t is_rep = 0;  // bit it: item it holds the first position of its key
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < m) {
            atomicAdd(&scnt32[rep[it] >> 1], 1u << (16 * (rep[it] & 1u)));
            if (rep[it] == pos) is_rep |= 1u << it;
        }
    }
    __syncthreads();
    // 2. one counting pass of the representatives over NB bins on their top
    // remaining key bits
    const int db = bits < DB ? bits : DB;
    const int dshift = bits - db;
    for (int j = tid; j < NB; j += NT) ht[j] = 0;
    if (tid == 0) s_fallback = 0;
    __syncthreads();
    uint32_t cnt[IPT], slot[IPT], dig[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        if ((is_rep >> it) & 1u) {
            const uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
            cnt[it] = scnt[pos];
            dig[it] = key_bits<W>(key[it], dshift, db);
            slot[it] = atomicAdd(&ht[dig[it]], 1u);
        }
    }
    __syncthreads();
    uint32_t U;
    {
        uint32_t v[BPT], sum = 0, big = 0;
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            v[b] = ht[tid * BPT + b];
            sum += v[b];
            big |= v[b] > (uint32_t)LOCAL_MAX;
        }
        // (ws is not written again in this kernel)
        uint32_t run = block_excl_scan<NT, uint32_t, false>(sum, ws, &U);
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            ht[tid * BPT + b] = run;
            run += v[b];
        }
        if (tid == NT - 1) ht[NB] = run;
        if (big) s_fallback = 1;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        if ((is_rep >> it) & 1u) {
            const uint32_t dst = ht[dig[it]] + slot[it];
#pragma unroll
            for (int r = 0; r < W; r++) skeys[dst * W + r] = key[it].w[r];
            scnt[dst] = (uint16_t)cnt[it];
        }
    }
    __syncthreads();
    if (s_fallback) {
        // a bin above LOCAL_MAX keys (skewed key bits): the bucket is listed
        // for the LSD kernel, which sorts it from its stored keys
        if (tid == 0) fb_list[atomicAdd(fb_count, 1u)] = bid;
        return;
    }
    {
        // 3. the place of each unique key inside its bin (at most LOCAL_MAX
        // keys): its rank among the bin's keys, four compares per step
        uint32_t fin[IPT];
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            fin[it] = 0;
            if ((is_rep >> it) & 1u) {
                const uint32_t lo = ht[dig[it]], hi = ht[dig[it] + 1];
                uint32_t rank = 0;
                for (uint32_t j = lo; j < hi; j += 4) {
                    Key<W> o[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t jj = j + q < hi ? j + q : hi - 1;
#pragma unroll
                        for (int r = 0; r < W; r++) o[q].w[r] = skeys[jj * W + r];
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++) rank += (j + q < hi && key_less<W>(o[q], key[it])) ? 1u : 0u;
                }
                fin[it] = lo + rank;
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            if ((is_rep >> it) & 1u) {
#pragma unroll
                for (int r = 0; r < W; r++) skeys[fin[it] * W + r] = key[it].w[r];
                scnt[fin[it]] = (uint16_t)cnt[it];
            }
        }
        __syncthreads();
    }
    for (uint32_t q2 = tid; q2 < U; q2 += NT) {
        uint64_t *d = X + (start + q2) * W;
#pragma unroll
        for (int r = 0; r < W; r++) d[r] = skeys[q2 * W + r];
        counts_tmp[start + q2] = scnt[q2];
    }
    if (tid == 0) ucount[bid] = U;
}

template <int W, int IPT, int NT>
__global__ __launch_bounds__(NT, hash_sort_waves(IPT)) void k_bucket_hash_sort(uint64_t *__restrict__ A,
                                                          uint64_t *__restrict__ B,
                                                          uint32_t *__restrict__ counts_tmp,
                                                          const Bucket *__restrict__ buckets,
                                                          const uint32_t *__restrict__ order,
                                                          uint32_t *__restrict__ ucount, int k,
                                                          uint32_t *__restrict__ fb_list,
                                                          uint32_t *__restrict__ fb_count) {
    hash_sort_bucket<W, IPT, NT>(A, B, counts_tmp, buckets, order[blockIdx.x], ucount, k, fb_list, fb_count);
}

template <int W>
__global__ __launch_bounds__(256) void k_compact(const uint64_t *__restrict__ A,
                                                 const uint64_t *__restrict__ B,
                                                 const uint32_t *__restrict__ counts_tmp,
                                                 const Bucket *__restrict__ buckets,
                                                 const uint64_t *__restrict__ uoff,
                                                 uint64_t *__restrict__ out_keys,
                                                 uint32_t *__restrict__ out_counts) {
    const Bucket bk = buckets[blockIdx.x];
    const uint64_t *X = bk.array ? B : A;
    const uint64_t o = uoff[blockIdx.x];
    const uint64_t u = uoff[blockIdx.x + 1] - o;
    for (uint64_t i = threadIdx.x; i < u; i += 256) {
#pragma unroll
        for (int r = 0; r < W; r++) out_keys[(o + i) * W + r] = X[(bk.start + i) * W + r];
        out_counts[o + i] = counts_tmp[bk.start + i];
    }
}

__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint32_t *__restrict__ idx,
                             uint32_t n, uint64_t *__restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---- bucket plan on the device ------------------------------------------------
// The node list (Bucket + partition per node) starts as one node per
// partition.  Every MSD level replaces the nodes above the largest LDS
// capacity by their non-empty children (in child order, so the list stays
// in key order within each partition); the host only reads back a few
// counts per level.

__global__ void k_init_nodes(const uint64_t *__restrict__ part_off, uint32_t P, Bucket *__restrict__ nodes,
                             uint32_t *__restrict__ npart) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    nodes[p] = Bucket{part_off[p], part_off[p + 1] - part_off[p], 0u, 0u};
    npart[p] = p;
}

// stats[0] = largest active size, stats[1] = keys in active nodes
__global__ void k_level_flags(const Bucket *__restrict__ nodes, uint32_t nn, uint64_t cap, int kbits, uint64_t tkl,
                              uint32_t *__restrict__ act, uint64_t *__restrict__ tiles,
                              unsigned long long *__restrict__ stats) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const Bucket b = nodes[i];
    const bool a = b.size > cap && (int)b.prefix < kbits;
    act[i] = a ? 1u : 0u;
    tiles[i] = a ? (b.size + tkl - 1) / tkl : 0;
    if (a) {
        atomicMax(&stats[0], (unsigned long long)b.size);
        atomicAdd(&stats[1], (unsigned long long)b.size);
    }
}

__global__ void k_level_segs(const Bucket *__restrict__ nodes, uint32_t nn, const uint32_t *__restrict__ act,
                             const uint64_t *__restrict__ act_rank, const uint64_t *__restrict__ tile_off, uint32_t nd,
                             Seg *__restrict__ segs, uint64_t *__restrict__ tile_pre, uint32_t *__restrict__ seg_node) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) tile_pre[act_rank[nn]] = tile_off[nn];
    if (i >= nn || !act[i]) return;
    const uint64_t r = act_rank[i];
    const Bucket b = nodes[i];
    segs[r] = Seg{b.start, b.size, r * nd};
    tile_pre[r] = tile_off[i];
    seg_node[r] = i;
}

// One workgroup per split segment: child start offsets (exclusive scan of
// the child counts from the segment start), each child's rank among the
// non-empty children and their number.
__device__ __forceinline__ void child_scan_seg(const Seg *__restrict__ segs, const uint32_t *__restrict__ child_hist,
                                               uint32_t nd, uint64_t *__restrict__ child_cur,
                                               uint32_t *__restrict__ nz_rank, uint32_t *__restrict__ nz_count,
                                               uint32_t r) {
    __shared__ uint64_t ws64[256 / 64 + 1];
    __shared__ uint32_t ws32[256 / 64 + 1];
    const uint64_t base = (uint64_t)r * nd;
    const uint32_t per = (nd + 255) / 256;
    const uint32_t g0 = threadIdx.x * per;
    uint64_t csum = 0;
    uint32_t zsum = 0;
    for (uint32_t j = 0; j < per; j++) {
        const uint32_t g = g0 + j;
        if (g < nd) {
            const uint32_t c = child_hist[base + g];
            csum += c;
            zsum += c ? 1u : 0u;
        }
    }
    uint64_t ctot;
    uint32_t ztot;
    uint64_t cex = block_excl_scan<256, uint64_t>(csum, ws64, &ctot);
    uint32_t zex = block_excl_scan<256, uint32_t>(zsum, ws32, &ztot);
    const uint64_t start = segs[r].start;
    for (uint32_t j = 0; j < per; j++) {
        const uint32_t g = g0 + j;
        if (g < nd) {
            const uint32_t c = child_hist[base + g];
            child_cur[base + g] = start + cex;
            nz_rank[base + g] = zex;
            cex += c;
            zex += c ? 1u : 0u;
        }
    }
    if (threadIdx.x == 0) nz_count[r] = ztot;
}

__global__ __launch_bounds__(256) void k_child_scan(const Seg *__restrict__ segs, const uint32_t *__restrict__ child_hist,
                                                    uint32_t nd, uint64_t *__restrict__ child_cur,
                                                    uint32_t *__restrict__ nz_rank, uint32_t *__restrict__ nz_count) {
    child_scan_seg(segs, child_hist, nd, child_cur, nz_rank, nz_count, blockIdx.x);
}

__global__ void k_level_counts(uint32_t nn, const uint32_t *__restrict__ act, const uint64_t *__restrict__ act_rank,
                               const uint32_t *__restrict__ nz_count, uint32_t *__restrict__ out_cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nn) out_cnt[i] = act[i] ? nz_count[act_rank[i]] : 1u;
}

__global__ void k_level_emit_nodes(const Bucket *__restrict__ nodes, const uint32_t *__restrict__ npart, uint32_t nn,
                                   const uint32_t *__restrict__ act, const uint64_t *__restrict__ new_base,
                                   Bucket *__restrict__ out, uint32_t *__restrict__ out_part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn || act[i]) return;
    const uint64_t o = new_base[i];
    out[o] = nodes[i];
    out_part[o] = npart[i];
}

// Children of split node seg_node[r]: runs at the child start offsets (read
// before the scatter pass advances them), one level deeper, in the other
// key array.
__global__ void k_level_emit_children(const uint32_t *__restrict__ seg_node, const uint64_t *__restrict__ new_base,
                                      const Bucket *__restrict__ nodes, const uint32_t *__restrict__ npart,
                                      const uint32_t *__restrict__ child_hist, const uint64_t *__restrict__ child_cur,
                                      const uint32_t *__restrict__ nz_rank, uint64_t nchild, uint32_t nd, int d,
                                      Bucket *__restrict__ out, uint32_t *__restrict__ out_part) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nchild) return;
    const uint32_t c = child_hist[id];
    if (!c) return;
    const uint32_t i = seg_node[id / nd];
    const Bucket b = nodes[i];
    const uint64_t o = new_base[i] + nz_rank[id];
    out[o] = Bucket{child_cur[id], c, b.prefix + (uint32_t)d, b.array ^ 1u};
    out_part[o] = npart[i];
}

struct ClassCaps {
    uint64_t cap[SORT_CLASSES];
};

// Size class of every bucket: the smallest LDS footprint that holds it.
// Counts are aggregated per workgroup of CT threads (wave ballots into LDS,
// then one global atomic per class per workgroup): the four counters would
// otherwise take one contended atomic per bucket, or per wave.
constexpr int CT = 1024;
__global__ __launch_bounds__(CT) void k_classify(const Bucket *__restrict__ nodes, uint32_t nb, int kbits,
                                                 ClassCaps caps, uint32_t *__restrict__ cls,
                                                 uint32_t *__restrict__ cls_count) {
    __shared__ uint32_t wc[SORT_CLASSES];
    if (threadIdx.x < SORT_CLASSES) wc[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    int c = -1;
    if (b < nb) {
        const Bucket bk = nodes[b];
        c = 0;
        if ((int)bk.prefix < kbits && bk.size > 1)
            while (c < SORT_CLASSES - 1 && bk.size > caps.cap[c]) c++;
        cls[b] = (uint32_t)c;
    }
#pragma unroll
    for (int q = 0; q < SORT_CLASSES; q++) {
        const uint64_t m = __ballot(c == q);
        if (m && __lane_id() == (unsigned)__ffsll((unsigned long long)m) - 1u) atomicAdd(&wc[q], (uint32_t)__popcll(m));
    }
    __syncthreads();
    if (threadIdx.x < SORT_CLASSES && wc[threadIdx.x]) atomicAdd(&cls_count[threadIdx.x], wc[threadIdx.x]);
}

// order[cls_off[c] ...]: the buckets of class c (any order inside a class);
// each workgroup reserves its share of every class with one cursor atomic
// and places its buckets by wave ballots.
__global__ __launch_bounds__(CT) void k_order_fill(const uint32_t *__restrict__ cls, uint32_t nb,
                                                   const uint32_t *__restrict__ cls_off,
                                                   uint32_t *__restrict__ cursor, uint32_t *__restrict__ order) {
    constexpr int NWV = CT / 64;
    __shared__ uint32_t wofs[NWV][SORT_CLASSES];  // per wave count, then its offset in the workgroup
    __shared__ uint32_t base[SORT_CLASSES];
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = b < nb ? (int)cls[b] : -1;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    uint64_t mine = 0;  // ballot of this thread's class
#pragma unroll
    for (int q = 0; q < SORT_CLASSES; q++) {
        const uint64_t m = __ballot(c == q);
        if (c == q) mine = m;
        if (lane == 0) wofs[wave][q] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < SORT_CLASSES) {
        const int q = (int)threadIdx.x;
        uint32_t run = 0;
        for (int w = 0; w < NWV; w++) {
            const uint32_t t = wofs[w][q];
            wofs[w][q] = run;
            run += t;
        }
        base[q] = run ? atomicAdd(&cursor[q], run) : 0u;
    }
    __syncthreads();
    if (c >= 0) order[cls_off[c] + base[c] + wofs[wave][c] + (uint32_t)__popcll(mine & lanemask_lt())] = b;
}

// First node of every partition p in [0, P] (node partitions ascend).
__global__ void k_part_first(const uint32_t *__restrict__ npart, uint32_t nb, uint32_t P, uint32_t *__restrict__ pfirst) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > P) return;
    uint32_t lo = 0, hi = nb;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (npart[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    pfirst[p] = lo;
}

// Small results go to the host through page-locked memory written by a
// kernel, not through the copy engines, so they never wait behind large
// transfers that other streams keep in flight.
template <class T>
__global__ void k_put_host(const T *__restrict__ src, uint32_t n, T *__restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Class offsets from the class counts; the counters become zeroed cursors.
__global__ void k_class_offsets(uint32_t *__restrict__ cls_count, uint32_t *__restrict__ cls_off) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t o = 0;
    for (int c = 0; c < SORT_CLASSES; c++) {
        cls_off[c] = o;
        o += cls_count[c];
        cls_count[c] = 0;
    }
    cls_off[SORT_CLASSES] = o;
}

// ---- device-planned count ----------------------------------------------------
// The kernels of a launch sequence fixed before the data is seen (so that a
// HIP graph can hold a whole count): every size the classic orchestration
// reads back is kept in the DevPlan in device memory and read there, grids
// are fixed and loop over the plan's counts, and every kernel returns at once
// when the plan records an overflow (a capacity exceeded: the caller then
// counts that input the classic way).  Grids that map tiles to XCDs are
// multiples of the XCD count, so a grid-stride step keeps a tile's XCD.

__global__ void k_dev_plan_init(const uint64_t *__restrict__ h_len, uint64_t cap_len, uint32_t P, uint32_t subs,
                                DevPlan *__restrict__ pl) {
    uint32_t *w = reinterpret_cast<uint32_t *>(pl);
    for (uint32_t i = threadIdx.x; i < sizeof(DevPlan) / 4; i += blockDim.x) w[i] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t len = *h_len;  // page-locked host word, set before the launch
        if (len > cap_len) pl->overflow |= 1u;
        pl->len = len > cap_len ? 0 : len;
        pl->n_sub = pl->len == 0 ? 1 : ceil_div(pl->len, SUB);
        pl->nwg = ceil_div(pl->n_sub, subs);
        pl->lv[0].nn = P;
    }
}

__global__ __launch_bounds__(XT) void k_nl_count_dev(const uint8_t *__restrict__ fq, const DevPlan *__restrict__ pl,
                                                      uint64_t n_sub_cap, uint32_t *__restrict__ out) {
    __shared__ uint32_t ws[XT / 64 + 1];
    const uint64_t len = pl->len, n_sub = pl->n_sub;
    for (uint64_t b = blockIdx.x; b < n_sub_cap; b += gridDim.x) {
        if (b < n_sub) nl_count_tile(fq, len, b, out, ws);
        else if (threadIdx.x == 0) out[b] = 0;  // past the input: the scan runs over the capacity
    }
}

// The k-mer total; beyond the key buffers, or with a partition of 2^32 keys
// or more (the one-tile split scatters keep 32-bit offsets within a
// partition), the count is void.
__global__ void k_dev_kmers(const uint64_t *__restrict__ part_off, uint32_t P, uint64_t n_cap,
                            DevPlan *__restrict__ pl) {
    const uint64_t n = part_off[P];
    pl->n = n;
    if (n > n_cap) pl->overflow |= 1u;
    for (uint32_t p = 0; p < P; p++)
        if (part_off[p + 1] - part_off[p] > 0xFFFFFFFFull) pl->overflow |= 1u;
}

__global__ void k_dev_init_nodes(const uint64_t *__restrict__ part_off, uint32_t P, Bucket *__restrict__ nodes,
                                 uint32_t *__restrict__ npart, const DevPlan *__restrict__ pl) {
    if (pl->overflow) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    nodes[p] = Bucket{part_off[p], part_off[p + 1] - part_off[p], 0u, 0u};
    npart[p] = p;
}

#define KITE_GRID_STRIDE(i, n) \
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

__global__ void k_dev_level_flags(const Bucket *__restrict__ nodes, DevPlan *__restrict__ pl, int l, uint64_t cap,
                                  int kbits, uint64_t tkl, uint32_t *__restrict__ act, uint64_t *__restrict__ tiles) {
    if (pl->overflow) return;
    DevLevel &lv = pl->lv[l];
    KITE_GRID_STRIDE(i, lv.nn) {
        const Bucket b = nodes[i];
        const bool a = b.size > cap && (int)b.prefix < kbits;
        act[i] = a ? 1u : 0u;
        tiles[i] = a ? (b.size + tkl - 1) / tkl : 0;
        if (a) {
            atomicMax(&lv.maxsz, (unsigned long long)b.size);
            atomicAdd(&lv.keys, (unsigned long long)b.size);
        }
    }
}

// The level's digit, as the classic orchestration picks it from the largest
// node: the fewest bits (up to maxd) that bring it to `target` keys.
__global__ void k_dev_level_plan(DevPlan *__restrict__ pl, int l, const uint64_t *__restrict__ act_rank,
                                 const uint64_t *__restrict__ tile_off, int kbits, int maxd, uint64_t target,
                                 int staged_ok, uint64_t nchild_cap) {
    if (pl->overflow) return;
    DevLevel &lv = pl->lv[l];
    lv.nact = act_rank[lv.nn];
    lv.tiles = tile_off[lv.nn];
    int d = 0;
    if (lv.nact) {
        d = 1;
        while (d < maxd && (lv.maxsz >> d) > target) d++;
        d = d < kbits - lv.prefix ? d : kbits - lv.prefix;
    }
    lv.d = d;
    lv.shift = kbits - lv.prefix - d;
    lv.nchild = lv.nact << d;
    lv.staged = staged_ok && d <= SPLIT_STAGED_BITS ? 1 : 0;
    if (lv.nchild > nchild_cap) pl->overflow |= 8u;
}

__global__ void k_dev_level_segs(const Bucket *__restrict__ nodes, const DevPlan *__restrict__ pl, int l,
                                 const uint32_t *__restrict__ act, const uint64_t *__restrict__ act_rank,
                                 const uint64_t *__restrict__ tile_off, Seg *__restrict__ segs,
                                 uint64_t *__restrict__ tile_pre, uint32_t *__restrict__ seg_node) {
    if (pl->overflow) return;
    const DevLevel &lv = pl->lv[l];
    const uint32_t nd = 1u << lv.d;
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_pre[lv.nact] = lv.tiles;
    KITE_GRID_STRIDE(i, lv.nn) {
        if (!act[i]) continue;
        const uint64_t r = act_rank[i];
        const Bucket b = nodes[i];
        segs[r] = Seg{b.start, b.size, r * nd};
        tile_pre[r] = tile_off[i];
        seg_node[r] = (uint32_t)i;
    }
}

__global__ void k_dev_zero_children(uint32_t *__restrict__ ch, const DevPlan *__restrict__ pl, int l) {
    if (pl->overflow) return;
    KITE_GRID_STRIDE(i, pl->lv[l].nchild) ch[i] = 0;
}

// Split passes over the level's tiles.  MODE 1 (per-lane stores) runs when
// the level's digit is too wide for the staged scatter.
template <int W, int MODE>
__global__ __launch_bounds__(ST) void k_split_dev(SplitArgs a, const DevPlan *__restrict__ pl, int l) {
    if (pl->overflow) return;
    const DevLevel &lv = pl->lv[l];
    if (lv.nact == 0 || (MODE == 1 && lv.staged)) return;
    a.nseg = (uint32_t)lv.nact;
    a.shift = lv.shift;
    a.dbits = lv.d;
    const uint64_t T = lv.tiles;
    for (uint64_t t = blockIdx.x; t < T; t += gridDim.x) {
        split_tile<W, MODE>(a, xcd_tile(t, T));
        __syncthreads();
    }
}

template <int W>
__global__ __launch_bounds__(SST) void k_split_staged_dev(SplitArgs a, const DevPlan *__restrict__ pl, int l) {
    if (pl->overflow) return;
    const DevLevel &lv = pl->lv[l];
    if (lv.nact == 0 || !lv.staged) return;
    a.nseg = (uint32_t)lv.nact;
    a.shift = lv.shift;
    a.dbits = lv.d;
    const uint64_t T = lv.tiles;
    for (uint64_t t = blockIdx.x; t < T; t += gridDim.x) {
        split_staged_tile<W>(a, xcd_tile(t, T));
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_dev_child_scan(const Seg *__restrict__ segs,
                                                        const uint32_t *__restrict__ child_hist,
                                                        const DevPlan *__restrict__ pl, int l,
                                                        uint64_t *__restrict__ child_cur,
                                                        uint32_t *__restrict__ nz_rank,
                                                        uint32_t *__restrict__ nz_count) {
    if (pl->overflow) return;
    const DevLevel &lv = pl->lv[l];
    for (uint64_t r = blockIdx.x; r < lv.nact; r += gridDim.x) {
        child_scan_seg(segs, child_hist, 1u << lv.d, child_cur, nz_rank, nz_count, (uint32_t)r);
        __syncthreads();
    }
}

__global__ void k_dev_level_counts(const DevPlan *__restrict__ pl, int l, const uint32_t *__restrict__ act,
                                   const uint64_t *__restrict__ act_rank, const uint32_t *__restrict__ nz_count,
                                   uint32_t *__restrict__ out_cnt) {
    if (pl->overflow) return;
    KITE_GRID_STRIDE(i, pl->lv[l].nn) out_cnt[i] = act[i] ? nz_count[act_rank[i]] : 1u;
}

__global__ void k_dev_level_next(DevPlan *__restrict__ pl, int l, const uint64_t *__restrict__ new_base,
                                 uint64_t nn_cap) {
    if (pl->overflow) return;
    const DevLevel &lv = pl->lv[l];
    const uint64_t nn = new_base[lv.nn];
    if (nn > nn_cap) pl->overflow |= 4u;
    pl->lv[l + 1].nn = nn > nn_cap ? 0 : nn;  // a count past the capacity is never stored
    pl->lv[l + 1].prefix = lv.prefix + lv.d;
}

__global__ void k_dev_emit_nodes(const Bucket *__restrict__ nodes, const uint32_t *__restrict__ npart,
                                 const DevPlan *__restrict__ pl, int l, const uint32_t *__restrict__ act,
                                 const uint64_t *__restrict__ new_base, Bucket *__restrict__ out,
                                 uint32_t *__restrict__ out_part) {
    if (pl->overflow) return;
    KITE_GRID_STRIDE(i, pl->lv[l].nn) {
        if (act[i]) continue;
        const uint64_t o = new_base[i];
        out[o] = nodes[i];
        out_part[o] = npart[i];
    }
}

__global__ void k_dev_emit_children(const uint32_t *__restrict__ seg_node, const uint64_t *__restrict__ new_base,
                                    const Bucket *__restrict__ nodes, const uint32_t *__restrict__ npart,
                                    const uint32_t *__restrict__ child_hist, const uint64_t *__restrict__ child_cur,
                                    const uint32_t *__restrict__ nz_rank, const DevPlan *__restrict__ pl, int l,
                                    Bucket *__restrict__ out, uint32_t *__restrict__ out_part) {
    if (pl->overflow) return;
    const DevLevel &lv = pl->lv[l];
    const uint32_t d = (uint32_t)lv.d;
    KITE_GRID_STRIDE(id, lv.nchild) {
        const uint32_t c = child_hist[id];
        if (!c) continue;
        const uint32_t i = seg_node[id >> d];
        const Bucket b = nodes[i];
        const uint64_t o = new_base[i] + nz_rank[id];
        out[o] = Bucket{child_cur[id], c, b.prefix + d, b.array ^ 1u};
        out_part[o] = npart[i];
    }
}

// After the last level no node may still need a split.
__global__ void k_dev_levels_done(DevPlan *__restrict__ pl) {
    if (pl->lv[DEV_LEVELS].maxsz) pl->overflow |= 2u;
}

__global__ void k_dev_part_first(const uint32_t *__restrict__ npart, const DevPlan *__restrict__ pl, uint32_t P,
                                 uint32_t *__restrict__ pfirst) {
    if (pl->overflow) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > P) return;
    uint64_t lo = 0, hi = pl->lv[DEV_LEVELS].nn;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (npart[mid] < p) lo = mid + 1;
        else hi = mid;
    }
    pfirst[p] = (uint32_t)lo;
}

__global__ __launch_bounds__(CT) void k_dev_classify(const Bucket *__restrict__ nodes, DevPlan *__restrict__ pl,
                                                     int kbits, ClassCaps caps, uint32_t *__restrict__ cls) {
    __shared__ uint32_t wc[SORT_CLASSES];
    if (pl->overflow) return;
    const uint64_t nb = pl->lv[DEV_LEVELS].nn;
    if (threadIdx.x < SORT_CLASSES) wc[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t b0 = (uint64_t)blockIdx.x * CT; b0 < nb; b0 += (uint64_t)gridDim.x * CT) {
        const uint64_t b = b0 + threadIdx.x;
        int c = -1;
        if (b < nb) {
            const Bucket bk = nodes[b];
            c = 0;
            if ((int)bk.prefix < kbits && bk.size > 1)
                while (c < SORT_CLASSES - 1 && bk.size > caps.cap[c]) c++;
            cls[b] = (uint32_t)c;
        }
#pragma unroll
        for (int q = 0; q < SORT_CLASSES; q++) {
            const uint64_t m = __ballot(c == q);
            if (m && __lane_id() == (unsigned)__ffsll((unsigned long long)m) - 1u)
                atomicAdd(&wc[q], (uint32_t)__popcll(m));
        }
    }
    __syncthreads();
    if (threadIdx.x < SORT_CLASSES && wc[threadIdx.x]) atomicAdd(&pl->cls_n[threadIdx.x], wc[threadIdx.x]);
}

__global__ void k_dev_class_offsets(DevPlan *__restrict__ pl) {
    if (pl->overflow) return;
    uint32_t o = 0;
    for (int c = 0; c < SORT_CLASSES; c++) {
        pl->cls_off[c] = o;
        o += pl->cls_n[c];
        pl->cls_cur[c] = 0;
    }
    pl->cls_off[SORT_CLASSES] = o;
}

__global__ __launch_bounds__(CT) void k_dev_order_fill(const uint32_t *__restrict__ cls, DevPlan *__restrict__ pl,
                                                       uint32_t *__restrict__ order) {
    constexpr int NWV = CT / 64;
    __shared__ uint32_t wofs[NWV][SORT_CLASSES];
    __shared__ uint32_t base[SORT_CLASSES];
    if (pl->overflow) return;
    const uint64_t nb = pl->lv[DEV_LEVELS].nn;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
    for (uint64_t b0 = (uint64_t)blockIdx.x * CT; b0 < nb; b0 += (uint64_t)gridDim.x * CT) {
        const uint64_t b = b0 + threadIdx.x;
        const int c = b < nb ? (int)cls[b] : -1;
        uint64_t mine = 0;
#pragma unroll
        for (int q = 0; q < SORT_CLASSES; q++) {
            const uint64_t m = __ballot(c == q);
            if (c == q) mine = m;
            if (lane == 0) wofs[wave][q] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (threadIdx.x < SORT_CLASSES) {
            const int q = (int)threadIdx.x;
            uint32_t run = 0;
            for (int w = 0; w < NWV; w++) {
                const uint32_t t = wofs[w][q];
                wofs[w][q] = run;
                run += t;
            }
            base[q] = run ? atomicAdd(&pl->cls_cur[q], run) : 0u;
        }
        __syncthreads();
        if (c >= 0)
            order[pl->cls_off[c] + base[c] + wofs[wave][c] + (uint32_t)__popcll(mine & lanemask_lt())] = (uint32_t)b;
        __syncthreads();
    }
}

// The buckets of class C: hash-then-bin, handing skewed buckets to the
// class's fallback list.
// Persistent workgroups take buckets from a cursor (the first gridDim.x by
// their index); the next bucket is claimed when a bucket starts, so the
// atomic's round trip overlaps the sort.
template <class F>
__device__ __forceinline__ void for_each_claimed(uint32_t cnt, uint32_t *cursor, F &&body) {
    __shared__ uint32_t s_next;
    uint32_t j = blockIdx.x;
    while (j < cnt) {
        uint32_t nj = 0;
        if (threadIdx.x == 0) nj = gridDim.x + atomicAdd(cursor, 1u);
        body(j);
        if (threadIdx.x == 0) s_next = nj;
        __syncthreads();
        j = s_next;
        __syncthreads();
    }
}

template <int W, int IPT, int NT, int C>
__global__ __launch_bounds__(NT, hash_sort_waves(IPT)) void k_bucket_hash_sort_dev(
    uint64_t *__restrict__ A, uint64_t *__restrict__ B, uint32_t *__restrict__ counts_tmp,
    const Bucket *__restrict__ buckets, const uint32_t *__restrict__ order, uint32_t *__restrict__ ucount, int k,
    uint32_t *__restrict__ fb, DevPlan *__restrict__ pl) {
    if (pl->overflow) return;
    const uint32_t cnt = pl->cls_n[C], first = pl->cls_off[C];
    for_each_claimed(cnt, &pl->sort_cur[C], [&](uint32_t j) {
        hash_sort_bucket<W, IPT, NT>(A, B, counts_tmp, buckets, order[first + j], ucount, k, fb + first,
                                     &pl->fb_n[C]);
    });
}

// The buckets of class C by LSD passes: the whole class (FB false) or the
// buckets its hash pass handed back (FB true).
template <int W, int IPT, int C, bool FB>
__global__ __launch_bounds__(BT) void k_bucket_sort_dev(uint64_t *__restrict__ A, uint64_t *__restrict__ B,
                                                         uint32_t *__restrict__ counts_tmp,
                                                         const Bucket *__restrict__ buckets,
                                                         const uint32_t *__restrict__ list,
                                                         uint32_t *__restrict__ ucount, int k,
                                                         DevPlan *__restrict__ pl) {
    if (pl->overflow) return;
    const uint32_t cnt = FB ? pl->fb_n[C] : pl->cls_n[C], first = pl->cls_off[C];
    for_each_claimed(cnt, &pl->sort_cur[SORT_CLASSES * (FB ? 2 : 1) + C], [&](uint32_t j) {
        lsd_sort_bucket<W, IPT>(A, B, counts_tmp, buckets, list[first + j], ucount, k);
    });
}

__global__ void k_dev_unique(DevPlan *__restrict__ pl, const uint64_t *__restrict__ uoff) {
    if (pl->overflow) return;
    pl->n_unique = uoff[pl->lv[DEV_LEVELS].nn];
}

template <int W>
__global__ __launch_bounds__(256) void k_compact_dev(const uint64_t *__restrict__ A, const uint64_t *__restrict__ B,
                                                     const uint32_t *__restrict__ counts_tmp,
                                                     const Bucket *__restrict__ buckets,
                                                     const uint64_t *__restrict__ uoff, const DevPlan *__restrict__ pl,
                                                     uint64_t *__restrict__ out_keys,
                                                     uint32_t *__restrict__ out_counts) {
    if (pl->overflow) return;
    // one bucket per wave: a fixed grid keeps many buckets' row copies in
    // flight (a workgroup per bucket in turn would wait on each bucket's
    // offsets before its copy)
    const uint64_t nb = pl->lv[DEV_LEVELS].nn;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t waves = (uint64_t)gridDim.x * (256 / 64);
    for (uint64_t b = (uint64_t)blockIdx.x * (256 / 64) + threadIdx.x / 64; b < nb; b += waves) {
        const Bucket bk = buckets[b];
        const uint64_t *X = bk.array ? B : A;
        const uint64_t o = uoff[b];
        const uint64_t u = uoff[b + 1] - o;
        for (uint64_t i = lane; i < u; i += 64) {
#pragma unroll
            for (int r = 0; r < W; r++) out_keys[(o + i) * W + r] = X[(bk.start + i) * W + r];
            out_counts[o + i] = counts_tmp[bk.start + i];
        }
    }
}

__global__ void k_dev_part_rows(const uint64_t *__restrict__ uoff, const uint32_t *__restrict__ pfirst, uint32_t P,
                                const DevPlan *__restrict__ pl, uint64_t *__restrict__ pu) {
    if (pl->overflow) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= P) pu[p] = uoff[pfirst[p]];
}

// The summary and the partition tables to page-locked host memory.
__global__ void k_dev_summary(const DevPlan *__restrict__ pl, uint32_t P, const uint64_t *__restrict__ pu,
                              const uint64_t *__restrict__ byte_off, const uint64_t *__restrict__ minmax,
                              DevSummary *__restrict__ h) {
    uint64_t *tab = reinterpret_cast<uint64_t *>(h + 1);
    const bool ok = pl->overflow == 0;
    KITE_GRID_STRIDE(i, 2ull * (P + 1) + 8ull * P) {
        uint64_t v = 0;
        if (ok) v = i <= P ? pu[i] : (i <= 2ull * P + 1 ? byte_off[i - (P + 1)] : minmax[i - 2ull * (P + 1)]);
        tab[i] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        DevSummary sm{};
        sm.len = pl->len;
        sm.overflow = pl->overflow;
        if (ok) {
            sm.n_kmers = pl->n;
            sm.n_unique = pl->n_unique;
            sm.bytes = pl->bytes;
            sm.buckets = pl->lv[DEV_LEVELS].nn;
            uint32_t lv = 0;
            for (int l = 0; l < DEV_LEVELS; l++) lv += pl->lv[l].nact ? 1u : 0u;
            sm.levels = lv;
            sm.fallback = (uint64_t)pl->fb_n[0] + pl->fb_n[1] + pl->fb_n[2] + pl->fb_n[3];
            sm.nwg = pl->nwg;
            for (int l = 0; l < DEV_LEVELS; l++) sm.split_keys += pl->lv[l].keys;
        }
        *h = sm;
    }
}

// ---- host orchestration -----------------------------------------------------
namespace {
struct EventTimer {
    bool on;
    hipStream_t s;
    hipEvent_t e[8];
    int n = 0;
    EventTimer(bool on_, hipStream_t s_) : on(on_), s(s_) {
        if (on)
            for (auto &x : e) KITE_HIP(hipEventCreate(&x));
    }
    ~EventTimer() {
        if (on)
            for (auto &x : e) (void)hipEventDestroy(x);
    }
    void mark() {
        if (on && n < 8) KITE_HIP(hipEventRecord(e[n++], s));
    }
    float ms(int a, int b) {
        float t = 0;
        if (on && b < n) KITE_HIP(hipEventElapsedTime(&t, e[a], e[b]));
        return t;
    }
};

int class_cap(int W, int c) { return BT * ((4 << c) / W); }
int sort_cap(int W) { return class_cap(W, SORT_CLASSES - 1); }
int split_tk(int W) { return W == 1 ? SplitCfg<1>::TK : (W == 2 ? SplitCfg<2>::TK : SplitCfg<4>::TK); }
}  // namespace

template <class F>
void CountEngine::launch(const char *name, uint64_t bytes, F &&f) {
    if (!kprof_) {
        f();
        KITE_CHECK_LAUNCH();
        return;
    }
    int idx = -1;
    for (size_t i = 0; i < kstats_.size(); i++)
        if (kstats_[i].name == name) idx = (int)i;
    if (idx < 0) {
        kstats_.push_back(KernelStat{name, 0, 0, 0});
        idx = (int)kstats_.size() - 1;
    }
    kstats_[idx].launches++;
    kstats_[idx].bytes += bytes;
    auto take = [&]() {
        hipEvent_t e;
        if (!ev_pool_.empty()) {
            e = ev_pool_.back();
            ev_pool_.pop_back();
        } else {
            KITE_HIP(hipEventCreate(&e));
        }
        return e;
    };
    hipEvent_t a = take(), b = take();
    KITE_HIP(hipEventRecord(a, stream_));
    f();
    KITE_CHECK_LAUNCH();
    KITE_HIP(hipEventRecord(b, stream_));
    pending_.push_back(PendingEv{idx, a, b});
}

void CountEngine::begin_kernel_stats() {
    kstats_.clear();
    for (auto &p : pending_) {
        ev_pool_.push_back(p.a);
        ev_pool_.push_back(p.b);
    }
    pending_.clear();
}

void CountEngine::collect_kernel_stats() {
    for (auto &p : pending_) {
        float ms = 0;
        KITE_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        kstats_[p.stat].ms += ms;
        ev_pool_.push_back(p.a);
        ev_pool_.push_back(p.b);
    }
    pending_.clear();
}

void CountEngine::release() {
    KITE_HIP(hipStreamSynchronize(stream_));
    for (int q = 0; q < OUT_SLOTS; q++) {
        out_keys_[q].release();
        out_counts_[q].release();
    }
    for (DevBuf *b : {&nl_, &line0_, &scan_tmp_, &hist_, &off_, &part_off_d_, &a_, &b_, &counts_tmp_, &segs_,
                      &tile_pre_, &child_hist_, &child_cur_,
                      &ucount_, &uoff_, &part_first_, &part_u_, &order_, &wg_hist_, &nodes_[0], &nodes_[1],
                      &npart_[0], &npart_[1], &act_, &tiles_, &act_rank_, &tile_off_, &lstats_, &seg_node_,
                      &nz_rank_, &nz_count_, &new_cnt_, &new_base_, &cls_, &cls_count_, &fallback_})
        b->release();
}

uint64_t *CountEngine::pinned_host(size_t bytes) {
    if (bytes > pin_bytes_) {
        if (pin_) {
            KITE_HIP(hipStreamSynchronize(stream_));
            KITE_HIP(hipHostFree(pin_));
            pin_ = nullptr;
        }
        KITE_HIP(hipHostMalloc(&pin_, bytes, hipHostMallocDefault));
        pin_bytes_ = bytes;
    }
    return static_cast<uint64_t *>(pin_);
}

CountEngine::~CountEngine() {
    (void)hipStreamSynchronize(stream_);
    for (auto &p : pending_) {
        ev_pool_.push_back(p.a);
        ev_pool_.push_back(p.b);
    }
    for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
    if (pin_) (void)hipHostFree(pin_);
}

uint64_t CountEngine::bytes_needed(uint64_t n, int words, uint64_t fastq, uint32_t partitions) {
    uint64_t kb = 8ull * (uint64_t)words;
    uint64_t n_sub = ceil_div(fastq, SUB);
    uint64_t nwg = n_sub;  // one sub-tile per workgroup: the largest matrix
    uint64_t mat = (uint64_t)partitions * nwg;
    return n * (3 * kb + 8) + mat * 12 + n_sub * 12 + (n / 256 + 1) * 64 + (64ull << 20);
}

// The settings of the next count: the base settings, overridden by the
// environment.
CountTunables CountEngine::resolved_tunables() const {
    CountTunables t = base_;
    if (const char *e = getenv("KITE_SPLIT_DIV")) t.split_target_div = std::max(1, atoi(e));
    if (const char *e = getenv("KITE_SORT_DEDUP")) t.dedup = atoi(e) != 0;
    if (const char *e = getenv("KITE_MAX_DBITS")) t.max_digit_bits = std::min(MAX_DBITS, std::max(1, atoi(e)));
    if (const char *e = getenv("KITE_HASH_CLASSES")) t.hash_classes = std::min(SORT_CLASSES, std::max(0, atoi(e)));
    if (const char *e = getenv("KITE_SPLIT_CHUNK")) t.split_chunk = std::min(64, std::max(1, atoi(e)));
    if (const char *e = getenv("KITE_EXTRACT_SUBS")) t.extract_subs = std::min(64, std::max(1, atoi(e)));
    if (const char *e = getenv("KITE_EXTRACT_STAGED")) t.extract_staged = atoi(e) != 0;
    if (const char *e = getenv("KITE_SPLIT_STAGED")) t.split_staged = atoi(e) != 0 ? 1 : 0;
    if (const char *e = getenv("KITE_SPLIT_ABOVE_DIV")) t.split_above_div = std::min(4, std::max(1, atoi(e)));
    return t;
}

void CountEngine::reserve(uint64_t max_kmers, int words) {
    const uint64_t n = std::max<uint64_t>(max_kmers, 1);
    a_.ensure(n * 8 * words);
    b_.ensure(n * 8 * words);
    counts_tmp_.ensure(n * 4);
}

void CountEngine::run(const uint8_t *d_fq, uint64_t len, int k, int canonical, uint32_t partitions) {
    if (k < 1 || k > 128) throw Error("invalid k-mer length: k must be between 1 and 128");
    if (partitions == 0 || (partitions & (partitions - 1)))
        throw Error("invalid partition count: must be a power of 2");
    if (partitions > MAX_PARTITIONS)
        throw Error("partitions above " + std::to_string(MAX_PARTITIONS) +
                    " are not supported by the GPU count engine");
    int pbits = __builtin_ctz(partitions);
    words_ = k <= 31 ? 1 : (k <= 64 ? 2 : 4);
    tun = resolved_tunables();
    if (words_ == 1) run_w<1>(d_fq, len, k, canonical, pbits);
    else if (words_ == 2) run_w<2>(d_fq, len, k, canonical, pbits);
    else run_w<4>(d_fq, len, k, canonical, pbits);
}

template <int W>
void CountEngine::run_w(const uint8_t *d_fq, uint64_t len, int k, int canonical, int pbits) {
    hipStream_t s = stream_;
    const uint32_t P = 1u << pbits;
    const int CAP = sort_cap(W);
    // measured: one-word keys do best with children up to half the largest
    // capacity, all of them hashed (the largest class on 1024 threads); wider
    // keys with a quarter and the LSD kernel for the largest class
    const int split_div = tun.split_target_div > 0 ? tun.split_target_div : (W == 1 ? 2 : 4);
    const int hash_classes = tun.hash_classes >= 0 ? tun.hash_classes : (W == 1 ? SORT_CLASSES : 3);
    // the regrouped split scatter pays for one-word keys (measured slower
    // for two-word keys, whose tiles give half as long runs per child)
    const bool split_staged = tun.split_staged >= 0 ? tun.split_staged != 0 : W == 1;
    EventTimer tm(timing_, s);
    tm.mark();
    part_off_h_.assign(P + 1, 0);
    n_kmers_ = n_unique_ = n_buckets_ = n_fallback_ = 0;
    kstats_.clear();
    for (auto &p : pending_) {  // events of a run whose stats were not collected
        ev_pool_.push_back(p.a);
        ev_pool_.push_back(p.b);
    }
    pending_.clear();
    levels_ = 0;

    const uint64_t n_sub = std::max<uint64_t>(1, ceil_div(len, SUB));
    const uint32_t subs = (uint32_t)std::max(1, tun.extract_subs);
    const uint32_t nwg = (uint32_t)ceil_div(n_sub, subs);
    const uint64_t mat = (uint64_t)P * nwg;

    // page-locked host scratch for small read-backs
    uint64_t *pin = pinned_host(((uint64_t)P + 1) * 8 + 256);
    auto to_host = [&](auto *dev, uint32_t n, auto *host) {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(dev)>>;
        k_put_host<T><<<dim3((n + 255) / 256), dim3(256), 0, s>>>(dev, n, reinterpret_cast<T *>(host));
        KITE_CHECK_LAUNCH();
    };

    // 1. line index at each sub-tile start
    uint32_t *nl = static_cast<uint32_t *>(nl_.ensure(n_sub * 4));
    uint64_t *line0 = static_cast<uint64_t *>(line0_.ensure((n_sub + 1) * 8));
    uint64_t scan_elems = std::max<uint64_t>(scan_temp_elems(mat + 1), scan_temp_elems(n_sub + 1));
    uint64_t *stmp = static_cast<uint64_t *>(scan_tmp_.ensure(scan_elems * 8));
    if (len > 0) {
        launch("k_nl_count", len, [&] { k_nl_count<<<dim3((unsigned)n_sub), dim3(XT), 0, s>>>(d_fq, len, nl); });
    } else {
        k_fill_u32<<<1, 64, 0, s>>>(nl, 1, 0);
        KITE_CHECK_LAUNCH();
    }
    scan_to_u64(nl, line0, n_sub, stmp, s);

    // 2. histogram of partitions per extract workgroup
    ExtractArgs ea{};
    ea.fq = d_fq;
    ea.len = len;
    ea.sub_line0 = line0;
    ea.n_sub = n_sub;
    ea.nwg = nwg;
    ea.subs = subs;
    ea.k = k;
    ea.canonical = canonical;
    ea.pbits = pbits;
    ea.hist = static_cast<uint32_t *>(hist_.ensure(mat * 4));
    size_t dyn_hist = (size_t)P * 4;
    size_t dyn_scat = (size_t)((P + 1) & ~1u) * 4 + (size_t)P * 8;
    // staged scatter when a chunk gives runs of two or more keys per partition
    const bool staged = tun.extract_staged && 2 * P <= (uint32_t)StageCfg<W>::SC;
    if (staged) dyn_scat += (size_t)P * 8;
    {
        static bool attr_set[3] = {false, false, false};
        const int wi = W == 1 ? 0 : (W == 2 ? 1 : 2);
        if (!attr_set[wi]) {
            const int max_dyn = (int)(((MAX_PARTITIONS + 1) & ~1u) * 4 + MAX_PARTITIONS * 8);
            KITE_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_extract<W, 0>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn));
            KITE_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_extract<W, 1>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn));
            attr_set[wi] = true;
        }
    }
    launch("k_extract_hist", len + mat * 4,
           [&] { k_extract<W, 0><<<dim3(nwg), dim3(XT), dyn_hist, s>>>(ea); });
    uint64_t *off = static_cast<uint64_t *>(off_.ensure((mat + 1) * 8));
    scan_to_u64(ea.hist, off, mat, stmp, s);
    uint64_t *part_off_d = static_cast<uint64_t *>(part_off_d_.ensure((P + 1) * 8));
    k_part_gather<<<dim3((P + 1 + 255) / 256), dim3(256), 0, s>>>(off, nwg, P, part_off_d);
    KITE_CHECK_LAUNCH();
    to_host(part_off_d, P + 1, pin);
    KITE_HIP(hipStreamSynchronize(s));
    std::copy(pin, pin + P + 1, part_off_h_.begin());
    const uint64_t n = part_off_h_[P];
    n_kmers_ = n;
    // the one-tile split scatter keeps 32-bit offsets inside a segment, and
    // a segment never outgrows its partition
    uint64_t max_part = 0;
    for (uint32_t p = 0; p < P; p++) max_part = std::max(max_part, part_off_h_[p + 1] - part_off_h_[p]);
    const uint32_t spl = (uint32_t)std::max(max_part > 0xFFFFFFFFull ? 2 : 1, tun.split_chunk);
    const uint64_t TK = (uint64_t)split_tk(W) * spl;  // keys per split workgroup

    // 3. scatter keys into partition order
    uint64_t *A = static_cast<uint64_t *>(a_.ensure(std::max<uint64_t>(n, 1) * 8 * W));
    uint64_t *B = static_cast<uint64_t *>(b_.ensure(std::max<uint64_t>(n, 1) * 8 * W));
    uint32_t *ctmp = static_cast<uint32_t *>(counts_tmp_.ensure(std::max<uint64_t>(n, 1) * 4));
    if (n > 0) {
        ea.off = off;
        ea.out = A;
        launch("k_extract_scatter", len + mat * 8 + n * 8 * W, [&] {
            if (staged) k_extract<W, 2><<<dim3(nwg), dim3(XT), dyn_scat, s>>>(ea);
            else k_extract<W, 1><<<dim3(nwg), dim3(XT), dyn_scat, s>>>(ea);
        });
    }
    tm.mark();

    // 4. MSD splits until every bucket fits one LDS sort (or is exhausted);
    // the node list is built on the device
    const int kbits = 2 * k;
    auto grid1 = [](uint64_t n) { return dim3((unsigned)std::max<uint64_t>(1, ceil_div(n, 256))); };
    uint32_t nn = P;
    int cur_list = 0;
    {
        Bucket *nodes = static_cast<Bucket *>(nodes_[0].ensure((uint64_t)P * sizeof(Bucket)));
        uint32_t *npart = static_cast<uint32_t *>(npart_[0].ensure((uint64_t)P * 4));
        k_init_nodes<<<grid1(P), 256, 0, s>>>(part_off_d, P, nodes, npart);
        KITE_CHECK_LAUNCH();
    }
    int prefix = 0;
    uint32_t src_arr = 0;
    uint64_t *hv = pin + P + 1;  // small read-backs (page-locked host memory)
    for (;;) {
        Bucket *nodes = nodes_[cur_list].as<Bucket>();
        uint32_t *npart = npart_[cur_list].as<uint32_t>();
        uint32_t *act = static_cast<uint32_t *>(act_.ensure((uint64_t)nn * 4));
        uint64_t *tiles = static_cast<uint64_t *>(tiles_.ensure((uint64_t)nn * 8));
        uint64_t *act_rank = static_cast<uint64_t *>(act_rank_.ensure(((uint64_t)nn + 1) * 8));
        uint64_t *tile_off = static_cast<uint64_t *>(tile_off_.ensure(((uint64_t)nn + 1) * 8));
        unsigned long long *lstats = static_cast<unsigned long long *>(lstats_.ensure(16));
        uint64_t *stmp = static_cast<uint64_t *>(scan_tmp_.ensure(
            std::max<uint64_t>(scan_elems, scan_temp_elems((uint64_t)nn + 1)) * 8));
        k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(lstats), 4, 0);
        KITE_CHECK_LAUNCH();
        k_level_flags<<<grid1(nn), 256, 0, s>>>(nodes, nn, (uint64_t)CAP / (uint64_t)std::max(1, tun.split_above_div),
                                                kbits, TK, act, tiles, lstats);
        KITE_CHECK_LAUNCH();
        scan_to_u64(act, act_rank, nn, stmp, s);
        scan_to_u64(tiles, tile_off, nn, stmp, s);
        to_host(act_rank + nn, 1, hv + 0);
        to_host(tile_off + nn, 1, hv + 1);
        to_host(lstats, 2, reinterpret_cast<unsigned long long *>(hv + 2));
        KITE_HIP(hipStreamSynchronize(s));
        const uint64_t nact = hv[0], tiles_total = hv[1], maxsz = hv[2], split_keys = hv[3];
        if (nact == 0) break;
        int d = 1;
        while (d < tun.max_digit_bits && (maxsz >> d) > (uint64_t)(CAP / split_div)) d++;
        d = std::min(d, kbits - prefix);
        const uint32_t nd = 1u << d;
        const uint64_t nchild = nact * nd;
        Seg *segs = static_cast<Seg *>(segs_.ensure(nact * sizeof(Seg)));
        uint64_t *tpre = static_cast<uint64_t *>(tile_pre_.ensure((nact + 1) * 8));
        uint32_t *seg_node = static_cast<uint32_t *>(seg_node_.ensure(nact * 4));
        k_level_segs<<<grid1(nn), 256, 0, s>>>(nodes, nn, act, act_rank, tile_off, nd, segs, tpre, seg_node);
        KITE_CHECK_LAUNCH();
        uint32_t *ch = static_cast<uint32_t *>(child_hist_.ensure(nchild * 4));
        k_fill_u32<<<grid1(nchild), 256, 0, s>>>(ch, nchild, 0);
        KITE_CHECK_LAUNCH();
        SplitArgs sa{};
        sa.src = src_arr ? B : A;
        sa.dst = src_arr ? A : B;
        sa.segs = segs;
        sa.nseg = (uint32_t)nact;
        sa.tile_pre = tpre;
        sa.shift = kbits - prefix - d;
        sa.dbits = d;
        const uint32_t chunk = spl;
        sa.chunk = chunk;
        sa.child_hist = ch;
        sa.wg_hist = chunk > 1 ? static_cast<uint32_t *>(wg_hist_.ensure(tiles_total * nd * 4)) : nullptr;
        launch("k_split_hist", split_keys * 8 * W,
               [&] { k_split<W, 0><<<dim3((unsigned)tiles_total), dim3(ST), 0, s>>>(sa); });
        uint64_t *ccur = static_cast<uint64_t *>(child_cur_.ensure(nchild * 8));
        uint32_t *nzr = static_cast<uint32_t *>(nz_rank_.ensure(nchild * 4));
        uint32_t *nzc = static_cast<uint32_t *>(nz_count_.ensure(nact * 4));
        k_child_scan<<<dim3((unsigned)nact), 256, 0, s>>>(segs, ch, nd, ccur, nzr, nzc);
        KITE_CHECK_LAUNCH();
        uint32_t *ncnt = static_cast<uint32_t *>(new_cnt_.ensure((uint64_t)nn * 4));
        uint64_t *nbase = static_cast<uint64_t *>(new_base_.ensure(((uint64_t)nn + 1) * 8));
        k_level_counts<<<grid1(nn), 256, 0, s>>>(nn, act, act_rank, nzc, ncnt);
        KITE_CHECK_LAUNCH();
        scan_to_u64(ncnt, nbase, nn, stmp, s);
        to_host(nbase + nn, 1, hv + 4);
        KITE_HIP(hipStreamSynchronize(s));
        const uint64_t nn_next = hv[4];
        if (nn_next > 0xFFFFFFFFull) throw Error("internal: bucket list above 2^32 entries");
        const int nxt = cur_list ^ 1;
        Bucket *out_nodes = static_cast<Bucket *>(nodes_[nxt].ensure(std::max<uint64_t>(nn_next, 1) * sizeof(Bucket)));
        uint32_t *out_part = static_cast<uint32_t *>(npart_[nxt].ensure(std::max<uint64_t>(nn_next, 1) * 4));
        k_level_emit_nodes<<<grid1(nn), 256, 0, s>>>(nodes, npart, nn, act, nbase, out_nodes, out_part);
        KITE_CHECK_LAUNCH();
        k_level_emit_children<<<grid1(nchild), 256, 0, s>>>(seg_node, nbase, nodes, npart, ch, ccur, nzr, nchild, nd, d,
                                                            out_nodes, out_part);
        KITE_CHECK_LAUNCH();
        sa.child_cur = ccur;
        launch("k_split_scatter", split_keys * 16 * W,
               [&] {
                   if (chunk == 1 && split_staged && d <= SPLIT_STAGED_BITS)
                       k_split_staged<W><<<dim3((unsigned)tiles_total), dim3(SST), 0, s>>>(sa);
                   else if (chunk == 1) k_split<W, 1><<<dim3((unsigned)tiles_total), dim3(ST), 0, s>>>(sa);
                   else k_split<W, 2><<<dim3((unsigned)tiles_total), dim3(ST), 0, s>>>(sa);
               });
        levels_++;
        nn = (uint32_t)nn_next;
        cur_list = nxt;
        prefix += d;
        src_arr ^= 1u;
    }
    tm.mark();

    // 5. sort + run-length count every bucket
    const uint64_t nb = nn;
    n_buckets_ = nb;
    Bucket *bkts = nodes_[cur_list].as<Bucket>();
    uint32_t *npart_fin = npart_[cur_list].as<uint32_t>();
    uint64_t *uoff = static_cast<uint64_t *>(uoff_.ensure((nb + 1) * 8));
    uint32_t *ucount = static_cast<uint32_t *>(ucount_.ensure(std::max<uint64_t>(nb, 1) * 4));
    uint32_t *pfirst = static_cast<uint32_t *>(part_first_.ensure((uint64_t)(P + 1) * 4));
    k_part_first<<<grid1(P + 1), 256, 0, s>>>(npart_fin, (uint32_t)nb, P, pfirst);
    KITE_CHECK_LAUNCH();
    if (nb > 0) {
        // size classes: the smallest LDS footprint that holds each bucket
        uint32_t *cls = static_cast<uint32_t *>(cls_.ensure(nb * 4));
        uint32_t *cls_count = static_cast<uint32_t *>(cls_count_.ensure(4 * (SORT_CLASSES + 1) * 2));
        k_fill_u32<<<1, 64, 0, s>>>(cls_count, 2 * (SORT_CLASSES + 1), 0);
        KITE_CHECK_LAUNCH();
        ClassCaps caps;
        for (int c = 0; c < SORT_CLASSES; c++) caps.cap[c] = (uint64_t)class_cap(W, c);
        k_classify<<<dim3((unsigned)ceil_div(nb, CT)), CT, 0, s>>>(bkts, (uint32_t)nb, kbits, caps, cls, cls_count);
        KITE_CHECK_LAUNCH();
        uint32_t *hcls = reinterpret_cast<uint32_t *>(hv + 5);
        to_host(cls_count, SORT_CLASSES, hcls);
        uint32_t *dcls_off = cls_count + (SORT_CLASSES + 1);
        k_class_offsets<<<1, 1, 0, s>>>(cls_count, dcls_off);
        KITE_CHECK_LAUNCH();
        KITE_HIP(hipStreamSynchronize(s));
        uint32_t cls_off[SORT_CLASSES + 1] = {0};
        for (int c = 0; c < SORT_CLASSES; c++) cls_off[c + 1] = cls_off[c] + hcls[c];
        uint32_t *ordp = static_cast<uint32_t *>(order_.ensure(nb * 4));
        k_order_fill<<<dim3((unsigned)ceil_div(nb, CT)), CT, 0, s>>>(cls, (uint32_t)nb, dcls_off, cls_count, ordp);
        KITE_CHECK_LAUNCH();
        const uint32_t *ord = order_.as<uint32_t>();
        // buckets the hash kernel hands back (a bin above its local limit),
        // per class in [cls_off[c], ...), counted in fbc[c]
        uint32_t *fb = static_cast<uint32_t *>(fallback_.ensure(nb * 4 + 16));
        uint32_t *fbc = fb + nb;
        k_fill_u32<<<1, 64, 0, s>>>(fbc, 4, 0);
        KITE_CHECK_LAUNCH();
        bool hashed = false;
        auto go = [&](auto cls, uint32_t first, uint32_t cnt) {
            constexpr int C = decltype(cls)::value;
            if (!cnt) return;
            // hash kernel: 4 / W items per thread, so a bigger class gets more
            // threads (up to 1024) rather than more items per thread; the
            // largest class takes 8 / W items on 1024 threads
            {
                if (tun.dedup && C < hash_classes) {
                    constexpr int IPTH = (C < 3 ? 4 : 8) / W > 0 ? (C < 3 ? 4 : 8) / W : 1;
                    constexpr int NTH = SortClass<W, C>::CAP / IPTH;
                    static_assert(NTH <= 1024, "one workgroup per bucket");
                    launch("k_bucket_sort", 0, [&] {
                        k_bucket_hash_sort<W, IPTH, NTH><<<dim3(cnt), dim3(NTH), 0, s>>>(
                            A, B, ctmp, bkts, ord + first, ucount, k, fb + first, fbc + C);
                    });
                    hashed = true;
                    return;
                }
            }
            launch("k_bucket_sort", 0, [&] {
                k_bucket_sort<W, SortClass<W, C>::IPT><<<dim3(cnt), dim3(BT), 0, s>>>(
                    A, B, ctmp, bkts, ord + first, ucount, k);
            });
        };
        go(std::integral_constant<int, 0>{}, cls_off[0], cls_off[1] - cls_off[0]);
        go(std::integral_constant<int, 1>{}, cls_off[1], cls_off[2] - cls_off[1]);
        go(std::integral_constant<int, 2>{}, cls_off[2], cls_off[3] - cls_off[2]);
        go(std::integral_constant<int, 3>{}, cls_off[3], cls_off[4] - cls_off[3]);
        if (hashed) {
            // the handed-back buckets of each class through the LSD kernel
            uint32_t *hfb = reinterpret_cast<uint32_t *>(hv + 10);
            to_host(fbc, SORT_CLASSES, hfb);
            KITE_HIP(hipStreamSynchronize(s));
            const uint32_t nfb[SORT_CLASSES] = {hfb[0], hfb[1], hfb[2], hfb[3]};
            n_fallback_ = nfb[0] + nfb[1] + nfb[2] + nfb[3];
            auto lsd = [&](auto cls, uint32_t first, uint32_t cnt) {
                constexpr int C = decltype(cls)::value;
                if (!cnt) return;
                launch("k_bucket_sort", 0, [&] {
                    k_bucket_sort<W, SortClass<W, C>::IPT><<<dim3(cnt), dim3(BT), 0, s>>>(A, B, ctmp, bkts, fb + first,
                                                                                         ucount, k);
                });
            };
            lsd(std::integral_constant<int, 0>{}, cls_off[0], nfb[0]);
            lsd(std::integral_constant<int, 1>{}, cls_off[1], nfb[1]);
            lsd(std::integral_constant<int, 2>{}, cls_off[2], nfb[2]);
            lsd(std::integral_constant<int, 3>{}, cls_off[3], nfb[3]);
        }
    }
    tm.mark();
    uint64_t *stmp2 = static_cast<uint64_t *>(scan_tmp_.ensure(
        std::max<uint64_t>(scan_elems, scan_temp_elems(nb + 1)) * 8));
    scan_to_u64(ucount, uoff, nb, stmp2, s);
    to_host(uoff + nb, 1, hv + 8);
    KITE_HIP(hipStreamSynchronize(s));
    n_unique_ = hv[8];
    uint64_t *ok = static_cast<uint64_t *>(out_keys_[slot_].ensure(std::max<uint64_t>(n_unique_, 1) * 8 * W));
    uint32_t *oc = static_cast<uint32_t *>(out_counts_[slot_].ensure(std::max<uint64_t>(n_unique_, 1) * 4));
    if (nb > 0) {
        launch("k_compact", n_unique_ * (8 * W + 4) * 2, [&] {
            k_compact<W><<<dim3((unsigned)nb), dim3(256), 0, s>>>(A, B, ctmp, bkts, uoff, ok, oc);
        });
    }
    // per-partition output offsets
    uint64_t *pu = static_cast<uint64_t *>(part_u_.ensure((P + 1) * 8));
    k_gather_u64<<<dim3((P + 1 + 255) / 256), dim3(256), 0, s>>>(uoff, pfirst, P + 1, pu);
    KITE_CHECK_LAUNCH();
    to_host(pu, P + 1, pin);
    tm.mark();
    KITE_HIP(hipStreamSynchronize(s));
    std::copy(pin, pin + P + 1, part_off_h_.begin());
    if (kprof_) {
        collect_kernel_stats();
        for (auto &ks : kstats_)  // sort reads every key once, writes the unique runs
            if (ks.name == "k_bucket_sort") ks.bytes = n * 8 * W + n_unique_ * (8 * W + 4);
    }
    if (timing_) {
        times_.extract_ms = tm.ms(0, 1);
        times_.split_ms = tm.ms(1, 2);
        times_.sort_ms = tm.ms(2, 3);
        times_.compact_ms = tm.ms(3, 4);
        times_.total_ms = tm.ms(0, 4);
    }
}

template void CountEngine::run_w<1>(const uint8_t *, uint64_t, int, int, int);
template void CountEngine::run_w<2>(const uint8_t *, uint64_t, int, int, int);
template void CountEngine::run_w<4>(const uint8_t *, uint64_t, int, int, int);

// ---- device-planned count: host side ---------------------------------------------

namespace {

// Workgroups of `f` resident on the whole device at once (a loop over work
// items spread over exactly these keeps every workgroup busy to the end).
int resident_grid(const void *f, int threads, size_t dyn_lds) {
    static std::map<std::tuple<const void *, int, size_t>, int> cache;
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        KITE_HIP(hipGetDevice(&dev));
        KITE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const auto key = std::make_tuple(f, threads, dyn_lds);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int per = 0;
    KITE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, threads, dyn_lds));
    const int g = std::max(1, per) * cus;
    cache[key] = g;
    return g;
}

dim3 capped_grid(uint64_t items, uint32_t per_wg, uint32_t cap) {
    return dim3((unsigned)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(std::max<uint64_t>(items, 1), per_wg), cap)));
}

}  // namespace

void CountEngine::dev_prepare(uint64_t cap_len, int k, uint32_t partitions, EncodeBufs &enc) {
    if (k < 1 || k > 128) throw Error("invalid k-mer length: k must be between 1 and 128");
    if (partitions == 0 || (partitions & (partitions - 1)) || partitions > MAX_PARTITIONS)
        throw Error("invalid partition count for the device-planned count");
    const uint32_t P = partitions;
    const int W = k <= 31 ? 1 : (k <= 64 ? 2 : 4);
    DevCaps c;
    c.len = cap_len;
    c.n_sub = std::max<uint64_t>(1, ceil_div(cap_len, SUB));
    c.nwg = c.n_sub;  // one sub-tile per workgroup: the widest matrix any setting makes
    // a k-mer ends on a sequence byte, and a record's quality line is as
    // long as its sequence line
    c.n = cap_len / 2 + 4096;
    c.nodes = (uint64_t)P + c.n / 256 + 65536;
    // (KITE_PLAN_NODES lowers the node capacity, so that tests reach the
    // classic path a planned count hands its input to)
    if (const char *e = getenv("KITE_PLAN_NODES"))
        c.nodes = std::max<uint64_t>((uint64_t)P + 1, std::min<uint64_t>(c.nodes, strtoull(e, nullptr, 10)));
    c.children = (uint64_t)std::max<uint32_t>(P, 2048) * (1u << MAX_DBITS);
    dcap_ = c;
    dP_ = P;
    dW_ = W;
    const uint64_t mat = (uint64_t)P * c.nwg, kb = 8ull * W;
    dv_.plan.ensure(sizeof(DevPlan));
    dv_.nl.ensure(c.n_sub * 4);
    dv_.line0.ensure((c.n_sub + 1) * 8);
    dv_.stmp.ensure(scan_temp_elems(std::max<uint64_t>(std::max<uint64_t>(mat, c.n_sub), c.nodes) + 1) * 8);
    dv_.hist.ensure(mat * 4);
    dv_.off.ensure((mat + 1) * 8);
    dv_.part_off.ensure(((uint64_t)P + 1) * 8);
    dv_.a.ensure(c.n * kb);
    dv_.b.ensure(c.n * kb);
    dv_.ctmp.ensure(c.n * 4);
    for (int q = 0; q < 2; q++) {
        dv_.nodes[q].ensure(c.nodes * sizeof(Bucket));
        dv_.npart[q].ensure(c.nodes * 4);
    }
    dv_.act.ensure(c.nodes * 4);
    dv_.tiles.ensure(c.nodes * 8);
    dv_.act_rank.ensure((c.nodes + 1) * 8);
    dv_.tile_off.ensure((c.nodes + 1) * 8);
    dv_.segs.ensure(c.nodes * sizeof(Seg));
    dv_.tpre.ensure((c.nodes + 1) * 8);
    dv_.seg_node.ensure(c.nodes * 4);
    dv_.nzc.ensure(c.nodes * 4);
    dv_.ch.ensure(c.children * 4);
    dv_.ccur.ensure(c.children * 8);
    dv_.nzr.ensure(c.children * 4);
    dv_.ncnt.ensure(c.nodes * 4);
    dv_.nbase.ensure((c.nodes + 1) * 8);
    dv_.cls.ensure(c.nodes * 4);
    dv_.order.ensure(c.nodes * 4);
    dv_.fb.ensure(c.nodes * 4);
    dv_.ucount.ensure(c.nodes * 4);
    dv_.uoff.ensure((c.nodes + 1) * 8);
    dv_.pfirst.ensure(((uint64_t)P + 1) * 4);
    dv_.pu.ensure(((uint64_t)P + 1) * 8);
    dv_.keys.ensure(c.n * kb);
    dv_.counts.ensure(c.n * 4);
    encode_prepare_dev(c.n, P, enc);
}

__global__ void k_store_u64(uint64_t *p, uint64_t v) { *p = v; }

void CountEngine::dev_set_len(uint64_t *d_len, uint64_t len) {
    k_store_u64<<<1, 1, 0, stream_>>>(d_len, len);
    KITE_CHECK_LAUNCH();
}

void CountEngine::dev_enqueue_count(const uint8_t *d_fq, const uint64_t *len, uint64_t max_len, int k, int canonical,
                                    uint32_t partitions) {
    const int W = k <= 31 ? 1 : (k <= 64 ? 2 : 4);
    if (partitions != dP_ || W != dW_) throw Error("internal: device-planned count prepared for another shape");
    if (max_len > dcap_.len) throw Error("internal: device-planned count above its prepared capacity");
    const int pbits = __builtin_ctz(partitions);
    if (W == 1) dev_enqueue_w<1>(d_fq, len, max_len, k, canonical, pbits);
    else if (W == 2) dev_enqueue_w<2>(d_fq, len, max_len, k, canonical, pbits);
    else dev_enqueue_w<4>(d_fq, len, max_len, k, canonical, pbits);
}

template <int W>
void CountEngine::dev_enqueue_w(const uint8_t *d_fq, const uint64_t *h_len, uint64_t max_len, int k, int canonical,
                                int pbits) {
    hipStream_t s = stream_;
    // buffers for the prepared capacity, grids for inputs up to max_len
    DevCaps c = dcap_;
    c.len = max_len;
    c.n_sub = std::max<uint64_t>(1, ceil_div(max_len, SUB));
    const uint32_t P = 1u << pbits;
    const int kbits = 2 * k;
    const CountTunables t = resolved_tunables();
    const int CAP = sort_cap(W);
    const int split_div = t.split_target_div > 0 ? t.split_target_div : (W == 1 ? 2 : 4);
    const int hash_classes = t.hash_classes >= 0 ? t.hash_classes : (W == 1 ? SORT_CLASSES : 3);
    const bool split_staged = t.split_staged >= 0 ? t.split_staged != 0 : W == 1;
    const uint32_t subs = (uint32_t)std::max(1, t.extract_subs);
    const uint64_t nwg = ceil_div(c.n_sub, subs);
    const uint64_t mat = (uint64_t)P * nwg;
    DevPlan *pl = dv_.plan.as<DevPlan>();
    uint64_t *stmp = dv_.stmp.as<uint64_t>();
    uint64_t *A = dv_.a.as<uint64_t>(), *B = dv_.b.as<uint64_t>();
    uint32_t *ctmp = dv_.ctmp.as<uint32_t>();
    auto chk = [] { KITE_CHECK_LAUNCH(); };

    // 1. plan, line index over the capacity (zero past the input)
    k_dev_plan_init<<<1, 256, 0, s>>>(h_len, c.len, P, subs, pl);
    chk();
    launch("k_nl_count", 0, [&] {
        k_nl_count_dev<<<capped_grid(c.n_sub, 1, 4096), XT, 0, s>>>(d_fq, pl, c.n_sub, dv_.nl.as<uint32_t>());
    });
    uint64_t *line0 = dv_.line0.as<uint64_t>();
    scan_to_u64(dv_.nl.as<uint32_t>(), line0, c.n_sub, stmp, s);

    // 2. partition histogram per extract workgroup, 3. scatter
    ExtractArgs ea{};
    ea.fq = d_fq;
    ea.sub_line0 = line0;
    ea.nwg = (uint32_t)nwg;
    ea.subs = subs;
    ea.k = k;
    ea.canonical = canonical;
    ea.pbits = pbits;
    ea.hist = dv_.hist.as<uint32_t>();
    ea.pl = pl;
    size_t dyn_hist = (size_t)P * 4;
    size_t dyn_scat = (size_t)((P + 1) & ~1u) * 4 + (size_t)P * 8;
    const bool staged = t.extract_staged && 2 * P <= (uint32_t)StageCfg<W>::SC;
    if (staged) dyn_scat += (size_t)P * 8;
    {
        const int max_dyn = (int)(((MAX_PARTITIONS + 1) & ~1u) * 4 + MAX_PARTITIONS * 8);
        KITE_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_extract<W, 0>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn));
        KITE_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_extract<W, 1>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn));
    }
    launch("k_extract_hist", 0, [&] { k_extract<W, 0><<<dim3((unsigned)nwg), dim3(XT), dyn_hist, s>>>(ea); });
    uint64_t *off = dv_.off.as<uint64_t>();
    scan_to_u64(ea.hist, off, mat, stmp, s);
    uint64_t *part_off = dv_.part_off.as<uint64_t>();
    k_part_gather<<<dim3((P + 1 + 255) / 256), dim3(256), 0, s>>>(off, (uint32_t)nwg, P, part_off);
    chk();
    k_dev_kmers<<<1, 1, 0, s>>>(part_off, P, c.n, pl);
    chk();
    ea.off = off;
    ea.out = A;
    launch("k_extract_scatter", 0, [&] {
        if (staged) k_extract<W, 2><<<dim3((unsigned)nwg), dim3(XT), dyn_scat, s>>>(ea);
        else k_extract<W, 1><<<dim3((unsigned)nwg), dim3(XT), dyn_scat, s>>>(ea);
    });

    // 4. MSD split levels, every one launched; a level with no node to split
    // copies the node list as it is
    Bucket *nodes[2] = {dv_.nodes[0].as<Bucket>(), dv_.nodes[1].as<Bucket>()};
    uint32_t *npart[2] = {dv_.npart[0].as<uint32_t>(), dv_.npart[1].as<uint32_t>()};
    k_dev_init_nodes<<<dim3((P + 255) / 256), 256, 0, s>>>(part_off, P, nodes[0], npart[0], pl);
    chk();
    uint32_t *act = dv_.act.as<uint32_t>(), *nzc = dv_.nzc.as<uint32_t>(), *ncnt = dv_.ncnt.as<uint32_t>();
    uint64_t *tiles = dv_.tiles.as<uint64_t>(), *act_rank = dv_.act_rank.as<uint64_t>();
    uint64_t *tile_off = dv_.tile_off.as<uint64_t>(), *tpre = dv_.tpre.as<uint64_t>(), *nbase = dv_.nbase.as<uint64_t>();
    Seg *segs = dv_.segs.as<Seg>();
    uint32_t *seg_node = dv_.seg_node.as<uint32_t>(), *ch = dv_.ch.as<uint32_t>(), *nzr = dv_.nzr.as<uint32_t>();
    uint64_t *ccur = dv_.ccur.as<uint64_t>();
    const uint64_t tk = (uint64_t)split_tk(W);
    const uint64_t cap_above = (uint64_t)CAP / (uint64_t)std::max(1, t.split_above_div);
    const dim3 gnode = capped_grid(c.nodes, 256, 2048), gchild = capped_grid(c.children, 256, 4096);
    const int g_hist = resident_grid(reinterpret_cast<const void *>(&k_split_dev<W, 0>), ST, 0);
    const int g_scat = resident_grid(reinterpret_cast<const void *>(&k_split_dev<W, 1>), ST, 0);
    const int g_stag = resident_grid(reinterpret_cast<const void *>(&k_split_staged_dev<W>), SST, 0);
    const int g_cscan = resident_grid(reinterpret_cast<const void *>(&k_dev_child_scan), 256, 0);
    for (int l = 0; l < DEV_LEVELS; l++) {
        const int cur = l & 1, nxt = cur ^ 1;
        k_dev_level_flags<<<gnode, 256, 0, s>>>(nodes[cur], pl, l, cap_above, kbits, tk, act, tiles);
        chk();
        scan_to_u64_dev(act, act_rank, &pl->lv[l].nn, c.nodes, stmp, s);
        scan_to_u64_dev(tiles, tile_off, &pl->lv[l].nn, c.nodes, stmp, s);
        k_dev_level_plan<<<1, 1, 0, s>>>(pl, l, act_rank, tile_off, kbits, t.max_digit_bits,
                                         (uint64_t)(CAP / split_div), split_staged ? 1 : 0, c.children);
        chk();
        k_dev_level_segs<<<gnode, 256, 0, s>>>(nodes[cur], pl, l, act, act_rank, tile_off, segs, tpre, seg_node);
        chk();
        k_dev_zero_children<<<gchild, 256, 0, s>>>(ch, pl, l);
        chk();
        SplitArgs sa{};
        sa.src = cur ? B : A;
        sa.dst = cur ? A : B;
        sa.segs = segs;
        sa.tile_pre = tpre;
        sa.chunk = 1;
        sa.child_hist = ch;
        launch("k_split_hist", 0, [&] { k_split_dev<W, 0><<<dim3((unsigned)g_hist), dim3(ST), 0, s>>>(sa, pl, l); });
        k_dev_child_scan<<<dim3((unsigned)g_cscan), 256, 0, s>>>(segs, ch, pl, l, ccur, nzr, nzc);
        chk();
        k_dev_level_counts<<<gnode, 256, 0, s>>>(pl, l, act, act_rank, nzc, ncnt);
        chk();
        scan_to_u64_dev(ncnt, nbase, &pl->lv[l].nn, c.nodes, stmp, s);
        k_dev_level_next<<<1, 1, 0, s>>>(pl, l, nbase, c.nodes);
        chk();
        k_dev_emit_nodes<<<gnode, 256, 0, s>>>(nodes[cur], npart[cur], pl, l, act, nbase, nodes[nxt], npart[nxt]);
        chk();
        k_dev_emit_children<<<gchild, 256, 0, s>>>(seg_node, nbase, nodes[cur], npart[cur], ch, ccur, nzr, pl, l,
                                                   nodes[nxt], npart[nxt]);
        chk();
        sa.child_cur = ccur;
        launch("k_split_scatter", 0, [&] {
            k_split_staged_dev<W><<<dim3((unsigned)g_stag), dim3(SST), 0, s>>>(sa, pl, l);
            k_split_dev<W, 1><<<dim3((unsigned)g_scat), dim3(ST), 0, s>>>(sa, pl, l);
        });
    }
    const int fin = DEV_LEVELS & 1;
    Bucket *bkts = nodes[fin];
    // no node may still need a split
    k_dev_level_flags<<<gnode, 256, 0, s>>>(bkts, pl, DEV_LEVELS, cap_above, kbits, tk, act, tiles);
    chk();
    k_dev_levels_done<<<1, 1, 0, s>>>(pl);
    chk();

    // 5. sort + run-length count every bucket, by size class
    uint32_t *pfirst = dv_.pfirst.as<uint32_t>(), *cls = dv_.cls.as<uint32_t>(), *ord = dv_.order.as<uint32_t>();
    uint32_t *fb = dv_.fb.as<uint32_t>(), *ucount = dv_.ucount.as<uint32_t>();
    k_dev_part_first<<<dim3((P + 1 + 255) / 256), 256, 0, s>>>(npart[fin], pl, P, pfirst);
    chk();
    ClassCaps caps;
    for (int q = 0; q < SORT_CLASSES; q++) caps.cap[q] = (uint64_t)class_cap(W, q);
    const dim3 gcls = capped_grid(c.nodes, CT, 512);
    k_dev_classify<<<gcls, CT, 0, s>>>(bkts, pl, kbits, caps, cls);
    chk();
    k_dev_class_offsets<<<1, 1, 0, s>>>(pl);
    chk();
    k_dev_order_fill<<<gcls, CT, 0, s>>>(cls, pl, ord);
    chk();
    bool hashed[SORT_CLASSES] = {false, false, false, false};
    auto sort_class = [&](auto cls_c) {
        constexpr int C = decltype(cls_c)::value;
        constexpr int IPT = SortClass<W, C>::IPT;
        if (t.dedup && C < hash_classes) {
            constexpr int IPTH = (C < 3 ? 4 : 8) / W > 0 ? (C < 3 ? 4 : 8) / W : 1;
            constexpr int NTH = SortClass<W, C>::CAP / IPTH;
            const void *f = reinterpret_cast<const void *>(&k_bucket_hash_sort_dev<W, IPTH, NTH, C>);
            const dim3 g((unsigned)resident_grid(f, NTH, 0));
            launch("k_bucket_sort", 0, [&] {
                k_bucket_hash_sort_dev<W, IPTH, NTH, C><<<g, dim3(NTH), 0, s>>>(A, B, ctmp, bkts, ord, ucount, k, fb,
                                                                                 pl);
            });
            hashed[C] = true;
        } else {
            const void *f = reinterpret_cast<const void *>(&k_bucket_sort_dev<W, IPT, C, false>);
            const dim3 g((unsigned)resident_grid(f, BT, 0));
            launch("k_bucket_sort", 0, [&] {
                k_bucket_sort_dev<W, IPT, C, false><<<g, dim3(BT), 0, s>>>(A, B, ctmp, bkts, ord, ucount, k, pl);
            });
        }
    };
    auto fallback_class = [&](auto cls_c) {
        constexpr int C = decltype(cls_c)::value;
        constexpr int IPT = SortClass<W, C>::IPT;
        if (!hashed[C]) return;
        const void *f = reinterpret_cast<const void *>(&k_bucket_sort_dev<W, IPT, C, true>);
        const dim3 g((unsigned)resident_grid(f, BT, 0));
        launch("k_bucket_sort", 0, [&] {
            k_bucket_sort_dev<W, IPT, C, true><<<g, dim3(BT), 0, s>>>(A, B, ctmp, bkts, fb, ucount, k, pl);
        });
    };
    sort_class(std::integral_constant<int, 0>{});
    sort_class(std::integral_constant<int, 1>{});
    sort_class(std::integral_constant<int, 2>{});
    sort_class(std::integral_constant<int, 3>{});
    fallback_class(std::integral_constant<int, 0>{});
    fallback_class(std::integral_constant<int, 1>{});
    fallback_class(std::integral_constant<int, 2>{});
    fallback_class(std::integral_constant<int, 3>{});

    // 6. contiguous per-partition output, 7. chunk encode, 8. summary
    uint64_t *uoff = dv_.uoff.as<uint64_t>(), *pu = dv_.pu.as<uint64_t>();
    scan_to_u64_dev(ucount, uoff, &pl->lv[DEV_LEVELS].nn, c.nodes, stmp, s);
    k_dev_unique<<<1, 1, 0, s>>>(pl, uoff);
    chk();
    uint64_t *okeys = dv_.keys.as<uint64_t>();
    uint32_t *ocounts = dv_.counts.as<uint32_t>();
    launch("k_compact", 0, [&] {
        k_compact_dev<W><<<capped_grid(c.nodes, 1, 4096), 256, 0, s>>>(A, B, ctmp, bkts, uoff, pl, okeys, ocounts);
    });
    k_dev_part_rows<<<dim3((P + 1 + 255) / 256), 256, 0, s>>>(uoff, pfirst, P, pl, pu);
    chk();
}

void CountEngine::dev_enqueue_encode(EncodeBufs &enc, uint8_t *bytes, uint64_t bytes_cap, void *h_summary) {
    hipStream_t s = stream_;
    const uint32_t P = dP_;
    DevPlan *pl = dv_.plan.as<DevPlan>();
    uint64_t *pu = dv_.pu.as<uint64_t>();
    encode_dev(dv_.keys.as<uint64_t>(), dv_.counts.as<uint32_t>(), pu, &pl->n_unique, dcap_.n, P, dW_, s, enc, bytes,
               bytes_cap, &pl->bytes, &pl->overflow);
    k_dev_summary<<<capped_grid(2ull * (P + 1) + 8ull * P, 256, 64), 256, 0, s>>>(
        pl, P, pu, enc.byte_off.as<uint64_t>(), enc.minmax.as<uint64_t>(), static_cast<DevSummary *>(h_summary));
    KITE_CHECK_LAUNCH();
}

template void CountEngine::dev_enqueue_w<1>(const uint8_t *, const uint64_t *, uint64_t, int, int, int);
template void CountEngine::dev_enqueue_w<2>(const uint8_t *, const uint64_t *, uint64_t, int, int, int);
template void CountEngine::dev_enqueue_w<4>(const uint8_t *, const uint64_t *, uint64_t, int, int, int);

}  // namespace kite
