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

#include "count.hpp"

namespace kite {

// ---- extract geometry -------------------------------------------------------
constexpr int XT = 256;                 // extract threads
constexpr int SUB = 8192;               // FASTQ bytes per sub-tile
constexpr int SEGB = 32;                // bytes per segment (one thread)
constexpr int HALO = 128;               // >= k_max - 1 bytes before each sub-tile
constexpr int HSEG = HALO / SEGB;       // 4 halo segments
constexpr int NSEG = SUB / SEGB + HSEG; // 260 segments per sub-tile
constexpr int SUBS = 4;                 // sub-tiles per extract workgroup
constexpr int EXT_U4 = (SUB + HALO) / 16;
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

__device__ __forceinline__ uint32_t is_acgt(uint32_t b) {
    uint32_t l = b | 0x20u;
    return (l == 'a') | (l == 'c') | (l == 'g') | (l == 't');
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
    return pbits ? (uint32_t)((f * 0x9E3779B97F4A7C15ull) >> (64 - pbits)) : 0u;
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
__global__ __launch_bounds__(XT) void k_nl_count(const uint8_t *__restrict__ fq, uint64_t len,
                                                  uint32_t *__restrict__ out) {
    __shared__ uint32_t ws[XT / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SUB + (uint64_t)threadIdx.x * SEGB;
    uint32_t c = 0;
    if (base + SEGB <= len) {
        const uint4 *p = reinterpret_cast<const uint4 *>(fq + base);
        uint4 a = p[0], b = p[1];
        c = newlines_in(a.x) + newlines_in(a.y) + newlines_in(a.z) + newlines_in(a.w) +
            newlines_in(b.x) + newlines_in(b.y) + newlines_in(b.z) + newlines_in(b.w);
    } else {
        for (uint64_t i = base; i < len && i < base + SEGB; i++) c += fq[i] == '\n';
    }
    uint32_t total;
    (void)block_excl_scan<XT, uint32_t>(c, ws, &total);
    if (threadIdx.x == 0) out[blockIdx.x] = total;
}

struct ExtractArgs {
    const uint8_t *fq;
    uint64_t len;
    const uint64_t *sub_line0;  // lines before each sub-tile
    uint64_t n_sub;
    uint32_t nwg;               // extract workgroups = matrix columns
    int k, canonical, pbits;
    uint32_t *hist;             // [P][nwg] counts (histogram pass)
    const uint64_t *off;        // scanned [P][nwg] (scatter pass)
    uint64_t *out;              // partition-ordered keys (scatter pass)
};

// Both passes walk identical k-mers; the scatter pass places each key at
// off[p][wg] + (slot within this workgroup's share of partition p).
template <int W, bool SCATTER>
__global__ __launch_bounds__(XT) void k_extract(ExtractArgs a) {
    __shared__ uint4 sb4[EXT_U4];
    __shared__ uint64_t scode[NSEG];
    __shared__ uint32_t svalid[NSEG];
    __shared__ uint32_t snl[NSEG];
    __shared__ uint32_t spre[NSEG];
    __shared__ uint32_t ws[XT / 64 + 1];
    extern __shared__ uint32_t dyn[];  // P u32 counters, then (scatter) P u64 bases
    const uint32_t P = 1u << a.pbits;
    uint32_t *cnt = dyn;
    uint64_t *pbase = reinterpret_cast<uint64_t *>(dyn + ((P + 1) & ~1u));
    const int tid = threadIdx.x;
    const uint32_t wg = (uint32_t)xcd_tile(blockIdx.x, gridDim.x);
    const uint8_t *sbytes = reinterpret_cast<const uint8_t *>(sb4);
    const uint32_t *sb = reinterpret_cast<const uint32_t *>(sb4);
    const bool live = a.len > 0 && a.fq[0] == '@';

    for (uint32_t p = tid; p < P; p += XT) {
        cnt[p] = 0;
        if constexpr (SCATTER) pbase[p] = a.off[(uint64_t)p * a.nwg + wg];
    }

    // key masks and reverse-complement shift (constant per launch)
    uint64_t kmask[W];
#pragma unroll
    for (int r = 0; r < W; r++) {
        int bits = 2 * a.k - 64 * r;
        kmask[r] = bits >= 64 ? ~0ull : (bits <= 0 ? 0ull : ((1ull << bits) - 1));
    }
    const int rs = 64 * W - 2 * a.k;
    const int rq = rs >> 6, rb = rs & 63;

    for (int si = 0; si < SUBS; si++) {
        uint64_t st = (uint64_t)wg * SUBS + si;
        if (st >= a.n_sub) break;
        const int64_t base = (int64_t)(st * SUB);
        __syncthreads();
        for (int i = tid; i < EXT_U4; i += XT) {
            int64_t gb = base - HALO + 16 * (int64_t)i;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (gb >= 0 && gb + 16 <= (int64_t)a.len) {
                v = *reinterpret_cast<const uint4 *>(a.fq + gb);
            } else if (gb + 16 > 0 && gb < (int64_t)a.len) {
                uint8_t tmp[16];
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    int64_t g = gb + q;
                    tmp[q] = (g >= 0 && g < (int64_t)a.len) ? a.fq[g] : 0;
                }
                memcpy(&v, tmp, 16);
            }
            sb4[i] = v;
        }
        __syncthreads();
        for (int s = tid; s < NSEG; s += XT) {
            uint32_t c = 0;
#pragma unroll
            for (int q = 0; q < SEGB / 4; q++) c += newlines_in(sb[s * (SEGB / 4) + q]);
            snl[s] = c;
        }
        __syncthreads();
        {
            uint32_t total;
            uint32_t ex = block_excl_scan<XT, uint32_t>(snl[tid], ws, &total);
            spre[tid] = ex;
            if (tid < NSEG - XT) {
                uint32_t t = total;
                for (int q = 0; q < tid; q++) t += snl[XT + q];
                spre[XT + tid] = t;
            }
        }
        __syncthreads();
        // line index (mod 4) at the first halo byte
        const uint32_t l0 = (uint32_t)(a.sub_line0[st] & 3u) - spre[HSEG];
        for (int s = tid; s < NSEG; s += XT) {
            uint32_t line = (l0 + spre[s]) & 3u;
            uint64_t code = 0;
            uint32_t valid = 0;
#pragma unroll
            for (int j = 0; j < SEGB; j++) {
                uint32_t b = sbytes[s * SEGB + j];
                code |= (uint64_t)(((b >> 1) ^ (b >> 2)) & 3u) << (62 - 2 * j);
                valid |= ((line == 1u) & is_acgt(b)) << j;
                line = (line + (b == '\n')) & 3u;
            }
            scode[s] = code;
            svalid[s] = live ? valid : 0u;
        }
        __syncthreads();

        const int s = tid + HSEG;
        uint32_t run = 0;
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
        const uint32_t vm = svalid[s];
        if (vm == 0) continue;
        uint64_t wv[W + 1];
#pragma unroll
        for (int i = 0; i <= W; i++) wv[i] = scode[s - W + i];
        for (int j = 0; j < SEGB; j++) {
            run = ((vm >> j) & 1u) ? run + 1 : 0;
            if (run < (uint32_t)a.k) continue;
            const int sh = 62 - 2 * j;
            Key<W> f;
#pragma unroll
            for (int r = 0; r < W; r++) {
                uint64_t lo = wv[W - r] >> sh;
                uint64_t hi = sh ? (wv[W - r - 1] << (64 - sh)) : 0ull;
                f.w[r] = (lo | hi) & kmask[r];
            }
            Key<W> key = f;
            if (a.canonical) {
                uint64_t R[W];
#pragma unroll
                for (int r = 0; r < W; r++) R[r] = rev2(~f.w[W - 1 - r]);
                Key<W> rc;
#pragma unroll
                for (int r = 0; r < W; r++) {
                    uint64_t lo = 0, hi = 0;
#pragma unroll
                    for (int t = 0; t < W; t++) {
                        if (t == r + rq) lo = R[t];
                        if (t == r + rq + 1) hi = R[t];
                    }
                    rc.w[r] = (lo >> rb) | (rb ? (hi << (64 - rb)) : 0ull);
                }
                if (key_less<W>(rc, f)) key = rc;
            }
            const uint32_t p = key_partition<W>(key, a.pbits);
            if constexpr (!SCATTER) {
                atomicAdd(&cnt[p], 1u);
            } else {
                uint64_t slot = pbase[p] + atomicAdd(&cnt[p], 1u);
                uint64_t *dst = a.out + slot * W;
#pragma unroll
                for (int r = 0; r < W; r++) dst[r] = key.w[r];
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
    const uint64_t *tile_pre;  // nseg + 1 tile prefix
    int shift, dbits;
    uint32_t *child_hist;      // histogram pass
    uint64_t *child_cur;       // scatter pass: running destination per child
};

template <int W, bool SCATTER>
__global__ __launch_bounds__(ST) void k_split(SplitArgs a) {
    constexpr int IPT = SplitCfg<W>::IPT, TK = SplitCfg<W>::TK;
    __shared__ uint32_t hist[1 << MAX_DBITS];
    __shared__ uint64_t cbase[SCATTER ? (1 << MAX_DBITS) : 1];
    const int tid = threadIdx.x;
    const uint64_t b = xcd_tile(blockIdx.x, gridDim.x);
    uint32_t lo = 0, hi = a.nseg;  // largest s with tile_pre[s] <= b
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (a.tile_pre[mid] <= b) lo = mid;
        else hi = mid;
    }
    const Seg sg = a.segs[lo];
    const uint64_t t = b - a.tile_pre[lo];
    const uint64_t first = sg.start + t * TK;
    const uint64_t n = std::min<uint64_t>((uint64_t)TK, sg.size - t * TK);
    const uint32_t nd = 1u << a.dbits;
    for (uint32_t g = tid; g < nd; g += ST) hist[g] = 0;
    __syncthreads();
    Key<W> key[IPT];
    uint32_t dig[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint64_t i = (uint64_t)it * ST + tid;
        if (i < n) {
            const uint64_t *p = a.src + (first + i) * W;
#pragma unroll
            for (int r = 0; r < W; r++) key[it].w[r] = p[r];
            dig[it] = key_bits<W>(key[it], a.shift, a.dbits);
            atomicAdd(&hist[dig[it]], 1u);
        }
    }
    __syncthreads();
    if constexpr (!SCATTER) {
        for (uint32_t g = tid; g < nd; g += ST)
            if (hist[g]) atomicAdd(&a.child_hist[sg.child_base + g], hist[g]);
    } else {
        for (uint32_t g = tid; g < nd; g += ST) {
            uint32_t c = hist[g];
            if (c) cbase[g] = atomicAdd((unsigned long long *)&a.child_cur[sg.child_base + g],
                                        (unsigned long long)c);
            hist[g] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IPT; it++) {
            uint64_t i = (uint64_t)it * ST + tid;
            if (i < n) {
                uint64_t slot = cbase[dig[it]] + atomicAdd(&hist[dig[it]], 1u);
                uint64_t *d = a.dst + slot * W;
#pragma unroll
                for (int r = 0; r < W; r++) d[r] = key[it].w[r];
            }
        }
    }
}

// ---- per-bucket LSD radix sort in LDS + run-length counts ------------------
// One 256-thread workgroup per bucket; IPT (items per thread) sets the size
// class, so small buckets take little LDS and many workgroups share a CU.
// Keys are held warp-striped: wave w owns positions [w*64*ipt, (w+1)*64*ipt),
// item `it` of lane l is position w*64*ipt + it*64 + l.  Ranking inside a
// wave uses ballot-matched digits, processed item by item, so every pass is
// stable in position order; waves are combined by a (digit, wave) scan.
template <int W, int IPT>
__global__ __launch_bounds__(BT) void k_bucket_sort(uint64_t *__restrict__ A,
                                                     uint64_t *__restrict__ B,
                                                     uint32_t *__restrict__ counts_tmp,
                                                     const Bucket *__restrict__ buckets,
                                                     const uint32_t *__restrict__ order,
                                                     uint32_t *__restrict__ ucount, int k) {
    constexpr int CAP = BT * IPT;
    __shared__ uint64_t skeys[CAP * W];
    __shared__ uint32_t whist[256 * BNW];  // [digit][wave]
    __shared__ uint16_t upos[CAP + 1];
    __shared__ uint32_t ws[BNW + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const uint32_t bid = order[blockIdx.x];
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

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }

// Hash-then-bin variant: every key is inserted into an LDS open-addressing
// table (a slot holds the first position of its key, claimed by CAS) and
// duplicates add to that position's count.  The unique keys, carrying their
// counts, are then distributed by one counting pass over 2*CAP bins on
// their top remaining key bits (order inside a bin does not matter: keys
// are distinct), and each thread insertion-sorts the few keys of its own
// contiguous bins.  A bin above LOCAL_MAX keys (skewed key bits) sends the
// whole bucket through the LSD passes instead.  Output equals
// k_bucket_sort's: ascending unique keys with their run lengths.
template <int W, int IPT>
__global__ __launch_bounds__(BT) void k_bucket_hash_sort(uint64_t *__restrict__ A,
                                                          uint64_t *__restrict__ B,
                                                          uint32_t *__restrict__ counts_tmp,
                                                          const Bucket *__restrict__ buckets,
                                                          const uint32_t *__restrict__ order,
                                                          uint32_t *__restrict__ ucount, int k) {
    constexpr int CAP = BT * IPT;
    constexpr int HT = 2 * CAP;  // hash slots (load factor <= 1/2), then bins
    constexpr int BPT = HT / BT; // bins per thread
    constexpr int DB = ilog2c(HT);
    constexpr int LOCAL_MAX = 16;
    constexpr bool WH_IN_HT = HT >= 256 * BNW;  // the LSD histogram reuses the table when it fits
    __shared__ uint64_t skeys[CAP * W];
    __shared__ uint32_t scnt[CAP];
    __shared__ uint32_t ht[HT];
    __shared__ uint32_t whist_own[WH_IN_HT ? 1 : 256 * BNW];
    __shared__ uint32_t ws[BNW + 1];
    __shared__ uint32_t s_fallback;
    uint32_t *whist = WH_IN_HT ? ht : whist_own;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const uint32_t bid = order[blockIdx.x];
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
    int ipt = (int)((m + BT - 1) / BT);
    uint32_t wbase = (uint32_t)wave * 64u * (uint32_t)ipt;
    for (int j = tid; j < HT; j += BT) ht[j] = 0;
    for (int j = tid; j < CAP; j += BT) scnt[j] = 0;
    Key<W> key[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < m) {
            const uint64_t *p = X + (start + pos) * W;
#pragma unroll
            for (int r = 0; r < W; r++) {
                key[it].w[r] = p[r];
                skeys[pos * W + r] = key[it].w[r];
            }
        }
    }
    __syncthreads();
    // 1. hash dedup: the first position of every key collects its count
    uint32_t is_rep = 0;  // bit it: item it holds the first position of its key
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < m) {
            uint64_t f = key[it].w[0];
#pragma unroll
            for (int r = 1; r < W; r++) f ^= key[it].w[r] * 0xC2B2AE3D27D4EB4Full;
            uint32_t h = (uint32_t)((f * 0xD6E8FEB86659FD93ull) >> 40) & (HT - 1);
            uint32_t rep;
            for (;;) {
                uint32_t old = atomicCAS(&ht[h], 0u, pos + 1);
                if (old == 0) {
                    rep = pos;
                    break;
                }
                bool eq = true;
#pragma unroll
                for (int r = 0; r < W; r++) eq = eq && skeys[(old - 1) * W + r] == key[it].w[r];
                if (eq) {
                    rep = old - 1;
                    break;
                }
                h = (h + 1) & (HT - 1);
            }
            atomicAdd(&scnt[rep], 1u);
            if (rep == pos) is_rep |= 1u << it;
        }
    }
    __syncthreads();
    // compact the representatives with their counts
    uint32_t mine = __popc(is_rep), U;
    uint32_t q = block_excl_scan<BT, uint32_t>(mine, ws, &U);
    uint32_t cnt[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        cnt[it] = ((is_rep >> it) & 1u) ? scnt[pos] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        if ((is_rep >> it) & 1u) {
#pragma unroll
            for (int r = 0; r < W; r++) skeys[q * W + r] = key[it].w[r];
            scnt[q] = cnt[it];
            q++;
        }
    }
    __syncthreads();
    ipt = (int)((U + BT - 1) / BT);
    wbase = (uint32_t)wave * 64u * (uint32_t)ipt;
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < U) {
#pragma unroll
            for (int r = 0; r < W; r++) key[it].w[r] = skeys[pos * W + r];
            cnt[it] = scnt[pos];
        }
    }
    // 2. one counting pass over HT bins on the top remaining key bits
    const int db = bits < DB ? bits : DB;
    const int dshift = bits - db;
    for (int j = tid; j < HT; j += BT) ht[j] = 0;
    if (tid == 0) s_fallback = 0;
    __syncthreads();
    uint32_t slot[IPT], dig[IPT];
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < U) {
            dig[it] = key_bits<W>(key[it], dshift, db);
            slot[it] = atomicAdd(&ht[dig[it]], 1u);
        }
    }
    __syncthreads();
    {
        uint32_t v[BPT], sum = 0, big = 0;
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            v[b] = ht[tid * BPT + b];
            sum += v[b];
            big |= v[b] > (uint32_t)LOCAL_MAX;
        }
        uint32_t total;
        uint32_t run = block_excl_scan<BT, uint32_t>(sum, ws, &total);
#pragma unroll
        for (int b = 0; b < BPT; b++) {
            ht[tid * BPT + b] = run;
            run += v[b];
        }
        if (big) s_fallback = 1;
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IPT; it++) {
        uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
        if (it < ipt && pos < U) {
            uint32_t dst = ht[dig[it]] + slot[it];
#pragma unroll
            for (int r = 0; r < W; r++) skeys[dst * W + r] = key[it].w[r];
            scnt[dst] = cnt[it];
        }
    }
    __syncthreads();
    if (!s_fallback) {
        // 3. every thread sorts the keys of its own bins (at most LOCAL_MAX each)
        uint32_t lo = ht[tid * BPT];
        for (int b = 0; b < BPT; b++) {
            const uint32_t hi = b + 1 < BPT ? ht[tid * BPT + b + 1] : (tid + 1 < BT ? ht[(tid + 1) * BPT] : U);
            for (uint32_t i = lo + 1; i < hi; i++) {
                Key<W> kk;
#pragma unroll
                for (int r = 0; r < W; r++) kk.w[r] = skeys[i * W + r];
                const uint32_t c = scnt[i];
                uint32_t j = i;
                while (j > lo) {
                    Key<W> pk;
#pragma unroll
                    for (int r = 0; r < W; r++) pk.w[r] = skeys[(j - 1) * W + r];
                    if (!key_less<W>(kk, pk)) break;
#pragma unroll
                    for (int r = 0; r < W; r++) skeys[j * W + r] = pk.w[r];
                    scnt[j] = scnt[j - 1];
                    j--;
                }
#pragma unroll
                for (int r = 0; r < W; r++) skeys[j * W + r] = kk.w[r];
                scnt[j] = c;
            }
            lo = hi;
        }
        __syncthreads();
    } else {
        // skewed key bits: stable LSD passes over the unique keys
        for (int shift = 0; shift < bits && U > 1; shift += 8) {
            const int nb = std::min(8, bits - shift);
            for (int j = tid; j < 256 * BNW; j += BT) whist[j] = 0;
            __syncthreads();
            uint32_t rank[IPT];
#pragma unroll
            for (int it = 0; it < IPT; it++) {
                if (it < ipt) {
                    uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
                    bool valid = pos < U;
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
                if (it < ipt && pos < U) {
                    uint32_t dst = whist[dig[it] * BNW + wave] + rank[it];
#pragma unroll
                    for (int r = 0; r < W; r++) skeys[dst * W + r] = key[it].w[r];
                    scnt[dst] = cnt[it];
                }
            }
            __syncthreads();
#pragma unroll
            for (int it = 0; it < IPT; it++) {
                uint32_t pos = wbase + (uint32_t)it * 64u + (uint32_t)lane;
                if (it < ipt && pos < U) {
#pragma unroll
                    for (int r = 0; r < W; r++) key[it].w[r] = skeys[pos * W + r];
                    cnt[it] = scnt[pos];
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t q2 = tid; q2 < U; q2 += BT) {
        uint64_t *d = X + (start + q2) * W;
#pragma unroll
        for (int r = 0; r < W; r++) d[r] = skeys[q2 * W + r];
        counts_tmp[start + q2] = scnt[q2];
    }
    if (tid == 0) ucount[bid] = U;
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

template <class T>
void upload(DevBuf &buf, const std::vector<T> &v, hipStream_t s) {
    buf.ensure(v.size() * sizeof(T));
    if (!v.empty())
        KITE_HIP(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

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
    for (DevBuf *b : {&nl_, &line0_, &scan_tmp_, &hist_, &off_, &part_off_d_, &a_, &b_, &counts_tmp_, &out_keys_[0],
                      &out_keys_[1], &out_counts_[0], &out_counts_[1], &segs_, &tile_pre_, &child_hist_, &child_cur_,
                      &buckets_, &ucount_, &uoff_, &part_first_, &part_u_, &order_})
        b->release();
}

uint64_t CountEngine::bytes_needed(uint64_t n, int words, uint64_t fastq, uint32_t partitions) {
    uint64_t kb = 8ull * (uint64_t)words;
    uint64_t n_sub = ceil_div(fastq, SUB);
    uint64_t nwg = ceil_div(n_sub, SUBS);
    uint64_t mat = (uint64_t)partitions * nwg;
    return n * (3 * kb + 8) + mat * 12 + n_sub * 12 + (n / 256 + 1) * 64 + (64ull << 20);
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
    // tunables: built-in defaults, overridden per run by the environment
    tun = CountTunables{};
    if (const char *e = getenv("KITE_SPLIT_DIV")) tun.split_target_div = std::max(1, atoi(e));
    if (const char *e = getenv("KITE_SORT_DEDUP")) tun.dedup = atoi(e) != 0;
    if (const char *e = getenv("KITE_MAX_DBITS")) tun.max_digit_bits = std::min(MAX_DBITS, std::max(1, atoi(e)));
    if (words_ == 1) run_w<1>(d_fq, len, k, canonical, pbits);
    else if (words_ == 2) run_w<2>(d_fq, len, k, canonical, pbits);
    else run_w<4>(d_fq, len, k, canonical, pbits);
}

template <int W>
void CountEngine::run_w(const uint8_t *d_fq, uint64_t len, int k, int canonical, int pbits) {
    hipStream_t s = stream_;
    const uint32_t P = 1u << pbits;
    const int CAP = sort_cap(W);
    const uint64_t TK = (uint64_t)split_tk(W);
    EventTimer tm(timing_, s);
    tm.mark();
    part_off_h_.assign(P + 1, 0);
    n_kmers_ = n_unique_ = n_buckets_ = 0;
    kstats_.clear();
    pending_.clear();
    levels_ = 0;

    const uint64_t n_sub = std::max<uint64_t>(1, ceil_div(len, SUB));
    const uint32_t nwg = (uint32_t)ceil_div(n_sub, SUBS);
    const uint64_t mat = (uint64_t)P * nwg;

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
    ea.k = k;
    ea.canonical = canonical;
    ea.pbits = pbits;
    ea.hist = static_cast<uint32_t *>(hist_.ensure(mat * 4));
    size_t dyn_hist = (size_t)P * 4;
    size_t dyn_scat = (size_t)((P + 1) & ~1u) * 4 + (size_t)P * 8;
    {
        static bool attr_set[3] = {false, false, false};
        const int wi = W == 1 ? 0 : (W == 2 ? 1 : 2);
        if (!attr_set[wi]) {
            const int max_dyn = (int)(((MAX_PARTITIONS + 1) & ~1u) * 4 + MAX_PARTITIONS * 8);
            KITE_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_extract<W, false>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn));
            KITE_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_extract<W, true>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, max_dyn));
            attr_set[wi] = true;
        }
    }
    launch("k_extract_hist", len + mat * 4,
           [&] { k_extract<W, false><<<dim3(nwg), dim3(XT), dyn_hist, s>>>(ea); });
    uint64_t *off = static_cast<uint64_t *>(off_.ensure((mat + 1) * 8));
    scan_to_u64(ea.hist, off, mat, stmp, s);
    uint64_t *part_off_d = static_cast<uint64_t *>(part_off_d_.ensure((P + 1) * 8));
    k_part_gather<<<dim3((P + 1 + 255) / 256), dim3(256), 0, s>>>(off, nwg, P, part_off_d);
    KITE_CHECK_LAUNCH();
    KITE_HIP(hipMemcpyAsync(part_off_h_.data(), part_off_d, (P + 1) * 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    const uint64_t n = part_off_h_[P];
    n_kmers_ = n;

    // 3. scatter keys into partition order
    uint64_t *A = static_cast<uint64_t *>(a_.ensure(std::max<uint64_t>(n, 1) * 8 * W));
    uint64_t *B = static_cast<uint64_t *>(b_.ensure(std::max<uint64_t>(n, 1) * 8 * W));
    uint32_t *ctmp = static_cast<uint32_t *>(counts_tmp_.ensure(std::max<uint64_t>(n, 1) * 4));
    if (n > 0) {
        ea.off = off;
        ea.out = A;
        launch("k_extract_scatter", len + mat * 8 + n * 8 * W,
               [&] { k_extract<W, true><<<dim3(nwg), dim3(XT), dyn_scat, s>>>(ea); });
    }
    tm.mark();

    // 4. MSD splits until every bucket fits one LDS sort (or is exhausted)
    std::vector<Bucket> nodes;
    std::vector<uint32_t> node_part;
    nodes.reserve(P);
    for (uint32_t p = 0; p < P; p++) {
        uint64_t sz = part_off_h_[p + 1] - part_off_h_[p];
        if (sz == 0) continue;
        nodes.push_back(Bucket{part_off_h_[p], sz, 0u, 0u});
        node_part.push_back(p);
    }
    const int kbits = 2 * k;
    for (;;) {
        std::vector<uint32_t> act;
        uint64_t maxsz = 0;
        for (uint32_t i = 0; i < nodes.size(); i++)
            if (nodes[i].size > (uint64_t)CAP && (int)nodes[i].prefix < kbits) {
                act.push_back(i);
                maxsz = std::max(maxsz, nodes[i].size);
            }
        if (act.empty()) break;
        const int prefix = (int)nodes[act[0]].prefix;  // uniform within a level
        const uint32_t src_arr = nodes[act[0]].array;
        int d = 1;
        while (d < tun.max_digit_bits &&
               (maxsz >> d) > (uint64_t)(CAP / std::max(1, tun.split_target_div)))
            d++;
        d = std::min(d, kbits - prefix);
        const uint32_t nd = 1u << d;
        std::vector<Seg> segs(act.size());
        std::vector<uint64_t> tpre(act.size() + 1, 0);
        for (size_t i = 0; i < act.size(); i++) {
            const Bucket &nb = nodes[act[i]];
            segs[i] = Seg{nb.start, nb.size, (uint64_t)i * nd};
            tpre[i + 1] = tpre[i] + ceil_div(nb.size, TK);
        }
        const uint64_t tiles = tpre.back();
        const uint64_t nchild = (uint64_t)act.size() * nd;
        upload(segs_, segs, s);
        upload(tile_pre_, tpre, s);
        uint32_t *ch = static_cast<uint32_t *>(child_hist_.ensure(nchild * 4));
        k_fill_u32<<<dim3((unsigned)ceil_div(nchild, 256)), dim3(256), 0, s>>>(ch, nchild, 0);
        KITE_CHECK_LAUNCH();
        SplitArgs sa{};
        sa.src = src_arr ? B : A;
        sa.dst = src_arr ? A : B;
        sa.segs = segs_.as<Seg>();
        sa.nseg = (uint32_t)act.size();
        sa.tile_pre = tile_pre_.as<uint64_t>();
        sa.shift = kbits - prefix - d;
        sa.dbits = d;
        sa.child_hist = ch;
        uint64_t split_keys = 0;
        for (const Seg &sg : segs) split_keys += sg.size;
        launch("k_split_hist", split_keys * 8 * W,
               [&] { k_split<W, false><<<dim3((unsigned)tiles), dim3(ST), 0, s>>>(sa); });
        std::vector<uint32_t> chh(nchild);
        KITE_HIP(hipMemcpyAsync(chh.data(), ch, nchild * 4, hipMemcpyDeviceToHost, s));
        KITE_HIP(hipStreamSynchronize(s));
        std::vector<uint64_t> cur(nchild);
        for (size_t i = 0; i < act.size(); i++) {
            uint64_t o = segs[i].start;
            for (uint32_t g = 0; g < nd; g++) {
                cur[i * nd + g] = o;
                o += chh[i * nd + g];
            }
        }
        upload(child_cur_, cur, s);
        sa.child_cur = child_cur_.as<uint64_t>();
        launch("k_split_scatter", split_keys * 16 * W,
               [&] { k_split<W, true><<<dim3((unsigned)tiles), dim3(ST), 0, s>>>(sa); });
        levels_++;
        // expand the node list in order
        std::vector<Bucket> nn;
        std::vector<uint32_t> np;
        nn.reserve(nodes.size() + nchild);
        size_t ai = 0;
        for (uint32_t i = 0; i < nodes.size(); i++) {
            if (ai < act.size() && act[ai] == i) {
                for (uint32_t g = 0; g < nd; g++) {
                    uint64_t sz = chh[ai * nd + g];
                    if (!sz) continue;
                    nn.push_back(Bucket{cur[ai * nd + g], sz, (uint32_t)(prefix + d), src_arr ^ 1u});
                    np.push_back(node_part[i]);
                }
                ai++;
            } else {
                nn.push_back(nodes[i]);
                np.push_back(node_part[i]);
            }
        }
        nodes.swap(nn);
        node_part.swap(np);
        KITE_HIP(hipStreamSynchronize(s));  // host vectors of this level go out of scope
        // the next level only splits nodes produced by this one
        for (auto &nd2 : nodes)
            if (nd2.size > (uint64_t)CAP && (int)nd2.prefix < kbits && (int)nd2.prefix != prefix + d)
                throw Error("internal: mixed split levels");
    }
    tm.mark();

    // 5. sort + run-length count every bucket
    const uint64_t nb = nodes.size();
    n_buckets_ = nb;
    std::vector<uint32_t> pfirst(P + 1, (uint32_t)nb);
    {
        uint32_t j = 0;
        for (uint32_t p = 0; p <= P; p++) {
            while (j < nb && node_part[j] < p) j++;
            pfirst[p] = j;
        }
    }
    uint64_t *uoff = static_cast<uint64_t *>(uoff_.ensure((nb + 1) * 8));
    uint32_t *ucount = static_cast<uint32_t *>(ucount_.ensure(std::max<uint64_t>(nb, 1) * 4));
    if (nb > 0) {
        upload(buckets_, nodes, s);
        // size classes: the smallest LDS footprint that holds each bucket
        std::vector<uint32_t> order;
        order.reserve(nb);
        uint32_t cls_off[SORT_CLASSES + 1] = {0};
        for (int c = 0; c < SORT_CLASSES; c++) {
            cls_off[c] = (uint32_t)order.size();
            for (uint32_t b = 0; b < nb; b++) {
                const Bucket &bk = nodes[b];
                int cb = 0;
                if ((int)bk.prefix < kbits && bk.size > 1)
                    while (cb < SORT_CLASSES - 1 && bk.size > (uint64_t)class_cap(W, cb)) cb++;
                if (cb == c) order.push_back(b);
            }
        }
        cls_off[SORT_CLASSES] = (uint32_t)order.size();
        upload(order_, order, s);
        const uint32_t *ord = order_.as<uint32_t>();
        auto go = [&](auto cls, uint32_t first, uint32_t cnt) {
            constexpr int C = decltype(cls)::value;
            if (!cnt) return;
            // the hash kernel's table doubles the LDS footprint: above class 1
            // it would run one workgroup per CU, where the LSD kernel is faster
            if constexpr (C < 2) {
                if (tun.dedup) {
                    launch("k_bucket_sort", 0, [&] {
                        k_bucket_hash_sort<W, SortClass<W, C>::IPT><<<dim3(cnt), dim3(BT), 0, s>>>(
                            A, B, ctmp, buckets_.as<Bucket>(), ord + first, ucount, k);
                    });
                    return;
                }
            }
            launch("k_bucket_sort", 0, [&] {
                k_bucket_sort<W, SortClass<W, C>::IPT><<<dim3(cnt), dim3(BT), 0, s>>>(
                    A, B, ctmp, buckets_.as<Bucket>(), ord + first, ucount, k);
            });
        };
        go(std::integral_constant<int, 0>{}, cls_off[0], cls_off[1] - cls_off[0]);
        go(std::integral_constant<int, 1>{}, cls_off[1], cls_off[2] - cls_off[1]);
        go(std::integral_constant<int, 2>{}, cls_off[2], cls_off[3] - cls_off[2]);
        go(std::integral_constant<int, 3>{}, cls_off[3], cls_off[4] - cls_off[3]);
        KITE_HIP(hipStreamSynchronize(s));  // `order` is a local vector
    }
    tm.mark();
    uint64_t *stmp2 = static_cast<uint64_t *>(scan_tmp_.ensure(
        std::max<uint64_t>(scan_elems, scan_temp_elems(nb + 1)) * 8));
    scan_to_u64(ucount, uoff, nb, stmp2, s);
    KITE_HIP(hipMemcpyAsync(&n_unique_, uoff + nb, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    uint64_t *ok = static_cast<uint64_t *>(out_keys_[slot_].ensure(std::max<uint64_t>(n_unique_, 1) * 8 * W));
    uint32_t *oc = static_cast<uint32_t *>(out_counts_[slot_].ensure(std::max<uint64_t>(n_unique_, 1) * 4));
    if (nb > 0) {
        launch("k_compact", n_unique_ * (8 * W + 4) * 2, [&] {
            k_compact<W><<<dim3((unsigned)nb), dim3(256), 0, s>>>(A, B, ctmp, buckets_.as<Bucket>(),
                                                               uoff, ok, oc);
        });
    }
    // per-partition output offsets
    upload(part_first_, pfirst, s);
    uint64_t *pu = static_cast<uint64_t *>(part_u_.ensure((P + 1) * 8));
    k_gather_u64<<<dim3((P + 1 + 255) / 256), dim3(256), 0, s>>>(uoff, part_first_.as<uint32_t>(),
                                                               P + 1, pu);
    KITE_CHECK_LAUNCH();
    KITE_HIP(hipMemcpyAsync(part_off_h_.data(), pu, (P + 1) * 8, hipMemcpyDeviceToHost, s));
    tm.mark();
    KITE_HIP(hipStreamSynchronize(s));
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

}  // namespace kite
