// dump.hip -- `kite dump` on the device.
//
// Semantics (src/dump/merge.rs:13-240, 282-428; src/dump/decoder.rs:27-117;
// src/dump/loser_tree.rs:141-175; src/kiteops/mod.rs:44-56):
//   * a DVLE run is the first key (little-endian, 8 or 16 bytes) and its u32
//     count, then per row LEB128(key - previous key) and LEB128(count); the
//     count is read as u64 and truncated to u32; a RAW row is the key words
//     and a u32 count;
//   * the runs of one partition are merged in key order and equal keys sum
//     with u32 saturation;
//   * a row is printed when min <= count <= max (each bound optional), as
//     the k letters ACGT of the key, most significant base first, a tab, the
//     decimal count and a newline.
//
// Device plan for a group of partitions:
//   1. DVLE runs are cut into 4 KiB byte tiles.  Every byte without the
//      continuation bit ends a varint; per tile the ends are counted, a scan
//      gives every end its ordinal inside its run, and the odd ordinals (the
//      ends of count varints) mark where each row ends.
//   2. One thread per row parses its two varints between the previous row's
//      end and its own; RAW rows and first rows carry whole keys.
//   3. One workgroup per DVLE run adds the deltas up (a running 64- or
//      128-bit prefix sum) into keys.
//   4. Partitions with several runs: the r-th run of every partition forms
//      table r, and the tables are merged by the merge-path engine (u64
//      sums, clamped to u32 when printed, which equals u32 saturating adds).
//   5. Lines are formatted per tile of 256 rows in LDS, then copied out with
//      aligned word stores at the tile's offset from a scan of tile lengths.
// Malformed DVLE input (a varint with no end byte within 10 / 19 bytes, or
// a payload that ends before its rows) is reported as an error.
#include <algorithm>

#include "dump.hpp"

namespace kite {

namespace {

constexpr int DT = 256;                    // threads per dump workgroup
constexpr int TB = 4096;                   // DVLE bytes per tile (16 per thread)
constexpr uint64_t NO_END = ~0ull;
constexpr int MAX_LINE = 128 + 1 + 10 + 1;  // k <= 128, u32 count

enum : uint32_t { ERR_TRUNCATED = 1, ERR_VARINT = 2, ERR_ORDER = 4 };

__device__ __forceinline__ uint32_t term_mask16(uint4 v, int64_t rel0, uint64_t lo, uint64_t hi) {
    // bit q: byte q of the 16 loaded ends a varint and lies in [lo, hi)
    uint32_t m = 0;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const uint32_t b = (w[q >> 2] >> (8 * (q & 3))) & 0xffu;
        const int64_t rel = rel0 + q;
        m |= ((b & 0x80u) == 0 && rel >= (int64_t)lo && rel < (int64_t)hi ? 1u : 0u) << q;
    }
    return m;
}

template <int W>
constexpr uint64_t dvle_head() { return 8 * W + 4; }

// 1a. ends per tile
template <int W>
__global__ __launch_bounds__(DT) void k_dvle_count(const uint8_t *__restrict__ pay, const DumpRun *__restrict__ runs,
                                                   const uint32_t *__restrict__ tile_run, uint32_t *__restrict__ tcnt) {
    __shared__ uint32_t ws[DT / 64 + 1];
    const uint32_t t = blockIdx.x;
    const DumpRun r = runs[tile_run[t]];
    const int64_t rel0 = (int64_t)(t - r.tile0) * TB + 16 * threadIdx.x;
    uint32_t c = 0;
    if (rel0 < (int64_t)r.bytes) {
        const uint4 v = *reinterpret_cast<const uint4 *>(pay + r.src + rel0);
        c = __popc(term_mask16(v, rel0, dvle_head<W>(), r.bytes));
    }
    uint32_t total;
    (void)block_excl_scan<DT, uint32_t>(c, ws, &total);
    if (threadIdx.x == 0) tcnt[t] = total;
}

// 1b. the end of every row (the odd ordinals)
template <int W>
__global__ __launch_bounds__(DT) void k_dvle_ends(const uint8_t *__restrict__ pay, const DumpRun *__restrict__ runs,
                                                  const uint32_t *__restrict__ tile_run,
                                                  const uint64_t *__restrict__ tbase, uint64_t *__restrict__ ends) {
    __shared__ uint32_t ws[DT / 64 + 1];
    const uint32_t t = blockIdx.x;
    const DumpRun r = runs[tile_run[t]];
    const int64_t rel0 = (int64_t)(t - r.tile0) * TB + 16 * threadIdx.x;
    uint32_t m = 0;
    if (rel0 < (int64_t)r.bytes) {
        const uint4 v = *reinterpret_cast<const uint4 *>(pay + r.src + rel0);
        m = term_mask16(v, rel0, dvle_head<W>(), r.bytes);
    }
    uint32_t total;
    const uint32_t ex = block_excl_scan<DT, uint32_t>(__popc(m), ws, &total);
    uint64_t o = tbase[t] - tbase[r.tile0] + ex;  // ordinal of this thread's first end
    for (uint32_t e = m; e; e &= e - 1, o++) {
        if (o & 1) {
            const uint64_t j = (o + 1) >> 1;  // row j >= 1 ends here
            if (j < r.rows) ends[r.row0 + j] = r.src + (uint64_t)rel0 + (uint64_t)(__ffs(e) - 1);
        }
    }
}

__global__ void k_fill_u64(uint64_t *__restrict__ p, uint64_t n, uint64_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__device__ __forceinline__ uint32_t find_run(const DumpRun *__restrict__ runs, uint32_t nr, uint64_t row) {
    uint32_t lo = 0, hi = nr;  // largest r with runs[r].row0 <= row
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (runs[mid].row0 <= row) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t rd_le(const uint8_t *p, int n) {
    uint64_t v = 0;
    for (int b = 0; b < n; b++) v |= (uint64_t)p[b] << (8 * b);
    return v;
}

// 2. one thread per row: whole keys (first and RAW rows) or deltas, counts
template <int W>
__global__ __launch_bounds__(DT) void k_dump_rows(const uint8_t *__restrict__ pay, const DumpRun *__restrict__ runs,
                                                  uint32_t nr, uint64_t n, const uint64_t *__restrict__ ends,
                                                  uint32_t trunc32, uint64_t *__restrict__ vals,
                                                  uint64_t *__restrict__ counts, uint32_t *__restrict__ err) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DumpRun r = runs[find_run(runs, nr, i)];
    const uint64_t j = i - r.row0;
    uint64_t v[W];
    uint64_t c;
    if (r.codec == DUMP_CODEC_RAW || W == 4) {
        const uint8_t *p = pay + r.src + j * (8 * W + 4);
#pragma unroll
        for (int w = 0; w < W; w++) v[w] = rd_le(p + 8 * w, 8);
        c = (uint32_t)rd_le(p + 8 * W, 4);
        if (W == 4 && j > 0) {
            // 256-bit runs are merged as sorted unique runs; the reference
            // sorts them (src/dump/merge.rs:311-371), which is the same for
            // every run a writer produces, and anything else is refused
            bool gt = false, decided = false;
            for (int w = W - 1; w >= 0 && !decided; w--) {
                const uint64_t pv = rd_le(p - (8 * W + 4) + 8 * w, 8);
                if (v[w] != pv) {
                    gt = v[w] > pv;
                    decided = true;
                }
            }
            if (!gt) atomicOr(err, ERR_ORDER);
        }
    } else if (j == 0) {
        const uint8_t *p = pay + r.src;
#pragma unroll
        for (int w = 0; w < W; w++) v[w] = rd_le(p + 8 * w, 8);
        c = (uint32_t)rd_le(p + 8 * W, 4);
    } else {
        const uint64_t end = ends[i];
        if (end == NO_END) {
            atomicOr(err, ERR_TRUNCATED);
            return;
        }
        uint64_t pos = j == 1 ? r.src + dvle_head<W>() : ends[i - 1] + 1;
        // the key delta: at most 10 (64-bit) or 19 (128-bit) bytes
        constexpr int MAXD = W == 1 ? 10 : 19;
        uint64_t lo = 0, hi = 0;
        int q = 0;
        bool done = false;
        for (; q < MAXD && pos + q < end; q++) {
            const uint64_t b = pay[pos + q] & 0x7fu;
            const int sh = 7 * q;
            if (sh < 64) {
                lo |= b << sh;
                if (sh > 57) hi |= b >> (64 - sh);
            } else if (sh < 128) {
                hi |= b << (sh - 64);
            }
            if (!(pay[pos + q] & 0x80u)) {
                done = true;
                q++;
                break;
            }
        }
        if (!done) {
            atomicOr(err, ERR_VARINT);
            return;
        }
        pos += q;
        // the count: the rest up to and including the row's end byte
        const uint64_t len = end + 1 - pos;
        if (len == 0 || len > 10) {
            atomicOr(err, ERR_VARINT);
            return;
        }
        uint64_t cv = 0;
        for (uint64_t t = 0; t < len; t++) cv |= (uint64_t)(pay[pos + t] & 0x7fu) << (7 * t);
        c = trunc32 ? (uint64_t)(uint32_t)cv : cv;
        v[0] = lo;
        if constexpr (W >= 2) v[1] = hi;
    }
#pragma unroll
    for (int w = 0; w < W; w++) vals[i * W + w] = v[w];
    counts[i] = c;
}

// 3. running sums of the deltas of each DVLE run: a sum per tile of rows,
// a scan of those sums per run, then each tile on top of its run prefix
template <int W>
struct Acc {
    uint64_t w[W];
};

template <int W>
__device__ __forceinline__ Acc<W> acc_add(const Acc<W> &a, const Acc<W> &b) {
    Acc<W> r;
    r.w[0] = a.w[0] + b.w[0];
    if constexpr (W == 2) r.w[1] = a.w[1] + b.w[1] + (r.w[0] < a.w[0] ? 1ull : 0ull);
    return r;
}

template <int W>
__device__ __forceinline__ Acc<W> wave_incl_scan_acc(Acc<W> v) {
    const unsigned lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        Acc<W> o;
#pragma unroll
        for (int w = 0; w < W; w++) o.w[w] = __shfl_up(v.w[w], d, 64);
        if (lane >= (unsigned)d) v = acc_add<W>(o, v);
    }
    return v;
}

// One step of a run's running sum: rows [b0, b0 + DT * RPT) of the run
// starting at row0, each thread RPT consecutive rows, on top of `carry`.
// Returns the carry for the next step.
constexpr int PREFIX_RPT = 8;
constexpr uint64_t PREFIX_TILE = (uint64_t)DT * PREFIX_RPT;

template <int W>
__device__ Acc<W> prefix_step(uint64_t *__restrict__ vals, uint64_t row0, uint64_t rows, uint64_t b0, Acc<W> carry,
                              uint64_t (*wsum)[W]) {
    constexpr int RPT = PREFIX_RPT;
    constexpr int NW = DT / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Acc<W> x[RPT], s;
#pragma unroll
    for (int w = 0; w < W; w++) s.w[w] = 0;
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const uint64_t j = b0 + (uint64_t)tid * RPT + u;
#pragma unroll
        for (int w = 0; w < W; w++) x[u].w[w] = j < rows ? vals[(row0 + j) * W + w] : 0ull;
        s = acc_add<W>(s, x[u]);
    }
    const Acc<W> inc = wave_incl_scan_acc<W>(s);
    if (lane == 63)
#pragma unroll
        for (int w = 0; w < W; w++) wsum[wave][w] = inc.w[w];
    __syncthreads();
    Acc<W> pre = carry, tot = carry;
    for (int q = 0; q < NW; q++) {
        Acc<W> ws;
#pragma unroll
        for (int w = 0; w < W; w++) ws.w[w] = wsum[q][w];
        if (q < wave) pre = acc_add<W>(pre, ws);
        tot = acc_add<W>(tot, ws);
    }
    __syncthreads();
    // exclusive prefix of this thread = block prefix + wave-inclusive - own
    Acc<W> neg;
    neg.w[0] = inc.w[0] - s.w[0];
    if constexpr (W == 2) neg.w[1] = inc.w[1] - s.w[1] - (inc.w[0] < s.w[0] ? 1ull : 0ull);
    Acc<W> run = acc_add<W>(pre, neg);
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const uint64_t j = b0 + (uint64_t)tid * RPT + u;
        run = acc_add<W>(run, x[u]);
        if (j < rows)
#pragma unroll
            for (int w = 0; w < W; w++) vals[(row0 + j) * W + w] = run.w[w];
    }
    return tot;
}

// Running sums of whole runs, one workgroup per run (used on the per-tile
// sums below, whose runs are short)
template <int W>
__global__ __launch_bounds__(DT) void k_dvle_prefix(const DumpRun *__restrict__ runs, const uint32_t *__restrict__ dvle,
                                                    uint64_t *__restrict__ vals) {
    __shared__ uint64_t wsum[DT / 64][W];
    const DumpRun r = runs[dvle[blockIdx.x]];
    Acc<W> carry;
#pragma unroll
    for (int w = 0; w < W; w++) carry.w[w] = 0;
    for (uint64_t b0 = 0; b0 < r.rows; b0 += PREFIX_TILE) carry = prefix_step<W>(vals, r.row0, r.rows, b0, carry, wsum);
}

// A tile of PREFIX_TILE rows (fewer at a run's end) of one DVLE run
struct PrefixTile {
    uint64_t row0;  // first row, in the decoded row space
    uint64_t rows;
};

// The delta sum of every tile, twice: `incl` is then scanned per run in
// place, and `own` keeps the tile's own sum
template <int W>
__global__ __launch_bounds__(DT) void k_dvle_tile_sum(const PrefixTile *__restrict__ tiles,
                                                      const uint64_t *__restrict__ vals, uint64_t *__restrict__ incl,
                                                      uint64_t *__restrict__ own) {
    __shared__ uint64_t wsum[DT / 64][W];
    const PrefixTile t = tiles[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Acc<W> s;
#pragma unroll
    for (int w = 0; w < W; w++) s.w[w] = 0;
#pragma unroll
    for (int u = 0; u < PREFIX_RPT; u++) {
        const uint64_t j = (uint64_t)tid * PREFIX_RPT + u;
        Acc<W> x;
#pragma unroll
        for (int w = 0; w < W; w++) x.w[w] = j < t.rows ? vals[(t.row0 + j) * W + w] : 0ull;
        s = acc_add<W>(s, x);
    }
    const Acc<W> inc = wave_incl_scan_acc<W>(s);
    if (lane == 63)
#pragma unroll
        for (int w = 0; w < W; w++) wsum[wave][w] = inc.w[w];
    __syncthreads();
    if (tid == 0) {
        Acc<W> tot;
#pragma unroll
        for (int w = 0; w < W; w++) tot.w[w] = 0;
        for (int q = 0; q < DT / 64; q++) {
            Acc<W> ws;
#pragma unroll
            for (int w = 0; w < W; w++) ws.w[w] = wsum[q][w];
            tot = acc_add<W>(tot, ws);
        }
#pragma unroll
        for (int w = 0; w < W; w++) incl[(uint64_t)blockIdx.x * W + w] = own[(uint64_t)blockIdx.x * W + w] = tot.w[w];
    }
}

// Each tile's running sums on top of the sum of the run's earlier tiles
// (incl - own: the inclusive tile prefix less the tile itself)
template <int W>
__global__ __launch_bounds__(DT) void k_dvle_tile_apply(const PrefixTile *__restrict__ tiles,
                                                        const uint64_t *__restrict__ incl,
                                                        const uint64_t *__restrict__ own, uint64_t *__restrict__ vals) {
    __shared__ uint64_t wsum[DT / 64][W];
    const PrefixTile t = tiles[blockIdx.x];
    Acc<W> a, b, carry;
#pragma unroll
    for (int w = 0; w < W; w++) {
        a.w[w] = incl[(uint64_t)blockIdx.x * W + w];
        b.w[w] = own[(uint64_t)blockIdx.x * W + w];
    }
    carry.w[0] = a.w[0] - b.w[0];
    if constexpr (W == 2) carry.w[1] = a.w[1] - b.w[1] - (a.w[0] < b.w[0] ? 1ull : 0ull);
    (void)prefix_step<W>(vals, t.row0, t.rows, 0, carry, wsum);
}

// 4. rows of each run into its merge table
template <int W>
__global__ __launch_bounds__(DT) void k_dump_gather(const DumpRun *__restrict__ runs, uint32_t nr, uint64_t n,
                                                    const uint64_t *__restrict__ vals,
                                                    const uint64_t *__restrict__ counts,
                                                    uint64_t *const *__restrict__ tkeys,
                                                    uint64_t *const *__restrict__ tcounts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DumpRun r = runs[find_run(runs, nr, i)];
    const uint64_t d = r.dst + (i - r.row0);
    uint64_t *kd = tkeys[r.table];
#pragma unroll
    for (int w = 0; w < W; w++) kd[d * W + w] = vals[i * W + w];
    tcounts[r.table][d] = counts[i];
}

// 5. lines
__device__ __forceinline__ uint32_t dec_digits(uint32_t c) {
    uint32_t d = 1;
    for (uint32_t t = c; t >= 10; t /= 10) d++;
    return d;
}

__device__ __forceinline__ bool keep(uint32_t c, const DumpFilter &f) {
    return !((f.has_min && c < f.min) || (f.has_max && c > f.max));
}

__device__ __forceinline__ uint32_t clamp32(uint64_t c) { return c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c; }

__global__ __launch_bounds__(DT) void k_tsv_tile_len(const uint64_t *__restrict__ counts, uint64_t n, int k,
                                                     DumpFilter f, uint32_t *__restrict__ tlen,
                                                     uint32_t *__restrict__ tlines) {
    __shared__ uint32_t ws[DT / 64 + 1];
    const uint64_t i = (uint64_t)blockIdx.x * DT + threadIdx.x;
    uint32_t len = 0;
    if (i < n) {
        const uint32_t c = clamp32(counts[i]);
        if (keep(c, f)) len = (uint32_t)k + 2 + dec_digits(c);
    }
    uint32_t total, lines;
    (void)block_excl_scan<DT, uint32_t>(len, ws, &total);
    (void)block_excl_scan<DT, uint32_t>(len ? 1u : 0u, ws, &lines);
    if (threadIdx.x == 0) {
        tlen[blockIdx.x] = total;
        tlines[blockIdx.x] = lines;
    }
}

template <int W>
__device__ __forceinline__ uint32_t base_at(const uint64_t *key, int k, int i) {
    // letter i (0 = first) of the key: a 2k-bit integer for 1 and 2 words;
    // for 4 words the first hb = min(k, 64) letters are the low 2*hb bits of
    // the high half (words 2, 3) and the rest the low 2(k-hb) bits of the
    // low half (src/dump/decoder.rs:57-107)
    int bit;
    if constexpr (W == 4) {
        const int hb = k < 64 ? k : 64;
        bit = i < hb ? 128 + 2 * (hb - 1 - i) : 2 * (k - 1 - i);
    } else {
        bit = 2 * (k - 1 - i);
    }
    return (uint32_t)(key[bit >> 6] >> (bit & 63)) & 3u;
}

template <int W>
__global__ __launch_bounds__(DT) void k_tsv_write(const uint64_t *__restrict__ keys,
                                                  const uint64_t *__restrict__ counts, uint64_t n, int k,
                                                  DumpFilter f, const uint64_t *__restrict__ toff,
                                                  char *__restrict__ out) {
    __shared__ uint32_t ws[DT / 64 + 1];
    __shared__ uint32_t sbuf[DT * MAX_LINE / 4 + 4];
    uint8_t *sb = reinterpret_cast<uint8_t *>(sbuf);
    const uint64_t i = (uint64_t)blockIdx.x * DT + threadIdx.x;
    uint32_t len = 0, c = 0;
    if (i < n) {
        c = clamp32(counts[i]);
        if (keep(c, f)) len = (uint32_t)k + 2 + dec_digits(c);
    }
    uint32_t L;
    const uint32_t o = block_excl_scan<DT, uint32_t>(len, ws, &L);
    const uint64_t G = toff[blockIdx.x];
    // LDS holds the tile's text shifted by G % 4, so that the copy-out
    // reads aligned words for aligned global stores
    const uint32_t a = (uint32_t)(G & 3u);
    if (len) {
        uint64_t key[W];
#pragma unroll
        for (int w = 0; w < W; w++) key[w] = keys[i * W + w];
        uint8_t *p = sb + a + o;
        for (int q = 0; q < k; q++) p[q] = "ACGT"[base_at<W>(key, k, q)];
        p[k] = '\t';
        const uint32_t nd = len - (uint32_t)k - 2;
        uint32_t v = c;
        for (int q = (int)nd - 1; q >= 0; q--) {
            p[k + 1 + q] = (uint8_t)('0' + v % 10);
            v /= 10;
        }
        p[len - 1] = '\n';
    }
    __syncthreads();
    // words [G & ~3, G + L) of the output; the partial first and last words
    // are written bytewise (neighbouring tiles own their other bytes)
    const uint64_t w0 = G >> 2, w1 = (G + L + 3) >> 2;
    for (uint64_t w = w0 + threadIdx.x; w < w1; w += DT) {
        const uint32_t lw = (uint32_t)(w - w0);
        const uint64_t b0 = w << 2;
        if (b0 >= G && b0 + 4 <= G + L) {
            reinterpret_cast<uint32_t *>(out)[w] = sbuf[lw];
        } else {
            for (int q = 0; q < 4; q++) {
                const uint64_t b = b0 + q;
                if (b >= G && b < G + L) out[b] = (char)sb[4 * lw + q];
            }
        }
    }
}

// Frequencies of the totals that pass the filter: values below HISTO_SMALL
// in per-workgroup LDS bins then one u64 add per used bin, larger values
// appended to a list (big[0] is its length).
__global__ __launch_bounds__(DT) void k_histo(const uint64_t *__restrict__ counts, uint64_t n, DumpFilter f,
                                              uint64_t *__restrict__ small, uint64_t *__restrict__ big) {
    __shared__ uint32_t bins[HISTO_SMALL];
    for (uint32_t b = threadIdx.x; b < HISTO_SMALL; b += DT) bins[b] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * DT + threadIdx.x; i < n; i += (uint64_t)gridDim.x * DT) {
        const uint64_t c = counts[i];
        if ((f.has_min && c < f.min) || (f.has_max && c > f.max)) continue;
        if (c < HISTO_SMALL) {
            atomicAdd(&bins[c], 1u);
        } else {
            const uint64_t at = atomicAdd((unsigned long long *)big, 1ull);
            big[1 + at] = c;
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < HISTO_SMALL; b += DT)
        if (bins[b]) atomicAdd((unsigned long long *)&small[b], (unsigned long long)bins[b]);
}

inline dim3 grid_for(uint64_t n) { return dim3((unsigned)std::max<uint64_t>(1, ceil_div(n, DT))); }

}  // namespace

void DumpEngine::release() {
    for (DevBuf *b : {&runs_d_, &tile_run_, &tcnt_, &tbase_, &ends_, &vals_, &counts_, &scan_tmp_, &tlen_, &toff_,
                      &text_, &err_, &tabptr_, &lines_, &loff_, &hsmall_, &hbig_d_, &ptile_, &ptrun_, &psum_, &poff_})
        b->release();
    tabs_.clear();
    merged_.keys.release();
    merged_.counts.release();
    merged_.off.release();
    if (host_err_) (void)hipHostFree(host_err_);
    host_err_ = nullptr;
}

uint64_t DumpEngine::run(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int k, int W,
                         const DumpFilter &f, uint64_t *rows_out, uint64_t *lines_out, DumpTimes *times) {
    if (W == 1) return run_w<1>(d_payload, runs, np, k, f, rows_out, lines_out, times);
    if (W == 2) return run_w<2>(d_payload, runs, np, k, f, rows_out, lines_out, times);
    return run_w<4>(d_payload, runs, np, k, f, rows_out, lines_out, times);
}

namespace {
struct Events {
    hipEvent_t e[4];
    Events() {
        for (auto &x : e) KITE_HIP(hipEventCreate(&x));
    }
    ~Events() {
        for (auto &x : e) (void)hipEventDestroy(x);
    }
};
}  // namespace

template <int W>
uint64_t DumpEngine::decode_merge_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs_in, uint32_t np,
                                    bool trunc32, const uint64_t **keys_out, const uint64_t **counts_out,
                                    hipEvent_t ev[3]) {
    hipStream_t s = stream_;
    if (!host_err_) KITE_HIP(hipHostMalloc(reinterpret_cast<void **>(&host_err_), 64, hipHostMallocDefault));
    KITE_HIP(hipEventRecord(ev[0], s));

    // runs with rows, tiles of the DVLE ones
    std::vector<DumpRun> runs;
    runs.reserve(runs_in.size());
    std::vector<uint32_t> tile_run, dvle;
    uint64_t n = 0;
    uint32_t R = 0;
    for (const DumpRun &x : runs_in) {
        if (!x.rows) continue;
        DumpRun r = x;
        r.row0 = n;
        r.tile0 = (uint32_t)tile_run.size();
        n += r.rows;
        R = std::max(R, r.table + 1);
        if (W != 4 && r.codec == DUMP_CODEC_DVLE && r.rows > 1) {
            const uint64_t nt = ceil_div(r.bytes, TB);
            if (tile_run.size() + nt > 0xFFFFFFFFull) throw Error("dump group above 2^32 byte tiles");
            for (uint64_t t = 0; t < nt; t++) tile_run.push_back((uint32_t)runs.size());
        }
        if (W != 4 && r.codec == DUMP_CODEC_DVLE) dvle.push_back((uint32_t)runs.size());
        runs.push_back(r);
    }
    const uint32_t nr = (uint32_t)runs.size();
    const uint64_t T = tile_run.size();
    // prefix tiles of the DVLE runs; truns[d] spans run d's tiles in the
    // tile list, so a run-wise scan over the tile sums reuses k_dvle_prefix
    std::vector<PrefixTile> ptiles;
    std::vector<DumpRun> truns;
    std::vector<uint32_t> tids;
    for (uint32_t d : dvle) {
        const DumpRun &r = runs[d];
        DumpRun tr{};
        tr.row0 = ptiles.size();
        for (uint64_t b0 = 0; b0 < r.rows; b0 += PREFIX_TILE)
            ptiles.push_back(PrefixTile{r.row0 + b0, std::min<uint64_t>(PREFIX_TILE, r.rows - b0)});
        tr.rows = ptiles.size() - tr.row0;
        tids.push_back((uint32_t)truns.size());
        truns.push_back(tr);
    }
    if (ptiles.size() > 0x7FFFFFFFull) throw Error("dump group above 2^31 prefix tiles");
    DumpRun *druns = static_cast<DumpRun *>(runs_d_.ensure(std::max<size_t>(nr, 1) * sizeof(DumpRun)));
    uint32_t *dtile = static_cast<uint32_t *>(tile_run_.ensure(std::max<uint64_t>(T, 1) * 4 + dvle.size() * 4 + 4));
    uint32_t *ddvle = dtile + std::max<uint64_t>(T, 1);
    if (nr) KITE_HIP(hipMemcpyAsync(druns, runs.data(), nr * sizeof(DumpRun), hipMemcpyHostToDevice, s));
    if (T) KITE_HIP(hipMemcpyAsync(dtile, tile_run.data(), T * 4, hipMemcpyHostToDevice, s));
    if (!dvle.empty()) KITE_HIP(hipMemcpyAsync(ddvle, dvle.data(), dvle.size() * 4, hipMemcpyHostToDevice, s));
    uint32_t *derr = static_cast<uint32_t *>(err_.ensure(4));
    KITE_HIP(hipMemsetAsync(derr, 0, 4, s));

    uint64_t *vals = static_cast<uint64_t *>(vals_.ensure(std::max<uint64_t>(n, 1) * 8 * W));
    uint64_t *cnts = static_cast<uint64_t *>(counts_.ensure(std::max<uint64_t>(n, 1) * 8));
    uint64_t *ends = static_cast<uint64_t *>(ends_.ensure(std::max<uint64_t>(n, 1) * 8));
    if (T) {
        uint32_t *tcnt = static_cast<uint32_t *>(tcnt_.ensure(T * 4));
        uint64_t *tbase = static_cast<uint64_t *>(tbase_.ensure((T + 1) * 8));
        uint64_t *stmp = static_cast<uint64_t *>(scan_tmp_.ensure(scan_temp_elems(T + 1) * 8));
        k_fill_u64<<<grid_for(n), DT, 0, s>>>(ends, n, NO_END);
        KITE_CHECK_LAUNCH();
        k_dvle_count<W><<<dim3((unsigned)T), DT, 0, s>>>(d_payload, druns, dtile, tcnt);
        KITE_CHECK_LAUNCH();
        scan_to_u64(tcnt, tbase, T, stmp, s);
        k_dvle_ends<W><<<dim3((unsigned)T), DT, 0, s>>>(d_payload, druns, dtile, tbase, ends);
        KITE_CHECK_LAUNCH();
    }
    if (n) {
        k_dump_rows<W><<<grid_for(n), DT, 0, s>>>(d_payload, druns, nr, n, ends, trunc32 ? 1u : 0u, vals, cnts,
                                                  derr);
        KITE_CHECK_LAUNCH();
        if (!ptiles.empty()) {
            const uint64_t NPT = ptiles.size();
            PrefixTile *dpt = static_cast<PrefixTile *>(ptile_.ensure(NPT * sizeof(PrefixTile)));
            DumpRun *dtr = static_cast<DumpRun *>(ptrun_.ensure(truns.size() * sizeof(DumpRun) + truns.size() * 4));
            uint32_t *dtid = reinterpret_cast<uint32_t *>(dtr + truns.size());
            uint64_t *incl = static_cast<uint64_t *>(psum_.ensure(NPT * 2 * 8 * W));
            uint64_t *own = incl + NPT * W;
            KITE_HIP(hipMemcpyAsync(dpt, ptiles.data(), NPT * sizeof(PrefixTile), hipMemcpyHostToDevice, s));
            KITE_HIP(hipMemcpyAsync(dtr, truns.data(), truns.size() * sizeof(DumpRun), hipMemcpyHostToDevice, s));
            KITE_HIP(hipMemcpyAsync(dtid, tids.data(), tids.size() * 4, hipMemcpyHostToDevice, s));
            k_dvle_tile_sum<W><<<dim3((unsigned)NPT), DT, 0, s>>>(dpt, vals, incl, own);
            KITE_CHECK_LAUNCH();
            k_dvle_prefix<W><<<dim3((unsigned)truns.size()), DT, 0, s>>>(dtr, dtid, incl);
            KITE_CHECK_LAUNCH();
            k_dvle_tile_apply<W><<<dim3((unsigned)NPT), DT, 0, s>>>(dpt, incl, own, vals);
            KITE_CHECK_LAUNCH();
        }
    }
    KITE_HIP(hipMemcpyAsync(host_err_, derr, 4, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipEventRecord(ev[1], s));
    KITE_HIP(hipStreamSynchronize(s));
    if (*host_err_ & ERR_TRUNCATED) throw Error("DVLE payload ends before its rows");
    if (*host_err_ & ERR_VARINT) throw Error("invalid DVLE varint");
    if (*host_err_ & ERR_ORDER) throw Error("256-bit run not in ascending key order");

    // merge the runs of partitions that have several
    const uint64_t *keys = vals, *counts = cnts;
    uint64_t rows = n;
    if (R > 1) {
        std::vector<std::vector<uint64_t>> off(R, std::vector<uint64_t>(np + 1, 0));
        for (const DumpRun &r : runs) off[r.table][r.part + 1] += r.rows;
        for (uint32_t t = 0; t < R; t++)
            for (uint32_t p = 0; p < np; p++) off[t][p + 1] += off[t][p];
        for (DumpRun &r : runs) r.dst = off[r.table][r.part];
        KITE_HIP(hipMemcpyAsync(druns, runs.data(), nr * sizeof(DumpRun), hipMemcpyHostToDevice, s));
        if (tabs_.size() != R) {
            std::vector<DevTable> fresh(R);  // DevTable owns device memory and does not move
            tabs_.swap(fresh);
        }
        std::vector<uint64_t *> kp(R), cp(R);
        for (uint32_t t = 0; t < R; t++) {
            DevTable &d = tabs_[t];
            d.rows = off[t][np];
            kp[t] = static_cast<uint64_t *>(d.keys.ensure(std::max<uint64_t>(d.rows, 1) * 8 * W));
            cp[t] = static_cast<uint64_t *>(d.counts.ensure(std::max<uint64_t>(d.rows, 1) * 8));
            d.off.ensure((np + 1) * 8);
            KITE_HIP(hipMemcpyAsync(d.off.ptr, off[t].data(), (np + 1) * 8, hipMemcpyHostToDevice, s));
        }
        uint64_t **tp = static_cast<uint64_t **>(tabptr_.ensure(2 * R * sizeof(uint64_t *)));
        KITE_HIP(hipMemcpyAsync(tp, kp.data(), R * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
        KITE_HIP(hipMemcpyAsync(tp + R, cp.data(), R * sizeof(uint64_t *), hipMemcpyHostToDevice, s));
        k_dump_gather<W><<<grid_for(n), DT, 0, s>>>(druns, nr, n, vals, cnts, tp, tp + R);
        KITE_CHECK_LAUNCH();
        KITE_HIP(hipStreamSynchronize(s));
        merge_->merge_all(tabs_, np, W, merged_);
        keys = merged_.keys.as<uint64_t>();
        counts = merged_.counts.as<uint64_t>();
        rows = merged_.rows;
    }
    KITE_HIP(hipEventRecord(ev[2], s));
    *keys_out = keys;
    *counts_out = counts;
    return rows;
}

template <int W>
uint64_t DumpEngine::run_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs_in, uint32_t np, int k,
                           const DumpFilter &f, uint64_t *rows_out, uint64_t *lines_out, DumpTimes *times) {
    hipStream_t s = stream_;
    Events ev;
    const uint64_t *keys, *counts;
    // dump reads every stored count as u32 (src/dump/merge.rs:183, 193)
    const uint64_t rows = decode_merge_w<W>(d_payload, runs_in, np, true, &keys, &counts, ev.e);

    // lines
    const uint64_t nt = ceil_div(rows, DT);
    uint32_t *tlen = static_cast<uint32_t *>(tlen_.ensure(std::max<uint64_t>(nt, 1) * 4));
    uint32_t *tl = static_cast<uint32_t *>(lines_.ensure(std::max<uint64_t>(nt, 1) * 4));
    uint64_t *toff = static_cast<uint64_t *>(toff_.ensure((nt + 1) * 8));
    uint64_t *stmp = static_cast<uint64_t *>(scan_tmp_.ensure(scan_temp_elems(nt + 1) * 8));
    uint64_t *lo = static_cast<uint64_t *>(loff_.ensure((nt + 1) * 8));
    if (nt) {
        k_tsv_tile_len<<<dim3((unsigned)nt), DT, 0, s>>>(counts, rows, k, f, tlen, tl);
        KITE_CHECK_LAUNCH();
    }
    scan_to_u64(tlen, toff, nt, stmp, s);
    scan_to_u64(tl, lo, nt, stmp, s);
    uint64_t hv[2] = {0, 0};
    KITE_HIP(hipMemcpyAsync(&hv[0], toff + nt, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipMemcpyAsync(&hv[1], lo + nt, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    const uint64_t bytes = hv[0];
    char *text = static_cast<char *>(text_.ensure(bytes + 16));
    if (nt) {
        k_tsv_write<W><<<dim3((unsigned)nt), DT, 0, s>>>(keys, counts, rows, k, f, toff, text);
        KITE_CHECK_LAUNCH();
    }
    KITE_HIP(hipEventRecord(ev.e[3], s));
    KITE_HIP(hipStreamSynchronize(s));
    if (times) {
        KITE_HIP(hipEventElapsedTime(&times->decode_ms, ev.e[0], ev.e[1]));
        KITE_HIP(hipEventElapsedTime(&times->merge_ms, ev.e[1], ev.e[2]));
        KITE_HIP(hipEventElapsedTime(&times->format_ms, ev.e[2], ev.e[3]));
    }
    if (rows_out) *rows_out = rows;
    if (lines_out) *lines_out = hv[1];
    return bytes;
}

uint64_t DumpEngine::decode_sum(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int W,
                                const uint64_t **keys, const uint64_t **counts, const uint64_t **off,
                                DumpTimes *times) {
    if (W == 1) return decode_sum_w<1>(d_payload, runs, np, keys, counts, off, times);
    if (W == 2) return decode_sum_w<2>(d_payload, runs, np, keys, counts, off, times);
    return decode_sum_w<4>(d_payload, runs, np, keys, counts, off, times);
}

template <int W>
uint64_t DumpEngine::decode_sum_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np,
                                  const uint64_t **keys, const uint64_t **counts, const uint64_t **off,
                                  DumpTimes *times) {
    hipStream_t s = stream_;
    Events ev;
    // stored counts are read as u64 and summed with saturation
    const uint64_t rows = decode_merge_w<W>(d_payload, runs, np, false, keys, counts, ev.e);
    uint32_t R = 0;
    for (const DumpRun &r : runs)
        if (r.rows) R = std::max(R, r.table + 1);
    if (R > 1) {
        *off = merged_.off.as<uint64_t>();  // the merge of the runs wrote the offsets
    } else {
        // one run per partition: its offsets follow from the run sizes
        std::vector<uint64_t> h((size_t)np + 1, 0);
        for (const DumpRun &r : runs) h[r.part + 1] += r.rows;
        for (uint32_t p = 0; p < np; p++) h[p + 1] += h[p];
        uint64_t *d = static_cast<uint64_t *>(poff_.ensure(((uint64_t)np + 1) * 8));
        KITE_HIP(hipMemcpyAsync(d, h.data(), ((uint64_t)np + 1) * 8, hipMemcpyHostToDevice, s));
        *off = d;
    }
    KITE_HIP(hipStreamSynchronize(s));  // (h is a local)
    if (times) {
        KITE_HIP(hipEventElapsedTime(&times->decode_ms, ev.e[0], ev.e[1]));
        KITE_HIP(hipEventElapsedTime(&times->merge_ms, ev.e[1], ev.e[2]));
        times->format_ms = 0;
    }
    return rows;
}

void DumpEngine::histo_begin() {
    uint64_t *b = static_cast<uint64_t *>(hsmall_.ensure(HISTO_SMALL * 8));
    KITE_HIP(hipMemsetAsync(b, 0, HISTO_SMALL * 8, stream_));
    hbig_.clear();
}

void DumpEngine::histo(const uint8_t *d_payload, const std::vector<DumpRun> &runs, uint32_t np, int W,
                       const DumpFilter &f, uint64_t *rows_out, DumpTimes *times) {
    if (W == 1) histo_w<1>(d_payload, runs, np, f, rows_out, times);
    else if (W == 2) histo_w<2>(d_payload, runs, np, f, rows_out, times);
    else histo_w<4>(d_payload, runs, np, f, rows_out, times);
}

template <int W>
void DumpEngine::histo_w(const uint8_t *d_payload, const std::vector<DumpRun> &runs_in, uint32_t np,
                         const DumpFilter &f, uint64_t *rows_out, DumpTimes *times) {
    hipStream_t s = stream_;
    Events ev;
    const uint64_t *keys, *counts;
    // histo reads stored counts as u64 (src/kiteops/mod.rs:475-585)
    const uint64_t rows = decode_merge_w<W>(d_payload, runs_in, np, false, &keys, &counts, ev.e);
    uint64_t *small = hsmall_.as<uint64_t>();
    uint64_t *big = static_cast<uint64_t *>(hbig_d_.ensure((std::max<uint64_t>(rows, 1) + 1) * 8));
    KITE_HIP(hipMemsetAsync(big, 0, 8, s));
    if (rows) {
        const uint64_t nwg = std::min<uint64_t>(ceil_div(rows, DT * 16), 4096);
        k_histo<<<dim3((unsigned)nwg), DT, 0, s>>>(counts, rows, f, small, big);
        KITE_CHECK_LAUNCH();
    }
    uint64_t nbig = 0;
    KITE_HIP(hipMemcpyAsync(&nbig, big, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipEventRecord(ev.e[3], s));
    KITE_HIP(hipStreamSynchronize(s));
    if (nbig) {
        std::vector<uint64_t> v(nbig);
        KITE_HIP(hipMemcpyAsync(v.data(), big + 1, nbig * 8, hipMemcpyDeviceToHost, s));
        KITE_HIP(hipStreamSynchronize(s));
        for (uint64_t x : v) hbig_[x]++;
    }
    if (times) {
        KITE_HIP(hipEventElapsedTime(&times->decode_ms, ev.e[0], ev.e[1]));
        KITE_HIP(hipEventElapsedTime(&times->merge_ms, ev.e[1], ev.e[2]));
        KITE_HIP(hipEventElapsedTime(&times->format_ms, ev.e[2], ev.e[3]));
    }
    if (rows_out) *rows_out = rows;
}

void DumpEngine::histo_end(std::vector<uint64_t> &values, std::vector<uint64_t> &freqs) {
    std::vector<uint64_t> small(HISTO_SMALL);
    KITE_HIP(hipMemcpyAsync(small.data(), hsmall_.ptr, HISTO_SMALL * 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
    values.clear();
    freqs.clear();
    for (uint64_t v = 0; v < HISTO_SMALL; v++)
        if (small[v]) {
            values.push_back(v);
            freqs.push_back(small[v]);
        }
    for (const auto &kv : hbig_) {  // ascending, all >= HISTO_SMALL
        values.push_back(kv.first);
        freqs.push_back(kv.second);
    }
}

template uint64_t DumpEngine::run_w<1>(const uint8_t *, const std::vector<DumpRun> &, uint32_t, int,
                                       const DumpFilter &, uint64_t *, uint64_t *, DumpTimes *);
template uint64_t DumpEngine::run_w<2>(const uint8_t *, const std::vector<DumpRun> &, uint32_t, int,
                                       const DumpFilter &, uint64_t *, uint64_t *, DumpTimes *);
template uint64_t DumpEngine::run_w<4>(const uint8_t *, const std::vector<DumpRun> &, uint32_t, int,
                                       const DumpFilter &, uint64_t *, uint64_t *, DumpTimes *);

}  // namespace kite
