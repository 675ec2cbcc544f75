// merge.hip -- sorted union of per-partition (key, count) tables on the GPU.
//
// Semantics of merge_vectors_total / for_each_union_row
// (src/kiteops/mod.rs:587-802): inputs are per partition ascending and
// unique; the union is ascending and unique; equal keys sum with u64
// saturation.  Two tables are combined by a merge path (each thread finds
// its split of the two runs by binary search on its output diagonal, ties
// take the first table, then merges its items), then equal neighbours are
// folded by a flag / scan / compact pass.  N tables merge as a tournament.
// The union count matrix places every input row by binary search in the
// union of its partition.
#include <algorithm>
#include <vector>

#include "merge.hpp"

namespace kite {

constexpr int MT = 256;   // merge threads
constexpr int MIPT = 8;   // outputs per thread
constexpr int MTILE = MT * MIPT;

template <int W>
__device__ __forceinline__ bool le_at(const uint64_t *a, const uint64_t *b) {
#pragma unroll
    for (int r = W - 1; r >= 0; r--)
        if (a[r] != b[r]) return a[r] < b[r];
    return true;
}

template <int W>
__device__ __forceinline__ bool eq_at(const uint64_t *a, const uint64_t *b) {
    bool e = true;
#pragma unroll
    for (int r = 0; r < W; r++) e = e && a[r] == b[r];
    return e;
}

struct MergeArgs {
    const uint64_t *ak, *ac, *aoff;
    const uint64_t *bk, *bc, *boff;
    uint64_t *mk, *mc;
    const uint64_t *tile_pre;  // P + 1 prefix of tiles per partition
    uint32_t P;
};

template <int W>
__global__ __launch_bounds__(MT) void k_merge_path(MergeArgs a) {
    const uint64_t tile = blockIdx.x;
    uint32_t lo = 0, hi = a.P;  // partition: largest p with tile_pre[p] <= tile
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (a.tile_pre[mid] <= tile) lo = mid;
        else hi = mid;
    }
    const uint32_t p = lo;
    const uint64_t a0 = a.aoff[p], na = a.aoff[p + 1] - a0;
    const uint64_t b0 = a.boff[p], nb = a.boff[p + 1] - b0;
    const uint64_t tot = na + nb;
    const uint64_t d = (tile - a.tile_pre[p]) * MTILE + (uint64_t)threadIdx.x * MIPT;
    if (d >= tot) return;
    const uint64_t *A = a.ak + a0 * W, *B = a.bk + b0 * W;
    // i = number of A items among the first d outputs (A first on ties)
    uint64_t l = d > nb ? d - nb : 0, h = d < na ? d : na;
    while (l < h) {
        uint64_t mid = (l + h) / 2;
        if (le_at<W>(A + mid * W, B + (d - mid - 1) * W)) l = mid + 1;
        else h = mid;
    }
    uint64_t i = l, j = d - l;
    const uint64_t o = a0 + b0 + d;
    const uint64_t n = tot - d < (uint64_t)MIPT ? tot - d : (uint64_t)MIPT;
    for (uint64_t q = 0; q < n; q++) {
        bool take_a = j >= nb || (i < na && le_at<W>(A + i * W, B + j * W));
        const uint64_t *src = take_a ? A + i * W : B + j * W;
        uint64_t c = take_a ? a.ac[a0 + i] : a.bc[b0 + j];
#pragma unroll
        for (int r = 0; r < W; r++) a.mk[(o + q) * W + r] = src[r];
        a.mc[o + q] = c;
        if (take_a) i++;
        else j++;
    }
}

template <int W>
__global__ void k_merge_heads(const uint64_t *__restrict__ mk, uint64_t n, uint32_t *__restrict__ head) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || !eq_at<W>(mk + i * W, mk + (i - 1) * W)) ? 1u : 0u;
}

// Two-way folds: a key occurs at most once per table, so at most twice here.
template <int W>
__global__ void k_merge_fold(const uint64_t *__restrict__ mk, const uint64_t *__restrict__ mc, uint64_t n,
                             const uint32_t *__restrict__ head, const uint64_t *__restrict__ pos,
                             uint64_t *__restrict__ ok, uint64_t *__restrict__ oc) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    uint64_t c = mc[i];
    if (i + 1 < n && !head[i + 1]) {
        uint64_t s = c + mc[i + 1];
        c = s < c ? ~0ull : s;
    }
    uint64_t o = pos[i];
#pragma unroll
    for (int r = 0; r < W; r++) ok[o * W + r] = mk[i * W + r];
    oc[o] = c;
}

__global__ void k_merge_partoff(const uint64_t *__restrict__ aoff, const uint64_t *__restrict__ boff,
                                const uint64_t *__restrict__ pos, uint32_t P, uint64_t *__restrict__ ooff) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p <= P) ooff[p] = pos[aoff[p] + boff[p]];
}

// Union matrix: row `r` of input `col` lands at the union position of its key.
struct PlaceArgs {
    const uint64_t *uk, *uoff;   // union keys and partition offsets
    const uint64_t *ik, *ic, *ioff;  // one input table
    uint64_t rows;
    uint32_t P, col, ncols;
    void *matrix;                // union rows x ncols: u32 saturated, or u64 when wide
    uint32_t wide;
};

template <int W>
__global__ void k_union_place(PlaceArgs a) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= a.rows) return;
    uint32_t lo = 0, hi = a.P;  // partition of row r
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (a.ioff[mid] <= r) lo = mid;
        else hi = mid;
    }
    const uint64_t *key = a.ik + r * W;
    uint64_t l = a.uoff[lo], h = a.uoff[lo + 1];
    while (l < h) {
        uint64_t mid = (l + h) / 2;
        if (le_at<W>(a.uk + mid * W, key) && !eq_at<W>(a.uk + mid * W, key)) l = mid + 1;
        else h = mid;
    }
    uint64_t c = a.ic[r];
    if (a.wide) static_cast<uint64_t *>(a.matrix)[l * a.ncols + a.col] = c;
    else static_cast<uint32_t *>(a.matrix)[l * a.ncols + a.col] = c > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)c;
}

__global__ void k_max_u64(const uint64_t *__restrict__ p, uint64_t n, unsigned long long *__restrict__ out) {
    uint64_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = m > p[i] ? m : p[i];
    for (int d = 32; d > 0; d >>= 1) {
        const uint64_t o = __shfl_xor(m, d, 64);
        m = m > o ? m : o;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)m);
}

__global__ void k_zero_u32(uint32_t *p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// 1 for every row whose count lies within the filter's bounds.
__global__ void k_keep_flags(const uint64_t *__restrict__ counts, uint64_t n, DumpFilter f,
                             uint32_t *__restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = counts[i];
    flags[i] = ((!f.has_min || c >= f.min) && (!f.has_max || c <= f.max)) ? 1u : 0u;
}

// Kept rows to their scanned positions.
template <int W>
__global__ void k_keep_rows(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ counts, uint64_t n,
                            const uint32_t *__restrict__ flags, const uint64_t *__restrict__ pos,
                            uint64_t *__restrict__ out_keys, uint64_t *__restrict__ out_counts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flags[i]) return;
    const uint64_t o = pos[i];
#pragma unroll
    for (int r = 0; r < W; r++) out_keys[o * W + r] = keys[i * W + r];
    out_counts[o] = counts[i];
}

// out[p] = pos[off[p]]: partition offsets after the compaction (pos holds
// n + 1 entries, off[p] <= n).
__global__ void k_gather_pos(const uint64_t *__restrict__ off, uint32_t m, const uint64_t *__restrict__ pos,
                             uint64_t *__restrict__ out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) out[p] = pos[off[p]];
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// Counts narrowed to u32; the smallest row index whose count does not fit
// goes to first_bad.
__global__ void k_narrow_u32(const uint64_t *__restrict__ c, uint64_t n, uint32_t *__restrict__ out,
                             unsigned long long *__restrict__ first_bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t v = c[i];
    if (v > 0xFFFFFFFFull) atomicMin(first_bad, (unsigned long long)i);
    out[i] = v > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)v;
}

// ---- host ----------------------------------------------------------------
namespace {
uint32_t blocks(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }
}

void MergeEngine::upload_table(const HostTable &t, DevTable &d, int W) {
    const uint64_t n = t.rows();
    d.rows = n;
    d.keys.ensure(std::max<uint64_t>(n, 1) * 8 * W);
    d.counts.ensure(std::max<uint64_t>(n, 1) * 8);
    d.off.ensure((t.part_off.size()) * 8);
    if (n) {
        KITE_HIP(hipMemcpyAsync(d.keys.ptr, t.keys, n * 8 * W, hipMemcpyHostToDevice, stream_));
        KITE_HIP(hipMemcpyAsync(d.counts.ptr, t.counts, n * 8, hipMemcpyHostToDevice, stream_));
    }
    KITE_HIP(hipMemcpyAsync(d.off.ptr, t.part_off.data(), t.part_off.size() * 8, hipMemcpyHostToDevice, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
}

template <int W>
void MergeEngine::merge2(const DevTable &A, const DevTable &B, DevTable &out, uint32_t P) {
    hipStream_t s = stream_;
    std::vector<uint64_t> ao(P + 1), bo(P + 1);
    KITE_HIP(hipMemcpyAsync(ao.data(), A.off.ptr, (P + 1) * 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipMemcpyAsync(bo.data(), B.off.ptr, (P + 1) * 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    std::vector<uint64_t> tpre(P + 1, 0);
    for (uint32_t p = 0; p < P; p++) {
        uint64_t t = (ao[p + 1] - ao[p]) + (bo[p + 1] - bo[p]);
        tpre[p + 1] = tpre[p] + (t + MTILE - 1) / MTILE;
    }
    const uint64_t n = A.rows + B.rows;
    mk_.ensure(std::max<uint64_t>(n, 1) * 8 * W);
    mc_.ensure(std::max<uint64_t>(n, 1) * 8);
    head_.ensure(std::max<uint64_t>(n, 1) * 4);
    pos_.ensure((n + 1) * 8);
    scan_tmp_.ensure(scan_temp_elems(n + 1) * 8);
    tile_pre_.ensure((P + 1) * 8);
    KITE_HIP(hipMemcpyAsync(tile_pre_.ptr, tpre.data(), (P + 1) * 8, hipMemcpyHostToDevice, s));
    if (tpre[P]) {
        MergeArgs ma{A.keys.as<uint64_t>(), A.counts.as<uint64_t>(), A.off.as<uint64_t>(),
                     B.keys.as<uint64_t>(), B.counts.as<uint64_t>(), B.off.as<uint64_t>(),
                     mk_.as<uint64_t>(), mc_.as<uint64_t>(), tile_pre_.as<uint64_t>(), P};
        k_merge_path<W><<<dim3((unsigned)tpre[P]), dim3(MT), 0, s>>>(ma);
        KITE_CHECK_LAUNCH();
        k_merge_heads<W><<<blocks(n, 256), 256, 0, s>>>(mk_.as<uint64_t>(), n, head_.as<uint32_t>());
        KITE_CHECK_LAUNCH();
    }
    scan_to_u64(head_.as<uint32_t>(), pos_.as<uint64_t>(), n, scan_tmp_.as<uint64_t>(), s);
    uint64_t u = 0;
    KITE_HIP(hipMemcpyAsync(&u, pos_.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    out.rows = u;
    out.keys.ensure(std::max<uint64_t>(u, 1) * 8 * W);
    out.counts.ensure(std::max<uint64_t>(u, 1) * 8);
    out.off.ensure((P + 1) * 8);
    if (n) {
        k_merge_fold<W><<<blocks(n, 256), 256, 0, s>>>(mk_.as<uint64_t>(), mc_.as<uint64_t>(), n,
                                                       head_.as<uint32_t>(), pos_.as<uint64_t>(),
                                                       out.keys.as<uint64_t>(), out.counts.as<uint64_t>());
        KITE_CHECK_LAUNCH();
    }
    k_merge_partoff<<<blocks(P + 1, 256), 256, 0, s>>>(A.off.as<uint64_t>(), B.off.as<uint64_t>(),
                                                       pos_.as<uint64_t>(), P, out.off.as<uint64_t>());
    KITE_CHECK_LAUNCH();
    KITE_HIP(hipStreamSynchronize(s));
}

void MergeEngine::merge_all(std::vector<DevTable> &tabs, uint32_t P, int W, DevTable &result) {
    // tournament: pairwise merges until one table is left
    std::vector<DevTable *> cur;
    for (auto &t : tabs) cur.push_back(&t);
    std::vector<std::unique_ptr<DevTable>> owned;
    if (cur.size() == 1) {
        DevTable &src = *cur[0];
        result.rows = src.rows;
        result.keys.ensure(std::max<uint64_t>(src.rows, 1) * 8 * W);
        result.counts.ensure(std::max<uint64_t>(src.rows, 1) * 8);
        result.off.ensure((P + 1) * 8);
        if (src.rows) {
            KITE_HIP(hipMemcpyAsync(result.keys.ptr, src.keys.ptr, src.rows * 8 * W, hipMemcpyDeviceToDevice, stream_));
            KITE_HIP(hipMemcpyAsync(result.counts.ptr, src.counts.ptr, src.rows * 8, hipMemcpyDeviceToDevice, stream_));
        }
        KITE_HIP(hipMemcpyAsync(result.off.ptr, src.off.ptr, (P + 1) * 8, hipMemcpyDeviceToDevice, stream_));
        KITE_HIP(hipStreamSynchronize(stream_));
        return;
    }
    while (cur.size() > 1) {
        std::vector<DevTable *> nxt;
        for (size_t i = 0; i + 1 < cur.size(); i += 2) {
            DevTable *dst;
            if (cur.size() == 2) {
                dst = &result;
            } else {
                owned.emplace_back(new DevTable());
                dst = owned.back().get();
            }
            if (W == 1) merge2<1>(*cur[i], *cur[i + 1], *dst, P);
            else if (W == 2) merge2<2>(*cur[i], *cur[i + 1], *dst, P);
            else merge2<4>(*cur[i], *cur[i + 1], *dst, P);
            nxt.push_back(dst);
        }
        if (cur.size() % 2) nxt.push_back(cur.back());
        cur.swap(nxt);
    }
}

uint64_t MergeEngine::max_count(const std::vector<DevTable> &tabs) {
    DevBuf m;
    m.ensure(8);
    k_zero_u32<<<1, 64, 0, stream_>>>(m.as<uint32_t>(), 2);
    KITE_CHECK_LAUNCH();
    for (const DevTable &t : tabs)
        if (t.rows) {
            k_max_u64<<<(uint32_t)std::min<uint64_t>(2048, blocks(t.rows, 256)), 256, 0, stream_>>>(
                t.counts.as<uint64_t>(), t.rows, m.as<unsigned long long>());
            KITE_CHECK_LAUNCH();
        }
    uint64_t h = 0;
    KITE_HIP(hipMemcpyAsync(&h, m.ptr, 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
    return h;
}

void MergeEngine::union_matrix(const std::vector<DevTable> &tabs, const DevTable &u, uint32_t P, int W,
                               DevBuf &matrix, int cell_bytes) {
    const uint32_t nc = (uint32_t)tabs.size();
    const uint64_t cells = u.rows * nc;
    const bool wide = cell_bytes == 8;
    matrix.ensure(std::max<uint64_t>(cells, 1) * (wide ? 8 : 4));
    if (cells) {
        const uint64_t words = cells * (wide ? 2 : 1);
        k_zero_u32<<<blocks(words, 256), 256, 0, stream_>>>(matrix.as<uint32_t>(), words);
        KITE_CHECK_LAUNCH();
    }
    for (uint32_t c = 0; c < nc; c++) {
        const DevTable &t = tabs[c];
        if (!t.rows) continue;
        PlaceArgs pa{u.keys.as<uint64_t>(), u.off.as<uint64_t>(), t.keys.as<uint64_t>(), t.counts.as<uint64_t>(),
                     t.off.as<uint64_t>(), t.rows, P, c, nc, matrix.ptr, wide ? 1u : 0u};
        if (W == 1) k_union_place<1><<<blocks(t.rows, 256), 256, 0, stream_>>>(pa);
        else if (W == 2) k_union_place<2><<<blocks(t.rows, 256), 256, 0, stream_>>>(pa);
        else k_union_place<4><<<blocks(t.rows, 256), 256, 0, stream_>>>(pa);
        KITE_CHECK_LAUNCH();
    }
    KITE_HIP(hipStreamSynchronize(stream_));
}

void MergeEngine::download(const DevTable &t, int W, uint32_t P, uint64_t *keys, uint64_t *counts,
                           uint64_t *part_off) {
    if (t.rows && keys)
        KITE_HIP(hipMemcpyAsync(keys, t.keys.ptr, t.rows * 8 * W, hipMemcpyDeviceToHost, stream_));
    if (t.rows && counts)
        KITE_HIP(hipMemcpyAsync(counts, t.counts.ptr, t.rows * 8, hipMemcpyDeviceToHost, stream_));
    if (part_off) KITE_HIP(hipMemcpyAsync(part_off, t.off.ptr, (P + 1) * 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
}

void MergeEngine::filter_into(const uint64_t *keys, const uint64_t *counts, const uint64_t *off, uint64_t n,
                              uint32_t P, int W, const DumpFilter &f, DevTable &out) {
    hipStream_t s = stream_;
    uint64_t *ok = static_cast<uint64_t *>(out.keys.ensure(std::max<uint64_t>(n, 1) * 8 * W));
    uint64_t *oc = static_cast<uint64_t *>(out.counts.ensure(std::max<uint64_t>(n, 1) * 8));
    uint64_t *oo = static_cast<uint64_t *>(out.off.ensure(((uint64_t)P + 1) * 8));
    if (!f.has_min && !f.has_max) {
        if (n) {
            KITE_HIP(hipMemcpyAsync(ok, keys, n * 8 * W, hipMemcpyDeviceToDevice, s));
            KITE_HIP(hipMemcpyAsync(oc, counts, n * 8, hipMemcpyDeviceToDevice, s));
        }
        KITE_HIP(hipMemcpyAsync(oo, off, ((uint64_t)P + 1) * 8, hipMemcpyDeviceToDevice, s));
        out.rows = n;
        return;
    }
    uint32_t *flags = static_cast<uint32_t *>(head_.ensure(std::max<uint64_t>(n, 1) * 4));
    uint64_t *pos = static_cast<uint64_t *>(pos_.ensure((n + 1) * 8));
    uint64_t *tmp = static_cast<uint64_t *>(scan_tmp_.ensure(scan_temp_elems(n + 1) * 8));
    if (n) {
        k_keep_flags<<<blocks(n, 256), 256, 0, s>>>(counts, n, f, flags);
        KITE_CHECK_LAUNCH();
    }
    scan_to_u64(flags, pos, n, tmp, s);
    if (n) {
        if (W == 1) k_keep_rows<1><<<blocks(n, 256), 256, 0, s>>>(keys, counts, n, flags, pos, ok, oc);
        else if (W == 2) k_keep_rows<2><<<blocks(n, 256), 256, 0, s>>>(keys, counts, n, flags, pos, ok, oc);
        else k_keep_rows<4><<<blocks(n, 256), 256, 0, s>>>(keys, counts, n, flags, pos, ok, oc);
        KITE_CHECK_LAUNCH();
    }
    k_gather_pos<<<blocks(P + 1, 256), 256, 0, s>>>(off, P + 1, pos, oo);
    KITE_CHECK_LAUNCH();
    uint64_t kept = 0;
    KITE_HIP(hipMemcpyAsync(&kept, pos + n, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    out.rows = kept;
}

uint64_t MergeEngine::narrow_u32(const uint64_t *counts, uint64_t n, uint32_t *out) {
    hipStream_t s = stream_;
    unsigned long long *bad = static_cast<unsigned long long *>(bad_.ensure(8));
    k_set_u64<<<1, 1, 0, s>>>(bad, (unsigned long long)n);
    KITE_CHECK_LAUNCH();
    if (n) {
        k_narrow_u32<<<blocks(n, 256), 256, 0, s>>>(counts, n, out, bad);
        KITE_CHECK_LAUNCH();
    }
    unsigned long long h = 0;
    KITE_HIP(hipMemcpyAsync(&h, bad, 8, hipMemcpyDeviceToHost, s));
    KITE_HIP(hipStreamSynchronize(s));
    return (uint64_t)h;
}

template void MergeEngine::merge2<1>(const DevTable &, const DevTable &, DevTable &, uint32_t);
template void MergeEngine::merge2<2>(const DevTable &, const DevTable &, DevTable &, uint32_t);
template void MergeEngine::merge2<4>(const DevTable &, const DevTable &, DevTable &, uint32_t);

}  // namespace kite
