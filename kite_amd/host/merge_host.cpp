// merge_host.cpp -- see merge_host.hpp.  Partitions go through in batches
// sized to a host memory budget; each batch is one GPU union call, or with
// the host engine one pass of the worker threads over its partitions.
#include "merge_host.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>

#include "kite_gpu.h"

namespace kh {

void load_filtered(const Reader &r, uint32_t p0, uint32_t p1, unsigned threads, std::vector<uint64_t> &keys,
                   std::vector<uint64_t> &counts, std::vector<uint64_t> &off) {
    const uint32_t n = p1 - p0;
    const int W = r.words();
    const Header &h = r.header();
    std::vector<std::vector<Row>> parts(n);
    std::atomic<uint32_t> next{0};
    std::mutex mu;
    std::string err;
    auto work = [&] {
        for (;;) {
            uint32_t i = next++;
            if (i >= n) return;
            try {
                std::vector<Row> rows = r.load_partition(p0 + i);
                // apply_reader_filter_to_rows (src/kiteops/mod.rs:103-108)
                if (h.min_count || h.max_count) {
                    size_t u = 0;
                    for (auto &x : rows)
                        if (passes(x.count, h.min_count, h.max_count)) rows[u++] = x;
                    rows.resize(u);
                }
                parts[i].swap(rows);
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> g(mu);
                err = e.what();
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < std::max(1u, std::min<unsigned>(threads, n)); t++) th.emplace_back(work);
    for (auto &x : th) x.join();
    if (!err.empty()) throw Error(err);
    off.assign(n + 1, 0);
    for (uint32_t i = 0; i < n; i++) off[i + 1] = off[i] + parts[i].size();
    keys.resize(std::max<uint64_t>(off[n], 1) * W);
    counts.resize(std::max<uint64_t>(off[n], 1));
    for (uint32_t i = 0; i < n; i++) {
        uint64_t o = off[i];
        for (auto &x : parts[i]) {
            for (int w = 0; w < W; w++) keys[o * W + w] = x.key.w[w];
            counts[o++] = x.count;
        }
    }
}

namespace {

// The host engine's union of one batch of partitions: for every partition
// the inputs' ascending rows merged with u64 saturating sums
// (merge_vectors_total, src/kiteops/mod.rs:587-724), partitions spread over
// the worker threads; the same (key, total) rows as the device merge.
void merge_batch_host(const std::vector<std::vector<uint64_t>> &K, const std::vector<std::vector<uint64_t>> &C,
                      const std::vector<std::vector<uint64_t>> &O, uint32_t np, int W, unsigned threads,
                      std::vector<uint64_t> &ok, std::vector<uint64_t> &oc, std::vector<uint64_t> &oo) {
    const size_t n = K.size();
    // each partition writes its rows at the offset the sum of its inputs'
    // rows allows, then the rows are packed in order
    std::vector<uint64_t> cap(np + 1, 0), got(np, 0);
    for (uint32_t i = 0; i < np; i++) {
        uint64_t c = 0;
        for (size_t t = 0; t < n; t++) c += O[t][i + 1] - O[t][i];
        cap[i + 1] = cap[i] + c;
    }
    std::atomic<uint32_t> next{0};
    auto less = [W](const uint64_t *a, const uint64_t *b) {
        for (int q = W - 1; q >= 0; q--)
            if (a[q] != b[q]) return a[q] < b[q];
        return false;
    };
    auto work = [&] {
        std::vector<uint64_t> pos(n);
        for (;;) {
            const uint32_t i = next++;
            if (i >= np) return;
            for (size_t t = 0; t < n; t++) pos[t] = O[t][i];
            uint64_t o = cap[i];
            for (;;) {
                // the smallest head key of the inputs
                const uint64_t *best = nullptr;
                for (size_t t = 0; t < n; t++)
                    if (pos[t] < O[t][i + 1] && (!best || less(&K[t][pos[t] * W], best))) best = &K[t][pos[t] * W];
                if (!best) break;
                uint64_t key[4] = {0, 0, 0, 0};
                for (int q = 0; q < W; q++) key[q] = best[q];
                uint64_t sum = 0;
                for (size_t t = 0; t < n; t++) {
                    if (pos[t] >= O[t][i + 1] || less(key, &K[t][pos[t] * W]) || less(&K[t][pos[t] * W], key)) continue;
                    const uint64_t c = C[t][pos[t]++];
                    sum = sum > UINT64_MAX - c ? UINT64_MAX : sum + c;
                }
                for (int q = 0; q < W; q++) ok[o * W + q] = key[q];
                oc[o++] = sum;
            }
            got[i] = o - cap[i];
        }
    };
    ok.resize(std::max<uint64_t>(cap[np], 1) * W);
    oc.resize(std::max<uint64_t>(cap[np], 1));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < std::max(1u, std::min<unsigned>(threads, np)); t++) th.emplace_back(work);
    for (auto &x : th) x.join();
    oo.assign(np + 1, 0);
    for (uint32_t i = 0; i < np; i++) {
        oo[i + 1] = oo[i] + got[i];
        if (oo[i] != cap[i]) {
            std::copy(ok.begin() + cap[i] * W, ok.begin() + (cap[i] + got[i]) * W, ok.begin() + oo[i] * W);
            std::copy(oc.begin() + cap[i], oc.begin() + cap[i] + got[i], oc.begin() + oo[i]);
        }
    }
}

}  // namespace

MergeStats merge_files(const std::vector<std::unique_ptr<Reader>> &rs, Writer &w, unsigned threads, bool gpu) {
    MergeStats ms;
    ms.engine = gpu ? "gpu" : "cpu";
    const uint32_t P = rs[0]->partitions();
    const int W = rs[0]->words();
    kite_gpu_ctx *ctx = nullptr;
    if (gpu && kite_gpu_open(0, 0, &ctx))
        throw Error(std::string("failed to open the GPU engine: ") + kite_gpu_last_error());
    struct Closer {
        kite_gpu_ctx *c;
        ~Closer() {
            if (c) kite_gpu_close(c);
        }
    } closer{ctx};
    // batch partitions so the decoded rows of all inputs stay near 2 GiB
    uint64_t total_rows = 0;
    for (auto &r : rs) total_rows += r->header().total_rows;
    const uint64_t per_row = 8ull * W + 8;
    const uint64_t budget_rows = std::max<uint64_t>(1, (2ull << 30) / per_row);
    const uint32_t batches = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(P, (total_rows + budget_rows - 1) / budget_rows));
    const uint32_t step = (P + batches - 1) / batches;
    const size_t n = rs.size();
    for (uint32_t p0 = 0; p0 < P; p0 += step) {
        const uint32_t p1 = std::min(P, p0 + step), np = p1 - p0;
        std::vector<std::vector<uint64_t>> K(n), C(n), O(n);
        uint64_t sum = 0;
        for (size_t i = 0; i < n; i++) {
            load_filtered(*rs[i], p0, p1, threads, K[i], C[i], O[i]);
            sum += O[i][np];
        }
        std::vector<const uint64_t *> kp(n), cp(n), op(n);
        for (size_t i = 0; i < n; i++) {
            kp[i] = K[i].data();
            cp[i] = C[i].data();
            op[i] = O[i].data();
        }
        std::vector<uint64_t> ok, oc, oo(np + 1);
        auto t0 = std::chrono::steady_clock::now();
        if (gpu) {
            ok.resize(std::max<uint64_t>(sum, 1) * W);
            oc.resize(std::max<uint64_t>(sum, 1));
            if (kite_gpu_merge_tables(ctx, (uint32_t)W, np, (uint32_t)n, kp.data(), cp.data(), op.data(), ok.data(),
                                      oc.data(), oo.data()))
                throw Error(std::string("GPU merge failed: ") + kite_gpu_last_error());
        } else {
            merge_batch_host(K, C, O, np, W, threads, ok, oc, oo);
        }
        ms.merge_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (auto &v : K) std::vector<uint64_t>().swap(v);
        for (auto &v : C) std::vector<uint64_t>().swap(v);
        // serialize partitions in parallel, append in order
        std::vector<std::vector<uint8_t>> data(np);
        std::vector<uint32_t> codec(np, 0);
        std::atomic<uint32_t> next{0};
        std::mutex mu;
        std::string err;
        auto work = [&] {
            for (;;) {
                uint32_t i = next++;
                if (i >= np) return;
                try {
                    std::vector<Row> rows(oo[i + 1] - oo[i]);
                    for (uint64_t j = oo[i]; j < oo[i + 1]; j++) {
                        Row &x = rows[j - oo[i]];
                        for (int q = 0; q < W; q++) x.key.w[q] = ok[j * W + q];
                        x.count = oc[j];
                    }
                    data[i] = serialize_rows(rows, W, &codec[i]);
                } catch (const std::exception &e) {
                    std::lock_guard<std::mutex> g(mu);
                    err = e.what();
                }
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < std::max(1u, std::min<unsigned>(threads, np)); t++) th.emplace_back(work);
        for (auto &x : th) x.join();
        if (!err.empty()) throw Error(err);
        for (uint32_t i = 0; i < np; i++) {
            uint64_t rows = oo[i + 1] - oo[i];
            if (!rows) continue;
            uint64_t kmin[4] = {0, 0, 0, 0}, kmax[4] = {0, 0, 0, 0};
            for (int q = 0; q < W; q++) {
                kmin[q] = ok[oo[i] * W + q];
                kmax[q] = ok[(oo[i + 1] - 1) * W + q];
            }
            w.append_chunk(p0 + i, codec[i], rows, data[i].size(), kmin, kmax, data[i].data(), data[i].size());
        }
    }
    return ms;
}

// Exact count of `key` in rows [lo, hi) of one input's sorted table.
uint64_t exact_count(const std::vector<uint64_t> &K, const std::vector<uint64_t> &C, uint64_t lo, uint64_t hi,
                     const uint64_t *key, int W) {
    auto less = [&](const uint64_t *a, const uint64_t *b) {
        for (int q = W - 1; q >= 0; q--)
            if (a[q] != b[q]) return a[q] < b[q];
        return false;
    };
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (less(&K[mid * W], key))
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < C.size() && !less(key, &K[lo * W]) && !less(&K[lo * W], key)) return C[lo];
    throw Error("saturated union cell has no matching input row");
}

}  // namespace kh
