// merge_host.cpp -- see merge_host.hpp.  The device engine hands every
// input's stored runs to kite_gpu_merge_runs; the host engine takes the
// partitions in batches sized to a host memory budget, one pass of the
// worker threads over each batch.
#include "merge_host.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>

#include "kite_gpu.h"

namespace kh {

void collect_runs(const Reader &r, std::vector<kite_dump_run> &runs, std::vector<std::vector<uint8_t>> &decoded) {
    const Header &h = r.header();
    const int W = r.words();
    runs.reserve(r.chunk_count());
    for (uint32_t p = 0; p < h.partitions; p++) {
        auto [first, cnt] = r.partition_chunks(p);
        for (uint64_t c = first; c < first + cnt; c++) {
            const ChunkMeta &m = r.chunks()[c];
            kite_dump_run x{};
            x.partition = p;
            x.rows = m.rows;
            x.codec = (W == 4 || !(m.codec & CODEC_DVLE)) ? KITE_CODEC_RAW : KITE_CODEC_DVLE;
            if (m.codec & CODEC_ZSTD) {
                decoded.push_back(r.payload(m));
                x.data = decoded.back().data();
                x.bytes = decoded.back().size();
            } else {
                x.data = r.raw_payload(m);
                x.bytes = m.encoded;
            }
            runs.push_back(x);
        }
    }
}

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

namespace {

// The device engine: every input's stored runs go to kite_gpu_merge_runs,
// which returns the merged partitions as chunk payloads in partition order.
MergeStats merge_files_gpu(const std::vector<std::unique_ptr<Reader>> &rs, Writer &w) {
    MergeStats ms;
    ms.engine = "gpu";
    kite_gpu_ctx *ctx = nullptr;
    if (kite_gpu_open(0, 0, &ctx)) throw Error(std::string("failed to open the GPU engine: ") + kite_gpu_last_error());
    struct Closer {
        kite_gpu_ctx *c;
        ~Closer() { kite_gpu_close(c); }
    } closer{ctx};
    const size_t n = rs.size();
    std::vector<std::vector<kite_dump_run>> runs(n);
    std::vector<std::vector<std::vector<uint8_t>>> decoded(n);
    std::vector<kite_merge_input> in(n);
    for (size_t i = 0; i < n; i++) {
        collect_runs(*rs[i], runs[i], decoded[i]);
        const Header &h = rs[i]->header();
        in[i] = kite_merge_input{runs[i].data(), runs[i].size(), h.min_count.has_value() ? 1 : 0,
                                 h.min_count.value_or(0), h.max_count.has_value() ? 1 : 0, h.max_count.value_or(0)};
    }
    struct Sink {
        Writer *w;
        std::string err;
    } sink{&w, {}};
    auto append = [](void *user, uint32_t partition, uint32_t codec, uint64_t rows, uint64_t raw_bytes,
                     const uint64_t key_min[4], const uint64_t key_max[4], const uint8_t *data, size_t len) -> int {
        Sink *s = static_cast<Sink *>(user);
        try {
            s->w->append_chunk(partition, codec, rows, raw_bytes, key_min, key_max, data, len);
        } catch (const std::exception &e) {
            s->err = e.what();
            return 1;
        }
        return 0;
    };
    const Header &f = rs[0]->header();
    kite_merge_stats st{};
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = kite_gpu_merge_runs(ctx, f.key_width, f.partitions, in.data(), (uint32_t)n, append, &sink, &st);
    ms.merge_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!sink.err.empty()) throw Error(sink.err);
    if (rc == KITE_E_OVERFLOW) throw Error(kite_gpu_last_error());
    if (rc) throw Error(std::string("GPU merge failed: ") + kite_gpu_last_error());
    fprintf(stderr,
            "[kite merge] gpu: %llu runs, %llu rows in, %llu kept, %llu merged, %llu chunk bytes in %llu group(s); "
            "in %.1f ms, device decode %.1f ms, merge %.1f ms, encode %.1f ms, out %.1f ms\n",
            (unsigned long long)st.runs, (unsigned long long)st.rows_in, (unsigned long long)st.rows_kept,
            (unsigned long long)st.rows, (unsigned long long)st.bytes, (unsigned long long)st.groups, st.in_ms,
            st.decode_ms, st.merge_ms, st.encode_ms, st.out_ms);
    return ms;
}

}  // namespace

MergeStats merge_files(const std::vector<std::unique_ptr<Reader>> &rs, Writer &w, unsigned threads, bool gpu) {
    if (gpu) return merge_files_gpu(rs, w);
    MergeStats ms;
    ms.engine = "cpu";
    const uint32_t P = rs[0]->partitions();
    const int W = rs[0]->words();
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
        for (size_t i = 0; i < n; i++) load_filtered(*rs[i], p0, p1, threads, K[i], C[i], O[i]);
        std::vector<uint64_t> ok, oc, oo(np + 1);
        auto t0 = std::chrono::steady_clock::now();
        merge_batch_host(K, C, O, np, W, threads, ok, oc, oo);
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
