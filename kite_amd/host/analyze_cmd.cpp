// analyze_cmd.cpp -- `kite analyze`: union of the inputs' (filtered) counts
// and the Pearson / Spearman correlation and PCA tables, computed on the GPU
// (run_analyze, src/analyze/mod.rs:847-1326).  Files written:
// exact_count_matrix.tsv (:969-981, :1539-1550), sample_summary.tsv
// (:1236-1245), correlation_<method>.tsv (:825-845, :1097-1123) and
// embedding_pca.tsv (:1720-1753).  Kendall, cosine and the distance tables,
// the heatmap files, t-SNE / UMAP / PaCMAP and the Quarto report are not
// produced by this build.
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>

#include "cli.hpp"
#include "commands.hpp"
#include "kite_gpu.h"
#include "kitefmt.hpp"
#include "merge_host.hpp"

namespace kh {

namespace {
void write_square(const std::string &path, const std::vector<std::string> &labels, const std::vector<double> &m) {
    FILE *o = fopen(path.c_str(), "wb");
    if (!o) throw Error("failed to create " + path);
    fprintf(o, "sample");
    for (auto &l : labels) fprintf(o, "\t%s", l.c_str());
    fprintf(o, "\n");
    const size_t n = labels.size();
    for (size_t i = 0; i < n; i++) {
        fprintf(o, "%s", labels[i].c_str());
        for (size_t j = 0; j < n; j++) fprintf(o, "\t%.8f", m[i * n + j]);
        fprintf(o, "\n");
    }
    fclose(o);
}
}  // namespace

int run_analyze(int argc, char **argv) {
    Args a({{"input", 'i', true, true}, {"output-dir", 'o'}, {"label", 0, true, true}, {"group", 0, true, true},
            {"correlation", 0}, {"embedding", 0}, {"embedding-input", 0}, {"embedding-input-dims", 0},
            {"heatmap-kmeans", 0}, {"seed", 0}, {"heatmap-seed", 0}, {"cluster-samples", 0},
            {"tsne-perplexity", 0}, {"tsne-learning-rate", 0}, {"tsne-iterations", 0}, {"umap-neighbors", 0},
            {"umap-min-dist", 0}, {"umap-spread", 0}, {"pacmap-neighbors", 0}, {"pacmap-mn-end", 0},
            {"threads", 't'}, {"progress", 'p', false}, {"test-mode", 0, false}, {"fp64-check", 0, false}},
           argc, argv, 2);
    print_command_line("analyze", argc, argv);
    auto inputs = a.list("input");
    if (inputs.empty() || !a.has("output-dir"))
        throw UsageError("the following required arguments were not provided:\n  --input <FILE>...\n  --output-dir <DIR>");
    const std::string corr = a.get("correlation", "pearson");
    const std::string emb = a.get("embedding", "pca");
    static const char *cm[] = {"all", "pearson", "spearman", "kendall", "cosine", "euclidean", "manhattan", "mahalanobis"};
    static const char *em[] = {"all", "pca", "tsne", "umap", "pacmap", "pcamap"};
    if (std::find_if(std::begin(cm), std::end(cm), [&](const char *x) { return corr == x; }) == std::end(cm))
        throw UsageError("invalid value '" + corr + "' for '--correlation <MODE>'");
    if (std::find_if(std::begin(em), std::end(em), [&](const char *x) { return emb == x; }) == std::end(em))
        throw UsageError("invalid value '" + emb + "' for '--embedding <MODE>'");
    if (a.get_u64("embedding-input-dims", 50) == 0) throw Error("--embedding-input-dims must be >= 1");
    uint32_t methods = 0;
    if (corr == "pearson" || corr == "all") methods |= KITE_CORR_PEARSON;
    if (corr == "spearman" || corr == "all") methods |= KITE_CORR_SPEARMAN;
    if (!methods)
        throw Error("correlation mode '" + corr + "' is not provided by the GPU engine (pearson, spearman, all)");
    if (corr == "all")
        fprintf(stderr, "[kite analyze] correlation=all: writing pearson and spearman (other methods are not provided)\n");
    const bool want_pca = emb == "pca" || emb == "all";
    if (emb != "pca")
        fprintf(stderr, "[kite analyze] embedding=%s: only the PCA table is produced\n", emb.c_str());

    const std::string out_dir = a.get("output-dir");
    make_parent_dirs(out_dir + "/x");
    auto labels = a.list("label");
    if (!labels.empty() && labels.size() != inputs.size())
        throw Error("--label count (" + std::to_string(labels.size()) + ") must match input count (" +
                    std::to_string(inputs.size()) + ")");
    if (labels.empty())
        for (size_t i = 0; i < inputs.size(); i++) labels.push_back(infer_label(inputs[i], i));
    auto groups = a.list("group");
    if (!groups.empty() && groups.size() != inputs.size())
        throw Error("--group count (" + std::to_string(groups.size()) + ") must match input count (" +
                    std::to_string(inputs.size()) + ")");
    if (groups.empty()) groups = labels;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::unique_ptr<Reader>> rs;
    for (auto &p : inputs) rs.emplace_back(new Reader(p));
    {
        const Header &f = rs[0]->header();
        for (size_t i = 1; i < rs.size(); i++) {
            const Header &h = rs[i]->header();
            if (h.k != f.k || (h.canonical != 0) != (f.canonical != 0) || h.partitions != f.partitions ||
                h.key_width != f.key_width)
                throw Error("incompatible .kite inputs: input #" + std::to_string(i + 1) + " (" + inputs[i] +
                            ") does not match input #1 (" + inputs[0] + ") in k, canonical, partitions or key width");
        }
    }
    const unsigned nt = worker_threads((unsigned)a.get_u64("threads", 0));
    const uint32_t P = rs[0]->partitions();
    const int W = rs[0]->words(), k = rs[0]->header().k;
    const size_t n = rs.size();
    std::vector<std::vector<uint64_t>> K(n), C(n), O(n);
    uint64_t sum_rows = 0;
    for (size_t i = 0; i < n; i++) {
        load_filtered(*rs[i], 0, P, nt, K[i], C[i], O[i]);
        sum_rows += O[i][P];
    }
    std::vector<const uint64_t *> kp(n), cp(n), op(n);
    for (size_t i = 0; i < n; i++) {
        kp[i] = K[i].data();
        cp[i] = C[i].data();
        op[i] = O[i].data();
    }
    std::vector<uint64_t> uk(std::max<uint64_t>(sum_rows, 1) * W), uo(P + 1);
    std::vector<uint32_t> um(std::max<uint64_t>(sum_rows, 1) * n);
    std::vector<double> pear(n * n), spear(n * n), cov(n * n), p64(n * n), s64(n * n);
    kite_gpu_ctx *ctx = nullptr;
    if (kite_gpu_open(0, 0, &ctx)) throw Error(std::string("failed to open the GPU engine: ") + kite_gpu_last_error());
    struct Closer {
        kite_gpu_ctx *c;
        ~Closer() { kite_gpu_close(c); }
    } closer{ctx};
    kite_corr_stats st{};
    const uint32_t flags = a.flag("fp64-check") ? KITE_CORR_FP64_BASELINE : 0;
    if (kite_gpu_analyze_tables(ctx, (uint32_t)W, P, (uint32_t)n, kp.data(), cp.data(), op.data(),
                                methods | KITE_CORR_PEARSON, flags, uk.data(), um.data(), uo.data(), pear.data(),
                                spear.data(), cov.data(), p64.data(), s64.data(), &st))
        throw Error(std::string("GPU analyze failed: ") + kite_gpu_last_error());
    const uint64_t N = uo[P];
    // a cell the u32 matrix saturates prints the input's exact u64 sum
    // (src/kiteops/mod.rs:433-470 sums in u64): keep the loaded rows only
    // when some input has such a count
    bool wide = false;
    for (auto &c : C)
        for (uint64_t x : c) wide |= x > 0xFFFFFFFFull;
    if (!wide) {
        for (auto &v : K) std::vector<uint64_t>().swap(v);
        for (auto &v : C) std::vector<uint64_t>().swap(v);
    }
    auto cell = [&](uint64_t r, uint32_t p, size_t j) -> uint64_t {
        const uint32_t c = um[r * n + j];
        if (c != 0xFFFFFFFFu || !wide) return c;
        return exact_count(K[j], C[j], O[j][p], O[j][p + 1], &uk[r * W], W);
    };
    fprintf(stderr, "[kite analyze] union rows=%llu, samples=%zu, gram %.3f ms (all methods, one pass), codes=%u B\n",
            (unsigned long long)N, n, st.gram_ms_total, st.code_bytes);
    if (flags) {
        double ep = 0, es = 0;
        for (size_t i = 0; i < n * n; i++) {
            ep = std::max(ep, std::fabs(pear[i] - p64[i]));
            es = std::max(es, std::fabs(spear[i] - s64[i]));
        }
        fprintf(stderr, "[kite analyze] fp64 check: max|dr| pearson=%.3g spearman=%.3g, fp64 gram %.3f ms\n", ep,
                es, st.fp64_ms_total);
    }
    // exact_count_matrix.tsv, formatted in parallel slabs and written in order
    {
        FILE *o = fopen((out_dir + "/exact_count_matrix.tsv").c_str(), "wb");
        if (!o) throw Error("failed to create " + out_dir + "/exact_count_matrix.tsv");
        setvbuf(o, nullptr, _IOFBF, 16 << 20);
        fprintf(o, "kmer");
        for (auto &l : labels) fprintf(o, "\t%s", l.c_str());
        fprintf(o, "\n");
        const uint64_t slab = 1 << 18;
        const uint64_t nslab = (N + slab - 1) / slab;
        std::vector<std::string> buf(nslab);
        std::vector<std::atomic<int>> ready(nslab);
        for (auto &x : ready) x = 0;
        std::atomic<uint64_t> next{0};
        auto work = [&] {
            for (;;) {
                uint64_t s = next++;
                if (s >= nslab) return;
                std::string &b = buf[s];
                uint64_t r0 = s * slab, r1 = std::min(N, r0 + slab);
                b.resize((r1 - r0) * (k + 1 + n * 21));
                char *w = b.data();
                uint32_t p = (uint32_t)(std::upper_bound(uo.begin(), uo.end(), r0) - uo.begin()) - 1;
                for (uint64_t r = r0; r < r1; r++) {
                    while (r >= uo[p + 1]) p++;
                    Key key;
                    for (int q = 0; q < W; q++) key.w[q] = uk[r * W + q];
                    decode_key(key, W, k, w);
                    w += k;
                    for (size_t j = 0; j < n; j++) w += sprintf(w, "\t%llu", (unsigned long long)cell(r, p, j));
                    *w++ = '\n';
                }
                b.resize((size_t)(w - b.data()));
                ready[s] = 1;
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) th.emplace_back(work);
        for (uint64_t s = 0; s < nslab; s++) {
            while (!ready[s]) std::this_thread::sleep_for(std::chrono::microseconds(50));
            fwrite(buf[s].data(), 1, buf[s].size(), o);
            std::string().swap(buf[s]);
        }
        for (auto &x : th) x.join();
        fclose(o);
    }
    {
        FILE *o = fopen((out_dir + "/sample_summary.tsv").c_str(), "wb");
        if (!o) throw Error("failed to create sample_summary.tsv");
        fprintf(o, "sample\ttotal_raw_counts\tnonzero_kmers\n");
        for (size_t j = 0; j < n; j++) {
            uint64_t tot = 0, nz = 0;
            uint32_t p = 0;
            for (uint64_t r = 0; r < N; r++) {
                while (r >= uo[p + 1]) p++;
                const uint64_t c = cell(r, p, j);
                uint64_t s = tot + c;
                tot = s < tot ? ~0ull : s;
                nz += c > 0;
            }
            fprintf(o, "%s\t%llu\t%llu\n", labels[j].c_str(), (unsigned long long)tot, (unsigned long long)nz);
        }
        fclose(o);
    }
    if (methods & KITE_CORR_PEARSON) write_square(out_dir + "/correlation_pearson.tsv", labels, pear);
    if (methods & KITE_CORR_SPEARMAN) write_square(out_dir + "/correlation_spearman.tsv", labels, spear);
    if (want_pca) {
        const uint32_t dims = (uint32_t)std::min<size_t>(n, 10);
        std::vector<double> sc(n * std::max<uint32_t>(dims, 2), 0.0), ev(n);
        const uint32_t d2 = std::max<uint32_t>(dims, 2);
        if (n <= 1024) {
            if (kite_gpu_pca(ctx, cov.data(), (uint32_t)n, d2, sc.data(), ev.data()))
                throw Error(std::string("GPU PCA failed: ") + kite_gpu_last_error());
        } else {
            fprintf(stderr, "[kite analyze] warning: PCA is computed for up to 1024 samples; embedding left at 0\n");
        }
        FILE *o = fopen((out_dir + "/embedding_pca.tsv").c_str(), "wb");
        if (!o) throw Error("failed to create embedding_pca.tsv");
        fprintf(o, "sample\tgroup\tmethod\tdim1\tdim2\n");
        for (size_t i = 0; i < n; i++)
            fprintf(o, "%s\t%s\tpca\t%.8f\t%.8f\n", labels[i].c_str(), groups[i].c_str(), sc[i * d2 + 0],
                    sc[i * d2 + 1]);
        fclose(o);
        fprintf(stderr, "[kite analyze] PCA eigenvalues (top %u):", dims);
        for (uint32_t d = 0; d < dims; d++) fprintf(stderr, " %.6g", ev[d]);
        fprintf(stderr, "\n");
    }
    fprintf(stderr, "[kite analyze] done in %.3fs (output_dir=%s)\n",
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), out_dir.c_str());
    return 0;
}

}  // namespace kh
