// profile_cmd.cpp -- `kite profile`: one TSV row per k-mer of the union of
// the inputs, with each input's (header-filtered, merged) count, 0 where the
// input lacks the k-mer (run_profile / render_partition_rows,
// src/profile/mod.rs:14-169).  The union is built on the GPU
// (kite_gpu_union_tables, the device form of for_each_union_row,
// src/kiteops/mod.rs:625-802); its matrix saturates at u32::MAX, so the few
// cells at that value are re-read as exact u64 sums from the input's own rows
// (load_partition_counts sums with saturating u64 adds, :433-470).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "cli.hpp"
#include "commands.hpp"
#include "kite_gpu.h"
#include "kitefmt.hpp"
#include "merge_host.hpp"

namespace kh {

namespace {
std::string fmt_filter(const Header &h) {
    auto f = [](const std::optional<uint32_t> &v) { return v ? std::to_string(*v) : std::string("none"); };
    return "min=" + f(h.min_count) + ", max=" + f(h.max_count);
}
}  // namespace

int run_profile(int argc, char **argv) {
    Args a({{"input", 'i', true, true}, {"output", 'o'}, {"label", 0, true, true}, {"threads", 't'},
            {"progress", 'p', false}, {"test-mode", 0, false}},
           argc, argv, 2);
    print_command_line("profile", argc, argv);
    const auto t0 = std::chrono::steady_clock::now();
    auto inputs = a.list("input");
    if (inputs.empty() || !a.has("output"))
        throw UsageError("the following required arguments were not provided:\n  --input <FILE>...\n  --output <FILE>");
    const std::string out_path = a.get("output");
    // infer_sample_labels (src/kiteops/mod.rs:123-162)
    auto labels = a.list("label");
    const bool user_labels = !labels.empty();
    if (user_labels && labels.size() != inputs.size()) {
        std::string ls, is;
        for (size_t i = 0; i < labels.size(); i++) ls += (i ? ", " : "") + labels[i];
        for (size_t i = 0; i < inputs.size(); i++) is += (i ? ", " : "") + inputs[i];
        throw Error("--label count (" + std::to_string(labels.size()) + ") must match input count (" +
                    std::to_string(inputs.size()) + "). labels=[" + ls + "], inputs=[" + is + "]");
    }
    if (!user_labels)
        for (size_t i = 0; i < inputs.size(); i++) labels.push_back(infer_label(inputs[i], i));
    fprintf(stderr, "[kite profile] inputs=%zu, output=%s\n", inputs.size(), out_path.c_str());
    std::vector<std::unique_ptr<Reader>> rs;
    for (auto &p : inputs) rs.emplace_back(new Reader(p));
    {
        std::string j;
        for (size_t i = 0; i < labels.size(); i++) j += (i ? "," : "") + labels[i];
        fprintf(stderr, "[kite profile] labels=%s (%s)\n", j.c_str(),
                user_labels ? "user supplied" : "inferred from file names");
    }
    for (size_t i = 0; i < rs.size(); i++) {
        const Header &h = rs[i]->header();
        fprintf(stderr,
                "[kite profile] input #%2zu label=%s path=%s k=%u canonical=%s partitions=%u key_width=%u "
                "rows=%llu chunks=%llu filters: %s\n",
                i + 1, labels[i].c_str(), inputs[i].c_str(), (unsigned)h.k, h.canonical ? "true" : "false",
                (unsigned)h.partitions, (unsigned)h.key_width, (unsigned long long)h.total_rows,
                (unsigned long long)rs[i]->chunk_count(), fmt_filter(h).c_str());
    }
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
    fprintf(stderr, "[kite profile] worker_threads=%u (parallel partitions in batches)\n", nt);
    const uint32_t P = rs[0]->partitions();
    const int W = rs[0]->words(), k = rs[0]->header().k;
    const size_t n = rs.size();

    // create_output_writer (src/kiteops/mod.rs:110-121): "-" is stdout.
    FILE *o = stdout;
    if (out_path != "-") {
        make_parent_dirs(out_path);
        o = fopen(out_path.c_str(), "wb");
        if (!o) throw Error("failed to create " + out_path);
    }
    struct Closer {
        FILE *f;
        ~Closer() {
            if (f && f != stdout) fclose(f);
        }
    } closer{o};
    setvbuf(o, nullptr, _IOFBF, 16 << 20);
    fprintf(o, "kmer");
    for (auto &l : labels) fprintf(o, "\t%s", l.c_str());
    fprintf(o, "\n");

    kite_gpu_ctx *ctx = nullptr;
    if (kite_gpu_open(0, 0, &ctx)) throw Error(std::string("failed to open the GPU engine: ") + kite_gpu_last_error());
    struct CtxCloser {
        kite_gpu_ctx *c;
        ~CtxCloser() { kite_gpu_close(c); }
    } ctx_closer{ctx};
    // Partitions go in groups of about `group_rows` stored rows over all
    // inputs (the reference renders partitions in batches of its worker
    // count), so host memory stays bounded for any input size.
    uint64_t all_rows = 0;
    for (auto &r : rs) all_rows += r->header().total_rows;
    uint64_t group_rows = 64ull << 20;
    if (const char *e = getenv("KITE_PROFILE_GROUP_ROWS")) group_rows = std::max(1ull, strtoull(e, nullptr, 10));
    const uint32_t per_group = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(P, all_rows ? (group_rows * P) / all_rows : P));
    uint64_t N = 0;
    for (uint32_t p0 = 0; p0 < P; p0 += per_group) {
        const uint32_t p1 = std::min(P, p0 + per_group), G = p1 - p0;
        std::vector<std::vector<uint64_t>> K(n), C(n), O(n);
        uint64_t sum_rows = 0;
        for (size_t i = 0; i < n; i++) {
            load_filtered(*rs[i], p0, p1, nt, K[i], C[i], O[i]);
            sum_rows += O[i][G];
        }
        std::vector<const uint64_t *> kp(n), cp(n), op(n);
        for (size_t i = 0; i < n; i++) {
            kp[i] = K[i].data();
            cp[i] = C[i].data();
            op[i] = O[i].data();
        }
        std::vector<uint64_t> uk(std::max<uint64_t>(sum_rows, 1) * W), uo(G + 1);
        std::vector<uint32_t> um(std::max<uint64_t>(sum_rows, 1) * n);
        if (kite_gpu_union_tables(ctx, (uint32_t)W, G, (uint32_t)n, kp.data(), cp.data(), op.data(), uk.data(),
                                  um.data(), uo.data()))
            throw Error(std::string("GPU union failed: ") + kite_gpu_last_error());
        const uint64_t NG = uo[G];
        N += NG;
        // rows formatted in parallel slabs, written in union order (partition
        // then key order, as the reference's batches are)
        const uint64_t slab = 1 << 18;
        const uint64_t nslab = (NG + slab - 1) / slab;
        std::vector<std::string> buf(nslab);
        std::vector<std::atomic<int>> ready(nslab);
        for (auto &x : ready) x = 0;
        std::atomic<uint64_t> next{0};
        std::string werr;
        std::atomic<bool> failed{false};
        auto work = [&] {
            for (;;) {
                const uint64_t sl = next++;
                if (sl >= nslab) return;
                std::string &b = buf[sl];
                const uint64_t r0 = sl * slab, r1 = std::min(NG, r0 + slab);
                try {
                    b.resize((r1 - r0) * (k + 1 + n * 21));
                    char *w = b.data();
                    uint32_t p = (uint32_t)(std::upper_bound(uo.begin(), uo.end(), r0) - uo.begin()) - 1;
                    for (uint64_t r = r0; r < r1; r++) {
                        while (r >= uo[p + 1]) p++;
                        Key key;
                        for (int q = 0; q < W; q++) key.w[q] = uk[r * W + q];
                        decode_key(key, W, k, w);
                        w += k;
                        for (size_t j = 0; j < n; j++) {
                            const uint32_t c = um[r * n + j];
                            if (c != 0xFFFFFFFFu) {
                                w += sprintf(w, "\t%u", c);
                            } else {
                                const uint64_t e =
                                    exact_count(K[j], C[j], O[j][p], O[j][p + 1], &uk[r * W], W);
                                w += sprintf(w, "\t%llu", (unsigned long long)e);
                            }
                        }
                        *w++ = '\n';
                    }
                    b.resize((size_t)(w - b.data()));
                } catch (const std::exception &e) {
                    if (!failed.exchange(true)) werr = e.what();
                    b.clear();
                }
                ready[sl] = 1;
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < std::max(1u, nt); t++) th.emplace_back(work);
        for (uint64_t sl = 0; sl < nslab; sl++) {
            while (!ready[sl]) std::this_thread::sleep_for(std::chrono::microseconds(50));
            if (!failed) fwrite(buf[sl].data(), 1, buf[sl].size(), o);
            std::string().swap(buf[sl]);
        }
        for (auto &x : th) x.join();
        if (failed) throw Error(werr);
    }
    if (fflush(o) != 0 || ferror(o)) throw Error("failed to write " + out_path);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (a.flag("test-mode")) {
        std::string j;
        for (size_t i = 0; i < labels.size(); i++) j += (i ? "," : "") + labels[i];
        fprintf(stderr,
                "[kite profile] test-mode: inputs=%zu labels=%s threads=%llu worker_threads=%u partitions=%u "
                "rows_written=%llu output=%s elapsed=%.3fs\n",
                n, j.c_str(), (unsigned long long)a.get_u64("threads", 0), nt, P, (unsigned long long)N,
                out_path.c_str(), secs);
    }
    fprintf(stderr, "[kite profile] done in %.3fs (rows=%llu, output=%s)\n", secs, (unsigned long long)N,
            out_path.c_str());
    return 0;
}

}  // namespace kh
