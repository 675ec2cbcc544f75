// commands.cpp -- `kite dump`, `histo`, `merge` and `info` on the host.
//
// dump: partitions in order 0..P-1 (the reference's deterministic stdout
//   path, src/dump/mod.rs:293-348); runs of a partition are merged with
//   u32-saturating sums (src/dump/loser_tree.rs:141-175; KmerU256 files:
//   sort + RLE, src/dump/merge.rs:311-371); header filter; "KMER\tCOUNT\n".
// histo: u64 totals per key, header filter, union of bins
//   (src/histo/mod.rs:16-176).
// merge: per-input header filter, sorted union with u64 sums on the GPU
//   merge path, one chunk per partition appended in partition order
//   (src/merge/mod.rs:22-301, src/kiteops/mod.rs:587-859).
// info: header summary and in-place marker edits (src/info/mod.rs:14-176,
//   src/format/header.rs:166-183).
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

#include "cli.hpp"
#include "commands.hpp"
#include "kite_gpu.h"
#include "kitefmt.hpp"
#include "merge_host.hpp"

namespace kh {

void print_command_line(const char *cmd, int argc, char **argv) {
    std::string s;
    const char *exe = strrchr(argv[0], '/');
    s = exe ? exe + 1 : argv[0];
    for (int i = 1; i < argc; i++) {
        std::string v = argv[i];
        bool plain = !v.empty();
        for (char c : v)
            plain = plain && (isalnum((unsigned char)c) || strchr("_-./=:,@", c));
        if (!plain) {
            std::string e;
            for (char c : v) e += c == '\'' ? std::string("'\\''") : std::string(1, c);
            v = "'" + e + "'";
        }
        s += " " + v;
    }
    fprintf(stderr, "[kite %s] Command: %s\n", cmd, s.c_str());
}

void make_parent_dirs(const std::string &path) {
    size_t pos = path.find_last_of('/');
    if (pos == std::string::npos || pos == 0) return;
    std::string dir = path.substr(0, pos);
    std::string cur;
    size_t i = 0;
    while (i <= dir.size()) {
        size_t j = dir.find('/', i);
        if (j == std::string::npos) j = dir.size();
        cur = dir.substr(0, j);
        if (!cur.empty()) mkdir(cur.c_str(), 0755);
        i = j + 1;
    }
}

unsigned worker_threads(unsigned requested) {
    return requested ? requested : std::max(1u, std::thread::hardware_concurrency());
}

namespace {

size_t fmt_u64(char *o, uint64_t v) {
    char t[24];
    size_t n = 0;
    do {
        t[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    for (size_t i = 0; i < n; i++) o[i] = t[n - 1 - i];
    return n;
}

// Totals of one partition with the dump's u32 saturating sums.
std::vector<Row> dump_partition(const Reader &r, uint32_t p) {
    auto [first, cnt] = r.partition_chunks(p);
    std::vector<Row> rows;
    if (cnt == 1) {
        r.decode_run(r.chunks()[first], rows);
        // a single run: sorted unique as written (a loser tree over one stream)
        for (auto &x : rows) x.count = (uint32_t)x.count;
        return rows;
    }
    rows = r.load_partition(p);
    for (auto &x : rows) x.count = std::min<uint64_t>(x.count, 0xFFFFFFFFull);
    return rows;
}

FILE *open_out(const std::string &path) {
    if (path == "-") return stdout;
    make_parent_dirs(path);
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) throw Error("failed to create " + path + ": " + strerror(errno));
    setvbuf(f, nullptr, _IOFBF, 16 << 20);
    return f;
}


// Opens the device engine for a host command unless the environment
// variable `var` says "host"; "gpu" makes a missing device an error.
kite_gpu_ctx *open_engine(const char *var) {
    const char *ev = getenv(var);
    const std::string engine = ev ? ev : "auto";
    if (engine != "auto" && engine != "gpu" && engine != "host")
        throw UsageError(std::string(var) + " must be auto, gpu or host");
    kite_gpu_ctx *g = nullptr;
    if (engine != "host" && kite_gpu_open(0, 0, &g) != KITE_OK) {
        g = nullptr;
        if (engine == "gpu") throw Error(std::string("GPU engine unavailable: ") + kite_gpu_last_error());
    }
    return g;
}

// The whole dump on the device (kite_gpu_dump): runs of every partition in
// order; the text is written as it arrives.
void dump_on_gpu(kite_gpu_ctx *g, const Reader &r, FILE *out) {
    const Header &h = r.header();
    std::vector<kite_dump_run> runs;
    std::vector<std::vector<uint8_t>> decoded;
    collect_runs(r, runs, decoded);
    // pieces of tens of MB go straight to the descriptor: through stdio a
    // part of each would first be copied into the stream's buffer
    auto write = [](void *user, const char *text, uint64_t len) -> int {
        const int fd = fileno(static_cast<FILE *>(user));
        while (len) {
            const ssize_t w = ::write(fd, text, len);
            if (w < 0) {
                if (errno == EINTR) continue;
                return 1;
            }
            text += w;
            len -= (uint64_t)w;
        }
        return 0;
    };
    if (fflush(out) != 0) throw Error(std::string("write failed: ") + strerror(errno));
    kite_dump_stats st{};
    if (kite_gpu_dump(g, h.k, h.key_width, h.partitions, runs.data(), runs.size(), h.min_count.has_value(),
                      h.min_count.value_or(0), h.max_count.has_value(), h.max_count.value_or(0), write, out,
                      &st) != KITE_OK)
        throw Error(std::string("dump failed: ") + kite_gpu_last_error());
    fprintf(stderr,
            "[kite dump] gpu: %llu runs, %llu rows, %llu lines, %llu bytes in %llu group(s); device "
            "decode %.1f ms, merge %.1f ms, format %.1f ms; text %.1f ms\n",
            (unsigned long long)st.runs, (unsigned long long)st.rows, (unsigned long long)st.lines,
            (unsigned long long)st.bytes, (unsigned long long)st.groups, st.decode_ms, st.merge_ms, st.format_ms,
            st.copy_ms);
}

}  // namespace

int run_dump(int argc, char **argv) {
    Args a({{"input", 'i'}, {"output", 'o'}, {"threads", 't'}, {"buffer", 0}, {"progress", 'p', false},
            {"test-mode", 0, false}},
           argc, argv, 2);
    print_command_line("dump", argc, argv);
    if (!a.has("input") || !a.has("output"))
        throw UsageError("the following required arguments were not provided:\n  --input <FILE>\n  --output <FILE>");
    const auto t0 = std::chrono::steady_clock::now();
    Reader r(a.get("input"));
    const Header &h = r.header();
    fprintf(stderr, "[kite dump] opening: %s\n", a.get("input").c_str());
    fprintf(stderr, "[kite dump] output:  %s\n", a.get("output").c_str());
    fprintf(stderr, "[kite dump] k=%u, partitions=%u, total_chunks=%llu\n", h.k, h.partitions,
            (unsigned long long)r.chunk_count());
    FILE *out = open_out(a.get("output"));
    // engine: the GPU when a device opens (KITE_DUMP_ENGINE=gpu requires
    // it, KITE_DUMP_ENGINE=host keeps the partition workers on the CPU)
    kite_gpu_ctx *gctx = open_engine("KITE_DUMP_ENGINE");
    if (gctx) {
        fprintf(stderr, "[kite dump] engine: gpu\n");
        std::string err;
        try {
            dump_on_gpu(gctx, r, out);
        } catch (const std::exception &e) {
            err = e.what();
        }
        kite_gpu_close(gctx);
        if (out == stdout) fflush(out);
        else fclose(out);
        if (!err.empty()) throw Error(err);
        fprintf(stderr, "[kite dump] done in %.3fs\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        return 0;
    }
    const unsigned nt = std::min<unsigned>(worker_threads((unsigned)a.get_u64("threads", 0)),
                                           std::max<uint32_t>(1, h.partitions));
    const int W = r.words(), k = h.k;
    // partitions are formatted in parallel and written strictly in order
    const uint32_t P = h.partitions;
    std::vector<std::string> buf(P);
    std::vector<std::atomic<int>> ready(P);
    for (auto &x : ready) x = 0;
    std::atomic<uint32_t> next{0};
    std::mutex mu;
    std::string err;
    auto work = [&] {
        for (;;) {
            uint32_t p = next++;
            if (p >= P) return;
            try {
                std::vector<Row> rows = dump_partition(r, p);
                std::string &s = buf[p];
                s.resize(rows.size() * (size_t)(k + 12));
                char *o = s.data();
                for (const Row &x : rows) {
                    if (!passes(x.count, h.min_count, h.max_count)) continue;
                    decode_key(x.key, W, k, o);
                    o += k;
                    *o++ = '\t';
                    o += fmt_u64(o, x.count);
                    *o++ = '\n';
                }
                s.resize((size_t)(o - s.data()));
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> g(mu);
                err = e.what();
            }
            ready[p] = 1;
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) th.emplace_back(work);
    uint64_t bytes = 0;
    for (uint32_t p = 0; p < P; p++) {
        while (!ready[p]) std::this_thread::sleep_for(std::chrono::microseconds(50));
        if (!err.empty()) break;
        if (!buf[p].empty() && fwrite(buf[p].data(), 1, buf[p].size(), out) != buf[p].size()) {
            err = "write failed";
            break;
        }
        bytes += buf[p].size();
        std::string().swap(buf[p]);
    }
    next = P;
    for (auto &x : th) x.join();
    if (out == stdout) fflush(out);
    else fclose(out);
    if (!err.empty()) throw Error(err);
    fprintf(stderr, "[kite dump] done in %.3fs\n",
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return 0;
}

int run_histo(int argc, char **argv) {
    Args a({{"input", 'i', true, true}, {"output", 'o'}, {"label", 0, true, true}, {"threads", 't'},
            {"progress", 'p', false}, {"test-mode", 0, false}},
           argc, argv, 2);
    print_command_line("histo", argc, argv);
    auto inputs = a.list("input");
    if (inputs.empty() || !a.has("output"))
        throw UsageError("the following required arguments were not provided:\n  --input <FILE>...\n  --output <FILE>");
    auto labels = a.list("label");
    if (!labels.empty() && labels.size() != inputs.size())
        throw Error("--label count (" + std::to_string(labels.size()) + ") must match input count (" +
                    std::to_string(inputs.size()) + ")");
    if (labels.empty())
        for (size_t i = 0; i < inputs.size(); i++) labels.push_back(infer_label(inputs[i], i));
    const unsigned nt = worker_threads((unsigned)a.get_u64("threads", 0));
    std::vector<std::map<uint64_t, uint64_t>> hs(inputs.size());
    // engine: the GPU when a device opens (KITE_HISTO_ENGINE as for dump)
    struct Ctx {
        kite_gpu_ctx *g;
        ~Ctx() {
            if (g) kite_gpu_close(g);
        }
    } gctx{open_engine("KITE_HISTO_ENGINE")};
    if (gctx.g) fprintf(stderr, "[kite histo] engine: gpu\n");
    for (size_t fi = 0; fi < inputs.size(); fi++) {
        Reader r(inputs[fi]);
        const Header &h = r.header();
        if (gctx.g) {
            std::vector<kite_dump_run> runs;
            std::vector<std::vector<uint8_t>> decoded;
            collect_runs(r, runs, decoded);
            auto take = [](void *user, const uint64_t *v, const uint64_t *f, uint64_t n) -> int {
                auto *m = static_cast<std::map<uint64_t, uint64_t> *>(user);
                for (uint64_t q = 0; q < n; q++) (*m)[v[q]] += f[q];
                return 0;
            };
            kite_dump_stats st{};
            if (kite_gpu_histo(gctx.g, h.key_width, h.partitions, runs.data(), runs.size(), h.min_count.has_value(),
                               h.min_count.value_or(0), h.max_count.has_value(), h.max_count.value_or(0), take,
                               &hs[fi], &st) != KITE_OK)
                throw Error(std::string("histo failed: ") + kite_gpu_last_error());
            continue;
        }
        std::vector<std::map<uint64_t, uint64_t>> local(nt);
        std::atomic<uint32_t> next{0};
        std::string err;
        std::mutex mu;
        auto work = [&](unsigned t) {
            for (;;) {
                uint32_t p = next++;
                if (p >= h.partitions) return;
                try {
                    for (const Row &x : r.load_partition(p))
                        if (passes(x.count, h.min_count, h.max_count)) local[t][x.count]++;
                } catch (const std::exception &e) {
                    std::lock_guard<std::mutex> g(mu);
                    err = e.what();
                }
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) th.emplace_back(work, t);
        for (auto &x : th) x.join();
        if (!err.empty()) throw Error(err);
        for (auto &m : local)
            for (auto &kv : m) hs[fi][kv.first] += kv.second;
    }
    std::map<uint64_t, int> bins;
    for (auto &m : hs)
        for (auto &kv : m) bins[kv.first] = 1;
    FILE *out = open_out(a.get("output"));
    fprintf(out, "count");
    for (auto &l : labels) fprintf(out, "\t%s", l.c_str());
    fprintf(out, "\n");
    for (auto &b : bins) {
        fprintf(out, "%llu", (unsigned long long)b.first);
        for (auto &m : hs) {
            auto it = m.find(b.first);
            fprintf(out, "\t%llu", (unsigned long long)(it == m.end() ? 0 : it->second));
        }
        fprintf(out, "\n");
    }
    if (out == stdout) fflush(out);
    else fclose(out);
    fprintf(stderr, "[kite histo] done (bins=%zu, output=%s)\n", bins.size(), a.get("output").c_str());
    return 0;
}

namespace {
std::string fmt_opt(const std::optional<uint32_t> &v) { return v ? std::to_string(*v) : "none"; }

void check_compatible(const std::vector<std::unique_ptr<Reader>> &rs, const std::vector<std::string> &paths) {
    if (rs.empty()) throw Error("no input .kite files were provided.\nPass one or more .kite files with -i/--input.");
    const Header &f = rs[0]->header();
    for (size_t i = 1; i < rs.size(); i++) {
        const Header &h = rs[i]->header();
        std::string nm = "input #" + std::to_string(i + 1) + " (" + paths[i] + ")";
        std::string fn = "input #1 (" + paths[0] + ")";
        if (h.k != f.k)
            throw Error("incompatible .kite inputs: " + nm + " has k=" + std::to_string(h.k) + ", but " + fn +
                        " has k=" + std::to_string(f.k) + ".\nAll input .kite files must use the same k.");
        if ((h.canonical != 0) != (f.canonical != 0))
            throw Error("incompatible .kite inputs: " + nm + " has canonical=" + (h.canonical ? "true" : "false") +
                        ", but " + fn + " has canonical=" + (f.canonical ? "true" : "false") +
                        ".\nAll input .kite files must agree on canonical mode.");
        if (h.partitions != f.partitions)
            throw Error("incompatible .kite inputs: " + nm + " has partitions=" + std::to_string(h.partitions) +
                        ", but " + fn + " has partitions=" + std::to_string(f.partitions) +
                        ".\nAll input .kite files must use the same partition count.");
        if (h.key_width != f.key_width)
            throw Error("incompatible .kite inputs: " + nm + " has key_width=" + std::to_string(h.key_width) +
                        ", but " + fn + " has key_width=" + std::to_string(f.key_width) +
                        ".\nAll input .kite files must use the same key width.");
    }
}
}  // namespace

int run_merge(int argc, char **argv) {
    Args a({{"input", 'i', true, true}, {"output", 'o'}, {"threads", 't'}, {"progress", 'p', false},
            {"test-mode", 0, false}, {"engine", 0}},
           argc, argv, 2);
    print_command_line("merge", argc, argv);
    auto inputs = a.list("input");
    if (inputs.size() < 2)
        throw UsageError("2 values required for '--input <FILE>...' but " + std::to_string(inputs.size()) + " were provided");
    if (!a.has("output")) throw UsageError("the following required arguments were not provided:\n  --output <FILE>");
    const std::string output = a.get("output");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::unique_ptr<Reader>> rs;
    for (auto &p : inputs) rs.emplace_back(new Reader(p));
    fprintf(stderr, "[kite merge] inputs=%zu, output=%s\n", inputs.size(), output.c_str());
    check_compatible(rs, inputs);
    const Header &f = rs[0]->header();
    std::optional<uint32_t> omin = f.min_count, omax = f.max_count;
    for (auto &r : rs) {
        if (r->header().min_count != f.min_count) omin.reset();
        if (r->header().max_count != f.max_count) omax.reset();
    }
    const unsigned nt = worker_threads((unsigned)a.get_u64("threads", 0));
    make_parent_dirs(output);
    Header h;
    h.k = f.k;
    h.canonical = f.canonical ? 1 : 0;
    h.partitions = f.partitions;
    h.key_width = f.key_width;
    h.codec_mask = (uint16_t)(CODEC_DVLE | CODEC_RAW);
    h.min_count = omin;
    h.max_count = omax;
    Writer w(output, h, 16u << 20, Writer::suggest_write_lanes((int)std::max(1u, std::thread::hardware_concurrency())));
    const bool use_gpu = a.get("engine", "gpu") != "cpu";
    MergeStats ms = merge_files(rs, w, nt, use_gpu);
    auto [tc, tr] = w.finalize();
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // <stem>.merge_report.md (src/merge/mod.rs:216-301)
    {
        size_t sl = output.find_last_of('/');
        std::string dir = sl == std::string::npos ? "" : output.substr(0, sl + 1);
        std::string name = sl == std::string::npos ? output : output.substr(sl + 1);
        size_t dot = name.find_last_of('.');
        std::string stem = (dot == std::string::npos || dot == 0) ? name : name.substr(0, dot);
        if (stem.empty()) stem = "merge";
        std::string rp = dir + stem + ".merge_report.md";
        FILE *o = fopen(rp.c_str(), "wb");
        if (!o) throw Error("failed to create " + rp);
        fprintf(o, "# Kite Merge Report\n\n## Summary\n\n| Field | Value |\n|---|---:|\n");
        fprintf(o, "| Inputs | %zu |\n| Output rows | %llu |\n| Output chunks | %llu |\n| Partitions | %u |\n",
                inputs.size(), (unsigned long long)tr, (unsigned long long)tc, f.partitions);
        fprintf(o, "| Threads requested | %llu |\n| Worker threads used | %u |\n| Elapsed seconds | %.3f |\n\n",
                (unsigned long long)a.get_u64("threads", 0), nt, el);
        fprintf(o, "## Compatibility\n\n| Field | Value |\n|---|---:|\n| k | %u |\n| Canonical | %s |\n| Key width | %u |\n\n",
                f.k, f.canonical ? "true" : "false", f.key_width);
        fprintf(o,
                "## Downstream Filter Semantics\n\nInput header min/max filters are applied before merge. The output "
                "header retains a min or max marker only when all inputs share the same marker.\n\n");
        fprintf(o, "- Output min: %s\n- Output max: %s\n\n", fmt_opt(omin).c_str(), fmt_opt(omax).c_str());
        fprintf(o, "## Inputs\n\n| # | Path | k | Canonical | Partitions | Key width | Rows | Chunks | Min | Max |\n");
        fprintf(o, "|---:|---|---:|---:|---:|---:|---:|---:|---:|---:|\n");
        for (size_t i = 0; i < rs.size(); i++) {
            const Header &x = rs[i]->header();
            fprintf(o, "| %zu | `%s` | %u | %s | %u | %u | %llu | %llu | %s | %s |\n", i + 1, inputs[i].c_str(), x.k,
                    x.canonical ? "true" : "false", x.partitions, x.key_width, (unsigned long long)x.total_rows,
                    (unsigned long long)rs[i]->chunk_count(), fmt_opt(x.min_count).c_str(),
                    fmt_opt(x.max_count).c_str());
        }
        fprintf(o, "\n## Output\n\n- File: `%s`\n- Report: `%s`\n", output.c_str(), rp.c_str());
        fclose(o);
        fprintf(stderr, "[kite merge] wrote %llu rows across %llu chunks in %.3fs (report=%s, engine=%s, merge %.3fs)\n",
                (unsigned long long)tr, (unsigned long long)tc, el, rp.c_str(), ms.engine.c_str(), ms.merge_s);
    }
    return 0;
}

int run_info(int argc, char **argv) {
    Args a({{"input", 'i'}, {"set-min", 0}, {"set-max", 0}, {"clear-min", 0, false}, {"clear-max", 0, false},
            {"test-mode", 0, false}},
           argc, argv, 2);
    print_command_line("info", argc, argv);
    if (!a.has("input")) throw UsageError("the following required arguments were not provided:\n  --input <FILE>");
    const std::string path = a.get("input");
    const auto smin = a.get_u32("set-min"), smax = a.get_u32("set-max");
    const bool cmin = a.flag("clear-min"), cmax = a.flag("clear-max");
    if (smin && cmin) throw Error("--set-min and --clear-min cannot be used together");
    if (smax && cmax) throw Error("--set-max and --clear-max cannot be used together");
    fprintf(stderr, "[kite info] opening: %s\n", path.c_str());
    if (smin || smax || cmin || cmax) {
        Header cur = Reader(path).header();
        auto nmin = cmin ? std::nullopt : (smin ? smin : cur.min_count);
        auto nmax = cmax ? std::nullopt : (smax ? smax : cur.max_count);
        if (nmin && nmax && *nmin > *nmax)
            throw Error("header min (" + std::to_string(*nmin) + ") must be <= header max (" + std::to_string(*nmax) + ")");
        // update_count_filters_in_file (src/format/header.rs:166-183): only the 128-byte header changes
        FILE *f = fopen(path.c_str(), "r+b");
        if (!f) throw Error("failed to open " + path + ": " + strerror(errno));
        uint8_t hb[128];
        if (fread(hb, 1, 128, f) != 128) {
            fclose(f);
            throw Error("failed to read header");
        }
        Header h = Header::parse(hb);
        h.min_count = nmin;
        h.max_count = nmax;
        h.write(hb);
        fseek(f, 0, SEEK_SET);
        if (fwrite(hb, 1, 128, f) != 128) {
            fclose(f);
            throw Error("failed to write header");
        }
        fclose(f);
        fprintf(stderr, "[kite info] updated header count filters: min=%s, max=%s\n", fmt_opt(nmin).c_str(),
                fmt_opt(nmax).c_str());
    }
    // report layout of run_info (src/info/mod.rs:108-157)
    Reader r(path);
    const Header &h = r.header();
    uint64_t pbytes = 0, prows = 0, hot = 0, cold = UINT64_MAX, maxc = 0, minc = UINT64_MAX;
    for (uint32_t p = 0; p < h.partitions; p++) {
        auto [first, cnt] = r.partition_chunks(p);
        uint64_t rows = 0, bytes = 0;
        for (uint64_t c = first; c < first + cnt; c++) {
            rows += r.chunks()[c].rows;
            bytes += r.chunks()[c].encoded;
        }
        prows += rows;
        pbytes += bytes;
        hot = std::max(hot, rows);
        cold = std::min(cold, rows);
    }
    if (!h.partitions) cold = 0;
    uint64_t raw = 0, dvle = 0, zstd = 0;
    for (auto &c : r.chunks()) {
        maxc = std::max(maxc, c.encoded);
        minc = std::min(minc, c.encoded);
        raw += (c.codec & CODEC_RAW) != 0;
        dvle += (c.codec & CODEC_DVLE) != 0;
        zstd += (c.codec & CODEC_ZSTD) != 0;
    }
    if (r.chunks().empty()) minc = 0;
    auto U = [](uint64_t v) { return (unsigned long long)v; };
    printf("Kite Database\n  Path:               %s\n  File Size:          %llu bytes\n\n", path.c_str(), U(r.file_size()));
    printf("Format\n  Version:            1\n  K:                  %u\n  Canonical:          %s\n", h.k,
           h.canonical ? "true" : "false");
    printf("  Min Count:          %s\n  Max Count:          %s\n", fmt_opt(h.min_count).c_str(), fmt_opt(h.max_count).c_str());
    printf("  Filter Semantics:   downstream commands exclude rows outside Min/Max\n");
    printf("  Partitions:         %u\n  Key Width:          %u\n\n", h.partitions, h.key_width);
    printf("Content\n  Total Chunks:       %llu\n  Total Rows:         %llu\n  Partition Rows:     %llu\n"
           "  Partition Bytes:    %llu\n\n",
           U(r.chunk_count()), U(h.total_rows), U(prows), U(pbytes));
    printf("Codec Mix\n  RAW Chunks:         %llu\n  DVLE Chunks:        %llu\n  ZSTD Chunks:        %llu\n\n", U(raw),
           U(dvle), U(zstd));
    printf("Distribution\n  Largest Chunk:      %llu bytes\n  Smallest Chunk:     %llu bytes\n"
           "  Hottest Partition:  %llu rows\n  Coldest Partition:  %llu rows\n",
           U(maxc), U(minc), U(hot), U(cold));
    fflush(stdout);
    return 0;
}

}  // namespace kh
