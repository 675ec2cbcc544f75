// count_cmd.cpp -- `kite count`: flag validation and the partition planner of
// the reference (src/count/mod.rs:208-248, 624-1016; src/runtime.rs:6-142),
// record-aligned FASTQ chunking (src/count/decompress.rs:13-138), and the
// GPU engine in place of the worker pool, with its run chunks written
// through the .kite staging writer.
#include <fcntl.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <fstream>
#include <mutex>
#include <sstream>
#include <thread>

#include "cli.hpp"
#include "commands.hpp"
#include "kite_gpu.h"
#include "kitefmt.hpp"

namespace kh {

namespace {
constexpr uint64_t MIB = 1ull << 20, GIB = 1ull << 30, TIB = 1ull << 40;
constexpr uint64_t AUTO_MIN_P = 256, AUTO_MAX_P = 2048;
constexpr uint64_t WORKER_SCRATCH = 48 * MIB;

// ---- src/runtime.rs:6-142 ----------------------------------------------------
uint64_t parse_memory_spec(const std::string &spec) {
    bool digits = !spec.empty();
    for (char c : spec) digits = digits && c >= '0' && c <= '9';
    const std::string fmt = "expected formats like 512M, 64G, 1T, or raw bytes";
    if (digits) {
        uint64_t v;
        try {
            v = std::stoull(spec);
        } catch (...) {
            throw Error("invalid memory size '" + spec + "': " + fmt);
        }
        if (v == 0) throw Error("invalid memory size '" + spec + "': value must be greater than zero");
        return v;
    }
    if (spec.size() < 2) throw Error("invalid memory size '" + spec + "': " + fmt);
    char suf = spec.back();
    std::string num = spec.substr(0, spec.size() - 1);
    uint64_t mul = suf == 'M' ? MIB : suf == 'G' ? GIB : suf == 'T' ? TIB : 0;
    if (!mul) throw Error("invalid memory size '" + spec + "': suffix must be M, G, or T");
    bool ok = !num.empty();
    for (char c : num) ok = ok && c >= '0' && c <= '9';
    if (!ok) throw Error("invalid memory size '" + spec + "': numeric part must be a positive integer");
    uint64_t v = std::stoull(num);
    if (v == 0) throw Error("invalid memory size '" + spec + "': value must be greater than zero");
    if (v > UINT64_MAX / mul) throw Error("memory size '" + spec + "' exceeds supported range");
    return v * mul;
}

uint64_t resolve_memory_limit(const std::string &spec) {
    if (!spec.empty()) return parse_memory_spec(spec);
    std::ifstream mi("/proc/meminfo");
    std::string line;
    while (std::getline(mi, line)) {
        if (line.rfind("MemAvailable:", 0) == 0) {
            std::istringstream is(line.substr(13));
            uint64_t kib = 0;
            is >> kib;
            uint64_t a = kib * 1024 * 2 / 3;
            if (a) return a;
        }
    }
    uint64_t total = (uint64_t)sysconf(_SC_PHYS_PAGES) * (uint64_t)sysconf(_SC_PAGESIZE);
    uint64_t head = std::min<uint64_t>(std::max<uint64_t>(total / 100, 512 * MIB), 2 * GIB);
    return total > head ? total - head : 0;
}

enum class Comp { Plain, Gzip, Bgzf };

// src/count/reader.rs:22-51
Comp detect_compression(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw Error("failed to inspect input FASTQ compression.\nInput: " + path + "\nReason: " + strerror(errno));
    uint8_t h[18];
    size_t n = fread(h, 1, 18, f);
    fclose(f);
    if (n < 2 || h[0] != 0x1f || h[1] != 0x8b) return Comp::Plain;
    if (n >= 18 && (h[3] & 0x04)) {
        uint16_t xlen = (uint16_t)(h[10] | (h[11] << 8));
        if (xlen >= 6 && h[12] == 0x42 && h[13] == 0x43) return Comp::Bgzf;
    }
    return Comp::Gzip;
}

// A byte stream: plain file or (multi-member) gzip, read in large pieces.
class Source {
   public:
    Source(const std::string &path, Comp c) : comp_(c) {
        if (c == Comp::Plain) {
            fd_ = open(path.c_str(), O_RDONLY);
            if (fd_ < 0) throw Error("failed to open input " + path + ": " + strerror(errno));
        } else {
            gz_ = gzopen(path.c_str(), "rb");
            if (!gz_) throw Error("failed to open input " + path);
            gzbuffer(gz_, 1 << 20);
        }
    }
    ~Source() {
        if (fd_ >= 0) close(fd_);
        if (gz_) gzclose(gz_);
    }
    // read up to n bytes; returns bytes read (0 at end)
    size_t read(uint8_t *dst, size_t n) {
        size_t got = 0;
        while (got < n) {
            long r;
            if (fd_ >= 0) {
                r = ::read(fd_, dst + got, n - got);
            } else {
                unsigned want = (unsigned)std::min<size_t>(n - got, 1u << 30);
                r = gzread(gz_, dst + got, want);
            }
            if (r < 0) throw Error("read error on FASTQ input");
            if (r == 0) break;
            got += (size_t)r;
        }
        return got;
    }

   private:
    Comp comp_;
    int fd_ = -1;
    gzFile gz_ = nullptr;
};

// find_fastq_split_4lines (src/count/bgzf.rs:109-121)
size_t split_4lines(const uint8_t *d, size_t n) {
    size_t nl = 0, last = 0;
    const uint8_t *p = d, *e = d + n;
    while (p < e) {
        const void *q = memchr(p, '\n', (size_t)(e - p));
        if (!q) break;
        p = static_cast<const uint8_t *>(q) + 1;
        if (++nl % 4 == 0) last = (size_t)(p - d);
    }
    return last;
}

// PlainChunkedReader / GzipChunkedReader::next_chunk (decompress.rs:36-82, 104-138)
class Chunker {
   public:
    Chunker(Source &s, size_t chunk) : src_(s), chunk_(chunk) {}
    // the input is read straight into `out`, behind the partial record the
    // previous chunk left over (no zeroing: the buffer type leaves new bytes
    // uninitialised)
    bool next(Bytes &out) {
        out.assign(carry_.begin(), carry_.end());
        carry_.clear();
        const size_t have = out.size();
        out.resize(have + chunk_);
        const size_t got = src_.read(out.data() + have, chunk_);
        out.resize(have + got);
        if (out.empty()) return false;
        if (got < chunk_) return true;
        size_t sp = split_4lines(out.data(), out.size());
        if (sp == 0 || sp >= out.size()) return true;
        carry_.assign(out.begin() + sp, out.end());
        out.resize(sp);
        return true;
    }

   private:
    Source &src_;
    size_t chunk_;
    std::vector<uint8_t> carry_;
};

struct Plan {
    uint64_t partitions = 0, chunk = 0, staging = 0, workers = 0, queue_depth = 0;
    double avg_read_len = 150;
    uint64_t sampled_reads = 0, input_bytes = 0;
    std::vector<std::string> reasons;
};

uint64_t floor_pow2(uint64_t v) { return v <= 1 ? 1 : 1ull << (63 - __builtin_clzll(v)); }
uint64_t ceil_pow2(uint64_t v) {
    v = std::max<uint64_t>(v, 1);
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// sample_input (src/count/mod.rs:1018-1079): <= 200k reads / 32 MiB
void sample_input(const std::string &path, Comp comp, Plan &pl) {
    Source s(path, comp);
    Chunker ch(s, 8 * MIB);
    Bytes c;
    uint64_t reads = 0, bases = 0, decomp = 0;
    while (reads < 200000 && decomp < 32 * MIB) {
        if (!ch.next(c)) break;
        decomp += c.size();
        if (c.empty() || c[0] != '@') continue;
        int state = 0;
        size_t ss = 0;
        for (size_t i = 0; i < c.size(); i++) {
            if (c[i] != '\n') continue;
            if (state == 0) {
                state = 1;
                ss = i + 1;
            } else if (state == 1) {
                if (i > ss) {
                    reads++;
                    bases += i - ss;
                }
                state = 2;
            } else if (state == 2) {
                state = 3;
            } else {
                state = 0;
            }
        }
        if (state == 1 && ss < c.size()) {
            reads++;
            bases += c.size() - ss;
        }
    }
    pl.sampled_reads = reads;
    pl.avg_read_len = reads ? (double)bases / (double)reads : 150.0;
}

// derive_heuristic_plan (src/count/mod.rs:624-864) and helpers (866-1016)
Plan derive_plan(const std::string &input, Comp comp, int k, uint16_t kw, uint64_t threads, uint64_t M,
                 uint64_t req_partitions, uint64_t chunk_kmers, std::optional<uint64_t> buffer_mb) {
    Plan pl;
    FILE *f = fopen(input.c_str(), "rb");
    if (f) {
        fseeko(f, 0, SEEK_END);
        pl.input_bytes = (uint64_t)ftello(f);
        fclose(f);
    }
    sample_input(input, comp, pl);
    const uint64_t E = kw == 64 ? 1 : 2;
    uint64_t base_mb = kw == 64 ? 640 : (kw == 128 ? 448 : 320);
    uint64_t adj = base_mb;
    if (kw == 64 && k >= 22 && k <= 31) adj = 192;
    else if (k <= 21) adj = base_mb + 64;
    else if (k >= 52) adj = base_mb - 64;
    const uint64_t target_part = adj * MIB;
    uint64_t twb = (kw == 64 && k <= 31) ? 1250 * MIB : (kw == 128 ? 2 * GIB : 3 * GIB);
    const uint64_t min_worker_budget = std::min(twb, std::max(M / 2, 512 * MIB));
    uint64_t decomp = 1;
    if (comp == Comp::Bgzf) {
        uint64_t maxd = std::max<uint64_t>(threads / 2, 1);
        uint64_t div = k <= 31 ? 4 : (k <= 51 ? 5 : (k <= 75 ? 6 : 8));
        decomp = std::min(std::max<uint64_t>(threads / div, 1), maxd);
    }
    uint64_t untracked;
    {
        uint64_t base = M / 6, stacks = (threads + 3) * 8 * MIB, codec = decomp * 32 * MIB;
        untracked = std::min(std::max(base + stacks + codec, 2 * GIB), 12 * GIB);
        untracked = std::min(untracked, M / 3);
        if (comp == Comp::Bgzf) untracked += pl.input_bytes;
    }
    uint64_t workers = threads > decomp + 1 ? threads - (decomp + 1) : 1;
    workers = std::max<uint64_t>(workers, 1);
    uint64_t P;
    if (req_partitions) {
        P = req_partitions;
        pl.reasons.push_back("Using user-defined partitions=" + std::to_string(P));
    } else {
        double expansion = std::min(std::max(pl.avg_read_len / 10.0, 4.0), 24.0);
        uint64_t est = (uint64_t)((double)pl.input_bytes * expansion * (double)E);
        uint64_t raw = std::max<uint64_t>((est + target_part - 1) / target_part, 1);
        uint64_t derived = std::min(std::max(ceil_pow2(raw), AUTO_MIN_P), AUTO_MAX_P);
        uint64_t st_mb = std::max<uint64_t>(buffer_mb.value_or(8), 1);
        uint64_t per_part = st_mb * MIB + workers * 5 * 1024;
        uint64_t capped = std::max<uint64_t>((M / 4) / std::max<uint64_t>(per_part, 1), AUTO_MIN_P);
        capped = floor_pow2(capped);
        P = std::min(std::max(std::min(derived, capped), AUTO_MIN_P), AUTO_MAX_P);
        pl.reasons.push_back("Auto partitions=" + std::to_string(derived) + " capped to " + std::to_string(P));
    }
    std::vector<uint64_t> cands;
    if (buffer_mb) {
        uint64_t r = std::max<uint64_t>(*buffer_mb, 1);
        for (uint64_t v : {r, (uint64_t)16, (uint64_t)8, (uint64_t)4, (uint64_t)2, (uint64_t)1})
            if (v <= r) cands.push_back(v);
    } else {
        cands = {16, 8, 4, 2};
    }
    std::sort(cands.rbegin(), cands.rend());
    cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
    size_t si = 0;
    for (;;) {
        uint64_t staging = cands[si] * MIB;
        uint64_t chunk = chunk_kmers ? std::min(std::max(chunk_kmers * 2, 8 * MIB), 64 * MIB)
                                     : std::min(std::max(M / std::max<uint64_t>(workers * 96, 1), 8 * MIB), 64 * MIB);
        uint64_t qbase = comp == Comp::Plain ? 8 : std::min<uint64_t>(std::max<uint64_t>(decomp * 2, 8), 64);
        uint64_t qd = std::min(qbase, std::min<uint64_t>(std::max<uint64_t>(M / GIB, 8), 64));
        uint64_t dbudget = chunk * qd;
        uint64_t reserved = dbudget + P * staging + workers * WORKER_SCRATCH + untracked;
        uint64_t tracked = M > untracked + dbudget ? M - untracked - dbudget : 0;
        uint64_t wbudget = (M > reserved ? M - reserved : 0) / std::max<uint64_t>(workers, 1);
        if (reserved < M && tracked > 0 && wbudget >= min_worker_budget) {
            pl.partitions = P;
            pl.chunk = chunk;
            pl.staging = staging;
            pl.workers = workers;
            pl.queue_depth = qd;
            return pl;
        }
        if (si + 1 < cands.size()) {
            si++;
            continue;
        }
        if (workers > 1) {
            workers--;
            si = 0;
            continue;
        }
        if (!req_partitions && P > AUTO_MIN_P) {
            P /= 2;
            si = 0;
            continue;
        }
        throw Error("unable to derive count heuristic plan within memory budget " + std::to_string(M) + " bytes");
    }
}

// Chunk codec policy of flush_bucket (src/count/worker.rs:340-372): a DVLE
// or RAW payload is zstd-compressed at the mode's level (fast 0, balanced 2,
// compact 5; src/cli.rs:13-19) when it is large enough, and the compressed
// form is kept only when it saves enough.  The thresholds restate
// min_compression_input_bytes, min_compression_savings_bytes and
// compression_level_for (src/count/worker.rs:451-501).
size_t zstd_min_input(int k, bool dvle, int entry_size) {
    if (entry_size == 12 && dvle && k <= 31) return 16 * 1024;
    if (entry_size == 20 && dvle) return 32 * 1024;
    if (entry_size == 36) return 64 * 1024;
    return 32 * 1024;
}

size_t zstd_min_savings(size_t payload, size_t logical, bool dvle, int entry_size) {
    const size_t bias = std::min<size_t>((logical > payload ? logical - payload : 0) / 32, 24 * 1024);
    const size_t floor = entry_size == 12 && dvle ? 8 * 1024 : (entry_size == 20 && dvle ? 16 * 1024 : 32 * 1024);
    const size_t m = std::max(floor, payload / 100);
    return m > bias ? m - bias : 0;
}

int zstd_level_for(int base, int k, size_t payload, size_t logical, bool dvle, int entry_size) {
    if (base <= 0) return 0;
    int level = base;
    if (dvle && entry_size == 12 && k <= 31 && payload >= 1024 * 1024) level++;
    // logical >= 2 * payload, with the reference's saturating multiply
    if ((payload > SIZE_MAX / 2 ? logical == SIZE_MAX : logical >= payload * 2) && payload >= 4 * 1024 * 1024)
        level++;
    return std::min(level, 6);
}

// Receives the encoded chunks from the device.  With a zstd level, chunks
// that may compress go to a pool of host threads (the reference compresses
// on its count workers); the Writer takes appends from one thread at a time.
struct EmitCtx {
    Writer *w;
    int base_level = 0, k = 0, entry_size = 12;
    struct Job {
        uint32_t p, codec;
        uint64_t rows, raw, kmin[4], kmax[4];
        std::vector<uint8_t> data;
    };
    std::mutex wmu, qmu;
    std::condition_variable cv_job, cv_space;
    std::deque<Job> q;
    size_t q_bytes = 0, q_cap = 256u << 20;
    bool stop = false;
    std::string err;
    std::vector<std::thread> pool;
    uint64_t zstd_chunks = 0;

    void append(uint32_t p, uint32_t codec, uint64_t rows, uint64_t raw, const uint64_t *kmin, const uint64_t *kmax,
                const uint8_t *data, size_t len) {
        std::lock_guard<std::mutex> g(wmu);
        w->append_chunk(p, codec, rows, raw, kmin, kmax, data, len);
    }
    void compress_and_append(Job &j) {
        const bool dvle = (j.codec & CODEC_DVLE) != 0;
        const size_t payload = j.data.size(), logical = (size_t)j.rows * (size_t)entry_size;
        const int level = zstd_level_for(base_level, k, payload, logical, dvle, entry_size);
        if (level > 0 && payload >= zstd_min_input(k, dvle, entry_size)) {
            std::vector<uint8_t> z = zstd_compress(j.data.data(), payload, level);
            if (z.size() + zstd_min_savings(payload, logical, dvle, entry_size) < payload) {
                append(j.p, j.codec | CODEC_ZSTD, j.rows, j.raw, j.kmin, j.kmax, z.data(), z.size());
                std::lock_guard<std::mutex> g(qmu);
                zstd_chunks++;
                return;
            }
        }
        append(j.p, j.codec, j.rows, j.raw, j.kmin, j.kmax, j.data.data(), payload);
    }
    void start(unsigned threads) {
        for (unsigned t = 0; t < threads; t++)
            pool.emplace_back([this] {
                for (;;) {
                    Job j;
                    {
                        std::unique_lock<std::mutex> g(qmu);
                        cv_job.wait(g, [&] { return stop || !q.empty(); });
                        if (q.empty()) return;
                        j = std::move(q.front());
                        q.pop_front();
                        q_bytes -= j.data.size();
                    }
                    cv_space.notify_all();
                    try {
                        compress_and_append(j);
                    } catch (const std::exception &e) {
                        std::lock_guard<std::mutex> g(qmu);
                        if (err.empty()) err = e.what();
                    }
                }
            });
    }
    void push(Job &&j) {
        std::unique_lock<std::mutex> g(qmu);
        cv_space.wait(g, [&] { return q.empty() || q_bytes + j.data.size() <= q_cap; });
        q_bytes += j.data.size();
        q.push_back(std::move(j));
        g.unlock();
        cv_job.notify_one();
    }
    // Waits for every queued chunk; rethrows the first compression error.
    void drain() {
        {
            std::lock_guard<std::mutex> g(qmu);
            stop = true;
        }
        cv_job.notify_all();
        for (auto &t : pool) t.join();
        pool.clear();
        if (!err.empty()) throw Error(err);
    }
    ~EmitCtx() {
        if (!pool.empty()) {
            {
                std::lock_guard<std::mutex> g(qmu);
                stop = true;
            }
            cv_job.notify_all();
            for (auto &t : pool) t.join();
        }
    }
};

int emit_cb(void *user, uint32_t p, uint32_t codec, uint64_t rows, uint64_t raw, const uint64_t kmin[4],
            const uint64_t kmax[4], const uint8_t *data, size_t len) {
    EmitCtx &ec = *static_cast<EmitCtx *>(user);
    try {
        if (ec.pool.empty() || len < zstd_min_input(ec.k, (codec & CODEC_DVLE) != 0, ec.entry_size)) {
            ec.append(p, codec, rows, raw, kmin, kmax, data, len);
            return 0;
        }
        EmitCtx::Job j;
        j.p = p;
        j.codec = codec;
        j.rows = rows;
        j.raw = raw;
        std::copy(kmin, kmin + 4, j.kmin);
        std::copy(kmax, kmax + 4, j.kmax);
        j.data.assign(data, data + len);
        ec.push(std::move(j));
        return 0;
    } catch (...) {
        return 1;
    }
}

void gpu_check(int rc, const char *what) {
    if (rc) throw Error(std::string(what) + ": " + kite_gpu_last_error());
}
}  // namespace

int run_count(int argc, char **argv) {
    Args a({{"input", 'i'}, {"output", 'o'}, {"kmer", 'k'}, {"canonical", 0}, {"min", 0}, {"max", 0},
            {"mode", 0}, {"threads", 't'}, {"partitions", 0}, {"chunk-kmers", 0}, {"memory", 0},
            {"buffer", 0}, {"progress", 'p', false}, {"test-mode", 0, false}},
           argc, argv, 2);
    print_command_line("count", argc, argv);
    if (!a.has("input")) throw UsageError("the following required arguments were not provided:\n  --input <FILE>");
    if (!a.has("output")) throw UsageError("the following required arguments were not provided:\n  --output <FILE>");
    const std::string input = a.get("input"), output = a.get("output");
    uint64_t k64 = a.get_u64("kmer", 21);
    if (k64 > 255) throw UsageError("invalid value '" + a.get("kmer") + "' for '--kmer <k>': number too large to fit in target type");
    const int k = (int)k64;
    const bool canonical = a.get_bool("canonical", false);
    const auto mn = a.get_u32("min"), mx = a.get_u32("max");
    const std::string mode = a.get("mode", "fast");
    if (mode != "fast" && mode != "balanced" && mode != "compact")
        throw UsageError("invalid value '" + mode + "' for '--mode <MODE>'\n  [possible values: fast, balanced, compact]");
    const uint64_t threads_req = a.get_u64("threads", 0);
    const uint64_t parts_req = a.get_u64("partitions", 0);
    const uint64_t chunk_kmers = a.get_u64("chunk-kmers", 0);
    std::optional<uint64_t> buffer;
    if (a.has("buffer")) buffer = a.get_u64("buffer", 0);
    const bool test_mode = a.flag("test-mode");
    int zstd_level = mode == "balanced" ? 2 : (mode == "compact" ? 5 : 0);
    if (zstd_level > 0 && !zstd_available()) {
        fprintf(stderr, "[kite count] warning: libzstd.so.1 is not available; --mode %s writes uncompressed chunks\n",
                mode.c_str());
        zstd_level = 0;
    }

    if (k < 1 || k > 128)
        throw Error("invalid k-mer length: --k must be between 1 and 128.\nGot: " + std::to_string(k) + "\nExample: --k 21");
    if (parts_req != 0 && (parts_req & (parts_req - 1)))
        throw Error("invalid partition count: --partitions must be a power of 2 because kite uses partition bits for routing.\nGot: " +
                    std::to_string(parts_req) + "\nExamples: 256, 512, 1024");
    if (mn && mx && *mn > *mx)
        throw Error("invalid count filter range: --min must be less than or equal to --max.\nGot: --min " +
                    std::to_string(*mn) + " --max " + std::to_string(*mx));
    const bool to_null = output == "/dev/null";
    if (!to_null && (output.size() < 5 || output.compare(output.size() - 5, 5, ".kite") != 0))
        throw Error("invalid count output path: output must end with .kite, or be /dev/null for benchmark runs.\nGot: " +
                    output + "\nExample: -o result.kite");
    const uint16_t kw = width_for_k(k);
    const uint64_t threads = threads_req ? threads_req : std::max(1u, std::thread::hardware_concurrency());
    make_parent_dirs(output);
    const uint64_t M = resolve_memory_limit(a.get("memory"));
    const Comp comp = detect_compression(input);
    const auto t0 = std::chrono::steady_clock::now();
    Plan pl = derive_plan(input, comp, k, kw, threads, M, parts_req, chunk_kmers, buffer);

    fprintf(stderr, "[kite count] input:       %s\n", input.c_str());
    fprintf(stderr, "[kite count] output:      %s\n", output.c_str());
    fprintf(stderr, "[kite count] compression: %s\n",
            comp == Comp::Plain ? "plain" : (comp == Comp::Gzip ? "gzip" : "bgzf"));
    fprintf(stderr, "[kite count] k=%d, canonical=%s, partitions=%llu, min=%s, max=%s\n", k,
            canonical ? "true" : "false", (unsigned long long)pl.partitions,
            mn ? std::to_string(*mn).c_str() : "none", mx ? std::to_string(*mx).c_str() : "none");
    // codec_policy_label (src/count/mod.rs:952-959)
    fprintf(stderr, "[kite count] mode=%s, codec=%s, engine=gpu (gfx950)\n", mode.c_str(),
            kw == 256 ? (zstd_level > 0 ? "raw-u256+zstd-if-smaller" : "raw-u256")
                      : (zstd_level > 0 ? "dvle+zstd-if-smaller" : "dvle-only"));
    if (test_mode) {
        fprintf(stderr, "[kite count][auto] input=%.2fGB, read_len=%.1fbp from %llu reads\n",
                pl.input_bytes / (double)GIB, pl.avg_read_len, (unsigned long long)pl.sampled_reads);
        fprintf(stderr, "[kite count][auto] P=%llu, chunk=%lluMB, S=%lluMB\n", (unsigned long long)pl.partitions,
                (unsigned long long)(pl.chunk / MIB), (unsigned long long)(pl.staging / MIB));
        for (auto &r : pl.reasons) fprintf(stderr, "[kite count][auto] %s\n", r.c_str());
    }
    // KITE_COUNT_PLAN_ONLY=1 stops after the plan (no device is opened, no
    // output written): the planner can be checked on a host without a GPU
    if (const char *po = getenv("KITE_COUNT_PLAN_ONLY"); po && std::string(po) == "1") return 0;

    kite_gpu_ctx *ctx = nullptr;
    gpu_check(kite_gpu_open(0, 0, &ctx), "failed to open the GPU engine");
    struct Closer {
        kite_gpu_ctx *c;
        ~Closer() { kite_gpu_close(c); }
    } closer{ctx};
    gpu_check(kite_gpu_count_begin(ctx, (uint32_t)k, canonical, (uint32_t)pl.partitions), "count");
    uint64_t chunks = 0;
    // BGZF members are inflated on the device unless KITE_BGZF_ENGINE=host
    // (then zlib on this host, as for plain gzip)
    const char *be = getenv("KITE_BGZF_ENGINE");
    if (comp == Comp::Bgzf && !(be && std::string(be) == "host")) {
        int fd = open(input.c_str(), O_RDONLY);
        if (fd < 0) throw Error("failed to open input " + input + ": " + strerror(errno));
        struct Fd {
            int f;
            ~Fd() { close(f); }
        } fdc{fd};
        std::vector<uint8_t> buf(64 * MIB);
        size_t have = 0;
        for (;;) {
            ssize_t r = ::read(fd, buf.data() + have, buf.size() - have);
            if (r < 0) throw Error("read error on BGZF input");
            have += (size_t)r;
            size_t used = 0;
            gpu_check(kite_gpu_count_feed_bgzf(ctx, buf.data(), have, &used), "count");
            if (used) chunks++;
            memmove(buf.data(), buf.data() + used, have - used);
            have -= used;
            if (r == 0) break;  // a truncated last member is left unread (src/count/bgzf.rs:84-87)
            if (have == buf.size()) buf.resize(buf.size() * 2);
        }
    } else {
        // the next chunk is read on a second thread while this one is fed
        Source src(input, comp);
        Chunker ch(src, pl.chunk);
        Bytes c, nx;
        bool more = ch.next(c);
        while (more) {
            bool more_n = false;
            std::exception_ptr rerr;
            std::thread rd([&] {
                try {
                    more_n = ch.next(nx);
                } catch (...) {
                    rerr = std::current_exception();
                }
            });
            int rc;
            try {
                rc = kite_gpu_count_feed(ctx, c.data(), c.size());
            } catch (...) {
                rd.join();
                throw;
            }
            rd.join();
            gpu_check(rc, "count");
            if (rerr) std::rethrow_exception(rerr);
            chunks++;
            std::swap(c, nx);
            more = more_n;
        }
    }
    kite_count_stats st{};
    gpu_check(kite_gpu_count_finish(ctx, &st), "count");
    const auto t1 = std::chrono::steady_clock::now();
    uint64_t total_chunks = 0, total_rows = 0;
    if (!to_null) {
        Header h;
        h.k = (uint8_t)k;
        h.canonical = canonical ? 1 : 0;
        h.partitions = (uint32_t)pl.partitions;
        h.key_width = kw;
        h.codec_mask = (uint16_t)(CODEC_RAW | CODEC_DVLE | CODEC_ZSTD);
        h.min_count = mn;
        h.max_count = mx;
        Writer w(output, h, pl.staging, Writer::suggest_write_lanes((int)pl.workers));
        EmitCtx ec;
        ec.w = &w;
        ec.base_level = zstd_level;
        ec.k = k;
        ec.entry_size = kw == 64 ? 12 : (kw == 128 ? 20 : 36);
        if (zstd_level > 0) ec.start(std::max(1u, (unsigned)pl.workers));
        const int rc = kite_gpu_count_emit(ctx, emit_cb, &ec);
        ec.drain();
        gpu_check(rc, "count");
        std::tie(total_chunks, total_rows) = w.finalize();
        if (zstd_level > 0)
            fprintf(stderr, "[kite count] zstd chunks=%llu (level %d)\n", (unsigned long long)ec.zstd_chunks,
                    zstd_level);
    }
    const auto t2 = std::chrono::steady_clock::now();
    const double gpu_s = std::chrono::duration<double>(t1 - t0).count();
    const double wr_s = std::chrono::duration<double>(t2 - t1).count();
    fprintf(stderr,
            "[kite count] done: reads chunks=%llu (skipped %u), k-mers=%llu, unique=%llu, device %.1f ms, "
            "count %.3fs, write %.3fs, chunks=%llu rows=%llu\n",
            (unsigned long long)chunks, st.skipped_chunks, (unsigned long long)st.kmers,
            (unsigned long long)st.unique, st.device_ms, gpu_s, wr_s, (unsigned long long)total_chunks,
            (unsigned long long)total_rows);
    if (test_mode) {
        fprintf(stderr, "[kite count][test] extract=%.3fms split=%.3fms sort=%.3fms compact=%.3fms batches=%u\n",
                st.extract_ms, st.split_ms, st.sort_ms, st.compact_ms, st.batches);
    }
    return 0;
}

}  // namespace kh
