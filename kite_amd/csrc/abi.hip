// abi.hip -- the extern "C" boundary of libkite_gpu.so (declared in
// include/kite_gpu.h): context, count session, run-chunk emission.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/kite_gpu.h"
#include "count.hpp"
#include "corr.hpp"
#include "dump.hpp"
#include "encode.hpp"
#include "hostlink.hpp"
#include "inflate.hpp"
#include "merge.hpp"

namespace {

thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const kite::Error &e) {
        std::string m = e.what();
        int code = (m.find("out of memory") != std::string::npos) ? KITE_E_NOMEM : KITE_E_DEVICE;
        if (m.rfind("invalid", 0) == 0 || m.find("not supported") != std::string::npos)
            code = KITE_E_INVALID;
        return set_error(code, m);
    } catch (const std::bad_alloc &) {
        return set_error(KITE_E_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return set_error(KITE_E_DEVICE, e.what());
    }
}

// Page-locked host buffer that grows on demand.
struct PinnedBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void *ensure(size_t need) {
        if (need == 0) need = 256;
        if (need > bytes) {
            release();
            size_t want = need + need / 8;
            KITE_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
            bytes = want;
        }
        return ptr;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

// Depth of the pipelined batch's ring of input buffers and result slots
// (KITE_BATCH_RING overrides the default of 3 within 2..MAX_RING).
constexpr int MAX_RING = kite::CountEngine::OUT_SLOTS;

int batch_ring() {
    int r = 3;
    if (const char *e = getenv("KITE_BATCH_RING")) r = atoi(e);
    return std::min(MAX_RING, std::max(2, r));
}

// One result slot of the batched count: the host copy of a sample's keys,
// counts and partition offsets, and the event that ends its D2H copy.
struct BatchSlot {
    PinnedBuf keys, counts;
    kite::HostPinned enc_bytes, enc_minmax, enc_boff;  // chunk payloads (encoded mode)
    uint64_t enc_total = 0;
    std::vector<uint64_t> part_off;
    kite_count_stats st{};
    uint32_t sample = 0;
    hipEvent_t d2h_done = nullptr;
    hipEvent_t run_done = nullptr;  // the run that read input buffer [slot] ended
    hipEvent_t in_ready = nullptr;  // input buffer [slot] holds its sample
    hipEvent_t enc_done = nullptr;  // graph batch: the slot's chunk payload and summary are written
};

}  // namespace

// Working memory of kite_gpu_pca_noalloc, the entry point for callers that
// cannot allocate device memory: the matrix being rotated and the
// eigenvector accumulator (PCA_SCRATCH_N x PCA_SCRATCH_N FP64 each), then
// the eigenvalues.  Each GPU holds its own copy for the life of the process.
constexpr uint32_t PCA_SCRATCH_N = 64;
constexpr size_t PCA_SCRATCH_DOUBLES = 2ull * PCA_SCRATCH_N * PCA_SCRATCH_N + PCA_SCRATCH_N;
static __device__ __align__(256) double scratch[PCA_SCRATCH_DOUBLES];

namespace {
constexpr int SCRATCH_MAX_DEVICES = 256;
std::mutex g_scratch_init_mu;
// device address of `scratch` on each GPU (the host-side name is only a
// stand-in, never a device address)
double *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
// calls on one GPU share its array: they take that GPU's lock in turn
std::mutex g_scratch_use_mu[SCRATCH_MAX_DEVICES];

// Called with `device` current, once per device (later calls return at once).
void init_scratch(int device) {
    if (device < 0 || device >= SCRATCH_MAX_DEVICES)
        throw kite::Error("invalid device ordinal for the PCA working memory");
    std::lock_guard<std::mutex> g(g_scratch_init_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw kite::Error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(p);
}
}  // namespace

struct kite_gpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    std::unique_ptr<kite::CountEngine> count;
    std::unique_ptr<kite::MergeEngine> merge;
    std::unique_ptr<kite::CorrEngine> corr;
    std::unique_ptr<kite::DumpEngine> dump;

    // count session
    bool session = false;
    uint32_t k = 0, partitions = 0;
    int canonical = 0, words = 1;
    uint64_t batch_cap = 0;
    kite::DevBuf staging;
    uint64_t staged = 0;
    // plain chunks go through two page-locked buffers: a chunk is copied
    // (and its newlines counted) into one while the other's H2D is in flight
    kite::HostPinned feed_pin[2];
    hipEvent_t feed_done[2] = {nullptr, nullptr};
    int feed_cur = 0;
    const uint8_t *direct = nullptr;
    uint64_t direct_len = 0;
    bool device_result = false;
    std::vector<kite::EncodedRuns> host_runs;  // earlier batches, encoded on the device
    kite::EncodeBufs enc_bufs;
    kite_count_stats stats{};

    // batched count: copy streams, a ring of input buffers and result slots
    hipStream_t h2d = nullptr, d2h = nullptr;
    kite::DevBuf batch_in[MAX_RING];
    BatchSlot slots[MAX_RING];
    kite::EncodeBufs enc_slot[MAX_RING];  // device encode buffers of the result slots
    // profiling (kite_gpu_count_profile): per-kernel stats summed over the
    // samples of the last batch, and per-sample stage times
    bool profile = false;   // kernel events on the engine stream
    bool timeline = false;  // per-sample stage events on the batch's three streams
    bool kstats_from_batch = false;
    std::vector<kite::KernelStat> batch_kstats;
    std::vector<std::array<hipEvent_t, 6>> tl_events;  // grows, reused across batches
    hipEvent_t tl_start = nullptr;
    std::vector<double> tl_ms;  // 6 per sample of the last profiled batch
    // count settings chosen by kite_gpu_autotune, per (key words, partitions,
    // input size class)
    std::map<std::tuple<int, uint32_t, int>, kite::CountTunables> tuned;

    // dump: staged payloads of a partition group, pinned text pieces
    kite::HostPinned dump_stage, dump_text;
    kite::DevBuf dump_pay;

    // graph-captured batch (encoded chunks): one executable graph per (input
    // size class, ring slot), valid while the shape, buffers and settings
    // it was captured with stay the same
    struct BatchGraph {
        hipGraphExec_t count = nullptr, encode = nullptr;
        kite::CountTunables tun;
        const void *in = nullptr, *out = nullptr, *meta = nullptr, *len = nullptr;  // the slot's buffers they use
        uint64_t max_len = 0;  // the count's grids cover inputs up to this length
    };
    std::map<std::pair<int, int>, BatchGraph> graphs;
    std::tuple<uint32_t, int, uint32_t, uint64_t> graph_shape{0, -1, 0, 0};  // k, canonical, P, input capacity
    kite::EncodeBufs graph_enc;                // scratch of the planned chunk encode
    kite::DevBuf graph_bytes[MAX_RING];        // chunk payloads of each ring slot
    kite::HostPinned graph_meta[MAX_RING];     // per slot: a length word (warm-up count), then the summary
    kite::DevBuf graph_dlen;                   // per slot: the input length the count graph reads
    std::unique_ptr<kite::HostLink> link;      // D2H of the payloads on SDMA engines picked here
    bool link_tried = false;
    uint64_t graph_fallbacks = 0;              // samples a planned count handed to the classic path

    // merge of stored runs: per-input tables of a partition group, their
    // union, its u32 counts and chunk encode buffers, pinned bounce buffers
    std::vector<kite::DevTable> merge_tabs;
    kite::DevTable merge_res;
    kite::DevBuf merge_c32;
    kite::EncodeBufs merge_enc;
    kite::HostPinned merge_in[2], merge_out[2];

    // BGZF input inflated on the device into `staging`: the staged bytes
    // then hold whole records plus the start of the next one
    bool bgzf_feed = false;
    kite::HostPinned bgzf_host;
    kite::DevBuf bgzf_comp, bgzf_blocks, bgzf_misc, bgzf_tmp;
    double inflate_ms = 0;
};

namespace {

void bind(kite_gpu_ctx *c) { KITE_HIP(hipSetDevice(c->device)); }

// The context's device-to-host copy link (SDMA engines picked by the
// library), opened on first use; nullptr when KITE_D2H_ENGINES=0 or when no
// engine qualifies (copies then take hipMemcpyAsync).
kite::HostLink *host_link(kite_gpu_ctx *c) {
    const char *e = getenv("KITE_D2H_ENGINES");
    if (e && strcmp(e, "0") == 0) return nullptr;
    if (!c->link_tried) {
        c->link_tried = true;
        c->link = std::make_unique<kite::HostLink>(c->device, MAX_RING);
    }
    return c->link && c->link->usable() ? c->link.get() : nullptr;
}

// Hands the chunks of partitions [0, np) to fn in partition order (fn sees
// partition p0 + p): payload bytes on the device at d_bytes + boff[p], the
// byte and row offsets and key ranges on the host.  The device work that
// wrote the payloads has finished.  They leave through the context's two
// page-locked piece buffers in pieces of whole partitions up to 64 MiB (a
// larger partition is a piece of its own), two pieces in flight on the
// library's D2H engine while fn takes the one before; piece q + 2 is queued
// once fn has taken piece q.  Returns -1, or the partition (p0 + p) whose
// callback returned non-zero.
int64_t emit_device_chunks(kite_gpu_ctx *ctx, const uint8_t *d_bytes, const uint64_t *boff, const uint64_t *roff,
                           const uint64_t *mm, uint32_t np, uint32_t p0, uint32_t codec, kite_append_chunk_fn fn,
                           void *user) {
    hipStream_t s = ctx->stream;
    constexpr uint64_t PIECE = 64ull << 20;
    std::vector<std::pair<uint32_t, uint32_t>> pieces;
    for (uint32_t a = 0; a < np;) {
        uint32_t b = a + 1;
        while (b < np && boff[b + 1] - boff[a] <= PIECE) b++;
        pieces.emplace_back(a, b);
        a = b;
    }
    kite::HostLink *link = host_link(ctx);
    hipEvent_t done[2];
    for (auto &x : done) KITE_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    struct Ev2 {
        hipEvent_t *e;
        kite::HostLink *link;
        ~Ev2() {
            if (link) link->quiesce();
            (void)hipEventDestroy(e[0]);
            (void)hipEventDestroy(e[1]);
        }
    } ev2{done, link};
    for (int b = 0; b < 2; b++) ctx->merge_out[b].ensure(PIECE);
    auto issue = [&](size_t q) {
        const uint64_t a = boff[pieces[q].first], n = boff[pieces[q].second] - a;
        uint8_t *pin = static_cast<uint8_t *>(ctx->merge_out[q & 1].ensure(std::max<uint64_t>(n, PIECE)));
        if (link) {
            if (n) link->d2h((int)(q & 1), pin, d_bytes + a, n);
        } else {
            if (n) KITE_HIP(hipMemcpyAsync(pin, d_bytes + a, n, hipMemcpyDeviceToHost, s));
            KITE_HIP(hipEventRecord(done[q & 1], s));
        }
    };
    for (size_t q = 0; q < std::min<size_t>(2, pieces.size()); q++) issue(q);
    for (size_t q = 0; q < pieces.size(); q++) {
        if (link) link->wait((int)(q & 1));
        else KITE_HIP(hipEventSynchronize(done[q & 1]));
        const uint8_t *pin = ctx->merge_out[q & 1].as<uint8_t>();
        const uint64_t base = boff[pieces[q].first];
        for (uint32_t p = pieces[q].first; p < pieces[q].second; p++) {
            const uint64_t nr = roff[p + 1] - roff[p];
            if (!nr) continue;
            const uint64_t len = boff[p + 1] - boff[p];
            if (fn(user, p0 + p, codec, nr, len, &mm[(size_t)p * 8], &mm[(size_t)p * 8 + 4], pin + (boff[p] - base),
                   (size_t)len) != 0) {
                (void)hipStreamSynchronize(s);
                return (int64_t)(p0 + p);
            }
        }
        if (q + 2 < pieces.size()) issue(q + 2);
    }
    KITE_HIP(hipStreamSynchronize(s));
    return -1;
}

// Input size classes of the tuned settings: below 64 MiB, 512 MiB, 4 GiB, above.
int size_class(uint64_t len) {
    return len < (64ull << 20) ? 0 : (len < (512ull << 20) ? 1 : (len < (4ull << 30) ? 2 : 3));
}

// The settings a count of `len` input bytes starts from: the autotuner's
// choice for this key width, partition count and size class, if any.
void apply_tuned(kite_gpu_ctx *c, uint64_t len) {
    auto it = c->tuned.find(std::make_tuple(c->words, c->partitions, size_class(len)));
    c->count->set_base_tunables(it != c->tuned.end() ? it->second : kite::CountTunables{});
}

// Keep the staged FASTQ bytes when the staging buffer grows.
void staging_reserve(kite_gpu_ctx *c, uint64_t need) {
    if (need <= c->staging.bytes) return;
    uint64_t want = std::max<uint64_t>(need, std::min<uint64_t>(c->batch_cap, 2 * c->staging.bytes));
    want = std::max<uint64_t>(want, need);
    void *np = nullptr;
    KITE_HIP(hipMalloc(&np, want));
    if (c->staged)
        KITE_HIP(hipMemcpyAsync(np, c->staging.ptr, c->staged, hipMemcpyDeviceToDevice, c->stream));
    KITE_HIP(hipStreamSynchronize(c->stream));
    c->staging.release();
    c->staging.ptr = np;
    c->staging.bytes = want;
}

// Encodes the device result into chunk payloads (only those cross PCIe).
void encode_device_result(kite_gpu_ctx *c, kite::EncodedRuns &out) {
    kite::encode_runs_device(c->count->keys(), c->count->counts(), c->count->d_part_off(), c->count->part_off(),
                             c->partitions, c->words, c->stream, c->enc_bufs, out);
}

void stash_device_result(kite_gpu_ctx *c) {
    if (!c->device_result) return;
    kite::EncodedRuns r;
    encode_device_result(c, r);
    c->host_runs.push_back(std::move(r));
    c->device_result = false;
}

void run_batch(kite_gpu_ctx *c, const uint8_t *ptr, uint64_t len) {
    stash_device_result(c);
    c->kstats_from_batch = false;
    c->count->set_timing(true);
    apply_tuned(c, len);
    c->count->run(ptr, len, (int)c->k, c->canonical, c->partitions);
    c->device_result = true;
    const kite::StageTimes &t = c->count->times();
    c->stats.kmers += c->count->n_kmers();
    c->stats.unique += c->count->n_unique();
    c->stats.buckets = c->count->buckets();
    c->stats.split_levels = (uint32_t)c->count->levels();
    c->stats.batches++;
    c->stats.device_ms += t.total_ms;
    c->stats.extract_ms = t.extract_ms;
    c->stats.split_ms = t.split_ms;
    c->stats.sort_ms = t.sort_ms;
    c->stats.compact_ms = t.compact_ms;
}

// A batch of inflated BGZF text starts at a record boundary; like a reader
// chunk it yields no k-mers unless it starts with '@' (src/count/parser.rs:
// 24-27 with the error ignored at src/count/worker.rs:217).
void run_bgzf_batch(kite_gpu_ctx *c, uint64_t len) {
    uint8_t first = 0;
    KITE_HIP(hipMemcpy(&first, c->staging.ptr, 1, hipMemcpyDeviceToHost));
    if (first != '@') {
        c->stats.skipped_chunks++;
        return;
    }
    run_batch(c, c->staging.as<uint8_t>(), len);
}

void flush_pending(kite_gpu_ctx *c) {
    if (c->direct) {
        const uint8_t *p = c->direct;
        uint64_t n = c->direct_len;
        c->direct = nullptr;
        c->direct_len = 0;
        run_batch(c, p, n);
    } else if (c->staged) {
        if (c->bgzf_feed) run_bgzf_batch(c, c->staged);
        else run_batch(c, c->staging.as<uint8_t>(), c->staged);
        c->staged = 0;
    }
}

// Inflates the whole BGZF members ms[m0, m1) of `data` into dst (device);
// returns their output bytes.
uint64_t inflate_members(kite_gpu_ctx *c, const uint8_t *data, const std::vector<kite::BgzfMember> &ms, size_t m0,
                         size_t m1, uint8_t *dst) {
    if (m0 >= m1) return 0;
    const uint64_t a = ms[m0].offset, b = ms[m1 - 1].offset + ms[m1 - 1].size;
    uint8_t *h = static_cast<uint8_t *>(c->bgzf_host.ensure(b - a + 16));
    memcpy(h, data + a, b - a);
    memset(h + (b - a), 0, 16);
    uint8_t *dc = static_cast<uint8_t *>(c->bgzf_comp.ensure(b - a + 16));
    KITE_HIP(hipMemcpyAsync(dc, h, b - a + 16, hipMemcpyHostToDevice, c->stream));
    std::vector<kite::InflateBlock> bl;
    bl.reserve(m1 - m0);
    uint64_t out = 0;
    for (size_t m = m0; m < m1; m++) {
        bl.push_back(kite::InflateBlock{ms[m].offset - a + ms[m].data, out, ms[m].clen, ms[m].isize});
        out += ms[m].isize;
    }
    uint32_t *err = static_cast<uint32_t *>(c->bgzf_misc.ensure(64));
    KITE_HIP(hipMemsetAsync(err, 0, 4, c->stream));
    hipEvent_t e0, e1;
    KITE_HIP(hipEventCreate(&e0));
    KITE_HIP(hipEventCreate(&e1));
    KITE_HIP(hipEventRecord(e0, c->stream));
    kite::inflate_blocks(dc, bl, dst, err, c->bgzf_blocks, c->stream);
    KITE_HIP(hipEventRecord(e1, c->stream));
    uint32_t he = 0;
    KITE_HIP(hipMemcpyAsync(&he, err, 4, hipMemcpyDeviceToHost, c->stream));
    KITE_HIP(hipStreamSynchronize(c->stream));
    float ms_ = 0;
    (void)hipEventElapsedTime(&ms_, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->inflate_ms += ms_;
    if (he & 1u) throw kite::Error("invalid deflate data in a BGZF member");
    if (he & 2u) throw kite::Error("BGZF member output differs from its ISIZE");
    return out;
}

// Runs the staged BGZF text up to its last record boundary (the end of a
// 4-line group, find_fastq_split_4lines, src/count/bgzf.rs:109-121) and keeps
// the rest at the front of the staging buffer.  False if the staged text
// holds no whole record.
bool cut_and_run_bgzf(kite_gpu_ctx *c) {
    const uint64_t n = c->staged;
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(c->bgzf_misc.as<uint8_t>() + 8);
    KITE_HIP(hipMemsetAsync(cnt, 0, 8, c->stream));
    kite::count_newlines_device(c->staging.as<uint8_t>(), n, cnt, c->stream);
    unsigned long long total = 0;
    KITE_HIP(hipMemcpyAsync(&total, cnt, 8, hipMemcpyDeviceToHost, c->stream));
    KITE_HIP(hipStreamSynchronize(c->stream));
    if (total < 4) return false;
    // the cut follows the (r+1)-th newline from the end, r = total mod 4
    const uint64_t need = total % 4 + 1;
    uint64_t cut = 0;
    std::vector<uint8_t> win;
    for (uint64_t w = std::min<uint64_t>(n, 1 << 16);; w = std::min<uint64_t>(n, 4 * w)) {
        win.resize(w);
        KITE_HIP(hipMemcpy(win.data(), c->staging.as<uint8_t>() + (n - w), w, hipMemcpyDeviceToHost));
        uint64_t seen = 0;
        for (uint64_t q = w; q-- > 0;)
            if (win[q] == '\n' && ++seen == need) {
                cut = n - w + q + 1;
                break;
            }
        if (cut || w == n) break;
    }
    if (!cut) return false;
    run_bgzf_batch(c, cut);
    const uint64_t tail = n - cut;
    if (tail) {
        uint8_t *tmp = static_cast<uint8_t *>(c->bgzf_tmp.ensure(tail));
        KITE_HIP(hipMemcpyAsync(tmp, c->staging.as<uint8_t>() + cut, tail, hipMemcpyDeviceToDevice, c->stream));
        KITE_HIP(hipMemcpyAsync(c->staging.ptr, tmp, tail, hipMemcpyDeviceToDevice, c->stream));
        KITE_HIP(hipStreamSynchronize(c->stream));
    }
    c->staged = tail;
    return true;
}

// Copies n bytes from src to dst and returns the number of '\n' among them,
// eight bytes at a time (zero bytes of w ^ 0x0A.. counted exactly), on up
// to 16 host threads for large chunks.
uint64_t copy_count_newlines(uint8_t *dst, const uint8_t *src, size_t n) {
    auto part = [](uint8_t *d, const uint8_t *s, size_t len) -> uint64_t {
        uint64_t c = 0;
        size_t i = 0;
        for (; i + 8 <= len; i += 8) {
            uint64_t w;
            memcpy(&w, s + i, 8);
            memcpy(d + i, &w, 8);
            const uint64_t x = w ^ 0x0A0A0A0A0A0A0A0Aull;
            const uint64_t z = ~(((x & 0x7F7F7F7F7F7F7F7Full) + 0x7F7F7F7F7F7F7F7Full) | x | 0x7F7F7F7F7F7F7F7Full);
            c += (uint64_t)__builtin_popcountll(z);
        }
        for (; i < len; i++) {
            d[i] = s[i];
            c += s[i] == '\n';
        }
        return c;
    };
    const unsigned nt = (unsigned)std::min<uint64_t>(
        std::max<uint64_t>(1, n >> 23), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    if (nt == 1) return part(dst, src, n);
    std::vector<uint64_t> cnt(nt, 0);
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const size_t a = std::min(n, (size_t)t * per), b = std::min(n, a + per);
        th.emplace_back([&, t, a, b] { cnt[t] = part(dst + a, src + a, b - a); });
    }
    for (auto &x : th) x.join();
    uint64_t c = 0;
    for (uint64_t v : cnt) c += v;
    return c;
}

}  // namespace

extern "C" {

const char *kite_gpu_last_error(void) { return g_last_error.c_str(); }

int kite_gpu_abi_version(void) { return KITE_GPU_ABI_VERSION; }

int kite_gpu_open(int device, uint64_t hbm_budget_bytes, kite_gpu_ctx **out) {
    if (!out) return set_error(KITE_E_INVALID, "invalid argument: out is NULL");
    *out = nullptr;
    return guarded([&]() -> int {
        int n = 0;
        KITE_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n)
            return set_error(KITE_E_INVALID, "invalid device ordinal " + std::to_string(device));
        auto c = std::make_unique<kite_gpu_ctx>();
        c->device = device;
        KITE_HIP(hipSetDevice(device));
        init_scratch(device);
        KITE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        size_t free_b = 0, total_b = 0;
        KITE_HIP(hipMemGetInfo(&free_b, &total_b));
        c->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(free_b * 0.9);
        c->count = std::make_unique<kite::CountEngine>(c->stream);
        c->merge = std::make_unique<kite::MergeEngine>(c->stream);
        c->corr = std::make_unique<kite::CorrEngine>(c->stream);
        c->dump = std::make_unique<kite::DumpEngine>(c->stream, c->merge.get());
        *out = c.release();
        return KITE_OK;
    });
}

void kite_gpu_close(kite_gpu_ctx *ctx) {
    if (!ctx) return;
    try {
        bind(ctx);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->link.reset();
        for (auto &g : ctx->graphs)
            for (hipGraphExec_t e : {g.second.count, g.second.encode})
                if (e) (void)hipGraphExecDestroy(e);
        ctx->graphs.clear();
        ctx->count.reset();
        ctx->dump.reset();
        ctx->merge.reset();
        ctx->corr.reset();
        ctx->staging.release();
        for (int b = 0; b < 2; b++) {
            ctx->feed_pin[b].release();
            if (ctx->feed_done[b]) (void)hipEventDestroy(ctx->feed_done[b]);
        }
        ctx->bgzf_host.release();
        ctx->bgzf_comp.release();
        ctx->bgzf_blocks.release();
        ctx->bgzf_misc.release();
        ctx->bgzf_tmp.release();
        ctx->dump_stage.release();
        ctx->dump_text.release();
        ctx->dump_pay.release();
        for (auto &ev : ctx->tl_events)
            for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (ctx->tl_start) (void)hipEventDestroy(ctx->tl_start);
        for (int b = 0; b < MAX_RING; b++) {
            ctx->batch_in[b].release();
            ctx->slots[b].keys.release();
            ctx->slots[b].counts.release();
            ctx->slots[b].enc_bytes.release();
            ctx->slots[b].enc_minmax.release();
            ctx->slots[b].enc_boff.release();
            for (hipEvent_t e : {ctx->slots[b].d2h_done, ctx->slots[b].run_done, ctx->slots[b].in_ready,
                                 ctx->slots[b].enc_done})
                if (e) (void)hipEventDestroy(e);
        }
        if (ctx->h2d) (void)hipStreamDestroy(ctx->h2d);
        if (ctx->d2h) (void)hipStreamDestroy(ctx->d2h);
        (void)hipStreamDestroy(ctx->stream);
        kite::free_retired();
    } catch (...) {
    }
    delete ctx;
}

int kite_gpu_count_begin(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    if (k < 1 || k > 128)
        return set_error(KITE_E_INVALID,
                         "invalid k-mer length: --k must be between 1 and 128.\nGot: " + std::to_string(k));
    if (partitions == 0 || (partitions & (partitions - 1)))
        return set_error(KITE_E_INVALID,
                         "invalid partition count: --partitions must be a power of 2.\nGot: " +
                             std::to_string(partitions));
    return guarded([&]() -> int {
        bind(ctx);
        ctx->session = true;
        ctx->k = k;
        ctx->canonical = canonical ? 1 : 0;
        ctx->partitions = partitions;
        ctx->words = k <= 31 ? 1 : (k <= 64 ? 2 : 4);
        // device bytes per FASTQ byte, assuming at most one k-mer per two input bytes
        uint64_t per_byte = (3ull * 8 * ctx->words + 8) / 2 + 2;
        ctx->batch_cap = std::max<uint64_t>(1ull << 20, ctx->budget / per_byte);
        ctx->staged = 0;
        ctx->direct = nullptr;
        ctx->direct_len = 0;
        ctx->device_result = false;
        ctx->host_runs.clear();
        ctx->stats = kite_count_stats{};
        ctx->stats.words = (uint32_t)ctx->words;
        ctx->bgzf_feed = false;
        ctx->inflate_ms = 0;
        return KITE_OK;
    });
}

int kite_gpu_count_feed(kite_gpu_ctx *ctx, const uint8_t *fastq, size_t len) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    if (ctx->bgzf_feed && ctx->staged)
        return set_error(KITE_E_INVALID, "invalid state: plain chunks after BGZF input in one count session");
    if (len == 0) return KITE_OK;
    if (!fastq) return set_error(KITE_E_INVALID, "invalid argument: fastq is NULL");
    if (fastq[0] != '@') {
        ctx->stats.skipped_chunks++;
        return KITE_OK;
    }
    return guarded([&]() -> int {
        bind(ctx);
        if (ctx->direct) flush_pending(ctx);
        // this pinned buffer's previous H2D must be done before it is reused
        const int cur = ctx->feed_cur;
        if (ctx->feed_done[cur]) KITE_HIP(hipEventSynchronize(ctx->feed_done[cur]));
        uint8_t *pin = static_cast<uint8_t *>(ctx->feed_pin[cur].ensure(len + 4));
        const uint64_t nl = copy_count_newlines(pin, fastq, len);
        const bool ends_nl = fastq[len - 1] == '\n';
        uint32_t pad = (uint32_t)((4 - nl % 4) % 4);
        if (!ends_nl && pad == 0) pad = 4;
        for (uint32_t q = 0; q < pad; q++) pin[len + q] = '\n';
        const uint64_t total = len + pad;
        if (ctx->staged && ctx->staged + total > ctx->batch_cap) flush_pending(ctx);
        staging_reserve(ctx, ctx->staged + total);
        uint8_t *dst = ctx->staging.as<uint8_t>() + ctx->staged;
        // on the engine stream, ahead of the kernels that read the staging
        KITE_HIP(hipMemcpyAsync(dst, pin, total, hipMemcpyHostToDevice, ctx->stream));
        if (!ctx->feed_done[cur]) KITE_HIP(hipEventCreateWithFlags(&ctx->feed_done[cur], hipEventDisableTiming));
        KITE_HIP(hipEventRecord(ctx->feed_done[cur], ctx->stream));
        ctx->feed_cur = cur ^ 1;
        ctx->staged += total;
        ctx->stats.input_bytes += len;
        return KITE_OK;
    });
}

int kite_gpu_count_feed_device(kite_gpu_ctx *ctx, const void *d_fastq, size_t len) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    if (len == 0) return KITE_OK;
    if (!d_fastq) return set_error(KITE_E_INVALID, "invalid argument: d_fastq is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        flush_pending(ctx);
        ctx->direct = static_cast<const uint8_t *>(d_fastq);
        ctx->direct_len = len;
        ctx->stats.input_bytes += len;
        return KITE_OK;
    });
}

int kite_gpu_count_feed_bgzf(kite_gpu_ctx *ctx, const uint8_t *data, size_t len, size_t *consumed) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    if (consumed) *consumed = 0;
    if (len == 0) return KITE_OK;
    if (!data) return set_error(KITE_E_INVALID, "invalid argument: data is NULL");
    if (!ctx->bgzf_feed && (ctx->staged || ctx->direct))
        return set_error(KITE_E_INVALID, "invalid state: BGZF input after plain chunks in one count session");
    return guarded([&]() -> int {
        bind(ctx);
        size_t used = 0;
        const std::vector<kite::BgzfMember> ms = kite::scan_bgzf(data, len, &used);
        ctx->bgzf_feed = true;
        size_t m = 0;
        while (m < ms.size()) {
            size_t j = m;
            uint64_t add = 0;
            while (j < ms.size() && ctx->staged + add + ms[j].isize <= ctx->batch_cap) add += ms[j++].isize;
            if (j == m) {
                // the batch is full: run it up to its last record, or grow it
                // when it holds no whole record
                if (cut_and_run_bgzf(ctx)) continue;
                add = ms[m].isize;
                j = m + 1;
            }
            staging_reserve(ctx, ctx->staged + add + 16);
            const uint64_t out = inflate_members(ctx, data, ms, m, j, ctx->staging.as<uint8_t>() + ctx->staged);
            ctx->staged += out;
            ctx->stats.input_bytes += out;
            m = j;
        }
        if (consumed) *consumed = used;
        return KITE_OK;
    });
}

int kite_gpu_inflate_bgzf(kite_gpu_ctx *ctx, const uint8_t *data, size_t len, uint8_t *out, size_t cap,
                          size_t *out_len, double *device_ms) {
    if (!ctx || !out_len) return set_error(KITE_E_INVALID, "invalid argument: NULL pointer");
    if (len && !data) return set_error(KITE_E_INVALID, "invalid argument: data is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        const std::vector<kite::BgzfMember> ms = kite::scan_bgzf(data, len, nullptr);
        uint64_t total = 0;
        for (const auto &m : ms) total += m.isize;
        *out_len = total;
        if (total > cap) return set_error(KITE_E_INVALID, "invalid argument: output buffer too small");
        if (!out && total) return set_error(KITE_E_INVALID, "invalid argument: out is NULL");
        const double before = ctx->inflate_ms;
        uint8_t *dev = static_cast<uint8_t *>(ctx->bgzf_tmp.ensure(total + 16));
        inflate_members(ctx, data, ms, 0, ms.size(), dev);
        if (total) KITE_HIP(hipMemcpy(out, dev, total, hipMemcpyDeviceToHost));
        if (device_ms) *device_ms = ctx->inflate_ms - before;
        return KITE_OK;
    });
}

int kite_gpu_count_finish(kite_gpu_ctx *ctx, kite_count_stats *stats) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    return guarded([&]() -> int {
        bind(ctx);
        flush_pending(ctx);
        if (stats) *stats = ctx->stats;
        return KITE_OK;
    });
}

int kite_gpu_count_result_size(kite_gpu_ctx *ctx, uint64_t *n_unique, uint32_t *words,
                               uint32_t *partitions) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    if (n_unique) *n_unique = ctx->device_result ? ctx->count->n_unique() : 0;
    if (words) *words = (uint32_t)ctx->words;
    if (partitions) *partitions = ctx->partitions;
    return KITE_OK;
}

int kite_gpu_count_download(kite_gpu_ctx *ctx, uint64_t *keys, uint32_t *counts, uint64_t *part_off) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    return guarded([&]() -> int {
        bind(ctx);
        const uint64_t u = ctx->device_result ? ctx->count->n_unique() : 0;
        if (keys && u)
            KITE_HIP(hipMemcpy(keys, ctx->count->keys(), u * 8 * ctx->words, hipMemcpyDeviceToHost));
        if (counts && u)
            KITE_HIP(hipMemcpy(counts, ctx->count->counts(), u * 4, hipMemcpyDeviceToHost));
        if (part_off) {
            if (ctx->device_result) {
                const auto &po = ctx->count->part_off();
                std::copy(po.begin(), po.end(), part_off);
            } else {
                std::fill(part_off, part_off + ctx->partitions + 1, 0);
            }
        }
        return KITE_OK;
    });
}

int kite_gpu_count_device_result(kite_gpu_ctx *ctx, const uint64_t **d_keys, const uint32_t **d_counts,
                                 const uint64_t **part_off) {
    if (!ctx || !ctx->session || !ctx->device_result)
        return set_error(KITE_E_INVALID, "invalid state: no device result");
    if (d_keys) *d_keys = ctx->count->keys();
    if (d_counts) *d_counts = ctx->count->counts();
    if (part_off) *part_off = ctx->count->part_off().data();
    return KITE_OK;
}

int kite_gpu_count_emit(kite_gpu_ctx *ctx, kite_append_chunk_fn fn, void *user) {
    if (!ctx || !ctx->session) return set_error(KITE_E_INVALID, "invalid state: no count session");
    if (!fn) return set_error(KITE_E_INVALID, "invalid argument: callback is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        const uint32_t P = ctx->partitions;
        // earlier device batches, encoded into host memory when the next began
        for (const kite::EncodedRuns &r : ctx->host_runs) {
            for (uint32_t p = 0; p < P; p++) {
                const uint64_t rows = r.rows_off[p + 1] - r.rows_off[p];
                if (!rows) continue;
                const uint64_t b0 = r.byte_off[p], len = r.byte_off[p + 1] - b0;
                int rc = fn(user, p, r.codec, rows, len, &r.minmax[(size_t)p * 8], &r.minmax[(size_t)p * 8 + 4],
                            r.bytes.data() + b0, len);
                if (rc)
                    return set_error(KITE_E_CALLBACK,
                                     "append callback failed for partition " + std::to_string(p));
            }
        }
        // the last batch: encoded on the device, its chunks streamed out
        if (ctx->device_result) {
            const int W = ctx->words;
            const std::vector<uint64_t> &roff = ctx->count->part_off();
            uint64_t *hb = static_cast<uint64_t *>(ctx->enc_bufs.host_byte_off.ensure(((size_t)P + 1) * 8));
            const kite::EncodeDevice ed =
                kite::encode_on_device(ctx->count->keys(), ctx->count->counts(), ctx->count->d_part_off(),
                                       roff.empty() ? 0 : roff[P], P, W, ctx->stream, ctx->enc_bufs, hb);
            const std::vector<uint64_t> boff(hb, hb + P + 1);
            std::vector<uint64_t> mm((size_t)P * 8);
            KITE_HIP(hipMemcpyAsync(mm.data(), ed.minmax, (size_t)P * 64, hipMemcpyDeviceToHost, ctx->stream));
            KITE_HIP(hipStreamSynchronize(ctx->stream));
            const int64_t bad = emit_device_chunks(ctx, ed.bytes, boff.data(), roff.data(), mm.data(), P, 0,
                                                   W == 4 ? KITE_CODEC_RAW : KITE_CODEC_DVLE, fn, user);
            if (bad >= 0)
                return set_error(KITE_E_CALLBACK, "append callback failed for partition " + std::to_string(bad));
        }
        return KITE_OK;
    });
}

namespace {
int load_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t P, uint32_t n, const uint64_t *const *keys,
                const uint64_t *const *counts, const uint64_t *const *part_off, std::vector<kite::DevTable> &tabs) {
    if (words != 1 && words != 2 && words != 4) return set_error(KITE_E_INVALID, "invalid argument: words must be 1, 2 or 4");
    if (P == 0 || n == 0 || !keys || !counts || !part_off)
        return set_error(KITE_E_INVALID, "invalid argument: empty table list");
    tabs = std::vector<kite::DevTable>(n);
    for (uint32_t i = 0; i < n; i++) {
        kite::HostTable ht;
        ht.keys = keys[i];
        ht.counts = counts[i];
        ht.part_off.assign(part_off[i], part_off[i] + P + 1);
        if (ht.part_off[0] != 0) return set_error(KITE_E_INVALID, "invalid argument: part_off must start at 0");
        for (uint32_t p = 0; p < P; p++)
            if (ht.part_off[p + 1] < ht.part_off[p])
                return set_error(KITE_E_INVALID, "invalid argument: part_off must be non-decreasing");
        ctx->merge->upload_table(ht, tabs[i], (int)words);
    }
    return KITE_OK;
}
}  // namespace

int kite_gpu_merge_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t partitions, uint32_t n_inputs,
                          const uint64_t *const *keys, const uint64_t *const *counts,
                          const uint64_t *const *part_off, uint64_t *out_keys, uint64_t *out_counts,
                          uint64_t *out_part_off) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        std::vector<kite::DevTable> tabs;
        int rc = load_tables(ctx, words, partitions, n_inputs, keys, counts, part_off, tabs);
        if (rc) return rc;
        kite::DevTable u;
        ctx->merge->merge_all(tabs, partitions, (int)words, u);
        ctx->merge->download(u, (int)words, partitions, out_keys, out_counts, out_part_off);
        return KITE_OK;
    });
}

int kite_gpu_union_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t partitions, uint32_t n_inputs,
                          const uint64_t *const *keys, const uint64_t *const *counts,
                          const uint64_t *const *part_off, uint64_t *out_keys, uint32_t *out_matrix,
                          uint64_t *out_part_off) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        std::vector<kite::DevTable> tabs;
        int rc = load_tables(ctx, words, partitions, n_inputs, keys, counts, part_off, tabs);
        if (rc) return rc;
        kite::DevTable u;
        ctx->merge->merge_all(tabs, partitions, (int)words, u);
        kite::DevBuf m;
        ctx->merge->union_matrix(tabs, u, partitions, (int)words, m);
        ctx->merge->download(u, (int)words, partitions, out_keys, nullptr, out_part_off);
        if (out_matrix && u.rows)
            KITE_HIP(hipMemcpy(out_matrix, m.ptr, u.rows * n_inputs * 4, hipMemcpyDeviceToHost));
        return KITE_OK;
    });
}

namespace {
void copy_out(const std::vector<double> &v, double *dst) {
    if (dst && !v.empty()) std::copy(v.begin(), v.end(), dst);
}

// d_cells: N x n count matrix on the device, u32 (cell_bytes 4) or u64 (8).
int corr_device(kite_gpu_ctx *ctx, const void *d_cells, int cell_bytes, uint64_t N, uint32_t n, uint32_t methods,
                uint32_t flags, double *pearson, double *spearman, double *cov, double *pearson_fp64,
                double *spearman_fp64, kite_corr_stats *stats) {
    if (n < 1 || n > (uint32_t)kite::MAX_SAMPLES)
        return set_error(KITE_E_INVALID, "invalid argument: analyze supports 1.." +
                                             std::to_string(kite::MAX_SAMPLES) + " samples");
    // rank codes: the Gram reads 1, 2 or 4 bytes a cell
    kite::DevBuf codes;
    const int cb = ctx->corr->prepare(d_cells, cell_bytes, N, (int)n, codes);
    kite::CorrRequest req;
    req.methods = methods;
    req.fp64_baseline = (flags & KITE_CORR_FP64_BASELINE) != 0;
    req.with_lolo = (flags & KITE_CORR_LOLO) != 0;
    kite::CorrOutput out;
    ctx->corr->run(codes.ptr, N, (int)n, req, out);
    copy_out(out.pearson, pearson);
    copy_out(out.spearman, spearman);
    copy_out(out.cov, cov);
    copy_out(out.pearson_fp64, pearson_fp64);
    copy_out(out.spearman_fp64, spearman_fp64);
    if (stats) {
        stats->gram_ms_pearson = out.gram_ms_pearson;
        stats->gram_ms_spearman = out.gram_ms_spearman;
        stats->fp64_ms_pearson = out.fp64_ms_pearson;
        stats->fp64_ms_spearman = out.fp64_ms_spearman;
        stats->rows = N;
        stats->cols = n;
        stats->code_bytes = (uint32_t)cb;
        stats->gram_ms_total = out.gram_ms_total;
        stats->fp64_ms_total = out.fp64_ms_total;
        stats->passes = (uint32_t)out.passes;
        stats->fp64_passes = (uint32_t)out.fp64_passes;
    }
    return KITE_OK;
}
}  // namespace

int kite_gpu_corr_counts(kite_gpu_ctx *ctx, const uint32_t *counts, uint64_t N, uint32_t n, uint32_t methods,
                         uint32_t flags, double *pearson, double *spearman, double *cov, double *pearson_fp64,
                         double *spearman_fp64, kite_corr_stats *stats) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    if (!counts && N) return set_error(KITE_E_INVALID, "invalid argument: counts is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        kite::DevBuf d;
        d.ensure(std::max<uint64_t>(N * n, 1) * 4);
        if (N) KITE_HIP(hipMemcpy(d.ptr, counts, N * n * 4, hipMemcpyHostToDevice));
        return corr_device(ctx, d.ptr, 4, N, n, methods, flags, pearson, spearman, cov, pearson_fp64,
                           spearman_fp64, stats);
    });
}

int kite_gpu_analyze_tables(kite_gpu_ctx *ctx, uint32_t words, uint32_t partitions, uint32_t n_inputs,
                            const uint64_t *const *keys, const uint64_t *const *counts,
                            const uint64_t *const *part_off, uint32_t methods, uint32_t flags,
                            uint64_t *out_keys, uint32_t *out_matrix, uint64_t *out_part_off, double *pearson,
                            double *spearman, double *cov, double *pearson_fp64, double *spearman_fp64,
                            kite_corr_stats *stats) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    return guarded([&]() -> int {
        bind(ctx);
        std::vector<kite::DevTable> tabs;
        int rc = load_tables(ctx, words, partitions, n_inputs, keys, counts, part_off, tabs);
        if (rc) return rc;
        // counts above u32::MAX keep their exact u64 value in the matrix the
        // correlation reads (the reference sums in u64, kiteops/mod.rs:433-470)
        const bool wide = ctx->merge->max_count(tabs) > 0xFFFFFFFFull;
        kite::DevTable u;
        ctx->merge->merge_all(tabs, partitions, (int)words, u);
        kite::DevBuf m;
        ctx->merge->union_matrix(tabs, u, partitions, (int)words, m, wide ? 8 : 4);
        tabs.clear();
        ctx->merge->download(u, (int)words, partitions, out_keys, nullptr, out_part_off);
        if (out_matrix && u.rows) {
            if (!wide) {
                KITE_HIP(hipMemcpy(out_matrix, m.ptr, u.rows * n_inputs * 4, hipMemcpyDeviceToHost));
            } else {
                std::vector<uint64_t> h(u.rows * n_inputs);
                KITE_HIP(hipMemcpy(h.data(), m.ptr, h.size() * 8, hipMemcpyDeviceToHost));
                for (size_t i = 0; i < h.size(); i++)
                    out_matrix[i] = h[i] > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)h[i];
            }
        }
        return corr_device(ctx, m.ptr, wide ? 8 : 4, u.rows, n_inputs, methods, flags, pearson, spearman, cov,
                           pearson_fp64, spearman_fp64, stats);
    });
}

int kite_gpu_pca(kite_gpu_ctx *ctx, const double *cov, uint32_t n, uint32_t dims, double *scores,
                 double *eigenvalues) {
    if (!ctx || !cov) return set_error(KITE_E_INVALID, "invalid argument: NULL pointer");
    if (n < 1 || n > 1024) return set_error(KITE_E_INVALID, "invalid argument: PCA supports 1..1024 samples");
    return guarded([&]() -> int {
        bind(ctx);
        std::vector<double> c(cov, cov + (size_t)n * n), ev, V;
        ctx->corr->pca(c, (int)n, ev, V);
        std::vector<int> order(n);
        for (uint32_t i = 0; i < n; i++) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ev[a] > ev[b]; });
        if (eigenvalues)
            for (uint32_t d = 0; d < n; d++) eigenvalues[d] = ev[order[d]];
        if (scores)
            for (uint32_t d = 0; d < dims; d++) {
                const bool have = d < n;
                const double sc = have ? std::sqrt(std::max(0.0, ev[order[d]])) : 0.0;
                for (uint32_t i = 0; i < n; i++) scores[i * dims + d] = have ? V[i * n + order[d]] * sc : 0.0;
            }
        return KITE_OK;
    });
}

int kite_gpu_pca_noalloc(kite_gpu_ctx *ctx, const double *cov, uint32_t n, uint32_t dims, double *scores,
                         double *eigenvalues) {
    if (!ctx || !cov) return set_error(KITE_E_INVALID, "invalid argument: NULL pointer");
    if (n < 1 || n > PCA_SCRATCH_N)
        return set_error(KITE_E_INVALID, "invalid argument: kite_gpu_pca_noalloc supports 1.." +
                                             std::to_string(PCA_SCRATCH_N) + " samples (its fixed working memory)");
    return guarded([&]() -> int {
        bind(ctx);
        double *A = g_scratch_ptr[ctx->device];
        if (!A) return set_error(KITE_E_DEVICE, "PCA working memory not set up on this device");
        double *V = A + (size_t)PCA_SCRATCH_N * PCA_SCRATCH_N, *E = V + (size_t)PCA_SCRATCH_N * PCA_SCRATCH_N;
        // host copies on the stack: this path allocates nothing
        double ev[PCA_SCRATCH_N], vec[PCA_SCRATCH_N * PCA_SCRATCH_N];
        {
            std::lock_guard<std::mutex> g(g_scratch_use_mu[ctx->device]);
            KITE_HIP(hipMemcpyAsync(A, cov, (size_t)n * n * 8, hipMemcpyHostToDevice, ctx->stream));
            kite::jacobi_device(A, V, E, (int)n, ctx->stream);
            KITE_HIP(hipMemcpyAsync(ev, E, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
            KITE_HIP(hipMemcpyAsync(vec, V, (size_t)n * n * 8, hipMemcpyDeviceToHost, ctx->stream));
            KITE_HIP(hipStreamSynchronize(ctx->stream));
        }
        // eigenvalues descending, equal ones in index order (a stable sort)
        int order[PCA_SCRATCH_N];
        for (uint32_t i = 0; i < n; i++) {
            int x = (int)i, j = (int)i;
            while (j > 0 && ev[order[j - 1]] < ev[x]) {
                order[j] = order[j - 1];
                j--;
            }
            order[j] = x;
        }
        if (eigenvalues)
            for (uint32_t d = 0; d < n; d++) eigenvalues[d] = ev[order[d]];
        if (scores)
            for (uint32_t d = 0; d < dims; d++) {
                const bool have = d < n;
                const double sc = have ? std::sqrt(std::max(0.0, ev[order[d]])) : 0.0;
                for (uint32_t i = 0; i < n; i++) scores[i * dims + d] = have ? vec[i * n + order[d]] * sc : 0.0;
            }
        return KITE_OK;
    });
}

int kite_gpu_count_profile(kite_gpu_ctx *ctx, int enable) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    ctx->count->set_kernel_profile((enable & KITE_PROFILE_KERNELS) != 0);
    ctx->profile = (enable & KITE_PROFILE_KERNELS) != 0;
    ctx->timeline = (enable & KITE_PROFILE_TIMELINE) != 0;
    return KITE_OK;
}

int kite_gpu_batch_timeline(kite_gpu_ctx *ctx, double *out, uint32_t cap_samples) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    const uint32_t ns = (uint32_t)(ctx->tl_ms.size() / 6);
    if (out) std::copy(ctx->tl_ms.begin(), ctx->tl_ms.begin() + (size_t)std::min(ns, cap_samples) * 6, out);
    return (int)ns;
}

int kite_gpu_count_kernel_stats(kite_gpu_ctx *ctx, char *buf, size_t cap) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    std::string js = "{";
    bool first = true;
    for (const auto &ks : ctx->kstats_from_batch ? ctx->batch_kstats : ctx->count->kernel_stats()) {
        char tmp[256];
        snprintf(tmp, sizeof tmp, "%s\"%s\": {\"ms\": %.6f, \"launches\": %llu, \"bytes\": %llu}",
                 first ? "" : ", ", ks.name.c_str(), ks.ms, (unsigned long long)ks.launches,
                 (unsigned long long)ks.bytes);
        js += tmp;
        first = false;
    }
    js += "}";
    if (!buf || cap <= js.size()) return set_error(KITE_E_INVALID, "buffer too small");
    memcpy(buf, js.c_str(), js.size() + 1);
    return (int)js.size();
}

}  // extern "C"

// ---- batched count -----------------------------------------------------------
namespace {

void batch_stats(kite_gpu_ctx *c, const kite_batch_sample &smp, kite_count_stats &st) {
    const kite::StageTimes &t = c->count->times();
    st = kite_count_stats{};
    st.input_bytes = smp.len;
    st.kmers = c->count->n_kmers();
    st.unique = c->count->n_unique();
    st.buckets = c->count->buckets();
    st.batches = 1;
    st.split_levels = (uint32_t)c->count->levels();
    st.skipped_chunks = (smp.len > 0 && smp.fastq[0] != '@') ? 1u : 0u;
    st.words = (uint32_t)c->words;
    st.device_ms = t.total_ms;
    st.extract_ms = t.extract_ms;
    st.split_ms = t.split_ms;
    st.sort_ms = t.sort_ms;
    st.compact_ms = t.compact_ms;
}

// Adds the engine's per-kernel stats of the last run to the batch's sums.
void add_batch_kstats(kite_gpu_ctx *c) {
    if (!c->profile) return;
    for (const kite::KernelStat &ks : c->count->kernel_stats()) {
        auto it = std::find_if(c->batch_kstats.begin(), c->batch_kstats.end(),
                               [&](const kite::KernelStat &x) { return x.name == ks.name; });
        if (it == c->batch_kstats.end()) {
            c->batch_kstats.push_back(ks);
        } else {
            it->ms += ks.ms;
            it->launches += ks.launches;
            it->bytes += ks.bytes;
        }
    }
}

// Where a finished sample goes: raw rows (kite_batch_result_fn) or chunk
// payloads encoded on the device (kite_batch_chunks_fn).
struct BatchSink {
    kite_batch_result_fn raw = nullptr;
    kite_batch_chunks_fn chunks = nullptr;
    void *user = nullptr;
    uint32_t base = 0;  // the caller's index of the first sample passed in
    bool encoded() const { return chunks != nullptr; }
    int call(kite_gpu_ctx *c, BatchSlot &sl) const {
        if (chunks)
            return chunks(user, base + sl.sample, &sl.st, c->words == 4 ? KITE_CODEC_RAW : KITE_CODEC_DVLE,
                          sl.enc_bytes.as<uint8_t>(), sl.enc_boff.as<uint64_t>(), sl.part_off.data(),
                          sl.enc_minmax.as<uint64_t>());
        return raw(user, base + sl.sample, &sl.st, sl.keys.as<uint64_t>(), sl.counts.as<uint32_t>(),
                   sl.part_off.data());
    }
};

// Encodes the engine's result slot into `eb` and sizes the slot's host
// buffers; the payload stays on the device until the caller copies it.
kite::EncodeDevice encode_slot(kite_gpu_ctx *c, BatchSlot &sl, kite::EncodeBufs &eb) {
    const uint32_t P = c->partitions;
    uint64_t *hb = static_cast<uint64_t *>(sl.enc_boff.ensure(((size_t)P + 1) * 8));
    kite::EncodeDevice ed = kite::encode_on_device(c->count->keys(), c->count->counts(), c->count->d_part_off(),
                                                   c->count->n_unique(), P, c->words, c->stream, eb, hb);
    sl.enc_total = ed.total;
    sl.enc_bytes.ensure(std::max<uint64_t>(ed.total, 1));
    sl.enc_minmax.ensure((size_t)P * 64);
    return ed;
}

// One sample at a time, nothing overlapped.  reuse == false (KITE_BATCH_PLAIN):
// the device buffers are allocated and freed per sample, as separate count
// runs would; reuse == true (KITE_BATCH_SERIAL): persistent buffers.
void batch_plain(kite_gpu_ctx *c, const kite_batch_sample *smp, uint32_t n, const BatchSink &sink, bool reuse) {
    BatchSlot &sl = c->slots[0];
    const int W = c->words;
    if (reuse) {
        uint64_t maxlen = 0;
        for (uint32_t i = 0; i < n; i++) maxlen = std::max<uint64_t>(maxlen, smp[i].len);
        c->batch_in[0].ensure(maxlen);
    }
    for (uint32_t i = 0; i < n; i++) {
        void *d = nullptr;
        if (reuse) {
            d = c->batch_in[0].ptr;
            if (smp[i].len) KITE_HIP(hipMemcpy(d, smp[i].fastq, smp[i].len, hipMemcpyHostToDevice));
        } else {
            c->count->release();
            if (smp[i].len) {
                KITE_HIP(hipMalloc(&d, smp[i].len));
                KITE_HIP(hipMemcpy(d, smp[i].fastq, smp[i].len, hipMemcpyHostToDevice));
            }
        }
        c->count->set_slot(0);
        c->count->set_timing(true);
        try {
            apply_tuned(c, smp[i].len);
            c->count->run(static_cast<const uint8_t *>(d), smp[i].len, (int)c->k, c->canonical, c->partitions);
        } catch (...) {
            if (d && !reuse) (void)hipFree(d);
            throw;
        }
        add_batch_kstats(c);
        const uint64_t u = c->count->n_unique();
        if (sink.encoded()) {
            // the engine stream does not synchronize with the null stream:
            // the copies are queued behind the encode kernels on that stream
            const kite::EncodeDevice ed = encode_slot(c, sl, c->enc_slot[0]);
            if (ed.total)
                KITE_HIP(hipMemcpyAsync(sl.enc_bytes.ptr, ed.bytes, ed.total, hipMemcpyDeviceToHost, c->stream));
            KITE_HIP(hipMemcpyAsync(sl.enc_minmax.ptr, ed.minmax, (size_t)c->partitions * 64, hipMemcpyDeviceToHost,
                                    c->stream));
            KITE_HIP(hipStreamSynchronize(c->stream));
        } else {
            uint64_t *keys = static_cast<uint64_t *>(sl.keys.ensure(u * 8 * W));
            uint32_t *counts = static_cast<uint32_t *>(sl.counts.ensure(u * 4));
            if (u) {
                KITE_HIP(hipMemcpy(keys, c->count->keys(), u * 8 * W, hipMemcpyDeviceToHost));
                KITE_HIP(hipMemcpy(counts, c->count->counts(), u * 4, hipMemcpyDeviceToHost));
            }
        }
        if (d && !reuse) KITE_HIP(hipFree(d));
        batch_stats(c, smp[i], sl.st);
        sl.part_off = c->count->part_off();
        sl.sample = i;
        if (sink.call(c, sl) != 0) throw kite::Error("callback aborted the batch");
    }
}

// Hands finished result slots to the caller's callback on its own thread,
// in sample order, and frees each slot for reuse afterwards.
struct Consumer {
    kite_gpu_ctx *c;
    BatchSink sink;
    // set when the slots' D2H copies are queued through the context's
    // HostLink: their ends are its completion signals, not d2h_done
    const kite::HostLink *link = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> queue;
    bool free_slot[MAX_RING];
    bool done = false;
    std::string error;
    std::thread th;

    Consumer(kite_gpu_ctx *cc, const BatchSink &sk) : c(cc), sink(sk) {
        for (bool &f : free_slot) f = true;
        th = std::thread([this] { loop(); });
    }
    ~Consumer() { stop(); }
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu);
            done = true;
        }
        cv.notify_all();
        if (th.joinable()) th.join();
    }
    void loop() {
        for (;;) {
            int s;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return done || !queue.empty(); });
                if (queue.empty()) return;
                s = queue.front();
                queue.pop_front();
            }
            BatchSlot &sl = c->slots[s];
            std::string err;
            if (link) {
                try {
                    link->wait(s);
                } catch (const std::exception &x) {
                    err = x.what();
                }
            } else {
                hipError_t e = hipEventSynchronize(sl.d2h_done);
                if (e != hipSuccess) err = std::string("HIP error: ") + hipGetErrorString(e);
            }
            if (err.empty() && sink.call(c, sl) != 0) err = "callback aborted the batch";
            {
                std::lock_guard<std::mutex> g(mu);
                if (!err.empty() && error.empty()) error = err;
                free_slot[s] = true;
            }
            cv.notify_all();
        }
    }
    // Frees slot s without a callback (a sample whose result is dropped).
    void release(int s) {
        {
            std::lock_guard<std::mutex> g(mu);
            free_slot[s] = true;
        }
        cv.notify_all();
    }
    // Blocks until slot s is free; returns false once an error was recorded.
    bool acquire(int s) {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] { return free_slot[s] || !error.empty(); });
        if (!error.empty()) return false;
        free_slot[s] = false;
        return true;
    }
    void push(int s) {
        {
            std::lock_guard<std::mutex> g(mu);
            queue.push_back(s);
        }
        cv.notify_all();
    }
    // Waits for every queued slot to be consumed.
    void drain() {
        std::unique_lock<std::mutex> g(mu);
        cv.wait(g, [&] {
            if (!error.empty()) return true;
            if (!queue.empty()) return false;
            for (bool f : free_slot)
                if (!f) return false;
            return true;
        });
    }
    std::string err() {
        std::lock_guard<std::mutex> g(mu);
        return error;
    }
};

// The H2D side of a ring batch, on a thread of its own: the input of sample
// j goes into buffer j % R on the copy stream h2d once the run of sample
// j - R, which read that buffer, is queued (ran()), behind a stream wait for
// that run's end (slot j % R's run_done), and once the H2D of sample j - 1
// has ended; in_ready of the slot follows the copy.  One input copy is in
// flight at a time: with a single copy queued, the runtime keeps it on its
// preferred host-to-device engine instead of spreading queued copies over
// engines that also carry the D2H or move host traffic slower.
class Uploader {
   public:
    Uploader(kite_gpu_ctx *c, const kite_batch_sample *smp, uint32_t n, int R, bool prof)
        : c_(c), smp_(smp), n_(n), R_((uint32_t)R), prof_(prof) {}
    ~Uploader() { stop(); }
    void start() { th_ = std::thread([this] { loop(); }); }
    // the run of sample i is queued and its slot's run_done recorded
    void ran(uint32_t i) {
        {
            std::lock_guard<std::mutex> g(mu_);
            ran_ = i + 1;
        }
        cv_.notify_all();
    }
    // blocks until the H2D of sample i is queued; throws the uploader's error
    void wait_queued(uint32_t i) {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return queued_ > i || !err_.empty(); });
        if (!err_.empty()) throw kite::Error(err_);
    }
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }

   private:
    void loop() {
        try {
            for (uint32_t j = 0; j < n_; j++) {
                const uint32_t b = j % R_;
                {
                    std::unique_lock<std::mutex> g(mu_);
                    cv_.wait(g, [&] { return stop_ || j < R_ || ran_ > j - R_; });
                    if (stop_) return;
                }
                if (j > 0) KITE_HIP(hipEventSynchronize(c_->slots[(j - 1) % R_].in_ready));
                if (j >= R_) KITE_HIP(hipStreamWaitEvent(c_->h2d, c_->slots[b].run_done, 0));
                if (prof_) KITE_HIP(hipEventRecord(c_->tl_events[j][0], c_->h2d));
                if (smp_[j].len)
                    KITE_HIP(hipMemcpyAsync(c_->batch_in[b].ptr, smp_[j].fastq, smp_[j].len, hipMemcpyHostToDevice,
                                            c_->h2d));
                if (prof_) KITE_HIP(hipEventRecord(c_->tl_events[j][1], c_->h2d));
                KITE_HIP(hipEventRecord(c_->slots[b].in_ready, c_->h2d));
                {
                    std::lock_guard<std::mutex> g(mu_);
                    queued_ = j + 1;
                }
                cv_.notify_all();
            }
        } catch (const std::exception &e) {
            {
                std::lock_guard<std::mutex> g(mu_);
                if (err_.empty()) err_ = e.what();
            }
            cv_.notify_all();
        }
    }
    kite_gpu_ctx *c_;
    const kite_batch_sample *smp_;
    uint32_t n_, R_;
    bool prof_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_;
    uint32_t queued_ = 0, ran_ = 0;
    bool stop_ = false;
    std::string err_;
};

// Samples flow through a ring of R input buffers and R result slots (input
// buffer, engine output slot, encode buffers and host buffers of slot i % R
// belong to sample i), with the kernels launched one by one after the
// count's host read-backs:
//   uploader thread:   the H2D of sample j into buffer j % R (Uploader);
//   engine stream:     the count of sample i once its H2D is done (raw rows:
//                      also once the D2H of sample i - R has read the
//                      engine's output slot), then the chunk encode (once
//                      the D2H of sample i - R has read the encode buffers);
//   D2H of sample i after its run: on the context's HostLink from a drainer
//                      thread (the engine thread then waits on the host for
//                      the end of the D2H of sample i - R), or on the copy
//                      stream d2h when stage events are on, when
//                      KITE_D2H_ENGINES=0 or when no engine qualifies;
//   consumer thread:   the callback of sample i after its D2H, which frees
//                      the slot's host buffers for sample i + R.
void batch_pipelined(kite_gpu_ctx *c, const kite_batch_sample *smp, uint32_t n, const BatchSink &sink) {
    const int R = batch_ring();
    if (!c->h2d) KITE_HIP(hipStreamCreateWithFlags(&c->h2d, hipStreamNonBlocking));
    if (!c->d2h) KITE_HIP(hipStreamCreateWithFlags(&c->d2h, hipStreamNonBlocking));
    for (int b = 0; b < R; b++) {
        BatchSlot &sl = c->slots[b];
        for (hipEvent_t *e : {&sl.d2h_done, &sl.run_done, &sl.in_ready})
            if (!*e) KITE_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    const bool prof = c->timeline;
    if (prof) {
        if (!c->tl_start) KITE_HIP(hipEventCreate(&c->tl_start));
        while (c->tl_events.size() < n) {
            std::array<hipEvent_t, 6> ev{};
            for (hipEvent_t &e : ev) KITE_HIP(hipEventCreate(&e));
            c->tl_events.push_back(ev);
        }
    }
    uint64_t maxlen = 0;
    for (uint32_t i = 0; i < n; i++) maxlen = std::max<uint64_t>(maxlen, smp[i].len);
    // every input buffer is sized before any copy is queued into one, and
    // the count's key arrays for the longest input (a FASTQ record's
    // sequence line is at most half its bytes: at most len / 2 k-mers)
    for (int b = 0; b < std::min<int>(R, (int)n); b++) c->batch_in[b].ensure(maxlen);
    const int W = c->words;
    c->count->reserve(maxlen / 2, W);
    kite::HostLink *link = prof ? nullptr : host_link(c);

    Consumer cons(c, sink);
    cons.link = link;
    // drainer (HostLink copies): after sample i's run, the D2H of its results
    struct Pending {
        uint32_t i;
        int s;
        kite::CopyPiece piece[2];
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Pending> todo;
    uint32_t drained = 0;  // samples whose D2H is queued (or failed to queue)
    bool stop = false;
    std::string drain_err;
    if (prof) KITE_HIP(hipEventRecord(c->tl_start, c->h2d));
    Uploader up(c, smp, n, R, prof);
    up.start();
    auto drain = [&] {
        for (;;) {
            Pending t;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return stop || !todo.empty(); });
                if (todo.empty()) return;
                t = todo.front();
                todo.pop_front();
            }
            bool ok = false;
            try {
                KITE_HIP(hipEventSynchronize(c->slots[t.s].run_done));
                c->slots[t.s].st.d2h_engine = link->d2h_many(t.s, t.piece, 2);
                ok = true;
            } catch (const std::exception &e) {
                std::lock_guard<std::mutex> g(mu);
                if (drain_err.empty()) drain_err = e.what();
            }
            if (ok) cons.push(t.s);
            else cons.release(t.s);
            {
                std::lock_guard<std::mutex> g(mu);
                drained = t.i + 1;
            }
            cv.notify_all();
        }
    };
    std::thread drainer;
    if (link) drainer = std::thread(drain);
    auto finish_drainer = [&] {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv.notify_all();
        if (drainer.joinable()) drainer.join();
    };
    // before slot s is written again: the D2H of sample i - R has read it
    auto slot_read = [&](uint32_t i, int s) {
        if (i < (uint32_t)R) return;
        if (link) {
            std::unique_lock<std::mutex> g(mu);
            cv.wait(g, [&] { return drained > i - R; });
            g.unlock();
            link->wait(s);
        } else {
            KITE_HIP(hipStreamWaitEvent(c->stream, c->slots[s].d2h_done, 0));
        }
    };

    std::string failure;
    try {
        for (uint32_t i = 0; i < n; i++) {
            {
                std::lock_guard<std::mutex> g(mu);
                if (!drain_err.empty()) break;
            }
            const int s = (int)(i % (uint32_t)R);
            BatchSlot &sl = c->slots[s];
            up.wait_queued(i);
            KITE_HIP(hipStreamWaitEvent(c->stream, sl.in_ready, 0));
            // raw rows: the D2H of sample i - R reads the engine's output slot s
            if (!sink.encoded()) slot_read(i, s);
            if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][2], c->stream));
            c->count->set_slot(s);
            c->count->set_timing(true);
            apply_tuned(c, smp[i].len);
            c->count->run(c->batch_in[s].as<uint8_t>(), smp[i].len, (int)c->k, c->canonical, c->partitions);
            add_batch_kstats(c);
            // slot s's host buffers are free once the callback of sample i - R
            // ran, which was after the D2H of that sample had ended
            if (!cons.acquire(s)) break;
            const uint64_t u = c->count->n_unique();
            batch_stats(c, smp[i], sl.st);
            sl.part_off = c->count->part_off();
            sl.sample = i;
            Pending t{i, s, {}};
            if (sink.encoded()) {
                // encode buffer s was last read by the D2H of sample i - R
                slot_read(i, s);
                const kite::EncodeDevice ed = encode_slot(c, sl, c->enc_slot[s]);
                KITE_HIP(hipEventRecord(sl.run_done, c->stream));
                up.ran(i);
                if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][3], c->stream));
                t.piece[0] = kite::CopyPiece{sl.enc_bytes.ptr, ed.bytes, ed.total};
                t.piece[1] = kite::CopyPiece{sl.enc_minmax.ptr, ed.minmax, (size_t)c->partitions * 64};
            } else {
                KITE_HIP(hipEventRecord(sl.run_done, c->stream));
                up.ran(i);
                if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][3], c->stream));
                uint64_t *keys = static_cast<uint64_t *>(sl.keys.ensure(u * 8 * W));
                uint32_t *counts = static_cast<uint32_t *>(sl.counts.ensure(u * 4));
                t.piece[0] = kite::CopyPiece{keys, c->count->keys(), u * 8 * W};
                t.piece[1] = kite::CopyPiece{counts, c->count->counts(), u * 4};
            }
            if (link) {
                {
                    std::lock_guard<std::mutex> g(mu);
                    todo.push_back(t);
                }
                cv.notify_all();
            } else {
                KITE_HIP(hipStreamWaitEvent(c->d2h, sl.run_done, 0));
                if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][4], c->d2h));
                for (const kite::CopyPiece &q : t.piece)
                    if (q.bytes) KITE_HIP(hipMemcpyAsync(q.dst, q.src, q.bytes, hipMemcpyDeviceToHost, c->d2h));
                if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][5], c->d2h));
                KITE_HIP(hipEventRecord(sl.d2h_done, c->d2h));
                cons.push(s);
            }
        }
        finish_drainer();
        if (drain_err.empty()) cons.drain();
    } catch (const std::exception &e) {
        failure = e.what();
    }
    finish_drainer();
    up.stop();
    cons.stop();
    if (link) link->quiesce();
    (void)hipStreamSynchronize(c->h2d);
    (void)hipStreamSynchronize(c->d2h);
    (void)hipStreamSynchronize(c->stream);
    if (!failure.empty()) throw kite::Error(failure);
    if (!drain_err.empty()) throw kite::Error(drain_err);
    std::string ce = cons.err();
    if (!ce.empty()) throw kite::Error(ce);
    if (prof) {
        c->tl_ms.assign((size_t)n * 6, 0.0);
        for (uint32_t i = 0; i < n; i++)
            for (int q = 0; q < 6; q++) {
                float ms = 0;
                KITE_HIP(hipEventElapsedTime(&ms, c->tl_start, c->tl_events[i][q]));
                c->tl_ms[(size_t)i * 6 + q] = ms;
            }
    }
}

// Graph-captured batches unless KITE_BATCH_GRAPH=0 (the classic pipeline,
// with its host read-backs between launches, for comparison).
bool use_graphs() {
    const char *e = getenv("KITE_BATCH_GRAPH");
    return !(e && atoi(e) == 0);
}

bool same_tunables(const kite::CountTunables &a, const kite::CountTunables &b) {
    return a.split_target_div == b.split_target_div && a.max_digit_bits == b.max_digit_bits && a.dedup == b.dedup &&
           a.hash_classes == b.hash_classes && a.split_chunk == b.split_chunk && a.extract_subs == b.extract_subs &&
           a.extract_staged == b.extract_staged && a.split_staged == b.split_staged &&
           a.split_above_div == b.split_above_div;
}

void drop_graphs(kite_gpu_ctx *c) {
    for (auto &g : c->graphs)
        for (hipGraphExec_t e : {g.second.count, g.second.encode})
            if (e) KITE_HIP(hipGraphExecDestroy(e));
    c->graphs.clear();
}

// Captures what `enqueue` puts on the engine stream as an executable graph.
template <class F>
hipGraphExec_t capture(kite_gpu_ctx *c, F &&enqueue) {
    hipGraph_t graph = nullptr;
    KITE_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        (void)hipStreamEndCapture(c->stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    KITE_HIP(hipStreamEndCapture(c->stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    KITE_HIP(e);
    return exec;
}

// Bytes of chunk payload a ring slot holds for a planned count of up to n
// k-mers: beyond it the sample takes the classic path.
uint64_t graph_bytes_cap(uint64_t n, int W) { return n * (W == 1 ? 10 : (W == 2 ? 20 : 36)); }

// The pipelined batch of encoded chunks with every sample's device work --
// the count, then the chunk encode, each sized on the device -- as two
// launches of captured HIP graphs (a pair per input size class and ring
// slot, captured on first use and kept while the shape, buffers and
// settings hold):
//   uploader thread:  the H2D of sample j into buffer j % R on the copy
//                     stream h2d, once the count graph of sample j - R
//                     (which read that buffer) is queued, behind a stream
//                     wait for its end, and once the H2D of sample j - 1 has
//                     ended: one input copy in flight at a time;
//   engine stream:    the input length as a kernel argument into the
//                     slot's length word, and the count graph of sample i
//                     once its H2D ended; then, once the D2H of sample
//                     i - R has read the slot's payload buffer, the encode
//                     graph, which writes the chunk payload into that buffer
//                     and the sizes and partition tables into the slot's
//                     page-locked summary;
//   drainer thread:   after sample i's encode graph, the D2H of exactly its
//                     payload bytes, on an SDMA engine the context's
//                     HostLink picks (on the copy stream d2h, engine left to
//                     the runtime, when stage events are on, when
//                     KITE_D2H_ENGINES=0 or when no engine qualifies);
//   consumer thread:  the callback after that D2H, which frees the slot.
// The copy engines: the runtime puts a lone host-to-device copy on its
// preferred engine, but spreads copies queued together over more engines,
// some of which carry host traffic at a fraction of the rate or share the
// engine of the other direction; one H2D in flight, and the D2H on an
// engine of its own, keep both directions of the link busy at once.
// The engine thread reads nothing back between samples.  Before the encode
// graph of sample i it waits, on the host, for the drainer to have queued
// the D2H of sample i - R, then for that copy's end: on the host for a
// HostLink copy, with a stream wait on its event otherwise.  A sample whose
// data exceeds a capacity of the planned count (its summary says so) stops
// the graphs there: the samples from it on run through the classic
// pipeline.  Returns the first sample not delivered (n when all were).
uint32_t batch_graph(kite_gpu_ctx *c, const kite_batch_sample *smp, uint32_t n, const BatchSink &sink) {
    const int R = batch_ring();
    if (!c->h2d) KITE_HIP(hipStreamCreateWithFlags(&c->h2d, hipStreamNonBlocking));
    if (!c->d2h) KITE_HIP(hipStreamCreateWithFlags(&c->d2h, hipStreamNonBlocking));
    for (int b = 0; b < R; b++) {
        BatchSlot &sl = c->slots[b];
        for (hipEvent_t *e : {&sl.d2h_done, &sl.run_done, &sl.in_ready, &sl.enc_done})
            if (!*e) KITE_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    const bool prof = c->timeline;
    // payload D2H on engines picked here (stage events need the d2h stream)
    kite::HostLink *link = prof ? nullptr : host_link(c);
    if (prof) {
        if (!c->tl_start) KITE_HIP(hipEventCreate(&c->tl_start));
        while (c->tl_events.size() < n) {
            std::array<hipEvent_t, 6> ev{};
            for (hipEvent_t &e : ev) KITE_HIP(hipEventCreate(&e));
            c->tl_events.push_back(ev);
        }
    }
    uint64_t maxlen = 0;
    for (uint32_t i = 0; i < n; i++) maxlen = std::max<uint64_t>(maxlen, smp[i].len);
    const uint32_t P = c->partitions;
    const int W = c->words;
    // the shape the graphs hold: a larger input or another k, P re-prepares
    // the buffers (and so drops every graph)
    const bool same_kp = std::get<0>(c->graph_shape) == c->k && std::get<1>(c->graph_shape) == c->canonical &&
                         std::get<2>(c->graph_shape) == P;
    const uint64_t cap = same_kp ? std::max<uint64_t>(maxlen, std::get<3>(c->graph_shape)) : maxlen;
    for (int b = 0; b < std::min<int>(R, (int)n); b++) c->batch_in[b].ensure(cap);
    const auto shape = std::make_tuple(c->k, c->canonical, P, cap);
    const bool reshape = shape != c->graph_shape;
    if (reshape) {
        KITE_HIP(hipStreamSynchronize(c->stream));
        drop_graphs(c);
        c->count->dev_prepare(cap, (int)c->k, P, c->graph_enc);
    }
    const uint64_t bcap = graph_bytes_cap(c->count->dev_caps().n, W);
    for (int b = 0; b < R; b++) {
        c->graph_bytes[b].ensure(bcap);
        c->graph_meta[b].ensure(64 + kite::CountEngine::dev_summary_bytes(P));
    }
    uint64_t *dlen = static_cast<uint64_t *>(c->graph_dlen.ensure(MAX_RING * 8));
    auto len_word = [&](int b) { return c->graph_meta[b].as<uint64_t>(); };
    auto summary = [&](int b) { return reinterpret_cast<kite::DevSummary *>(c->graph_meta[b].as<uint8_t>() + 64); };
    if (reshape) {
        // one planned count of an empty input, outside any capture: it sets
        // the kernels' attributes and sizes the grids the captures reuse
        *len_word(0) = 0;
        c->count->dev_enqueue_count(c->batch_in[0].as<uint8_t>(), len_word(0), cap, (int)c->k, c->canonical, P);
        c->count->dev_enqueue_encode(c->graph_enc, c->graph_bytes[0].as<uint8_t>(), bcap, summary(0));
        KITE_HIP(hipStreamSynchronize(c->stream));
        c->graph_shape = shape;
    }
    // the graphs of (size class, slot), captured when missing or stale; the
    // count's grids are sized for the longest input of the class (so a small
    // sample does not pay for the largest one's grids)
    std::map<int, uint64_t> class_max;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t &m = class_max[size_class(smp[i].len)];
        m = std::max<uint64_t>(m, smp[i].len);
    }
    auto graphs_for = [&](uint64_t len, int b) -> kite_gpu_ctx::BatchGraph & {
        apply_tuned(c, len);
        const kite::CountTunables t = c->count->resolved_tunables();
        const int cls = size_class(len);
        kite_gpu_ctx::BatchGraph &g = c->graphs[std::make_pair(cls, b)];
        // a grid for the class's longest input this batch (kept when it covers it)
        if (g.count && g.encode && same_tunables(g.tun, t) && g.in == c->batch_in[b].ptr &&
            g.out == c->graph_bytes[b].ptr && g.meta == c->graph_meta[b].ptr && g.len == dlen + b &&
            g.max_len >= class_max[cls] && g.max_len <= cap)
            return g;
        const uint64_t grid_len = class_max[cls];
        for (hipGraphExec_t *e : {&g.count, &g.encode}) {
            if (*e) KITE_HIP(hipGraphExecDestroy(*e));
            *e = nullptr;
        }
        g.count = capture(c, [&] {
            c->count->dev_enqueue_count(c->batch_in[b].as<uint8_t>(), dlen + b, grid_len, (int)c->k, c->canonical, P);
        });
        g.encode = capture(c, [&] {
            c->count->dev_enqueue_encode(c->graph_enc, c->graph_bytes[b].as<uint8_t>(), bcap, summary(b));
        });
        g.tun = t;
        g.in = c->batch_in[b].ptr;
        g.out = c->graph_bytes[b].ptr;
        g.meta = c->graph_meta[b].ptr;
        g.len = dlen + b;
        g.max_len = grid_len;
        return g;
    };
    // every graph this batch needs, before any copy is queued; with kernel
    // profiling on, the same launches are made one by one instead, each
    // between two HIP events (a captured graph holds no timing)
    const bool eager = c->profile;
    std::vector<std::pair<hipGraphExec_t, hipGraphExec_t>> exec(n);
    if (eager) {
        c->count->begin_kernel_stats();
    } else {
        for (uint32_t i = 0; i < n; i++) {
            const kite_gpu_ctx::BatchGraph &g = graphs_for(smp[i].len, (int)(i % (uint32_t)R));
            exec[i] = std::make_pair(g.count, g.encode);
        }
    }
    std::vector<kite::DevSummary> sums(n);  // each sample's summary (the profiled bytes)

    // drainer: after sample i's encode graph, the D2H of its payload
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint32_t> launched;
    uint32_t drained = 0;  // samples whose D2H is queued (or that were dropped)
    bool stop = false;
    uint32_t overflow_at = n;  // the first sample whose planned count overflowed
    std::string drain_err;
    Consumer cons(c, sink);
    cons.link = link;
    auto mark_drained = [&](uint32_t i) {
        {
            std::lock_guard<std::mutex> g(mu);
            drained = i + 1;
        }
        cv.notify_all();
    };
    if (prof) KITE_HIP(hipEventRecord(c->tl_start, c->h2d));
    Uploader up(c, smp, n, R, prof);
    up.start();
    std::thread drainer([&] {
        for (;;) {
            uint32_t i;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return stop || !launched.empty(); });
                if (launched.empty()) return;
                i = launched.front();
                launched.pop_front();
            }
            const int s = (int)(i % (uint32_t)R);
            BatchSlot &sl = c->slots[s];
            bool taken = false;
            try {
                KITE_HIP(hipEventSynchronize(sl.enc_done));
                const kite::DevSummary sm = *summary(s);
                sums[i] = sm;
                bool skip;
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (sm.overflow && overflow_at == n) overflow_at = i;
                    skip = i >= overflow_at;
                }
                if (skip || !cons.acquire(s)) {
                    mark_drained(i);
                    continue;
                }
                taken = true;
                const uint64_t *tab = reinterpret_cast<const uint64_t *>(summary(s) + 1);
                sl.sample = i;
                sl.st = kite_count_stats{};
                sl.st.input_bytes = smp[i].len;
                sl.st.kmers = sm.n_kmers;
                sl.st.unique = sm.n_unique;
                sl.st.buckets = sm.buckets;
                sl.st.batches = 1;
                sl.st.split_levels = sm.levels;
                sl.st.skipped_chunks = (smp[i].len > 0 && smp[i].fastq[0] != '@') ? 1u : 0u;
                sl.st.words = (uint32_t)W;
                sl.st.graph = eager ? 0u : 1u;
                sl.part_off.assign(tab, tab + P + 1);
                uint64_t *boff = static_cast<uint64_t *>(sl.enc_boff.ensure(((size_t)P + 1) * 8));
                std::copy(tab + P + 1, tab + 2 * ((size_t)P + 1), boff);
                uint64_t *mm = static_cast<uint64_t *>(sl.enc_minmax.ensure((size_t)P * 64));
                std::copy(tab + 2 * ((size_t)P + 1), tab + 2 * ((size_t)P + 1) + 8 * (size_t)P, mm);
                sl.enc_total = sm.bytes;
                uint8_t *hb = static_cast<uint8_t *>(sl.enc_bytes.ensure(std::max<uint64_t>(sm.bytes, 1)));
                if (link) {
                    // the encode graph's writes are visible: enc_done was
                    // waited for above
                    if (sm.bytes) sl.st.d2h_engine = link->d2h(s, hb, c->graph_bytes[s].ptr, sm.bytes);
                } else {
                    if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][4], c->d2h));
                    if (sm.bytes)
                        KITE_HIP(hipMemcpyAsync(hb, c->graph_bytes[s].ptr, sm.bytes, hipMemcpyDeviceToHost, c->d2h));
                    if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][5], c->d2h));
                    KITE_HIP(hipEventRecord(sl.d2h_done, c->d2h));
                }
                mark_drained(i);
                taken = false;
                cons.push(s);
            } catch (const std::exception &e) {
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (drain_err.empty()) drain_err = e.what();
                    if (overflow_at == n) overflow_at = i;
                }
                if (taken) cons.release(s);
                mark_drained(i);
            }
        }
    });
    std::string failure;
    uint32_t launched_n = 0;
    try {
        for (uint32_t i = 0; i < n; i++) {
            {
                std::lock_guard<std::mutex> g(mu);
                if (i >= overflow_at || !cons.err().empty()) break;
            }
            up.wait_queued(i);
            const int s = (int)(i % (uint32_t)R);
            BatchSlot &sl = c->slots[s];
            KITE_HIP(hipStreamWaitEvent(c->stream, sl.in_ready, 0));
            // the length word of buffer s: the count graph of sample i - R,
            // earlier on this stream, has read it
            c->count->dev_set_len(dlen + s, smp[i].len);
            if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][2], c->stream));
            if (eager) {
                apply_tuned(c, smp[i].len);
                c->count->dev_enqueue_count(c->batch_in[s].as<uint8_t>(), dlen + s,
                                            class_max[size_class(smp[i].len)], (int)c->k, c->canonical, P);
            } else {
                KITE_HIP(hipGraphLaunch(exec[i].first, c->stream));
            }
            KITE_HIP(hipEventRecord(sl.run_done, c->stream));  // input buffer s is free again
            up.ran(i);
            if (i >= (uint32_t)R) {
                // the slot's payload buffer is read by the D2H of sample i - R:
                // once the drainer has queued that copy, wait for its end
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return drained > i - R; });
                if (i >= overflow_at) break;
                g.unlock();
                if (link) link->wait(s);
                else KITE_HIP(hipStreamWaitEvent(c->stream, sl.d2h_done, 0));
            }
            if (eager) c->count->dev_enqueue_encode(c->graph_enc, c->graph_bytes[s].as<uint8_t>(), bcap, summary(s));
            else KITE_HIP(hipGraphLaunch(exec[i].second, c->stream));
            if (prof) KITE_HIP(hipEventRecord(c->tl_events[i][3], c->stream));
            KITE_HIP(hipEventRecord(sl.enc_done, c->stream));
            launched_n = i + 1;
            {
                std::lock_guard<std::mutex> g(mu);
                launched.push_back(i);
            }
            cv.notify_all();
        }
    } catch (const std::exception &e) {
        failure = e.what();
    }
    {
        std::lock_guard<std::mutex> g(mu);
        stop = true;
    }
    cv.notify_all();
    drainer.join();
    up.stop();
    // (after a failure a slot may stay taken: nothing more is drained)
    if (failure.empty()) cons.drain();
    cons.stop();
    if (link) link->quiesce();
    (void)hipStreamSynchronize(c->h2d);
    (void)hipStreamSynchronize(c->d2h);
    (void)hipStreamSynchronize(c->stream);
    if (!failure.empty()) throw kite::Error(failure);
    if (!drain_err.empty()) throw kite::Error(drain_err);
    std::string ce = cons.err();
    if (!ce.empty()) throw kite::Error(ce);
    if (eager) {
        // per-kernel device time of the planned launches, with each kernel's
        // algorithmic bytes from the samples' summaries
        c->count->collect_kernel_stats();
        const uint64_t kb = 8ull * W;
        for (kite::KernelStat &ks : c->count->kernel_stats_mut()) {
            ks.bytes = 0;
            for (uint32_t i = 0; i < std::min(launched_n, overflow_at); i++) {
                const kite::DevSummary &m = sums[i];
                const uint64_t mat = (uint64_t)P * m.nwg;
                if (ks.name == "k_nl_count") ks.bytes += m.len;
                else if (ks.name == "k_extract_hist") ks.bytes += m.len + mat * 4;
                else if (ks.name == "k_extract_scatter") ks.bytes += m.len + mat * 8 + m.n_kmers * kb;
                else if (ks.name == "k_split_hist") ks.bytes += m.split_keys * kb;
                else if (ks.name == "k_split_scatter") ks.bytes += m.split_keys * 2 * kb;
                else if (ks.name == "k_bucket_sort") ks.bytes += m.n_kmers * kb + m.n_unique * (kb + 4);
                else if (ks.name == "k_compact") ks.bytes += 2 * m.n_unique * (kb + 4);
            }
        }
        c->batch_kstats = c->count->kernel_stats();
    }
    if (prof) {
        c->tl_ms.assign((size_t)n * 6, 0.0);
        for (uint32_t i = 0; i < std::min(launched_n, overflow_at); i++)
            for (int q = 0; q < 6; q++) {
                float ms = 0;
                KITE_HIP(hipEventElapsedTime(&ms, c->tl_start, c->tl_events[i][q]));
                c->tl_ms[(size_t)i * 6 + q] = ms;
            }
    }
    return std::min(overflow_at, n);
}

}  // namespace

extern "C" {

namespace {
int count_batch_impl(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions,
                     const kite_batch_sample *samples, uint32_t n_samples, int mode, const BatchSink &sink) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    if (!sink.raw && !sink.chunks) return set_error(KITE_E_INVALID, "invalid argument: callback is NULL");
    if (n_samples && !samples) return set_error(KITE_E_INVALID, "invalid argument: samples is NULL");
    if (mode != KITE_BATCH_PLAIN && mode != KITE_BATCH_PIPELINED && mode != KITE_BATCH_SERIAL)
        return set_error(KITE_E_INVALID, "invalid argument: unknown batch mode");
    for (uint32_t i = 0; i < n_samples; i++)
        if (samples[i].len && !samples[i].fastq)
            return set_error(KITE_E_INVALID, "invalid argument: sample " + std::to_string(i) + " has no data");
    int rc = kite_gpu_count_begin(ctx, k, canonical, partitions);
    if (rc) return rc;
    ctx->session = false;  // the batch does not leave a count session behind
    rc = guarded([&]() -> int {
        bind(ctx);
        ctx->batch_kstats.clear();
        ctx->kstats_from_batch = true;
        ctx->tl_ms.clear();
        for (uint32_t i = 0; i < n_samples; i++)
            if (samples[i].len > ctx->batch_cap)
                return set_error(KITE_E_NOMEM, "sample " + std::to_string(i) +
                                                   " exceeds the device budget of one batch");
        if (mode == KITE_BATCH_PIPELINED && sink.encoded() && use_graphs()) {
            // captured graphs; samples from the first one whose data exceeds
            // the planned count's capacities on go through the classic pipeline
            const uint32_t done = batch_graph(ctx, samples, n_samples, sink);
            if (done < n_samples) {
                ctx->graph_fallbacks += n_samples - done;
                BatchSink rest = sink;
                rest.base = done;
                batch_pipelined(ctx, samples + done, n_samples - done, rest);
            }
        } else if (mode == KITE_BATCH_PIPELINED) {
            batch_pipelined(ctx, samples, n_samples, sink);
        } else {
            batch_plain(ctx, samples, n_samples, sink, mode == KITE_BATCH_SERIAL);
        }
        return KITE_OK;
    });
    if (rc == KITE_E_DEVICE && g_last_error.find("callback aborted") != std::string::npos) rc = KITE_E_CALLBACK;
    return rc;
}
}  // namespace

int kite_gpu_count_batch(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions,
                         const kite_batch_sample *samples, uint32_t n_samples, int mode,
                         kite_batch_result_fn fn, void *user) {
    BatchSink sink;
    sink.raw = fn;
    sink.user = user;
    return count_batch_impl(ctx, k, canonical, partitions, samples, n_samples, mode, sink);
}

int kite_gpu_count_batch_chunks(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions,
                                const kite_batch_sample *samples, uint32_t n_samples, int mode,
                                kite_batch_chunks_fn fn, void *user) {
    BatchSink sink;
    sink.chunks = fn;
    sink.user = user;
    return count_batch_impl(ctx, k, canonical, partitions, samples, n_samples, mode, sink);
}

int kite_gpu_autotune(kite_gpu_ctx *ctx, uint32_t k, int canonical, uint32_t partitions, const uint8_t *fastq,
                      size_t len, char *report, size_t cap) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    if (k < 1 || k > 128) return set_error(KITE_E_INVALID, "invalid k-mer length: k must be between 1 and 128");
    if (partitions == 0 || (partitions & (partitions - 1)))
        return set_error(KITE_E_INVALID, "invalid partition count: must be a power of 2");
    if (!fastq || len == 0) return set_error(KITE_E_INVALID, "invalid argument: empty sample");
    return guarded([&]() -> int {
        bind(ctx);
        // a count session's device result is kept as encoded runs first
        if (ctx->session) stash_device_result(ctx);
        const int W = k <= 31 ? 1 : (k <= 64 ? 2 : 4);
        kite::DevBuf d;
        d.ensure(len);
        KITE_HIP(hipMemcpy(d.ptr, fastq, len, hipMemcpyHostToDevice));
        // the defaults, spelled out (the engine resolves -1 by key width)
        kite::CountTunables best;
        best.split_target_div = W == 1 ? 2 : 4;
        best.hash_classes = W == 1 ? 4 : 3;
        best.max_digit_bits = 12;
        best.extract_subs = 4;
        uint64_t ref_unique = 0, ref_kmers = 0;
        int trials = 0;
        std::string log;
        auto time_of = [&](const kite::CountTunables &t) -> double {
            ctx->count->set_base_tunables(t);
            ctx->count->set_timing(true);
            double m = 1e300;
            for (int rep = 0; rep < 3; rep++) {
                ctx->count->run(d.as<uint8_t>(), len, (int)k, canonical ? 1 : 0, partitions);
                if (!trials && !rep) {
                    ref_unique = ctx->count->n_unique();
                    ref_kmers = ctx->count->n_kmers();
                }
                if (ctx->count->n_unique() != ref_unique || ctx->count->n_kmers() != ref_kmers)
                    throw kite::Error("autotune: a setting changed the count result");
                m = std::min<double>(m, ctx->count->times().total_ms);
            }
            trials++;
            char buf[160];
            snprintf(buf, sizeof buf, "%s[%d, %d, %d, %d, %d, %.4f]", log.empty() ? "" : ", ", t.split_target_div,
                     t.max_digit_bits, t.hash_classes, t.extract_subs, t.split_above_div, m);
            log += buf;
            return m;
        };
        // coordinate descent: one knob at a time over its candidates, the
        // others held; a change must win by 1% to count as one
        const double default_ms = time_of(best);
        double best_ms = default_ms;
        for (int pass = 0; pass < 2; pass++) {
            bool changed = false;
            auto try_knob = [&](int kite::CountTunables::*field, std::initializer_list<int> cands) {
                for (int v : cands) {
                    if (best.*field == v) continue;
                    kite::CountTunables t = best;
                    t.*field = v;
                    const double m = time_of(t);
                    if (m < best_ms * 0.99) {
                        best = t;
                        best_ms = m;
                        changed = true;
                    }
                }
            };
            try_knob(&kite::CountTunables::split_target_div, {1, 2, 4, 8});
            try_knob(&kite::CountTunables::max_digit_bits, {8, 10, 11, 12});
            try_knob(&kite::CountTunables::hash_classes, {2, 3, 4});
            try_knob(&kite::CountTunables::extract_subs, {1, 2, 4, 8});
            try_knob(&kite::CountTunables::split_above_div, {1, 2});
            if (!changed) break;
        }
        const int cls = size_class(len);
        ctx->tuned[std::make_tuple(W, partitions, cls)] = best;
        ctx->count->set_base_tunables(kite::CountTunables{});
        if (report) {
            char head[512];
            snprintf(head, sizeof head,
                     "{\"words\": %d, \"partitions\": %u, \"size_class\": %d, \"input_bytes\": %llu, "
                     "\"default_ms\": %.4f, \"tuned_ms\": %.4f, \"chosen\": {\"split_target_div\": %d, "
                     "\"max_digit_bits\": %d, \"hash_classes\": %d, \"extract_subs\": %d, "
                     "\"split_above_div\": %d}, \"trials\": %d, \"trial_columns\": [\"split_target_div\", "
                     "\"max_digit_bits\", \"hash_classes\", \"extract_subs\", \"split_above_div\", \"ms\"], "
                     "\"trial_log\": [",
                     W, partitions, cls, (unsigned long long)len, default_ms, best_ms, best.split_target_div,
                     best.max_digit_bits, best.hash_classes, best.extract_subs, best.split_above_div, trials);
            const std::string js = std::string(head) + log + "]}";
            if (cap <= js.size()) return set_error(KITE_E_INVALID, "report buffer too small");
            memcpy(report, js.c_str(), js.size() + 1);
        }
        return KITE_OK;
    });
}

int kite_gpu_set_tuned(kite_gpu_ctx *ctx, uint32_t k, uint32_t partitions, int size_class, int split_target_div,
                       int max_digit_bits, int hash_classes, int extract_subs, int split_above_div) {
    if (!ctx) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL");
    if (k < 1 || k > 128) return set_error(KITE_E_INVALID, "invalid k-mer length: k must be between 1 and 128");
    if (partitions == 0 || (partitions & (partitions - 1)))
        return set_error(KITE_E_INVALID, "invalid partition count: must be a power of 2");
    auto one_of = [](int v, std::initializer_list<int> ok) {
        for (int o : ok)
            if (v == o) return true;
        return false;
    };
    if (size_class < 0 || size_class > 3 || !one_of(split_target_div, {1, 2, 4, 8}) || max_digit_bits < 8 ||
        max_digit_bits > 12 || hash_classes < 2 || hash_classes > 4 || !one_of(extract_subs, {1, 2, 4, 8}) ||
        !one_of(split_above_div, {1, 2}))
        return set_error(KITE_E_INVALID, "invalid count settings: outside the autotuner's candidates");
    return guarded([&]() -> int {
        kite::CountTunables t;
        t.split_target_div = split_target_div;
        t.max_digit_bits = max_digit_bits;
        t.hash_classes = hash_classes;
        t.extract_subs = extract_subs;
        t.split_above_div = split_above_div;
        const int W = k <= 31 ? 1 : (k <= 64 ? 2 : 4);
        ctx->tuned[std::make_tuple(W, partitions, size_class)] = t;
        return KITE_OK;
    });
}

int kite_gpu_copy_peak(kite_gpu_ctx *ctx, uint64_t bytes, double *h2d_gbps, double *d2h_gbps, double *both_gbps) {
    if (!ctx || bytes == 0) return set_error(KITE_E_INVALID, "invalid argument: ctx is NULL or bytes is 0");
    return guarded([&]() -> int {
        bind(ctx);
        PinnedBuf hs, hd;
        kite::DevBuf ds, dd;
        void *hsrc = hs.ensure(bytes), *hdst = hd.ensure(bytes);
        memset(hsrc, 1, bytes);
        memset(hdst, 2, bytes);
        void *dsrc = ds.ensure(bytes), *ddst = dd.ensure(bytes);
        hipStream_t a = nullptr, b = nullptr;
        KITE_HIP(hipStreamCreateWithFlags(&a, hipStreamNonBlocking));
        KITE_HIP(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
        struct Streams {
            hipStream_t x, y;
            ~Streams() {
                (void)hipStreamDestroy(x);
                (void)hipStreamDestroy(y);
            }
        } guard{a, b};
        // one untimed copy each way first (first use of the buffers and DMA
        // queues), then the best of five of each mode
        KITE_HIP(hipMemcpyAsync(ddst, hsrc, bytes, hipMemcpyHostToDevice, a));
        KITE_HIP(hipMemcpyAsync(hdst, dsrc, bytes, hipMemcpyDeviceToHost, b));
        KITE_HIP(hipStreamSynchronize(a));
        KITE_HIP(hipStreamSynchronize(b));
        double best[3] = {0, 0, 0};
        for (int rep = 0; rep < 5; rep++)
            for (int mode = 1; mode <= 3; mode++) {
                KITE_HIP(hipDeviceSynchronize());
                const auto t0 = std::chrono::steady_clock::now();
                if (mode & 1) KITE_HIP(hipMemcpyAsync(ddst, hsrc, bytes, hipMemcpyHostToDevice, a));
                if (mode & 2) KITE_HIP(hipMemcpyAsync(hdst, dsrc, bytes, hipMemcpyDeviceToHost, b));
                KITE_HIP(hipStreamSynchronize(a));
                KITE_HIP(hipStreamSynchronize(b));
                const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                const double gbps = (double)bytes * (mode == 3 ? 2 : 1) / sec / 1e9;
                best[mode - 1] = std::max(best[mode - 1], gbps);
            }
        if (h2d_gbps) *h2d_gbps = best[0];
        if (d2h_gbps) *d2h_gbps = best[1];
        if (both_gbps) *both_gbps = best[2];
        return KITE_OK;
    });
}

int kite_gpu_host_alloc(uint64_t bytes, void **out) {
    if (!out) return set_error(KITE_E_INVALID, "invalid argument: out is NULL");
    *out = nullptr;
    return guarded([&]() -> int {
        KITE_HIP(hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault));
        return KITE_OK;
    });
}

void kite_gpu_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}


}  // extern "C"

namespace {

// Validates the stored runs of a dump or histo call (the reader's checks
// on each run, src/kiteops/mod.rs:475-585).
int check_runs(uint32_t key_width, uint32_t partitions, const kite_dump_run *runs, uint64_t n_runs) {
    if (n_runs && !runs) return set_error(KITE_E_INVALID, "invalid argument: runs is NULL");
    if (key_width != 64 && key_width != 128 && key_width != 256)
        return set_error(KITE_E_INVALID, "invalid key width " + std::to_string(key_width));
    const int W = (int)(key_width / 64);
    const uint64_t es = 8 * (uint64_t)W + 4;
    for (uint64_t i = 0; i < n_runs; i++) {
        const kite_dump_run &r = runs[i];
        if (r.partition >= partitions) return set_error(KITE_E_INVALID, "invalid run: partition out of range");
        if (i && r.partition < runs[i - 1].partition)
            return set_error(KITE_E_INVALID, "invalid run order: partitions must be non-decreasing");
        if (r.codec != KITE_CODEC_DVLE && r.codec != KITE_CODEC_RAW)
            return set_error(KITE_E_INVALID, "invalid run codec " + std::to_string(r.codec) +
                                                 " (zstd payloads are decoded by the caller)");
        if (r.bytes && !r.data) return set_error(KITE_E_INVALID, "invalid run: data is NULL");
        const bool raw = W == 4 || r.codec == KITE_CODEC_RAW;
        if (r.rows && raw && r.bytes / es < r.rows)
            return set_error(KITE_E_INVALID, "invalid run: raw row payload truncated");
        if (r.rows && !raw && r.bytes < es)
            return set_error(KITE_E_INVALID, W == 2 ? "invalid run: u128 DVLE payload too small"
                                                    : "invalid run: u64 DVLE payload too small");
    }
    return KITE_OK;
}

// Stages the runs group by group (whole partitions, bounded by payload
// bytes and rows) in device memory and calls fn(d_payload, runs, np,
// rows_in, n_runs_in_group) for each group.
template <class F>
void for_each_run_group(kite_gpu_ctx *ctx, int W, const kite_dump_run *runs, uint64_t n_runs, F &&fn) {
    // (KITE_DUMP_GROUP_ROWS lowers the row bound, to exercise many groups)
    constexpr uint64_t PAY_CAP = 1ull << 30;
    uint64_t ROW_CAP = 96ull << 20;
    if (const char *e = getenv("KITE_DUMP_GROUP_ROWS")) ROW_CAP = std::max<uint64_t>(1, strtoull(e, nullptr, 10));
    hipStream_t s = ctx->stream;
    uint64_t i = 0;
    while (i < n_runs) {
        uint64_t j = i, pay = 0, rows = 0;
        while (j < n_runs) {
            // the whole partition of run j joins, or the group ends here
            uint64_t e = j, pb = 0, pr = 0;
            while (e < n_runs && runs[e].partition == runs[j].partition) {
                pb += (runs[e].bytes + 15) & ~15ull;
                pr += runs[e].rows;
                e++;
            }
            if (j > i && (pay + pb > PAY_CAP || rows + pr > ROW_CAP)) break;
            pay += pb;
            rows += pr;
            j = e;
        }
        const uint32_t p0 = runs[i].partition, np = runs[j - 1].partition - p0 + 1;
        uint8_t *host = static_cast<uint8_t *>(ctx->dump_stage.ensure(pay + 16));
        std::vector<kite::DumpRun> dr;
        dr.reserve(j - i);
        uint64_t off = 0;
        uint32_t rank = 0;
        for (uint64_t q = i; q < j; q++) {
            const kite_dump_run &r = runs[q];
            rank = (q > i && r.partition == runs[q - 1].partition) ? rank + 1 : 0;
            kite::DumpRun d{};
            d.src = off;
            d.bytes = r.bytes;
            d.rows = r.rows;
            d.codec = (W == 4 || r.codec == KITE_CODEC_RAW) ? kite::DUMP_CODEC_RAW : kite::DUMP_CODEC_DVLE;
            d.part = r.partition - p0;
            d.table = rank;
            dr.push_back(d);
            off += (r.bytes + 15) & ~15ull;
        }
        memset(host + off, 0, 16);
        // the payloads into page-locked staging on several host threads
        // (one thread copies a few GB/s out of a memory-mapped file)
        {
            const uint64_t nrun = j - i;
            const unsigned nt = (unsigned)std::min<uint64_t>(
                std::max<uint64_t>(1, pay >> 26), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
            std::atomic<uint64_t> next{0};
            auto copy = [&] {
                for (uint64_t q; (q = next.fetch_add(1)) < nrun;)
                    if (runs[i + q].bytes) memcpy(host + dr[q].src, runs[i + q].data, runs[i + q].bytes);
            };
            std::vector<std::thread> th;
            for (unsigned t = 1; t < nt; t++) th.emplace_back(copy);
            copy();
            for (auto &x : th) x.join();
        }
        uint8_t *dev = static_cast<uint8_t *>(ctx->dump_pay.ensure(pay + 16));
        KITE_HIP(hipMemcpyAsync(dev, host, pay + 16, hipMemcpyHostToDevice, s));
        fn(dev, dr, np, rows, j - i);
        i = j;
    }
}

kite::DumpFilter make_filter(int has_min, uint32_t min_count, int has_max, uint32_t max_count) {
    kite::DumpFilter f;
    f.has_min = has_min ? 1u : 0u;
    f.min = min_count;
    f.has_max = has_max ? 1u : 0u;
    f.max = max_count;
    return f;
}

}  // namespace

extern "C" {

int kite_gpu_dump(kite_gpu_ctx *ctx, uint32_t k, uint32_t key_width, uint32_t partitions, const kite_dump_run *runs,
                  uint64_t n_runs, int has_min, uint32_t min_count, int has_max, uint32_t max_count,
                  kite_dump_write_fn fn, void *user, kite_dump_stats *stats) {
    if (!ctx || !fn) return set_error(KITE_E_INVALID, "invalid argument: NULL pointer");
    if (int rc = check_runs(key_width, partitions, runs, n_runs)) return rc;
    const int W = (int)(key_width / 64);
    if (k < 1 || k > 32u * (uint32_t)W)
        return set_error(KITE_E_INVALID, "invalid k " + std::to_string(k) + " for key width " +
                                             std::to_string(key_width));
    return guarded([&]() -> int {
        bind(ctx);
        kite_dump_stats st{};
        const kite::DumpFilter f = make_filter(has_min, min_count, has_max, max_count);
        constexpr uint64_t PIECE = 64ull << 20;  // text bytes per D2H piece
        hipStream_t s = ctx->stream;
        bool aborted = false;
        for_each_run_group(ctx, W, runs, n_runs, [&](const uint8_t *dev, const std::vector<kite::DumpRun> &dr,
                                                     uint32_t np, uint64_t rows_in, uint64_t nruns) {
            if (aborted) return;
            uint64_t out_rows = 0, lines = 0;
            kite::DumpTimes t;
            const uint64_t bytes = ctx->dump->run(dev, dr, np, (int)k, W, f, &out_rows, &lines, &t);
            st.decode_ms += t.decode_ms;
            st.merge_ms += t.merge_ms;
            st.format_ms += t.format_ms;
            // text to the caller in pieces through two page-locked buffers:
            // pieces q and q + 1 are in flight while the caller takes piece
            // q - 1, and piece q + 2 is queued once the caller returns piece
            // q (its buffer), so the copy engine does not wait on the caller
            char *pin = static_cast<char *>(ctx->dump_text.ensure(2 * PIECE));
            const char *text = ctx->dump->text();
            const uint64_t pieces = kite::ceil_div(bytes, PIECE);
            kite::HostLink *link = host_link(ctx);
            hipEvent_t done[2];
            for (auto *e : {&done[0], &done[1]}) KITE_HIP(hipEventCreate(e));
            struct Ev {
                hipEvent_t *e[2];
                kite::HostLink *link;
                ~Ev() {
                    if (link) link->quiesce();
                    for (auto *x : e) (void)hipEventDestroy(*x);
                }
            } evs{{&done[0], &done[1]}, link};
            // the text is complete before an SDMA copy reads it
            if (link) KITE_HIP(hipStreamSynchronize(s));
            const auto c0 = std::chrono::steady_clock::now();
            auto issue = [&](uint64_t q) {
                const uint64_t a = q * PIECE, n = std::min<uint64_t>(PIECE, bytes - a);
                if (link) {
                    link->d2h((int)(q & 1), pin + (q & 1) * PIECE, text + a, n);
                } else {
                    KITE_HIP(hipMemcpyAsync(pin + (q & 1) * PIECE, text + a, n, hipMemcpyDeviceToHost, s));
                    KITE_HIP(hipEventRecord(done[q & 1], s));
                }
            };
            for (uint64_t q = 0; q < std::min<uint64_t>(2, pieces); q++) issue(q);
            for (uint64_t q = 0; q < pieces; q++) {
                if (link) link->wait((int)(q & 1));
                else KITE_HIP(hipEventSynchronize(done[q & 1]));
                const uint64_t a = q * PIECE, n = std::min<uint64_t>(PIECE, bytes - a);
                if (fn(user, pin + (q & 1) * PIECE, n) != 0) {
                    (void)hipStreamSynchronize(s);
                    aborted = true;
                    return;
                }
                if (q + 2 < pieces) issue(q + 2);
            }
            KITE_HIP(hipStreamSynchronize(s));
            st.copy_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
            st.runs += nruns;
            st.rows_in += rows_in;
            st.rows += out_rows;
            st.lines += lines;
            st.bytes += bytes;
            st.groups++;
        });
        if (aborted) return set_error(KITE_E_CALLBACK, "dump writer callback aborted");
        if (stats) *stats = st;
        return KITE_OK;
    });
}

int kite_gpu_histo(kite_gpu_ctx *ctx, uint32_t key_width, uint32_t partitions, const kite_dump_run *runs,
                   uint64_t n_runs, int has_min, uint32_t min_count, int has_max, uint32_t max_count,
                   kite_histo_fn fn, void *user, kite_dump_stats *stats) {
    if (!ctx || !fn) return set_error(KITE_E_INVALID, "invalid argument: NULL pointer");
    if (int rc = check_runs(key_width, partitions, runs, n_runs)) return rc;
    const int W = (int)(key_width / 64);
    return guarded([&]() -> int {
        bind(ctx);
        kite_dump_stats st{};
        const kite::DumpFilter f = make_filter(has_min, min_count, has_max, max_count);
        ctx->dump->histo_begin();
        for_each_run_group(ctx, W, runs, n_runs, [&](const uint8_t *dev, const std::vector<kite::DumpRun> &dr,
                                                     uint32_t np, uint64_t rows_in, uint64_t nruns) {
            uint64_t out_rows = 0;
            kite::DumpTimes t;
            ctx->dump->histo(dev, dr, np, W, f, &out_rows, &t);
            st.decode_ms += t.decode_ms;
            st.merge_ms += t.merge_ms;
            st.format_ms += t.format_ms;
            st.runs += nruns;
            st.rows_in += rows_in;
            st.rows += out_rows;
            st.groups++;
        });
        std::vector<uint64_t> values, freqs;
        ctx->dump->histo_end(values, freqs);
        for (uint64_t q : freqs) st.lines += q;
        if (stats) *stats = st;
        if (fn(user, values.data(), freqs.data(), values.size()) != 0)
            return set_error(KITE_E_CALLBACK, "histo callback aborted");
        return KITE_OK;
    });
}

}  // extern "C"

namespace {

// One host byte range bound for device memory: src[0, bytes) to dst.
struct HostRange {
    const uint8_t *src;
    uint64_t bytes;
    uint64_t dst;  // byte offset in the device buffer
};

// Copies the ranges (ascending, disjoint destinations) into dev[0, total)
// through two page-locked bounce buffers of PIECE bytes: up to 16 host
// threads fill one piece while the H2D of the other runs.  Gaps between
// ranges are left as they are.
void h2d_ranges(kite_gpu_ctx *ctx, uint8_t *dev, uint64_t total, const std::vector<HostRange> &rs) {
    constexpr uint64_t PIECE = 256ull << 20;
    hipStream_t s = ctx->stream;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    hipEvent_t done[2];
    for (auto &e : done) KITE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    struct Ev {
        hipEvent_t *e;
        ~Ev() {
            (void)hipEventDestroy(e[0]);
            (void)hipEventDestroy(e[1]);
        }
    } ev{done};
    size_t r0 = 0;  // first range that may overlap the current piece
    for (uint64_t q = 0, a = 0; a < total; q++, a += PIECE) {
        const uint64_t b = std::min(total, a + PIECE);
        uint8_t *pin = static_cast<uint8_t *>(ctx->merge_in[q & 1].ensure(PIECE));
        if (q >= 2) KITE_HIP(hipEventSynchronize(done[q & 1]));  // its last H2D has read it
        // the parts of the ranges inside [a, b), cut in 4 MiB copies
        struct Seg {
            const uint8_t *src;
            uint64_t bytes, at;
        };
        std::vector<Seg> seg;
        while (r0 < rs.size() && rs[r0].dst + rs[r0].bytes <= a) r0++;
        for (size_t r = r0; r < rs.size() && rs[r].dst < b; r++) {
            const uint64_t lo = std::max(a, rs[r].dst), hi = std::min(b, rs[r].dst + rs[r].bytes);
            for (uint64_t x = lo; x < hi; x += 4ull << 20)
                seg.push_back(Seg{rs[r].src + (x - rs[r].dst), std::min<uint64_t>(4ull << 20, hi - x), x - a});
        }
        std::atomic<size_t> next{0};
        auto copy = [&] {
            for (size_t i; (i = next.fetch_add(1)) < seg.size();) memcpy(pin + seg[i].at, seg[i].src, seg[i].bytes);
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < std::min<size_t>(nt, seg.size()); t++) th.emplace_back(copy);
        copy();
        for (auto &x : th) x.join();
        KITE_HIP(hipMemcpyAsync(dev + a, pin, b - a, hipMemcpyHostToDevice, s));
        KITE_HIP(hipEventRecord(done[q & 1], s));
    }
}

std::string u128_decimal(uint64_t lo, uint64_t hi) {
    unsigned __int128 v = ((unsigned __int128)hi << 64) | lo;
    std::string d;
    do {
        d.push_back(char('0' + (int)(v % 10)));
        v /= 10;
    } while (v);
    return std::string(d.rbegin(), d.rend());
}

}  // namespace

extern "C" {

int kite_gpu_merge_runs(kite_gpu_ctx *ctx, uint32_t key_width, uint32_t partitions, const kite_merge_input *inputs,
                        uint32_t n_inputs, kite_append_chunk_fn fn, void *user, kite_merge_stats *stats) {
    if (!ctx || !fn || (n_inputs && !inputs)) return set_error(KITE_E_INVALID, "invalid argument: NULL pointer");
    if (n_inputs == 0) return set_error(KITE_E_INVALID, "invalid argument: no inputs");
    if (partitions == 0) return set_error(KITE_E_INVALID, "invalid partition count 0");
    for (uint32_t i = 0; i < n_inputs; i++)
        if (int rc = check_runs(key_width, partitions, inputs[i].runs, inputs[i].n_runs)) return rc;
    const int W = (int)(key_width / 64);
    return guarded([&]() -> int {
        bind(ctx);
        hipStream_t s = ctx->stream;
        kite_merge_stats st{};
        // first run of every partition, per input
        std::vector<std::vector<uint64_t>> first(n_inputs, std::vector<uint64_t>((size_t)partitions + 1, 0));
        for (uint32_t i = 0; i < n_inputs; i++) {
            const kite_merge_input &in = inputs[i];
            uint64_t r = 0;
            for (uint32_t p = 0; p <= partitions; p++) {
                while (r < in.n_runs && in.runs[r].partition < p) r++;
                first[i][p] = r;
            }
            st.runs += in.n_runs;
        }
        // partition groups bounded by rows over all inputs (device memory is
        // about six copies of a row of key and u64 count, plus its chunk
        // bytes) and by stored payload bytes
        uint64_t row_cap = std::min<uint64_t>(384ull << 20, std::max<uint64_t>(1, ctx->budget / (6 * (8 * W + 8) + 64)));
        if (const char *e = getenv("KITE_MERGE_GROUP_ROWS")) row_cap = std::max<uint64_t>(1, strtoull(e, nullptr, 10));
        constexpr uint64_t PAY_CAP = 8ull << 30;
        if (ctx->merge_tabs.size() != n_inputs) {
            std::vector<kite::DevTable> fresh(n_inputs);  // DevTable owns device memory and does not move
            ctx->merge_tabs.swap(fresh);
        }
        const uint32_t codec = W == 4 ? KITE_CODEC_RAW : KITE_CODEC_DVLE;
        uint32_t p0 = 0;
        while (p0 < partitions) {
            uint32_t p1 = p0;
            uint64_t rows = 0, pay = 0;
            while (p1 < partitions) {
                uint64_t pr = 0, pb = 0;
                for (uint32_t i = 0; i < n_inputs; i++)
                    for (uint64_t r = first[i][p1]; r < first[i][p1 + 1]; r++) {
                        pr += inputs[i].runs[r].rows;
                        pb += (inputs[i].runs[r].bytes + 15) & ~15ull;
                    }
                if (p1 > p0 && (rows + pr > row_cap || pay + pb > PAY_CAP)) break;
                rows += pr;
                pay += pb;
                p1++;
            }
            const uint32_t np = p1 - p0;
            // stored payloads of the group: input by input, 16-byte aligned
            std::vector<std::vector<kite::DumpRun>> dr(n_inputs);
            std::vector<HostRange> ranges;
            uint64_t off = 0;
            for (uint32_t i = 0; i < n_inputs; i++) {
                uint32_t rank = 0;
                for (uint64_t r = first[i][p0]; r < first[i][p1]; r++) {
                    const kite_dump_run &x = inputs[i].runs[r];
                    rank = (r > first[i][p0] && x.partition == inputs[i].runs[r - 1].partition) ? rank + 1 : 0;
                    kite::DumpRun d{};
                    d.src = off;
                    d.bytes = x.bytes;
                    d.rows = x.rows;
                    d.codec = (W == 4 || x.codec == KITE_CODEC_RAW) ? kite::DUMP_CODEC_RAW : kite::DUMP_CODEC_DVLE;
                    d.part = x.partition - p0;
                    d.table = rank;
                    dr[i].push_back(d);
                    if (x.bytes) ranges.push_back(HostRange{x.data, x.bytes, off});
                    off += (x.bytes + 15) & ~15ull;
                }
            }
            const auto t_in = std::chrono::steady_clock::now();
            uint8_t *dev = static_cast<uint8_t *>(ctx->dump_pay.ensure(off + 16));
            h2d_ranges(ctx, dev, off + 16, ranges);
            KITE_HIP(hipStreamSynchronize(s));
            st.in_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_in).count();
            st.rows_in += rows;

            // every input decoded, summed per key and filtered into its table
            for (uint32_t i = 0; i < n_inputs; i++) {
                const uint64_t *k = nullptr, *c = nullptr, *o = nullptr;
                kite::DumpTimes dt;
                const uint64_t n = ctx->dump->decode_sum(dev, dr[i], np, W, &k, &c, &o, &dt);
                st.decode_ms += dt.decode_ms + dt.merge_ms;
                kite::DumpFilter f;
                f.has_min = inputs[i].has_min ? 1u : 0u;
                f.min = inputs[i].min_count;
                f.has_max = inputs[i].has_max ? 1u : 0u;
                f.max = inputs[i].max_count;
                ctx->merge->filter_into(k, c, o, n, np, W, f, ctx->merge_tabs[i]);
                st.rows_kept += ctx->merge_tabs[i].rows;
            }
            hipEvent_t e[4];
            for (auto &x : e) KITE_HIP(hipEventCreate(&x));
            struct Ev4 {
                hipEvent_t *e;
                ~Ev4() {
                    for (int q = 0; q < 4; q++) (void)hipEventDestroy(e[q]);
                }
            } ev4{e};
            KITE_HIP(hipEventRecord(e[0], s));
            kite::DevTable &u = ctx->merge_res;
            ctx->merge->merge_all(ctx->merge_tabs, np, W, u);
            KITE_HIP(hipEventRecord(e[1], s));
            // serialize_partition_chunk: every total must fit u32
            uint32_t *c32 = static_cast<uint32_t *>(ctx->merge_c32.ensure(std::max<uint64_t>(u.rows, 1) * 4));
            const uint64_t bad = ctx->merge->narrow_u32(u.counts.as<uint64_t>(), u.rows, c32);
            if (bad < u.rows) {
                uint64_t kw[4] = {0, 0, 0, 0};
                KITE_HIP(hipMemcpy(kw, u.keys.as<uint64_t>() + bad * W, 8 * W, hipMemcpyDeviceToHost));
                return set_error(KITE_E_OVERFLOW, W == 4 ? std::string("merged count exceeds u32")
                                                         : "merged count exceeds u32 for key " +
                                                               u128_decimal(kw[0], W == 2 ? kw[1] : 0));
            }
            uint64_t *hb = static_cast<uint64_t *>(ctx->merge_enc.host_byte_off.ensure(((size_t)np + 1) * 8));
            const kite::EncodeDevice ed = kite::encode_on_device(u.keys.as<uint64_t>(), c32, u.off.as<uint64_t>(),
                                                                 u.rows, np, W, s, ctx->merge_enc, hb);
            KITE_HIP(hipEventRecord(e[2], s));
            std::vector<uint64_t> boff(hb, hb + np + 1), roff((size_t)np + 1), mm((size_t)np * 8);
            KITE_HIP(hipMemcpyAsync(roff.data(), u.off.ptr, ((size_t)np + 1) * 8, hipMemcpyDeviceToHost, s));
            KITE_HIP(hipMemcpyAsync(mm.data(), ed.minmax, (size_t)np * 64, hipMemcpyDeviceToHost, s));
            KITE_HIP(hipStreamSynchronize(s));
            float ms = 0;
            KITE_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
            st.merge_ms += ms;
            KITE_HIP(hipEventElapsedTime(&ms, e[1], e[2]));
            st.encode_ms += ms;
            st.rows += u.rows;
            st.bytes += ed.total;

            const auto t_out = std::chrono::steady_clock::now();
            if (emit_device_chunks(ctx, ed.bytes, boff.data(), roff.data(), mm.data(), np, p0, codec, fn, user) >= 0)
                return set_error(KITE_E_CALLBACK, "merge chunk callback aborted");
            st.out_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_out).count();
            st.groups++;
            p0 = p1;
        }
        if (stats) *stats = st;
        return KITE_OK;
    });
}

}  // extern "C"

