// kitefmt.cpp -- .kite writer, reader, run codecs and key text.
#include "kitefmt.hpp"

#include <dlfcn.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <functional>

namespace kh {

namespace {
void le16(uint8_t *p, uint16_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}
void le32(uint8_t *p, uint32_t v) {
    for (int i = 0; i < 4; i++) p[i] = (uint8_t)(v >> (8 * i));
}
void le64(uint8_t *p, uint64_t v) {
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}
uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const uint8_t *p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}
uint64_t rd64(const uint8_t *p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}
}  // namespace

void Header::write(uint8_t h[128]) const {
    memset(h, 0, 128);
    memcpy(h, "KX2!", 4);
    le16(h + 4, 1);
    h[6] = k;
    h[7] = canonical;
    le32(h + 8, partitions);
    le16(h + 12, key_width);
    le16(h + 14, codec_mask);
    le32(h + 16, 0);
    le64(h + 20, chunk_index_offset);
    le64(h + 28, partition_table_offset);
    le64(h + 36, footer_offset);
    le64(h + 44, total_chunks);
    le64(h + 52, total_rows);
    if (min_count || max_count) {
        memcpy(h + 60, "KMT1", 4);
        le32(h + 64, (min_count ? 1u : 0u) | (max_count ? 2u : 0u));
        le32(h + 68, min_count.value_or(0));
        le32(h + 72, max_count.value_or(0));
    }
}

Header Header::parse(const uint8_t *h) {
    if (memcmp(h, "KX2!", 4) != 0) throw Error("invalid magic: expected KITE");
    uint16_t ver = rd16(h + 4);
    if (ver != 1) throw Error("unsupported version: " + std::to_string(ver));
    Header r;
    r.k = h[6];
    r.canonical = h[7];
    r.partitions = rd32(h + 8);
    r.key_width = rd16(h + 12);
    r.codec_mask = rd16(h + 14);
    r.chunk_index_offset = rd64(h + 20);
    r.partition_table_offset = rd64(h + 28);
    r.footer_offset = rd64(h + 36);
    r.total_chunks = rd64(h + 44);
    r.total_rows = rd64(h + 52);
    if (memcmp(h + 60, "KMT1", 4) == 0) {
        uint32_t fl = rd32(h + 64);
        if (fl & 1) r.min_count = rd32(h + 68);
        if (fl & 2) r.max_count = rd32(h + 72);
    }
    return r;
}

// ---- writer -----------------------------------------------------------------
namespace {

// Splits one large copy or file write over a few helper threads and the
// caller (appends come from one thread, and one core copies only a few GB/s
// out of page-locked memory).  Parts of at least PART bytes; smaller
// requests run on the caller alone.
class SplitPool {
   public:
    static constexpr size_t PART = 1u << 20;
    static SplitPool &get() {
        static SplitPool p;
        return p;
    }
    // fn(part offset, part length) over [0, len) in parts, the caller
    // taking its share; returns the first error message ("" when none)
    template <class F>
    std::string run(size_t len, F &&fn) {
        const size_t parts = std::min<size_t>(threads_.size() + 1, len / PART);
        if (parts <= 1) {
            try {
                fn(size_t(0), len);
            } catch (const std::exception &e) {
                return e.what();
            }
            return "";
        }
        std::lock_guard<std::mutex> one(caller_);  // one request at a time
        const size_t step = ((len + parts - 1) / parts + 63) & ~size_t(63);
        Task t;
        t.fn = [&](size_t q) {
            const size_t a = q * step;
            if (a < len) fn(a, std::min(step, len - a));
        };
        t.parts = parts;
        {
            std::lock_guard<std::mutex> g(mu_);
            task_ = &t;
            gen_++;
        }
        cv_.notify_all();
        work(t);
        // every part done and no helper still holding the task (it lives on
        // this stack); no helper can take it once task_ is cleared here
        std::unique_lock<std::mutex> g(mu_);
        done_cv_.wait(g, [&] { return t.done == t.parts && t.holders == 0; });
        task_ = nullptr;
        return t.err;
    }

   private:
    struct Task {
        std::function<void(size_t)> fn;
        size_t parts = 0, next = 0, done = 0, holders = 0;
        std::string err;
    };
    SplitPool() {
        const unsigned n = std::min(4u, std::max(1u, std::thread::hardware_concurrency()) - 1u);
        for (unsigned i = 0; i < n; i++) threads_.emplace_back([this] { loop(); });
    }
    ~SplitPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    // takes parts of t until none is left (mu_ not held on entry)
    void work(Task &t) {
        for (;;) {
            size_t q;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (t.next >= t.parts) return;
                q = t.next++;
            }
            std::string e;
            try {
                t.fn(q);
            } catch (const std::exception &x) {
                e = x.what();
            }
            bool last;
            {
                std::lock_guard<std::mutex> g(mu_);
                if (!e.empty() && t.err.empty()) t.err = e;
                last = ++t.done == t.parts;
            }
            if (last) done_cv_.notify_all();
        }
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            Task *t;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || (gen_ != seen && task_); });
                if (stop_) return;
                seen = gen_;
                t = task_;
                t->holders++;
            }
            work(*t);
            {
                std::lock_guard<std::mutex> g(mu_);
                t->holders--;
            }
            done_cv_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_, caller_;
    std::condition_variable cv_, done_cv_;
    Task *task_ = nullptr;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

}  // namespace

int Writer::suggest_write_lanes(int workers) {
    int l = (workers + 7) / 8;
    return std::min(6, std::max(2, l));
}

Writer::Writer(const std::string &path, const Header &h, size_t staging_bytes, int write_lanes)
    : h_(h), S_(staging_bytes) {
    f_ = fopen(path.c_str(), "w+b");
    if (!f_) throw Error("failed to create " + path + ": " + strerror(errno));
    struct stat sb;
    regular_ = fstat(fileno(f_), &sb) == 0 ? S_ISREG(sb.st_mode) : true;
    setvbuf(f_, nullptr, _IOFBF, 8 << 20);
    size_t hw = staging_bytes * (size_t)write_lanes * 4;
    high_ = std::max(hw, staging_bytes);
    low_ = high_ / 2;
    spare_cap_ = high_ / std::max<size_t>(1, staging_bytes) + 2 * (size_t)std::max(1, write_lanes) + 4;
    st_.resize(h.partitions);
    uint8_t hdr[128];
    Header prov = h;
    prov.chunk_index_offset = prov.partition_table_offset = prov.footer_offset = 0;
    prov.total_chunks = prov.total_rows = 0;
    prov.write(hdr);
    if (fwrite(hdr, 1, 128, f_) != 128) throw Error("write failed");
    if (regular_) {
        if (fflush(f_) != 0) throw Error("write failed");
        fd_ = fileno(f_);
        for (int i = 0; i < std::max(1, write_lanes); i++) pool_.emplace_back([this] { worker(); });
    }
}

Writer::~Writer() {
    stop_writers();
    if (f_) fclose(f_);
}

void Writer::write_at(const uint8_t *data, size_t len, uint64_t off) {
    while (len) {
        const ssize_t w = pwrite(fd_, data, len, (off_t)off);
        if (w <= 0) {
            if (w < 0 && errno == EINTR) continue;
            throw Error(std::string("write failed: ") + strerror(errno));
        }
        data += w;
        len -= (size_t)w;
        off += (uint64_t)w;
    }
}

void Writer::worker() {
    for (;;) {
        Job j;
        {
            std::unique_lock<std::mutex> g(mu_);
            cv_.wait(g, [&] { return stop_ || !jobs_.empty(); });
            if (jobs_.empty()) return;
            j = std::move(jobs_.front());
            jobs_.pop_front();
            busy_++;
        }
        std::string e;
        try {
            write_at(j.buf.data(), j.buf.size(), j.off);
        } catch (const std::exception &x) {
            e = x.what();
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            busy_--;
            queued_bytes_ -= j.buf.size();
            if (!e.empty() && werr_.empty()) werr_ = e;
            if (spare_.size() < spare_cap_) {
                j.buf.clear();
                spare_.push_back(std::move(j.buf));
            }
        }
        cv_done_.notify_all();
    }
}

void Writer::wait_writes() {
    std::unique_lock<std::mutex> g(mu_);
    cv_done_.wait(g, [&] { return jobs_.empty() && busy_ == 0; });
    if (!werr_.empty()) throw Error(werr_);
}

void Writer::stop_writers() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : pool_) t.join();
    pool_.clear();
}

uint64_t Writer::claim(const uint8_t *data, size_t len) {
    uint64_t off = next_offset_;
    if (fd_ >= 0) {
        // the caller's buffer: written now, in parts at their own offsets
        const std::string e =
            SplitPool::get().run(len, [&](size_t a, size_t n) { write_at(data + a, n, off + a); });
        if (!e.empty()) throw Error(e);
    } else if (len && fwrite(data, 1, len, f_) != len) {
        throw Error("write failed");
    }
    next_offset_ += len;
    return off;
}

void Writer::flush_partition(uint32_t p) {
    Stage &s = st_[p];
    if (s.buf.empty()) return;
    const size_t n = s.buf.size();
    uint64_t off;
    if (fd_ >= 0) {
        // the staging buffer goes to a writer thread; at most high_ bytes queued
        off = next_offset_;
        next_offset_ += n;
        std::unique_lock<std::mutex> g(mu_);
        cv_done_.wait(g, [&] { return queued_bytes_ < high_ || !werr_.empty(); });
        if (!werr_.empty()) throw Error(werr_);
        queued_bytes_ += n;
        jobs_.push_back(Job{off, std::move(s.buf)});
        g.unlock();
        cv_.notify_one();
        s.buf = Bytes();
    } else {
        off = claim(s.buf.data(), s.buf.size());
    }
    staged_ -= n;
    s.buf.clear();
    for (ChunkMeta m : s.pending) {
        m.offset += off;
        s.committed.push_back(m);
    }
    s.pending.clear();
}

bool Writer::drain_one() {
    size_t best = 0;
    int64_t bp = -1;
    for (uint32_t p = 0; p < h_.partitions; p++)
        if (st_[p].buf.size() > best) {
            best = st_[p].buf.size();
            bp = p;
        }
    if (bp < 0) return false;
    flush_partition((uint32_t)bp);
    return true;
}

void Writer::append_chunk(uint32_t p, uint32_t codec, uint64_t rows, uint64_t raw_bytes,
                          const uint64_t kmin[4], const uint64_t kmax[4], const uint8_t *data,
                          size_t len) {
    if (p >= h_.partitions) throw Error("partition id out of range");
    Stage &s = st_[p];
    ChunkMeta m;
    m.partition = p;
    m.codec = codec;
    m.rows = rows;
    m.raw = raw_bytes;
    m.encoded = len;
    for (int i = 0; i < 4; i++) {
        m.kmin[i] = kmin[i];
        m.kmax[i] = kmax[i];
    }
    if (!s.buf.empty() && s.buf.size() + len > S_) flush_partition(p);
    if (len >= S_) {
        m.offset = claim(data, len);
        s.committed.push_back(m);
        return;
    }
    m.offset = s.buf.size();
    if (s.buf.capacity() == 0) {
        std::lock_guard<std::mutex> g(mu_);
        if (!spare_.empty()) {
            s.buf.swap(spare_.back());
            spare_.pop_back();
        }
    }
    if (s.buf.capacity() < s.buf.size() + len) s.buf.reserve(std::max(s.buf.size() + len, std::min(S_, 2 * s.buf.capacity())));
    const size_t at = s.buf.size();
    s.buf.resize(at + len);
    uint8_t *dst = s.buf.data() + at;
    (void)SplitPool::get().run(len, [&](size_t a, size_t n) { memcpy(dst + a, data + a, n); });
    s.pending.push_back(m);
    staged_ += len;
    if (s.buf.size() >= S_) flush_partition(p);
    if (staged_ >= high_) {
        int attempts = 0;
        while (staged_ > low_ && attempts < 4) {
            if (!drain_one()) break;
            attempts++;
        }
    }
}

std::pair<uint64_t, uint64_t> Writer::finalize() {
    if (finalized_) throw Error("writer finalized twice");
    finalized_ = true;
    uint64_t total_chunks = 0, total_rows = 0;
    for (uint32_t p = 0; p < h_.partitions; p++) flush_partition(p);
    if (fd_ >= 0) {
        wait_writes();
        stop_writers();
        if (fseeko(f_, (off_t)next_offset_, SEEK_SET) != 0) throw Error("write failed");
    }
    for (uint32_t p = 0; p < h_.partitions; p++) {
        auto &c = st_[p].committed;
        std::sort(c.begin(), c.end(), [](const ChunkMeta &a, const ChunkMeta &b) { return a.offset < b.offset; });
        total_chunks += c.size();
        for (auto &m : c) total_rows += m.rows;
    }
    const uint64_t data_end = next_offset_;
    Header h = h_;
    h.chunk_index_offset = data_end;
    h.partition_table_offset = data_end + total_chunks * 128;
    h.footer_offset = h.partition_table_offset + (uint64_t)h_.partitions * 40;
    h.total_chunks = total_chunks;
    h.total_rows = total_rows;
    if (regular_) {
        std::vector<uint8_t> tr;
        tr.reserve(total_chunks * 128 + h_.partitions * 40 + 24);
        uint8_t e[128];
        for (uint32_t p = 0; p < h_.partitions; p++)
            for (auto &m : st_[p].committed) {
                memset(e, 0, 128);
                le32(e, p);
                le32(e + 4, m.codec);
                le64(e + 8, m.offset);
                le64(e + 16, m.encoded);
                le64(e + 24, m.raw);
                le64(e + 32, m.rows);
                for (int i = 0; i < 4; i++) le64(e + 40 + 8 * i, m.kmin[i]);
                for (int i = 0; i < 4; i++) le64(e + 72 + 8 * i, m.kmax[i]);
                tr.insert(tr.end(), e, e + 128);
            }
        uint64_t first = 0;
        for (uint32_t p = 0; p < h_.partitions; p++) {
            uint8_t pe[40];
            uint64_t rows = 0, bytes = 0, dmin = 0;
            const auto &c = st_[p].committed;
            for (size_t i = 0; i < c.size(); i++) {
                rows += c[i].rows;
                bytes += c[i].encoded;
                if (i == 0 || c[i].offset < dmin) dmin = c[i].offset;
            }
            le64(pe, first);
            le64(pe + 8, c.size());
            le64(pe + 16, rows);
            le64(pe + 24, dmin);
            le64(pe + 32, bytes);
            tr.insert(tr.end(), pe, pe + 40);
            first += c.size();
        }
        uint8_t ft[24];
        le64(ft, total_chunks);
        le64(ft + 8, total_rows);
        memcpy(ft + 16, "KX2FOOT!", 8);
        tr.insert(tr.end(), ft, ft + 24);
        if (fwrite(tr.data(), 1, tr.size(), f_) != tr.size()) throw Error("write failed");
        uint8_t hdr[128];
        h.write(hdr);
        fflush(f_);
        if (fseeko(f_, 0, SEEK_SET) != 0 || fwrite(hdr, 1, 128, f_) != 128) throw Error("write failed");
        fflush(f_);
        if (ftruncate(fileno(f_), (off_t)(h.footer_offset + 24)) != 0) throw Error("truncate failed");
    }
    fclose(f_);
    f_ = nullptr;
    return {total_chunks, total_rows};
}

// ---- reader -----------------------------------------------------------------
Reader::Reader(const std::string &path) {
    int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) throw Error("failed to open " + path + ": " + strerror(errno));
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        throw Error("failed to stat " + path);
    }
    size_ = (size_t)sb.st_size;
    if (size_ < 128 + 24) {
        close(fd);
        throw Error("file too small");
    }
    void *m = mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) throw Error("failed to map " + path);
    base_ = static_cast<const uint8_t *>(m);
    try {
        h_ = Header::parse(base_);
        const uint64_t P = h_.partitions;
        if (h_.footer_offset + 24 > size_) throw Error("footer out of bounds");
        if (memcmp(base_ + h_.footer_offset + 16, "KX2FOOT!", 8) != 0) throw Error("invalid footer magic");
        if (h_.partition_table_offset + P * 40 > size_) throw Error("partition table out of bounds");
        if (h_.partition_table_offset < h_.chunk_index_offset) throw Error("invalid chunk index offsets");
        uint64_t cb = h_.partition_table_offset - h_.chunk_index_offset;
        if (cb % 128) throw Error("chunk index size misaligned");
        uint64_t nc = cb / 128;
        chunks_.resize(nc);
        for (uint64_t i = 0; i < nc; i++) {
            const uint8_t *e = base_ + h_.chunk_index_offset + i * 128;
            ChunkMeta &c = chunks_[i];
            c.partition = rd32(e);
            c.codec = rd32(e + 4);
            c.offset = rd64(e + 8);
            c.encoded = rd64(e + 16);
            c.raw = rd64(e + 24);
            c.rows = rd64(e + 32);
            for (int j = 0; j < 4; j++) c.kmin[j] = rd64(e + 40 + 8 * j);
            for (int j = 0; j < 4; j++) c.kmax[j] = rd64(e + 72 + 8 * j);
        }
        part_.resize(P);
        for (uint64_t p = 0; p < P; p++) {
            const uint8_t *e = base_ + h_.partition_table_offset + p * 40;
            uint64_t first = rd64(e), cnt = rd64(e + 8);
            if (first + cnt > nc) throw Error("partition " + std::to_string(p) + " chunk range out of bounds");
            for (uint64_t c = first; c < first + cnt; c++) {
                const ChunkMeta &cm = chunks_[c];
                if (cm.encoded == 0 || cm.offset + cm.encoded > size_)
                    throw Error("partition " + std::to_string(p) + " chunk out of bounds");
                if (cm.codec & ~(CODEC_DVLE | CODEC_RAW | CODEC_ZSTD))
                    throw Error("partition " + std::to_string(p) + " chunk uses unsupported codec mask");
            }
            part_[p] = {first, cnt};
        }
    } catch (...) {
        munmap(const_cast<uint8_t *>(base_), size_);
        throw;
    }
}

Reader::~Reader() {
    if (base_) munmap(const_cast<uint8_t *>(base_), size_);
}

std::vector<uint8_t> Reader::payload(const ChunkMeta &c) const {
    const uint8_t *p = base_ + c.offset;
    if (c.codec & CODEC_ZSTD) {
        auto out = zstd_decompress(p, c.encoded, c.raw);
        if (out.size() != c.raw) throw Error("decoded payload size mismatch");
        return out;
    }
    return std::vector<uint8_t>(p, p + c.encoded);
}

namespace {
size_t get_varint(const uint8_t *d, size_t avail, size_t maxb, uint64_t &lo, uint64_t &hi) {
    unsigned __int128 r = 0;
    unsigned sh = 0;
    size_t i = 0;
    for (; i < avail && i < maxb; i++) {
        if (sh < 128) r |= (unsigned __int128)(d[i] & 0x7f) << sh;
        if (!(d[i] & 0x80)) {
            lo = (uint64_t)r;
            hi = (uint64_t)(r >> 64);
            return i + 1;
        }
        sh += 7;
    }
    lo = (uint64_t)r;
    hi = (uint64_t)(r >> 64);
    return i;
}
}  // namespace

void Reader::decode_run(const ChunkMeta &c, std::vector<Row> &out) const {
    std::vector<uint8_t> tmp;
    const uint8_t *d;
    size_t len;
    if (c.codec & CODEC_ZSTD) {
        tmp = payload(c);
        d = tmp.data();
        len = tmp.size();
    } else {
        d = base_ + c.offset;
        len = c.encoded;
    }
    const int W = words();
    if (W == 4 || !(c.codec & CODEC_DVLE)) {
        const int es = entry_size(h_.key_width);
        for (uint64_t i = 0; i < c.rows; i++) {
            if ((i + 1) * es > len) throw Error("raw row payload truncated at row " + std::to_string(i));
            Row r;
            for (int w = 0; w < W; w++) r.key.w[w] = rd64(d + i * es + 8 * w);
            r.count = rd32(d + i * es + 8 * W);
            out.push_back(r);
        }
        return;
    }
    if (c.rows == 0) return;
    const size_t head = 8 * (size_t)W + 4;
    if (len < head) throw Error(W == 2 ? "u128 DVLE payload too small" : "u64 DVLE payload too small");
    Row r;
    r.key.w[0] = rd64(d);
    if (W == 2) r.key.w[1] = rd64(d + 8);
    r.count = rd32(d + 8 * W);
    out.push_back(r);
    size_t pos = head;
    for (uint64_t i = 1; i < c.rows; i++) {
        uint64_t dl, dh, cl, ch;
        size_t u = get_varint(d + pos, len - pos, W == 2 ? 19 : 10, dl, dh);
        if (!u) throw Error("invalid DVLE delta");
        pos += u;
        size_t v = get_varint(d + pos, len - pos, 10, cl, ch);
        if (!v) throw Error("invalid DVLE count");
        pos += v;
        if (W == 1) {
            r.key.w[0] += dl;
        } else {
            unsigned __int128 kv = ((unsigned __int128)r.key.w[1] << 64) | r.key.w[0];
            kv += ((unsigned __int128)dh << 64) | dl;
            r.key.w[0] = (uint64_t)kv;
            r.key.w[1] = (uint64_t)(kv >> 64);
        }
        r.count = cl;
        out.push_back(r);
    }
}

std::vector<Row> Reader::load_partition(uint32_t p) const {
    std::vector<Row> rows;
    auto [first, cnt] = part_[p];
    uint64_t tot = 0;
    for (uint64_t c = first; c < first + cnt; c++) tot += chunks_[c].rows;
    rows.reserve(tot);
    for (uint64_t c = first; c < first + cnt; c++) decode_run(chunks_[c], rows);
    if (cnt > 1) {
        std::sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.key < b.key; });
    }
    size_t u = 0;
    for (size_t i = 0; i < rows.size(); i++) {
        if (u && rows[u - 1].key == rows[i].key) {
            uint64_t s = rows[u - 1].count + rows[i].count;
            rows[u - 1].count = s < rows[i].count ? ~0ull : s;
        } else {
            rows[u++] = rows[i];
        }
    }
    rows.resize(u);
    return rows;
}

// ---- codecs and text --------------------------------------------------------
size_t put_varint(uint8_t *o, uint64_t lo, uint64_t hi) {
    size_t n = 0;
    for (;;) {
        uint8_t b = (uint8_t)(lo & 0x7f);
        lo = (lo >> 7) | (hi << 57);
        hi >>= 7;
        if (!lo && !hi) {
            o[n++] = b;
            return n;
        }
        o[n++] = b | 0x80;
    }
}

std::vector<uint8_t> serialize_rows(const std::vector<Row> &rows, int W, uint32_t *codec) {
    std::vector<uint8_t> out;
    if (rows.empty()) return out;
    for (const Row &r : rows)
        if (r.count > 0xFFFFFFFFull) throw Error("merged count exceeds u32 for key " + std::to_string(r.key.w[0]));
    if (W == 4) {
        out.resize(rows.size() * 36);
        for (size_t i = 0; i < rows.size(); i++) {
            memcpy(out.data() + 36 * i, rows[i].key.w, 32);
            uint32_t c = (uint32_t)rows[i].count;
            memcpy(out.data() + 36 * i + 32, &c, 4);
        }
        *codec = CODEC_RAW;
        return out;
    }
    out.resize(8 * W + 4 + (rows.size() - 1) * (W == 1 ? 15 : 24));
    size_t p = 8 * W + 4;
    memcpy(out.data(), rows[0].key.w, 8 * W);
    uint32_t c0 = (uint32_t)rows[0].count;
    memcpy(out.data() + 8 * W, &c0, 4);
    for (size_t i = 1; i < rows.size(); i++) {
        unsigned __int128 a = ((unsigned __int128)rows[i - 1].key.w[1] << 64) | rows[i - 1].key.w[0];
        unsigned __int128 b = ((unsigned __int128)rows[i].key.w[1] << 64) | rows[i].key.w[0];
        unsigned __int128 dlt = b - a;
        if (W == 1) dlt = (uint64_t)(rows[i].key.w[0] - rows[i - 1].key.w[0]);
        p += put_varint(out.data() + p, (uint64_t)dlt, (uint64_t)(dlt >> 64));
        p += put_varint(out.data() + p, rows[i].count, 0);
    }
    out.resize(p);
    *codec = CODEC_DVLE;
    return out;
}

void decode_key(const Key &k, int words, int kk, char *buf) {
    static const char L[4] = {'A', 'C', 'G', 'T'};
    if (words <= 2) {
        unsigned __int128 v = ((unsigned __int128)k.w[1] << 64) | k.w[0];
        if (words == 1) v = k.w[0];
        for (int i = 0; i < kk; i++) buf[i] = L[(uint32_t)(v >> (2 * (kk - 1 - i))) & 3];
        return;
    }
    // 64 letters MSB-first from all 128 bits of hi, then the low 2(k-64) bits of lo
    unsigned __int128 lo = ((unsigned __int128)k.w[1] << 64) | k.w[0];
    unsigned __int128 hi = ((unsigned __int128)k.w[3] << 64) | k.w[2];
    int hb = kk < 64 ? kk : 64, lb = kk - hb;
    unsigned __int128 hs = hb == 64 ? hi : (hi << (128 - 2 * hb));
    for (int i = 0; i < hb; i++) buf[i] = L[(uint32_t)(hs >> (126 - 2 * i)) & 3];
    for (int j = 0; j < lb; j++) buf[hb + j] = L[(uint32_t)(lo >> (2 * (lb - 1 - j))) & 3];
}

std::string infer_label(const std::string &path, size_t index) {
    size_t s = path.find_last_of('/');
    std::string name = s == std::string::npos ? path : path.substr(s + 1);
    static const char *suf[] = {".fastq.gz", ".fq.gz", ".fastq", ".fq", ".kite", ".tsv", ".txt", ".gz"};
    for (const char *x : suf) {
        size_t n = strlen(x);
        if (name.size() >= n && name.compare(name.size() - n, n, x) == 0) {
            name.resize(name.size() - n);
            break;
        }
    }
    if (name.empty()) name = "sample" + std::to_string(index + 1);
    return name;
}

// ---- zstd (dlopen) ------------------------------------------------------------
namespace {
struct Zstd {
    void *h = nullptr;
    size_t (*decompress)(void *, size_t, const void *, size_t) = nullptr;
    size_t (*compress)(void *, size_t, const void *, size_t, int) = nullptr;
    size_t (*bound)(size_t) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    Zstd() {
        h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) return;
        decompress = (decltype(decompress))dlsym(h, "ZSTD_decompress");
        compress = (decltype(compress))dlsym(h, "ZSTD_compress");
        bound = (decltype(bound))dlsym(h, "ZSTD_compressBound");
        is_error = (decltype(is_error))dlsym(h, "ZSTD_isError");
        if (!decompress || !compress || !bound || !is_error) h = nullptr;
    }
};
Zstd &zstd() {
    static Zstd z;
    return z;
}
}  // namespace

bool zstd_available() { return zstd().h != nullptr; }

std::vector<uint8_t> zstd_decompress(const uint8_t *src, size_t len, size_t raw_size) {
    Zstd &z = zstd();
    if (!z.h) throw Error("zstd decompress failed: libzstd.so.1 is not available");
    std::vector<uint8_t> out(raw_size);
    size_t r = z.decompress(out.data(), raw_size, src, len);
    if (z.is_error(r)) throw Error("zstd decompress failed");
    out.resize(r);
    return out;
}

std::vector<uint8_t> zstd_compress(const uint8_t *src, size_t len, int level) {
    Zstd &z = zstd();
    if (!z.h) throw Error("zstd compress failed: libzstd.so.1 is not available");
    std::vector<uint8_t> out(z.bound(len));
    size_t r = z.compress(out.data(), out.size(), src, len, level);
    if (z.is_error(r)) throw Error("zstd compress failed");
    out.resize(r);
    return out;
}

}  // namespace kh
