/*
 * frame_io.cpp -- host side of the decode path: the .frame container.
 *
 * pvvc_load_frame / pvvc_parse_frame replace io::LoadFrame
 * (src/cuda/common/slice_io.cu:143-275).  The layout is the one
 * io::ChangeSliceToFrame writes (src/io/slice_io.cc:258-315):
 *
 *   i32 timestamp; u32 S; i32 index[S]; u8 type[S]; u32 size[S]; u8 qp[S];
 *   f32 mv[S][16] (row-major); u32 geo_size[S]; geo blobs (non-empty only);
 *   u32 color_size[S]; colour blobs (non-empty only).
 *
 * A blob whose slice type has bit 3 (geometry) or bit 4 (colour) set is a
 * zstd frame and is unwrapped on load, as the reference loader does
 * (slice_io.cu:233-240,258-265).  Unlike the reference (one `new[]` per
 * slice), every unwrapped blob lands in one host arena, 16-byte aligned,
 * so that a frame can be staged to the device with a single copy.
 */
#include "pvvc_internal.h"

#include <zstd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <filesystem>
#include <functional>
#include <mutex>
#include <new>
#include <regex>
#include <thread>

namespace {

/* Host worker pool for the zstd unwrap: parallel_for(n, fn) runs fn(i) for
 * i in [0, n) on the workers and the calling thread; several callers may
 * share the pool at once. */
class HostPool {
  public:
    void set_threads(int n) {
        std::lock_guard<std::mutex> lk(cfg_mu_);
        want_ = std::max(1, std::min(n, 64));
    }
    int threads() {
        std::lock_guard<std::mutex> lk(cfg_mu_);
        return want_;
    }
    void parallel_for(size_t n, const std::function<void(size_t)>& fn) {
        ensure_started();
        if (n == 0) return;
        if (n == 1 || threads() <= 1) {
            for (size_t i = 0; i < n; ++i) fn(i);
            return;
        }
        Job job;
        job.fn = &fn;
        job.n = n;
        {
            std::lock_guard<std::mutex> lk(mu_);
            q_.push_back(&job);
        }
        cv_.notify_all();
        run(job);
        /* the job lives on this stack: wait until every index ran and no
         * worker still holds it */
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return job.done.load() == n && job.holders == 0; });
        auto it = std::find(q_.begin(), q_.end(), &job);
        if (it != q_.end()) q_.erase(it);
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }

  private:
    struct Job {
        const std::function<void(size_t)>* fn = nullptr;
        size_t n = 0;
        std::atomic<size_t> next{0}, done{0};
        int holders = 0;   /* workers between taking the job and leaving run(); under mu_ */
    };
    void run(Job& j) {
        size_t i;
        while ((i = j.next.fetch_add(1)) < j.n) {
            (*j.fn)(i);
            if (j.done.fetch_add(1) + 1 == j.n) {
                std::lock_guard<std::mutex> lk(mu_);
                done_cv_.notify_all();
            }
        }
    }
    void worker() {
        for (;;) {
            Job* j = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] {
                    while (!q_.empty() && q_.front()->next.load() >= q_.front()->n) q_.pop_front();
                    return stop_ || !q_.empty();
                });
                if (stop_) return;
                j = q_.front();
                j->holders++;
            }
            run(*j);
            std::lock_guard<std::mutex> lk(mu_);
            j->holders--;
            done_cv_.notify_all();
        }
    }
    void ensure_started() {
        std::lock_guard<std::mutex> lk(cfg_mu_);
        while (static_cast<int>(th_.size()) < want_ - 1) th_.emplace_back([this] { worker(); });
    }
    std::mutex cfg_mu_, mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<Job*> q_;
    std::vector<std::thread> th_;
    int want_ = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
    bool stop_ = false;
};

HostPool& host_pool() {
    static HostPool pool;
    return pool;
}

/* Bounds-checked little-endian reader over a memory image. */
struct Cursor {
    const uint8_t* p;
    size_t n, at = 0;
    bool take(void* dst, size_t bytes) {
        if (n - at < bytes) return false;
        std::memcpy(dst, p + at, bytes);
        at += bytes;
        return true;
    }
    const uint8_t* span(size_t bytes) {
        if (n - at < bytes) return nullptr;
        const uint8_t* r = p + at;
        at += bytes;
        return r;
    }
};

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

/* ZstdDecoder::Decode semantics (src/common/entropy_codec.cc:341-360): the
 * frame header states the size; 0 or unknown is an error. */
bool zstd_size(const uint8_t* src, size_t n, size_t& sz) {
    unsigned long long v = ZSTD_getFrameContentSize(src, n);
    if (v == 0 || v == ZSTD_CONTENTSIZE_UNKNOWN || v == ZSTD_CONTENTSIZE_ERROR) return false;
    if (v > (1ull << 32)) return false;
    sz = static_cast<size_t>(v);
    return true;
}

/* Decompresses into dst[0, cap); returns the size or -1. */
int64_t zstd_into(const uint8_t* src, size_t n, uint8_t* dst, size_t cap) {
    thread_local ZSTD_DCtx* dctx = ZSTD_createDCtx();
    if (!dctx) return -1;
    size_t r = ZSTD_decompressDCtx(dctx, dst, cap, src, n);
    return ZSTD_isError(r) ? -1 : static_cast<int64_t>(r);
}

}  // namespace

int pvvc_frame_parse_image(const uint8_t* bytes, size_t nbytes, pvvc_frame& f) {
    Cursor c{bytes, nbytes};
    if (!c.take(&f.timestamp, 4) || !c.take(&f.slice_cnt, 4)) return 5;
    const uint32_t S = f.slice_cnt;
    /* Refuse a count the image cannot hold before allocating for it. */
    if (static_cast<uint64_t>(S) * (4 + 1 + 4 + 1 + 64 + 4 + 4) > nbytes) return 5;
    f.index.resize(S);
    f.type.resize(S);
    f.size.resize(S);
    f.qp.resize(S);
    f.mv.resize(size_t(S) * 16);
    f.geo_size.resize(S);
    f.col_size.resize(S);
    f.geo_off.assign(S, 0);
    f.col_off.assign(S, 0);
    if (!c.take(f.index.data(), 4 * size_t(S)) || !c.take(f.type.data(), S) || !c.take(f.size.data(), 4 * size_t(S)) ||
        !c.take(f.qp.data(), S) || !c.take(f.mv.data(), 64 * size_t(S)))
        return 5;
    std::vector<uint32_t> raw_geo(S), raw_col(S);
    if (!c.take(raw_geo.data(), 4 * size_t(S))) return 5;
    std::vector<const uint8_t*> gsrc(S, nullptr), csrc(S, nullptr);
    for (uint32_t i = 0; i < S; ++i) {
        if (!raw_geo[i]) continue;
        if (!(gsrc[i] = c.span(raw_geo[i]))) return 5;
    }
    if (!c.take(raw_col.data(), 4 * size_t(S))) return 5;
    for (uint32_t i = 0; i < S; ++i) {
        if (!raw_col[i]) continue;
        if (!(csrc[i] = c.span(raw_col[i]))) return 5;
    }
    /* Unwrap into one arena; sizes become the unwrapped sizes (slice_io.cu:239,264).
     * The zstd frame headers give every blob's size, so the arena is laid out
     * first and each blob is decompressed (in parallel) straight into place. */
    std::vector<uint8_t> gz(S), cz(S);
    size_t total = 0;
    for (uint32_t i = 0; i < S; ++i) {
        size_t gs = raw_geo[i], cs = raw_col[i];
        gz[i] = gsrc[i] && (f.type[i] & PVVC_TYPE_GEO_ZSTD);
        cz[i] = csrc[i] && (f.type[i] & PVVC_TYPE_COLOR_ZSTD);
        if (gz[i] && !zstd_size(gsrc[i], raw_geo[i], gs)) return 6;
        if (cz[i] && !zstd_size(csrc[i], raw_col[i], cs)) return 6;
        f.geo_off[i] = total;
        f.geo_size[i] = static_cast<uint32_t>(gs);
        total += align16(gs);
        f.col_off[i] = total;
        f.col_size[i] = static_cast<uint32_t>(cs);
        total += align16(cs);
    }
    /* a zstd header may claim any size: refuse arenas far beyond the image */
    if (total > (size_t{1} << 30) && total / 1024 > nbytes) return 6;
    try {
        f.data.assign(total + 16, 0);
    } catch (const std::bad_alloc&) {
        return 6;
    }
    std::atomic<int> bad{0};
    auto blob = [&](size_t i) {
        uint8_t* base = f.data.data();
        if (gsrc[i]) {
            if (gz[i]) {
                const int64_t r = zstd_into(gsrc[i], raw_geo[i], base + f.geo_off[i], f.geo_size[i]);
                if (r < 0) bad = 1;
                else f.geo_size[i] = static_cast<uint32_t>(r);
            } else {
                std::memcpy(base + f.geo_off[i], gsrc[i], f.geo_size[i]);
            }
        }
        if (csrc[i]) {
            if (cz[i]) {
                const int64_t r = zstd_into(csrc[i], raw_col[i], base + f.col_off[i], f.col_size[i]);
                if (r < 0) bad = 1;
                else f.col_size[i] = static_cast<uint32_t>(r);
            } else {
                std::memcpy(base + f.col_off[i], csrc[i], f.col_size[i]);
            }
        }
    };
    /* tasks of 64 slices: a blob is ~1 KB, too little to hand out alone */
    constexpr size_t kChunk = 64;
    host_pool().parallel_for((S + kChunk - 1) / kChunk, [&](size_t t) {
        const size_t hi = std::min<size_t>(S, (t + 1) * kChunk);
        for (size_t i = t * kChunk; i < hi; ++i) blob(i);
    });
    if (bad.load()) return 6;
    f.points = 0;
    for (uint32_t i = 0; i < S; ++i) f.points += f.size[i];
    return 0;
}

extern "C" {

int pvvc_parse_frame(const uint8_t* bytes, size_t nbytes, pvvc_frame** out) {
    if (!out || (!bytes && nbytes)) return PVVC_ERR_INVALID_ARG;
    auto* f = new pvvc_frame();
    int rc = pvvc_frame_parse_image(bytes, nbytes, *f);
    if (rc) {
        delete f;
        *out = nullptr;
        return rc;
    }
    *out = f;
    return 0;
}

int pvvc_load_frame(const char* path, pvvc_frame** out) {
    if (!out) return PVVC_ERR_INVALID_ARG;
    *out = nullptr;
    /* slice_io.cu:144-148: the name must end in ".frame" */
    static const std::regex name_type{"^.*\\.frame"};
    if (!path || !std::regex_match(path, name_type)) return 1;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        switch (errno) {
            case ENOENT: return 2;
            case EACCES: return 3;
            default: return 4;
        }
    }
    std::vector<uint8_t> img;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) img.insert(img.end(), buf, buf + got);
    std::fclose(fp);
    return pvvc_parse_frame(img.data(), img.size(), out);
}

void pvvc_frame_free(pvvc_frame* f) { delete f; }

int pvvc_set_host_threads(int n) {
    if (n <= 0) return PVVC_ERR_INVALID_ARG;
    host_pool().set_threads(n);
    return PVVC_OK;
}
int pvvc_frame_timestamp(const pvvc_frame* f) { return f ? f->timestamp : -1; }
uint32_t pvvc_frame_slice_cnt(const pvvc_frame* f) { return f ? f->slice_cnt : 0; }
uint64_t pvvc_frame_points(const pvvc_frame* f) { return f ? f->points : 0; }

void pvvc_frame_meta(const pvvc_frame* f, int* index, uint8_t* type, uint32_t* size, uint8_t* qp, float* mv,
                     uint32_t* geometry_size, uint32_t* color_size) {
    if (!f) return;
    const uint32_t S = f->slice_cnt;
    if (index) std::memcpy(index, f->index.data(), 4 * size_t(S));
    if (type) std::memcpy(type, f->type.data(), S);
    if (size) std::memcpy(size, f->size.data(), 4 * size_t(S));
    if (qp) std::memcpy(qp, f->qp.data(), S);
    if (mv) std::memcpy(mv, f->mv.data(), 64 * size_t(S));
    if (geometry_size) std::memcpy(geometry_size, f->geo_size.data(), 4 * size_t(S));
    if (color_size) std::memcpy(color_size, f->col_size.data(), 4 * size_t(S));
}

const uint8_t* pvvc_frame_geometry(const pvvc_frame* f, uint32_t s) {
    if (!f || s >= f->slice_cnt || !f->geo_size[s]) return nullptr;
    return f->data.data() + f->geo_off[s];
}

const uint8_t* pvvc_frame_color(const pvvc_frame* f, uint32_t s) {
    if (!f || s >= f->slice_cnt || !f->col_size[s]) return nullptr;
    return f->data.data() + f->col_off[s];
}

/* ------------------------------------------------------------------------
 * .slice files: one patch slice each (io::SaveSlice / io::LoadSlice,
 * src/io/slice_io.cc:20-206), and their assembly into a frame
 * (io::ChangeSliceToFrame, :208-315):
 *
 *   u8 type; i32 timestamp; i32 index; u64 size; u8 qp;
 *   f32 mv[16] (Eigen::Matrix4f storage: column-major);
 *   if not predictive: u64 geometry size (> 0), geometry bytes;
 *   if not skip:       u64 colour size (> 0), colour bytes.
 *
 * Bytes after the colour blob are ignored (the reference warns).  Slices
 * are ordered by patch index (a stable order among equal indices, which the
 * reference leaves unspecified); the frame's timestamp is the first ordered
 * slice's; sizes are written as u32; mv is written row-major.
 * ------------------------------------------------------------------------ */
}  // extern "C"

namespace {

struct SliceRec {
    uint8_t type = 0;
    int timestamp = -1, index = -1;
    uint64_t size = 0;
    uint8_t qp = 1;
    float mv_cm[16] = {};  /* column-major, as stored */
    bool has_geo = false, has_col = false;
    std::vector<uint8_t> geo, col;
};

int open_code(int e) { return e == ENOENT ? 2 : e == EACCES ? 3 : 4; }

/* LoadSlice (slice_io.cc:124-206) with its exceptions as codes. */
int read_slice(const char* path, SliceRec& r) {
    static const std::regex name_type{"^.*\\.slice"};
    if (!path || !std::regex_match(path, name_type)) return 1;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return open_code(errno);
    /* a blob longer than the bytes left is a short read, found before any allocation */
    uint64_t fsize = 0;
    if (std::fseek(fp, 0, SEEK_END) == 0) {
        const long e = std::ftell(fp);
        fsize = e > 0 ? static_cast<uint64_t>(e) : 0;
    }
    std::rewind(fp);
    auto left = [&]() -> uint64_t {
        const long at = std::ftell(fp);
        return at >= 0 && static_cast<uint64_t>(at) <= fsize ? fsize - static_cast<uint64_t>(at) : 0;
    };
    auto rd = [&](void* p, size_t n) { return std::fread(p, 1, n, fp) == n; };
    int rc = 0;
    if (!rd(&r.type, 1)) rc = 5;
    else if (!(r.type & 0x01)) rc = 7;
    else if (!rd(&r.timestamp, 4) || !rd(&r.index, 4) || !rd(&r.size, 8) || !rd(&r.qp, 1) || !rd(r.mv_cm, 64)) rc = 5;
    if (!rc && !(r.type & 0x02)) {
        uint64_t g = 0;
        if (!rd(&g, 8)) rc = 5;
        else if (!g) rc = 8;
        else if (g > left()) rc = 5;
        else {
            r.geo.resize(g);
            if (!rd(r.geo.data(), g)) rc = 5;
            r.has_geo = !rc;
        }
    }
    if (!rc && !(r.type & 0x04)) {
        uint64_t c = 0;
        if (!rd(&c, 8)) rc = 5;
        else if (!c) rc = 8;
        else if (c > left()) rc = 5;
        else {
            r.col.resize(c);
            if (!rd(r.col.data(), c)) rc = 5;
            r.has_col = !rc;
        }
    }
    std::fclose(fp);
    return rc;
}

template <class T>
void put(std::vector<uint8_t>& o, const T& v) {
    const auto* p = reinterpret_cast<const uint8_t*>(&v);
    o.insert(o.end(), p, p + sizeof(T));
}

/* The .frame image ChangeSliceToFrame writes for these slices
 * (slice_io.cc:238-310, common.h:378-412). */
int frame_image(std::vector<SliceRec>& sl, std::vector<uint8_t>& img) {
    if (sl.empty()) return 9;
    std::stable_sort(sl.begin(), sl.end(), [](const SliceRec& a, const SliceRec& b) { return a.index < b.index; });
    const uint32_t S = static_cast<uint32_t>(sl.size());
    img.clear();
    put(img, sl.front().timestamp);
    put(img, S);
    for (auto& r : sl) put(img, r.index);
    for (auto& r : sl) put(img, r.type);
    for (auto& r : sl) put(img, static_cast<uint32_t>(r.size));
    for (auto& r : sl) put(img, r.qp);
    for (auto& r : sl)
        for (int row = 0; row < 4; ++row)
            for (int col = 0; col < 4; ++col) put(img, r.mv_cm[4 * col + row]);
    for (auto& r : sl) put(img, static_cast<uint32_t>(r.has_geo ? r.geo.size() : 0));
    for (auto& r : sl)
        if (r.has_geo) img.insert(img.end(), r.geo.begin(), r.geo.end());
    for (auto& r : sl) put(img, static_cast<uint32_t>(r.has_col ? r.col.size() : 0));
    for (auto& r : sl)
        if (r.has_col) img.insert(img.end(), r.col.begin(), r.col.end());
    return 0;
}

int read_slices(const char* const* paths, int n, std::vector<SliceRec>& sl) {
    if (n <= 0 || !paths) return 9;
    sl.assign(static_cast<size_t>(n), SliceRec{});
    for (int i = 0; i < n; ++i)
        if (int rc = read_slice(paths[i], sl[i])) return rc;
    return 0;
}

int write_file(const char* path, const std::vector<uint8_t>& img) {
    static const std::regex out_type{"^.*\\.frame"};
    if (!path || !std::regex_match(path, out_type)) return 1;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return open_code(errno);
    const bool ok = img.empty() || std::fwrite(img.data(), 1, img.size(), fp) == img.size();
    return (std::fclose(fp) == 0 && ok) ? 0 : 10;
}

}  // namespace

extern "C" {

int64_t pvvc_save_slice(const char* path, const pvvc_slice_desc* s) {
    static const std::regex name_type{"^.*\\.slice"};
    if (!path || !s || !std::regex_match(path, name_type)) return -1;
    if (!(s->type & 0x01)) return -7;
    const bool geo = !(s->type & 0x02), col = !(s->type & 0x04);
    if ((geo && (!s->geometry || !s->geometry_size)) || (col && (!s->color || !s->color_size))) return -8;
    std::vector<uint8_t> o;
    put(o, s->type);
    put(o, s->timestamp);
    put(o, s->index);
    put(o, s->size);
    put(o, s->qp);
    for (int col_i = 0; col_i < 4; ++col_i)
        for (int row = 0; row < 4; ++row) put(o, s->mv[4 * row + col_i]); /* stored column-major */
    if (geo) {
        put(o, s->geometry_size);
        o.insert(o.end(), s->geometry, s->geometry + s->geometry_size);
    }
    if (col) {
        put(o, s->color_size);
        o.insert(o.end(), s->color, s->color + s->color_size);
    }
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return -open_code(errno);
    const bool ok = std::fwrite(o.data(), 1, o.size(), fp) == o.size();
    if (std::fclose(fp) != 0 || !ok) return -10;
    return static_cast<int64_t>(o.size());
}

int pvvc_load_slices(const char* const* paths, int n, pvvc_frame** out) {
    if (!out) return PVVC_ERR_INVALID_ARG;
    *out = nullptr;
    std::vector<SliceRec> sl;
    std::vector<uint8_t> img;
    int rc = read_slices(paths, n, sl);
    if (!rc) rc = frame_image(sl, img);
    if (rc) return rc;
    return pvvc_parse_frame(img.data(), img.size(), out);
}

int pvvc_slices_to_frame_file(const char* const* paths, int n, const char* frame_path) {
    std::vector<SliceRec> sl;
    std::vector<uint8_t> img;
    int rc = read_slices(paths, n, sl);
    if (!rc) rc = frame_image(sl, img);
    return rc ? rc : write_file(frame_path, img);
}

int pvvc_change_slice_to_frame(const char* input_dir, const char* frame_path) {
    static const std::regex out_type{"^.*\\.frame"};
    if (!frame_path || !std::regex_match(frame_path, out_type)) return 1;
    std::error_code ec;
    if (!input_dir || !std::filesystem::is_directory(input_dir, ec)) return 4;
    static const std::regex slice_name{"^.*_time_[0-9]+_slice_[0-9]+\\.slice$"};
    std::vector<std::string> names;
    for (const auto& e : std::filesystem::directory_iterator(input_dir, ec)) {
        const std::string nm = e.path().string();
        if (std::regex_match(nm, slice_name)) names.push_back(nm);
    }
    if (ec) return 4;
    if (names.empty()) return 2;
    std::sort(names.begin(), names.end()); /* a fixed order before the sort by index */
    std::vector<const char*> p;
    for (auto& nm : names) p.push_back(nm.c_str());
    return pvvc_slices_to_frame_file(p.data(), static_cast<int>(p.size()), frame_path);
}

}  // extern "C"
