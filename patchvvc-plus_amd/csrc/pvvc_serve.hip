/*
 * pvvc_serve.hip -- multi-stream serving loop of the per-patch GoP decoder.
 *
 * The reference client decodes one stream: a loader thread fills a ring of
 * frames while a decoder thread calls launch_cudaProcess once per frame and
 * synchronises (src/cuda/common/manager.cc:145-310).  This server decodes
 * many streams at once with continuous batching:
 *
 *  - streams are admitted by HBM footprint (their patch state slots plus an
 *    output ring); a stream that does not fit waits, first in first out,
 *    and is admitted when closed streams free memory;
 *  - each stream's GoP state (intra geometry, reference colours) stays
 *    resident in HBM across its P-frames;
 *  - a scheduler thread takes every frame that has arrived, from every
 *    stream with no frame in flight and a free output slot, into one batch,
 *    and launches it (one H2D copy of the packed batch, then the persistent
 *    decode) on a free lane; each lane has its own HIP stream, batch buffers,
 *    queue counter, slab and stage, so batches overlap on the device;
 *  - a completion thread per lane waits for its batch and publishes
 *    per-frame completions with arrival, launch and completion times.
 *
 * A stream never has frames in two batches at once, so batches carry no
 * dependencies between them; a stream's frames decode in submission order.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "pvvc_runtime.h"

using namespace pvvc;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

uint64_t al256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

/* First-fit allocator over [0, size). */
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;  /* start -> length */
    void init(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
    }
    int64_t alloc(int64_t len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t at = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest) free_[at + len] = rest;
                return at;
            }
        }
        return -1;
    }
    void release(int64_t at, int64_t len) {
        auto it = free_.emplace(at, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    const pvvc_frame* frame;
    double arrival;
    uint64_t tag;
    int64_t seq;
};

enum { ST_PENDING = 0, ST_ACTIVE = 1, ST_CLOSING = 2, ST_CLOSED = 3 };

struct StreamRec {
    int id = 0;
    int state = ST_PENDING;
    int max_patches = 0, max_pts_patch = 0;
    int64_t max_pts_frame = 0;
    int ring = 0;
    uint32_t geo_cap = 0;
    int slot_base = -1;
    int64_t ring_base = -1;
    uint8_t* mem = nullptr;        /* geometry + reference colours of its slots */
    uint64_t footprint = 0;
    std::vector<char> ring_busy;
    std::deque<Req> pending;
    int inflight = 0;
    int64_t next_seq = 0;
    int64_t frames_done = 0;
};

struct FrameRec {
    int stream;
    int64_t seq;
    uint64_t tag;
    double arrival;
    int ring_slot;
    int64_t points;
    int64_t out_point;             /* first point in the output pool */
    uint32_t first_item, n_items;
};

struct Lane {
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    uint8_t* dev = nullptr;        /* batch image (device) */
    uint8_t* host = nullptr;       /* batch image (pinned host) */
    size_t cap = 0;
    int* status_host = nullptr;    /* pinned */
    size_t status_cap = 0;
    uint8_t* slab = nullptr;
    uint64_t slab_bytes = 0;       /* per workgroup */
    uint64_t slab_total = 0;
    float* stage = nullptr;
    uint32_t* queue = nullptr;
    bool busy = false;
    bool has_work = false;
    double launch_t = 0;
    std::vector<FrameRec> frames;
    std::thread reaper;
    std::condition_variable cv;
};

}  // namespace

struct pvvc_server {
    pvvc_server_config cfg{};
    uint64_t budget = 0, used = 0;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::map<int, StreamRec> streams;
    std::deque<int> admit_q;
    int next_id = 1;
    RangeAlloc slot_free, pool_free;
    SlotDev* slots = nullptr;
    float* pool_xyz = nullptr;
    uint8_t* pool_rgb = nullptr;
    std::vector<Lane> lanes;
    std::deque<pvvc_completion> done;
    std::thread sched;
    bool stop = false;
    uint32_t launch_id = 0;
    size_t rr = 0;
    int64_t frames_done = 0, points_done = 0, batches = 0;
    double host_s = 0;
    std::vector<double> lat_ms;
    double first_arrival = -1, last_done = 0;
};

namespace {

bool ok(hipError_t e) { return rt_hip_ok(e); }

uint64_t stream_state_bytes(int max_patches, int max_pts_patch, uint32_t geo_cap) {
    return static_cast<uint64_t>(max_patches) * (al256(geo_cap) + al256(12ull * max_pts_patch));
}

/* Admits stream `st` if its footprint and ranges fit; caller holds mu. */
bool try_admit(pvvc_server* sv, StreamRec& st) {
    if (sv->used + st.footprint > sv->budget) return false;
    const int64_t sb = sv->slot_free.alloc(st.max_patches);
    if (sb < 0) return false;
    const int64_t ring_pts = st.max_pts_frame * st.ring;
    const int64_t rb = sv->pool_free.alloc(ring_pts);
    if (rb < 0) {
        sv->slot_free.release(sb, st.max_patches);
        return false;
    }
    const uint64_t sbytes = stream_state_bytes(st.max_patches, st.max_pts_patch, st.geo_cap);
    uint8_t* mem = nullptr;
    if (!ok(hipMalloc(&mem, std::max<uint64_t>(sbytes, 256)))) {
        sv->slot_free.release(sb, st.max_patches);
        sv->pool_free.release(rb, ring_pts);
        return false;
    }
    std::vector<SlotDev> h(st.max_patches);
    const uint64_t stride = al256(st.geo_cap) + al256(12ull * st.max_pts_patch);
    for (int i = 0; i < st.max_patches; ++i) {
        SlotDev& s = h[i];
        std::memset(&s, 0, sizeof(s));
        uint8_t* b = mem + stride * i;
        s.geo = b;
        s.ref = reinterpret_cast<float*>(b + al256(st.geo_cap));
        s.ws = nullptr;
        s.ws_cap = 0;
        s.geo_cap = st.geo_cap;
        s.n_cap = static_cast<uint32_t>(st.max_pts_patch);
    }
    /* no batch in flight uses this slot range: its previous owner retired */
    if (!ok(hipMemcpy(sv->slots + sb, h.data(), sizeof(SlotDev) * st.max_patches, hipMemcpyHostToDevice))) {
        hipFree(mem);
        sv->slot_free.release(sb, st.max_patches);
        sv->pool_free.release(rb, ring_pts);
        return false;
    }
    st.mem = mem;
    st.slot_base = static_cast<int>(sb);
    st.ring_base = rb;
    st.ring_busy.assign(st.ring, 0);
    st.state = ST_ACTIVE;
    sv->used += st.footprint;
    return true;
}

/* Admits waiting streams in arrival order while they fit, and wakes the
 * scheduler for the frames they queued while waiting; caller holds mu. */
void admit_waiting(pvvc_server* sv) {
    bool admitted = false;
    while (!sv->admit_q.empty()) {
        auto it = sv->streams.find(sv->admit_q.front());
        if (it == sv->streams.end() || it->second.state != ST_PENDING) {
            sv->admit_q.pop_front();
            continue;
        }
        if (!try_admit(sv, it->second)) break;
        sv->admit_q.pop_front();
        admitted = true;
    }
    if (admitted) sv->cv_sched.notify_one();
}

/* Releases a closing stream once nothing of it is queued or in flight; caller holds mu. */
void maybe_retire(pvvc_server* sv, StreamRec& st) {
    if (st.state != ST_CLOSING || st.inflight || !st.pending.empty()) return;
    if (st.mem) hipFree(st.mem);
    st.mem = nullptr;
    if (st.slot_base >= 0) sv->slot_free.release(st.slot_base, st.max_patches);
    if (st.ring_base >= 0) sv->pool_free.release(st.ring_base, st.max_pts_frame * st.ring);
    st.slot_base = -1;
    st.ring_base = -1;
    sv->used -= st.footprint;
    st.state = ST_CLOSED;
    admit_waiting(sv);
}

int free_lane(pvvc_server* sv) {
    for (size_t i = 0; i < sv->lanes.size(); ++i)
        if (!sv->lanes[i].busy) return static_cast<int>(i);
    return -1;
}

bool stream_ready(const StreamRec& st) {
    if (st.state != ST_ACTIVE && st.state != ST_CLOSING) return false;
    if (st.inflight || st.pending.empty()) return false;
    for (char b : st.ring_busy)
        if (!b) return true;
    return false;
}

bool any_ready(pvvc_server* sv) {
    for (auto& kv : sv->streams)
        if (stream_ready(kv.second)) return true;
    return false;
}

bool grow_lane(Lane& ln, size_t bytes, size_t n_items) {
    if (bytes > ln.cap) {
        const size_t cap = std::max(bytes, ln.cap + ln.cap / 2);
        if (ln.dev) hipFree(ln.dev);
        if (ln.host) hipHostFree(ln.host);
        ln.dev = nullptr;
        ln.host = nullptr;
        ln.cap = 0;
        if (!ok(hipMalloc(&ln.dev, cap))) return false;
        if (!ok(hipHostMalloc(reinterpret_cast<void**>(&ln.host), cap, hipHostMallocDefault))) return false;
        ln.cap = cap;
    }
    if (n_items > ln.status_cap) {
        const size_t cap = std::max(n_items, ln.status_cap * 2);
        if (ln.status_host) hipHostFree(ln.status_host);
        ln.status_host = nullptr;
        ln.status_cap = 0;
        if (!ok(hipHostMalloc(reinterpret_cast<void**>(&ln.status_host), 4 * cap, hipHostMallocDefault)))
            return false;
        ln.status_cap = cap;
    }
    return true;
}

/* Publishes completions of frames that never reached the device; caller holds mu. */
void fail_frames(pvvc_server* sv, const std::vector<FrameRec>& frs, int status) {
    const double t = now_s();
    for (const FrameRec& fr : frs) {
        auto it = sv->streams.find(fr.stream);
        if (it == sv->streams.end()) continue;
        StreamRec& st = it->second;
        pvvc_completion c{};
        c.stream_id = fr.stream;
        c.tag = fr.tag;
        c.seq = fr.seq;
        c.status = status;
        c.points = fr.points;
        c.out_slot = fr.ring_slot;
        c.arrival_s = fr.arrival;
        c.launch_s = t;
        c.done_s = t;
        sv->done.push_back(c);
        st.inflight--;
        maybe_retire(sv, st);
    }
    sv->cv_done.notify_all();
}

void scheduler(pvvc_server* sv) {
    rt_init_kernels();
    for (;;) {
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_sched.wait(lk, [&] { return sv->stop || (free_lane(sv) >= 0 && any_ready(sv)); });
        if (sv->stop) return;
        const int li = free_lane(sv);
        Lane& ln = sv->lanes[li];
        /* take every arrived frame of every ready stream, round robin, within the caps */
        std::vector<FrameRec> frs;
        std::vector<const pvvc_frame*> frames;
        std::vector<int> sbase, scount;
        std::vector<uint32_t> gcap;
        int64_t pts = 0;
        const size_t ns = sv->streams.size();
        std::vector<int> ids;
        ids.reserve(ns);
        for (auto& kv : sv->streams) ids.push_back(kv.first);
        const size_t start = ns ? sv->rr % ns : 0;
        sv->rr++;
        for (size_t k = 0; k < ns; ++k) {
            StreamRec& st = sv->streams[ids[(start + k) % ns]];
            if (!stream_ready(st)) continue;
            while (!st.pending.empty()) {
                const Req& r = st.pending.front();
                const int64_t fp = static_cast<int64_t>(r.frame->points);
                if (!frs.empty() && pts + fp > sv->cfg.max_batch_points) break;
                int slot = -1;
                for (int j = 0; j < st.ring; ++j)
                    if (!st.ring_busy[j]) {
                        slot = j;
                        break;
                    }
                if (slot < 0) break;
                st.ring_busy[slot] = 1;
                st.inflight++;
                FrameRec fr{st.id, r.seq, r.tag, r.arrival, slot, fp, st.ring_base + slot * st.max_pts_frame, 0, 0};
                frs.push_back(fr);
                frames.push_back(r.frame);
                sbase.push_back(st.slot_base);
                scount.push_back(st.max_patches);
                gcap.push_back(st.geo_cap);
                pts += fp;
                st.pending.pop_front();
            }
            if (pts >= sv->cfg.max_batch_points) break;
        }
        if (frs.empty()) continue;
        const double t_host0 = now_s();
        ln.busy = true;
        const pvvc_tuning tun = sv->cfg.tuning;
        lk.unlock();

        BatchPlan plan;
        int rc = plan_batch(frames.data(), static_cast<int>(frames.size()), sbase.data(), scount.data(),
                            gcap.data(), nullptr, plan);
        if (rc == PVVC_OK && !grow_lane(ln, plan.total, plan.n_items)) rc = PVVC_ERR_HIP;
        bool fits = true;
        for (const Segment& sg : plan.segs) fits &= sg.ws_need <= ln.slab_bytes;
        if (rc == PVVC_OK && !fits) rc = PVVC_ERR_CAPACITY;
        if (rc != PVVC_OK) {
            lk.lock();
            ln.busy = false;
            fail_frames(sv, frs, rc);
            sv->cv_sched.notify_one();
            continue;
        }
        /* outputs go to each frame's ring slot in the output pool */
        for (size_t f = 0; f < frs.size(); ++f) {
            frs[f].first_item = plan.frame_first[f];
            frs[f].n_items = plan.frame_count[f];
            const int64_t delta = frs[f].out_point - plan.frame_off[f];
            for (uint32_t i = 0; i < plan.frame_count[f]; ++i)
                plan.items[plan.frame_first[f] + i].out_off += static_cast<uint64_t>(delta);
        }
        fill_batch_image(plan, frames.data(), ln.dev, ln.host);
        bool good = ok(hipMemcpyAsync(ln.dev, ln.host, plan.total, hipMemcpyHostToDevice, ln.s));
        const BatchDev bd{reinterpret_cast<const ItemDev*>(ln.dev + plan.o_items),
                          reinterpret_cast<const uint32_t*>(ln.dev + plan.o_order),
                          reinterpret_cast<const uint32_t*>(ln.dev + plan.o_defer),
                          reinterpret_cast<int*>(ln.dev + plan.o_status)};
        const LaunchRes res{sv->slots, ln.slab, ln.slab_bytes, ln.slab_total, ln.stage, ln.queue, 0u};
        const Sink sink{sv->pool_xyz, sv->pool_rgb, nullptr, PVVC_OUT_NATIVE, 0};
        rt_timing_reset();
        for (const Segment& sg : plan.segs) {
            uint32_t id;
            {
                std::lock_guard<std::mutex> g(sv->mu);
                id = ++sv->launch_id;
            }
            if (launch_segment(bd, sg, sink, tun, res, id, ln.s) != PVVC_OK) good = false;
        }
        good &= ok(hipGetLastError());
        good &= ok(hipMemcpyAsync(ln.status_host, ln.dev + plan.o_status, 4ull * plan.n_items, hipMemcpyDeviceToHost,
                                  ln.s));
        good &= ok(hipEventRecord(ln.done, ln.s));
        lk.lock();
        if (!good) {
            /* nothing reliable was enqueued: wait for the lane, then fail the frames */
            hipStreamSynchronize(ln.s);
            ln.busy = false;
            fail_frames(sv, frs, PVVC_ERR_HIP);
            sv->cv_sched.notify_one();
            continue;
        }
        ln.frames = std::move(frs);
        ln.launch_t = now_s();
        ln.has_work = true;
        sv->batches++;
        sv->host_s += ln.launch_t - t_host0;
        ln.cv.notify_one();
    }
}

void reaper(pvvc_server* sv, int li) {
    Lane& ln = sv->lanes[li];
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            ln.cv.wait(lk, [&] { return ln.has_work || sv->stop; });
            if (!ln.has_work) return;
        }
        const bool good = ok(hipEventSynchronize(ln.done));
        const double t = now_s();
        std::lock_guard<std::mutex> lk(sv->mu);
        for (const FrameRec& fr : ln.frames) {
            int status = good ? PVVC_OK : PVVC_ERR_HIP;
            for (uint32_t i = 0; good && i < fr.n_items; ++i)
                if (ln.status_host[fr.first_item + i] != PVVC_OK) {
                    status = ln.status_host[fr.first_item + i];
                    break;
                }
            pvvc_completion c{};
            c.stream_id = fr.stream;
            c.tag = fr.tag;
            c.seq = fr.seq;
            c.status = status;
            c.points = fr.points;
            c.xyz = sv->pool_xyz + 3 * fr.out_point;
            c.rgb = sv->pool_rgb + 3 * fr.out_point;
            c.out_slot = fr.ring_slot;
            c.arrival_s = fr.arrival;
            c.launch_s = ln.launch_t;
            c.done_s = t;
            sv->done.push_back(c);
            sv->frames_done++;
            sv->points_done += fr.points;
            sv->lat_ms.push_back((t - fr.arrival) * 1e3);
            if (sv->first_arrival < 0 || fr.arrival < sv->first_arrival) sv->first_arrival = fr.arrival;
            sv->last_done = std::max(sv->last_done, t);
            auto it = sv->streams.find(fr.stream);
            if (it != sv->streams.end()) {
                it->second.inflight--;
                it->second.frames_done++;
                maybe_retire(sv, it->second);
            }
        }
        ln.frames.clear();
        ln.has_work = false;
        ln.busy = false;
        sv->cv_sched.notify_one();
        sv->cv_done.notify_all();
    }
}

void destroy_lanes(pvvc_server* sv) {
    for (Lane& ln : sv->lanes) {
        if (ln.s) hipStreamSynchronize(ln.s);
        if (ln.dev) hipFree(ln.dev);
        if (ln.host) hipHostFree(ln.host);
        if (ln.status_host) hipHostFree(ln.status_host);
        if (ln.slab) hipFree(ln.slab);
        if (ln.stage) hipFree(ln.stage);
        if (ln.queue) hipFree(ln.queue);
        if (ln.done) hipEventDestroy(ln.done);
        if (ln.s) hipStreamDestroy(ln.s);
    }
}

}  // namespace

extern "C" {

double pvvc_server_clock(void) { return now_s(); }

int pvvc_server_create(const pvvc_server_config* cfg_in, pvvc_server** out) {
    if (!out) return PVVC_ERR_INVALID_ARG;
    *out = nullptr;
    pvvc_server_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    if (cfg.lanes <= 0) cfg.lanes = 3;
    if (cfg.max_slots <= 0) cfg.max_slots = 1 << 18;
    if (cfg.max_points_per_patch <= 0) cfg.max_points_per_patch = 4096;
    if (cfg.out_pool_points <= 0) cfg.out_pool_points = int64_t{256} << 20;
    if (cfg.max_batch_points <= 0) cfg.max_batch_points = int64_t{16} << 20;
    if (cfg.tuning.waves_per_patch <= 0) cfg.tuning.waves_per_patch = 1;
    if (cfg.tuning.lds_bytes <= 0) cfg.tuning.lds_bytes = 8 * 1024;
    if (cfg.tuning.wgs_per_cu <= 0) cfg.tuning.wgs_per_cu = 16;
    if (cfg.tuning.queue_chunk <= 0) cfg.tuning.queue_chunk = 1;
    if (cfg.lanes > 8 || cfg.max_batch_points >= (int64_t{1} << 31)) return PVVC_ERR_INVALID_ARG;
    auto* sv = new pvvc_server();
    sv->cfg = cfg;
    sv->slot_free.init(cfg.max_slots);
    sv->pool_free.init(cfg.out_pool_points);
    bool good = ok(hipMalloc(&sv->slots, sizeof(SlotDev) * cfg.max_slots)) &&
                ok(hipMalloc(&sv->pool_xyz, 12ull * cfg.out_pool_points)) &&
                ok(hipMalloc(&sv->pool_rgb, 3ull * cfg.out_pool_points));
    const uint32_t max_geo = static_cast<uint32_t>(64 + 8ull * cfg.max_points_per_patch + 25);
    const uint64_t need = al256(item_ws_bound(static_cast<uint32_t>(cfg.max_points_per_patch), max_geo));
    const uint32_t grid = persistent_grid(cfg.tuning, 0xffffffffu);
    sv->lanes = std::vector<Lane>(cfg.lanes);
    for (Lane& ln : sv->lanes) {
        if (!good) break;
        good = ok(hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking)) &&
               ok(hipEventCreateWithFlags(&ln.done, sv->cfg.lanes_spin ? hipEventDisableTiming
                                                                         : (hipEventDisableTiming | hipEventBlockingSync))) &&
               ok(hipMalloc(&ln.slab, need * grid)) && ok(hipMalloc(&ln.stage, 12ull * cfg.max_batch_points)) &&
               ok(hipMalloc(&ln.queue, 64));
        ln.slab_bytes = need;
        ln.slab_total = need * grid;
    }
    /* first launches load the code object and first-use buffers grow: pay
     * both here rather than in the first frames' latency */
    for (Lane& ln : sv->lanes) {
        if (!good) break;
        good = rt_warm_kernels(cfg.tuning, ln.queue, ln.s) && grow_lane(ln, size_t{16} << 20, size_t{1} << 16);
    }
    if (good) {
        size_t free_b = 0, total_b = 0;
        good = ok(hipMemGetInfo(&free_b, &total_b));
        sv->budget = cfg.hbm_budget ? cfg.hbm_budget : static_cast<uint64_t>(0.8 * static_cast<double>(free_b));
    }
    if (!good) {
        destroy_lanes(sv);
        if (sv->slots) hipFree(sv->slots);
        if (sv->pool_xyz) hipFree(sv->pool_xyz);
        if (sv->pool_rgb) hipFree(sv->pool_rgb);
        delete sv;
        return PVVC_ERR_HIP;
    }
    sv->sched = std::thread(scheduler, sv);
    for (int i = 0; i < cfg.lanes; ++i) sv->lanes[i].reaper = std::thread(reaper, sv, i);
    *out = sv;
    return PVVC_OK;
}

int pvvc_server_destroy(pvvc_server* sv) {
    if (!sv) return PVVC_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_sched.notify_all();
    if (sv->sched.joinable()) sv->sched.join();
    for (Lane& ln : sv->lanes) {
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            ln.cv.notify_all();
        }
        if (ln.reaper.joinable()) ln.reaper.join();
    }
    destroy_lanes(sv);
    for (auto& kv : sv->streams)
        if (kv.second.mem) hipFree(kv.second.mem);
    hipFree(sv->slots);
    hipFree(sv->pool_xyz);
    hipFree(sv->pool_rgb);
    delete sv;
    return PVVC_OK;
}

int pvvc_server_open_stream(pvvc_server* sv, int max_patches, int max_points_per_patch, int64_t max_points_per_frame,
                            int out_ring, int* stream_id) {
    if (!sv || !stream_id || max_patches <= 0 || max_points_per_patch <= 0 || max_points_per_frame <= 0)
        return PVVC_ERR_INVALID_ARG;
    if (max_points_per_patch > sv->cfg.max_points_per_patch || max_patches > sv->cfg.max_slots)
        return PVVC_ERR_CAPACITY;
    if (max_points_per_frame > sv->cfg.max_batch_points) return PVVC_ERR_CAPACITY;
    out_ring = out_ring <= 0 ? 2 : std::min(out_ring, 64);
    if (max_points_per_frame * out_ring > sv->cfg.out_pool_points) return PVVC_ERR_CAPACITY;
    std::lock_guard<std::mutex> lk(sv->mu);
    StreamRec st;
    st.id = sv->next_id++;
    st.max_patches = max_patches;
    st.max_pts_patch = max_points_per_patch;
    st.max_pts_frame = max_points_per_frame;
    st.ring = out_ring;
    st.geo_cap = static_cast<uint32_t>(64 + 8ull * max_points_per_patch + 25);
    st.footprint = stream_state_bytes(max_patches, max_points_per_patch, st.geo_cap) +
                   15ull * static_cast<uint64_t>(max_points_per_frame) * out_ring;
    if (st.footprint > sv->budget) return PVVC_ERR_CAPACITY;
    const int id = st.id;
    StreamRec& rec = sv->streams.emplace(id, std::move(st)).first->second;
    if (!(sv->admit_q.empty() && try_admit(sv, rec))) sv->admit_q.push_back(id);
    *stream_id = id;
    return PVVC_OK;
}

int pvvc_server_stream_state(pvvc_server* sv, int stream_id) {
    if (!sv) return -1;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->streams.find(stream_id);
    return it == sv->streams.end() ? -1 : it->second.state;
}

int pvvc_server_submit(pvvc_server* sv, int stream_id, const pvvc_frame* frame, double arrival_s, uint64_t tag) {
    if (!sv || !frame) return PVVC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->streams.find(stream_id);
    if (it == sv->streams.end()) return PVVC_ERR_INVALID_ARG;
    StreamRec& st = it->second;
    if (st.state != ST_PENDING && st.state != ST_ACTIVE) return PVVC_ERR_INVALID_ARG;
    if (static_cast<int64_t>(frame->points) > st.max_pts_frame) return PVVC_ERR_CAPACITY;
    for (uint32_t s = 0; s < frame->slice_cnt; ++s)
        if (frame->index[s] < 0 || frame->index[s] >= st.max_patches) return PVVC_ERR_INVALID_ARG;
    st.pending.push_back(Req{frame, arrival_s, tag, st.next_seq++});
    sv->cv_sched.notify_one();
    return PVVC_OK;
}

int pvvc_server_close_stream(pvvc_server* sv, int stream_id) {
    if (!sv) return PVVC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->streams.find(stream_id);
    if (it == sv->streams.end()) return PVVC_ERR_INVALID_ARG;
    StreamRec& st = it->second;
    if (st.state == ST_PENDING) {
        st.state = ST_CLOSED;
        st.pending.clear();
        admit_waiting(sv);
        return PVVC_OK;
    }
    if (st.state == ST_ACTIVE) st.state = ST_CLOSING;
    maybe_retire(sv, st);
    return PVVC_OK;
}

int pvvc_server_poll(pvvc_server* sv, pvvc_completion* out, int max, int timeout_ms) {
    if (!sv || !out || max <= 0) return -PVVC_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(sv->mu);
    if (sv->done.empty() && timeout_ms != 0) {
        auto pred = [&] { return !sv->done.empty() || sv->stop; };
        if (timeout_ms < 0) sv->cv_done.wait(lk, pred);
        else sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    }
    int k = 0;
    while (k < max && !sv->done.empty()) {
        out[k++] = sv->done.front();
        sv->done.pop_front();
    }
    return k;
}

int pvvc_server_release(pvvc_server* sv, int stream_id, int out_slot) {
    if (!sv) return PVVC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->streams.find(stream_id);
    if (it == sv->streams.end()) return PVVC_ERR_INVALID_ARG;
    StreamRec& st = it->second;
    if (out_slot < 0 || out_slot >= static_cast<int>(st.ring_busy.size())) return PVVC_ERR_INVALID_ARG;
    st.ring_busy[out_slot] = 0;
    sv->cv_sched.notify_one();
    return PVVC_OK;
}

int pvvc_server_get_stats(pvvc_server* sv, pvvc_server_stats* out) {
    if (!sv || !out) return PVVC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    std::memset(out, 0, sizeof(*out));
    out->frames_done = sv->frames_done;
    out->points_done = sv->points_done;
    out->batches = sv->batches;
    out->hbm_budget = sv->budget;
    out->hbm_used = sv->used;
    for (auto& kv : sv->streams) {
        if (kv.second.state == ST_ACTIVE || kv.second.state == ST_CLOSING) out->streams_active++;
        if (kv.second.state == ST_PENDING) out->streams_waiting++;
        out->backlog_frames += static_cast<int64_t>(kv.second.pending.size());
    }
    out->host_ms_per_batch = sv->batches ? sv->host_s * 1e3 / static_cast<double>(sv->batches) : 0.0;
    std::vector<double> v = sv->lat_ms;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto q = [&](double p) {
            const double x = p * static_cast<double>(v.size() - 1);
            const size_t i = static_cast<size_t>(x);
            const double fr = x - static_cast<double>(i);
            return i + 1 < v.size() ? v[i] * (1 - fr) + v[i + 1] * fr : v[i];
        };
        out->p50_ms = q(0.50);
        out->p99_ms = q(0.99);
        out->max_ms = v.back();
        if (sv->last_done > sv->first_arrival && sv->first_arrival >= 0)
            out->mpts_per_s = static_cast<double>(sv->points_done) / (sv->last_done - sv->first_arrival) / 1e6;
    }
    return PVVC_OK;
}

int pvvc_server_reset_stats(pvvc_server* sv) {
    if (!sv) return PVVC_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->frames_done = sv->points_done = sv->batches = 0;
    sv->host_s = 0;
    sv->lat_ms.clear();
    sv->first_arrival = -1;
    sv->last_done = 0;
    return PVVC_OK;
}

}  // extern "C"
