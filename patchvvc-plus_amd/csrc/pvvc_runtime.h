/*
 * pvvc_runtime.h -- host-side internals shared by the decoder library's
 * sources: batch planning, persistent launches, device helpers.
 */
#ifndef PVVC_RUNTIME_H_
#define PVVC_RUNTIME_H_

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pvvc_internal.h"
#include "pvvc_kernels.h"

namespace pvvc {

/* Queue segment: one persistent launch. */
struct Segment {
    uint32_t begin, count;       /* range in order[] */
    uint32_t dbegin, dcount;     /* deferred predictive items, range in defer[] */
    uint64_t ws_need;            /* per-item workspace bound */
};

/* Host-side plan of a batch image: [items][order][status][defer][frame blobs],
 * 256-B aligned sections.  Frame f's slices use state slots
 * slot_base[f] + index[s]. */
struct BatchPlan {
    std::vector<ItemDev> items;
    std::vector<uint64_t> geo_at, col_at;  /* blob offsets in the image (UINT64_MAX: none) */
    std::vector<uint32_t> order, defer;
    std::vector<Segment> segs;
    std::vector<int64_t> frame_off, frame_pts;
    std::vector<uint32_t> frame_first, frame_count;
    std::vector<size_t> frame_data_at;
    uint32_t n_items = 0;
    int64_t total_points = 0, input_bytes = 0, state_bytes = 0;
    size_t o_items = 0, o_order = 0, o_status = 0, o_defer = 0, o_data = 0, total = 0;
};

/* Workspace bound of one item (coefficients, first-child and weight arrays,
 * top-down buffers); geo_bytes is the geometry blob it expands. */
inline uint64_t item_ws_bound(uint32_t n, uint64_t geo_bytes) {
    return 12ull * n + 2 * (4 * (geo_bytes + 1) + 16) + 48ull * n + 512;
}

int plan_batch(const pvvc_frame* const* frames, int n_frames, const int* slot_base, const int* slot_count,
               const uint32_t* geo_cap, uint32_t* intra_geo, BatchPlan& p);
void fill_batch_image(const BatchPlan& p, const pvvc_frame* const* frames, uint8_t* dev_base, uint8_t* host);

/* Device addresses of a batch image. */
struct BatchDev {
    const ItemDev* items;
    const uint32_t* order;
    const uint32_t* defer;
    int* status;
};

/* Per-launch resources: state slots, per-workgroup slab, stage, queue counter. */
struct LaunchRes {
    SlotDev* slots;
    uint8_t* slab;
    uint64_t slab_bytes;   /* per workgroup: the stride between workgroups' workspaces */
    uint64_t slab_total;   /* bytes of the whole slab */
    float* stage;
    uint32_t* queue;
    uint32_t grid_cap;     /* at most this many workgroups (0: no cap) */
    const float* pre_coef = nullptr;  /* coefficients decoded by lane_rlgr_kernel (decode_slice) */
    const int* pre_status = nullptr;
    bool timed = true;     /* record the launch's HIP events for pvvc_last_decode_kernel_times */
};

/* Workgroups of a persistent launch of `count` items. */
uint32_t persistent_grid(const pvvc_tuning& t, uint32_t count);
/* The launch shape of a segment of `count` items: `t`, or, when auto_small
 * and the segment has at most 4 items per CU, 4 wavefronts per item with
 * 32 KiB of LDS and 4 workgroups per CU. */
pvvc_tuning segment_tuning(const pvvc_tuning& t, bool auto_small, uint32_t count);
/* Enqueues one segment: persistent_kernel (its last workgroup re-arms the
 * queue) and the finish pass.  PVVC_ERR_CAPACITY, with nothing enqueued, when
 * the segment's workspace bound exceeds the slab stride or the grid's
 * workspaces exceed the slab. */
int launch_segment(const BatchDev& b, const Segment& sg, const Sink& sink, const pvvc_tuning& t,
                    const LaunchRes& r, uint32_t launch_id, hipStream_t s);
/* Records a HIP failure in the library's sticky error; true on success. */
bool rt_hip_ok(hipError_t e);
void rt_set_error(int e);
/* The library stream (pvvc_set_stream; null stream by default). */
hipStream_t rt_stream();
/* Sets the dynamic-LDS limits of the decode kernels (once per process). */
void rt_init_kernels();
/* Forgets the calling thread's launch-time records (their events are reused). */
void rt_timing_reset();
/* Loads the decode kernels on stream s ahead of real work: an empty
 * persistent launch (its queue is at or past the item count, so every
 * workgroup exits at once) with the given launch shape. */
bool rt_warm_kernels(const pvvc_tuning& t, uint32_t* queue, hipStream_t s);

}  // namespace pvvc

#endif
