/*
 * pvvc_hip.h -- C ABI of the MI355X-native patchVVC per-patch GoP decoder.
 *
 * Drop-in surface: the reference client's decode-frame entry
 * (`launch_cudaProcess`, include/cuda/render.cuh:41-57, defined in
 * src/cuda/render/CUDA2OpenGL.cu:68-74) and its frame loader
 * (`io::LoadFrame`, src/cuda/common/slice_io.cu:143-275), plus native
 * status-returning calls for the batched persistent decode and the
 * multi-stream serving loop.  All device pointers are plain `void*`/typed
 * pointers into HIP device memory; streams are passed as `void*`
 * (a hipStream_t).  No call allocates from the device heap.
 */
#ifndef PVVC_HIP_H_
#define PVVC_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes.  Per-slice codes follow the reference CPU decoder's         */
/* exceptions (src/octree/invert_raht_octree.cc:25-40,188-190,264-266;       */
/* src/common/entropy_codec.cc:345-353) instead of aborting the process.     */
/* ------------------------------------------------------------------------ */
enum pvvc_status {
    PVVC_OK = 0,
    PVVC_ERR_BAD_SLICE = 1,        /* invalid type bit, no data, short geometry, height 0 or > 64 */
    PVVC_ERR_EMPTY_REFERENCE = 2,  /* predictive slice without a decoded intra slice */
    PVVC_ERR_UNMATCHED_SIZE = 3,   /* leaf count != size, or P size != intra size */
    PVVC_ERR_ZSTD = 4,             /* zstd frame could not be unwrapped */
    PVVC_ERR_BITSTREAM = 5,        /* RLGR read past the colour bytes, or run exponent > 30 */
    PVVC_ERR_SYMBOL_RANGE = 6,     /* RLGR symbol outside int32 */
    PVVC_ERR_CAPACITY = 7,         /* patch larger than its state slot / workspace */
    PVVC_ERR_INVALID_ARG = 8,
    PVVC_ERR_HIP = 9,              /* a HIP runtime call failed */
    PVVC_ERR_TIMEOUT = 10,         /* reserved, never returned: predictive slices never wait in a
                                      kernel (a finish pass adds their reference colours) */
    PVVC_ERR_IO = 11,              /* file error (see pvvc_load_frame for the reference codes) */
    PVVC_ERR_BUSY = 12             /* serving: no slot / memory to admit a stream */
};

/* ------------------------------------------------------------------------ */
/* Drop-in decode-frame entry                                                */
/* ------------------------------------------------------------------------ */

/* Output point, identical to the reference's `Points`
 * (include/cuda/base.cuh:50-58): xyz after the patch's rigid motion, RGB as
 * unrounded float / 255 (src/cuda/render/CUDA2OpenGL.cu:46-51). */
typedef struct Points {
    float x, y, z;
    float r, g, b;
} Points;

/* Opaque device-resident per-patch decoder state: replaces the array of
 * `vvc::client::octree::InvertRAHTOctree` objects
 * (include/cuda/octree.cuh:130-171) that the reference caller allocates
 * with cudaMalloc and fills with default-constructed host copies
 * (src/cuda/common/manager.cc:235-238, test/test_decode_only.cc:51-55).
 * Holds, per patch index, the intra geometry bytes, the reference colours
 * and a decode workspace; nothing is taken from the device heap. */
typedef struct pvvc_state pvvc_state;

/* Allocates state for patch indices [0, n_patches) with up to max_points
 * points per patch.  Returns NULL on failure (pvvc_last_error() says why). */
pvvc_state* pvvc_state_alloc(int n_patches, int max_points);
void pvvc_state_free(pvvc_state* state);
/* Status of patch `patch` from the last decode that touched it.  Synchronises
 * the state's stream and, like pvvc_state_synchronize, moves the state's
 * pending launch error into the calling thread's pvvc_last_error(). */
int pvvc_state_patch_status(const pvvc_state* state, int patch);

/* Per-state settings of the drop-in call.  Each state object carries its own
 * stream, launch shape and white-small flag, so host threads that each drive
 * their own state never share mutable library state; a setting a state has
 * not made follows the process-wide default (pvvc_set_stream,
 * pvvc_set_dropin_tuning, pvvc_set_white_small_patches).  The reference
 * creates and destroys a stream inside every call
 * (src/cuda/render/CUDA2OpenGL.cu:69-73), so its callers never share one. */
int pvvc_state_set_stream(pvvc_state* state, void* stream);
/* waves 0 returns the state to the process default. */
int pvvc_state_set_dropin_tuning(pvvc_state* state, int waves, int lds_bytes);
/* enable -1 returns the state to the process default. */
int pvvc_state_set_white_small_patches(pvvc_state* state, int enable);
/* Waits for the state's stream, then returns (and records in the calling
 * thread's pvvc_last_error) the status of the first slice that failed in the
 * drop-in calls since the previous read: PVVC_ERR_INVALID_ARG for an index[s]
 * outside [0, n_patches), else that slice's per-patch code; PVVC_OK if none.
 * The reference reports nothing (its kernel fails silently). */
int pvvc_state_synchronize(pvvc_state* state);

/* Drop-in for launch_cudaProcess (include/cuda/render.cuh:41-57,
 * src/cuda/render/CUDA2OpenGL.cu:68-74).  Same argument meaning: every
 * pointer is caller-owned device memory; geometry[s]/color[s] hold the
 * already zstd-unwrapped bytes of slice s; mv[s] is 16 row-major floats;
 * inner_offset[s] is the output offset of slice s; the state array is
 * indexed by index[s].  grid and block are accepted and ignored (the
 * library picks the launch shape).  Enqueued asynchronously on the state's
 * stream (pvvc_state_set_stream; else the process default of
 * pvvc_set_stream; else the null stream); the caller synchronises, as the
 * reference callers do (test/test_decode_only.cc:147).  Calls on different
 * state objects are independent and may come from different threads.
 * The reference returns nothing and reports nothing; here a slice's failure
 * is recorded per patch (pvvc_state_patch_status), and the first failure of
 * the state's calls (including an index[s] out of range, whose slice is
 * skipped) is returned by pvvc_state_synchronize and folded into the calling
 * thread's pvvc_last_error() by pvvc_state_synchronize,
 * pvvc_state_patch_status and, for states on the default stream,
 * pvvc_synchronize. */
void launch_cudaProcess(int grid, int block, Points* cudaData, int timestamp, int* inner_offset, int* index,
                        uint8_t* type, float** mv, uint32_t* size, uint8_t* qp, uint8_t** geometry,
                        uint32_t* geometry_size, uint8_t** color, uint32_t* color_size, pvvc_state* invertRAHTOctree_gpu,
                        int patch_size);

/* Process-wide default launch shape of launch_cudaProcess, for states that
 * set none: wavefronts per slice (1, 2, 4) and the LDS workspace per slice in
 * bytes (arrays that do not fit use the state slot's global workspace).
 * Default: 4 wavefronts, 32 KiB. */
int pvvc_set_dropin_tuning(int waves, int lds_bytes);
/* Process-wide default stream (a hipStream_t) of the states and decoders that
 * set none, of the encoder, the metric and the helper copies. */
int pvvc_set_stream(void* stream);
/* Last error recorded by a call made on the calling thread (cleared when
 * read).  Each host thread has its own. */
int pvvc_last_error(void);
/* Reference-only rendering artefact: patches with fewer than 10 points
 * forced to white (src/cuda/render/CUDA2OpenGL.cu:54-58).  Process-wide
 * default, off. */
int pvvc_set_white_small_patches(int enable);

/* ------------------------------------------------------------------------ */
/* Container (host): replaces io::LoadFrame (src/cuda/common/slice_io.cu:    */
/* 143-275).  The file layout is the one ChangeSliceToFrame writes           */
/* (src/io/slice_io.cc:258-315).  Blobs flagged zstd (type bit 3 geometry,   */
/* bit 4 colour) are unwrapped on load, as the reference loader does.        */
/* ------------------------------------------------------------------------ */
typedef struct pvvc_frame pvvc_frame;

/* Returns 0, or the reference loader's codes: 1 wrong file name (not
 * "*.frame"), 2 no such file, 3 permission denied, 4 other open error,
 * 5 short read; plus 6 for a zstd frame that fails to unwrap. */
int pvvc_load_frame(const char* path, pvvc_frame** out);
/* Parses a .frame image already in host memory (same codes, 1-4 unused). */
int pvvc_parse_frame(const uint8_t* bytes, size_t nbytes, pvvc_frame** out);
void pvvc_frame_free(pvvc_frame* frame);
/* Host threads that unwrap a frame's zstd blobs in parallel (default:
 * min(16, hardware threads); 1 unwraps on the calling thread only).  The
 * reference loader unwraps slice by slice on one thread
 * (src/cuda/common/slice_io.cu:225-270). */
int pvvc_set_host_threads(int n);
int pvvc_frame_timestamp(const pvvc_frame* frame);
uint32_t pvvc_frame_slice_cnt(const pvvc_frame* frame);
uint64_t pvvc_frame_points(const pvvc_frame* frame);
/* Copies per-slice metadata out (any pointer may be NULL).  mv is [S][16]. */
void pvvc_frame_meta(const pvvc_frame* frame, int* index, uint8_t* type, uint32_t* size, uint8_t* qp, float* mv,
                     uint32_t* geometry_size, uint32_t* color_size);
const uint8_t* pvvc_frame_geometry(const pvvc_frame* frame, uint32_t slice);
const uint8_t* pvvc_frame_color(const pvvc_frame* frame, uint32_t slice);

/* ------------------------------------------------------------------------ */
/* .slice files (host): one patch slice per file, the layout of io::SaveSlice */
/* / io::LoadSlice (src/io/slice_io.cc:20-206):                              */
/*   u8 type; i32 timestamp; i32 index; u64 size; u8 qp;                     */
/*   f32 mv[16] column-major (Eigen::Matrix4f storage);                      */
/*   if not predictive (type bit 1): u64 geometry size > 0, geometry bytes;  */
/*   if not skip (type bit 2): u64 colour size > 0, colour bytes.            */
/* Codes: 1 wrong file name ("*.slice" in, "*.frame" out), 2 no such file    */
/* (or no slice file in the directory), 3 permission denied, 4 other open    */
/* error or not a directory, 5 short read, 6 zstd failure, 7 type without   */
/* the valid bit, 8 empty geometry or colour blob, 9 no slices, 10 write     */
/* error.                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct pvvc_slice_desc {
    uint8_t type;
    int32_t timestamp;
    int32_t index;
    uint64_t size;           /* points */
    uint8_t qp;
    float mv[16];            /* row-major, as in a .frame; stored transposed */
    const uint8_t* geometry; /* written unless the slice is predictive */
    uint64_t geometry_size;
    const uint8_t* color;    /* written unless the slice is a skip */
    uint64_t color_size;
} pvvc_slice_desc;

/* io::SaveSlice (slice_io.cc:20-120): returns the bytes written, or minus a
 * code above. */
int64_t pvvc_save_slice(const char* path, const pvvc_slice_desc* slice);
/* io::LoadSlice of each path (slice_io.cc:124-206), then the frame that
 * io::ChangeSliceToFrame would write for them (slice_io.cc:208-315: slices
 * ordered by patch index, stable among equal indices; the first ordered
 * slice's timestamp), parsed and zstd-unwrapped as pvvc_parse_frame does.
 * The reference decoder harness reads slices this way
 * (PVVCDecompression::LoadSlices, src/codec/pvvc_decompression.cc:34-84). */
int pvvc_load_slices(const char* const* paths, int n, pvvc_frame** out);
/* Writes that frame as a .frame file (the bytes ChangeSliceToFrame writes). */
int pvvc_slices_to_frame_file(const char* const* paths, int n, const char* frame_path);
/* io::ChangeSliceToFrame (slice_io.cc:208-315): every "*_time_<t>_slice_<i>.slice"
 * file of input_dir into one .frame file. */
int pvvc_change_slice_to_frame(const char* input_dir, const char* frame_path);

/* Stages a frame for launch_cudaProcess the way test/test_decode_only.cc:
 * 97-123 does, but with one packed host->device copy: fills the device
 * arrays of the drop-in call.  `dev` must come from pvvc_dropin_alloc. */
typedef struct pvvc_dropin_frame {
    int* inner_offset;
    int* index;
    uint8_t* type;
    float** mv;
    uint32_t* size;
    uint8_t* qp;
    uint8_t** geometry;
    uint32_t* geometry_size;
    uint8_t** color;
    uint32_t* color_size;
    int slice_cnt;
    int64_t point_cnt;
    void* blob;        /* device allocation holding all of the above */
    size_t blob_bytes;
} pvvc_dropin_frame;
int pvvc_dropin_stage(const pvvc_frame* frame, pvvc_dropin_frame* dev);
void pvvc_dropin_release(pvvc_dropin_frame* dev);
/* Bytes of the packed image pvvc_dropin_stage_to needs for this frame. */
size_t pvvc_dropin_stage_bytes(const pvvc_frame* frame);
/* Stages into caller buffers, allocating nothing: packs the image into
 * `host` (pinned memory from pvvc_host_alloc makes the copy asynchronous)
 * and copies it to `dev` on the default stream.  `host` may be reused once
 * the stream has passed the copy.  PVVC_ERR_CAPACITY if either buffer is
 * smaller than pvvc_dropin_stage_bytes(frame). */
int pvvc_dropin_stage_to(const pvvc_frame* frame, void* dev, size_t dev_bytes, void* host, size_t host_bytes,
                         pvvc_dropin_frame* out);
/* Pinned host memory (hipHostMalloc). */
void* pvvc_host_alloc(size_t bytes);
void pvvc_host_free(void* ptr);

/* ------------------------------------------------------------------------ */
/* Native decoder: batched persistent decode                                 */
/* ------------------------------------------------------------------------ */
typedef struct pvvc_decoder pvvc_decoder;
typedef struct pvvc_batch pvvc_batch;

/* Tunables (autotuner output).  pvvc_decoder_set_tuning keeps the current value of a zero field. */
typedef struct pvvc_tuning {
    int waves_per_patch;   /* 1, 2 or 4 waves per work item */
    int lds_bytes;         /* LDS workspace per work item (4..64 KiB) */
    int wgs_per_cu;        /* persistent workgroups per CU */
    int queue_chunk;       /* items taken per queue pop */
    int priority_items;    /* the first N items of the length-sorted queue run at the highest
                              wave priority (s_setprio 3), the rest of the queue in thirds at
                              2, 1 and 0; -1 = all at 0; default: N = one item per persistent
                              workgroup (the items that start together) */
    int lane_max_points;   /* persistent mode: slices of at most this many points decode their
                              RLGR one slice per lane (vector unit) on a second stream of the
                              decoder, beside the scalar decoder of the longer slices, and join
                              the caller's stream before the call returns; 0 = off (default),
                              and -1 switches it off in pvvc_decoder_set_tuning */
} pvvc_tuning;

int pvvc_decoder_create(pvvc_decoder** out, int max_patches, int max_points_per_patch);
int pvvc_decoder_destroy(pvvc_decoder* dec);
/* Until a tuning is set (or the autotuner runs), a persistent segment of at most 4 slices per CU
 * runs 4 wavefronts per slice with 32 KiB of LDS and 4 workgroups per CU; larger ones use the
 * decoder's tuning (default 1 wavefront, 8 KiB, 16 workgroups per CU, queue chunk 1). */
int pvvc_decoder_set_tuning(pvvc_decoder* dec, const pvvc_tuning* tuning);
int pvvc_decoder_get_tuning(const pvvc_decoder* dec, pvvc_tuning* tuning);
/* The decoder's own stream (a hipStream_t) for every call on it and its
 * batches; until set, the process default (pvvc_set_stream). */
int pvvc_decoder_set_stream(pvvc_decoder* dec, void* stream);
/* The decoder's own white-small setting (-1: the process default). */
int pvvc_decoder_set_white_small_patches(pvvc_decoder* dec, int enable);

/* Packs frames [0, n_frames) (in decode order) into one device-resident
 * batch: a single host->device copy.  Work items are all slices of all
 * frames; patches whose intra slice is in the batch are ordered after it. */
int pvvc_batch_create(pvvc_decoder* dec, const pvvc_frame* const* frames, int n_frames, pvvc_batch** out);
void pvvc_batch_free(pvvc_batch* batch);
int pvvc_batch_frames(const pvvc_batch* batch);
int64_t pvvc_batch_frame_points(const pvvc_batch* batch, int frame);
int64_t pvvc_batch_frame_offset(const pvvc_batch* batch, int frame);
/* Bytes the device reads from the batch's input: 82 B of slice metadata per
 * slice plus the unwrapped geometry and colour blobs. */
int64_t pvvc_batch_input_bytes(const pvvc_batch* batch);
/* Bytes of resident GoP state the batch's predictive slices read: their
 * patch's intra geometry blob plus 12 B/pt of reference YUV. */
int64_t pvvc_batch_state_bytes(const pvvc_batch* batch);
int pvvc_batch_items(const pvvc_batch* batch);

/* Output modes. */
enum pvvc_output { PVVC_OUT_NATIVE = 0, PVVC_OUT_POINTS = 1 };
/* Decode modes. */
enum pvvc_mode {
    PVVC_MODE_PER_FRAME = 0,   /* one launch per frame, one workgroup per slice */
    PVVC_MODE_PERSISTENT = 1,  /* whole batch in one persistent launch, length-sorted queue */
    PVVC_MODE_GRAPH = 2        /* per-frame launches replayed from one captured hipGraph per frame */
};

/* Decodes the batch asynchronously on the decoder's stream.  Output for
 * frame f starts at point pvvc_batch_frame_offset(f).  Native output:
 * xyz (float x3) and rgb (uint8 x3) per point; Points output: `points`. */
int pvvc_decode_batch(pvvc_decoder* dec, pvvc_batch* batch, int mode, int output, float* xyz, uint8_t* rgb,
                      Points* points);
/* Per-item status of the last decode of this batch (synchronises). */
int pvvc_batch_status(pvvc_batch* batch, int* status, int n);

/* One-call decodes (pack, decode, wait): the per-frame step of the client
 * decoder thread (Manager::StartDecoder, src/client/manager.cc:257, which
 * calls Render::CUDADecode, src/cuda/render/render.cc:270-289) and a whole
 * GoP in one persistent launch.  Outputs are device buffers as for
 * pvvc_decode_batch; frame f of a GoP starts after the points of frames
 * [0, f).  Returns the call's own error, else the first non-OK slice status
 * (frames in order, slices in file order), else PVVC_OK; `status`
 * (optional, status_cap entries) receives every slice's status. */
int pvvc_decode_frame(pvvc_decoder* dec, const pvvc_frame* frame, int output, float* xyz, uint8_t* rgb,
                      Points* points, int* status, int status_cap);
int pvvc_decode_gop(pvvc_decoder* dec, const pvvc_frame* const* frames, int n_frames, int output, float* xyz,
                    uint8_t* rgb, Points* points, int* status, int status_cap);
/* Resets every patch state slot (drops GoP state). */
int pvvc_decoder_reset(pvvc_decoder* dec);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The PVVC_MODE_PERSISTENT decode of pvvc_decode_batch,
 * identical output, with no allocation of any kind: the queue counters, the
 * stage of deferred predictive slices and the per-workgroup workspaces come
 * from one library-owned device array of 256 MiB per GPU (held until the
 * process ends), and the launch's workgroups are cut to the workspaces that
 * fit there.  PVVC_ERR_CAPACITY, with nothing enqueued, when the batch's
 * stage (12 B per point of the batch) and one workspace exceed it (about
 * 21 M points per batch).  The decoder and the batch are the caller's, made
 * beforehand (pvvc_decoder_create / pvvc_batch_create allocate).  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other threads, other decoders) race, so the caller serialises
 * them.  Asynchronous on the decoder's stream; no kernel times are recorded. */
int pvvc_decode_batch_noalloc(pvvc_decoder* dec, pvvc_batch* batch, int output, float* xyz, uint8_t* rgb,
                              Points* points);

/* Autotuner: picks waves per slice (1/2/4), the LDS budget per slice
 * (4-64 KiB) and, for the persistent mode, workgroups per CU, the queue
 * chunk and the lane decoder's slice size limit (off, 256, 1000, 2000, 4000 points)
 * by coordinate descent, timing `reps` decodes of `batch` per
 * candidate (HIP events; median).  Sets the decoder's tuning to the best
 * and returns it with its time; the outputs hold that configuration's
 * decode.  Synchronises the decoder's stream. */
int pvvc_decoder_autotune(pvvc_decoder* dec, pvvc_batch* batch, int mode, int output, float* xyz, uint8_t* rgb,
                          Points* points, int reps, pvvc_tuning* best, double* best_ms);

/* ------------------------------------------------------------------------ */
/* Forward path: colour encoder (RAHTOctree::RAHT, raht_octree.cc:189-247;  */
/* quantisation, GoP_encoding.cc:172-205; RLGREncoder::Encode,              */
/* entropy_codec.cc:65-159) on the GPU, one wavefront per colour stream.     */
/* ------------------------------------------------------------------------ */
typedef struct pvvc_color_job {
    const uint8_t* geometry;    /* host: the patch's octree bytes (as in an intra slice, unwrapped) */
    uint32_t geometry_size;
    const float* yuv;           /* host: n leaf colours, YUV AoS, in leaf (decoder output) order; the
                                   reference averages the points of a leaf (raht_octree.cc:189-200) */
    uint32_t n;
    uint8_t qp;                 /* quantisation step (> 0) */
    uint8_t* out;               /* host: receives the RLGR bytes (not zstd-wrapped) */
    uint32_t out_cap;
    uint32_t out_size;          /* result: bytes written */
    int status;                 /* result: PVVC_OK, or BAD_SLICE / UNMATCHED_SIZE (geometry),
                                   SYMBOL_RANGE (coefficient outside int32), CAPACITY (out_cap) */
} pvvc_color_job;
/* Encodes every job's colours (one launch); blocks until the bytes are back.
 * kernel_ms (optional) receives the encode kernel's time.  The device image
 * and pinned staging buffers are kept for the next call (grow-only); calls
 * from several threads are serialised. */
int pvvc_encode_colors(pvvc_color_job* jobs, int n_jobs, double* kernel_ms);
/* Frees the encoder's kept buffers (the next pvvc_encode_colors makes them again). */
void pvvc_encoder_release(void);

/* ------------------------------------------------------------------------ */
/* Quality metric: symmetric nearest-neighbour MSE of two coloured clouds,   */
/* replacing vvc::common::MSE (SetClouds + Compute + Get*MSEs,               */
/* include/registration/registration.h:48-100, registration_base.cc:159-238). */
/* ------------------------------------------------------------------------ */
typedef struct pvvc_mse {
    /* {geo, y, u, v}: means over the points of p of the squared distance to
     * the nearest point of q and of the squared Y/U/V difference to it
     * (MSE::Compute's `.first`); q_to_p the same from q (`.second`).  -1 for
     * an empty cloud's direction; the other direction is then left at 0. */
    double p_to_q[4];
    double q_to_p[4];
} pvvc_mse;
/* p_xyz/q_xyz: device f32 x3 per point; p_rgb/q_rgb: device RGB8 x3 (the
 * decoder's native output layout).  Neighbour = least squared distance in
 * FP32, (dx*dx + dy*dy) + dz*dz; among equidistant points the lowest index
 * (the reference's FLANN k-d tree leaves that order unspecified).  YUV of an
 * RGB8 point as ColorYUV(PointXYZRGB) (include/common/common.h:95-99).
 * Sums are FP64 in a fixed order, so results are deterministic; the
 * reference keeps float running sums.  nn_p / nn_q (optional, device int64
 * per point) receive each point's neighbour index.  Blocks until done;
 * kernel_ms (optional) receives the device time of the whole computation. */
int pvvc_cloud_mse(const float* p_xyz, const uint8_t* p_rgb, uint64_t np, const float* q_xyz,
                   const uint8_t* q_rgb, uint64_t nq, int64_t* nn_p, int64_t* nn_q, pvvc_mse* out,
                   double* kernel_ms);
/* PSNR as PVVCDecompression prints it (pvvc_decompression.cc:156-159), in
 * FP32: 10 log10(peak^2 / max(a, b)); the reference's peaks are 1024 for
 * geometry and 256 for Y. */
float pvvc_psnr(double mse_a, double mse_b, float peak);

/* ------------------------------------------------------------------------ */
/* Serving loop: many concurrent streams, continuous batching.  The          */
/* reference client decodes one stream with a loader thread and a decoder   */
/* thread around launch_cudaProcess (src/cuda/common/manager.cc:145-310);    */
/* the server generalises that decoder thread to many streams.              */
/* ------------------------------------------------------------------------ */
typedef struct pvvc_server pvvc_server;

typedef struct pvvc_server_config {
    uint64_t hbm_budget;        /* bytes for admitted streams' state + output rings
                                   (0: 80% of the free HBM after the fixed buffers) */
    int lanes;                  /* batches in flight at once, each on its own HIP stream (default 3) */
    int max_slots;              /* patch state slots over all streams (default 262144) */
    int max_points_per_patch;   /* largest patch any stream may send (default 4096) */
    int64_t out_pool_points;    /* output pool shared by the streams' rings (default 256 Mi points,
                                   15 B each) */
    int64_t max_batch_points;   /* points per batch (default 16 Mi) */
    pvvc_tuning tuning;         /* launch shape of the persistent decode */
    int lanes_spin;             /* 1: completion threads spin on their lane's event (lowest wake-up
                                   latency, one host core each); 0: they block */
} pvvc_server_config;

typedef struct pvvc_completion {
    int stream_id;
    int status;                 /* first non-OK slice status of the frame, or OK */
    int64_t seq;                /* frame number within its stream (submission order) */
    uint64_t tag;               /* caller's tag from pvvc_server_submit */
    int64_t points;
    const float* xyz;           /* device: points x 3 floats, valid until pvvc_server_release */
    const uint8_t* rgb;         /* device: points x 3 bytes */
    int out_slot;               /* ring slot to hand back with pvvc_server_release */
    double arrival_s, launch_s, done_s;  /* pvvc_server_clock() times */
} pvvc_completion;

typedef struct pvvc_server_stats {
    int64_t frames_done, points_done, batches;
    uint64_t hbm_budget, hbm_used;
    int streams_active, streams_waiting;
    double p50_ms, p99_ms, max_ms;  /* arrival -> completion, per frame */
    double mpts_per_s;              /* points done / (last completion - first arrival) */
    double host_ms_per_batch;       /* scheduler time to plan, pack and launch a batch */
    int64_t backlog_frames;         /* frames queued, not yet in a batch */
} pvvc_server_stats;

/* Monotonic clock (seconds) for arrival stamps. */
double pvvc_server_clock(void);
int pvvc_server_create(const pvvc_server_config* cfg, pvvc_server** out);
/* Stops the loop after the batches in flight; frees everything. */
int pvvc_server_destroy(pvvc_server* srv);
/* Opens a stream.  It is admitted at once if its footprint (state slots for
 * max_patches patches of up to max_points_per_patch points, plus out_ring
 * output frames of max_points_per_frame points) fits the budget and no
 * stream waits before it; else it waits (state 0) and is admitted, first in
 * first out, when closed streams free memory.  Frames may be submitted
 * while it waits.  PVVC_ERR_CAPACITY if it could never fit. */
int pvvc_server_open_stream(pvvc_server* srv, int max_patches, int max_points_per_patch, int64_t max_points_per_frame,
                            int out_ring, int* stream_id);
/* 0 waiting for admission, 1 active, 2 closing, 3 closed; -1 unknown id. */
int pvvc_server_stream_state(pvvc_server* srv, int stream_id);
/* Queues a frame (decode order).  The frame must stay alive until its
 * completion has been polled.  arrival_s stamps the latency start. */
int pvvc_server_submit(pvvc_server* srv, int stream_id, const pvvc_frame* frame, double arrival_s, uint64_t tag);
/* Closes a stream: its queued frames still decode, then its memory is freed
 * and waiting streams are admitted. */
int pvvc_server_close_stream(pvvc_server* srv, int stream_id);
/* Takes up to `max` completions, waiting up to timeout_ms (-1: forever). */
int pvvc_server_poll(pvvc_server* srv, pvvc_completion* out, int max, int timeout_ms);
/* Hands an output ring slot back; the stream's next frame may then use it. */
int pvvc_server_release(pvvc_server* srv, int stream_id, int out_slot);
int pvvc_server_get_stats(pvvc_server* srv, pvvc_server_stats* out);
int pvvc_server_reset_stats(pvvc_server* srv);

/* ------------------------------------------------------------------------ */
/* Device memory / timing helpers (plumbing for callers without a HIP        */
/* binding of their own).                                                    */
/* ------------------------------------------------------------------------ */
void* pvvc_dev_alloc(size_t bytes);
void pvvc_dev_free(void* ptr);
int pvvc_memcpy_h2d(void* dst, const void* src, size_t bytes);
int pvvc_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy on the default stream (asynchronous). */
int pvvc_memcpy_d2d(void* dst, const void* src, size_t bytes);
int pvvc_memset(void* dst, int value, size_t bytes);
/* Waits for the default stream, then folds the pending launch errors of the
 * states on that stream into the calling thread's pvvc_last_error(). */
int pvvc_synchronize(void);
/* A new non-blocking HIP stream at the default priority (for callers without a
 * HIP binding of their own); destroy it after its last use. */
int pvvc_stream_create(void** stream);
int pvvc_stream_destroy(void* stream);
int pvvc_stream_synchronize(void* stream);
/* Kernel timing for the calling thread's decode calls: HIP events recorded
 * on the launch stream around each launch, which the three queries below
 * read.  Off by default (no events, the queries report 0 launches): each
 * event is a marker between kernels that costs several microseconds, ~15 us
 * per persistent GoP decode.  Returns the previous setting.  No counterpart
 * in the reference, which times its harness from the host. */
int pvvc_set_timing(int on);
/* Kernel time of the last decode call made on the calling thread, from HIP
 * events recorded on the launch stream around its launches (milliseconds;
 * includes the queue-reset and finish launches of a persistent decode), when
 * timing was on for it. */
double pvvc_last_kernel_ms(void);
/* Per-launch times (ms) of the last decode call on the calling thread;
 * returns the launch count (fills at most `max`). */
int pvvc_last_launch_times(double* ms, int max);
/* Times (ms) of the decode kernels alone of the last decode call on the
 * calling thread: each persistent_kernel, frame_kernel or dropin_kernel
 * launch, events recorded directly before and after it on its stream
 * (none for a graph replay).  Returns the launch count. */
int pvvc_last_decode_kernel_times(double* ms, int max);
/* Device properties used for launch shapes: CU count, LDS per workgroup. */
int pvvc_device_info(int* cu_count, int* lds_per_block, int* warp_size, char* name, int name_len);

#ifdef __cplusplus
}
#endif
#endif /* PVVC_HIP_H_ */
