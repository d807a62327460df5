/*
 * pvvc_kernels.h -- device code of the per-patch GoP decoder for CDNA4 (gfx950).
 *
 * One work item = one slice (one patch of one frame).  A workgroup of
 * WAVES wavefronts decodes it end to end, fused:
 *
 *   R  RLGR colour decode (src/common/entropy_codec.cc:163-283), run on the
 *      scalar unit of one wavefront: the adaptive state, the 64-bit bit
 *      window and the loop control are wave-uniform (SGPRs).  The colour
 *      bytes arrive 256 at a time through one coalesced vector load per
 *      chunk (prefetched one chunk ahead) and are read into the window with
 *      v_readlane.  Each symbol is written into lane (n mod 64) of a VGPR
 *      accumulator; every 64 symbols the wave dequantises them
 *      (invert_raht_octree.cc:63-69) and stores them with one coalesced
 *      store.  Zero runs only advance the symbol counter.
 *   T  octree expansion (invert_raht_octree.cc:122-254): breadth-first over
 *      the occupancy bytes, one lane per node, byte popcount and a
 *      ballot/mbcnt exclusive scan giving each node's first child; subtree
 *      leaf counts bottom-up from the child ranges.
 *   H  inverse RAHT (invert_raht_octree.cc:256-317, octree_base.cc:84-100,
 *      186-206), level by level top-down, one lane per node, in FP32 with
 *      the reference's operation order.  The AC coefficients of node j on
 *      level l sit at reversed stream position N_l + firstchild(j) - j.
 *      Leaf centres follow the reference recurrence SubSpaceCenter
 *      (octree_base.cc:102-120) in the same pass.
 *   M  leaf colour (+ reference colour for predictive slices), the patch's
 *      rigid motion (pvvc_decompression.cc:141-149) and YUV->RGB
 *      (common.h:154-158), written straight to the output.
 *
 * Intra slices keep their geometry bytes and decoded colours in the patch's
 * state slot for the predictive slices of the GoP.  A predictive slice whose
 * intra slice is decoded in the same launch does not wait for it: it writes
 * its unreferenced leaf colours to a staging buffer and a finish pass after
 * the launch adds the reference (no cross-workgroup synchronisation, no
 * device-wide fences inside a launch).  Workspace arrays live in LDS when
 * they fit the per-item LDS budget, else in a global slab.
 *
 * Arithmetic is compiled with -ffp-contract=off and correctly rounded
 * divide/sqrt so that FP32 results equal the reference CPU decoder's.
 */
#ifndef PVVC_KERNELS_H_
#define PVVC_KERNELS_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pvvc_hip.h"

namespace pvvc {

constexpr int kMaxHeight = 64;

/* Per-patch state slot (device).  `ws` is this patch's private workspace
 * for launches that decode each patch at most once. */
struct SlotDev {
    uint8_t* geo;      /* intra geometry bytes, geo_cap */
    float* ref;        /* reference colours, planar Y[n_cap] U[n_cap] V[n_cap] */
    uint8_t* ws;       /* workspace, ws_cap bytes */
    uint64_t ws_cap;
    uint32_t geo_cap;
    uint32_t n_cap;
    uint32_t n;        /* points of the current GoP, 0 = no intra slice */
    uint32_t geo_size;
    int32_t status;    /* status of the last slice decoded for this patch */
    uint32_t ready;    /* reserved */
    uint32_t pad[2];
};

/* Work item of the native batched decode. */
struct ItemDev {
    const uint8_t* geo;   /* intra: own geometry; predictive: intra geometry in the
                             same batch, or nullptr to use the state slot */
    const uint8_t* col;
    uint32_t geo_size;
    uint32_t col_size;
    uint32_t n;
    int32_t patch;
    uint8_t type;
    uint8_t qp;
    uint8_t wait;         /* predictive item whose intra item is in the same batch:
                             its reference colours are added after the launch */
    uint8_t pad0;
    uint32_t stage_off;   /* first staged point of a deferred item (batch-relative) */
    uint64_t out_off;     /* first output point */
    float mv[12];         /* rows 0..2 of the row-major 4x4 */
};

/* Output sink: native (xyz f32x3 + rgb u8x3) or reference Points. */
struct Sink {
    float* xyz;
    uint8_t* rgb;
    Points* pts;
    int mode;          /* PVVC_OUT_NATIVE / PVVC_OUT_POINTS */
    int white_small;   /* reference artefact, CUDA2OpenGL.cu:54-58 */
};

/* Uniform description of the slice a workgroup decodes. */
struct SliceArgs {
    const uint8_t* geo;
    const uint8_t* col;
    uint32_t geo_size;
    uint32_t col_size;
    uint32_t n;
    uint32_t type;
    uint32_t qp;
    uint32_t wait;
    uint64_t out_off;
    uint32_t stage_off;
    float mv[12];
    float* stage;      /* predictive slice whose intra slice is decoded in the same
                          launch: unreferenced leaf YUV goes here (3 floats at
                          3 * (stage_off + leaf)); the reference is added by the
                          finish pass after the launch */
};

/* Level metadata shared through LDS. */
struct LevelMeta {
    uint32_t cnt[kMaxHeight + 1];   /* nodes per level */
    uint32_t base[kMaxHeight + 1];  /* first internal-node index of the level */
    uint32_t height;
    uint32_t internal;              /* internal node count */
    int32_t tree_status;
    int32_t rlgr_status;
    uint32_t pad[3];
};

}  // namespace pvvc

#endif
