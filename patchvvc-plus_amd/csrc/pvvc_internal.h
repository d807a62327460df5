/*
 * pvvc_internal.h -- types shared by the host library sources.
 */
#ifndef PVVC_INTERNAL_H_
#define PVVC_INTERNAL_H_

#include <cstdint>
#include <vector>

#include "../../include/pvvc_hip.h"

/* Slice type bits: include/common/common.h:32-56 (valid, predict, skip,
 * geometry zstd, colour zstd); defaults I = 0x01, P = 0x03. */
constexpr uint8_t PVVC_TYPE_VALID = 0x01;
constexpr uint8_t PVVC_TYPE_PREDICT = 0x02;
constexpr uint8_t PVVC_TYPE_SKIP = 0x04;
constexpr uint8_t PVVC_TYPE_GEO_ZSTD = 0x08;
constexpr uint8_t PVVC_TYPE_COLOR_ZSTD = 0x10;

/* A loaded frame: the reference's Frame_t (include/cuda/base.cuh:227-293)
 * as SoA vectors, with all unwrapped blobs in one 16-byte-aligned arena. */
struct pvvc_frame {
    int timestamp = -1;
    uint32_t slice_cnt = 0;
    std::vector<int> index;
    std::vector<uint8_t> type, qp;
    std::vector<uint32_t> size, geo_size, col_size;
    std::vector<float> mv; /* [S][16] row-major */
    std::vector<uint64_t> geo_off, col_off;
    std::vector<uint8_t> data;
    uint64_t points = 0;
};

int pvvc_frame_parse_image(const uint8_t* bytes, size_t nbytes, pvvc_frame& f);

#endif
