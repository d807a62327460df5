/* Host build of the lane RLGR decoder (patchvvc-plus_amd/csrc/pvvc_lane.h),
 * for checking it against the oracle on the CPU: straight from memory, and
 * through the LDS ring's indexing (one lane's column of a host array). */
#include <vector>

#include "../patchvvc-plus_amd/csrc/pvvc_lane.h"

extern "C" int lane_decode(const uint8_t* src, uint32_t nbytes, uint32_t total, uint32_t qp, float* coef) {
    return pvvc::rlgr_decode_lane(src, nbytes, total, qp, coef);
}

extern "C" int lane_decode_ring(const uint8_t* src, uint32_t nbytes, uint32_t total, uint32_t qp, float* coef) {
    std::vector<uint32_t> ring((pvvc::kLaneRing + 1u) * 64u, 0xDEADBEEFu);
    const uintptr_t a = reinterpret_cast<uintptr_t>(src);
    const uint32_t skip = 8u * static_cast<uint32_t>(a & 3u);
    const uint32_t end = skip + 8u * nbytes;
    const pvvc::LaneRing r{reinterpret_cast<const uint32_t*>(a & ~uintptr_t(3)), end, end ? (end - 1u) / 32u : 0u,
                           0u, ring.data() + 17};
    return pvvc::rlgr_decode_lane_from(r, skip, nbytes, total, qp, coef);
}
