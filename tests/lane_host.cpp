/* Host build of the lane RLGR decoder (patchvvc-plus_amd/csrc/pvvc_lane.h),
 * for checking it against the oracle on the CPU. */
#include "../patchvvc-plus_amd/csrc/pvvc_lane.h"

extern "C" int lane_decode(const uint8_t* src, uint32_t nbytes, uint32_t total, uint32_t qp, float* coef) {
    return pvvc::rlgr_decode_lane(src, nbytes, total, qp, coef);
}
