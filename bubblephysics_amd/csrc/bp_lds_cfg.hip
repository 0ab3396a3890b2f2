// bp_lds_cfg.hip -- one configuration of the LDS-resident world kernel, chosen at build time by
// BP_LDS_CFG (the configuration's id in bp_lds_configs.h).
#include "bp_world_lds.h"
#include "bp_lds_configs.h"

#ifndef BP_LDS_CFG
#error "BP_LDS_CFG must name a configuration id"
#endif

#define BP_CAT2(a, b) a##b
#define BP_CAT(a, b) BP_CAT2(a, b)

namespace {
// groups: workgroups of the launch (count when zero or above count); cursor: the zeroed device
// word the workgroups take further worlds from when groups < count; placed: the word they count
// themselves in as they start (may be null).
template <int ID, int B, int P, int R, int SC, int BC, int VC, int TC>
hipError_t launch_if(const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats, const uint32_t *order,
                     uint32_t count, uint32_t steps, uint32_t statsStride, uint32_t notify, uint32_t groups,
                     uint32_t *cursor, uint32_t *placed, hipStream_t stream, bool &done) {
    if constexpr (ID == BP_LDS_CFG) {
        done = true;
        const bool pooled = cursor && groups > 0 && groups < count;
        hipLaunchKernelGGL((bpd::world_lds_kernel<B, P, R, SC, BC, VC, TC>), dim3(pooled ? groups : count), dim3(B), 0, stream,
                           arena, descs, stats, order, count, steps, statsStride, notify, pooled ? cursor : nullptr,
                           pooled ? placed : nullptr);
        return hipGetLastError();
    } else {
        return hipSuccess;
    }
}

template <int ID, int B, int P, int R, int SC, int BC, int VC, int TC>
void residency_if(int &blocks, bool &done) {
    if constexpr (ID == BP_LDS_CFG) {
        done = true;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, bpd::world_lds_kernel<B, P, R, SC, BC, VC, TC>, B, 0) !=
            hipSuccess)
            blocks = 0;
    }
}
}  // namespace

extern "C" __attribute__((visibility("hidden"))) hipError_t BP_CAT(bp_launch_lds_cfg_, BP_LDS_CFG)(
    const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats, const uint32_t *order, uint32_t count,
    uint32_t steps, uint32_t statsStride, uint32_t notify, uint32_t groups, uint32_t *cursor, uint32_t *placed,
    hipStream_t stream) {
    bool done = false;
    hipError_t e = hipSuccess;
#define BP_LDS_ONE(ID, B, P, R, SC, BC, VC, TC)                                                                            \
    if (!done) e = launch_if<ID, B, P, R, SC, BC, VC, TC>(arena, descs, stats, order, count, steps, statsStride, notify, groups, \
                                                          cursor, placed, stream, done);
    BP_LDS_CONFIGS(BP_LDS_ONE)
#undef BP_LDS_ONE
    return done ? e : hipErrorInvalidValue;
}

// Workgroups of this configuration one CU holds at once (0 when the runtime cannot tell).
extern "C" __attribute__((visibility("hidden"))) int BP_CAT(bp_residency_lds_cfg_, BP_LDS_CFG)() {
    int blocks = 0;
    bool done = false;
#define BP_LDS_ONE(ID, B, P, R, SC, BC, VC, TC) \
    if (!done) residency_if<ID, B, P, R, SC, BC, VC, TC>(blocks, done);
    BP_LDS_CONFIGS(BP_LDS_ONE)
#undef BP_LDS_ONE
    return blocks;
}
