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
template <int ID, int B, int P, int R, int SC, int BC, int VC, int TC>
hipError_t launch_if(const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats, const uint32_t *order,
                     uint32_t count, uint32_t steps, uint32_t statsStride, uint32_t notify, hipStream_t stream, bool &done) {
    if constexpr (ID == BP_LDS_CFG) {
        done = true;
        hipLaunchKernelGGL((bpd::world_lds_kernel<B, P, R, SC, BC, VC, TC>), dim3(count), dim3(B), 0, stream, arena, descs,
                           stats, order, count, steps, statsStride, notify);
        return hipGetLastError();
    } else {
        return hipSuccess;
    }
}
}  // namespace

extern "C" __attribute__((visibility("hidden"))) hipError_t BP_CAT(bp_launch_lds_cfg_, BP_LDS_CFG)(
    const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats, const uint32_t *order, uint32_t count,
    uint32_t steps, uint32_t statsStride, uint32_t notify, hipStream_t stream) {
    bool done = false;
    hipError_t e = hipSuccess;
#define BP_LDS_ONE(ID, B, P, R, SC, BC, VC, TC) \
    if (!done) e = launch_if<ID, B, P, R, SC, BC, VC, TC>(arena, descs, stats, order, count, steps, statsStride, notify, stream, done);
    BP_LDS_CONFIGS(BP_LDS_ONE)
#undef BP_LDS_ONE
    return done ? e : hipErrorInvalidValue;
}
