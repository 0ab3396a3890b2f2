// bp_world_lds.hip -- configuration table and launch dispatch of the LDS-resident world kernel.
// Each configuration is compiled in its own translation unit (bp_lds_cfg.hip built once per
// configuration), so the instantiations build in parallel.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "bp_device.h"
#include "bp_lds_configs.h"

#define BP_LDS_DECLARE(ID, B, P, R, SC, BC, VC, TC)                                                                 \
    extern "C" hipError_t bp_launch_lds_cfg_##ID(const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats, \
                                                 const uint32_t *order, uint32_t count, uint32_t steps,         \
                                                 uint32_t statsStride, uint32_t notify, uint32_t groups,        \
                                                 uint32_t *cursor, uint32_t *placed, hipStream_t stream);        \
    extern "C" int bp_residency_lds_cfg_##ID();
BP_LDS_CONFIGS(BP_LDS_DECLARE)
#undef BP_LDS_DECLARE

extern "C" __attribute__((visibility("hidden"))) int bp_lds_config_count() { return 10; }

// caps[0..3] = segment, body, vertex and triangle capacities.
extern "C" __attribute__((visibility("hidden"))) void bp_lds_config(int id, uint32_t *threads, uint32_t *per, uint32_t *rowcap,
                                                                   uint32_t *caps) {
#define BP_LDS_DESCRIBE(ID, B, P, R, SC, BC, VC, TC)                               \
    if (id == ID) {                                                               \
        *threads = B; *per = P; *rowcap = R;                                      \
        if (caps) { caps[0] = SC; caps[1] = BC; caps[2] = VC; caps[3] = TC; }      \
        return;                                                                   \
    }
    BP_LDS_CONFIGS(BP_LDS_DESCRIBE)
#undef BP_LDS_DESCRIBE
    *threads = 0; *per = 0; *rowcap = 0;
    if (caps) caps[0] = caps[1] = caps[2] = caps[3] = 0;
}

// Workgroups of configuration `id` one CU holds at once (0 when unknown).
extern "C" __attribute__((visibility("hidden"))) int bp_lds_config_residency(int id) {
#define BP_LDS_RESIDENCY(ID, B, P, R, SC, BC, VC, TC) \
    if (id == ID) return bp_residency_lds_cfg_##ID();
    BP_LDS_CONFIGS(BP_LDS_RESIDENCY)
#undef BP_LDS_RESIDENCY
    return 0;
}

// `steps` steps of every world in one launch; the status of step k of world w is stats[k * statsStride + w].
// With `notify`, each status record is marked complete with a system-scope release (post_status).
// With `groups` below `count` and a zeroed `cursor` word, `groups` workgroups share the worlds and
// count themselves in `placed` (may be null) as they start.
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_world_lds(int id, const bpd::Arena *arena, bpd::WorldDesc *descs,
                                                                           bpd::WorldStatus *stats, const uint32_t *order,
                                                                           uint32_t count, uint32_t steps, uint32_t statsStride,
                                                                           uint32_t notify, uint32_t groups, uint32_t *cursor,
                                                                           uint32_t *placed, hipStream_t stream) {
    if (count == 0) return hipSuccess;
#define BP_LDS_LAUNCH(ID, B, P, R, SC, BC, VC, TC) \
    if (id == ID) return bp_launch_lds_cfg_##ID(arena, descs, stats, order, count, steps, statsStride, notify, groups, cursor, placed, \
                                          stream);
    BP_LDS_CONFIGS(BP_LDS_LAUNCH)
#undef BP_LDS_LAUNCH
    return hipErrorInvalidValue;
}
