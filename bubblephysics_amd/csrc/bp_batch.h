// bp_batch.h -- internal: the batch of worlds that every host entry point drives (the arena,
// the kernel classes, the streams) and the helpers shared by bp_api.cpp and bp_server.cpp.
#ifndef BP_BATCH_H
#define BP_BATCH_H
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bubblephysics.h"
#include "bp_device.h"

extern "C" hipError_t bp_launch_world_step(int block, const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats,
                                           const uint32_t *order, uint32_t count, uint32_t notify, hipStream_t stream);
extern "C" hipError_t bp_launch_world_lds(int id, const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats,
                                          const uint32_t *order, uint32_t count, uint32_t steps, uint32_t statsStride,
                                          uint32_t notify, uint32_t groups, uint32_t *cursor, uint32_t *placed,
                                          hipStream_t stream);
extern "C" hipError_t bp_launch_gate(const uint32_t *placed, uint32_t target, uint32_t maxTicks, hipStream_t stream);
extern "C" int bp_lds_config_count();
extern "C" int bp_lds_config_residency(int id);
extern "C" void bp_lds_config(int id, uint32_t *threads, uint32_t *per, uint32_t *rowcap, uint32_t *caps);
extern "C" uint32_t bp_coop_max_groups();
extern "C" hipError_t bp_launch_gather_world(const bpd::Arena *arena, const bpd::WorldDesc *desc, bp_bubble *out, uint32_t count,
                                             hipStream_t stream);
extern "C" hipError_t bp_launch_gather_world_full(const bpd::Arena *arena, const bpd::WorldDesc *desc, bp_bubble *out,
                                                  uint32_t count, hipStream_t stream);
extern "C" hipError_t bp_launch_multi_copy(const char *blob, const bpd::CopyOp *ops, uint32_t count, hipStream_t stream);
extern "C" hipError_t bp_launch_gather_worlds_full(const bpd::Arena *arena, const bpd::WorldMove *moves, uint32_t count,
                                                   uint32_t maxRecords, hipStream_t stream);
extern "C" hipError_t bp_launch_scatter_worlds(const bpd::Arena *arena, const bpd::WorldMove *moves, uint32_t count,
                                               uint32_t maxRecords, hipStream_t stream);
extern "C" hipError_t bp_launch_scatter_world(const bpd::Arena *arena, uint32_t offset, uint32_t count, const bp_bubble *in,
                                              const int32_t *motorIdx, const int32_t *bandIdx, hipStream_t stream);
extern "C" uint32_t bp_coop_block();
extern "C" hipError_t bp_launch_fill_u32(uint32_t *p, size_t words, uint32_t value, hipStream_t stream);
extern "C" hipError_t bp_launch_drop_removed(const bpd::Arena *arena, bpd::WorldDesc *descs, const bpd::DropRecord *recs,
                                             uint32_t count, int64_t *removedIds, hipStream_t stream);
extern "C" hipError_t bp_launch_audit(const bpd::Arena *arena, const bpd::WorldDesc *descs, uint32_t worlds, void *out,
                                      hipStream_t stream);
extern "C" hipError_t bp_contacts_scratch(void **out, uint32_t *maxBubbles, uint32_t *maxRecords);
extern "C" hipError_t bp_launch_contacts(const bpd::Arena *arena, const bpd::WorldDesc *descs, uint32_t world,
                                         uint32_t *offsets, void *out, uint32_t capacity, uint32_t *total,
                                         hipStream_t stream);
extern "C" hipError_t bp_launch_body_contacts(const bpd::Arena *arena, const bpd::WorldDesc *descs, uint32_t world,
                                              uint32_t *offsets, void *out, uint32_t capacity, uint32_t *total,
                                              hipStream_t stream);
extern "C" hipError_t bp_launch_contours(const bpd::Arena *arena, const bpd::WorldDesc *descs, uint32_t world,
                                         uint32_t bubbles, bp_contour_config cfg, uint32_t stride, float2 *points,
                                         bp_bubble_contour *shapes, hipStream_t stream);
extern "C" uint32_t bp_contour_capacity();
extern "C" hipError_t bp_launch_world_coop(const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats,
                                           const uint32_t *orderEntry, uint32_t world, uint32_t groups, hipStream_t stream);

namespace bph {

inline thread_local std::string g_lastError;

inline void set_error(const std::string &text) { g_lastError = text; }

#define BP_HIP(expr)                                                                                   \
    do {                                                                                               \
        hipError_t bp_e = (expr);                                                                      \
        if (bp_e != hipSuccess) {                                                                      \
            bph::set_error(std::string(#expr) + ": " + hipGetErrorString(bp_e));                         \
            return BP_ERR_COMMAND_FAILED;                                                              \
        }                                                                                              \
    } while (0)

constexpr uint32_t kAdjPerBubble = 64;       // CSR capacity per bubble (grows on overflow)
constexpr int kClassCount = 12;              // ten LDS configurations, the global-state kernel, the cooperative kernel
constexpr int kGlobalClass = 10;
constexpr int kCoopClass = 11;
constexpr uint32_t kLdsLargest = 4096;       // bubbles of the widest LDS configuration
// Geometry every world's device ranges hold at least (when the batch has worlds that large), so
// that a served world can move into an evicted world's ranges.
constexpr uint32_t kReserveSegments = 64, kReserveBodies = 32, kReserveVertices = 256, kReserveTriangles = 256;
constexpr int kStreams = kClassCount;          // one stream per kernel class: classes never queue behind each other

inline uint32_t class_threads(int k) {
    if (k == kGlobalClass) return 1024;
    if (k == kCoopClass) return bp_coop_block();
    uint32_t t, p, r;
    bp_lds_config(k, &t, &p, &r, nullptr);
    return t;
}

inline uint32_t class_capacity(int k) {
    if (k == kGlobalClass || k == kCoopClass) return 0xffffffffu;
    uint32_t t, p, r;
    bp_lds_config(k, &t, &p, &r, nullptr);
    return t * p;
}

// Whether a world of n bubbles with the given geometry fits configuration k (the global-state
// kernel takes every world).
inline bool class_fits(int k, size_t n, size_t segments, size_t bodies, size_t vertices, size_t triangles) {
    if (k == kGlobalClass || k == kCoopClass) return true;
    uint32_t t, p, r, caps[4];
    bp_lds_config(k, &t, &p, &r, caps);
    return n <= (size_t)t * p && segments <= caps[0] && bodies <= caps[1] && vertices <= caps[2] && triangles <= caps[3];
}

inline uint32_t auto_buckets(uint32_t n) {
    uint32_t b = 16;
    while (b < n && b < bpd::kMaxBuckets) b <<= 1;
    return b;
}

struct HostWorld {
    bp_world_config config{};
    std::vector<bp_bubble> bubbles;
    std::vector<bp_segment> segments;
    std::vector<bp_body> bodies;
    std::vector<float> vertices;        // float2 pairs
    std::vector<uint32_t> triangles;    // 4 per triangle
    std::vector<bp_motor> motors;
    std::vector<bp_band> bands;
    std::vector<bp_field> fields;
    std::vector<float> fieldVertices;   // float2 pairs
    std::vector<int64_t> fieldIds;      // filter ids, ascending within each field's range
    uint32_t adjPerBubble = kAdjPerBubble;
    int forcedClass = -1;               // kernel configuration chosen by the tuner, -1 = default
    uint32_t exportAdjacency = 0;
    bool resident = true;               // false: held in host memory only (serving eviction)
    uint32_t capacity = 0;              // bubble slots of its device range (>= bubbles), set by pack
    int packedClass = -1;               // kernel configuration of the last pack (-1 = not on the device)
    bool failed = false;                // failed a step: not advanced again until new input (kept across packs)
    std::vector<int64_t> removedIds;    // bubbles the region removed during the last stepping call
};

inline bool class_fits(int k, const HostWorld &w) {
    return class_fits(k, w.bubbles.size(), w.segments.size(), w.bodies.size(), w.vertices.size() / 2, w.triangles.size() / 4);
}

// Default kernel configuration of a world: the cooperative kernel for worlds above the widest
// LDS configuration (or on request, as the reference's multi-dispatch path), else the narrowest
// LDS configuration whose bubble and geometry capacities hold it (classes 0-9 are the LDS
// configurations of bp_world_lds.hip), else the single-workgroup global-state kernel.
inline int default_class_for(const HostWorld &w, bool cooperative) {
    if (cooperative && (w.config.force_cooperative || w.bubbles.size() > kLdsLargest)) return kCoopClass;
    static const int kBySize[] = {0, 1, 2, 3, 4, 7, 9};
    for (int k : kBySize)
        if (class_fits(k, w)) return k;
    return kGlobalClass;
}

inline uint32_t next_pow2(uint32_t v, uint32_t lowest) {
    uint32_t p = lowest;
    while (p < v && p < (1u << 24)) p <<= 1;
    return p;
}

template <class T> struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;
};

}  // namespace bph

struct bp_batch {
    std::vector<bph::HostWorld> worlds;
    bool packed = false;
    bool deviceValid = false;          // device state holds the published state of every world
    std::vector<void *> allocations;
    bpd::Arena arena{};
    bpd::Arena *arenaDev = nullptr;
    bpd::WorldDesc *descs = nullptr;
    std::vector<bpd::WorldDesc> hostDescs;
    // The lengths of each world's device ranges (set by pack; a world moved into another world's
    // ranges in an in-place swap takes them over), which bound what a world moved in may hold.
    struct Region {
        uint32_t bubbles, segments, bodies, vertices, triangles, motors, bands, fields, fieldVertices, fieldIds, adjacency;
    };
    std::vector<Region> regions;
    uint32_t *orders[bph::kClassCount] = {};
    std::vector<uint32_t> classWorlds[bph::kClassCount];
    std::vector<uint32_t> coopGroups;   // per world: workgroups of its cooperative launch (0 = not cooperative)
    std::vector<uint64_t> workEstimate; // per world: device ticks of its last step (launch order key)
    // Launch order of the classes: the class holding the longest-running world first, so that
    // the world on the critical path of the batch gets a compute unit before the others.
    int classOrder[bph::kClassCount] = {11, 10, 9, 8, 7, 6, 5, 4, 3, 2, 1, 0};
    bool orderDirty = false;            // estimates changed since the class order lists were uploaded
    bool estimatesMeasured = false;     // workEstimate holds device ticks (after a whole-batch step)
    // Workgroup pools: a class launched with fewer workgroups than worlds, each workgroup taking
    // the next world of the class's longest-first order when it finishes one. The pools divide
    // the CUs among the classes in proportion to their measured work, so that the large
    // workgroups, which need a whole CU, never wait behind small ones scattered over the CUs.
    bool pooling = true;
    uint32_t *cursors = nullptr;        // per class: the next-world word of a pooled launch; then the
                                        // count of pooled workgroups that have started
    uint32_t poolGroups[bph::kClassCount] = {};   // per class: workgroups of a pooled launch (0 = one per world)
    int residency[bph::kClassCount] = {};         // per class: workgroups one CU holds
    int deviceCUs = 0;
    hipStream_t streams[bph::kStreams] = {};
    hipEvent_t startEvent = nullptr, endEvents[bph::kStreams] = {};
    hipEvent_t gateEvent = nullptr;     // the gate of the classes launched after the pools
    hipEvent_t classStart[bph::kClassCount] = {}, classEnd[bph::kClassCount] = {};
    bpd::WorldStatus *stats = nullptr;
    size_t statsCapacity = 0;          // in WorldStatus records
    std::vector<bpd::WorldStatus> hostStats;
    float lastGpuMs = 0.0f;
    float lastClassMs[bph::kClassCount] = {};
    bool inFlight = false;
    bool serving = false;              // a bp_server drives the batch; its own entry points refuse
    bool pending = false;              // launches queued by launch() and not yet collected
    uint32_t pendingSteps = 0;
    std::vector<uint32_t> pendingList; // worlds of the pending launches, in the caller's order
    std::vector<uint32_t> subsetHost;  // per class, the listed worlds (longest first)
    uint32_t *subsetOrders = nullptr;  // device copy of subsetHost
    bp_bubble *xferBubbles = nullptr;  // one world's records in transit (transactions)
    int32_t *xferIndices = nullptr;    // their motor and band indices
    uint32_t xferCapacity = 0;
    bp_bubble *pinnedRecs = nullptr;   // pinned host staging of the records of a swap round
    void *pinnedBlobDev = nullptr;       // device address of pinnedBlob
    int32_t *pinnedIdx = nullptr;
    uint32_t pinnedCapacity = 0;
    void *pinnedBlob = nullptr;        // pinned host staging of one world's geometry
    size_t pinnedBlobBytes = 0;
    char *deviceBlob = nullptr;        // its device copy (released with the arena)
    size_t deviceBlobBytes = 0;
    uint32_t launchOffset[bph::kClassCount] = {}, launchCount[bph::kClassCount] = {};
    uint32_t totalBubbles = 0;
    uint32_t totalBodies = 0;          // body records of the arena (every world's body range)
    std::mutex mutex;
    // Captured step chain (bp_batch_capture): the launches of `graphSteps` steps of every world.
    // One graph per kernel class, each replayed on the class's own stream (so the classes run
    // concurrently, as direct launches do); the cooperative class's graph runs last on stream 0.
    hipGraph_t graphs[bph::kClassCount] = {};
    hipGraphExec_t graphExecs[bph::kClassCount] = {};
    bool graphPooled = false;
    uint32_t graphSteps = 0;
    hipEvent_t forkEvent = nullptr;     // replay(): the classes' streams wait for the cleared records
    // Region removals: a world whose step removed bubbles is held on the device until they are
    // deleted. A batch deletes them and runs the rest of the call's steps (the public world's
    // bookkeeping, NativePublicWorld.swift:264-267); a plain world (bp_world) keeps them in its
    // published frame, as the reference's solver does, and leaves the deletion to the caller.
    bool dropOnHold = true;
    bool settling = false;
    bpd::DropRecord *dropRecs = nullptr;
    int64_t *dropIds = nullptr;
    size_t dropCapacity = 0, dropIdCapacity = 0;

    ~bp_batch() {
        release();
        destroy_streams();
        if (pinnedRecs) (void)hipHostFree(pinnedRecs);
        if (pinnedIdx) (void)hipHostFree(pinnedIdx);
        if (pinnedBlob) (void)hipHostFree(pinnedBlob);
        free_swap_round();
    }

    void release_graph() {
        for (int k = 0; k < bph::kClassCount; ++k) {
            if (graphExecs[k]) (void)hipGraphExecDestroy(graphExecs[k]);
            if (graphs[k]) (void)hipGraphDestroy(graphs[k]);
            graphExecs[k] = nullptr; graphs[k] = nullptr;
        }
        graphSteps = 0;
        graphPooled = false;
    }

    // Streams and events live as long as the batch (a repack keeps them, so a server that launches
    // on the batch's streams keeps valid handles across its repacks).
    bp_status create_streams() {
        for (int k = 0; k < bph::kStreams; ++k) {
            if (!streams[k]) BP_HIP(hipStreamCreateWithFlags(&streams[k], hipStreamNonBlocking));
            if (!endEvents[k]) BP_HIP(hipEventCreate(&endEvents[k]));
        }
        for (int k = 0; k < bph::kClassCount; ++k) {
            if (!classStart[k]) BP_HIP(hipEventCreate(&classStart[k]));
            if (!classEnd[k]) BP_HIP(hipEventCreate(&classEnd[k]));
        }
        if (!startEvent) BP_HIP(hipEventCreate(&startEvent));
        if (!gateEvent) BP_HIP(hipEventCreateWithFlags(&gateEvent, hipEventDisableTiming));
        return BP_OK;
    }
    void destroy_streams() {
        for (int k = 0; k < bph::kStreams; ++k) {
            if (streams[k]) (void)hipStreamDestroy(streams[k]);
            if (endEvents[k]) (void)hipEventDestroy(endEvents[k]);
            streams[k] = nullptr; endEvents[k] = nullptr;
        }
        for (int k = 0; k < bph::kClassCount; ++k) {
            if (classStart[k]) (void)hipEventDestroy(classStart[k]);
            if (classEnd[k]) (void)hipEventDestroy(classEnd[k]);
            classStart[k] = classEnd[k] = nullptr;
        }
        if (startEvent) (void)hipEventDestroy(startEvent);
        if (gateEvent) (void)hipEventDestroy(gateEvent);
        gateEvent = nullptr;
        startEvent = nullptr;
    }

    void release() {
        release_graph();
        if (forkEvent) (void)hipEventDestroy(forkEvent);
        forkEvent = nullptr;
        for (void *p : allocations) (void)hipFree(p);
        allocations.clear();
        stats = nullptr; statsCapacity = 0; descs = nullptr; arenaDev = nullptr; subsetOrders = nullptr; cursors = nullptr;
        xferBubbles = nullptr; xferIndices = nullptr; xferCapacity = 0;
        deviceBlob = nullptr; deviceBlobBytes = 0;
        dropRecs = nullptr; dropIds = nullptr; dropCapacity = dropIdCapacity = 0;
        pending = false;
        for (auto &o : orders) o = nullptr;
        packed = false;
    }

    template <class T> bp_status alloc(T *&out, size_t count) {
        void *p = nullptr;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) { bph::set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return BP_ERR_ALLOCATION_FAILED; }
        allocations.push_back(p);
        out = static_cast<T *>(p);
        return BP_OK;
    }
    template <class T> bp_status upload(T *dst, const std::vector<T> &src) {
        if (src.empty()) return BP_OK;
        BP_HIP(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
        return BP_OK;
    }

    bp_status pack();
    // Copies the published state of every resident world into its host mirror; `full` also
    // copies the properties and the device's bubble count (transactions applied on the device).
    bp_status download_state(bool full = false);
    bp_status download_world(uint32_t wi);
    bp_status write_bubbles(uint32_t wi, const bp_bubble *bubbles, uint32_t count);
    void motor_band_indices(uint32_t wi, const bp_bubble *bubbles, uint32_t count, int32_t *indices) const;
    void written_desc(uint32_t wi, uint32_t count, bpd::WorldDesc &d) const;
    bp_status adopt_written(uint32_t wi, const bp_bubble *bubbles, uint32_t count, const bpd::WorldDesc &d);
    uint64_t footprint(uint32_t wi) const;
    // Serving residency without a repack: whether the held world `in` fits the device ranges of
    // the resident world `out`, and the exchange (out's published state is read back into its
    // host copy, in's host copy is written into the ranges). Both worlds must be idle.
    // A swap round exchanges the pairs (outs[i], ins[i]) (all of them distinct worlds), with one
    // read-back and one upload for as many pairs as the transfer buffer holds; results[i] ends
    // BP_OK, BP_ERR_CAPACITY_OVERFLOW (the world does not fit) or BP_ERR_UNAVAILABLE (the
    // resident world holds region-removed bubbles, or the pair waits for a later round), and the
    // return values are device or allocation errors.
    bool fits_ranges_of(uint32_t out, uint32_t in) const;
    // mirrors (null, or one per pair; capacity 0 for none) are the serving mirror ranges the
    // descriptor of a world moved in points to.
    // A round runs in three calls, each after the stream's work of the previous one has finished:
    // swap_begin queues the read-back (results[i]: BP_ERR_UNAVAILABLE also for a pair left for a
    // later round because the transfer buffer is full), swap_upload fills the host copies of the
    // worlds leaving and queues the upload of the worlds moving in (swapUploads of them; none
    // when every resident world of the round is held), swap_finish records the new residency.
    struct MirrorRange { void *mirror; float4 *bodyMirror; int32_t *mirrorIdx; uint32_t capacity; };
    struct SwapPair;
    struct SwapRound;
    SwapRound *swapRound = nullptr;
    size_t swapReadAt = 0;
    uint32_t swapRecords = 0, swapUploads = 0;
    SwapRound &swap_round();
    void free_swap_round();
    bp_status ensure_pinned_blob(size_t bytes);
    bp_status swap_begin(const uint32_t *outs, const uint32_t *ins, uint32_t count, const MirrorRange *mirrors,
                         bp_status *results, hipStream_t stream);
    bp_status swap_upload(bp_status *results, hipStream_t stream);
    void swap_finish();
    bp_status launch(const uint32_t *subset, uint32_t subsetCount, uint32_t steps);
    void plan_pools();
    bp_status refresh_order();
    bp_status launch_classes(bool pooled, uint32_t steps, bool subset, bool timed);
    bp_status finish(bp_frame_info *infos, bool wait);
    bp_status step(uint32_t steps, bp_frame_info *infos);
    bp_status audit(bp_audit *out, uint32_t count);
    bp_status contacts(uint32_t world, bp_contact *out, uint32_t capacity, uint32_t *count, bool bodies = false);
    bp_status contour_stride(uint32_t world, const bp_contour_config *cfg, uint32_t *stride);
    bp_status contours(uint32_t world, const bp_contour_config *cfg, uint32_t stride, float *points, bp_bubble_contour *shapes,
                       uint32_t count);
    bp_status capture(uint32_t steps);
    bp_status replay(bp_frame_info *infos);
    bp_status collect(uint32_t steps, bp_frame_info *infos, const std::vector<uint32_t> &listed, float ms);
    bp_status settle_held(uint32_t steps, bp_frame_info *infos, const std::vector<uint32_t> &listed);
    bp_status drop_removed(const std::vector<uint32_t> &held, const std::vector<uint32_t> &removals);
    bp_status region_results(uint32_t wi, int64_t *removed, uint32_t removedCapacity, uint32_t *removedCount,
                             int64_t *wrapped, uint32_t wrappedCapacity, uint32_t *wrappedCount);
    bp_status remove_bubbles(uint32_t wi, const int64_t *ids, uint32_t count);
};

#endif  // BP_BATCH_H
