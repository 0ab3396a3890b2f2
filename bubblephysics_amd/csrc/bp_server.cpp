// bp_server.cpp -- native request loop over the worlds of a batch (continuous batching).
//
// One dispatcher thread owns every device call of the server. Requests (one step of one world,
// optionally after a transaction) queue per world. A world with a waiting request and nothing
// on the device is "ready" in the kernel class its bubble count selects. Whenever a class has
// ready worlds, fewer than `max_inflight_per_class` launches in flight and a stream is idle, one
// launch takes all of them: the transaction kernel, the class's step kernel (both read the world
// list and the transaction records from pinned host memory), then an event. Each launch takes a
// stream of its own, so different classes and consecutive launches of one class run
// concurrently, and requests that arrive while others
// are on the device join the next launch of their class. Each world's result is posted as soon
// as its own workgroup has finished: the step kernels write the result record into pinned host
// memory and mark it complete with a system-scope release, and the dispatcher polls the marks.
// A world takes part in a new launch only once its previous launch has completed (its event),
// so a step always reads the state that the previous step published.
//
// Residency: worlds are admitted against a device-memory budget in index order. A request for a
// world held in host memory takes the device ranges of an idle world whose ranges hold it (the
// one with the fewest bubble slots, least recently used among equals; its state is read back to
// host memory first), with no other world moved and nothing drained; when no such world exists,
// the request waits until the launches in flight drain, then the least recently used idle worlds
// are evicted and the batch is repacked.
//
// The reference has no counterpart: its game loop steps one world per call
// (BubbleWorld.submit/advance, Sources/BubblePhysicsPublic/BubbleWorld.swift:252-283).
#include "bp_batch.h"

#include <atomic>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

using namespace bph;

extern "C" hipError_t bp_launch_transactions(const bpd::Arena *arena, bpd::WorldDesc *descs, const bpd::TxRecord *txs,
                                             uint32_t count, bpd::WorldStatus *stats, hipStream_t stream);
extern "C" hipError_t bp_launch_gather_full_world(const bpd::Arena *arena, const bpd::WorldDesc *desc, bp_bubble *out,
                                                  uint32_t count, hipStream_t stream);

namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
    bp_request req;
    double arrival;          // server clock
    int64_t sink;            // index into the trace results, or -1 for the poll queue
    bool staged = false;     // the transaction was decided on the host (applied or refused)
    uint32_t txResult = 0;   // its outcome (BP_REJECT_*) when staged
    bool replace = false;    // staged on the mirror: the launch replaces the world's records
    uint32_t replaceCount = 0;
};

struct Launch {
    int cls = -1;
    int stream = -1;
    hipEvent_t done = nullptr;
    uint32_t count = 0, txCount = 0;
    std::vector<uint32_t> worlds;
    std::vector<Pending> reqs;
    std::vector<uint32_t> bubbles;
    std::vector<uint8_t> posted;                              // result already handed to the caller
    uint32_t postedCount = 0;
    // World list and transaction records in coherent pinned memory, read by the kernels through
    // their device addresses (no copy per launch).
    uint32_t *hOrder = nullptr, *dOrder = nullptr;
    bpd::TxRecord *hTx = nullptr, *dTx = nullptr;
};

// An add, merge or split being staged on the host: the world's published records are read back
// (phase 1), the reference's staging runs on the host, and the staged list is written back in
// place (phase 2), each phase ending with an event the dispatcher polls, so a staging world never
// holds up the launches of the others. Buffers grow with the largest world staged.
struct StageJob {
    uint32_t world = 0;
    int phase = 0;                      // 0 free, 1 reading the world, 2 writing the staged list
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bp_bubble *dRecs = nullptr, *hRecs = nullptr;        // device and pinned host records
    int32_t *dIdx = nullptr, *hIdx = nullptr;            // motor and band indices of the staged list
    bpd::WorldDesc *hDesc = nullptr;                     // pinned host descriptor
    float4 *hPose = nullptr;                             // body poses of both published sides
    uint32_t capacity = 0, poseCapacity = 0;
    std::vector<bp_bubble> staged;
};

constexpr int kMaxLaunches = 64;
constexpr int kStageJobs = 16;
constexpr double kUrgentWait = 0.02;             // seconds after which a waiting world may displace a world whose requests only wait
constexpr double kRepackDelay = 0.25;            // seconds a world waits for in-place admission before a repack
constexpr double kAdmissionInterval = 5e-4;      // seconds between in-place admission rounds (a round's exchanges share one round trip)
// Classes a served world may run in: the LDS configurations and the one-workgroup global kernel
// (a cooperative grid needs the whole device and would stall every other launch).
const int kServeClasses[] = {0, 1, 2, 3, 4, 7, 9};

}  // namespace

struct bp_server {
    bp_batch *batch = nullptr;
    bp_server_config cfg{};
    Clock::time_point t0;
    int device = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Pending> incoming;               // submitted, not yet taken by the dispatcher
    std::vector<bp_server_result> results;      // completed, not yet polled
    bp_server_result *sinkResults = nullptr;    // results of the trace being replayed
    uint64_t submitted = 0, completed = 0;      // guarded by mu
    bool stop = false;
    bp_status fatal = BP_OK;
    std::string fatalText;
    bp_server_stats st{};                       // published counters, guarded by mu
    bp_server_stats local{};                    // the dispatcher's own counters, published under mu

    // Dispatcher state.
    uint32_t W = 0;
    std::vector<std::deque<Pending>> queue;
    std::vector<uint8_t> busy, queued, resident, wantAdmission;
    std::vector<uint8_t> moving;                // worlds of the in-place exchanges in flight
    struct SwapRoundState {
        int phase = 0;                          // 0: none in flight, 1: read-back queued, 2: upload queued
        std::vector<uint32_t> outs, ins;
        std::vector<uint8_t> outIdle;
        std::vector<bp_batch::MirrorRange> mirrors;
        std::vector<bp_status> results;
    } swapState;
    hipEvent_t swapDone = nullptr;
    std::vector<uint8_t> worldFailed;           // a step failed: the world takes no more steps or transactions
    std::vector<uint32_t> dropPending;          // bubbles the world's last step removed from the region
    bpd::DropRecord *dropDev = nullptr;
    std::vector<uint32_t> count, headCount;
    std::vector<int> headClass;
    std::vector<uint64_t> lastUse, footprintOf;
    uint64_t useClock = 0, used = 0;
    std::deque<uint32_t> ready[kClassCount];
    std::vector<uint32_t> admissions;
    std::vector<double> waitingSince;           // per world: when it began to wait for admission
    double lastAdmissionRound = -1.0;           // server clock of the last in-place admission round
    bool admissionRetry = true;                 // worlds became idle or began to wait since the last attempt
    bool launchAfterAdmission = false;          // the worlds ready after an admission round launch first
    int inflightPerClass[kClassCount] = {};
    Launch launches[kMaxLaunches];
    std::vector<int> freeLaunch, inflight;
    std::vector<hipStream_t> streams;
    std::vector<int> streamLoad;                // launches in flight per stream (0 or 1)
    int idleStreams = 0;
    bpd::WorldStatus *hostStats = nullptr;      // one record per world in pinned host memory (each world is
    bpd::WorldStatus *stats = nullptr;          // in one launch at a time); `stats` is its device address
    size_t ownedFrom = 0;                       // streams[ownedFrom..] were created by the server
    hipStream_t readStream = nullptr;
    hipStream_t stageStream = nullptr;          // the staging jobs' copies and kernels
    bp_bubble *readBuffer = nullptr;            // bp_server_read_bubbles
    uint32_t readCapacity = 0;
    StageJob stageJobs[kStageJobs];
    std::vector<int> freeStage, activeStage;
    std::deque<uint32_t> stageWaiting;          // worlds whose staging waits for a free job
    std::vector<uint8_t> staging;               // per world: its head transaction is being staged
    // Mirrors of the worlds' published records and body poses in pinned, mapped host memory: the
    // step kernels write them after every published step, so an add, merge or split is staged on
    // the host from the mirror, written into the mirror and applied by the launch's transaction
    // kernel, with no device round trip of its own. A world whose mirror is not current (after a
    // repack deleted region-removed bubbles, or beyond its mirror's slots) is staged through the
    // device instead (begin_stage).
    bp_bubble *mirrorHost = nullptr, *mirrorDev = nullptr;
    int32_t *mirrorIdxHost = nullptr, *mirrorIdxDev = nullptr;
    float4 *bodyMirrorHost = nullptr, *bodyMirrorDev = nullptr;
    std::vector<uint64_t> mirrorOffset, bodyMirrorOffset;
    std::vector<uint32_t> mirrorCap;
    std::vector<uint8_t> mirrorValid, mirrorAttached;
    std::vector<void *> deviceAllocs, hostAllocs;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    bp_status init();
    void release();
    void run();
    void enqueue(const Pending &p);
    void prepare_head(uint32_t w);
    void finish_request(const Pending &p, bp_status status, uint32_t substeps, uint32_t bubbles, uint32_t launchWorlds,
                        int cls, float deviceMs, uint32_t txResult = 0);
    bp_status attach_mirrors();
    void detach_mirrors();
    bp_status stage_from_mirror(uint32_t w, Pending &p);
    bp_status stage_scene(uint32_t w, const Pending &p, const bp_bubble *recs, uint32_t n, const float4 *poses,
                          std::vector<bp_bubble> &out, uint32_t &outCount, bp_stage_rejection &why);
    bp_status begin_stage(uint32_t w);
    bp_status stage_on_host(StageJob &job);
    bool poll_stages();
    void free_stage_job(int j);
    bp_status drop_removed(const std::vector<uint32_t> &ws);
    bool poll_launches();
    bool form_launches();
    bp_status start_launch(int cls);
    bp_status admit_waiting();
    bp_status start_swap_round();
    bp_status advance_swap_round(bool wait);
    bp_status finish_swap_round();
    int class_for(uint32_t w, uint32_t n) const;
    uint32_t buckets_for(uint32_t w, uint32_t n) const;
    void fail(bp_status s);
    void publish_counters() {          // with mu held
        st.launches = local.launches; st.launched_worlds = local.launched_worlds; st.transactions = local.transactions;
        st.rejected = local.rejected; st.evictions = local.evictions; st.admissions = local.admissions;
        st.repacks = local.repacks; st.resident_bytes = local.resident_bytes;
        st.admission_microseconds = local.admission_microseconds;
    }
};

namespace {

template <class T> bp_status dev_alloc(bp_server *s, T *&out, size_t count) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) { set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return BP_ERR_ALLOCATION_FAILED; }
    s->deviceAllocs.push_back(p);
    out = static_cast<T *>(p);
    return BP_OK;
}

// Coherent pinned host memory and its device address.
template <class T> bp_status host_alloc(bp_server *s, T *&out, T *&device, size_t count) {
    void *p = nullptr, *dp = nullptr;
    hipError_t e = hipHostMalloc(&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) { set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e)); return BP_ERR_ALLOCATION_FAILED; }
    s->hostAllocs.push_back(p);
    e = hipHostGetDevicePointer(&dp, p, 0);
    if (e != hipSuccess) { set_error(std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e)); return BP_ERR_ALLOCATION_FAILED; }
    out = static_cast<T *>(p);
    device = static_cast<T *>(dp);
    return BP_OK;
}

}  // namespace

int bp_server::class_for(uint32_t w, uint32_t n) const {
    const HostWorld &hw = batch->worlds[w];
    const size_t ns = hw.segments.size(), nb = hw.bodies.size(), nv = hw.vertices.size() / 2, nt = hw.triangles.size() / 4;
    if (hw.forcedClass >= 0 && hw.forcedClass != kCoopClass && class_fits(hw.forcedClass, n, ns, nb, nv, nt)) return hw.forcedClass;
    if (n > cfg.big_world) return kGlobalClass;
    for (int k : kServeClasses)
        if (class_fits(k, n, ns, nb, nv, nt)) return k;
    return kGlobalClass;
}

uint32_t bp_server::buckets_for(uint32_t w, uint32_t n) const {
    const bp_world_config &c = batch->worlds[w].config;
    return c.grid_bucket_count == 0 ? auto_buckets(n) : std::min(c.grid_bucket_count, bpd::kMaxBuckets);
}

void bp_server::fail(bp_status s) {
    std::lock_guard<std::mutex> lock(mu);
    if (fatal == BP_OK) { fatal = s; fatalText = g_lastError; }
    cvDone.notify_all();
}

bp_status bp_server::init() {
    bp_batch &b = *batch;
    bp_status s = BP_OK;
    W = (uint32_t)b.worlds.size();
    // Residency against the budget, in index order; the rest wait in host memory.
    resident.assign(W, 0);
    footprintOf.assign(W, 0);
    bool changed = !b.packed;
    for (uint32_t w = 0; w < W; ++w) {
        footprintOf[w] = b.footprint(w);
        const bool fits = cfg.memory_budget == 0 || used + footprintOf[w] <= cfg.memory_budget;
        if (fits) { resident[w] = 1; used += footprintOf[w]; }
        if (b.worlds[w].resident != (fits != 0)) {
            if (!changed) {
                bp_status s = b.download_state();
                if (s != BP_OK) return s;
            }
            b.worlds[w].resident = fits;
            changed = true;
        }
    }
    if (changed || !b.packed) {
        b.packed = false;
        if ((s = b.pack()) != BP_OK) return s;
    } else if ((s = b.download_state(true)) != BP_OK) {    // the host mirrors seed the serving mirrors
        return s;
    }
    for (uint32_t w = 0; w < W; ++w) if (resident[w]) footprintOf[w] = b.footprint(w);
    {
        // Mirror slots per world: its device slots, or the slots a pack would give it.
        mirrorOffset.assign(W, 0); bodyMirrorOffset.assign(W, 0); mirrorCap.assign(W, 0); mirrorValid.assign(W, 0);
        mirrorAttached.assign(W, 0);
        uint64_t records = 0, poses = 0;
        for (uint32_t w = 0; w < W; ++w) {
            const uint32_t n = (uint32_t)b.worlds[w].bubbles.size();
            const uint32_t cap = std::max<uint32_t>(resident[w] ? b.hostDescs[w].bubbleCapacity : 0u, n + std::max<uint32_t>(8, n / 8));
            mirrorOffset[w] = records; mirrorCap[w] = cap; records += cap;
            bodyMirrorOffset[w] = poses; poses += b.worlds[w].bodies.size();
        }
        if ((s = host_alloc(this, mirrorHost, mirrorDev, records)) != BP_OK ||
            (s = host_alloc(this, mirrorIdxHost, mirrorIdxDev, 2 * records)) != BP_OK ||
            (s = host_alloc(this, bodyMirrorHost, bodyMirrorDev, std::max<uint64_t>(poses, 1))) != BP_OK)
            return s;
    }
    queue.assign(W, {});
    busy.assign(W, 0); queued.assign(W, 0); wantAdmission.assign(W, 0); moving.assign(W, 0); waitingSince.assign(W, 0.0);
    worldFailed.assign(W, 0); dropPending.assign(W, 0);
    for (uint32_t w = 0; w < W; ++w) worldFailed[w] = b.worlds[w].failed ? 1 : 0;
    count.resize(W); headCount.assign(W, 0); headClass.assign(W, -1); lastUse.assign(W, 0);
    for (uint32_t w = 0; w < W; ++w) count[w] = (uint32_t)b.worlds[w].bubbles.size();
    if ((s = attach_mirrors()) != BP_OK) return s;
    {
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, std::max<size_t>(W, 1) * sizeof(bpd::WorldStatus), hipHostMallocCoherent | hipHostMallocMapped);
        if (e != hipSuccess) { set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e)); return BP_ERR_ALLOCATION_FAILED; }
        hostAllocs.push_back(p);
        hostStats = static_cast<bpd::WorldStatus *>(p);
        std::memset(hostStats, 0, std::max<size_t>(W, 1) * sizeof(bpd::WorldStatus));
        void *dp = nullptr;
        BP_HIP(hipHostGetDevicePointer(&dp, p, 0));
        stats = static_cast<bpd::WorldStatus *>(dp);
    }
    const uint32_t cap = std::max<uint32_t>(1, std::min(cfg.max_worlds_per_launch, W));
    for (int k = 0; k < kMaxLaunches; ++k) {
        Launch &L = launches[k];
        if ((s = host_alloc(this, L.hOrder, L.dOrder, cap)) != BP_OK || (s = host_alloc(this, L.hTx, L.dTx, cap)) != BP_OK)
            return s;
        BP_HIP(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
        L.worlds.reserve(cap); L.reqs.reserve(cap); L.bubbles.reserve(cap); L.posted.reserve(cap);
        freeLaunch.push_back(k);
    }
    if ((s = dev_alloc(this, dropDev, std::max<uint32_t>(W, 1))) != BP_OK) return s;
    // Streams: `cfg.streams` in all, each on a hardware queue of its own when the process has that
    // many (GPU_MAX_HW_QUEUES): two for reads and staging, the rest for launches, starting with the
    // batch's own streams (idle while the server drives the batch). Two streams on one hardware
    // queue run in order, so a staging copy would otherwise wait behind a long step launch.
    const uint32_t total = std::max<uint32_t>(3, cfg.streams), launchStreams = total - 2;
    streams.clear();
    for (int k = 0; k < kStreams && streams.size() < launchStreams; ++k) streams.push_back(b.streams[k]);
    ownedFrom = streams.size();
    while (streams.size() < launchStreams) {
        hipStream_t st_ = nullptr;
        BP_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
        streams.push_back(st_);
    }
    streamLoad.assign(streams.size(), 0);
    idleStreams = (int)streams.size();
    BP_HIP(hipStreamCreateWithFlags(&readStream, hipStreamNonBlocking));
    BP_HIP(hipStreamCreateWithFlags(&stageStream, hipStreamNonBlocking));
    BP_HIP(hipEventCreateWithFlags(&swapDone, hipEventDisableTiming));
    staging.assign(W, 0);
    for (int j = 0; j < kStageJobs; ++j) {
        stageJobs[j].stream = stageStream;
        BP_HIP(hipEventCreateWithFlags(&stageJobs[j].done, hipEventDisableTiming));
        freeStage.push_back(j);
    }
    local.resident_bytes = used;
    st.resident_bytes = used;
    return BP_OK;
}

void bp_server::release() {
    detach_mirrors();
    for (int k = 0; k < kMaxLaunches; ++k)
        if (launches[k].done) { (void)hipEventDestroy(launches[k].done); launches[k].done = nullptr; }
    for (size_t k = ownedFrom; k < streams.size(); ++k) if (streams[k]) (void)hipStreamDestroy(streams[k]);
    streams.clear();
    if (readStream) (void)hipStreamDestroy(readStream);
    readStream = nullptr;
    if (stageStream) (void)hipStreamDestroy(stageStream);
    if (swapDone) (void)hipEventDestroy(swapDone);
    swapDone = nullptr;
    stageStream = nullptr;
    for (StageJob &j : stageJobs) {
        if (j.done) (void)hipEventDestroy(j.done);
        if (j.dRecs) (void)hipFree(j.dRecs);
        if (j.dIdx) (void)hipFree(j.dIdx);
        if (j.hRecs) (void)hipHostFree(j.hRecs);
        if (j.hIdx) (void)hipHostFree(j.hIdx);
        if (j.hDesc) (void)hipHostFree(j.hDesc);
        if (j.hPose) (void)hipHostFree(j.hPose);
        j = StageJob{};
    }
    for (void *p : deviceAllocs) (void)hipFree(p);
    for (void *p : hostAllocs) (void)hipHostFree(p);
    deviceAllocs.clear(); hostAllocs.clear();
}

// A new request: it waits behind the world's earlier requests.
void bp_server::enqueue(const Pending &p) {
    const uint32_t w = p.req.world;
    if (w >= W) {
        set_error("request names an unknown world");
        finish_request(p, BP_ERR_INVALID_ARGUMENT, 0, 0, 0, -1, 0.0f);
        return;
    }
    queue[w].push_back(p);
    if (!busy[w] && !queued[w] && !staging[w]) prepare_head(w);
}

// Makes the world's oldest request ready: decides the transaction's effect on the bubble count
// (rejecting one that needs more slots than the world holds) and the class of the step.
void bp_server::prepare_head(uint32_t w) {
    if (moving[w]) return;                      // the round that moves it calls back here
    while (!queue[w].empty()) {
        if (!resident[w]) {
            if (!wantAdmission[w]) { wantAdmission[w] = 1; admissions.push_back(w); waitingSince[w] = now(); admissionRetry = true; }
            return;
        }
        Pending &p = queue[w].front();
        const uint32_t n = count[w];
        if (worldFailed[w]) {
            // worldFailed is final (BubbleWorld.swift:296-299, 424-430): no transaction, no step.
            Pending done = p; queue[w].pop_front();
            finish_request(done, BP_ERR_UNAVAILABLE, 0, n, 0, -1, 0.0f);
            continue;
        }
        if (p.req.transaction > BP_TX_SPLIT) {
            set_error("unknown transaction kind");
            Pending done = p; queue[w].pop_front();
            finish_request(done, BP_ERR_INVALID_ARGUMENT, 0, n, 0, -1, 0.0f);
            continue;
        }
        uint32_t next = n;
        const uint32_t kind = p.req.transaction;
        if ((kind == BP_TX_ADD || kind == BP_TX_MERGE || kind == BP_TX_SPLIT) && !p.staged && n > 0) {
            // Structural transactions are staged while the world is idle: its published state is
            // final, and the staged list is written back before the world's next launch. The world
            // comes back here when its staging has finished (poll_stages).
            if (mirrorValid[w]) {
                bp_status s = stage_from_mirror(w, p);
                if (s == BP_ERR_CAPACITY_OVERFLOW) {
                    Pending done = p; queue[w].pop_front();
                    ++local.rejected;
                    finish_request(done, BP_ERR_CAPACITY_OVERFLOW, 0, n, 0, -1, 0.0f);
                    continue;
                }
                if (s != BP_OK) { fail(s); return; }
            } else {
                if (!staging[w]) {
                    staging[w] = 1;
                    bp_status s = begin_stage(w);
                    if (s != BP_OK) { fail(s); return; }
                }
                return;
            }
        }
        if (p.replace) {
            next = p.replaceCount;
        } else if (kind == BP_TX_ADD || kind == BP_TX_MERGE || kind == BP_TX_SPLIT) {
            next = count[w];
        } else if (kind == BP_TX_REMOVE) {
            next = n > 1 ? n - 1 : n;
        }
        headCount[w] = next;
        headClass[w] = class_for(w, next);
        ready[headClass[w]].push_back(w);
        queued[w] = 1;
        return;
    }
}

void bp_server::finish_request(const Pending &p, bp_status status, uint32_t substeps, uint32_t bubbles, uint32_t launchWorlds,
                               int cls, float deviceMs, uint32_t txResult) {
    bp_server_result r{};
    r.tag = p.req.tag;
    r.world = p.req.world;
    r.status = status;
    r.substeps = substeps;
    r.bubbles = bubbles;
    r.arrival = p.arrival;
    r.completed = now();
    r.launch_worlds = launchWorlds;
    r.kernel_class = cls < 0 ? 0xffffffffu : (uint32_t)cls;
    r.device_milliseconds = deviceMs;
    r.transaction_result = txResult;
    std::lock_guard<std::mutex> lock(mu);
    if (p.sink >= 0 && sinkResults) sinkResults[p.sink] = r;
    else results.push_back(r);
    ++completed;
    st.completed = completed;
    publish_counters();
    cvDone.notify_all();
}

// Phase 1 of staging an add, merge or split: the world's published records, its descriptor and
// its bodies' poses are read back on a stream of the staging job (a job is taken when one is free;
// otherwise the world waits its turn). Nothing here waits for the device.
bp_status bp_server::begin_stage(uint32_t w) {
    if (freeStage.empty()) { stageWaiting.push_back(w); return BP_OK; }
    const int ji = freeStage.back();
    freeStage.pop_back();
    StageJob &job = stageJobs[ji];
    bp_batch &b = *batch;
    const HostWorld &hw = b.worlds[w];
    const uint32_t n = count[w], need = n + 2, nb = (uint32_t)hw.bodies.size();
    if (job.capacity < need) {
        if (job.dRecs) (void)hipFree(job.dRecs);
        if (job.dIdx) (void)hipFree(job.dIdx);
        if (job.hRecs) (void)hipHostFree(job.hRecs);
        if (job.hIdx) (void)hipHostFree(job.hIdx);
        job.dRecs = job.hRecs = nullptr; job.dIdx = job.hIdx = nullptr; job.capacity = 0;
        const uint32_t cap = std::max<uint32_t>(need, 1024);
        BP_HIP(hipMalloc(&job.dRecs, cap * sizeof(bp_bubble)));
        BP_HIP(hipMalloc(&job.dIdx, 2 * (size_t)cap * sizeof(int32_t)));
        BP_HIP(hipHostMalloc(&job.hRecs, cap * sizeof(bp_bubble), hipHostMallocDefault));
        BP_HIP(hipHostMalloc(&job.hIdx, 2 * (size_t)cap * sizeof(int32_t), hipHostMallocDefault));
        job.capacity = cap;
    }
    if (!job.hDesc) BP_HIP(hipHostMalloc(&job.hDesc, sizeof(bpd::WorldDesc), hipHostMallocDefault));
    if (job.poseCapacity < nb) {
        if (job.hPose) (void)hipHostFree(job.hPose);
        job.hPose = nullptr; job.poseCapacity = 0;
        BP_HIP(hipHostMalloc(&job.hPose, 2 * (size_t)std::max<uint32_t>(nb, 64) * sizeof(float4), hipHostMallocDefault));
        job.poseCapacity = std::max<uint32_t>(nb, 64);
    }
    job.world = w;
    job.phase = 1;
    hipError_t e = bp_launch_gather_full_world(b.arenaDev, b.descs + w, job.dRecs, n, job.stream);
    if (e != hipSuccess) { set_error(std::string("gather: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    BP_HIP(hipMemcpyAsync(job.hRecs, job.dRecs, n * sizeof(bp_bubble), hipMemcpyDeviceToHost, job.stream));
    BP_HIP(hipMemcpyAsync(job.hDesc, b.descs + w, sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost, job.stream));
    if (nb) {
        const uint32_t off = b.hostDescs[w].bodyOffset;
        BP_HIP(hipMemcpyAsync(job.hPose, b.arena.bodyPose[0] + off, nb * sizeof(float4), hipMemcpyDeviceToHost, job.stream));
        BP_HIP(hipMemcpyAsync(job.hPose + nb, b.arena.bodyPose[1] + off, nb * sizeof(float4), hipMemcpyDeviceToHost, job.stream));
    }
    BP_HIP(hipEventRecord(job.done, job.stream));
    activeStage.push_back(ji);
    return BP_OK;
}

// Phase 2: the reference's staging on the host (bp_stage_add, bp_stage_merge, bp_stage_split:
// TransactionStager.swift:406-421, 630-735) against the world's published records and its bodies
// at their current poses; a refused transaction leaves the world as it is (the request's
// txResult holds the reason), one that needs more slots than the world has is rejected, and an
// applied one is written back in place (scatter kernel and descriptor) on the job's stream.
bp_status bp_server::stage_on_host(StageJob &job) {
    bp_batch &b = *batch;
    const uint32_t w = job.world;
    HostWorld &hw = b.worlds[w];
    Pending &p = queue[w].front();
    const uint32_t n = count[w];
    const uint32_t nb = (uint32_t)hw.bodies.size(), pub = job.hDesc->pub;
    uint32_t outCount = 0;
    bp_stage_rejection why{};
    bp_status st = stage_scene(w, p, job.hRecs, n, job.hPose + (size_t)pub * nb, job.staged, outCount, why);
    if (st == BP_ERR_INVALID_ARGUMENT && why.reason != BP_REJECT_NONE) {
        p.staged = true;                               // refused: the world is unchanged, the step runs
        p.txResult = why.reason;
        job.phase = 0;
        return BP_OK;
    }
    if (st != BP_OK) return st;
    if (outCount > hw.capacity) {
        --local.transactions;
        ++local.rejected;
        Pending done = p; queue[w].pop_front();
        finish_request(done, BP_ERR_CAPACITY_OVERFLOW, 0, n, 0, -1, 0.0f);
        job.phase = 0;
        return BP_OK;
    }
    job.staged.resize(outCount);
    std::memcpy(job.hRecs, job.staged.data(), outCount * sizeof(bp_bubble));
    b.motor_band_indices(w, job.staged.data(), outCount, job.hIdx);
    b.written_desc(w, outCount, *job.hDesc);
    if (outCount) {
        BP_HIP(hipMemcpyAsync(job.dRecs, job.hRecs, outCount * sizeof(bp_bubble), hipMemcpyHostToDevice, job.stream));
        BP_HIP(hipMemcpyAsync(job.dIdx, job.hIdx, 2 * (size_t)outCount * sizeof(int32_t), hipMemcpyHostToDevice, job.stream));
        hipError_t e = bp_launch_scatter_world(b.arenaDev, job.hDesc->bubbleOffset, outCount, job.dRecs, job.dIdx,
                                               job.dIdx + outCount, job.stream);
        if (e != hipSuccess) { set_error(std::string("scatter: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    }
    BP_HIP(hipMemcpyAsync(b.descs + w, job.hDesc, sizeof(bpd::WorldDesc), hipMemcpyHostToDevice, job.stream));
    BP_HIP(hipEventRecord(job.done, job.stream));
    job.phase = 2;
    return BP_OK;
}

// The reference's staging of the head request's add, merge or split (bp_stage_add,
// bp_stage_merge, bp_stage_split: TransactionStager.swift:406-421, 630-735) against the world's
// published records `recs` (n) and its bodies at the poses `poses` (x, y, angle per body); the
// staged list goes to `out`. Counts the transaction.
bp_status bp_server::stage_scene(uint32_t w, const Pending &p, const bp_bubble *recs, uint32_t n, const float4 *poses,
                                 std::vector<bp_bubble> &out, uint32_t &outCount, bp_stage_rejection &why) {
    const HostWorld &hw = batch->worlds[w];
    std::vector<bp_body> bodies = hw.bodies;
    for (size_t k = 0; k < bodies.size(); ++k) {
        bodies[k].pose[0] = poses[k].x; bodies[k].pose[1] = poses[k].y; bodies[k].pose[2] = poses[k].z;
    }
    bp_stage_scene scene{};
    scene.bubbles = recs;
    scene.bubble_count = n;
    scene.bodies = bodies.empty() ? nullptr : bodies.data();
    scene.body_count = (uint32_t)bodies.size();
    scene.vertices = hw.vertices.empty() ? nullptr : hw.vertices.data();
    scene.vertex_count = (uint32_t)(hw.vertices.size() / 2);
    for (int k = 0; k < 4; ++k) scene.region[k] = hw.config.region[k];
    const uint32_t i = p.req.index % n;
    const int64_t newId = recs[n - 1].id + 1;
    out.resize(n + 2);
    outCount = 0;
    ++local.transactions;
    if (p.req.transaction == BP_TX_ADD) {
        bp_bubble nb = recs[i];
        nb.id = newId;
        const float offset = 2.0f * recs[i].radius;
        nb.center[0] = nb.previous_center[0] = recs[i].center[0] + offset * p.req.direction[0];
        nb.center[1] = nb.previous_center[1] = recs[i].center[1] + offset * p.req.direction[1];
        nb.velocity[0] = nb.velocity[1] = 0.0f;
        nb.rotation = nb.angular_velocity = 0.0f;
        return bp_stage_add(&scene, &nb, out.data(), (uint32_t)out.size(), &outCount, &why);
    }
    if (p.req.transaction == BP_TX_MERGE) {
        if (n < 2) {
            why.reason = BP_REJECT_INVALID_STRUCTURAL_OPERATION;
            return BP_ERR_INVALID_ARGUMENT;
        }
        const uint32_t k = i + 1 < n ? i + 1 : i - 1, a = std::min(i, k), c = std::max(i, k);
        const int64_t sources[2] = {recs[a].id, recs[c].id};
        return bp_stage_merge(&scene, sources, 2, &recs[a], 0, out.data(), (uint32_t)out.size(), &outCount, &why);
    }
    bp_bubble first = recs[i], second = recs[i];
    second.id = newId;
    return bp_stage_split(&scene, recs[i].id, &first, 0.5f, &second, 0.5f, out.data(), (uint32_t)out.size(), &outCount, &why);
}

// Stages the head request's add, merge or split from the world's mirror (its published state,
// current once its last launch has completed) and writes the staged list and its motor and band
// indices into the mirror, for the launch's transaction kernel to apply before the step.
bp_status bp_server::stage_from_mirror(uint32_t w, Pending &p) {
    const HostWorld &hw = batch->worlds[w];
    const uint32_t n = count[w], cap = mirrorCap[w];
    std::vector<bp_bubble> out;
    uint32_t outCount = 0;
    bp_stage_rejection why{};
    bp_status st = stage_scene(w, p, mirrorHost + mirrorOffset[w], n, bodyMirrorHost + bodyMirrorOffset[w], out, outCount, why);
    p.staged = true;
    if (st == BP_ERR_INVALID_ARGUMENT && why.reason != BP_REJECT_NONE) {
        p.txResult = why.reason;                       // refused: the world is unchanged, the step runs
        return BP_OK;
    }
    if (st != BP_OK) return st;
    if (outCount > hw.capacity || outCount > cap) {
        p.staged = false;
        --local.transactions;
        return BP_ERR_CAPACITY_OVERFLOW;
    }
    std::memcpy(mirrorHost + mirrorOffset[w], out.data(), outCount * sizeof(bp_bubble));
    std::vector<int32_t> idx(2 * (size_t)outCount);
    batch->motor_band_indices(w, out.data(), outCount, idx.data());
    int32_t *dst = mirrorIdxHost + 2 * mirrorOffset[w];
    for (uint32_t k = 0; k < outCount; ++k) { dst[k] = idx[k]; dst[cap + k] = idx[outCount + k]; }
    p.txResult = BP_REJECT_NONE;
    p.replace = true;
    p.replaceCount = outCount;
    return BP_OK;
}

// Points every resident world's descriptor at its mirror and fills the mirror from the batch's
// host copy of the world (current after a pack or a read-back); a world beyond its mirror's
// slots gets none. Runs with nothing in flight.
bp_status bp_server::attach_mirrors() {
    bp_batch &b = *batch;
    BP_HIP(hipMemcpy(b.hostDescs.data(), b.descs, W * sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost));
    for (uint32_t w = 0; w < W; ++w) {
        bpd::WorldDesc &d = b.hostDescs[w];
        const HostWorld &hw = b.worlds[w];
        const uint32_t n = (uint32_t)hw.bubbles.size();
        mirrorValid[w] = 0;
        mirrorAttached[w] = 0;
        d.mirror = nullptr; d.bodyMirror = nullptr; d.mirrorIdx = nullptr; d.mirrorCapacity = 0;
        // A resident world's head request staged on its mirror loses its staged list here (the
        // mirror is refilled or dropped): it is staged again from the new state.
        if (resident[w] && !queue.empty() && !queue[w].empty() && queue[w].front().replace) {
            Pending &p = queue[w].front();
            p.staged = false; p.replace = false; p.replaceCount = 0;
            --local.transactions;
        }
        if (!resident[w] || n > mirrorCap[w] || n != d.bubbleCount) continue;
        mirrorAttached[w] = 1;
        d.mirror = mirrorDev + mirrorOffset[w];
        d.bodyMirror = bodyMirrorDev + bodyMirrorOffset[w];
        d.mirrorIdx = mirrorIdxDev + 2 * mirrorOffset[w];
        d.mirrorCapacity = mirrorCap[w];
        if (n) std::memcpy(mirrorHost + mirrorOffset[w], hw.bubbles.data(), n * sizeof(bp_bubble));
        for (size_t k = 0; k < hw.bodies.size(); ++k) {
            const bp_body &bd = hw.bodies[k];
            bodyMirrorHost[bodyMirrorOffset[w] + k] = make_float4(bd.pose[0], bd.pose[1], bd.pose[2], bd.pose[2]);
        }
        mirrorValid[w] = 1;
    }
    BP_HIP(hipMemcpy(b.descs, b.hostDescs.data(), W * sizeof(bpd::WorldDesc), hipMemcpyHostToDevice));
    return BP_OK;
}

// Clears the mirror addresses from the descriptors before the mirrors are freed.
void bp_server::detach_mirrors() {
    if (!mirrorHost || !batch || !batch->packed || !batch->descs) return;
    bp_batch &b = *batch;
    if (hipMemcpy(b.hostDescs.data(), b.descs, W * sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost) != hipSuccess) return;
    for (auto &d : b.hostDescs) { d.mirror = nullptr; d.bodyMirror = nullptr; d.mirrorIdx = nullptr; d.mirrorCapacity = 0; }
    (void)hipMemcpy(b.descs, b.hostDescs.data(), W * sizeof(bpd::WorldDesc), hipMemcpyHostToDevice);
}

void bp_server::free_stage_job(int j) {
    StageJob &job = stageJobs[j];
    staging[job.world] = 0;
    job.phase = 0;
    freeStage.push_back(j);
}

// Advances the staging jobs whose event has completed; a world whose staging ended goes back to
// prepare_head (its request then runs as an ordinary step). Starts the staging of waiting worlds
// as jobs free up.
bool bp_server::poll_stages() {
    bool any = false;
    for (size_t k = 0; k < activeStage.size();) {
        const int ji = activeStage[k];
        StageJob &job = stageJobs[ji];
        hipError_t e = hipEventQuery(job.done);
        if (e == hipErrorNotReady) { ++k; continue; }
        if (e != hipSuccess) { set_error(std::string("hipEventQuery: ") + hipGetErrorString(e)); fail(BP_ERR_COMMAND_FAILED); return false; }
        const uint32_t w = job.world;
        if (job.phase == 1) {
            bp_status s = stage_on_host(job);
            if (s != BP_OK) { fail(s); return false; }
            if (job.phase == 2) { ++k; any = true; continue; }           // written back next
        } else {
            const uint32_t outCount = (uint32_t)job.staged.size();
            bp_status s = batch->adopt_written(w, job.staged.data(), outCount, *job.hDesc);
            if (s != BP_OK) { fail(s); return false; }
            count[w] = outCount;
            Pending &p = queue[w].front();
            p.staged = true;
            p.txResult = BP_REJECT_NONE;
        }
        activeStage[k] = activeStage.back();
        activeStage.pop_back();
        free_stage_job(ji);
        prepare_head(w);
        any = true;
    }
    while (!freeStage.empty() && !stageWaiting.empty()) {
        const uint32_t w = stageWaiting.front();
        stageWaiting.pop_front();
        if (!resident[w] || queue[w].empty()) { staging[w] = 0; prepare_head(w); continue; }
        bp_status s = begin_stage(w);
        if (s != BP_OK) { fail(s); return false; }
        any = true;
    }
    return any;
}

// Deletes, on the device, the bubbles the last step of each listed world removed from the region
// (the public world deletes them before the next frame, NativePublicWorld.swift:264-267).
bp_status bp_server::drop_removed(const std::vector<uint32_t> &ws) {
    if (ws.empty()) return BP_OK;
    std::vector<bpd::DropRecord> recs(ws.size());
    for (size_t j = 0; j < ws.size(); ++j) {
        const uint32_t w = ws[j];
        const uint32_t n = count[w] >= dropPending[w] ? count[w] - dropPending[w] : 0u;
        recs[j] = bpd::DropRecord{w, buckets_for(w, n), 0u, 0u};
    }
    BP_HIP(hipMemcpyAsync(dropDev, recs.data(), recs.size() * sizeof(bpd::DropRecord), hipMemcpyHostToDevice, readStream));
    hipError_t e = bp_launch_drop_removed(batch->arenaDev, batch->descs, dropDev, (uint32_t)ws.size(), nullptr, readStream);
    if (e != hipSuccess) { set_error(std::string("drop kernel: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    BP_HIP(hipStreamSynchronize(readStream));
    for (uint32_t w : ws) {
        count[w] = count[w] >= dropPending[w] ? count[w] - dropPending[w] : 0u;
        dropPending[w] = 0;
        mirrorValid[w] = 0;                          // the mirror still holds the deleted bubbles
    }
    return BP_OK;
}

// Posts the result of every world whose record is marked complete, and frees the worlds of every
// launch whose event has completed.
bool bp_server::poll_launches() {
    bool any = false;
    for (size_t k = 0; k < inflight.size();) {
        Launch &L = launches[inflight[k]];
        auto post = [&](uint32_t j) {
            const uint32_t w = L.worlds[j];
            const bpd::WorldStatus &s = hostStats[w];
            const bp_status status = s.executed ? (bp_status)s.status : BP_ERR_UNAVAILABLE;
            if (s.executed && status != BP_OK && status != BP_ERR_CAPACITY_OVERFLOW) worldFailed[w] = 1;
            if (s.executed && status == BP_OK && s.removals) dropPending[w] = s.removals;
            const uint32_t txr = L.reqs[j].staged ? L.reqs[j].txResult : s.txResult;
            finish_request(L.reqs[j], status, s.executed ? s.substeps : 0, L.bubbles[j], L.count, L.cls,
                           s.executed ? (float)s.ticks * 1e-5f : 0.0f, txr);
            L.posted[j] = 1;
            ++L.postedCount;
            any = true;
        };
        for (uint32_t j = 0; j < L.count && L.postedCount < L.count; ++j)
            if (!L.posted[j] && __atomic_load_n(&hostStats[L.worlds[j]].seq, __ATOMIC_ACQUIRE) == 1u) post(j);
        hipError_t e = hipEventQuery(L.done);
        if (e == hipErrorNotReady) { ++k; continue; }
        if (e != hipSuccess) { set_error(std::string("hipEventQuery: ") + hipGetErrorString(e)); fail(BP_ERR_COMMAND_FAILED); return false; }
        std::vector<uint32_t> held;
        for (uint32_t j = 0; j < L.count; ++j) {
            if (!L.posted[j]) post(j);                       // the launch is over: every record is final
            if (dropPending[L.worlds[j]]) held.push_back(L.worlds[j]);
        }
        if (bp_status ds = drop_removed(held); ds != BP_OK) { fail(ds); return false; }
        for (uint32_t j = 0; j < L.count; ++j) {
            const uint32_t w = L.worlds[j];
            busy[w] = 0;
            // A published step without region removals leaves the mirror current (the kernels write
            // it when the world's descriptor holds one and the world fits it); any other outcome
            // leaves it stale (an unpublished state, or bubbles the drop deleted).
            const bpd::WorldStatus &r = hostStats[w];
            mirrorValid[w] = mirrorAttached[w] && r.executed && r.status == BP_OK && r.removals == 0 && count[w] <= mirrorCap[w];
            prepare_head(w);
        }
        --inflightPerClass[L.cls];
        --streamLoad[L.stream];
        ++idleStreams;
        freeLaunch.push_back(inflight[k]);
        inflight[k] = inflight.back();
        inflight.pop_back();
        any = true;
    }
    return any;
}

// Queues one launch of class `cls` with its ready worlds.
bp_status bp_server::start_launch(int cls) {
    const int li = freeLaunch.back();
    freeLaunch.pop_back();
    Launch &L = launches[li];
    L.cls = cls;
    L.worlds.clear(); L.reqs.clear(); L.bubbles.clear();
    L.txCount = 0;
    const uint32_t cap = std::max<uint32_t>(1, std::min(cfg.max_worlds_per_launch, W));
    auto &rq = ready[cls];
    while (!rq.empty() && L.worlds.size() < cap) {
        const uint32_t w = rq.front();
        rq.pop_front();
        queued[w] = 0;
        busy[w] = 1;
        Pending p = queue[w].front();
        queue[w].pop_front();
        hostStats[w].txResult = BP_REJECT_NONE;
        const uint32_t kind = p.req.transaction;
        if (p.replace) {
            bpd::TxRecord &t = L.hTx[L.txCount++];
            t.world = w; t.kind = bpd::kTxReplace; t.index = p.replaceCount; t.buckets = buckets_for(w, p.replaceCount);
            t.impulse[0] = t.impulse[1] = 0.0f; t.direction[0] = t.direction[1] = 0.0f;
        } else if ((kind == BP_TX_IMPULSE || kind == BP_TX_REMOVE) && count[w] > 0) {
            ++local.transactions;
            if (kind == BP_TX_IMPULSE && !(std::isfinite(p.req.impulse[0]) && std::isfinite(p.req.impulse[1]))) {
                p.staged = true;                             // requireFinite(impulse): refused on the host
                p.txResult = BP_REJECT_INVALID_VALUE;
            } else {
                bpd::TxRecord &t = L.hTx[L.txCount++];
                t.world = w; t.kind = kind; t.index = p.req.index; t.buckets = buckets_for(w, headCount[w]);
                t.impulse[0] = p.req.impulse[0]; t.impulse[1] = p.req.impulse[1];
                t.direction[0] = p.req.direction[0]; t.direction[1] = p.req.direction[1];
            }
        }
        count[w] = headCount[w];
        lastUse[w] = ++useClock;
        L.worlds.push_back(w);
        L.reqs.push_back(p);
        L.bubbles.push_back(count[w]);
    }
    // Longest world first inside the launch (its workgroup starts first).
    std::vector<uint32_t> idx(L.worlds.size());
    for (uint32_t j = 0; j < idx.size(); ++j) idx[j] = j;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return L.bubbles[a] > L.bubbles[b]; });
    std::vector<uint32_t> ws(idx.size()), nb(idx.size());
    std::vector<Pending> rs(idx.size());
    for (uint32_t j = 0; j < idx.size(); ++j) { ws[j] = L.worlds[idx[j]]; nb[j] = L.bubbles[idx[j]]; rs[j] = L.reqs[idx[j]]; }
    L.worlds.swap(ws); L.bubbles.swap(nb); L.reqs.swap(rs);
    L.count = (uint32_t)L.worlds.size();
    L.posted.assign(L.count, 0);
    L.postedCount = 0;
    for (uint32_t j = 0; j < L.count; ++j) {
        L.hOrder[j] = L.worlds[j];
        __atomic_store_n(&hostStats[L.worlds[j]].seq, 0u, __ATOMIC_RELEASE);   // the kernel sets it to 1
    }
    // An idle stream (form_launches checked that one exists): a launch never queues behind
    // another one, so a world with a very long step holds up only the worlds of its own launch.
    int si = 0;
    while (streamLoad[si] != 0) ++si;
    L.stream = si;
    hipStream_t s = streams[si];
    bp_batch &b = *batch;
    if (L.txCount) {
        hipError_t e = bp_launch_transactions(b.arenaDev, b.descs, L.dTx, L.txCount, stats, s);
        if (e != hipSuccess) { set_error(std::string("transaction launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    }
    hipError_t e = cls == kGlobalClass ? bp_launch_world_step(1024, b.arenaDev, b.descs, stats, L.dOrder, L.count, 1u, s)
                                       : bp_launch_world_lds(cls, b.arenaDev, b.descs, stats, L.dOrder, L.count, 1, W, 1u, 0u, nullptr, nullptr, s);
    if (e != hipSuccess) { set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    BP_HIP(hipEventRecord(L.done, s));
    ++inflightPerClass[cls];
    ++streamLoad[si];
    --idleStreams;
    inflight.push_back(li);
    ++local.launches;
    local.launched_worlds += L.count;
    return BP_OK;
}

// Admits the worlds whose requests wait in host memory. Runs with nothing in flight: the
// newest state of every world is read back, idle worlds are evicted least recently used first
// until the admitted ones fit the budget, and the batch is repacked.
bp_status bp_server::admit_waiting() {
    bp_batch &b = *batch;
    bp_status s = b.download_state(true);
    if (s != BP_OK) return s;
    std::vector<uint8_t> admitted(W, 0);
    std::vector<uint32_t> deferred;                  // waiting worlds left for the next round
    uint32_t admittedCount = 0;
    for (uint32_t w : admissions) {
        if (resident[w]) continue;
        if (!deferred.empty()) { deferred.push_back(w); continue; }
        const uint64_t need = b.footprint(w);
        // Whether evicting every resident world not admitted in this round makes room.
        uint64_t freeable = 0;
        for (uint32_t v = 0; v < W; ++v) if (resident[v] && !admitted[v]) freeable += footprintOf[v];
        if (cfg.memory_budget && used - freeable + need > cfg.memory_budget) {
            if (admittedCount == 0) { set_error("memory budget too small for the waiting worlds"); return BP_ERR_ALLOCATION_FAILED; }
            deferred.push_back(w);                    // the worlds admitted now go first
            continue;
        }
        while (cfg.memory_budget && used + need > cfg.memory_budget) {
            // Least recently used resident world with no waiting request and not admitted now.
            int64_t victim = -1;
            for (uint32_t v = 0; v < W; ++v)
                if (resident[v] && !admitted[v] && queue[v].empty() && (victim < 0 || lastUse[v] < lastUse[victim])) victim = v;
            if (victim < 0)
                for (uint32_t v = 0; v < W; ++v)
                    if (resident[v] && !admitted[v] && (victim < 0 || lastUse[v] < lastUse[victim])) victim = v;
            if (victim < 0) { set_error("memory budget too small for the waiting worlds"); return BP_ERR_ALLOCATION_FAILED; }
            resident[victim] = 0;
            b.worlds[victim].resident = false;
            used -= footprintOf[victim];
            // Requests already prepared for the victim go back to waiting for admission.
            if (queued[victim]) {
                auto &rq = ready[headClass[victim]];
                rq.erase(std::remove(rq.begin(), rq.end(), victim), rq.end());
                queued[victim] = 0;
            }
            ++local.evictions;
        }
        resident[w] = 1;
        admitted[w] = 1;
        ++admittedCount;
        b.worlds[w].resident = true;
        footprintOf[w] = need;
        used += need;
        ++local.admissions;
    }
    b.packed = false;
    if ((s = b.pack()) != BP_OK) return s;
    ++local.repacks;
    if ((s = attach_mirrors()) != BP_OK) return s;
    for (uint32_t w = 0; w < W; ++w) {
        count[w] = (uint32_t)b.worlds[w].bubbles.size();
        dropPending[w] = 0;                          // the repack deleted the region-removed bubbles
        if (resident[w]) footprintOf[w] = b.footprint(w);
    }
    used = 0;
    for (uint32_t w = 0; w < W; ++w) if (resident[w]) used += footprintOf[w];
    local.resident_bytes = used;
    std::vector<uint32_t> waiting;
    waiting.swap(admissions);
    for (uint32_t w : waiting) wantAdmission[w] = 0;
    for (uint32_t w : deferred) { wantAdmission[w] = 1; admissions.push_back(w); waitingSince[w] = now(); }
    // Ready lists hold class choices made with the old layout; rebuild them.
    for (auto &rq : ready) rq.clear();
    for (uint32_t w = 0; w < W; ++w) {
        queued[w] = 0;
        if (!queue[w].empty()) prepare_head(w);
    }
    return BP_OK;
}

// Admits waiting worlds into the device ranges of evicted worlds: for each, the least recently
// used resident world that is idle (no launch, request, staging or pending removal) and whose
// ranges hold it, within the budget. A round's exchanges share one read-back and one upload, and
// the dispatcher goes on launching while they are on the device: both worlds of an exchange are
// `moving` (neither launches nor takes part in another exchange) until the round ends. Worlds for
// which no victim qualifies stay in the admission list for the repacking round.
bp_status bp_server::start_swap_round() {
    bp_batch &b = *batch;
    std::vector<uint32_t> left;
    SwapRoundState &r = swapState;
    r.outs.clear(); r.ins.clear(); r.outIdle.clear(); r.mirrors.clear();
    // The worlds that may be displaced, gathered once per round.
    struct Candidate { uint32_t v, slots; uint64_t use; uint8_t idle; };
    std::vector<Candidate> cands;
    cands.reserve(W);
    for (uint32_t v = 0; v < W; ++v) {
        if (!resident[v] || busy[v] || staging[v] || dropPending[v] || moving[v]) continue;
        cands.push_back(Candidate{v, b.regions[v].bubbles, lastUse[v], (uint8_t)(!queued[v] && queue[v].empty())});
    }
    // The best candidate whose ranges hold w: idle before waiting (those only when `urgent`), then
    // the fewest bubble slots, then the least recently used; ranges are checked only for a
    // candidate that beats the best so far.
    auto pick = [&](uint32_t need, uint32_t w, bool urgent) -> int64_t {
        int64_t best = -1;
        for (size_t i = 0; i < cands.size(); ++i) {
            const Candidate &c = cands[i];
            if (c.slots < need || (!c.idle && !urgent) || moving[c.v]) continue;
            if (best >= 0) {
                const Candidate &o = cands[(size_t)best];
                if (o.idle && !c.idle) continue;
                if (o.idle == c.idle && (c.slots > o.slots || (c.slots == o.slots && c.use >= o.use))) continue;
            }
            if (b.fits_ranges_of(c.v, w)) best = (int64_t)i;
        }
        return best;
    };
    for (uint32_t w : admissions) {
        if (resident[w]) { wantAdmission[w] = 0; continue; }
        if (moving[w]) { left.push_back(w); continue; }
        // The world takes over the victim's device memory, so the device footprint does not grow.
        // Among the idle worlds whose ranges hold it, the one with the fewest bubble slots (so that
        // large ranges stay for large worlds), the least recently used of those. A world that has
        // waited kUrgentWait may also displace a world whose requests only wait (none on the device
        // or being staged): those go back to waiting for admission.
        const uint32_t n = (uint32_t)b.worlds[w].bubbles.size(), need = n + std::max<uint32_t>(8, n / 8);
        const int64_t at = pick(need, w, now() - waitingSince[w] >= kUrgentWait);
        if (at < 0) {
            left.push_back(w);
            continue;
        }
        const uint32_t v = cands[(size_t)at].v;
        moving[v] = 1;
        moving[w] = 1;
        if (queued[v]) {                                 // a displaced world launches no more until the round ends
            auto &rq = ready[headClass[v]];
            rq.erase(std::remove(rq.begin(), rq.end(), v), rq.end());
            queued[v] = 0;
        }
        r.outs.push_back(v);
        r.ins.push_back(w);
        r.outIdle.push_back(cands[(size_t)at].idle);
        // The mirror of the world moved in, when it has room: its descriptor goes up pointing there.
        r.mirrors.push_back(n <= mirrorCap[w] ? bp_batch::MirrorRange{mirrorDev + mirrorOffset[w], bodyMirrorDev + bodyMirrorOffset[w],
                                                                      mirrorIdxDev + 2 * mirrorOffset[w], mirrorCap[w]}
                                              : bp_batch::MirrorRange{nullptr, nullptr, nullptr, 0u});
    }
    admissions.swap(left);
    if (r.outs.empty()) return BP_OK;
    r.results.assign(r.outs.size(), BP_OK);
    bp_status s = b.swap_begin(r.outs.data(), r.ins.data(), (uint32_t)r.outs.size(), r.mirrors.data(), r.results.data(), readStream);
    if (s != BP_OK) return s;
    if (b.swapRecords == 0) return finish_swap_round();
    BP_HIP(hipEventRecord(swapDone, readStream));
    r.phase = 1;
    return BP_OK;
}

// Advances the round in flight when its device work has finished.
bp_status bp_server::advance_swap_round(bool wait) {
    SwapRoundState &r = swapState;
    if (r.phase == 0) return BP_OK;
    if (wait) {
        BP_HIP(hipEventSynchronize(swapDone));
    } else {
        const hipError_t e = hipEventQuery(swapDone);
        if (e == hipErrorNotReady) return BP_OK;
        BP_HIP(e);
    }
    if (r.phase == 1) {
        bp_batch &b = *batch;
        bp_status s = b.swap_upload(r.results.data(), readStream);
        if (s != BP_OK) return s;
        if (b.swapUploads) {
            BP_HIP(hipEventRecord(swapDone, readStream));
            r.phase = 2;
            return wait ? advance_swap_round(true) : BP_OK;
        }
    }
    return finish_swap_round();
}

bp_status bp_server::finish_swap_round() {
    bp_batch &b = *batch;
    SwapRoundState &r = swapState;
    b.swap_finish();
    r.phase = 0;
    admissionRetry = true;
    for (size_t i = 0; i < r.outs.size(); ++i) {
        const uint32_t v = r.outs[i], w = r.ins[i];
        moving[v] = 0;
        moving[w] = 0;
        if (r.results[i] != BP_OK) {
            // Not exchanged (held, or left for a later round): both go on as they were.
            if (r.results[i] != BP_ERR_UNAVAILABLE && r.results[i] != BP_ERR_CAPACITY_OVERFLOW) return r.results[i];
            admissions.push_back(w);
            if (!busy[v] && !queued[v] && !staging[v]) prepare_head(v);
            continue;
        }
        resident[v] = 0;
        used -= footprintOf[v];
        footprintOf[v] = b.footprint(v);
        mirrorAttached[v] = 0;
        mirrorValid[v] = 0;
        count[v] = (uint32_t)b.worlds[v].bubbles.size();
        ++local.evictions;
        if (!queue[v].empty()) {
            // The displaced world's waiting requests wait for its admission; a transaction staged
            // on its mirror is staged again from its state after admission.
            Pending &p = queue[v].front();
            if (p.replace) { p.staged = false; p.replace = false; p.replaceCount = 0; --local.transactions; }
            prepare_head(v);
        }
        resident[w] = 1;
        footprintOf[w] = b.footprint(w);
        used += footprintOf[w];                          // at most the victim's: its ranges hold the world
        count[w] = (uint32_t)b.worlds[w].bubbles.size();
        dropPending[w] = 0;
        wantAdmission[w] = 0;
        ++local.admissions;
        // The world's mirror (its descriptor on the device points there), filled from its host
        // copy: the state just written to the device.
        if (r.mirrors[i].capacity) {
            const HostWorld &hw = b.worlds[w];
            const uint32_t n = (uint32_t)hw.bubbles.size();
            if (n) std::memcpy(mirrorHost + mirrorOffset[w], hw.bubbles.data(), n * sizeof(bp_bubble));
            for (size_t k = 0; k < hw.bodies.size(); ++k) {
                const bp_body &bd = hw.bodies[k];
                bodyMirrorHost[bodyMirrorOffset[w] + k] = make_float4(bd.pose[0], bd.pose[1], bd.pose[2], bd.pose[2]);
            }
            mirrorAttached[w] = 1;
            mirrorValid[w] = 1;
        }
        prepare_head(w);
    }
    local.resident_bytes = used;
    return BP_OK;
}

// Launches the ready worlds. While worlds wait for admission, no new launches start (the repack
// needs an idle device) except right after an admission round: then every world that is ready
// is launched once before the next round, so each round makes progress even when the budget
// forces the next admission to evict a world that still has requests.
bool bp_server::form_launches() {
    if (swapState.phase != 0) {
        const double t0 = now();
        bp_status s = advance_swap_round(false);
        local.admission_microseconds += (uint64_t)((now() - t0) * 1e6);
        if (s != BP_OK) { fail(s); return false; }
    }
    if (swapState.phase == 0 && !admissions.empty() && admissionRetry && now() - lastAdmissionRound >= kAdmissionInterval) {
        admissionRetry = false;
        const double t0 = now();
        lastAdmissionRound = t0;
        bp_status s = start_swap_round();
        local.admission_microseconds += (uint64_t)((now() - t0) * 1e6);
        if (s != BP_OK) { fail(s); return false; }
    }
    // A world that no idle world's ranges hold usually finds one as launches complete; the device
    // drains for a repacking round only once the oldest of them has waited kRepackDelay.
    bool repack = false;
    if (!admissions.empty() && !launchAfterAdmission && swapState.phase == 0) {
        double oldest = now();
        for (uint32_t w : admissions) oldest = std::min(oldest, waitingSince[w]);
        repack = now() - oldest >= kRepackDelay;
    }
    if (repack) {
        if (!inflight.empty() || !activeStage.empty()) return false;   // drain first: the repack moves every world
        const double t0 = now();
        bp_status s = admit_waiting();
        local.admission_microseconds += (uint64_t)((now() - t0) * 1e6);
        if (s != BP_OK) { fail(s); return false; }
        launchAfterAdmission = true;
        return true;
    }
    bool any = false;
    // Classes in the order of their oldest ready request, each while an idle stream is left.
    int order[kClassCount];
    int classes = 0;
    for (int k = 0; k < kClassCount; ++k) if (!ready[k].empty()) order[classes++] = k;
    std::sort(order, order + classes, [&](int a, int b) {
        return queue[ready[a].front()].front().arrival < queue[ready[b].front()].front().arrival;
    });
    for (int x = 0; x < classes; ++x) {
        const int k = order[x];
        while (!ready[k].empty() && inflightPerClass[k] < (int)cfg.max_inflight_per_class && !freeLaunch.empty() &&
               idleStreams > 0) {
            bp_status s = start_launch(k);
            if (s != BP_OK) { fail(s); return false; }
            any = true;
        }
    }
    bool readyLeft = false;
    for (int k = 0; k < kClassCount && !readyLeft; ++k) readyLeft = !ready[k].empty();
    if (!readyLeft) launchAfterAdmission = false;
    return any;
}

void bp_server::run() {
    (void)hipSetDevice(device);
    std::vector<Pending> taken;
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(mu);
            if (fatal != BP_OK) break;
            bool idle = inflight.empty() && activeStage.empty();
            if (idle)
                for (int k = 0; k < kClassCount && idle; ++k) idle = ready[k].empty();
            idle = idle && admissions.empty() && swapState.phase == 0;
            if (stop && inflight.empty()) break;
            if (incoming.empty() && idle && !stop) cvWork.wait_for(lock, std::chrono::milliseconds(2));
            taken.assign(incoming.begin(), incoming.end());
            incoming.clear();
            publish_counters();
        }
        for (const Pending &p : taken) enqueue(p);
        bool progressed = poll_launches();
        progressed |= poll_stages();
        if (progressed) admissionRetry = true;           // worlds may have become idle
        {
            std::lock_guard<std::mutex> lock(mu);
            if (fatal != BP_OK) break;
            if (stop) {                             // no new launches once stopping
                if (!progressed) std::this_thread::yield();
                continue;
            }
        }
        progressed |= form_launches();
        if (!progressed && (!inflight.empty() || !activeStage.empty() || !admissions.empty())) std::this_thread::yield();
    }
    // Drain whatever is still on the device before the dispatcher ends.
    for (int id : inflight) (void)hipEventSynchronize(launches[id].done);
    for (int j : activeStage) (void)hipEventSynchronize(stageJobs[j].done);
    if (swapState.phase != 0 && advance_swap_round(true) != BP_OK) swapState.phase = 0;
    poll_launches();
    std::lock_guard<std::mutex> lock(mu);
    publish_counters();
}

// ---------------------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------------------
extern "C" {

void bp_server_record_sizes(uint32_t *out) {
    if (!out) return;
    out[0] = sizeof(bp_request); out[1] = sizeof(bp_server_result); out[2] = sizeof(bp_server_config);
    out[3] = sizeof(bp_server_stats);
}

void bp_server_config_default(bp_server_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->memory_budget = 0;
    c->streams = 32;
    c->max_inflight_per_class = 32;
    c->max_worlds_per_launch = 1024;
    c->big_world = 4096;
}

bp_status bp_server_create(bp_batch *batch, const bp_server_config *config, bp_server **out) {
    if (!batch || !out) return BP_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lock(batch->mutex);
        if (batch->inFlight || batch->pending || batch->serving) return BP_ERR_FRAME_IN_FLIGHT;
        batch->serving = true;
    }
    bp_server *s = new bp_server();
    s->batch = batch;
    bp_server_config_default(&s->cfg);
    if (config) {
        s->cfg = *config;
        if (s->cfg.streams == 0) s->cfg.streams = 32;
        if (s->cfg.max_inflight_per_class == 0) s->cfg.max_inflight_per_class = 32;
        if (s->cfg.max_worlds_per_launch == 0) s->cfg.max_worlds_per_launch = 1024;
        if (s->cfg.big_world == 0) s->cfg.big_world = 4096;
        s->cfg.streams = std::min<uint32_t>(s->cfg.streams, 32);
    }
    (void)hipGetDevice(&s->device);
    s->t0 = Clock::now();
    bp_status st = s->init();
    if (st != BP_OK) {
        s->release();
        batch->serving = false;
        delete s;
        return st;
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return BP_OK;
}

void bp_server_destroy(bp_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cvWork.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipDeviceSynchronize();
    bp_batch &b = *s->batch;
    // Hand back host mirrors with the device's counts and properties (transactions ran on the
    // device); the next batch step repacks with the classes the new counts select.
    if (b.download_state(true) == BP_OK) b.packed = false;
    s->release();
    b.serving = false;
    delete s;
}

bp_status bp_server_submit(bp_server *s, const bp_request *r) {
    if (!s || !r) return BP_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        if (s->fatal != BP_OK) { set_error(s->fatalText); return s->fatal; }
        if (s->stop) return BP_ERR_UNAVAILABLE;
        s->incoming.push_back(Pending{*r, s->now(), -1});
        ++s->submitted;
        s->st.submitted = s->submitted;
    }
    s->cvWork.notify_one();
    return BP_OK;
}

bp_status bp_server_poll(bp_server *s, bp_server_result *out, uint32_t capacity, uint32_t *out_count) {
    if (!s || (capacity && !out)) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    const uint32_t n = (uint32_t)std::min<size_t>(capacity, s->results.size());
    std::copy(s->results.begin(), s->results.begin() + n, out);
    s->results.erase(s->results.begin(), s->results.begin() + n);
    if (out_count) *out_count = n;
    if (s->fatal != BP_OK) { set_error(s->fatalText); return s->fatal; }
    return BP_OK;
}

bp_status bp_server_wait_idle(bp_server *s) {
    if (!s) return BP_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(s->mu);
    s->cvDone.wait(lock, [&] { return s->completed >= s->submitted || s->fatal != BP_OK; });
    if (s->fatal != BP_OK) { set_error(s->fatalText); return s->fatal; }
    return BP_OK;
}

bp_status bp_server_run_trace(bp_server *s, const bp_request *trace, uint32_t count, bp_server_result *results) {
    if (!s || (count && (!trace || !results))) return BP_ERR_INVALID_ARGUMENT;
    {
        std::unique_lock<std::mutex> lock(s->mu);
        if (s->sinkResults) return BP_ERR_FRAME_IN_FLIGHT;
        s->cvDone.wait(lock, [&] { return s->completed >= s->submitted || s->fatal != BP_OK; });
        s->sinkResults = results;
    }
    const double base = s->now();
    for (uint32_t i = 0; i < count; ++i) {
        const double due = base + trace[i].arrival;
        for (;;) {
            const double left = due - s->now();
            if (left <= 0.0) break;
            if (left > 3e-4) std::this_thread::sleep_for(std::chrono::duration<double>(left - 2e-4));
            else std::this_thread::yield();
        }
        {
            std::lock_guard<std::mutex> lock(s->mu);
            if (s->fatal != BP_OK) break;
            s->incoming.push_back(Pending{trace[i], due, (int64_t)i});
            ++s->submitted;
            s->st.submitted = s->submitted;
        }
        s->cvWork.notify_one();
    }
    std::unique_lock<std::mutex> lock(s->mu);
    s->cvDone.wait(lock, [&] { return s->completed >= s->submitted || s->fatal != BP_OK; });
    s->sinkResults = nullptr;
    if (s->fatal != BP_OK) { set_error(s->fatalText); return s->fatal; }
    return BP_OK;
}

bp_status bp_server_read_bubbles(bp_server *s, uint32_t world, bp_bubble *out, uint32_t capacity, uint32_t *out_count) {
    if (!s || world >= s->W) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->completed < s->submitted || !s->incoming.empty()) return BP_ERR_FRAME_IN_FLIGHT;
    bp_batch &b = *s->batch;
    const uint32_t n = s->count[world];
    if (out_count) *out_count = n;
    if (!out) return BP_OK;
    if (capacity < n) return BP_ERR_CAPACITY_OVERFLOW;
    if (!s->resident[world]) {
        std::copy(b.worlds[world].bubbles.begin(), b.worlds[world].bubbles.end(), out);
        return BP_OK;
    }
    if (n == 0) return BP_OK;
    if (s->readCapacity < n) {
        bp_status st = dev_alloc(s, s->readBuffer, n);
        if (st != BP_OK) return st;
        s->readCapacity = n;
    }
    hipError_t e = bp_launch_gather_full_world(b.arenaDev, b.descs + world, s->readBuffer, n, s->readStream);
    if (e != hipSuccess) { set_error(std::string("gather: ") + hipGetErrorString(e)); return BP_ERR_COMMAND_FAILED; }
    BP_HIP(hipMemcpyAsync(out, s->readBuffer, n * sizeof(bp_bubble), hipMemcpyDeviceToHost, s->readStream));
    BP_HIP(hipStreamSynchronize(s->readStream));
    return BP_OK;
}

bp_status bp_server_get_stats(bp_server *s, bp_server_stats *out) {
    if (!s || !out) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->st;
    return BP_OK;
}

double bp_server_clock(const bp_server *s) { return s ? s->now() : 0.0; }

}  // extern "C"
