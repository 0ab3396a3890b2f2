// bp_tune.cpp -- autotuner of a batch's kernel configuration per world size.
//
// Worlds are binned by size (bin b holds the worlds of 2^b to 2^(b+1) - 1 bubbles). The tuner
// times whole-batch steps: the kernel classes of a batch run concurrently on their streams, so a
// bin's best configuration is the one that makes the whole batch's step shortest, not the one
// that runs the bin fastest alone. Bins are visited by decreasing work (bubbles); for each, every
// kernel configuration that holds all of its worlds is tried (the LDS-resident configurations,
// workgroup width x bubbles per lane with their row, segment, body and polygon capacities, and
// the one-workgroup global-state kernel, so the choice covers the workgroup shape and the
// cutover from LDS-resident to HBM-resident state), then the chosen one with half and twice the
// automatic hashed-grid bucket count (the cell width itself is fixed by the candidate rule,
// 2 (r + m) + separation tolerance, the smallest width whose 3 x 3 cells hold every candidate;
// the bucket count sets how many cells share a bucket). A change is kept only when the median
// whole-batch rate improves by more than 5 %, so a tuned batch is never slower than the default
// beyond timing noise. Every trial starts from the same published state: a step writes the
// unpublished side and moves only the world descriptors' published-side mark, so restoring the
// descriptors restores the worlds. Results do not depend on the choice (every configuration gives
// the same bits), only the time does.
//
// The reference has no counterpart: its single-dispatch path has one fixed workgroup width
// (NativeMetalBackend.swift:17-21) and a fixed bucket count.
#include "bp_batch.h"

using namespace bph;

namespace {

constexpr uint32_t kBins = 24;
// A choice is kept when it beats the incumbent's median whole-batch rate by this factor: the
// rate of one step of a batch varies by several percent from launch to launch (placement of
// workgroups on CUs), so a smaller margin adopts noise.
constexpr double kKeep = 1.05;

uint32_t bin_of(uint32_t n) {
    uint32_t b = 0;
    while ((2u << b) <= n && b + 1 < kBins) ++b;
    return b;
}

uint32_t shifted_buckets(uint32_t n, int shift) {
    const uint32_t nb = auto_buckets(n);
    return shift < 0 ? std::max<uint32_t>(16, nb >> 1) : shift > 0 ? std::min(bpd::kMaxBuckets, nb << 1) : nb;
}

// Moves the listed worlds to class `cls` (the class only selects the kernel; the arena layout is
// the same for every class).
void move_to(bp_batch &b, const std::vector<uint32_t> &list, const std::vector<int> &cls) {
    for (size_t k = 0; k < list.size(); ++k) {
        const uint32_t w = list[k];
        const int from = b.worlds[w].packedClass, to = cls[k];
        if (from == to) continue;
        auto &v = b.classWorlds[from];
        v.erase(std::remove(v.begin(), v.end(), w), v.end());
        b.classWorlds[to].push_back(w);
        b.worlds[w].packedClass = to;
    }
    b.orderDirty = true;
}

}  // namespace

// One trial: one step of every resident world from the saved published state with the working
// descriptors (bucket counts). Returns the batch's bubble-substeps per second of device time.
static bp_status tune_trial(bp_batch &b, const std::vector<bpd::WorldDesc> &working, double &rate) {
    BP_HIP(hipMemcpy(b.descs, working.data(), working.size() * sizeof(bpd::WorldDesc), hipMemcpyHostToDevice));
    bp_status s = b.launch(nullptr, 0, 1);
    if (s == BP_OK) s = b.finish(nullptr, true);
    if (s != BP_OK) return s;
    double work = 0.0;
    for (size_t w = 0; w < working.size(); ++w) {
        const bpd::WorldStatus &st = b.hostStats[w];
        if (st.executed && st.status == BP_OK) work += (double)working[w].bubbleCount * st.substeps;
    }
    rate = b.lastGpuMs > 0.0f ? work / (b.lastGpuMs * 1e-3) : 0.0;
    return BP_OK;
}

static bp_status median_trial(bp_batch &b, const std::vector<bpd::WorldDesc> &working, uint32_t repeats, double &rate) {
    std::vector<double> rates(repeats);
    for (uint32_t r = 0; r < repeats; ++r) {
        bp_status s = tune_trial(b, working, rates[r]);
        if (s != BP_OK) return s;
    }
    std::sort(rates.begin(), rates.end());
    rate = rates[repeats / 2];
    return BP_OK;
}

extern "C" {

bp_status bp_batch_autotune(bp_batch *b, uint32_t repeats, bp_tune_bin *out, uint32_t capacity, uint32_t *out_count) {
    if (!b || (capacity && !out)) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    repeats = std::max<uint32_t>(1, repeats);
    b->inFlight = true;
    bp_status s = BP_OK;
    if (!b->packed) s = b->pack();
    b->release_graph();                              // trials move worlds between the captured class lists
    if (s == BP_OK) s = b->download_state();        // host mirrors hold the state the batch resumes from
    std::vector<bpd::WorldDesc> saved;
    if (s == BP_OK) {
        saved.resize(b->worlds.size());
        hipError_t e = hipMemcpy(saved.data(), b->descs, saved.size() * sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_error(std::string("descriptor read: ") + hipGetErrorString(e)); s = BP_ERR_COMMAND_FAILED; }
    }
    std::vector<std::vector<uint32_t>> bins(kBins);
    std::vector<double> binWork(kBins, 0.0);
    if (s == BP_OK)
        for (uint32_t w = 0; w < b->worlds.size(); ++w) {
            const HostWorld &hw = b->worlds[w];
            const uint32_t n = (uint32_t)hw.bubbles.size();
            if (!hw.resident || n == 0 || hw.packedClass == kCoopClass || hw.packedClass < 0) continue;
            bins[bin_of(n)].push_back(w);
            binWork[bin_of(n)] += n;
        }
    std::vector<uint32_t> order;
    for (uint32_t bin = 0; bin < kBins; ++bin)
        if (!bins[bin].empty()) order.push_back(bin);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return binWork[x] > binWork[y]; });
    std::vector<bpd::WorldDesc> working = saved;
    std::vector<bp_tune_bin> records(kBins);
    std::vector<int> bestClass(kBins, -1), bestShift(kBins, 0);
    double incumbent = 0.0;
    if (s == BP_OK) s = median_trial(*b, working, repeats, incumbent);
    for (uint32_t bin : order) {
        if (s != BP_OK) break;
        const auto &list = bins[bin];
        bp_tune_bin &rec = records[bin];
        rec.min_bubbles = 1u << bin;
        rec.max_bubbles = bin + 1 < kBins ? (2u << bin) - 1 : 0xffffffffu;
        rec.worlds = (uint32_t)list.size();
        std::vector<int> original(list.size());
        for (size_t k = 0; k < list.size(); ++k) original[k] = b->worlds[list[k]].packedClass;
        double best = incumbent;
        int chosen = -1;
        for (int k = 0; k <= kGlobalClass && s == BP_OK; ++k) {
            bool fits = true;
            for (uint32_t w : list) fits = fits && class_fits(k, b->worlds[w]);
            if (!fits) continue;
            move_to(*b, list, std::vector<int>(list.size(), k));
            double r = 0.0;
            s = median_trial(*b, working, repeats, r);
            rec.class_rate[k] = (float)r;
            if (r > best * kKeep) { best = r; chosen = k; }
        }
        if (s != BP_OK) break;
        move_to(*b, list, chosen >= 0 ? std::vector<int>(list.size(), chosen) : original);
        for (int shift : {-1, 1}) {
            if (s != BP_OK) break;
            std::vector<bpd::WorldDesc> trial = working;
            for (uint32_t w : list) trial[w].buckets = shifted_buckets(trial[w].bubbleCount, shift);
            double r = 0.0;
            s = median_trial(*b, trial, repeats, r);
            rec.bucket_rate[shift < 0 ? 0 : 1] = (float)r;
            if (r > best * kKeep) { best = r; rec.best_bucket_shift = shift; }
        }
        if (rec.best_bucket_shift != 0)
            for (uint32_t w : list) working[w].buckets = shifted_buckets(working[w].bubbleCount, rec.best_bucket_shift);
        rec.best_class = chosen;                        // -1: the automatic classes stay
        rec.best_rate = (float)best;
        bestClass[bin] = chosen;
        bestShift[bin] = rec.best_bucket_shift;
        incumbent = best;
    }
    uint32_t written = 0;
    for (uint32_t bin = 0; bin < kBins; ++bin) {
        if (bins[bin].empty()) continue;
        if (written < capacity) out[written] = records[bin];
        ++written;
    }
    // Restore the published state and adopt the choices: the next step repacks from the host
    // mirrors with each world's tuned class and bucket count.
    if (!saved.empty()) {
        hipError_t e = hipMemcpy(b->descs, saved.data(), saved.size() * sizeof(bpd::WorldDesc), hipMemcpyHostToDevice);
        if (e != hipSuccess && s == BP_OK) { set_error(std::string("descriptor restore: ") + hipGetErrorString(e)); s = BP_ERR_COMMAND_FAILED; }
    }
    if (s == BP_OK) {
        for (uint32_t bin = 0; bin < kBins; ++bin)
            for (uint32_t w : bins[bin]) {
                HostWorld &hw = b->worlds[w];
                if (bestClass[bin] >= 0) hw.forcedClass = bestClass[bin];
                if (bestShift[bin] != 0)
                    hw.config.grid_bucket_count = shifted_buckets((uint32_t)hw.bubbles.size(), bestShift[bin]);
            }
    }
    b->packed = false;
    b->inFlight = false;
    if (out_count) *out_count = written;
    return s;
}

bp_status bp_batch_apply_tuning(bp_batch *b, const bp_tune_bin *bins, uint32_t count) {
    if (!b || (count && !bins)) return BP_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(b->mutex);
    if (b->inFlight || b->pending || b->serving) return BP_ERR_FRAME_IN_FLIGHT;
    bp_status s = b->download_state();
    if (s != BP_OK) return s;
    for (HostWorld &hw : b->worlds) {
        const uint32_t n = (uint32_t)hw.bubbles.size();
        for (uint32_t k = 0; k < count; ++k) {
            const bp_tune_bin &t = bins[k];
            if (n < t.min_bubbles || n > t.max_bubbles) continue;
            if (t.best_class >= 0 && t.best_class < kClassCount && t.best_class != kCoopClass && class_fits(t.best_class, hw))
                hw.forcedClass = t.best_class;
            if (t.best_bucket_shift != 0) {
                const uint32_t nb = auto_buckets(n);
                hw.config.grid_bucket_count = t.best_bucket_shift < 0 ? std::max<uint32_t>(16, nb >> 1)
                                                                    : std::min(bpd::kMaxBuckets, nb << 1);
            }
            break;
        }
    }
    b->packed = false;
    return BP_OK;
}

}  // extern "C"
