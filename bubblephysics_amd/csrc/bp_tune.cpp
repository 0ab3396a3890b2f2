// bp_tune.cpp -- autotuner of a batch's kernel configuration per world size.
//
// Worlds are binned by size (bin b holds the worlds of 2^b to 2^(b+1) - 1 bubbles). For each bin
// the tuner times one step of the bin's worlds in every kernel configuration that holds all of
// them: the LDS-resident configurations (workgroup width x bubbles per lane, with their row,
// segment, body and polygon capacities) and the one-workgroup global-state kernel, so the
// choice covers the workgroup shape and the cutover from LDS-resident to HBM-resident state.
// It then times the best configuration with half and twice the automatic hashed-grid bucket
// count (the cell width itself is fixed by the candidate rule, 2 (r + margin) + separation
// tolerance, the smallest width whose 3 x 3 cells hold every candidate; the bucket count sets
// how many cells share a bucket). Every trial starts from the same published state: a step
// writes the unpublished side and moves only the world descriptors' published-side mark, so
// restoring the descriptors restores the worlds. Results do not depend on the choice (every
// configuration gives the same bits), only the time does.
//
// The reference has no counterpart: its single-dispatch path has one fixed workgroup width
// (NativeMetalBackend.swift:17-21) and a fixed bucket count.
#include "bp_batch.h"

using namespace bph;

namespace {

constexpr uint32_t kBins = 24;

uint32_t bin_of(uint32_t n) {
    uint32_t b = 0;
    while ((2u << b) <= n && b + 1 < kBins) ++b;
    return b;
}

}  // namespace

// One trial: `steps`=1 of the listed worlds in class `cls` with the given bucket counts,
// from the saved descriptors. Returns bubble-substeps per second (0 when nothing ran).
static bp_status tune_trial(bp_batch &b, const std::vector<bpd::WorldDesc> &saved, const std::vector<uint32_t> &list,
                            int cls, int shift, double &rate) {
    std::vector<bpd::WorldDesc> descs = saved;
    for (uint32_t w : list) {
        uint32_t nb = auto_buckets(descs[w].bubbleCount);
        nb = shift < 0 ? std::max<uint32_t>(16, nb >> 1) : shift > 0 ? std::min(bpd::kMaxBuckets, nb << 1) : nb;
        descs[w].buckets = nb;
    }
    BP_HIP(hipMemcpy(b.descs, descs.data(), descs.size() * sizeof(bpd::WorldDesc), hipMemcpyHostToDevice));
    // The listed worlds move to class `cls` for this launch.
    for (uint32_t w : list) {
        auto &from = b.classWorlds[b.worlds[w].packedClass];
        from.erase(std::remove(from.begin(), from.end(), w), from.end());
        b.classWorlds[cls].push_back(w);
        b.worlds[w].packedClass = cls;
    }
    bp_status s = b.launch(list.data(), (uint32_t)list.size(), 1);
    if (s == BP_OK) s = b.finish(nullptr, true);
    if (s != BP_OK) return s;
    double work = 0.0;
    const size_t W = b.worlds.size();
    for (uint32_t w : list) {
        const bpd::WorldStatus &st = b.hostStats[w];
        (void)W;
        if (st.executed) work += (double)descs[w].bubbleCount * st.substeps;
    }
    const double ms = b.lastClassMs[cls];
    rate = ms > 0.0 ? work / (ms * 1e-3) : 0.0;
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
    if (s == BP_OK) s = b->download_state();        // host mirrors hold the state the batch resumes from
    std::vector<bpd::WorldDesc> saved;
    if (s == BP_OK) {
        saved.resize(b->worlds.size());
        hipError_t e = hipMemcpy(saved.data(), b->descs, saved.size() * sizeof(bpd::WorldDesc), hipMemcpyDeviceToHost);
        if (e != hipSuccess) { set_error(std::string("descriptor read: ") + hipGetErrorString(e)); s = BP_ERR_COMMAND_FAILED; }
    }
    std::vector<std::vector<uint32_t>> bins(kBins);
    if (s == BP_OK)
        for (uint32_t w = 0; w < b->worlds.size(); ++w) {
            const HostWorld &hw = b->worlds[w];
            const uint32_t n = (uint32_t)hw.bubbles.size();
            if (!hw.resident || n == 0 || hw.packedClass == kCoopClass || hw.packedClass < 0) continue;
            bins[bin_of(n)].push_back(w);
        }
    uint32_t written = 0;
    std::vector<int> bestClass(kBins, -1), bestShift(kBins, 0);
    for (uint32_t bin = 0; bin < kBins && s == BP_OK; ++bin) {
        const auto &list = bins[bin];
        if (list.empty()) continue;
        bp_tune_bin rec{};
        rec.min_bubbles = 1u << bin;
        rec.max_bubbles = bin + 1 < kBins ? (2u << bin) - 1 : 0xffffffffu;
        rec.worlds = (uint32_t)list.size();
        rec.best_class = -1;
        double best = -1.0;
        for (int k = 0; k <= kGlobalClass && s == BP_OK; ++k) {
            bool fits = true;
            for (uint32_t w : list) fits = fits && class_fits(k, b->worlds[w]);
            if (!fits) continue;
            std::vector<double> rates(repeats);
            for (uint32_t r = 0; r < repeats && s == BP_OK; ++r) s = tune_trial(*b, saved, list, k, 0, rates[r]);
            std::sort(rates.begin(), rates.end());
            const double med = rates[repeats / 2];
            rec.class_rate[k] = (float)med;
            if (med > best) { best = med; rec.best_class = k; }
        }
        rec.best_rate = (float)best;
        rec.best_bucket_shift = 0;
        for (int shift : {-1, 1}) {
            if (s != BP_OK || rec.best_class < 0) break;
            std::vector<double> rates(repeats);
            for (uint32_t r = 0; r < repeats && s == BP_OK; ++r) s = tune_trial(*b, saved, list, rec.best_class, shift, rates[r]);
            std::sort(rates.begin(), rates.end());
            const double med = rates[repeats / 2];
            rec.bucket_rate[shift < 0 ? 0 : 1] = (float)med;
            if (med > best * 1.02) { best = med; rec.best_bucket_shift = shift; rec.best_rate = (float)med; }
        }
        bestClass[bin] = rec.best_class;
        bestShift[bin] = rec.best_bucket_shift;
        if (written < capacity) out[written] = rec;
        ++written;
    }
    // Restore the published state and adopt the choices: the next step repacks from the host
    // mirrors with each world's tuned class and bucket count.
    if (s == BP_OK) {
        hipError_t e = hipMemcpy(b->descs, saved.data(), saved.size() * sizeof(bpd::WorldDesc), hipMemcpyHostToDevice);
        if (e != hipSuccess) { set_error(std::string("descriptor restore: ") + hipGetErrorString(e)); s = BP_ERR_COMMAND_FAILED; }
    }
    if (s == BP_OK) {
        for (uint32_t bin = 0; bin < kBins; ++bin)
            for (uint32_t w : bins[bin]) {
                HostWorld &hw = b->worlds[w];
                hw.forcedClass = bestClass[bin];
                if (bestShift[bin] != 0) {
                    const uint32_t nb = auto_buckets((uint32_t)hw.bubbles.size());
                    hw.config.grid_bucket_count = bestShift[bin] < 0 ? std::max<uint32_t>(16, nb >> 1)
                                                                    : std::min(bpd::kMaxBuckets, nb << 1);
                }
            }
        b->packed = false;
    }
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
