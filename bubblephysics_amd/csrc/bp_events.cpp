// bp_events.cpp -- contact transitions of one step (bp_contact_transitions): the contactBegan and
// contactEnded events the public world derives from the step's contact records
// (NativePublicWorld.swift:340-391, contactTransitions).
//
// Per record of the step (in record order): the pair key is (bubble A, bubble B) for a bubble pair
// and (bubble A, the segment's owner) for a bubble-segment contact, ordered by raw ObjectID
// (Identifiers.swift:56-65); a segment without an owner gives no key. A record counts when
// penetration > 0 or its impulse > 0, and only when an end of the key is of interest. Per key the
// sums run in record order: point += pointQ * w and weight += w with w = max(penetration, 0) +
// 1e-4, normal += the normal from the key's first object to its second (a record's normal points
// from A to B, or from the segment to A), impulse += F dt^2 / dt. Began events, for keys not in
// the previous set, come in key order with point = point sum / weight sum, normal = the sum
// normalised (zero when the sum is zero) and the impulse sum; ended events follow, in key order,
// for keys of the previous set that are not current. The current keys become the new set. With
// no object of interest and an empty previous set there are no events and the set stays empty.
// Bubble-body records (BP_CONTACT_BODY) and the rigid solver's body-pair and wall records
// (BP_CONTACT_RIGID) are the reference's rigid records (:377-380): key (first, second), weight 1,
// their normal from first to second, impulse / dt, and no penetration test.
#include <algorithm>
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "../../include/bubblephysics.h"

namespace {

using Key = std::pair<uint64_t, uint64_t>;

struct Sum {
    float px = 0.0f, py = 0.0f, weight = 0.0f, nx = 0.0f, ny = 0.0f, impulse = 0.0f;
};

}  // namespace

extern "C" bp_status bp_contact_transitions(const bp_contact *contacts, uint32_t count, const bp_segment *segments,
                                            uint32_t segment_count, const uint64_t *interested, uint32_t interested_count,
                                            float step_duration, uint64_t *history, uint32_t history_count,
                                            uint32_t history_capacity, uint32_t *out_history_count,
                                            bp_contact_event *events, uint32_t event_capacity, uint32_t *out_event_count) {
    if ((count && !contacts) || (segment_count && !segments) || (interested_count && !interested) ||
        (history_count && !history) || !out_history_count || !out_event_count || (event_capacity && !events))
        return BP_ERR_INVALID_ARGUMENT;
    *out_history_count = 0;
    *out_event_count = 0;
    if (interested_count == 0 && history_count == 0) return BP_OK;
    std::vector<uint64_t> interest(interested, interested + interested_count);
    std::sort(interest.begin(), interest.end());
    auto is_interesting = [&](uint64_t id) { return std::binary_search(interest.begin(), interest.end(), id); };
    // Segments are sorted by id within a world; look owners up by id.
    auto owner_of = [&](int64_t segmentId, uint64_t &owner) -> bool {
        const bp_segment *end = segments + segment_count;
        const bp_segment *s = std::lower_bound(segments, end, segmentId, [](const bp_segment &a, int64_t id) { return a.id < id; });
        if (s == end || s->id != segmentId || !(s->flags & BP_SEGMENT_HAS_OWNER)) return false;
        owner = (uint64_t)s->owner_id;
        return true;
    };
    std::map<Key, Sum> sums;
    for (uint32_t k = 0; k < count; ++k) {
        const bp_contact &c = contacts[k];
        const uint64_t first = (uint64_t)c.first;
        const bool rigid = c.kind == BP_CONTACT_BODY || c.kind == BP_CONTACT_RIGID;
        uint64_t second;
        if (c.kind == BP_CONTACT_PAIR || rigid) second = (uint64_t)c.second;
        else if (!owner_of(c.second, second)) continue;
        if (!rigid && !(c.penetration > 0.0f || c.impulse > 0.0f)) continue;
        if (!is_interesting(first) && !is_interesting(second)) continue;
        const float w = rigid ? 1.0f : std::fmax(c.penetration, 0.0f) + 1e-4f;
        float nx = c.normal[0], ny = c.normal[1];
        if (c.kind == BP_CONTACT_SEGMENT) { nx = -nx; ny = -ny; }       // from the segment to A: A to the owner
        const Key key = first <= second ? Key(first, second) : Key(second, first);
        if (key.first != first) { nx = -nx; ny = -ny; }
        Sum &s = sums[key];
        s.px = s.px + c.point[0] * w;
        s.py = s.py + c.point[1] * w;
        s.weight = s.weight + w;
        s.nx = s.nx + nx;
        s.ny = s.ny + ny;
        s.impulse = s.impulse + c.impulse / step_duration;
    }
    std::vector<Key> previous(history_count);
    for (uint32_t k = 0; k < history_count; ++k) previous[k] = Key(history[2 * k], history[2 * k + 1]);
    std::sort(previous.begin(), previous.end());
    uint32_t written = 0, total = 0;
    for (const auto &entry : sums) {                                  // std::map iterates in key order
        if (std::binary_search(previous.begin(), previous.end(), entry.first)) continue;
        const Sum &s = entry.second;
        const float length = std::sqrt(s.nx * s.nx + s.ny * s.ny);
        if (written < event_capacity) {
            bp_contact_event &e = events[written++];
            e = bp_contact_event{};
            e.kind = BP_EVENT_CONTACT_BEGAN;
            e.first = entry.first.first;
            e.second = entry.first.second;
            e.point[0] = s.px / s.weight;
            e.point[1] = s.py / s.weight;
            e.normal[0] = length > 0.0f ? s.nx / length : 0.0f;
            e.normal[1] = length > 0.0f ? s.ny / length : 0.0f;
            e.impulse = s.impulse;
        }
        ++total;
    }
    for (const Key &key : previous) {
        if (sums.count(key)) continue;
        if (written < event_capacity) {
            bp_contact_event &e = events[written++];
            e = bp_contact_event{};
            e.kind = BP_EVENT_CONTACT_ENDED;
            e.first = key.first;
            e.second = key.second;
        }
        ++total;
    }
    *out_event_count = total;
    uint32_t h = 0;
    for (const auto &entry : sums) {
        if (h < history_capacity) { history[2 * h] = entry.first.first; history[2 * h + 1] = entry.first.second; }
        ++h;
    }
    *out_history_count = h;
    return h > history_capacity ? BP_ERR_CAPACITY_OVERFLOW : BP_OK;
}
