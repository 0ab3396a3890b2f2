// bp_stager.cpp -- structural staging of merge and split on the host.
//
// The reference stages these transactions on the CPU before the snapshot reaches the solver
// (TransactionStager.swift:630-735, StructuralOperations.swift:10-151); the result's centre is
// checked, and moved when needed, by the placement rules of PlacementGeometry.swift:42-159. This
// file states the same arithmetic (FP64 where the reference uses Double, FP32 elsewhere, in the
// same order) over the C records of one world, and returns the world's new bubble list. The
// contour intent the reference attaches (structuralShape) belongs to rendering and is not kept.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bubblephysics.h"

static_assert(sizeof(bp_stage_scene) == 56 && sizeof(bp_stage_rejection) == 16, "staging records changed size");

namespace {

struct P2 { float x, y; };

bool collides(uint32_t catA, uint32_t maskA, uint32_t catB, uint32_t maskB) {   // Identifiers.swift:50-52
    return (catA & maskB) != 0u && (catB & maskA) != 0u;
}

// Scene of one world for the placement tests: bubbles that stay, bodies, the region.
struct Scene {
    const bp_stage_scene *s;
    std::vector<const bp_bubble *> bubbles;     // the bubbles a placement is tested against
};

float segment_distance(P2 p, P2 a, P2 b) {                       // PlacementGeometry.swift:96-101
    const P2 e{b.x - a.x, b.y - a.y};
    const float sq = e.x * e.x + e.y * e.y;
    const float t = sq > 0.0f ? std::max(0.0f, std::min(1.0f, ((p.x - a.x) * e.x + (p.y - a.y) * e.y) / sq)) : 0.0f;
    const P2 q{p.x - (a.x + e.x * t), p.y - (a.y + e.y * t)};
    return std::sqrt(q.x * q.x + q.y * q.y);
}

// Point in a body's world shape, boundary included (PlacementGeometry.swift:82-158).
bool body_contains(const bp_stage_scene &s, const bp_body &b, P2 p) {
    const float c = std::cos(b.pose[2]), sn = std::sin(b.pose[2]);
    auto world = [&](float x, float y) { return P2{b.pose[0] + x * c - y * sn, b.pose[1] + x * sn + y * c}; };
    if (b.shape == BP_SHAPE_CIRCLE) {
        const P2 d{p.x - b.pose[0], p.y - b.pose[1]};
        return d.x * d.x + d.y * d.y <= b.core_radius * b.core_radius;
    }
    if (b.shape == BP_SHAPE_CAPSULE)
        return segment_distance(p, world(-b.half_length, 0.0f), world(b.half_length, 0.0f)) <= b.core_radius;
    const uint32_t nv = b.vertex_count;
    if (nv == 0 || !s.vertices || (uint64_t)b.vertex_offset + nv > s.vertex_count) return false;
    std::vector<P2> v(nv);
    for (uint32_t k = 0; k < nv; ++k) v[k] = world(s.vertices[2 * (b.vertex_offset + k)], s.vertices[2 * (b.vertex_offset + k) + 1]);
    for (uint32_t i = 0, j = nv - 1; i < nv; j = i++) {           // on an edge or a vertex
        const P2 a = v[i], e = v[j];
        const float cross = (e.x - a.x) * (p.y - a.y) - (e.y - a.y) * (p.x - a.x);
        if (cross == 0.0f && p.x >= std::min(a.x, e.x) && p.x <= std::max(a.x, e.x) && p.y >= std::min(a.y, e.y) &&
            p.y <= std::max(a.y, e.y))
            return true;
    }
    bool inside = false;                                           // crossing parity
    for (uint32_t i = 0, j = nv - 1; i < nv; j = i++) {
        const P2 a = v[i], e = v[j];
        if ((a.y > p.y) != (e.y > p.y) && p.x <= (e.x - a.x) * (p.y - a.y) / (e.y - a.y) + a.x) inside = !inside;
    }
    return inside;
}

// First reason a bubble of radius r and filter (cat, mask) cannot stand at p: region, then a
// rigid body, then another bubble's centre (PlacementGeometry.swift:51-94).
uint32_t placement_issue(const Scene &sc, P2 p, float r, uint32_t cat, uint32_t mask) {
    const bp_stage_scene &s = *sc.s;
    if (!(p.x >= s.region[0] && p.x <= s.region[2] && p.y >= s.region[1] && p.y <= s.region[3]))
        return BP_PLACEMENT_OUTSIDE_REGION;
    for (uint32_t k = 0; k < s.body_count; ++k)
        if (collides(cat, mask, s.bodies[k].category, s.bodies[k].mask) && body_contains(s, s.bodies[k], p))
            return BP_PLACEMENT_INSIDE_RIGID_BODY;
    for (const bp_bubble *o : sc.bubbles) {
        if (!collides(cat, mask, o->category, o->mask)) continue;
        const P2 d{p.x - o->center[0], p.y - o->center[1]};
        const float tol = std::max(1e-4f * std::min(r, o->radius), 1e-4f);   // coincidentCenterTolerance
        if (std::sqrt(d.x * d.x + d.y * d.y) < tol) return BP_PLACEMENT_COINCIDENT_CENTER;
    }
    return BP_PLACEMENT_OK;
}

// 8 rings of 16 points at radii j * step around c (StructuralOperations.swift:40-49).
void spiral(P2 c, float step, std::vector<P2> &out) {
    for (int ring = 1; ring <= 8; ++ring)
        for (int k = 0; k < 16; ++k) {
            const float angle = (float)k * (2.0f * (float)M_PI / 16.0f), radius = (float)ring * step;
            out.push_back(P2{c.x + std::cos(angle) * radius, c.y + std::sin(angle) * radius});
        }
}

bp_status reject(bp_stage_rejection *why, uint32_t reason, uint32_t issue = BP_PLACEMENT_OK, int64_t id = 0) {
    if (why) { why->reason = reason; why->issue = issue; why->object_id = id; }
    return BP_ERR_INVALID_ARGUMENT;
}

const bp_bubble *find(const bp_stage_scene &s, int64_t id) {
    for (uint32_t i = 0; i < s.bubble_count; ++i)
        if (s.bubbles[i].id == id) return &s.bubbles[i];
    return nullptr;
}

// A new bubble from a result template (resultDescriptor, TransactionStager.swift:621-629).
bp_bubble make_result(const bp_bubble &tpl, P2 c, P2 v, float radius, float mass) {
    bp_bubble b = tpl;
    b.center[0] = b.previous_center[0] = c.x;
    b.center[1] = b.previous_center[1] = c.y;
    b.velocity[0] = v.x; b.velocity[1] = v.y;
    b.rotation = 0.0f; b.angular_velocity = 0.0f;
    b.radius = radius; b.mass = mass;
    b.inverse_mass = (tpl.flags & BP_BUBBLE_STATIC) || !(mass > 0.0f) ? 0.0f : 1.0f / mass;
    return b;
}

// The world's list without `removed`, plus `added`, sorted by id; a duplicate id is refused.
bp_status assemble(const bp_stage_scene &s, const std::vector<int64_t> &removed, const std::vector<bp_bubble> &added,
                   bp_bubble *out, uint32_t capacity, uint32_t *count, bp_stage_rejection *why) {
    std::vector<bp_bubble> all;
    all.reserve(s.bubble_count + added.size());
    for (uint32_t i = 0; i < s.bubble_count; ++i)
        if (std::find(removed.begin(), removed.end(), s.bubbles[i].id) == removed.end()) all.push_back(s.bubbles[i]);
    for (const bp_bubble &a : added) {
        for (const bp_bubble &b : all)
            if (b.id == a.id) return reject(why, BP_REJECT_DUPLICATE_ID, BP_PLACEMENT_OK, a.id);
        all.push_back(a);
    }
    std::stable_sort(all.begin(), all.end(), [](const bp_bubble &a, const bp_bubble &b) { return a.id < b.id; });
    if (count) *count = (uint32_t)all.size();
    if (!out || capacity < all.size()) return BP_ERR_CAPACITY_OVERFLOW;
    std::memcpy(out, all.data(), all.size() * sizeof(bp_bubble));
    if (why) std::memset(why, 0, sizeof *why);
    return BP_OK;
}

}  // namespace

extern "C" bp_status bp_stage_add(const bp_stage_scene *scene, const bp_bubble *bubble, bp_bubble *out, uint32_t out_capacity,
                                  uint32_t *out_count, bp_stage_rejection *why) {
    if (!scene || !bubble || (scene->bubble_count && !scene->bubbles) || (scene->body_count && !scene->bodies))
        return BP_ERR_INVALID_ARGUMENT;
    if (find(*scene, bubble->id)) return reject(why, BP_REJECT_DUPLICATE_ID, BP_PLACEMENT_OK, bubble->id);   // :407
    Scene sc{scene, {}};
    for (uint32_t i = 0; i < scene->bubble_count; ++i) sc.bubbles.push_back(&scene->bubbles[i]);
    const uint32_t issue = placement_issue(sc, P2{bubble->center[0], bubble->center[1]}, bubble->radius, bubble->category,
                                           bubble->mask);                                                   // :415-420
    if (issue != BP_PLACEMENT_OK) return reject(why, BP_REJECT_INVALID_PLACEMENT, issue, bubble->id);
    return assemble(*scene, {}, {*bubble}, out, out_capacity, out_count, why);
}

extern "C" bp_status bp_stage_merge(const bp_stage_scene *scene, const int64_t *sources, uint32_t source_count,
                                    const bp_bubble *result, int32_t result_mass_given, bp_bubble *out,
                                    uint32_t out_capacity, uint32_t *out_count, bp_stage_rejection *why) {
    if (!scene || !sources || !result || (scene->bubble_count && !scene->bubbles) || (scene->body_count && !scene->bodies))
        return BP_ERR_INVALID_ARGUMENT;
    const bp_stage_scene &s = *scene;
    std::vector<int64_t> ids(sources, sources + source_count);
    std::vector<int64_t> unique = ids;
    std::sort(unique.begin(), unique.end());
    if (ids.size() < 2 || std::adjacent_find(unique.begin(), unique.end()) != unique.end())
        return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    double mass = 0.0, radiusSquared = 0.0, mx = 0.0, my = 0.0;           // TransactionStager.swift:632-640
    for (int64_t id : ids) {
        const bp_bubble *b = find(s, id);
        if (!b) return reject(why, BP_REJECT_UNKNOWN_OBJECT, BP_PLACEMENT_OK, id);
        const double w = (double)b->mass;
        mass += w;
        radiusSquared += (double)b->radius * (double)b->radius;
        mx += (double)b->velocity[0] * w;
        my += (double)b->velocity[1] * w;
    }
    std::vector<const bp_bubble *> in;                                     // sources in id order
    for (int64_t id : unique) in.push_back(find(s, id));
    double cx = 0.0, cy = 0.0, pairs = 0.0;                                // mergeCenter, StructuralOperations.swift:10-22
    for (size_t i = 0; i < in.size(); ++i)
        for (size_t j = i + 1; j < in.size(); ++j) {
            const bp_bubble &a = *in[i], &b = *in[j];
            const P2 d{b.center[0] - a.center[0], b.center[1] - a.center[1]};
            const float dist = std::hypot(d.x, d.y);
            const P2 n = dist > 0.0f ? P2{d.x / dist, d.y / dist} : P2{1.0f, 0.0f};
            const float dr = a.radius - b.radius;
            const P2 q{(a.center[0] + b.center[0] + n.x * dr) * 0.5f, (a.center[1] + b.center[1] + n.y * dr) * 0.5f};
            cx += (double)q.x; cy += (double)q.y; pairs += 1.0;
        }
    const P2 contact{(float)(cx / pairs), (float)(cy / pairs)};
    const float resultRadius = (float)std::sqrt(radiusSquared);
    const P2 velocity{(float)(mx / mass), (float)(my / mass)};
    Scene sc{scene, {}};
    for (uint32_t i = 0; i < s.bubble_count; ++i)
        if (std::find(ids.begin(), ids.end(), s.bubbles[i].id) == ids.end()) sc.bubbles.push_back(&s.bubbles[i]);
    auto issue = [&](P2 p) { return placement_issue(sc, p, resultRadius, result->category, result->mask); };
    // resolvedMergeCenter and mergeCandidates (StructuralOperations.swift:23-58): the contact
    // point, then P + t (S_i - P) for t = 0.25, 0.5, 0.75 per source, the sources' centres, then
    // 8 rings around the heaviest source (the first of equal masses).
    P2 center = contact;
    const uint32_t first = issue(contact);
    if (first != BP_PLACEMENT_OK) {
        std::vector<P2> cand;
        for (const bp_bubble *b : in)
            for (float t : {0.25f, 0.5f, 0.75f})
                cand.push_back(P2{contact.x + (b->center[0] - contact.x) * t, contact.y + (b->center[1] - contact.y) * t});
        for (const bp_bubble *b : in) cand.push_back(P2{b->center[0], b->center[1]});
        const bp_bubble *heaviest = in[0];
        for (size_t k = 1; k < in.size(); ++k)
            if (in[k]->mass > heaviest->mass) heaviest = in[k];
        spiral(P2{heaviest->center[0], heaviest->center[1]}, resultRadius / 4.0f, cand);
        bool found = false;
        for (P2 p : cand)
            if (issue(p) == BP_PLACEMENT_OK) { center = p; found = true; break; }
        if (!found) return reject(why, BP_REJECT_INVALID_PLACEMENT, first, result->id);
    }
    bp_bubble merged = make_result(*result, center, velocity, resultRadius, result_mass_given ? result->mass : (float)mass);
    if (result_mass_given) merged.inverse_mass = (result->flags & BP_BUBBLE_STATIC) || !(result->mass > 0.0f) ? 0.0f : 1.0f / result->mass;
    // initializeMerge (StructuralOperations.swift:59-69): angular velocity from the sources'
    // angular momentum about the result's centre and the sources' total mass.
    double L = 0.0, sourceMass = 0.0;
    for (const bp_bubble *b : in) {
        const double m = (double)b->mass, r = (double)b->radius;
        const P2 d{b->center[0] - center.x, b->center[1] - center.y};
        sourceMass += m;
        L += 0.5 * m * r * r * (double)b->angular_velocity + m * ((double)d.x * (double)b->velocity[1] - (double)d.y * (double)b->velocity[0]);
    }
    merged.angular_velocity = (float)(L / (0.5 * sourceMass * (double)resultRadius * (double)resultRadius));
    return assemble(s, ids, {merged}, out, out_capacity, out_count, why);
}

extern "C" bp_status bp_stage_split(const bp_stage_scene *scene, int64_t source, const bp_bubble *first, float first_share,
                                    const bp_bubble *second, float second_share, bp_bubble *out, uint32_t out_capacity,
                                    uint32_t *out_count, bp_stage_rejection *why) {
    if (!scene || !first || !second || (scene->bubble_count && !scene->bubbles) || (scene->body_count && !scene->bodies))
        return BP_ERR_INVALID_ARGUMENT;
    const bp_stage_scene &s = *scene;
    const double total = (double)first_share + (double)second_share;     // TransactionStager.swift:675-685
    if (!std::isfinite(first_share) || !std::isfinite(second_share) || !(first_share > 0.0f) || !(second_share > 0.0f) ||
        !(std::fabs(total - 1.0) <= 1e-6))
        return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    const double s0 = (double)first_share / total, s1 = (double)second_share / total;
    if (!(s0 > 0.0 && s0 < 1.0 && s1 > 0.0 && s1 < 1.0)) return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    const bp_bubble *srcp = find(s, source);
    if (!srcp) return reject(why, BP_REJECT_UNKNOWN_OBJECT, BP_PLACEMENT_OK, source);
    const bp_bubble src = *srcp;
    const P2 sc0{src.center[0], src.center[1]}, sv{src.velocity[0], src.velocity[1]};
    bp_bubble a = make_result(*first, sc0, sv, (float)((double)src.radius * std::sqrt(s0)), (float)((double)src.mass * s0));
    bp_bubble b = make_result(*second, sc0, sv, (float)((double)src.radius * std::sqrt(s1)), (float)((double)src.mass * s1));
    if (a.id == b.id) return reject(why, BP_REJECT_DUPLICATE_ID, BP_PLACEMENT_OK, a.id);
    for (uint32_t i = 0; i < s.bubble_count; ++i)
        if (s.bubbles[i].id != source && (s.bubbles[i].id == a.id || s.bubbles[i].id == b.id))
            return reject(why, BP_REJECT_DUPLICATE_ID, BP_PLACEMENT_OK, s.bubbles[i].id);
    // initializeSplit (StructuralOperations.swift:70-112): the chord that cuts the disc into the
    // two shares (48 bisection steps), the centroids of the two segments on the source's axis,
    // and the angular velocity that keeps the source's angular momentum.
    const double r = (double)src.radius, area = M_PI * r * r;
    double low = -r, high = r;
    for (int k = 0; k < 48; ++k) {
        const double x = (low + high) * 0.5;
        const double left = r * r * (std::asin(x / r) + M_PI / 2) + x * std::sqrt(std::max(0.0, r * r - x * x));
        if (left < area * s0) low = x; else high = x;
    }
    const double cut = (low + high) * 0.5, moment = 2.0 / 3 * std::pow(std::max(0.0, r * r - cut * cut), 1.5);
    const double x1 = -moment / (area * s0), x2 = moment / (area * s1);
    if (!std::isfinite(x1) || !std::isfinite(x2) || !std::isfinite((float)x1) || !std::isfinite((float)x2) || !(x1 < x2))
        return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    const P2 axis{std::cos(src.rotation), std::sin(src.rotation)};
    P2 pa{sc0.x + axis.x * (float)x1, sc0.y + axis.y * (float)x1}, pb{sc0.x + axis.x * (float)x2, sc0.y + axis.y * (float)x2};
    const double ownInertia = 0.5 * ((double)a.mass * (double)a.radius * (double)a.radius + (double)b.mass * (double)b.radius * (double)b.radius);
    const double inertia = ownInertia + (double)a.mass * x1 * x1 + (double)b.mass * x2 * x2;
    float omega = (float)(0.5 * (double)src.mass * (double)src.radius * (double)src.radius * (double)src.angular_velocity / inertia);
    if (!std::isfinite(omega)) return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    const P2 tangent{-axis.y, axis.x};
    P2 va{sv.x + tangent.x * ((float)x1 * omega), sv.y + tangent.y * ((float)x1 * omega)};
    P2 vb{sv.x + tangent.x * ((float)x2 * omega), sv.y + tangent.y * ((float)x2 * omega)};
    if (!std::isfinite(pa.x) || !std::isfinite(pa.y) || !std::isfinite(pb.x) || !std::isfinite(pb.y) || !std::isfinite(va.x) ||
        !std::isfinite(va.y) || !std::isfinite(vb.x) || !std::isfinite(vb.y) || (pa.x == pb.x && pa.y == pb.y))
        return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    // resolveSplitCenters (TransactionStager.swift:700-734): the first part is tested against the
    // world without the source and the parts, the second also against the first part's final centre.
    const P2 centroidA = pa, centroidB = pb;
    Scene scA{scene, {}};
    for (uint32_t i = 0; i < s.bubble_count; ++i)
        if (s.bubbles[i].id != source) scA.bubbles.push_back(&s.bubbles[i]);
    auto resolve = [&](const Scene &sc, const bp_bubble &part, P2 centroid, P2 &at) -> uint32_t {
        auto issue = [&](P2 p) { return placement_issue(sc, p, part.radius, part.category, part.mask); };
        const uint32_t cause = issue(centroid);
        if (cause == BP_PLACEMENT_OK) { at = centroid; return BP_PLACEMENT_OK; }
        std::vector<P2> cand;                                            // resolvedSplitCenter, :113-124
        for (float t : {0.25f, 0.5f, 0.75f}) cand.push_back(P2{centroid.x + (sc0.x - centroid.x) * t, centroid.y + (sc0.y - centroid.y) * t});
        cand.push_back(sc0);
        spiral(sc0, part.radius / 4.0f, cand);
        for (P2 p : cand)
            if (issue(p) == BP_PLACEMENT_OK) { at = p; return BP_PLACEMENT_OK; }
        return cause;
    };
    uint32_t cause = resolve(scA, a, centroidA, pa);
    if (cause != BP_PLACEMENT_OK) return reject(why, BP_REJECT_INVALID_PLACEMENT, cause, a.id);
    a.center[0] = pa.x; a.center[1] = pa.y;
    Scene scB = scA;
    scB.bubbles.push_back(&a);
    cause = resolve(scB, b, centroidB, pb);
    if (cause != BP_PLACEMENT_OK) return reject(why, BP_REJECT_INVALID_PLACEMENT, cause, b.id);
    if (pa.x != centroidA.x || pa.y != centroidA.y || pb.x != centroidB.x || pb.y != centroidB.y) {
        // relocateSplitParts (StructuralOperations.swift:125-151): velocities about the parts'
        // centre of mass keep the source's momentum and angular momentum.
        const double m1 = (double)a.mass, m2 = (double)b.mass;
        const double d1x = (double)pa.x - (double)sc0.x, d1y = (double)pa.y - (double)sc0.y;
        const double d2x = (double)pb.x - (double)sc0.x, d2y = (double)pb.y - (double)sc0.y;
        const double comx = (d1x * m1 + d2x * m2) / (m1 + m2), comy = (d1y * m1 + d2y * m2) / (m1 + m2);
        const double r1x = d1x - comx, r1y = d1y - comy, r2x = d2x - comx, r2y = d2y - comy;
        const double inertia2 = ownInertia + m1 * (r1x * r1x + r1y * r1y) + m2 * (r2x * r2x + r2y * r2y);
        omega = (float)(0.5 * (double)src.mass * (double)src.radius * (double)src.radius * (double)src.angular_velocity / inertia2);
        va = P2{sv.x + (float)(-r1y) * omega, sv.y + (float)r1x * omega};
        vb = P2{sv.x + (float)(-r2y) * omega, sv.y + (float)r2x * omega};
        if (!std::isfinite(omega) || !std::isfinite(va.x) || !std::isfinite(va.y) || !std::isfinite(vb.x) || !std::isfinite(vb.y))
            return reject(why, BP_REJECT_INVALID_STRUCTURAL_OPERATION);
    }
    a.center[0] = a.previous_center[0] = pa.x; a.center[1] = a.previous_center[1] = pa.y;
    b.center[0] = b.previous_center[0] = pb.x; b.center[1] = b.previous_center[1] = pb.y;
    a.velocity[0] = va.x; a.velocity[1] = va.y; b.velocity[0] = vb.x; b.velocity[1] = vb.y;
    a.angular_velocity = omega; b.angular_velocity = omega;
    return assemble(s, {source}, {a, b}, out, out_capacity, out_count, why);
}
