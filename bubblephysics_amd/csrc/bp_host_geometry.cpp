// bp_host_geometry.cpp -- host-side preparation of static and rigid geometry.
//
// * bp_static_polygon_edges: a static polygon becomes one-sided solver segments
//   (Sources/BubblePhysicsNativeMetal/NativeStaticPolygonEdges.swift:16-72): signed area and
//   convexity decided in double, collinear merge when |e1 x e2| <= 1e-5 |e1||e2| and e1.e2 > 0,
//   outward sign of the left normal, the previous edge's start for the vertex region.
// * bp_polygon_body_geometry: a polygon rigid shape becomes centroid-relative vertices and an
//   ear-clipped triangulation (NativeRigidGeometry.init, NativeRigidPolygonState.swift:89-133).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/bubblephysics.h"

namespace {

struct P2 { float x, y; };

double cross_d(P2 u, P2 v) { return (double)u.x * (double)v.y - (double)u.y * (double)v.x; }
P2 sub(P2 a, P2 b) { return P2{a.x - b.x, a.y - b.y}; }

double signed_area(const std::vector<P2> &p) {
    if (p.size() < 3) return 0.0;
    double total = 0.0;
    for (size_t i = 1; i + 1 < p.size(); ++i) total += cross_d(sub(p[i], p[0]), sub(p[i + 1], p[0]));
    return total;
}

std::vector<P2> merge_collinear(std::vector<P2> ring) {
    const double tolerance = 1e-5;
    bool removed = true;
    while (removed && ring.size() > 3) {
        removed = false;
        size_t index = 0;
        while (index < ring.size() && ring.size() > 3) {
            size_t count = ring.size();
            P2 incoming = sub(ring[index], ring[(index + count - 1) % count]);
            P2 outgoing = sub(ring[(index + 1) % count], ring[index]);
            double lengths = (double)std::sqrt(incoming.x * incoming.x + incoming.y * incoming.y) *
                             (double)std::sqrt(outgoing.x * outgoing.x + outgoing.y * outgoing.y);
            double dot = (double)incoming.x * outgoing.x + (double)incoming.y * outgoing.y;
            if (lengths > 0 && dot > 0 && std::fabs(cross_d(incoming, outgoing)) <= tolerance * lengths) {
                ring.erase(ring.begin() + (long)index);
                removed = true;
            } else {
                ++index;
            }
        }
    }
    return ring;
}

}  // namespace

extern "C" {

bp_status bp_static_polygon_edges(const float *points, uint32_t count, int64_t first_id, int64_t owner_id, uint32_t has_owner,
                                  bp_segment *out, uint32_t capacity, uint32_t *out_count) {
    if (!points || count == 0 || !out_count) return BP_ERR_INVALID_ARGUMENT;
    std::vector<P2> p(count);
    for (uint32_t i = 0; i < count; ++i) p[i] = P2{points[2 * i], points[2 * i + 1]};
    auto base = [&](int64_t id, P2 a, P2 b) {
        bp_segment s;
        std::memset(&s, 0, sizeof s);
        s.id = id;
        s.owner_id = has_owner ? owner_id : 0;
        s.a[0] = a.x; s.a[1] = a.y; s.b[0] = b.x; s.b[1] = b.y;
        s.flags = has_owner ? BP_SEGMENT_HAS_OWNER : 0u;
        s.category = 1u; s.mask = 0xffffffffu;
        return s;
    };
    double area = signed_area(p);
    if (count < 3 || area == 0.0) {
        *out_count = count;
        if (capacity < count || !out) return BP_ERR_CAPACITY_OVERFLOW;
        for (uint32_t i = 0; i < count; ++i) out[i] = base(first_id + i, p[i], p[(i + 1) % count]);
        return BP_OK;
    }
    std::vector<P2> ring = merge_collinear(p);
    double orientation = area > 0 ? 1.0 : -1.0;
    uint32_t n = (uint32_t)ring.size();
    *out_count = n;
    if (capacity < n || !out) return BP_ERR_CAPACITY_OVERFLOW;
    for (uint32_t i = 0; i < n; ++i) {
        P2 previous = ring[(i + n - 1) % n], a = ring[i], b = ring[(i + 1) % n], next = ring[(i + 2) % n];
        bp_segment s = base(first_id + i, a, b);
        bool convexStart = cross_d(sub(a, previous), sub(b, a)) * orientation > 0;
        bool convexEnd = cross_d(sub(b, a), sub(next, b)) * orientation > 0;
        s.flags |= BP_SEGMENT_POLYGON_EDGE | (convexStart ? BP_SEGMENT_CONVEX_START : 0u) | (convexEnd ? BP_SEGMENT_CONVEX_END : 0u);
        s.outward_sign = area > 0 ? -1.0f : 1.0f;
        s.previous_start[0] = previous.x; s.previous_start[1] = previous.y;
        out[i] = s;
    }
    return BP_OK;
}

// Polygon geometry of a rigid body: centroid (world offset of the local origin), vertices
// relative to the centroid in counter-clockwise order, triangles (a, b, c, 0), bounding radius.
bp_status bp_polygon_body_geometry(const float *local_points, uint32_t count, float *out_centroid, float *out_vertices,
                                   uint32_t *out_vertex_count, uint32_t *out_triangles, uint32_t *out_triangle_count,
                                   float *out_radius) {
    if (!local_points || count < 3 || !out_centroid || !out_vertices || !out_vertex_count || !out_triangles ||
        !out_triangle_count || !out_radius)
        return BP_ERR_INVALID_ARGUMENT;
    struct D2 { double x, y; };
    auto cross = [](D2 a, D2 b) { return a.x * b.y - a.y * b.x; };
    std::vector<D2> pts(count);
    D2 origin{local_points[0], local_points[1]};
    for (uint32_t i = 0; i < count; ++i) pts[i] = D2{(double)local_points[2 * i] - origin.x, (double)local_points[2 * i + 1] - origin.y};
    double area = 0.0;
    D2 weighted{0.0, 0.0};
    for (uint32_t i = 0; i < count; ++i) {
        D2 a = pts[i], b = pts[(i + 1) % count];
        double weight = cross(a, b);
        area += weight;
        weighted.x += (a.x + b.x) * weight; weighted.y += (a.y + b.y) * weight;
    }
    D2 center{weighted.x / (3 * area), weighted.y / (3 * area)};
    if (area < 0) std::reverse(pts.begin(), pts.end());
    bool changed = true;
    while (changed && pts.size() > 3) {
        changed = false;
        for (size_t i = 0; i < pts.size(); ++i) {
            D2 a = pts[(i + pts.size() - 1) % pts.size()], b = pts[i], c = pts[(i + 1) % pts.size()];
            if (cross(D2{b.x - a.x, b.y - a.y}, D2{c.x - b.x, c.y - b.y}) == 0) { pts.erase(pts.begin() + (long)i); changed = true; break; }
        }
    }
    std::vector<uint32_t> remaining(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) remaining[i] = (uint32_t)i;
    std::vector<uint32_t> pieces;
    auto sub = [](D2 a, D2 b) { return D2{a.x - b.x, a.y - b.y}; };
    while (remaining.size() > 3) {
        long ear = -1;
        for (size_t i = 0; i < remaining.size(); ++i) {
            uint32_t a = remaining[(i + remaining.size() - 1) % remaining.size()], b = remaining[i], c = remaining[(i + 1) % remaining.size()];
            if (!(cross(sub(pts[b], pts[a]), sub(pts[c], pts[a])) > 0)) continue;
            bool contains = false;
            for (uint32_t j : remaining) {
                if (j == a || j == b || j == c) continue;
                if (cross(sub(pts[b], pts[a]), sub(pts[j], pts[a])) >= 0 && cross(sub(pts[c], pts[b]), sub(pts[j], pts[b])) >= 0 &&
                    cross(sub(pts[a], pts[c]), sub(pts[j], pts[c])) >= 0) { contains = true; break; }
            }
            if (!contains) { pieces.insert(pieces.end(), {a, b, c, 0u}); ear = (long)i; break; }
        }
        if (ear < 0) return BP_ERR_INVALID_SNAPSHOT;
        remaining.erase(remaining.begin() + ear);
    }
    pieces.insert(pieces.end(), {remaining[0], remaining[1], remaining[2], 0u});
    out_centroid[0] = (float)(center.x + origin.x);
    out_centroid[1] = (float)(center.y + origin.y);
    float radius = 0.0f;
    for (size_t i = 0; i < pts.size(); ++i) {
        float vx = (float)(pts[i].x - center.x), vy = (float)(pts[i].y - center.y);
        out_vertices[2 * i] = vx; out_vertices[2 * i + 1] = vy;
        radius = std::fmax(radius, std::sqrt(vx * vx + vy * vy));
    }
    *out_vertex_count = (uint32_t)pts.size();
    std::memcpy(out_triangles, pieces.data(), pieces.size() * sizeof(uint32_t));
    *out_triangle_count = (uint32_t)(pieces.size() / 4);
    *out_radius = radius;
    if (!std::isfinite(out_centroid[0]) || !std::isfinite(out_centroid[1]) || !std::isfinite(radius) || !(radius > 0))
        return BP_ERR_INVALID_SNAPSHOT;
    return BP_OK;
}

}  // extern "C"
