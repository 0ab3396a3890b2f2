// bp_oracle_rigid.cpp -- CPU oracle of the rigid solver for worlds without bubbles.
//
// TEST INFRASTRUCTURE ONLY. Loaded by the tests as the checker of the HIP library's rigid
// solver (bp_rigid_batch_*); never by the product path.
//
// It restates, one world at a time and sequentially, the rigid-body kernels of the reference
// (paths relative to the reference repository):
//   Sources/BubblePhysicsNativeMetal/Shaders/NativePostSolveKernels.metal-source ("PostSolve")
//     nativeRigidSweepCrossing :29-62, nativeRigidPredict :64-112, nativeRigidGuardWalls :116-178,
//     nativeRigidFinish :180-206, nativeRigidOverlapAt :210-215, nativeRigidTimeOfImpact :220-238,
//     nativeRigidResolveCCD :243-257, nativeRigidApplyCCD :260-271
//   Sources/BubblePhysicsNativeMetal/Shaders/NativeContactKernels.metal-source ("Contact")
//     nativeTriangleContact :246-270, nativeRoundPolygonContact :276-339,
//     nativeRoundCoreCrossesPolygon :343-351, nativeRigidPairContact :356-378,
//     nativeRigidContacts :382-433, nativeRigidApplyBodies :435-449, nativeJointEnd :459-470,
//     nativeSolveJoints :474-548, nativeApplyJointBodies :551-568
//   Sources/BubblePhysicsNativeMetal/Shaders/NativeGeometryKernels.metal-source ("Geometry")
//     nativeFilterAccepts :130-132, nativeRigidImmovable :134, nativeRotate :136-138,
//     nativeRigidVertex :139-141, nativeClampLinearSpeed / AngularSpeed :151-158,
//     nativeInsertCandidate :230-251, nativeJointExcluded :254-258,
//     nativeEmitBodyCandidates (body-pair mode) :350-415, nativeClosestPoint :436-440,
//     nativeSegmentDistance :444-462, nativeRoundCoreOf :467-473, nativeRigidInside :477-490,
//     nativeRoundContact :546-580, nativeCenterCrossing :583-622
// in the dispatch order of NativeMetalSolver.encodeStepPrefix (NativeMetalSolver.swift:766-799):
// predict, body-pair candidates, CCD resolve and apply, then `iterations` rounds of contacts,
// apply, joints, joint apply and wall guard; nativeRigidFinish at the end gate (:917-921); the
// status precedence of NativeMetalSolver.swift:1031-1040 (nonFinite before guardFailure).
// Every kernel of the reference reads a state no other thread of the same dispatch writes (each
// body, row slot or joint owns its output), so running its threads one after another in index
// order gives the dispatch's result; a pass that reads what another thread writes is split in
// two here as the reference splits it into two dispatches.
//
// Before the step, the optional motion prelude restates the runtime's velocity update for role
// .dynamic (MotionControllers.swift:57-77, 103-106): v = v * exp(-damping dt), v += F / m * dt.
//
// Arithmetic: FP32 with no contraction (-ffp-contract=off), correctly rounded division and
// square root, and the deterministic sin, cos and exp of bp_detmath.h that the kernels share;
// expressions are evaluated in the reference's left-to-right order.
//
// Pinning (tests/test_oracle_rigid.py): the reference's rigid scenes (RigidCCDTests,
// RigidShapeContactTests, JointWorldTests, VelocityLimitTests, NativeMetalRigidPolygonTests) and
// closed forms (free flight, angular damping, speed clamps, the region rule's shift, SAT depth
// of overlapping boxes, round-core distances, the time of impact of two approaching boxes),
// with planted mistakes that these checks catch.
#include "../include/bubblephysics.h"
#include "../bubblephysics_amd/csrc/bp_detmath.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace bpr {

// Planted mistakes for the tests that show the pins catch them (0 = none):
//   1 SAT: the normal of the deepest axis taken with the wrong orientation;
//   2 CCD: the time of impact ignored (bodies keep their predicted poses);
//   3 joints: the distance correction applied with the wrong sign;
//   4 finish: the wrap shift without the region size;
//   5 predict: the angular damping factor not applied.
int g_plant = 0;

struct F2 {
    float x, y;
};
static inline F2 f2(float x, float y) { return F2{x, y}; }
static inline F2 operator+(F2 a, F2 b) { return f2(a.x + b.x, a.y + b.y); }
static inline F2 operator-(F2 a, F2 b) { return f2(a.x - b.x, a.y - b.y); }
static inline F2 operator-(F2 a) { return f2(-a.x, -a.y); }
static inline F2 operator*(F2 a, float s) { return f2(a.x * s, a.y * s); }
static inline F2 operator*(float s, F2 a) { return f2(s * a.x, s * a.y); }
static inline F2 operator/(F2 a, float s) { return f2(a.x / s, a.y / s); }
static inline float dot(F2 a, F2 b) { return a.x * b.x + a.y * b.y; }
static inline float cross(F2 a, F2 b) { return a.x * b.y - a.y * b.x; }
static inline float length(F2 a) { return std::sqrt(a.x * a.x + a.y * a.y); }
static inline float distance(F2 a, F2 b) { return length(a - b); }
static inline F2 normalize(F2 a) { return a / length(a); }
static inline float fmin_(float a, float b) { return std::fmin(a, b); }
static inline float fmax_(float a, float b) { return std::fmax(a, b); }
static inline float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
static inline F2 min2(F2 a, F2 b) { return f2(fmin_(a.x, b.x), fmin_(a.y, b.y)); }
static inline F2 max2(F2 a, F2 b) { return f2(fmax_(a.x, b.x), fmax_(a.y, b.y)); }
static inline float mixf(float x, float y, float t) { return x + (y - x) * t; }
static inline F2 mix2(F2 x, F2 y, float t) { return f2(mixf(x.x, y.x, t), mixf(x.y, y.y, t)); }
static inline bool finite4(const float *v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]) && std::isfinite(v[3]);
}
static inline float bits_to_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
static inline uint32_t float_to_bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }

// Geometry:136-138
static inline F2 rotate(F2 p, float angle) {
    float c = bpm::cosf_det(angle), s = bpm::sinf_det(angle);
    return f2(p.x * c - p.y * s, p.x * s + p.y * c);
}

struct Pose3 {
    float x, y, a;
};

struct Cand {       // NativeCandidate (Geometry:27) of a body row
    int64_t first, second;
    uint32_t kind, low, high;
    float impulse;  // metadata.w: sum of the iterations' normal corrections
};
struct Corr {       // NativeRigidCorrection (Geometry:51)
    float first[4], second[4], vfirst[4], vsecond[4];
};
struct Wall {       // NativeRigidWall (Geometry:54)
    int64_t body, owner;
    float geometry[4];
    float extra[4];
};

struct World {
    bp_rigid_config cfg;
    std::vector<bp_rigid_body> b;
    const F2 *verts;
    const uint32_t *tris;  // uint4 records
    const bp_segment *segs;
    uint32_t ns;
    std::vector<bp_joint> joints;
    const uint32_t *refs;    // (joint, side) pairs
    const uint32_t *ranges;  // uint4 per body
    const uint64_t *partners;
    uint32_t k;
    std::vector<Cand> rows;
    std::vector<uint32_t> rowCount;
    std::vector<Corr> corr;
    std::vector<Wall> walls;
    std::vector<Corr> jcorr;
    uint32_t safety = 0;
    uint32_t overflow = 0, maxPer = 0;

    uint32_t n() const { return (uint32_t)b.size(); }
    F2 vert(uint32_t i) const { return verts[i]; }
    F2 pose_xy(const bp_rigid_body &x) const { return f2(x.pose[0], x.pose[1]); }
    // Geometry:139-141
    F2 vertex(const bp_rigid_body &x, uint32_t i) const {
        return pose_xy(x) + rotate(vert(x.ranges[0] + i), x.pose[2]);
    }
    bool excluded(uint32_t object, uint64_t partner) const {  // Geometry:254-258
        if (!ranges) return false;
        const uint32_t *r = ranges + 4 * object;
        for (uint32_t i = 0; i < r[3]; ++i)
            if (partners[r[2] + i] == partner) return true;
        return false;
    }
};

static inline bool filter_accepts(const uint32_t *a, const uint32_t *b) {  // Geometry:130-132
    return (a[0] & b[1]) != 0 && (b[0] & a[1]) != 0;
}
static inline bool immovable(const bp_rigid_body &x) { return x.flags[0] != 0 || x.filter[2] != 0; }  // Geometry:134
static inline bool is_round(const bp_rigid_body &x) { return x.shape[0] > 0.5f; }                   // Geometry:143

// Geometry:151-158
static inline F2 clamp_linear(F2 v, float limit) {
    float speed = length(v);
    return limit > 0 && speed > limit ? v * (limit / speed) : v;
}
static inline float clamp_angular(float w, float limit) { return limit > 0 ? clampf(w, -limit, limit) : w; }

// Geometry:436-440
static inline F2 closest_point(F2 p, F2 a, F2 b) {
    F2 edge = b - a;
    float square = dot(edge, edge);
    return a + edge * (square > FLT_MIN ? clampf(dot(p - a, edge) / square, 0.0f, 1.0f) : 0.0f);
}

// Geometry:444-462
static float segment_distance(F2 a, F2 b, F2 c, F2 d, F2 &pa, F2 &pb) {
    F2 r = b - a, s = d - c;
    float denominator = cross(r, s);
    if (std::fabs(denominator) > 1e-12f) {
        F2 qp = c - a;
        float t = cross(qp, s) / denominator, u = cross(qp, r) / denominator;
        if (t >= 0 && t <= 1 && u >= 0 && u <= 1) {
            pa = a + t * r;
            pb = pa;
            return 0;
        }
    }
    float best = INFINITY;
    F2 cp = closest_point(a, c, d);
    float value = distance(a, cp);
    if (value < best) { best = value; pa = a; pb = cp; }
    cp = closest_point(b, c, d);
    value = distance(b, cp);
    if (value < best) { best = value; pa = b; pb = cp; }
    cp = closest_point(c, a, b);
    value = distance(c, cp);
    if (value < best) { best = value; pa = cp; pb = c; }
    cp = closest_point(d, a, b);
    value = distance(d, cp);
    if (value < best) { best = value; pa = cp; pb = d; }
    return best;
}

struct Core {
    F2 a, b;
    float radius;
};
// Geometry:467-473
static Core core_of(const World &w, const bp_rigid_body &x) {
    Core c;
    c.a = w.vertex(x, 0);
    c.b = x.ranges[1] > 1 ? w.vertex(x, 1) : c.a;
    c.radius = x.shape[1];
    return c;
}

// Geometry:477-490
static bool rigid_inside(const World &w, F2 p, const bp_rigid_body &x) {
    if (is_round(x)) {
        Core c = core_of(w, x);
        return distance(p, closest_point(p, c.a, c.b)) <= c.radius;
    }
    F2 local = rotate(p - w.pose_xy(x), -x.pose[2]);
    for (uint32_t i = 0; i < x.ranges[3]; ++i) {
        const uint32_t *t = w.tris + 4 * (x.ranges[2] + i);
        F2 a = w.vert(x.ranges[0] + t[0]), b = w.vert(x.ranges[0] + t[1]), c = w.vert(x.ranges[0] + t[2]);
        if (cross(b - a, local - a) >= 0 && cross(c - b, local - b) >= 0 && cross(a - c, local - c) >= 0) return true;
    }
    return false;
}

// Geometry:546-580
static bool round_contact(const Core &A, const Core &B, F2 fallback, F2 &normal, float &depth, F2 &point) {
    F2 pa, pb;
    float separation = segment_distance(A.a, A.b, B.a, B.b, pa, pb);
    if (separation > 1e-4f) {
        normal = (pb - pa) / separation;
        depth = A.radius + B.radius - separation;
        point = 0.5f * ((pa + normal * A.radius) + (pb - normal * B.radius));
        return true;
    }
    F2 axes[2];
    uint32_t axisCount = 0;
    F2 ea = A.b - A.a, eb = B.b - B.a;
    float la = length(ea), lb = length(eb);
    if (la > 1e-6f) axes[axisCount++] = f2(-ea.y, ea.x) / la;
    if (lb > 1e-6f) axes[axisCount++] = f2(-eb.y, eb.x) / lb;
    F2 axis = fallback;
    float overlap = 0, orientation = 1;
    float best = INFINITY;
    for (uint32_t i = 0; i < axisCount; ++i) {
        F2 candidate = axes[i];
        float amin = fmin_(dot(A.a, candidate), dot(A.b, candidate)), amax = fmax_(dot(A.a, candidate), dot(A.b, candidate));
        float bmin = fmin_(dot(B.a, candidate), dot(B.b, candidate)), bmax = fmax_(dot(B.a, candidate), dot(B.b, candidate));
        float plus = amax - bmin, minus = bmax - amin;
        float value = fmin_(plus, minus);
        if (value < best) {
            best = value;
            axis = candidate;
            overlap = value;
            if (std::fabs(plus - minus) < 1e-6f) orientation = dot(candidate, fallback) >= 0 ? 1.0f : -1.0f;
            else orientation = plus < minus ? 1.0f : -1.0f;
        }
    }
    normal = orientation * axis;
    depth = overlap + A.radius + B.radius;
    point = 0.5f * (pa + pb);
    return true;
}

// Segment moving linearly from (pa, pb) to (ca, cb) over the step.
struct MovingSeg {
    F2 pa, pb, ca, cb;
};
// Geometry:583-622
static bool center_crossing(F2 start, F2 end, const MovingSeg &s) {
    F2 e0 = s.pb - s.pa;
    F2 de = (s.cb - s.ca) - e0;
    F2 r0 = start - s.pa;
    F2 dr = (end - start) - (s.ca - s.pa);
    float a = cross(de, dr), b = cross(e0, dr) + cross(de, r0);
    float c = cross(e0, r0);
    if (a == 0 && b == 0 && c == 0) {
        float q0 = dot(r0, r0 - e0);
        if (q0 <= 0) return false;
        float qa = dot(dr, dr - de);
        float qb = dot(dr, r0 - e0) + dot(r0, dr - de);
        float minimum = fmin_(q0, qa + qb + q0);
        if (qa > 0) {
            float t = clampf(-qb / (2 * qa), 0.0f, 1.0f);
            minimum = fmin_(minimum, (qa * t + qb) * t + q0);
        }
        return minimum <= 0;
    }
    float roots[2] = {-1, -1};
    if (std::fabs(a) <= FLT_EPSILON * fmax_(std::fabs(b), std::fabs(c))) {
        if (std::fabs(b) > FLT_MIN) roots[0] = -c / b;
    } else {
        float d = b * b - 4 * a * c;
        if (d > 0) {
            float q = -0.5f * (b + std::copysign(std::sqrt(d), b));
            roots[0] = q / a;
            roots[1] = std::fabs(q) > FLT_MIN ? c / q : -1;
        }
    }
    for (int k = 0; k < 2; ++k) {
        float t = roots[k];
        if (!(t > 0 && t < 1)) continue;
        F2 edge = e0 + de * t, relative = r0 + dr * t;
        float square = dot(edge, edge), along = dot(relative, edge);
        if (square > FLT_MIN && along >= 0 && along <= square) return true;
    }
    return false;
}

// Contact: nativeTriangleContact :246-270
static bool triangle_contact(const F2 *a, const F2 *b, F2 &normal, float &depth, F2 &point) {
    depth = INFINITY;
    for (int shape = 0; shape < 2; ++shape)
        for (int edge = 0; edge < 3; ++edge) {
            F2 delta = shape == 0 ? a[(edge + 1) % 3] - a[edge] : b[(edge + 1) % 3] - b[edge];
            float magnitude = length(delta);
            if (magnitude <= FLT_MIN) return false;
            F2 axis = f2(-delta.y, delta.x) / magnitude;
            float amin = INFINITY, amax = -INFINITY, bmin = INFINITY, bmax = -INFINITY;
            for (int k = 0; k < 3; ++k) {
                float pa = dot(a[k], axis), pb = dot(b[k], axis);
                amin = fmin_(amin, pa);
                amax = fmax_(amax, pa);
                bmin = fmin_(bmin, pb);
                bmax = fmax_(bmax, pb);
            }
            if (amax < bmin || bmax < amin) return false;
            float plus = amax - bmin, minus = bmax - amin;
            float overlap = fmin_(plus, minus);
            if (overlap < depth) {
                depth = overlap;
                bool forward = plus < minus;
                if (g_plant == 1) forward = !forward;
                normal = forward ? axis : -axis;
            }
        }
    float maxA = -INFINITY, minB = INFINITY;
    F2 pa = f2(0, 0), pb = f2(0, 0);
    uint32_t na = 0, nb = 0;
    for (int k = 0; k < 3; ++k) {
        maxA = fmax_(maxA, dot(a[k], normal));
        minB = fmin_(minB, dot(b[k], normal));
    }
    for (int k = 0; k < 3; ++k) {
        if (dot(a[k], normal) >= maxA - 1e-5f) { pa = pa + a[k]; ++na; }
        if (dot(b[k], normal) <= minB + 1e-5f) { pb = pb + b[k]; ++nb; }
    }
    point = (pa / float(na) + pb / float(nb)) * 0.5f;
    return true;
}

// Contact: nativeRoundPolygonContact :276-339 (normal from R to P)
static bool round_polygon_contact(const World &w, const Core &R, const bp_rigid_body &P, F2 &normal, float &depth,
                                  F2 &point) {
    uint32_t count = P.ranges[1];
    float insideDepth = -1;
    F2 insideOutward = f2(0, 0), insidePoint = f2(0, 0);
    uint32_t endpoints = (R.b.x != R.a.x || R.b.y != R.a.y) ? 2u : 1u, insideEndpoints = 0;
    for (uint32_t k = 0; k < endpoints; ++k) {
        F2 endpoint = k == 0 ? R.a : R.b;
        if (!rigid_inside(w, endpoint, P)) continue;
        ++insideEndpoints;
        float nearest = INFINITY;
        F2 nearestPoint = f2(0, 0), nearestOutward = f2(0, 0);
        for (uint32_t e = 0; e < count; ++e) {
            F2 p = w.vertex(P, e), q = w.vertex(P, (e + 1) % count);
            F2 closest = closest_point(endpoint, p, q);
            float d = distance(endpoint, closest);
            if (d < nearest) {
                nearest = d;
                nearestPoint = closest;
                F2 edge = q - p;
                float size = length(edge);
                nearestOutward = size > FLT_MIN ? f2(edge.y, -edge.x) / size : f2(0, 1);
            }
        }
        if (nearest > insideDepth) {
            insideDepth = nearest;
            insidePoint = nearestPoint;
            insideOutward = nearestOutward;
        }
    }
    if (insideEndpoints == endpoints) {
        normal = -insideOutward;
        depth = R.radius + insideDepth;
        point = insidePoint;
        return true;
    }
    float best = INFINITY;
    F2 pc = f2(0, 0), pe = f2(0, 0);
    for (uint32_t e = 0; e < count; ++e) {
        F2 p = w.vertex(P, e), q = w.vertex(P, (e + 1) % count);
        F2 ca, ce;
        float d = segment_distance(R.a, R.b, p, q, ca, ce);
        if (d < best) {
            best = d;
            pc = ca;
            pe = ce;
        }
    }
    if (best > 1e-4f) {
        normal = (pe - pc) / best;
        depth = R.radius - best;
        point = pe;
        return true;
    }
    float minimum = INFINITY;
    F2 minimumOutward = f2(0, 1), minimumPoint = pe;
    for (uint32_t e = 0; e < count; ++e) {
        F2 p = w.vertex(P, e), q = w.vertex(P, (e + 1) % count);
        F2 edge = q - p;
        float size = length(edge);
        if (size <= FLT_MIN) continue;
        F2 outward = f2(edge.y, -edge.x) / size;
        float da = dot(R.a - p, outward), db = dot(R.b - p, outward);
        float d = R.radius - fmin_(da, db);
        if (d < minimum) {
            minimum = d;
            minimumOutward = outward;
            minimumPoint = (da <= db ? R.a : R.b) - outward * fmin_(da, db);
        }
    }
    normal = -minimumOutward;
    depth = std::isfinite(minimum) ? minimum : R.radius;
    point = minimumPoint;
    return true;
}

// Contact: nativeRoundCoreCrossesPolygon :343-351
static bool round_core_crosses_polygon(const World &w, const Core &R, const bp_rigid_body &P) {
    if (rigid_inside(w, R.a, P) || rigid_inside(w, R.b, P)) return true;
    for (uint32_t e = 0; e < P.ranges[1]; ++e) {
        F2 p = w.vertex(P, e), q = w.vertex(P, (e + 1) % P.ranges[1]);
        F2 ca, ce;
        if (segment_distance(R.a, R.b, p, q, ca, ce) <= 1e-4f) return true;
    }
    return false;
}

// Contact: nativeRigidPairContact :356-378 (normal from a to b)
static bool pair_contact(const World &w, const bp_rigid_body &a, const bp_rigid_body &b, F2 &n, float &depth, F2 &point) {
    bool aRound = is_round(a), bRound = is_round(b);
    if (aRound && bRound) return round_contact(core_of(w, a), core_of(w, b), f2(1, 0), n, depth, point);
    if (aRound != bRound) {
        const bp_rigid_body &round = aRound ? a : b;
        const bp_rigid_body &polygon = aRound ? b : a;
        F2 normal;
        if (!round_polygon_contact(w, core_of(w, round), polygon, normal, depth, point)) return false;
        n = aRound ? normal : -normal;
        return true;
    }
    bool found = false;
    for (uint32_t i = 0; i < a.ranges[3]; ++i)
        for (uint32_t j = 0; j < b.ranges[3]; ++j) {
            const uint32_t *ta = w.tris + 4 * (a.ranges[2] + i), *tb = w.tris + 4 * (b.ranges[2] + j);
            F2 av[3] = {w.vertex(a, ta[0]), w.vertex(a, ta[1]), w.vertex(a, ta[2])};
            F2 bv[3] = {w.vertex(b, tb[0]), w.vertex(b, tb[1]), w.vertex(b, tb[2])};
            F2 normal, p;
            float d;
            if (triangle_contact(av, bv, normal, d, p) && (!found || d > depth)) {
                found = true;
                depth = d;
                n = normal;
                point = p;
            }
        }
    return found;
}

// PostSolve:29-62
static bool sweep_crossing(const World &w, const bp_rigid_body &b, Pose3 p, Pose3 q, uint32_t entity, F2 &contactPoint,
                           F2 &contactNormal, int64_t &contactOwner) {
    for (uint32_t j = 0; j < w.ns; ++j) {
        const bp_segment &wall = w.segs[j];
        bool owned = (wall.flags & BP_SEGMENT_HAS_OWNER) != 0;
        if ((uint64_t)wall.owner_id == b.id && owned) continue;
        if (owned && w.excluded(entity, (uint64_t)wall.owner_id)) continue;
        uint32_t wf[2] = {wall.category, wall.mask};
        if (!filter_accepts(b.filter, wf)) continue;
        F2 wa = f2(wall.a[0], wall.a[1]), wb = f2(wall.b[0], wall.b[1]);
        MovingSeg still{wa, wb, wa, wb};
        for (uint32_t k = 0; k < b.ranges[1]; ++k) {
            F2 v = w.vert(b.ranges[0] + k);
            F2 startVertex = f2(p.x, p.y) + rotate(v, p.a), endVertex = f2(q.x, q.y) + rotate(v, q.a);
            bool hit = center_crossing(startVertex, endVertex, still);
            if (!hit) {
                F2 next = w.vert(b.ranges[0] + (k + 1) % b.ranges[1]);
                MovingSeg edge{f2(p.x, p.y) + rotate(v, p.a), f2(p.x, p.y) + rotate(next, p.a),
                               f2(q.x, q.y) + rotate(v, q.a), f2(q.x, q.y) + rotate(next, q.a)};
                hit = center_crossing(wa, wa, edge) || center_crossing(wb, wb, edge);
            }
            if (hit) {
                contactPoint = closest_point(endVertex, wa, wb);
                F2 e = wb - wa;
                contactNormal = normalize(f2(e.y, -e.x));
                if (dot(startVertex - wa, contactNormal) < 0) contactNormal = -contactNormal;
                contactOwner = wall.owner_id;
                return true;
            }
        }
    }
    return false;
}

static inline Pose3 mix3(Pose3 s, Pose3 e, float t) { return Pose3{mixf(s.x, e.x, t), mixf(s.y, e.y, t), mixf(s.a, e.a, t)}; }

// Runtime velocity update of a dynamic body before the step (MotionControllers.swift:57-77,
// 103-106; no fields, no motor).
static void motion_prelude(World &w) {
    if (!w.cfg.motion_prelude) return;
    float dt = w.cfg.time_step;
    for (auto &x : w.b) {
        if (x.flags[0] || x.filter[2]) continue;
        float factor = bpm::expf_det(-x.motion[3] * dt);
        x.motion[0] = x.motion[0] * factor;
        x.motion[1] = x.motion[1] * factor;
        x.motion[0] = x.motion[0] + w.cfg.global_force[0] / x.mass * dt;
        x.motion[1] = x.motion[1] + w.cfg.global_force[1] / x.mass * dt;
    }
}

// PostSolve:64-112
static void predict(World &w) {
    float dt = w.cfg.time_step;
    for (uint32_t index = 0; index < w.n(); ++index) {
        bp_rigid_body b = w.b[index];
        uint32_t entity = index;
        b.previous[0] = b.pose[0];
        b.previous[1] = b.pose[1];
        b.pose[3] = b.pose[2];
        b.flags[2] = 0;
        if (b.flags[0]) {
            b.motion[0] = b.motion[1] = b.motion[2] = 0;
            w.b[index] = b;
            continue;
        }
        if (g_plant != 5) b.motion[2] *= bpm::expf_det(-b.motion[3] * dt);
        F2 v = clamp_linear(f2(b.motion[0], b.motion[1]), b.limits[0]);
        b.motion[0] = v.x;
        b.motion[1] = v.y;
        b.motion[2] = clamp_angular(b.motion[2], b.limits[1]);
        Pose3 start{b.pose[0], b.pose[1], b.pose[2]};
        b.pose[0] += b.motion[0] * dt;
        b.pose[1] += b.motion[1] * dt;
        b.pose[2] += b.motion[2] * dt;
        if (!finite4(b.pose) || !finite4(b.motion)) {
            w.safety |= 1;
            w.b[index] = b;
            continue;
        }
        if (b.filter[2]) {
            w.b[index] = b;
            continue;
        }
        float turns = std::ceil(std::fabs(b.pose[2] - start.a) / 0.05f);
        if (turns > 128.0f) {
            w.safety |= 2 | 16;
            w.b[index] = b;
            continue;
        }
        uint32_t steps = std::max(1u, (uint32_t)turns);
        bool crossed = false;
        F2 contactPoint = f2(0, 0), contactNormal = f2(0, 0);
        int64_t contactOwner = 0;
        Pose3 end{b.pose[0], b.pose[1], b.pose[2]};
        Pose3 stopPose = start;
        for (uint32_t sample = 0; sample < steps && !crossed; ++sample) {
            Pose3 p = mix3(start, end, float(sample) / float(steps));
            Pose3 q = mix3(start, end, float(sample + 1) / float(steps));
            crossed = sweep_crossing(w, b, p, q, entity, contactPoint, contactNormal, contactOwner);
            if (!crossed) continue;
            Pose3 last = p, hit = q;
            for (int bisection = 0; bisection < 12; ++bisection) {
                Pose3 middle{0.5f * (last.x + hit.x), 0.5f * (last.y + hit.y), 0.5f * (last.a + hit.a)};
                F2 point, normal;
                int64_t owner;
                if (sweep_crossing(w, b, p, middle, entity, point, normal, owner)) hit = middle;
                else last = middle;
            }
            stopPose = last;
        }
        if (crossed) {
            float closing = fmax_(0.0f, -dot(f2(b.motion[0], b.motion[1]), contactNormal));
            float stopped = b.physical[0] > 0 ? closing / b.physical[0] * dt : 0.0f;
            b.pose[0] = stopPose.x;
            b.pose[1] = stopPose.y;
            b.pose[2] = stopPose.a;
            b.motion[0] = b.motion[1] = b.motion[2] = 0;
            for (uint32_t j = 0; j < w.n(); ++j)
                if (j != index && w.b[j].flags[0] && w.b[j].id == (uint64_t)contactOwner) {
                    Wall &r = w.walls[index];
                    r.body = (int64_t)b.id;
                    r.owner = contactOwner;
                    r.geometry[0] = contactNormal.x;
                    r.geometry[1] = contactNormal.y;
                    r.geometry[2] = contactPoint.x;
                    r.geometry[3] = contactPoint.y;
                    r.extra[0] = stopped;
                    r.extra[1] = r.extra[2] = r.extra[3] = 0;
                    break;
                }
        }
        w.b[index] = b;
    }
}

// Geometry:230-251 (key: kind, then participants as signed 64-bit)
static inline bool key_less(const Cand &a, const Cand &b) {
    if (a.kind != b.kind) return a.kind < b.kind;
    if (a.first != b.first) return a.first < b.first;
    return a.second < b.second;
}
static void insert_candidate(Cand *row, uint32_t &count, uint32_t &dropped, uint32_t limit, const Cand &value) {
    uint32_t location;
    if (count < limit) location = count++;
    else {
        ++dropped;
        if (!key_less(value, row[limit - 1])) return;
        location = limit - 1;
    }
    while (location > 0 && key_less(value, row[location - 1])) {
        row[location] = row[location - 1];
        --location;
    }
    row[location] = value;
}

// Geometry:350-415, body-pair mode (counts.w = 1)
static void emit_rows(World &w) {
    uint32_t limit = w.k;
    for (uint32_t index = 0; index < w.n(); ++index) {
        Cand *row = w.rows.data() + (size_t)index * limit;
        uint32_t count = 0, dropped = 0;
        for (uint32_t other = 0; other < w.n(); ++other) {
            if (other == index) continue;
            uint32_t low = std::min(index, other), high = std::max(index, other);
            const bp_rigid_body &a = w.b[low], &c = w.b[high];
            if (immovable(a) && immovable(c)) continue;
            if (!filter_accepts(a.filter, c.filter)) continue;
            if (w.excluded(index, w.b[other].id)) continue;
            F2 start = f2(a.previous[0], a.previous[1]) - f2(c.previous[0], c.previous[1]);
            F2 end = w.pose_xy(a) - w.pose_xy(c), travel = end - start;
            float travelSquared = dot(travel, travel);
            float t = travelSquared > FLT_MIN ? clampf(-dot(start, travel) / travelSquared, 0.0f, 1.0f) : 0.0f;
            float radius = 1.25f * (a.physical[2] + c.physical[2]);
            F2 closest = start + t * travel;
            if (!(dot(closest, closest) <= radius * radius)) continue;
            Cand cand{(int64_t)a.id, (int64_t)c.id, 3u, low, high, 0.0f};
            insert_candidate(row, count, dropped, limit, cand);
        }
        w.rowCount[index] = count;
        w.overflow += dropped;
        w.maxPer = std::max(w.maxPer, count + dropped);
    }
}

// PostSolve:210-215
static bool overlap_at(const World &w, bp_rigid_body a, bp_rigid_body b, float t) {
    F2 ap = mix2(f2(a.previous[0], a.previous[1]), w.pose_xy(a), t);
    a.pose[0] = ap.x;
    a.pose[1] = ap.y;
    a.pose[2] = mixf(a.pose[3], a.pose[2], t);
    F2 bp = mix2(f2(b.previous[0], b.previous[1]), w.pose_xy(b), t);
    b.pose[0] = bp.x;
    b.pose[1] = bp.y;
    b.pose[2] = mixf(b.pose[3], b.pose[2], t);
    F2 normal, point;
    float depth;
    return pair_contact(w, a, b, normal, depth, point) && depth > 1e-3f;
}

// PostSolve:220-238
static float time_of_impact(const World &w, const bp_rigid_body &a, const bp_rigid_body &b) {
    float travel = length((w.pose_xy(a) - f2(a.previous[0], a.previous[1])) - (w.pose_xy(b) - f2(b.previous[0], b.previous[1]))) +
                   std::fabs(a.pose[2] - a.pose[3]) * a.physical[2] + std::fabs(b.pose[2] - b.pose[3]) * b.physical[2];
    if (!(travel > 0.5f)) return 2;
    if (overlap_at(w, a, b, 0)) return 2;
    // min(64, max(2, uint(ceil(travel / 0.5)))) with the float-to-uint conversion saturating.
    float slices = std::ceil(travel / 0.5f);
    uint32_t steps = slices >= 64.0f ? 64u : std::max(2u, (uint32_t)slices);
    float last = 0;
    for (uint32_t step = 1; step <= steps; ++step) {
        float t = float(step) / float(steps);
        if (!overlap_at(w, a, b, t)) {
            last = t;
            continue;
        }
        float low = last, high = t;
        for (int bisection = 0; bisection < 12; ++bisection) {
            float middle = 0.5f * (low + high);
            if (overlap_at(w, a, b, middle)) high = middle;
            else low = middle;
        }
        return low;
    }
    return 2;
}

// PostSolve:243-271
static void ccd(World &w) {
    std::vector<float> marks(w.n());
    for (uint32_t index = 0; index < w.n(); ++index) {
        const bp_rigid_body &body = w.b[index];
        float earliest = 2;
        if (!immovable(body))
            for (uint32_t slot = 0; slot < w.rowCount[index]; ++slot) {
                const Cand &c = w.rows[(size_t)index * w.k + slot];
                if (c.kind != 3) continue;
                earliest = fmin_(earliest, time_of_impact(w, w.b[c.low], w.b[c.high]));
            }
        marks[index] = earliest <= 1 ? 1 + earliest : 0;
    }
    for (uint32_t index = 0; index < w.n(); ++index) {
        bp_rigid_body &body = w.b[index];
        body.limits[2] = marks[index];
        if (body.limits[2] >= 1 && g_plant != 2) {
            float t = body.limits[2] - 1;
            F2 p = mix2(f2(body.previous[0], body.previous[1]), w.pose_xy(body), t);
            body.pose[0] = p.x;
            body.pose[1] = p.y;
            body.pose[2] = mixf(body.pose[3], body.pose[2], t);
        }
        body.limits[2] = 0;
    }
}

// Contact:382-433, one row slot
static void contact_slot(World &w, uint32_t owner, uint32_t slot) {
    size_t index = (size_t)owner * w.k + slot;
    Cand &candidate = w.rows[index];
    const Corr previous = w.corr[index];
    Corr result{};
    result.first[3] = previous.first[3];
    result.second[3] = previous.second[3];
    result.vfirst[3] = previous.vfirst[3];
    result.vsecond[3] = previous.vsecond[3];
    if (candidate.kind != 3) { w.corr[index] = result; return; }
    uint32_t ai = candidate.low, bi = candidate.high;
    const bp_rigid_body &a = w.b[ai], &b = w.b[bi];
    if (ai >= bi || (immovable(a) && immovable(b))) { w.corr[index] = result; return; }
    F2 end = w.pose_xy(a) - w.pose_xy(b);
    float radius = a.physical[2] + b.physical[2], radiusSquared = radius * radius;
    if (dot(end, end) > radiusSquared) { w.corr[index] = result; return; }
    F2 n, point;
    float depth;
    if (!pair_contact(w, a, b, n, depth, point)) { w.corr[index] = result; return; }
    float wa = a.physical[0], wb = b.physical[0], ia = a.physical[1], ib = b.physical[1];
    F2 ra = point - w.pose_xy(a), rb = point - w.pose_xy(b);
    F2 va = f2(a.motion[0], a.motion[1]), vb = f2(b.motion[0], b.motion[1]);
    float spinA = a.motion[2], spinB = b.motion[2];
    float frictionA = a.physical[3], frictionB = b.physical[3];
    float restitutionA = bits_to_float(a.flags[3]), restitutionB = bits_to_float(b.flags[3]);
    float ca = cross(ra, n), cb = cross(rb, n), denom = wa + wb + ia * ca * ca + ib * cb * cb;
    if (depth < -0.001f || denom <= FLT_MIN) { w.corr[index] = result; return; }
    float lambda = fmax_(0.0f, depth) / denom * 0.5f;
    candidate.impulse = candidate.impulse + lambda;
    F2 m = -n * wa * lambda;
    result.first[0] = m.x; result.first[1] = m.y; result.first[2] = -ca * ia * lambda; result.first[3] = n.x;
    m = n * wb * lambda;
    result.second[0] = m.x; result.second[1] = m.y; result.second[2] = cb * ib * lambda; result.second[3] = n.y;
    float closing = dot(vb - va, n) + spinB * cb - spinA * ca;
    float friction = std::sqrt(fmax_(0.0f, frictionA * frictionB));
    float restitution = fmax_(restitutionA, restitutionB);
    float impulse = fmax_(0.0f, -closing) * (1.0f + restitution) / denom * 0.5f;
    F2 tangent = f2(-n.y, n.x);
    float tangentSpeed = dot(vb - va, tangent) + spinB * cross(rb, tangent) - spinA * cross(ra, tangent);
    float tangentImpulse = -tangentSpeed * friction / denom * 0.5f;
    m = -n * wa * impulse - tangent * wa * tangentImpulse;
    result.vfirst[0] = m.x; result.vfirst[1] = m.y;
    result.vfirst[2] = -ca * ia * impulse - cross(ra, tangent) * ia * tangentImpulse;
    result.vfirst[3] = point.x;
    m = n * wb * impulse + tangent * wb * tangentImpulse;
    result.vsecond[0] = m.x; result.vsecond[1] = m.y;
    result.vsecond[2] = cb * ib * impulse + cross(rb, tangent) * ib * tangentImpulse;
    result.vsecond[3] = point.y;
    w.corr[index] = result;
}

// Contact:435-449
static void apply_bodies(World &w) {
    for (uint32_t index = 0; index < w.n(); ++index) {
        float delta[3] = {0, 0, 0}, velocity[3] = {0, 0, 0};
        for (uint32_t slot = 0; slot < w.rowCount[index]; ++slot) {
            size_t at = (size_t)index * w.k + slot;
            const Cand &c = w.rows[at];
            const Corr &r = w.corr[at];
            bool first = c.kind == 3 && c.low == index;
            for (int d = 0; d < 3; ++d) {
                delta[d] += first ? r.first[d] : r.second[d];
                velocity[d] += first ? r.vfirst[d] : r.vsecond[d];
            }
        }
        bp_rigid_body &x = w.b[index];
        for (int d = 0; d < 3; ++d) {
            x.pose[d] += delta[d];
            x.motion[d] += velocity[d];
        }
    }
}

struct JointEnd {
    F2 point, arm, velocity;
    float inverseMass, inverseInertia, angle, omega;
};
// Contact:459-470
static JointEnd joint_end(const World &w, uint32_t entity, F2 anchor) {
    JointEnd e;
    e.point = anchor;
    e.arm = f2(0, 0);
    e.velocity = f2(0, 0);
    e.inverseMass = e.inverseInertia = e.angle = e.omega = 0;
    if (entity == BP_JOINT_WORLD_END) return e;
    const bp_rigid_body &b = w.b[entity];
    e.arm = rotate(anchor - f2(b.previous[2], b.previous[3]), b.pose[2]);
    e.point = w.pose_xy(b) + e.arm;
    e.velocity = f2(b.motion[0], b.motion[1]);
    e.omega = b.motion[2];
    e.angle = b.pose[2];
    e.inverseMass = b.physical[0];
    e.inverseInertia = b.physical[1];
    return e;
}
static inline F2 perpendicular(F2 v) { return f2(-v.y, v.x); }

// Contact:474-548
static void solve_joints(World &w) {
    float dt = fmax_(w.cfg.time_step, 1e-6f);
    uint32_t iterations = std::max(w.cfg.iterations, 1u);
    for (uint32_t index = 0; index < w.joints.size(); ++index) {
        bp_joint &joint = w.joints[index];
        JointEnd a = joint_end(w, joint.ends[0], f2(joint.anchors[0], joint.anchors[1]));
        JointEnd b = joint_end(w, joint.ends[1], f2(joint.anchors[2], joint.anchors[3]));
        F2 moveA = f2(0, 0), moveB = f2(0, 0), speedA = f2(0, 0), speedB = f2(0, 0);
        float turnA = 0, turnB = 0, spinA = 0, spinB = 0;
        F2 pointA = perpendicular(a.arm), pointB = perpendicular(b.arm);
        uint32_t kind = joint.ends[2];
        if (kind == 0) {
            F2 delta = b.point - a.point;
            float length_ = length(delta);
            F2 n = length_ > 1e-6f ? delta / length_ : f2(1, 0);
            float ca = cross(a.arm, n), cb = cross(b.arm, n);
            float wsum = a.inverseMass + b.inverseMass + a.inverseInertia * ca * ca + b.inverseInertia * cb * cb;
            if (wsum > FLT_MIN) {
                float alpha = joint.parameters[1] / (dt * dt);
                float lambda = (-(length_ - joint.parameters[0]) - alpha * joint.state[0]) / (wsum + alpha);
                if (g_plant == 3) lambda = -lambda;
                joint.state[0] += lambda;
                moveA = -n * a.inverseMass * lambda;
                turnA = -ca * a.inverseInertia * lambda;
                moveB = n * b.inverseMass * lambda;
                turnB = cb * b.inverseInertia * lambda;
                if (joint.parameters[2] > 0) {
                    float relative = dot((b.velocity + b.omega * pointB) - (a.velocity + a.omega * pointA), n);
                    float impulse = -relative * (1.0f - bpm::expf_det(-joint.parameters[2] * dt / float(iterations))) / wsum;
                    speedA = speedA - n * a.inverseMass * impulse;
                    spinA = spinA - ca * a.inverseInertia * impulse;
                    speedB = speedB + n * b.inverseMass * impulse;
                    spinB = spinB + cb * b.inverseInertia * impulse;
                }
            }
        } else {
            float weight = a.inverseMass + b.inverseMass;
            float k11 = weight + a.inverseInertia * pointA.x * pointA.x + b.inverseInertia * pointB.x * pointB.x;
            float k12 = a.inverseInertia * pointA.x * pointA.y + b.inverseInertia * pointB.x * pointB.y;
            float k22 = weight + a.inverseInertia * pointA.y * pointA.y + b.inverseInertia * pointB.y * pointB.y;
            float determinant = k11 * k22 - k12 * k12;
            if (determinant > FLT_MIN) {
                F2 gap = b.point - a.point;
                F2 impulse = f2(-(k22 * gap.x - k12 * gap.y), -(-k12 * gap.x + k11 * gap.y)) / determinant;
                moveA = -impulse * a.inverseMass;
                turnA = -a.inverseInertia * dot(pointA, impulse);
                moveB = impulse * b.inverseMass;
                turnB = b.inverseInertia * dot(pointB, impulse);
            }
            float inertia = a.inverseInertia + b.inverseInertia;
            if (inertia > FLT_MIN) {
                float angle = (b.angle + turnB) - (a.angle + turnA) - joint.parameters[3];
                float violation = 0;
                if (kind == 2) violation = angle;
                else if (joint.ends[3] & 1u)
                    violation = angle < joint.limits[0] ? angle - joint.limits[0]
                                                        : (angle > joint.limits[1] ? angle - joint.limits[1] : 0.0f);
                if (violation != 0) {
                    float lambda = -violation / inertia;
                    turnA = turnA - a.inverseInertia * lambda;
                    turnB = turnB + b.inverseInertia * lambda;
                }
                if (kind == 1 && (joint.ends[3] & 2u)) {
                    float relative = (b.omega + turnB / dt) - (a.omega + turnA / dt) - joint.limits[2];
                    float limit = joint.limits[3] * dt;
                    float total = clampf(joint.state[1] - relative / inertia, -limit, limit);
                    float lambda = total - joint.state[1];
                    joint.state[1] = total;
                    spinA = spinA - a.inverseInertia * lambda;
                    spinB = spinB + b.inverseInertia * lambda;
                }
            }
        }
        Corr &r = w.jcorr[index];
        r.first[0] = moveA.x; r.first[1] = moveA.y; r.first[2] = turnA; r.first[3] = 0;
        r.second[0] = moveB.x; r.second[1] = moveB.y; r.second[2] = turnB; r.second[3] = 0;
        F2 v = moveA / dt + speedA;
        r.vfirst[0] = v.x; r.vfirst[1] = v.y; r.vfirst[2] = turnA / dt + spinA; r.vfirst[3] = 0;
        v = moveB / dt + speedB;
        r.vsecond[0] = v.x; r.vsecond[1] = v.y; r.vsecond[2] = turnB / dt + spinB; r.vsecond[3] = 0;
    }
}

// Contact:551-568
static void apply_joint_bodies(World &w) {
    for (uint32_t index = 0; index < w.n(); ++index) {
        const uint32_t *range = w.ranges + 4 * index;
        if (range[1] == 0) continue;
        float delta[3] = {0, 0, 0}, velocity[3] = {0, 0, 0};
        for (uint32_t i = 0; i < range[1]; ++i) {
            const uint32_t *ref = w.refs + 2 * (range[0] + i);
            const Corr &c = w.jcorr[ref[0]];
            for (int d = 0; d < 3; ++d) {
                delta[d] += ref[1] ? c.second[d] : c.first[d];
                velocity[d] += ref[1] ? c.vsecond[d] : c.vfirst[d];
            }
        }
        float relax = 1.0f / float(range[1]);
        bp_rigid_body &x = w.b[index];
        for (int d = 0; d < 3; ++d) {
            x.pose[d] += delta[d] * relax;
            x.motion[d] += velocity[d] * relax;
        }
    }
}

// PostSolve:116-178
static void guard_walls(World &w) {
    for (uint32_t index = 0; index < w.n(); ++index) {
        if (w.b[index].flags[0] || w.b[index].filter[2]) continue;
        bp_rigid_body b = w.b[index];
        for (uint32_t j = 0; j < w.n(); ++j) {
            const bp_rigid_body &wall = w.b[j];
            if (!wall.flags[0] || is_round(wall) || !filter_accepts(b.filter, wall.filter)) continue;
            if (w.excluded(index, wall.id)) continue;
            uint32_t count = wall.ranges[1];
            for (uint32_t k = 0; k < b.ranges[1]; ++k) {
                F2 p = w.vertex(b, k);
                F2 start = f2(b.previous[0], b.previous[1]) + rotate(w.vert(b.ranges[0] + k), b.pose[3]);
                bool inside = rigid_inside(w, p, wall);
                bool startedInside = rigid_inside(w, start, wall);
                if (startedInside)
                    for (uint32_t e = 0; e < count; ++e) {
                        F2 ea = w.vertex(wall, e), eb = w.vertex(wall, (e + 1) % count);
                        if (distance(start, closest_point(start, ea, eb)) < 1e-5f) {
                            startedInside = false;
                            break;
                        }
                    }
                float nearest = INFINITY;
                F2 normal = f2(0, 0);
                float depth = 0;
                for (uint32_t e = 0; e < count; ++e) {
                    F2 ea = w.vertex(wall, e), eb = w.vertex(wall, (e + 1) % count);
                    MovingSeg edge{ea, eb, ea, eb};
                    F2 direction = eb - ea, out = normalize(f2(direction.y, -direction.x));
                    float signedStart = dot(start - ea, out);
                    bool onBoundary = distance(start, closest_point(start, ea, eb)) < 1e-5f;
                    bool crossing = signedStart >= -1e-5f &&
                                    (center_crossing(start, p, edge) || (onBoundary && dot(p - ea, out) < 0));
                    if (!inside && !crossing) continue;
                    if (inside && !startedInside && signedStart < -1e-5f) continue;
                    F2 closest = closest_point(p, ea, eb);
                    float d = (crossing || !startedInside) ? distance(start, closest_point(start, ea, eb)) : distance(p, closest);
                    if (d < nearest) {
                        nearest = d;
                        normal = out;
                        depth = fmax_(0.0f, -dot(p - closest, out));
                    }
                }
                if (std::isfinite(nearest) && depth > 0) {
                    if (is_round(b)) {
                        bp_rigid_body old = b;
                        old.pose[0] = b.previous[0];
                        old.pose[1] = b.previous[1];
                        old.pose[2] = b.pose[3];
                        if (round_core_crosses_polygon(w, core_of(w, old), wall)) break;
                    }
                    F2 shift = normal * (depth + 1e-5f);
                    b.pose[0] += shift.x;
                    b.pose[1] += shift.y;
                    F2 r = p - w.pose_xy(b);
                    float lever = cross(r, normal);
                    float speed = dot(f2(b.motion[0], b.motion[1]), normal) + b.motion[2] * lever;
                    float denom = b.physical[0] + b.physical[1] * lever * lever;
                    float impulse = fmax_(0.0f, -speed) / fmax_(denom, FLT_MIN);
                    F2 dv = normal * b.physical[0] * impulse;
                    b.motion[0] += dv.x;
                    b.motion[1] += dv.y;
                    b.motion[2] += lever * b.physical[1] * impulse;
                }
            }
        }
        w.b[index] = b;
    }
}

// PostSolve:180-206
static void finish(World &w) {
    const float *region = w.cfg.region;
    for (uint32_t index = 0; index < w.n(); ++index) {
        bp_rigid_body b = w.b[index];
        if (!finite4(b.pose) || !finite4(b.previous) || !finite4(b.motion)) {
            w.safety |= 1;
            continue;
        }
        if (b.flags[0]) continue;
        F2 v = clamp_linear(f2(b.motion[0], b.motion[1]), b.limits[0]);
        b.motion[0] = v.x;
        b.motion[1] = v.y;
        b.motion[2] = clamp_angular(b.motion[2], b.limits[1]);
        F2 low = f2(INFINITY, INFINITY), high = f2(-INFINITY, -INFINITY);
        for (uint32_t i = 0; i < b.ranges[1]; ++i) {
            F2 p = w.vertex(b, i);
            low = min2(low, p);
            high = max2(high, p);
        }
        low = low - f2(b.shape[1], b.shape[1]);
        high = high + f2(b.shape[1], b.shape[1]);
        F2 shift = f2(0, 0), size = f2(region[2] - region[0], region[3] - region[1]);
        if (g_plant == 4) size = f2(1, 1);
        bool outside = high.x < region[0] || high.y < region[1] || low.x > region[2] || low.y > region[3];
        if (outside) {
            if (!b.flags[1]) b.flags[2] = 1;
            else {
                if (high.x < region[0]) shift.x = std::ceil((region[0] - high.x) / size.x) * size.x;
                if (low.x > region[2]) shift.x = -std::ceil((low.x - region[2]) / size.x) * size.x;
                if (high.y < region[1]) shift.y = std::ceil((region[1] - high.y) / size.y) * size.y;
                if (low.y > region[3]) shift.y = -std::ceil((low.y - region[3]) / size.y) * size.y;
                b.pose[0] += shift.x;
                b.pose[1] += shift.y;
                b.previous[0] += shift.x;
                b.previous[1] += shift.y;
                b.flags[2] = 2;
            }
        }
        w.b[index] = b;
    }
}

// One step of a world; returns the status (NativeMetalSolver.swift:1031-1040).
static bp_status step(World &w, bp_rigid_frame_info *info) {
    w.safety = 0;
    w.overflow = 0;
    w.maxPer = 0;
    std::fill(w.corr.begin(), w.corr.end(), Corr{});
    for (auto &wall : w.walls) wall = Wall{};
    for (auto &j : w.joints) j.state[0] = j.state[1] = j.state[2] = j.state[3] = 0;
    motion_prelude(w);
    predict(w);
    emit_rows(w);
    ccd(w);
    for (uint32_t it = 0; it < w.cfg.iterations; ++it) {
        for (uint32_t owner = 0; owner < w.n(); ++owner)
            for (uint32_t slot = 0; slot < w.rowCount[owner]; ++slot) contact_slot(w, owner, slot);
        apply_bodies(w);
        if (!w.joints.empty()) {
            solve_joints(w);
            apply_joint_bodies(w);
        }
        guard_walls(w);
    }
    finish(w);
    bp_status status = BP_OK;
    uint32_t cause = BP_GUARD_NONE;
    if (w.safety & 1u) status = BP_ERR_NON_FINITE;
    else if (w.safety & 2u) {
        status = BP_ERR_GUARD_FAILURE;
        cause = (w.safety & 16u) ? BP_GUARD_RIGID_ROTATION_BUDGET : BP_GUARD_NONE;
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->status = status;
        info->guard_cause = cause;
        uint32_t pairs = 0;
        for (uint32_t c : w.rowCount) pairs += c;
        info->candidate_pairs = pairs;
        info->candidate_overflow = w.overflow;
        info->max_candidates_per_object = w.maxPer;
        for (const auto &x : w.b) {
            if (x.flags[2] == 1) ++info->removed;
            if (x.flags[2] == 2) ++info->wrapped;
        }
    }
    return status;
}

// NativeRigidPolygonState.contactRecords (NativeRigidPolygonState.swift:367-428), body part.
static uint32_t contact_records(const World &w, bp_contact *out, uint32_t capacity) {
    uint32_t count = 0;
    auto push = [&](const bp_contact &c) {
        if (count < capacity) out[count] = c;
        ++count;
    };
    std::vector<std::pair<uint64_t, uint64_t>> emitted;
    for (uint32_t body = 0; body < w.n(); ++body)
        for (uint32_t slot = 0; slot < w.rowCount[body]; ++slot) {
            size_t index = (size_t)body * w.k + slot;
            const Cand &c = w.rows[index];
            if (c.kind == 3 && c.low != body) continue;
            float nx = w.corr[index].first[3], ny = w.corr[index].second[3];
            if (!(nx != 0 || ny != 0)) continue;
            bp_contact r{};
            r.first = c.first;
            r.second = c.second;
            r.kind = BP_CONTACT_RIGID;
            r.normal[0] = nx;
            r.normal[1] = ny;
            r.point[0] = w.corr[index].vfirst[3];
            r.point[1] = w.corr[index].vsecond[3];
            r.impulse = c.impulse;
            r.index_b = c.high;
            emitted.push_back({(uint64_t)c.first, (uint64_t)c.second});
            push(r);
        }
    for (uint32_t body = 0; body < w.n(); ++body) {
        const Wall &wall = w.walls[body];
        if (!(wall.geometry[0] != 0 || wall.geometry[1] != 0)) continue;
        int32_t other = -1;
        for (uint32_t j = 0; j < w.n(); ++j)
            if (w.b[j].id == (uint64_t)wall.owner) { other = (int32_t)j; break; }
        if (other < 0) continue;
        bool lower = body < (uint32_t)other;
        uint64_t first = lower ? w.b[body].id : (uint64_t)wall.owner;
        uint64_t second = lower ? (uint64_t)wall.owner : w.b[body].id;
        bool seen = false;
        for (auto &e : emitted)
            if (e.first == first && e.second == second) { seen = true; break; }
        if (seen) continue;
        emitted.push_back({first, second});
        bp_contact r{};
        r.first = (int64_t)first;
        r.second = (int64_t)second;
        r.kind = BP_CONTACT_RIGID;
        r.normal[0] = lower ? -wall.geometry[0] : wall.geometry[0];
        r.normal[1] = lower ? -wall.geometry[1] : wall.geometry[1];
        r.point[0] = wall.geometry[2];
        r.point[1] = wall.geometry[3];
        r.impulse = wall.extra[0];
        r.index_b = lower ? (uint32_t)other : body;
        push(r);
    }
    return count;
}

}  // namespace bpr

extern "C" {

void bpo_rigid_set_plant(int code) { bpr::g_plant = code; }

// Runs `steps` steps of one world from `bodies` (updated in place with the last published state:
// a failed step publishes nothing and ends the run). joints (in/out: the last step's lambdas),
// infos[steps] and contacts (the last published step's records) may be null. Returns the number
// of steps run (the failing one included).
uint32_t bpo_rigid_run(const bp_rigid_config *config, bp_rigid_body *bodies, uint32_t body_count, const float *vertices,
                       const uint32_t *triangles, const bp_segment *segments, uint32_t segment_count, bp_joint *joints,
                       uint32_t joint_count, const uint32_t *references, const uint32_t *ranges, const uint64_t *partners,
                       uint32_t steps, bp_rigid_frame_info *infos, bp_contact *contacts, uint32_t contact_capacity,
                       uint32_t *out_contact_count) {
    using namespace bpr;
    World w;
    w.cfg = *config;
    w.b.assign(bodies, bodies + body_count);
    w.verts = reinterpret_cast<const F2 *>(vertices);
    w.tris = triangles;
    w.segs = segments;
    w.ns = segment_count;
    if (joints && joint_count) w.joints.assign(joints, joints + joint_count);
    w.refs = references;
    w.ranges = ranges;
    w.partners = partners;
    w.k = config->candidate_limit == 0 ? 128u : std::min(config->candidate_limit, 128u);
    w.rows.assign((size_t)body_count * w.k, Cand{});
    w.rowCount.assign(body_count, 0);
    w.corr.assign((size_t)body_count * w.k, Corr{});
    w.walls.assign(body_count, Wall{});
    w.jcorr.assign(w.joints.size(), Corr{});
    uint32_t ran = 0;
    if (out_contact_count) *out_contact_count = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        std::vector<bp_rigid_body> published = w.b;
        std::vector<bp_joint> publishedJoints = w.joints;
        bp_status st = step(w, infos ? infos + s : nullptr);
        ++ran;
        if (st != BP_OK) {
            w.b = published;
            w.joints = publishedJoints;
            break;
        }
        if (contacts || out_contact_count) {
            uint32_t n = contact_records(w, contacts, contacts ? contact_capacity : 0);
            if (out_contact_count) *out_contact_count = n;
        }
    }
    std::copy(w.b.begin(), w.b.end(), bodies);
    if (joints && joint_count) std::copy(w.joints.begin(), w.joints.end(), joints);
    return ran;
}

}  // extern "C"
