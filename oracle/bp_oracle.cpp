// bp_oracle.cpp -- CPU oracle for the BubblePhysics 4.0.0 native-metal world step.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker of the HIP library: it is loaded by
// the tests, by __graft_entry__.smoke() and by the CPU-baseline leg of bench.py, and never
// by the product path (bubblephysics_amd), which fails loudly without its HIP library.
//
// It restates, loop by loop and in the same arithmetic order, the Metal kernels of the
// reference (paths relative to the reference repository):
//   Sources/BubblePhysicsNativeMetal/Shaders/NativeForceKernels.metal-source     ("Force")
//   Sources/BubblePhysicsNativeMetal/Shaders/NativeGeometryKernels.metal-source  ("Geometry")
//   Sources/BubblePhysicsNativeMetal/Shaders/NativeContactKernels.metal-source   ("Contact")
//   Sources/BubblePhysicsNativeMetal/Shaders/NativePostSolveKernels.metal-source ("PostSolve")
// and the host control flow of NativeMetalSolver.executeSnapshot and NativeMetalSubstepLoop.run.
// Each function cites the lines it follows. The restatement is templated on the scalar:
// float reproduces the FP32 step, double gives the FP64 run used for the drift statement.
//
// Pinning (see tests/): the contact law against the Double ContactLawReference restated in
// the tests and against closed forms; candidate rows against brute force over all pairs;
// the sorted-grid and hashed-grid broadphases against each other; the integrator against
// exact half-impulse and terminal-velocity solutions; the h rule against the 1D wall scene.
//
// Build: g++ -O2 -ffp-contract=off -fno-fast-math -shared -fPIC (oracle/Makefile).
#include "../include/bubblephysics.h"
#include "../bubblephysics_amd/csrc/bp_detmath.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace bpo {

// Planted mistakes, switched on only by the tests that show the pins catch them (0 = none):
//   1 round core: distance to the core axis without subtracting the core radius;
//   2 polygon obstacle: the nearest vertex taken for the nearest point of the edge;
//   3 region wrap: the period without the collider's diameter;
//   4 edge-side assertion: the side after the move taken at the step's start point;
//   5 body prediction: the angular velocity ignored.
int g_plant = 0;
// Host threads of the per-bubble phases (1 = the plain sequential restatement).
int g_threads = 1;

// ---------------------------------------------------------------------------------------
// Scalar helpers. Float uses the deterministic transcendentals shared with the kernels;
// double uses libm (the FP64 run is a drift reference, not a bitwise one).
// ---------------------------------------------------------------------------------------
template <class T> struct Lim;
template <> struct Lim<float> {
    static float min_normal() { return FLT_MIN; }
    static float epsilon() { return FLT_EPSILON; }
    static float max() { return FLT_MAX; }
    static float sin(float x) { return bpm::sinf_det(x); }
    static float cos(float x) { return bpm::cosf_det(x); }
    static float exp(float x) { return bpm::expf_det(x); }
    static float pow(float b, float e) { return bpm::powf_det(b, e); }
};
template <> struct Lim<double> {
    static double min_normal() { return DBL_MIN; }
    static double epsilon() { return DBL_EPSILON; }
    static double max() { return DBL_MAX; }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double exp(double x) { return std::exp(x); }
    static double pow(double b, double e) { return std::pow(b, e); }
};

template <class T> inline T vmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T vmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T vclamp(T x, T lo, T hi) { return vmin(vmax(x, lo), hi); }
template <class T> inline T inf() { return std::numeric_limits<T>::infinity(); }

template <class T> struct V2 {
    T x, y;
};
template <class T> inline V2<T> v2(T x, T y) { return V2<T>{x, y}; }
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return v2(a.x + b.x, a.y + b.y); }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return v2(a.x - b.x, a.y - b.y); }
template <class T> inline V2<T> operator-(V2<T> a) { return v2(-a.x, -a.y); }
template <class T> inline V2<T> operator*(V2<T> a, T s) { return v2(a.x * s, a.y * s); }
template <class T> inline V2<T> operator*(T s, V2<T> a) { return v2(s * a.x, s * a.y); }
template <class T> inline V2<T> operator/(V2<T> a, T s) { return v2(a.x / s, a.y / s); }
template <class T> inline V2<T> operator-(V2<T> a, T s) { return v2(a.x - s, a.y - s); }
template <class T> inline V2<T> operator+(V2<T> a, T s) { return v2(a.x + s, a.y + s); }
template <class T> inline T dot(V2<T> a, V2<T> b) { return a.x * b.x + a.y * b.y; }
template <class T> inline T length(V2<T> a) { return std::sqrt(a.x * a.x + a.y * a.y); }
template <class T> inline T distance(V2<T> a, V2<T> b) { return length(a - b); }
template <class T> inline V2<T> vmin2(V2<T> a, V2<T> b) { return v2(vmin(a.x, b.x), vmin(a.y, b.y)); }
template <class T> inline V2<T> vmax2(V2<T> a, V2<T> b) { return v2(vmax(a.x, b.x), vmax(a.y, b.y)); }
template <class T> inline bool finite2(V2<T> a) { return std::isfinite(a.x) && std::isfinite(a.y); }
template <class T> inline T cross(V2<T> a, V2<T> b) { return a.x * b.y - a.y * b.x; }

// ---------------------------------------------------------------------------------------
// Records (the fields of the Metal records that the bubble step reads or writes).
// ---------------------------------------------------------------------------------------
template <class T> struct Bubble {            // NativeBubble, Geometry:4-17
    int64_t id;
    V2<T> pos, start;                         // position.xy, position.zw
    V2<T> vel;                                // motion.xy
    T rot, omega;                             // motion.z, motion.w
    T mass, invm, radius;                     // physical.xyz
    T Kc, gamma;                              // contact.y, material.x
    bool forceDriven;                         // contact.w > 0.5
    uint32_t enabled, isStatic, wrap, region; // publicFlags
    uint32_t cat, mask;                       // filter.xy
};

template <class T> struct Segment {           // NativeSegment, Geometry:18-25
    int64_t id, owner;
    V2<T> a, b;                               // current
    V2<T> prevStart;                          // region.xy
    T outward;                                // region.z
    uint32_t flags;                           // BP_SEGMENT_*
    uint32_t cat, mask;
};

template <class T> struct Body {              // NativeRigidBody, Geometry:35-41
    uint64_t id;
    T pose[4];                                // x, y, angle, previous angle
    V2<T> prev;                               // previous.xy
    T motion[4];                              // vx, vy, omega, damping (0 for kinematic)
    T boundR, coreR, halfLen, maxLin, maxAng;
    uint32_t shape, isStatic, kinematic;
    uint32_t cat, mask;
    uint32_t vOff, vCnt, tOff, tCnt;
};

template <class T> struct Motor { int64_t id; V2<T> force; T torque, inertia; };
template <class T> struct Band { int64_t id; V2<T> target; T K, maxForce; };
template <class T> struct Field {              // NativeForceField, Force:11-17
    int64_t id;
    V2<T> position; T rotation, falloff;
    V2<T> direction; T strength, radius;
    uint32_t kind, vOff, vCnt, fOff, fCnt, hasFilter;
};

struct Candidate {                            // NativeCandidate, Geometry:30
    int64_t pa, pb;                           // participants
    uint32_t kind, ia, ib;                    // metadata.xyz
};
struct Edge { uint32_t cand, neighbour, kind, side; };   // adjacency uint4

template <class T> struct SubstepControl {    // NativeSubstepControl, Force:22-26
    T time[4];                                // dt, elapsed, h, split
    uint32_t counts[4];                       // pair substeps, frame substeps, rebuild, limit
    uint64_t limiter[2];                      // object, term
};

struct Assertion {                            // NativeContactAssertion, Force:841-845
    uint32_t kind = 0;
    int64_t segmentId = 0;
    uint64_t objects[2] = {0, 0};
    double detail = 0;
};

template <class T> struct Law { T force, epsA, epsB, stiffness; };    // Force:41-46

constexpr uint32_t kFrameLimit = 1024;        // NativeMetalSubstep.frameLimit
constexpr uint32_t kBucketLimit = 2048;       // nativeGroupBucketLimit, Force:995

// ---------------------------------------------------------------------------------------
// Geometry helpers (Geometry:130-143, 428-490, 661-681).
// ---------------------------------------------------------------------------------------
inline bool filter_accepts(uint32_t ac, uint32_t am, uint32_t bc, uint32_t bm) {   // Geometry:130-132
    return (ac & bm) != 0 && (bc & am) != 0;
}
template <class T> inline bool rigid_immovable(const Body<T> &b) { return b.isStatic != 0 || b.kinematic != 0; }   // :134
template <class T> inline bool rigid_is_round(const Body<T> &b) { return b.shape != BP_SHAPE_POLYGON; }            // :143

template <class T> inline V2<T> rotate(V2<T> p, T angle) {                          // Geometry:136-138
    T c = Lim<T>::cos(angle), s = Lim<T>::sin(angle);
    return v2(p.x * c - p.y * s, p.x * s + p.y * c);
}
template <class T> inline V2<T> rigid_vertex(const Body<T> &b, const std::vector<V2<T>> &verts, uint32_t i) {   // :139-141
    return v2(b.pose[0], b.pose[1]) + rotate(verts[b.vOff + i], b.pose[2]);
}
template <class T> inline V2<T> closest_point(V2<T> p, V2<T> a, V2<T> b) {          // Geometry:436-440
    V2<T> edge = b - a;
    T square = dot(edge, edge);
    T t = square > Lim<T>::min_normal() ? vclamp(dot(p - a, edge) / square, T(0), T(1)) : T(0);
    return a + edge * t;
}
template <class T> struct RoundCore { V2<T> a, b; T radius; };
template <class T> inline RoundCore<T> round_core_of(const Body<T> &b, const std::vector<V2<T>> &verts) {   // :467-473
    RoundCore<T> c;
    c.a = rigid_vertex(b, verts, 0);
    c.b = b.vCnt > 1 ? rigid_vertex(b, verts, 1) : c.a;
    c.radius = b.coreR;
    return c;
}
template <class T>
inline bool rigid_inside(V2<T> p, const Body<T> &b, const std::vector<V2<T>> &verts,
                         const std::vector<uint32_t> &tris) {                        // Geometry:477-490
    if (rigid_is_round(b)) {
        RoundCore<T> core = round_core_of(b, verts);
        return distance(p, closest_point(p, core.a, core.b)) <= core.radius;
    }
    V2<T> local = rotate(p - v2(b.pose[0], b.pose[1]), -b.pose[2]);
    for (uint32_t i = 0; i < b.tCnt; ++i) {
        const uint32_t *t = &tris[4 * (b.tOff + i)];
        V2<T> a = verts[b.vOff + t[0]], bb = verts[b.vOff + t[1]], c = verts[b.vOff + t[2]];
        if (cross(bb - a, local - a) >= 0 && cross(c - bb, local - bb) >= 0 && cross(a - c, local - c) >= 0) return true;
    }
    return false;
}
template <class T> inline bool past_segment_end(V2<T> center, V2<T> from, V2<T> to) {   // Geometry:665-668
    V2<T> edge = to - from, rel = center - from;
    return std::fma(rel.x, edge.x, rel.y * edge.y) >= std::fma(edge.x, edge.x, edge.y * edge.y);
}

// ---------------------------------------------------------------------------------------
// Contact law (Force:28-102).
// ---------------------------------------------------------------------------------------
template <class T> inline T epsilon_from_force_ratio(T q) {                         // Force:50-53
    if (!(q > 0)) return 0;
    return 2 * q / ((2 * q + 1) + std::sqrt(4 * q + 1));
}
template <class T> inline T rest_from_force_ratio(T q) {                            // Force:55-57
    return 2 / (1 + std::sqrt(vmax(4 * q, T(0)) + 1));
}
template <class T> inline Law<T> wall_contact_law(T K, T r, T distanceValue) {     // Force:60-69
    Law<T> law{0, 0, 0, 0};
    T d = vmax(distanceValue, r * T(1e-6));
    if (!(K > 0) || !(r > 0) || !(d < r)) return law;
    T rest = d / r;
    law.epsA = (r - d) / r;
    law.force = K * (r - d) * r / (d * d);
    law.stiffness = K * (1 + law.epsA) / (r * rest * rest * rest);
    return law;
}
template <class T> inline Law<T> pair_contact_law(T Ka, T ra, T Kb, T rb, T distanceValue) {   // Force:75-102
    Law<T> law{0, 0, 0, 0};
    T R = ra + rb;
    T d = vmax(distanceValue, R * T(1e-6));
    T overlap = R - d;
    if (!(Ka > 0) || !(Kb > 0) || !(overlap > 0)) return law;
    T lowest = vmin(Ka, Kb) * overlap * R / (d * d);
    T force = lowest;
    if (Ka != Kb) {
        for (int step = 0; step < 12; ++step) {
            T qa = force / Ka, qb = force / Kb;
            T ea = epsilon_from_force_ratio(qa), eb = epsilon_from_force_ratio(qb);
            T rsa = rest_from_force_ratio(qa), rsb = rest_from_force_ratio(qb);
            T residual = ra * ea + rb * eb - overlap;
            T slope = ra * rsa * rsa * rsa / ((1 + ea) * Ka) + rb * rsb * rsb * rsb / ((1 + eb) * Kb);
            if (!(slope > Lim<T>::min_normal())) break;
            force = vmax(force - residual / slope, lowest);
        }
    }
    T qa = force / Ka, qb = force / Kb;
    T ea = epsilon_from_force_ratio(qa), eb = epsilon_from_force_ratio(qb);
    T rsa = rest_from_force_ratio(qa), rsb = rest_from_force_ratio(qb);
    T compliance = ra * rsa * rsa * rsa / ((1 + ea) * Ka) + rb * rsb * rsb * rsb / ((1 + eb) * Kb);
    law.force = force; law.epsA = ea; law.epsB = eb;
    law.stiffness = compliance > Lim<T>::min_normal() ? 1 / compliance : 0;
    return law;
}

// Wall region and contact, Force:112-140.
template <class T>
inline bool wall_region(const Segment<T> &s, V2<T> center, V2<T> &normal, T &distanceValue) {
    V2<T> start = s.a, end = s.b, edge = end - start;
    T edgeLength = length(edge);
    V2<T> base = edgeLength > 0 ? v2(-edge.y, edge.x) / edgeLength : v2(T(0), T(1));
    if (s.flags & BP_SEGMENT_POLYGON_EDGE) {
        V2<T> outward = base * s.outward;
        T along = dot(center - start, edge);
        bool convexStart = (s.flags & BP_SEGMENT_CONVEX_START) != 0;
        if (convexStart && along <= 0 && past_segment_end(center, s.prevStart, start)) {
            V2<T> delta = center - start;
            distanceValue = length(delta);
            normal = distanceValue > 0 ? delta / distanceValue : outward;
        } else if (along > 0 && !past_segment_end(center, start, end) && dot(center - start, outward) >= 0) {
            normal = outward;
            distanceValue = dot(center - start, outward);
        } else {
            return false;
        }
    } else {
        V2<T> delta = center - closest_point(center, s.a, s.b);
        distanceValue = length(delta);
        normal = distanceValue > 0 ? delta / distanceValue : base;
    }
    return true;
}
template <class T>
inline bool wall_contact(const Segment<T> &s, V2<T> center, T radius, V2<T> &normal, T &distanceValue) {
    return wall_region(s, center, normal, distanceValue) && distanceValue < radius;
}

// Body obstacle, Force:151-210.
template <class T> inline T step_fraction(const SubstepControl<T> &c) {             // Force:151-153
    return c.time[0] > 0 ? vclamp(c.time[1] / c.time[0], T(0), T(1)) : T(1);
}
template <class T> inline Body<T> body_at_fraction(Body<T> body, T fraction) {      // Force:155-159
    body.pose[0] = body.prev.x + (body.pose[0] - body.prev.x) * fraction;
    body.pose[1] = body.prev.y + (body.pose[1] - body.prev.y) * fraction;
    body.pose[2] = body.pose[3] + (body.pose[2] - body.pose[3]) * fraction;
    return body;
}
template <class T> struct Obstacle { V2<T> normal; T distanceValue; V2<T> velocity; T speedBound; };
template <class T>
inline Obstacle<T> body_obstacle(const Body<T> &source, T fraction, const std::vector<V2<T>> &verts,
                                 const std::vector<uint32_t> &tris, V2<T> center) {   // Force:170-210
    Body<T> body = body_at_fraction(source, fraction);
    Obstacle<T> result;
    V2<T> point, normal;
    T distanceValue;
    if (rigid_is_round(body)) {
        RoundCore<T> core = round_core_of(body, verts);
        point = closest_point(center, core.a, core.b);
        V2<T> delta = center - point;
        T away = length(delta);
        V2<T> axis = core.b - core.a;
        T axisLength = length(axis);
        V2<T> fallback = axisLength > T(1e-6) ? v2(-axis.y, axis.x) / length(v2(-axis.y, axis.x)) : v2(T(0), T(1));
        normal = away > T(1e-7) ? delta / away : fallback;
        distanceValue = g_plant == 1 ? away : away - core.radius;
        point = point + normal * core.radius;
    } else {
        T best = inf<T>();
        V2<T> nearest = v2(T(0), T(0)), outward = v2(T(0), T(1));
        for (uint32_t e = 0; e < body.vCnt; ++e) {
            V2<T> p = rigid_vertex(body, verts, e), q = rigid_vertex(body, verts, (e + 1) % body.vCnt);
            V2<T> candidate = g_plant == 2 ? p : closest_point(center, p, q);
            T d = distance(center, candidate);
            if (d < best) {
                best = d; nearest = candidate;
                V2<T> edge = q - p; T size = length(edge);
                outward = size > Lim<T>::min_normal() ? v2(edge.y, -edge.x) / size : v2(T(0), T(1));
            }
        }
        bool inside = rigid_inside(center, body, verts, tris);
        normal = inside ? outward : (best > T(1e-7) ? (center - nearest) / best : outward);
        distanceValue = inside ? -best : best;
        point = nearest;
    }
    V2<T> arm = point - v2(body.pose[0], body.pose[1]);
    result.normal = normal;
    result.distanceValue = distanceValue;
    result.velocity = v2(body.motion[0], body.motion[1]) + body.motion[2] * v2(-arm.y, arm.x);
    result.speedBound = length(v2(body.motion[0], body.motion[1])) + std::fabs(body.motion[2]) * length(arm);
    return result;
}

// ---------------------------------------------------------------------------------------
// The world and its step.
// ---------------------------------------------------------------------------------------
struct Config {
    double dt, contactTol, sepTol;
    double region[4];
    double globalForce[2];
    uint32_t applyRegion, limit, split, minimumMarginOnly, buckets, assertEverySubstep;
    uint32_t injectionEnabled, injectionAfter;
    int64_t injectionObject;
    double injectionPosition[2];
    uint32_t hashedGrid;      // 1 = in-group hashed grid (single dispatch), 0 = sorted grid
    uint32_t assertMode;      // 0 = all pairs (reference loop), 1 = grid-pruned, same verdict
};

struct Frame {
    int32_t status = BP_OK;
    uint32_t substeps = 0, rebuilds = 0, candidatePairs = 0, candidateOverflow = 0, maxCandidates = 0;
    uint32_t removals = 0, wraps = 0;
    double elapsed = 0;
    uint64_t limiterObject = 0, limiterTerm = BP_TERM_NONE;
    Assertion assertion;
};

template <class T> struct World {
    Config cfg{};
    std::vector<Bubble<T>> b;
    std::vector<Segment<T>> s;
    std::vector<Body<T>> bodies;
    std::vector<V2<T>> verts;
    std::vector<uint32_t> tris;               // 4 per triangle
    std::vector<Motor<T>> motors;             // sorted by id
    std::vector<Band<T>> bands;               // sorted by id
    std::vector<Field<T>> fields;             // sorted by id
    std::vector<V2<T>> fieldVerts;
    std::vector<uint64_t> fieldIds;           // ascending within each field's filter range

    // Working state of the step (NativeMetalPersistentState buffers).
    std::vector<T> margin;
    std::vector<V2<T>> ref, a0, acc, cforce;
    std::vector<T> reduceTerms;                 // per-bubble terms of one h reduction pass
    std::vector<std::vector<Edge>> adjacency;
    std::vector<uint32_t> demanded;
    std::vector<Candidate> candidates;        // compacted candidate list
    std::vector<T> stepHistory;
    std::vector<uint32_t> stepTerms;          // limiting term of each substep of the step
    std::vector<uint64_t> stepObjects;        // and the id of the bubble that set it (0 for the step's rest)
    SubstepControl<T> sc{};
    uint32_t overflow[4] = {0, 0, 0, 0};
    uint32_t safety = 0;
    uint32_t candidateCount = 0;
    Assertion assertion;
    T cellWidth = 0;
    uint32_t limit = 128;
    bool captureOn = false;                   // record the state after every substep (captureSubsteps)
    std::vector<V2<T>> captured;              // per substep: (position, velocity) of every bubble

    bool is_static(const Bubble<T> &x) const { return !(x.invm > 0) || (x.enabled && x.isStatic); }   // Force:381
    T inverse_mass(const Bubble<T> &x) const { return is_static(x) ? T(0) : x.invm; }                   // :382
    V2<T> effective_velocity(const Bubble<T> &x) const { return is_static(x) ? v2(T(0), T(0)) : x.vel; }   // :385

    const Band<T> *find_band(int64_t id) const {                             // Force:423-430
        auto it = std::lower_bound(bands.begin(), bands.end(), id, [](const Band<T> &m, int64_t v) { return m.id < v; });
        return it != bands.end() && it->id == id ? &*it : nullptr;
    }
    const Motor<T> *find_motor(int64_t id) const {                           // Force:431-438
        auto it = std::lower_bound(motors.begin(), motors.end(), id, [](const Motor<T> &m, int64_t v) { return m.id < v; });
        return it != motors.end() && it->id == id ? &*it : nullptr;
    }
    V2<T> band_force(const Band<T> &band, V2<T> position) const {           // Force:441-446
        V2<T> force = (band.target - position) * band.K;
        T lim = band.maxForce, magnitude = length(force);
        if (lim > 0 && magnitude > lim) force = force * (lim / magnitude);
        return force;
    }
    // Field force at a point, Force:389-409.
    V2<T> field_force(const Field<T> &f, V2<T> point) const {
        V2<T> delta = point - f.position;
        T c = Lim<T>::cos(f.rotation), s = Lim<T>::sin(f.rotation);
        V2<T> local = v2(c * delta.x + s * delta.y, -s * delta.x + c * delta.y);
        T distanceValue = length(local), radius = f.radius;
        if (f.vCnt == 0) {
            if (!(distanceValue <= radius)) return v2(T(0), T(0));
        } else {
            bool inside = false;
            for (uint32_t i = 0; i < f.vCnt; ++i) {
                V2<T> a = fieldVerts[f.vOff + i], b = fieldVerts[f.vOff + (i + 1) % f.vCnt];
                if ((a.y > local.y) != (b.y > local.y) && local.x < (b.x - a.x) * (local.y - a.y) / (b.y - a.y) + a.x) inside = !inside;
            }
            if (!inside) return v2(T(0), T(0));
        }
        if (f.kind == 0) return v2(c * f.direction.x - s * f.direction.y, s * f.direction.x + c * f.direction.y);
        if (!(distanceValue > 0)) return v2(T(0), T(0));
        T profile = f.falloff == 0 ? T(1) : Lim<T>::pow(vmax(T(0), 1 - distanceValue / radius), f.falloff);
        T scale = f.strength * profile / distanceValue;
        return f.kind == 2 ? v2(-delta.y, delta.x) * scale : delta * scale;
    }
    bool field_affects(const Field<T> &f, uint64_t id) const {              // Force:412-421
        if (!f.hasFilter) return true;
        auto begin = fieldIds.begin() + f.fOff, end = begin + f.fCnt;
        auto it = std::lower_bound(begin, end, id);
        return it != end && *it == id;
    }
    V2<T> non_contact_acceleration(const Bubble<T> &x, V2<T> position) const {   // Force:450-470
        T invm = inverse_mass(x);
        if (!x.forceDriven || !(invm > 0)) return v2(T(0), T(0));
        V2<T> force = v2(T(cfg.globalForce[0]), T(cfg.globalForce[1]));
        for (const Field<T> &f : fields)
            if (field_affects(f, (uint64_t)x.id)) force = force + field_force(f, position);
        if (!bands.empty()) {
            if (const Band<T> *band = find_band(x.id)) force = force + band_force(*band, position);
        }
        if (!motors.empty()) {
            if (const Motor<T> *motor = find_motor(x.id)) force = force + motor->force;
        }
        return force * invm;
    }

    // Begin step, Force:285-293.
    void begin_step() {
        for (auto &x : b) {
            x.region = 0;
            x.start = x.pos;
            if (x.enabled && x.isStatic) { x.vel = v2(T(0), T(0)); x.omega = 0; }
        }
    }

    // Kinematic and static bodies at the start of the step: nativeRigidPredict for the bodies a
    // bubble world may hold (PostSolve:66-82).
    void rigid_predict() {
        T dt = T(cfg.dt);
        for (auto &body : bodies) {
            body.prev = v2(body.pose[0], body.pose[1]);
            body.pose[3] = body.pose[2];
            if (body.isStatic) { body.motion[0] = body.motion[1] = body.motion[2] = 0; continue; }
            body.motion[2] = body.motion[2] * Lim<T>::exp(-body.motion[3] * dt);
            V2<T> lin = v2(body.motion[0], body.motion[1]);
            T speed = length(lin);
            if (body.maxLin > 0 && speed > body.maxLin) lin = lin * (body.maxLin / speed);
            body.motion[0] = lin.x; body.motion[1] = lin.y;
            if (body.maxAng > 0) body.motion[2] = vclamp(body.motion[2], -body.maxAng, body.maxAng);
            body.pose[0] = body.pose[0] + body.motion[0] * dt;
            body.pose[1] = body.pose[1] + body.motion[1] * dt;
            if (g_plant != 5) body.pose[2] = body.pose[2] + body.motion[2] * dt;
            if (!std::isfinite(body.pose[0]) || !std::isfinite(body.pose[1]) || !std::isfinite(body.pose[2]) ||
                !std::isfinite(body.pose[3]) || !std::isfinite(body.motion[0]) || !std::isfinite(body.motion[1]) ||
                !std::isfinite(body.motion[2]))
                safety |= 1;
        }
    }

    // Margin and x0, Force:492-499.
    void bubble_margin(size_t i) {
        const Bubble<T> &x = b[i];
        T t = vmax(sc.time[0] - sc.time[1], T(0));
        T travel = length(effective_velocity(x)) * t + T(0.5) * length(a0[i]) * t * t;
        margin[i] = cfg.minimumMarginOnly ? T(0.25) * x.radius : vmax(T(1.05) * travel, T(0.25) * x.radius);
        ref[i] = x.pos;
    }

    // k-smallest candidate row, Geometry:230-251.
    static bool key_less(const Candidate &a, const Candidate &b) {
        if (a.kind != b.kind) return a.kind < b.kind;
        if (a.pa != b.pa) return a.pa < b.pa;
        return a.pb < b.pb;
    }
    void insert_candidate(std::vector<Candidate> &row, uint32_t &dropped, uint32_t lim, const Candidate &value) {
        size_t location;
        if (row.size() < lim) { row.push_back(value); location = row.size() - 1; }
        else {
            ++dropped;
            if (!key_less(value, row[lim - 1])) return;
            location = lim - 1;
        }
        while (location > 0 && key_less(value, row[location - 1])) { row[location] = row[location - 1]; --location; }
        row[location] = value;
    }

    // Pair candidate, Geometry:262-273 (joints are not part of a bubble world here).
    void bubble_pair_candidate(uint32_t index, uint32_t j, std::vector<Candidate> &row, uint32_t &dropped) {
        const Bubble<T> &a = b[index], &o = b[j];
        if (!filter_accepts(a.cat, a.mask, o.cat, o.mask)) return;
        T reach = a.radius + o.radius + margin[index] + margin[j] + T(cfg.sepTol);
        if (std::fabs(a.pos.x - o.pos.x) <= reach && std::fabs(a.pos.y - o.pos.y) <= reach)
            insert_candidate(row, dropped, limit, Candidate{a.id, o.id, 0u, index, j});
    }

    // Rest of the row: segments, split of the row, overflow counters, Geometry:276-298. The
    // counters are added by the caller (sums and maxima, so the order of rows does not matter).
    void finish_bubble_row(uint32_t index, std::vector<Candidate> &row, uint32_t dropped, std::vector<uint32_t> &rowCounts,
                           uint32_t &droppedTotal, uint32_t &widest) {
        const Bubble<T> &a = b[index];
        for (uint32_t j = 0; j < s.size(); ++j) {
            const Segment<T> &seg = s[j];
            if (!filter_accepts(a.cat, a.mask, seg.cat, seg.mask)) continue;
            T reach = a.radius + margin[index] + T(cfg.sepTol);
            V2<T> lower = vmin2(seg.a, seg.b) - reach;
            V2<T> upper = vmax2(seg.a, seg.b) + reach;
            if (a.pos.x >= lower.x && a.pos.y >= lower.y && a.pos.x <= upper.x && a.pos.y <= upper.y)
                insert_candidate(row, dropped, limit, Candidate{a.id, seg.id, 1u, index, j});
        }
        uint32_t bubbleCount = 0, count = (uint32_t)row.size();
        while (bubbleCount < count && row[bubbleCount].kind == 0) ++bubbleCount;
        rowCounts[index * 2] = bubbleCount;
        rowCounts[index * 2 + 1] = count - bubbleCount;
        droppedTotal += dropped;
        widest = std::max(widest, count + dropped);
    }

    static V2<T> floor_cell_pos(V2<T> p) { return p; }
    struct Cell { int64_t x, y; };
    Cell cell_of(V2<T> p) const {
        return Cell{(int64_t)std::floor(p.x / cellWidth), (int64_t)std::floor(p.y / cellWidth)};
    }

    // Sorted grid: entries in the total order of nativeGridLess (Geometry:191-196); candidate rows
    // from three binary-searched columns (Geometry:300-330).
    void emit_candidates_sorted(std::vector<std::vector<Candidate>> &rows, std::vector<uint32_t> &rowCounts) {
        struct Entry { int64_t cx, cy; uint32_t index, valid; };
        size_t n = b.size();
        std::vector<Entry> grid(n);
        for (size_t i = 0; i < n; ++i) {                                       // nativeBuildGrid, Geometry:179-189
            Entry e{0, 0, (uint32_t)i, 1};
            if (finite2(b[i].pos)) { Cell c = cell_of(b[i].pos); e.cx = c.x; e.cy = c.y; }
            else e.valid = 0;
            grid[i] = e;
        }
        std::sort(grid.begin(), grid.end(), [](const Entry &a, const Entry &c) {   // nativeGridLess
            if (a.valid != c.valid) return a.valid > c.valid;
            if (a.cx != c.cx) return a.cx < c.cx;
            if (a.cy != c.cy) return a.cy < c.cy;
            return a.index < c.index;
        });
        uint32_t droppedTotal = 0, widest = 0;
        for (uint32_t index = 0; index < n; ++index) {
            const Bubble<T> &a = b[index];
            rows[index].clear();
            if (!finite2(a.pos) || !finite2(a.start)) { rowCounts[index * 2] = rowCounts[index * 2 + 1] = 0; continue; }
            Cell cell = cell_of(a.pos);
            uint32_t dropped = 0;
            for (int dx = -1; dx <= 1; ++dx) {
                int64_t tx = cell.x + dx, ty = cell.y - 1;
                size_t low = 0, high = n;
                while (low < high) {
                    size_t mid = low + (high - low) / 2;
                    const Entry &e = grid[mid];
                    if (e.cx < tx || (e.cx == tx && e.cy < ty)) low = mid + 1; else high = mid;
                }
                for (size_t k = low; k < n && grid[k].cx == tx && grid[k].cy <= ty + 2; ++k) {
                    uint32_t j = grid[k].index;
                    if (!grid[k].valid || j <= index) continue;
                    bubble_pair_candidate(index, j, rows[index], dropped);
                }
            }
            finish_bubble_row(index, rows[index], dropped, rowCounts, droppedTotal, widest);
        }
        overflow[0] += droppedTotal;
        overflow[1] = std::max(overflow[1], widest);
    }

    // Hashed grid of the single-workgroup kernel, Force:999-1035, 1066-1108.
    static uint32_t group_bucket(int64_t cx, int64_t cy, uint32_t buckets) {
        return ((uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u) % buckets;
    }
    void emit_candidates_hashed(std::vector<std::vector<Candidate>> &rows, std::vector<uint32_t> &rowCounts) {
        uint32_t buckets = std::min(std::max(cfg.buckets, 1u), kBucketLimit);
        size_t n = b.size();
        std::vector<std::vector<uint32_t>> bucket(buckets);
        std::vector<Cell> cells(n);
        for (uint32_t i = 0; i < n; ++i) {
            if (!finite2(b[i].pos)) continue;
            cells[i] = cell_of(b[i].pos);
            bucket[group_bucket(cells[i].x, cells[i].y, buckets)].push_back(i);
        }
        uint32_t droppedTotal = 0, widest = 0;
        const int ni = (int)n;
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1) \
    reduction(+ : droppedTotal) reduction(max : widest)
        for (int ii = 0; ii < ni; ++ii) {
            const uint32_t index = (uint32_t)ii;
            const Bubble<T> &a = b[index];
            rows[index].clear();
            if (!finite2(a.pos) || !finite2(a.start)) { rowCounts[index * 2] = rowCounts[index * 2 + 1] = 0; continue; }
            Cell cell = cell_of(a.pos);
            uint32_t dropped = 0;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    int64_t tx = cell.x + dx, ty = cell.y + dy;
                    for (uint32_t j : bucket[group_bucket(tx, ty, buckets)]) {
                        if (cells[j].x != tx || cells[j].y != ty) continue;
                        if (j <= index) continue;
                        bubble_pair_candidate(index, j, rows[index], dropped);
                    }
                }
            finish_bubble_row(index, rows[index], dropped, rowCounts, droppedTotal, widest);
        }
        overflow[0] += droppedTotal;
        overflow[1] = std::max(overflow[1], widest);
    }

    // Bubble-body candidate rows (mode "bubbles"), Geometry:350-392 and 410-415.
    void emit_body_candidates(std::vector<std::vector<Candidate>> &bodyRows) {
        for (uint32_t index = 0; index < bodies.size(); ++index) {
            const Body<T> &body = bodies[index];
            std::vector<Candidate> &row = bodyRows[index];
            row.clear();
            uint32_t dropped = 0;
            bool bubbleObstacle = body.isStatic ? rigid_is_round(body) : (body.kinematic != 0);
            if (bubbleObstacle) {
                V2<T> pose = v2(body.pose[0], body.pose[1]);
                V2<T> bodyLow = vmin2(pose, body.prev) - body.boundR;
                V2<T> bodyHigh = vmax2(pose, body.prev) + body.boundR;
                for (uint32_t ai = 0; ai < b.size(); ++ai) {
                    const Bubble<T> &a = b[ai];
                    if (!filter_accepts(a.cat, a.mask, body.cat, body.mask)) continue;
                    T reach = a.radius + vmax(margin[ai], T(0));
                    V2<T> lo = a.pos - reach, hi = a.pos + reach;
                    bool hit = !(hi.x < bodyLow.x || hi.y < bodyLow.y || bodyHigh.x < lo.x || bodyHigh.y < lo.y);
                    if (hit) insert_candidate(row, dropped, limit, Candidate{a.id, (int64_t)body.id, 2u, ai, index});
                }
            }
            uint32_t count = (uint32_t)row.size();
            overflow[0] += dropped;
            overflow[1] = std::max(overflow[1], count + dropped);
        }
    }

    void append_adjacency(uint32_t bubble, Edge e) {                       // Contact:190-195
        uint32_t slot = demanded[bubble]++;
        if (slot < limit) adjacency[bubble].push_back(e);
        else overflow[0] += 1;
    }

    // Whole grid build of one workgroup (Force:1040-1134), sequentially.
    void broadphase() {
        size_t n = b.size();
        for (size_t i = 0; i < n; ++i) {                                       // a0, margin, x0: Force:1046-1049
            a0[i] = non_contact_acceleration(b[i], b[i].pos);
            bubble_margin(i);
        }
        T width = T(cfg.sepTol);                                              // cell width: Force:1052-1064
        for (size_t i = 0; i < n; ++i) width = vmax(width, 2 * (b[i].radius + margin[i]) + T(cfg.sepTol));
        cellWidth = vmax(width, Lim<T>::min_normal());
        std::vector<std::vector<Candidate>> rows(n);
        std::vector<uint32_t> rowCounts(2 * n, 0);
        if (cfg.hashedGrid) emit_candidates_hashed(rows, rowCounts);
        else emit_candidates_sorted(rows, rowCounts);
        std::vector<std::vector<Candidate>> bodyRows(bodies.size());
        if (!bodies.empty()) emit_body_candidates(bodyRows);
        // Scan of 2n row lengths and compaction (Geometry:625-659, Contact:78-92, 106-136).
        uint64_t total = 0;
        for (size_t r = 0; r < 2 * n; ++r) {
            uint32_t kind = (uint32_t)(r / n), bubble = (uint32_t)(r % n);
            total += rowCounts[bubble * 2 + kind];
        }
        uint64_t capacity = (uint64_t)std::max<size_t>(16, n) * 128;          // contactCapacity
        candidateCount = (uint32_t)std::min<uint64_t>(total, capacity);
        if (total > capacity) overflow[0] += (uint32_t)(total - capacity);
        candidates.clear();
        for (uint32_t kind = 0; kind < 2; ++kind)
            for (uint32_t bubble = 0; bubble < n; ++bubble) {
                uint32_t start = kind == 0 ? 0 : rowCounts[bubble * 2];
                uint32_t cnt = rowCounts[bubble * 2 + kind];
                for (uint32_t k = 0; k < cnt && candidates.size() < candidateCount; ++k) candidates.push_back(rows[bubble][start + k]);
            }
        // Adjacency (Force:298-331, Contact:190-238).
        for (size_t i = 0; i < n; ++i) { adjacency[i].clear(); demanded[i] = 0; }
        for (uint32_t idx = 0; idx < candidates.size(); ++idx) {
            const Candidate &c = candidates[idx];
            append_adjacency(c.ia, Edge{idx + 1, c.ib, c.kind, c.kind == 0 ? 0u : 1u});
            if (c.kind == 0) append_adjacency(c.ib, Edge{idx + 1, c.ia, 0u, 1u});
        }
        for (uint32_t body = 0; body < bodyRows.size(); ++body)
            for (const Candidate &c : bodyRows[body])
                if (c.kind == 2 && c.ia < n) append_adjacency(c.ia, Edge{1u, body, 2u, 1u});
        for (size_t i = 0; i < n; ++i) {                                       // nativeBuildAdjacencyRow
            std::vector<Edge> &row = adjacency[i];
            for (size_t k = 1; k < row.size(); ++k) {
                Edge value = row[k];
                size_t loc = k;
                while (loc > 0 && (value.kind < row[loc - 1].kind ||
                                   (value.kind == row[loc - 1].kind && value.neighbour < row[loc - 1].neighbour))) {
                    row[loc] = row[loc - 1]; --loc;
                }
                row[loc] = value;
            }
            for (Edge &e : row) e.cand -= 1;
            overflow[1] = std::max(overflow[1], demanded[i]);
        }
    }

    // Contact force of one bubble: sum over its adjacency row, Force:216-243.
    V2<T> bubble_contact_force(uint32_t index, T fraction) const {
        const Bubble<T> &bubble = b[index];
        T radius = bubble.radius, K = bubble.Kc;
        V2<T> total = v2(T(0), T(0));
        for (const Edge &e : adjacency[index]) {
            if (e.kind == 0) {
                uint32_t low = std::min(index, e.neighbour), high = std::max(index, e.neighbour);
                const Bubble<T> &first = b[low], &second = b[high];
                V2<T> delta = second.pos - first.pos;
                T d = length(delta);
                V2<T> normal = d > 0 ? delta / d : v2(T(1), T(0));
                Law<T> law = pair_contact_law(first.Kc, first.radius, second.Kc, second.radius, d);
                total = total + (index == low ? -law.force : law.force) * normal;
            } else if (e.kind == 1) {
                V2<T> normal; T d;
                if (!wall_contact(s[e.neighbour], bubble.pos, radius, normal, d)) continue;
                total = total + wall_contact_law(K, radius, d).force * normal;
            } else {
                Obstacle<T> ob = body_obstacle(bodies[e.neighbour], fraction, verts, tris, bubble.pos);
                total = total + wall_contact_law(K, radius, ob.distanceValue).force * ob.normal;
            }
        }
        return total;
    }

    // omega^2, Force:510-541.
    T omega_squared(uint32_t index, T fraction) const {
        const Bubble<T> &bubble = b[index];
        T invm = inverse_mass(bubble);
        if (!(invm > 0)) return 0;
        T radius = bubble.radius, K = bubble.Kc, total = 0;
        if (!bands.empty()) {
            if (const Band<T> *band = find_band(bubble.id)) total = total + band->K * invm;
        }
        for (const Edge &e : adjacency[index]) {
            if (e.kind == 0) {
                uint32_t low = std::min(index, e.neighbour), high = std::max(index, e.neighbour);
                const Bubble<T> &first = b[low], &second = b[high];
                T d = length(second.pos - first.pos);
                Law<T> law = pair_contact_law(first.Kc, first.radius, second.Kc, second.radius, d);
                total = total + law.stiffness * (invm + std::sqrt(invm * inverse_mass(b[e.neighbour])));
            } else if (e.kind == 1) {
                V2<T> normal; T d;
                if (!wall_contact(s[e.neighbour], bubble.pos, radius, normal, d)) continue;
                total = total + wall_contact_law(K, radius, d).stiffness * invm;
            } else {
                Obstacle<T> ob = body_obstacle(bodies[e.neighbour], fraction, verts, tris, bubble.pos);
                total = total + wall_contact_law(K, radius, ob.distanceValue).stiffness * invm;
            }
        }
        return total;
    }

    // d/4 bound, Force:550-598.
    T distance_bound(uint32_t index, T hMax, T fraction) const {
        const Bubble<T> &bubble = b[index];
        if (!(inverse_mass(bubble) > 0)) return inf<T>();
        V2<T> velocity = bubble.vel;
        T radius = bubble.radius, best = inf<T>();
        for (const Edge &e : adjacency[index]) {
            T d, approach, accelerationSum, reach, bodySpeed = 0;
            if (e.kind == 0) {
                const Bubble<T> &other = b[e.neighbour];
                V2<T> delta = other.pos - bubble.pos;
                d = length(delta);
                V2<T> normal = d > 0 ? delta / d : v2(T(1), T(0));
                approach = dot(velocity - effective_velocity(other), normal);
                accelerationSum = length(acc[index]) + length(acc[e.neighbour]);
                reach = radius + other.radius;
            } else if (e.kind == 1) {
                V2<T> normal;
                if (!wall_region(s[e.neighbour], bubble.pos, normal, d)) continue;
                approach = -dot(velocity, normal);
                accelerationSum = length(acc[index]);
                reach = radius;
            } else {
                Obstacle<T> ob = body_obstacle(bodies[e.neighbour], fraction, verts, tris, bubble.pos);
                d = ob.distanceValue;
                approach = -dot(velocity, ob.normal);
                bodySpeed = ob.speedBound;
                accelerationSum = length(acc[index]);
                reach = radius;
            }
            bool active = d < reach;
            T dd = vmax(d, reach * T(1e-6));
            T speed;
            if (active) speed = std::fabs(approach) + bodySpeed;
            else {
                if (approach + bodySpeed + accelerationSum * hMax * T(0.5) <= 0) continue;
                speed = approach + bodySpeed;
            }
            T sq = std::sqrt(speed * speed + accelerationSum * dd * T(0.5));
            T denominator = sq + speed;
            if (!(denominator > 0)) continue;
            best = vmin(best, dd / (2 * denominator));
        }
        return best;
    }

    // Margin bound, Force:604-614.
    T margin_bound(uint32_t index) const {
        const Bubble<T> &bubble = b[index];
        if (!(inverse_mass(bubble) > 0)) return inf<T>();
        T left = margin[index] - length(bubble.pos - ref[index]);
        if (std::isnan(left)) return inf<T>();
        if (left <= 0) return 0;
        T speed = length(bubble.vel), a = length(acc[index]);
        T denominator = speed + std::sqrt(speed * speed + 2 * a * left);
        return denominator > 0 ? 2 * left / denominator : inf<T>();
    }

    // Reduction of h for the whole world, Force:625-736.
    void reduce_substep() {
        T remaining = sc.time[0] - sc.time[1];
        bool over = !(remaining > 0), stopped = sc.counts[2] != 0 || sc.counts[3] != 0;
        bool limited = sc.counts[1] >= kFrameLimit;
        if (over || stopped || limited) {
            if (limited && !over && !stopped) sc.counts[3] = 1;
            sc.time[2] = 0;
            return;
        }
        uint32_t n = (uint32_t)b.size();
        T fraction = step_fraction(sc);
        // Each pass evaluates every bubble's term (in parallel with g_threads), then takes the
        // extreme in bubble order: the lowest index wins ties, whatever the thread count.
        std::vector<T> &terms = reduceTerms;
        terms.resize(n);
        const int ni = (int)n;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int i = 0; i < ni; ++i) terms[i] = omega_squared((uint32_t)i, fraction);
        T best = -1; uint32_t owner = 0;                                       // pass 1: max omega^2
        for (uint32_t i = 0; i < n; ++i) { T v = terms[i]; if (v > best) { best = v; owner = i; } }
        T omegaSquared = best; uint32_t omegaOwner = owner;
        T stiffnessBound = omegaSquared > 0 ? 1 / std::sqrt(omegaSquared) : inf<T>();
        T hMax = vmin(remaining, stiffnessBound);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int i = 0; i < ni; ++i) terms[i] = distance_bound((uint32_t)i, hMax, fraction);
        best = inf<T>(); owner = 0;                                            // pass 2: min d/4 term
        for (uint32_t i = 0; i < n; ++i) { T v = terms[i]; if (v < best) { best = v; owner = i; } }
        T distanceBoundV = best; uint32_t distanceOwner = owner;
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int i = 0; i < ni; ++i) terms[i] = margin_bound((uint32_t)i);
        best = inf<T>(); owner = 0;                                            // pass 3: min margin term
        for (uint32_t i = 0; i < n; ++i) { T v = terms[i]; if (v < best) { best = v; owner = i; } }
        T marginBoundV = best; uint32_t marginOwner = owner;
        if (marginBoundV <= 0) {
            sc.counts[2] = 1;
            sc.time[2] = 0;
            sc.limiter[0] = (uint64_t)b[marginOwner].id; sc.limiter[1] = BP_TERM_MARGIN;
            return;
        }
        T split = vmax(sc.time[3], T(1));
        T h = remaining;
        uint64_t term = BP_TERM_REMAINING, object = 0;
        if (split > 1 && sc.time[0] / split < h) h = sc.time[0] / split;
        if (stiffnessBound / split < h) { h = stiffnessBound / split; term = BP_TERM_STIFFNESS; object = (uint64_t)b[omegaOwner].id; }
        if (distanceBoundV / split < h) { h = distanceBoundV / split; term = BP_TERM_DISTANCE; object = (uint64_t)b[distanceOwner].id; }
        if (marginBoundV < h) { h = marginBoundV; term = BP_TERM_MARGIN; object = (uint64_t)b[marginOwner].id; sc.counts[2] = 1; }
        h = vmax(h, sc.time[0] * T(1e-7));
        uint32_t frameSubstep = sc.counts[1];
        if (h >= remaining || remaining - h <= sc.time[0] * T(1e-5)) { h = remaining; sc.time[1] = sc.time[0]; }
        else sc.time[1] = sc.time[1] + h;
        sc.time[2] = h;
        sc.counts[0] += 1;
        sc.counts[1] = frameSubstep + 1;
        sc.limiter[0] = object; sc.limiter[1] = term;
        if (frameSubstep < stepHistory.size()) {
            stepHistory[frameSubstep] = h;
            stepTerms[frameSubstep] = (uint32_t)term;
            stepObjects[frameSubstep] = object;
        }
    }

    // Exact half impulse with drag, Force:754-758.
    static T half_impulse_factor(T gamma, T halfStep) {
        T x = gamma * halfStep;
        if (x < T(0.05))
            return halfStep * (1 - x * T(0.5) + x * x * (T(1) / 6) - x * x * x * (T(1) / 24) + x * x * x * x * (T(1) / 120));
        return (1 - Lim<T>::exp(-x)) / gamma;
    }

    // Leapfrog phases, Force:765-802.
    void integrate_phase(uint32_t index, uint32_t phase) {
        Bubble<T> &bubble = b[index];
        T invm = inverse_mass(bubble);
        if (phase == 2) { acc[index] = cforce[index] * invm + a0[index]; return; }
        T h = sc.time[2];
        if (!(h > 0)) return;
        T halfStep = T(0.5) * h;
        T gamma = bubble.forceDriven ? vmax(bubble.gamma, T(0)) : T(0);
        T decay = Lim<T>::exp(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
        if (phase == 0) {
            if (!(invm > 0)) return;
            V2<T> velocity = bubble.vel * decay + acc[index] * factor;
            T angular = bubble.omega * Lim<T>::exp(-gamma * h);
            if (!motors.empty() && bubble.forceDriven) {
                const Motor<T> *motor = find_motor(bubble.id);
                if (motor && motor->inertia > 0) angular = angular + motor->torque / motor->inertia * half_impulse_factor(gamma, h);
            }
            V2<T> position = bubble.pos;
            T rot = bubble.rot, omega = bubble.omega;
            bubble.pos = position + velocity * h;
            bubble.vel = velocity;
            bubble.rot = rot + omega * h;
            bubble.omega = angular;
            return;
        }
        if (invm > 0) {
            V2<T> a = cforce[index] * invm + non_contact_acceleration(bubble, bubble.pos);
            bubble.vel = bubble.vel * decay + a * factor;
            acc[index] = a;
        }
    }

    // Contact assertion of one bubble, Force:865-918 (first violation in bubble order wins).
    bool assert_contacts_of(uint32_t index, const std::vector<uint32_t> *pairCandidates) {
        const Bubble<T> &bubble = b[index];
        V2<T> center = bubble.pos, start = bubble.start;
        if (!finite2(center) || bubble.region == 1) return false;
        bool movable = inverse_mass(bubble) > 0;
        auto pair_check = [&](uint32_t other) -> bool {
            const Bubble<T> &nb = b[other];
            if (nb.region == 1 || !filter_accepts(bubble.cat, bubble.mask, nb.cat, nb.mask)) return false;
            if (!movable && !(inverse_mass(nb) > 0)) return false;
            T d = length(nb.pos - center);
            if (d <= 0) {
                assertion.kind = BP_VIOLATION_PAIR; assertion.objects[0] = (uint64_t)bubble.id;
                assertion.objects[1] = (uint64_t)nb.id; assertion.detail = (double)d;
                return true;
            }
            return false;
        };
        if (pairCandidates) {
            for (uint32_t other : *pairCandidates) if (pair_check(other)) return true;
        } else {
            for (uint32_t other = index + 1; other < b.size(); ++other) if (pair_check(other)) return true;
        }
        if (!movable) return false;
        T fraction = step_fraction(sc);
        for (const Body<T> &body : bodies) {
            if (!rigid_immovable(body) || !filter_accepts(bubble.cat, bubble.mask, body.cat, body.mask)) continue;
            if (rigid_inside(center, body_at_fraction(body, fraction), verts, tris)) {
                assertion.kind = BP_VIOLATION_INSIDE_BODY; assertion.objects[0] = (uint64_t)bubble.id;
                assertion.objects[1] = body.id; assertion.detail = 0;
                return true;
            }
        }
        for (const Segment<T> &seg : s) {
            if (!(seg.flags & BP_SEGMENT_POLYGON_EDGE) || !filter_accepts(bubble.cat, bubble.mask, seg.cat, seg.mask)) continue;
            V2<T> edge = seg.b - seg.a;
            T lengthSquared = dot(edge, edge);
            if (!(lengthSquared > 0)) continue;
            V2<T> outward = v2(-edge.y, edge.x) / std::sqrt(lengthSquared) * seg.outward;
            T before = dot(start - seg.a, outward), after = dot((g_plant == 4 ? start : center) - seg.a, outward);
            if (!(before > 0) || !(after <= 0)) continue;
            V2<T> crossing = start + (center - start) * (before / (before - after));
            T along = dot(crossing - seg.a, edge) / lengthSquared;
            if (along < 0 || along > 1) continue;
            assertion.kind = BP_VIOLATION_EDGE_SIDE; assertion.segmentId = seg.id;
            assertion.objects[0] = (uint64_t)bubble.id; assertion.objects[1] = (uint64_t)seg.owner;
            assertion.detail = (double)after;
            return true;
        }
        return false;
    }

    void assert_contacts() {
        if (assertion.kind != 0) return;                                       // sticky within the frame
        size_t n = b.size();
        if (cfg.assertMode == 1) {
            // Same verdict with the pair loop restricted to the 3x3 cells around the bubble: d <= 0
            // needs dx*dx + dy*dy == 0, i.e. |dx|, |dy| below 4e-23, far below any cell width.
            T w = cellWidth > 0 ? cellWidth : T(1);
            std::vector<std::pair<std::pair<int64_t, int64_t>, uint32_t>> cells;
            for (uint32_t i = 0; i < n; ++i)
                if (finite2(b[i].pos))
                    cells.push_back({{(int64_t)std::floor(b[i].pos.x / w), (int64_t)std::floor(b[i].pos.y / w)}, i});
            std::sort(cells.begin(), cells.end());
            std::vector<std::vector<uint32_t>> later(n);
            for (const auto &entry : cells) {
                uint32_t i = entry.second;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        std::pair<int64_t, int64_t> key{entry.first.first + dx, entry.first.second + dy};
                        auto it = std::lower_bound(cells.begin(), cells.end(), std::make_pair(key, 0u));
                        for (; it != cells.end() && it->first == key; ++it)
                            if (it->second > i) later[i].push_back(it->second);
                    }
                std::sort(later[i].begin(), later[i].end());
            }
            for (uint32_t i = 0; i < n; ++i) if (assert_contacts_of(i, &later[i])) return;
            return;
        }
        for (uint32_t i = 0; i < n; ++i) if (assert_contacts_of(i, nullptr)) return;
    }

    // Region removal and wrap, PostSolve:536-566.
    void apply_public_region() {
        V2<T> lo = v2(T(cfg.region[0]), T(cfg.region[1])), hi = v2(T(cfg.region[2]), T(cfg.region[3]));
        for (auto &x : b) {
            if (!x.enabled || x.isStatic) continue;
            V2<T> nearest = v2(vclamp(x.pos.x, lo.x, hi.x), vclamp(x.pos.y, lo.y, hi.y));
            if (length(x.pos - nearest) <= x.radius) continue;
            if (!x.wrap) { x.region = 1; continue; }
            V2<T> delta = v2(T(0), T(0));
            V2<T> span = g_plant == 3 ? hi - lo : hi - lo + 2 * x.radius;
            T *dp[2] = {&delta.x, &delta.y};
            T pos[2] = {x.pos.x, x.pos.y}, rlo[2] = {lo.x, lo.y}, rhi[2] = {hi.x, hi.y}, sp[2] = {span.x, span.y};
            for (int axis = 0; axis < 2; ++axis) {
                if (pos[axis] > rhi[axis])
                    *dp[axis] = -sp[axis] * vmax(T(1), std::ceil((pos[axis] - x.radius - rhi[axis]) / sp[axis]));
                else if (pos[axis] < rlo[axis])
                    *dp[axis] = sp[axis] * vmax(T(1), std::ceil((rlo[axis] - x.radius - pos[axis]) / sp[axis]));
            }
            V2<T> wrapped = x.pos + delta;
            V2<T> nearestWrapped = v2(vclamp(wrapped.x, lo.x, hi.x), vclamp(wrapped.y, lo.y, hi.y));
            if (length(wrapped - nearestWrapped) > x.radius) delta = delta + (nearestWrapped - wrapped);
            V2<T> movedPos = x.pos + delta, movedStart = x.start + delta;
            if (!finite2(movedPos) || !finite2(movedStart)) { safety |= 1; continue; }   // record left unwritten
            x.pos = movedPos;
            x.start = movedStart;
            x.region = 2;
        }
    }

    // NaN / Inf guard on bubbles and segments, PostSolve:5-27.
    void check_finite() {
        for (const auto &x : b) {
            if (!finite2(x.pos) || !finite2(x.start) || !finite2(x.vel) || !std::isfinite(x.rot) || !std::isfinite(x.omega) ||
                !std::isfinite(x.mass) || !std::isfinite(x.invm) || !std::isfinite(x.radius))
                safety |= 1;
        }
        for (const auto &seg : s) {
            V2<T> e = seg.b - seg.a;
            if (!finite2(seg.a) || !finite2(seg.b) || !finite2(e) || !std::isfinite(dot(e, e))) safety |= 1;
        }
    }

    // The per-bubble phases below run on g_threads host threads when it is above 1 (the per-phase
    // CPU baseline): each iteration writes only its own bubble's slots and reads state no other
    // iteration of the same loop writes, so any thread count gives the same bits.
    void contact_forces_all(T fraction) {
        const int n = (int)b.size();
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int i = 0; i < n; ++i) cforce[i] = bubble_contact_force((uint32_t)i, fraction);
    }
    void integrate_all(uint32_t phase) {
        const int n = (int)b.size();
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int i = 0; i < n; ++i) integrate_phase((uint32_t)i, phase);
    }

    // One step: executeSnapshot prologue (NativeMetalSolver.swift:467-498), step prefix
    // (:768-799), the single-workgroup substep loop (Force:1236-1307), end of step (:837-945)
    // and the host checks in error precedence (:1026-1067).
    Frame step() {
        Frame frame;
        size_t n = b.size();
        limit = std::min(std::max(cfg.limit, 1u), 128u);
        margin.assign(n, 0); ref.assign(n, v2(T(0), T(0))); a0.assign(n, v2(T(0), T(0)));
        acc.assign(n, v2(T(0), T(0))); cforce.assign(n, v2(T(0), T(0)));
        adjacency.assign(n, {}); demanded.assign(n, 0);
        stepHistory.assign(kFrameLimit, T(0));
        stepTerms.assign(kFrameLimit, 0u);
        stepObjects.assign(kFrameLimit, 0u);
        std::memset(overflow, 0, sizeof overflow);
        safety = 0; assertion = Assertion{}; candidateCount = 0; candidates.clear(); captured.clear();
        sc = SubstepControl<T>{};
        sc.time[0] = T(cfg.dt);
        sc.time[1] = n == 0 ? T(cfg.dt) : T(0);
        sc.time[3] = T(std::max(cfg.split, 1u));

        begin_step();
        rigid_predict();
        if (n > 0) {
            broadphase();
            T fraction = step_fraction(sc);
            contact_forces_all(fraction);
            integrate_all(2);
            for (uint32_t iteration = 0; iteration < 2 * kFrameLimit + 2; ++iteration) {
                reduce_substep();
                T h = sc.time[2];
                bool limitedFlag = sc.counts[3] != 0, finished = !(sc.time[0] - sc.time[1] > 0), rebuild = sc.counts[2] != 0;
                if (!(h > 0)) {
                    if (limitedFlag || finished || !rebuild) break;
                    sc.counts[0] = 0; sc.counts[2] = 0;
                    overflow[2] = std::max(overflow[2], overflow[0]);
                    overflow[0] = 0;
                    overflow[3] += 1;
                    broadphase();
                    contact_forces_all(step_fraction(sc));
                    integrate_all(2);
                    continue;
                }
                integrate_all(0);
                contact_forces_all(step_fraction(sc));
                integrate_all(1);
                if (cfg.injectionEnabled) {
                    for (auto &x : b)
                        if (sc.counts[1] == cfg.injectionAfter + 1 && x.id == cfg.injectionObject)
                            x.pos = v2(T(cfg.injectionPosition[0]), T(cfg.injectionPosition[1]));
                }
                if (cfg.assertEverySubstep) assert_contacts();
                if (captureOn)
                    for (const auto &x : b) { captured.push_back(x.pos); captured.push_back(x.vel); }
            }
            overflow[2] = std::max(overflow[2], overflow[0]);
        }
        bool finished = sc.counts[3] == 0 && !(sc.time[0] - sc.time[1] > 0);
        if (finished) {                                                        // gated end of step
            if (cfg.applyRegion) apply_public_region();
            check_finite();
            assert_contacts();
        }
        frame.substeps = sc.counts[1];
        frame.rebuilds = overflow[3];
        frame.candidatePairs = candidateCount;
        frame.candidateOverflow = overflow[2];
        frame.maxCandidates = overflow[1];
        frame.elapsed = (double)sc.time[1];
        frame.limiterObject = sc.limiter[0];
        frame.limiterTerm = (uint32_t)sc.limiter[1];
        for (const auto &x : b) { frame.removals += x.region == 1; frame.wraps += x.region == 2; }
        frame.assertion = assertion;
        if (sc.counts[3] != 0) frame.status = BP_ERR_SUBSTEP_LIMIT_EXCEEDED;
        else if (safety & 1) frame.status = BP_ERR_NON_FINITE;
        else if (assertion.kind != 0) frame.status = BP_ERR_CONTACT_VIOLATION;
        else {
            for (const auto &x : b)
                if (!finite2(x.pos) || !finite2(x.start) || !finite2(x.vel) || !std::isfinite(x.rot) || !std::isfinite(x.omega)) {
                    frame.status = BP_ERR_NON_FINITE;
                    break;
                }
        }
        return frame;
    }
};

// ---------------------------------------------------------------------------------------
// Conversion between the C ABI records and the oracle records.
// ---------------------------------------------------------------------------------------
template <class T> void load(World<T> &w, const bp_world_config *config, const bp_bubble *bubbles, uint32_t n,
                             const bp_segment *segments, uint32_t ns, const bp_body *bodies, uint32_t nb,
                             const float *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt,
                             const bp_motor *motors, uint32_t nm, const bp_band *bands, uint32_t nbands,
                             const bp_field *fields, uint32_t nf, const float *fieldVertices, uint32_t nfv,
                             const int64_t *fieldIds, uint32_t nfi) {
    Config &c = w.cfg;
    c.dt = config->time_step; c.contactTol = config->contact_tolerance; c.sepTol = config->separation_tolerance;
    for (int k = 0; k < 4; ++k) c.region[k] = config->region[k];
    c.globalForce[0] = config->global_force[0]; c.globalForce[1] = config->global_force[1];
    c.applyRegion = config->apply_region; c.limit = config->candidate_limit; c.split = config->substep_split;
    c.minimumMarginOnly = config->minimum_margin_only; c.buckets = config->grid_bucket_count;
    c.assertEverySubstep = config->assert_every_substep; c.injectionEnabled = config->injection_enabled;
    c.injectionAfter = config->injection_after_substep; c.injectionObject = config->injection_object;
    c.injectionPosition[0] = config->injection_position[0]; c.injectionPosition[1] = config->injection_position[1];
    c.hashedGrid = (n <= 1024 && !config->force_cooperative) ? 1u : 0u;
    w.b.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const bp_bubble &x = bubbles[i];
        Bubble<T> &y = w.b[i];
        y.id = x.id;
        y.pos = v2(T(x.center[0]), T(x.center[1])); y.start = v2(T(x.previous_center[0]), T(x.previous_center[1]));
        y.vel = v2(T(x.velocity[0]), T(x.velocity[1]));
        y.rot = x.rotation; y.omega = x.angular_velocity;
        y.mass = x.mass; y.invm = x.inverse_mass; y.radius = x.radius;
        y.Kc = x.contact_stiffness; y.gamma = x.damping;
        y.forceDriven = (x.flags & BP_BUBBLE_FORCE_DRIVEN) != 0;
        y.enabled = (x.flags & BP_BUBBLE_ENABLED) ? 1 : 0; y.isStatic = (x.flags & BP_BUBBLE_STATIC) ? 1 : 0;
        y.wrap = (x.flags & BP_BUBBLE_WRAP) ? 1 : 0; y.region = 0;
        y.cat = x.category; y.mask = x.mask;
    }
    w.s.resize(ns);
    for (uint32_t i = 0; i < ns; ++i) {
        const bp_segment &x = segments[i];
        Segment<T> &y = w.s[i];
        y.id = x.id; y.owner = x.owner_id;
        y.a = v2(T(x.a[0]), T(x.a[1])); y.b = v2(T(x.b[0]), T(x.b[1]));
        y.prevStart = v2(T(x.previous_start[0]), T(x.previous_start[1]));
        y.outward = x.outward_sign; y.flags = x.flags; y.cat = x.category; y.mask = x.mask;
    }
    w.bodies.resize(nb);
    for (uint32_t i = 0; i < nb; ++i) {
        const bp_body &x = bodies[i];
        Body<T> &y = w.bodies[i];
        y.id = x.id;
        y.pose[0] = x.pose[0]; y.pose[1] = x.pose[1]; y.pose[2] = x.pose[2]; y.pose[3] = x.previous[2];
        y.prev = v2(T(x.previous[0]), T(x.previous[1]));
        y.motion[0] = x.velocity[0]; y.motion[1] = x.velocity[1]; y.motion[2] = x.velocity[2]; y.motion[3] = 0;
        y.boundR = x.bounding_radius; y.coreR = x.core_radius; y.halfLen = x.half_length;
        y.maxLin = x.max_linear_speed; y.maxAng = x.max_angular_speed;
        y.shape = x.shape; y.isStatic = (x.flags & BP_BODY_STATIC) ? 1 : 0; y.kinematic = (x.flags & BP_BODY_KINEMATIC) ? 1 : 0;
        y.cat = x.category; y.mask = x.mask;
        y.vOff = x.vertex_offset; y.vCnt = x.vertex_count; y.tOff = x.triangle_offset; y.tCnt = x.triangle_count;
    }
    w.verts.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) w.verts[i] = v2(T(vertices[2 * i]), T(vertices[2 * i + 1]));
    w.tris.assign(triangles, triangles + 4 * (size_t)nt);
    w.motors.resize(nm);
    for (uint32_t i = 0; i < nm; ++i)
        w.motors[i] = Motor<T>{motors[i].bubble_id, v2(T(motors[i].force[0]), T(motors[i].force[1])), T(motors[i].torque), T(motors[i].inertia)};
    w.bands.resize(nbands);
    for (uint32_t i = 0; i < nbands; ++i)
        w.bands[i] = Band<T>{bands[i].bubble_id, v2(T(bands[i].target[0]), T(bands[i].target[1])), T(bands[i].stiffness), T(bands[i].maximum_force)};
    std::sort(w.motors.begin(), w.motors.end(), [](const Motor<T> &a, const Motor<T> &b) { return a.id < b.id; });
    std::sort(w.bands.begin(), w.bands.end(), [](const Band<T> &a, const Band<T> &b) { return a.id < b.id; });
    w.fields.resize(nf);
    for (uint32_t i = 0; i < nf; ++i) {
        const bp_field &x = fields[i];
        w.fields[i] = Field<T>{x.id, v2(T(x.position[0]), T(x.position[1])), T(x.rotation), T(x.falloff),
                               v2(T(x.direction[0]), T(x.direction[1])), T(x.strength), T(x.radius),
                               x.kind, x.vertex_offset, x.vertex_count, x.filter_offset, x.filter_count, x.has_filter ? 1u : 0u};
    }
    std::stable_sort(w.fields.begin(), w.fields.end(), [](const Field<T> &a, const Field<T> &b) { return a.id < b.id; });
    w.fieldVerts.resize(nfv);
    for (uint32_t i = 0; i < nfv; ++i) w.fieldVerts[i] = v2(T(fieldVertices[2 * i]), T(fieldVertices[2 * i + 1]));
    w.fieldIds.resize(nfi);
    for (uint32_t i = 0; i < nfi; ++i) w.fieldIds[i] = (uint64_t)fieldIds[i];
}

template <class T> void store(const World<T> &w, bp_bubble *bubbles, bp_body *bodies) {
    for (size_t i = 0; i < w.b.size(); ++i) {
        const Bubble<T> &y = w.b[i];
        bp_bubble &x = bubbles[i];
        x.center[0] = (float)y.pos.x; x.center[1] = (float)y.pos.y;
        x.previous_center[0] = (float)y.start.x; x.previous_center[1] = (float)y.start.y;
        x.velocity[0] = (float)y.vel.x; x.velocity[1] = (float)y.vel.y;
        x.rotation = (float)y.rot; x.angular_velocity = (float)y.omega;
    }
    if (bodies)
        for (size_t i = 0; i < w.bodies.size(); ++i) {
            const Body<T> &y = w.bodies[i];
            bp_body &x = bodies[i];
            x.pose[0] = (float)y.pose[0]; x.pose[1] = (float)y.pose[1]; x.pose[2] = (float)y.pose[2];
            x.previous[0] = (float)y.prev.x; x.previous[1] = (float)y.prev.y; x.previous[2] = (float)y.pose[3];
            x.velocity[0] = (float)y.motion[0]; x.velocity[1] = (float)y.motion[1]; x.velocity[2] = (float)y.motion[2];
        }
}

}  // namespace bpo

// ---------------------------------------------------------------------------------------
// C entry points used by the tests (ctypes).
// ---------------------------------------------------------------------------------------
extern "C" {

typedef struct bpo_frame {
    int32_t status;
    uint32_t substeps, rebuilds, candidate_pairs, candidate_overflow, max_candidates;
    uint32_t removals, wraps;
    double elapsed;
    uint64_t limiter_object;
    uint32_t limiter_term;
    uint32_t violation_kind;
    int64_t violation_segment;
    uint64_t violation_objects[2];
    double violation_detail;
} bpo_frame;

// Runs `steps` steps of one world in FP32 (precision 32) or FP64 (precision 64), in place on
// the bubble and body arrays. frames[steps]; stops at the first failing step (its status is
// reported, the arrays then hold the failed state, unlike the published state of the library).
// adjacency_out (optional): row offsets [n+1] and packed entries of the last grid build.
// assert_mode: 0 = all-pairs loop of the reference, 1 = pruned with the same verdict.
int bpo_world_run(const bp_world_config *config, bp_bubble *bubbles, uint32_t n, const bp_segment *segments, uint32_t ns,
                  bp_body *bodies, uint32_t nb, const float *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt,
                  const bp_motor *motors, uint32_t nm, const bp_band *bands, uint32_t nbands,
                  const bp_field *fields, uint32_t nf, const float *field_vertices, uint32_t nfv,
                  const int64_t *field_ids, uint32_t nfi, uint32_t steps,
                  int precision, uint32_t assert_mode, uint32_t hashed_override, bpo_frame *frames,
                  uint32_t *row_offsets, uint32_t *entries, uint32_t entries_capacity, float *step_history,
                  uint32_t *step_terms, uint64_t *step_objects, uint32_t *region_out, float *capture_out,
                  uint32_t capture_capacity, uint32_t *capture_count) {
    auto run = [&](auto &w) -> int {
        bpo::load(w, config, bubbles, n, segments, ns, bodies, nb, vertices, nv, triangles, nt, motors, nm, bands, nbands,
                  fields, nf, field_vertices, nfv, field_ids, nfi);
        w.cfg.assertMode = assert_mode;
        w.captureOn = capture_out != nullptr;
        if (hashed_override == 1) w.cfg.hashedGrid = 1;
        if (hashed_override == 2) w.cfg.hashedGrid = 0;
        int executed = 0;
        for (uint32_t k = 0; k < steps; ++k) {
            bpo::Frame f = w.step();
            executed = (int)k + 1;
            if (frames) {
                bpo_frame &o = frames[k];
                o.status = f.status; o.substeps = f.substeps; o.rebuilds = f.rebuilds; o.candidate_pairs = f.candidatePairs;
                o.candidate_overflow = f.candidateOverflow; o.max_candidates = f.maxCandidates; o.removals = f.removals;
                o.wraps = f.wraps; o.elapsed = f.elapsed; o.limiter_object = f.limiterObject; o.limiter_term = f.limiterTerm;
                o.violation_kind = f.assertion.kind; o.violation_segment = f.assertion.segmentId;
                o.violation_objects[0] = f.assertion.objects[0]; o.violation_objects[1] = f.assertion.objects[1];
                o.violation_detail = f.assertion.detail;
            }
            if (f.status != BP_OK) break;
        }
        bpo::store(w, bubbles, bodies);
        if (region_out)
            for (uint32_t i = 0; i < n; ++i) region_out[i] = w.b[i].region;
        if (capture_out) {
            // Substeps of the last executed step, each as n (x, y, vx, vy) records.
            uint32_t floats = (uint32_t)std::min<size_t>(2 * w.captured.size(), capture_capacity);
            for (uint32_t k = 0; k < floats / 2; ++k) {
                capture_out[2 * k] = (float)w.captured[k].x;
                capture_out[2 * k + 1] = (float)w.captured[k].y;
            }
            if (capture_count) *capture_count = (uint32_t)(w.captured.size() / 2);
        }
        if (row_offsets) {
            uint32_t total = 0;
            for (uint32_t i = 0; i < n; ++i) {
                row_offsets[i] = total;
                for (const auto &e : w.adjacency[i]) {
                    if (entries && total < entries_capacity) entries[total] = (e.kind << 30) | e.neighbour;
                    ++total;
                }
            }
            row_offsets[n] = total;
        }
        if (step_history)
            for (uint32_t k = 0; k < bpo::kFrameLimit; ++k) step_history[k] = (float)w.stepHistory[k];
        if (step_terms)
            for (uint32_t k = 0; k < bpo::kFrameLimit; ++k) step_terms[k] = w.stepTerms[k];
        if (step_objects)
            for (uint32_t k = 0; k < bpo::kFrameLimit; ++k) step_objects[k] = w.stepObjects[k];
        return executed;
    };
    if (precision == 64) { bpo::World<double> w; return run(w); }
    bpo::World<float> w;
    return run(w);
}

// FP64 audit of a world state by brute force over every pair, segment and body: momentum,
// angular momentum, kinetic energy and the elastic energy of every contact the step's force
// law acts on. The contact energy per deflected bubble is U = K r (eps/(1-eps) + ln(1-eps))
// (ContactLawReference.swift:26-31); a pair of different materials splits its overlap by
// r_a eps(F/K_a) + r_b eps(F/K_b) = overlap, solved for F by bisection (:68-82). Distances are
// the FP32 ones of the step's contact law (pair centre distance in FP64 from FP32 centres);
// which pairs, segments and bodies count follows the candidate predicates (filters, body
// kinds: Geometry:262-298, 350-392).
namespace {
double audit_epsilon(double q) { return q > 0 ? 2 * q / ((2 * q + 1) + std::sqrt(4 * q + 1)) : 0.0; }
double audit_spring(double eps, double K, double r) { return eps > 0 ? K * r * (eps / (1 - eps) + std::log1p(-eps)) : 0.0; }
double audit_wall(double K, double r, double d, uint32_t &contacts) {
    double dd = std::fmax(d, r * 1e-6);
    if (!(K > 0) || !(r > 0) || !(dd < r)) return 0.0;
    ++contacts;
    return audit_spring((r - dd) / r, K, r);
}
double audit_pair(double Ka, double ra, double Kb, double rb, double d, uint32_t &contacts) {
    double R = ra + rb, dd = std::fmax(d, R * 1e-6), overlap = R - dd;
    if (!(Ka > 0) || !(Kb > 0) || !(overlap > 0)) return 0.0;
    ++contacts;
    if (Ka == Kb) return audit_spring(overlap / R, Ka, R);
    double lo = 0, hi = std::fmax(Ka, Kb);
    while (ra * audit_epsilon(hi / Ka) + rb * audit_epsilon(hi / Kb) < overlap) hi *= 2;
    for (int k = 0; k < 100; ++k) {
        double mid = 0.5 * (lo + hi);
        if (ra * audit_epsilon(mid / Ka) + rb * audit_epsilon(mid / Kb) < overlap) lo = mid; else hi = mid;
    }
    double f = 0.5 * (lo + hi);
    return audit_spring(audit_epsilon(f / Ka), Ka, ra) + audit_spring(audit_epsilon(f / Kb), Kb, rb);
}
}  // namespace

void bpo_audit(const bp_world_config *config, const bp_bubble *bubbles, uint32_t n, const bp_segment *segments, uint32_t ns,
               const bp_body *bodies, uint32_t nb, const float *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt,
               bp_audit *out) {
    bpo::World<float> w;
    bpo::load(w, config, bubbles, n, segments, ns, bodies, nb, vertices, nv, triangles, nt, nullptr, 0, nullptr, 0, nullptr, 0,
              nullptr, 0, nullptr, 0);
    double px = 0, py = 0, ang = 0, kin = 0, pot = 0;
    uint32_t contacts = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const auto &a = w.b[i];
        if (!w.is_static(a)) {
            double m = a.mass, vx = a.vel.x, vy = a.vel.y;
            px += m * vx; py += m * vy;
            ang += m * ((double)a.pos.x * vy - (double)a.pos.y * vx);
            kin += 0.5 * m * (vx * vx + vy * vy);
        }
        if (!bpo::finite2(a.pos)) continue;
        for (uint32_t j = i + 1; j < n; ++j) {
            const auto &o = w.b[j];
            if (!bpo::finite2(o.pos) || !bpo::filter_accepts(a.cat, a.mask, o.cat, o.mask)) continue;
            double dx = (double)o.pos.x - (double)a.pos.x, dy = (double)o.pos.y - (double)a.pos.y;
            pot += audit_pair(a.Kc, a.radius, o.Kc, o.radius, std::sqrt(dx * dx + dy * dy), contacts);
        }
        for (uint32_t j = 0; j < ns; ++j) {
            const auto &s = w.s[j];
            if (!bpo::filter_accepts(a.cat, a.mask, s.cat, s.mask)) continue;
            bpo::V2<float> normal; float d;
            if (!bpo::wall_region(s, a.pos, normal, d)) continue;
            pot += audit_wall(a.Kc, a.radius, d, contacts);
        }
        for (uint32_t k = 0; k < nb; ++k) {
            const auto &body = w.bodies[k];
            bool obstacle = body.isStatic ? bpo::rigid_is_round(body) : (body.kinematic != 0);
            if (!obstacle || !bpo::filter_accepts(a.cat, a.mask, body.cat, body.mask)) continue;
            auto ob = bpo::body_obstacle(body, 1.0f, w.verts, w.tris, a.pos);
            pot += audit_wall(a.Kc, a.radius, ob.distanceValue, contacts);
        }
    }
    out->momentum[0] = px; out->momentum[1] = py; out->angular_momentum = ang; out->kinetic_energy = kin;
    out->contact_energy = pot; out->contacts = contacts; out->bubbles = n;
}

// End-of-step contact records of a world state (the records NativeMetalSolver.swift:846-871 builds
// after the region pass and returns as NativeSolverContact, :1041-1060):
//   nativeBuildContacts (Geometry:683-728) per candidate, with nativePolygonEdgeFeature (:674-681);
//   kept iff penetration > -separation tolerance (Contact:146-153), compacted in candidate order
//   (Contact:155-160), the candidate order being every bubble's pair candidates, then every
//   bubble's segment candidates (Geometry:625-651);
//   impulse = the law's force at the distance the record implies times dt^2 (Force:261-280);
//   id = the FNV-1a steps of NativeMetalSolver.swift:1047-1052 over (kind, first, second).
// The candidates are read from the rows of the last grid build (row_offsets[n + 1], entries
// (kind << 30) | index): bubble i's pair candidates are its pair entries above i, its segment
// candidates its segment entries. Returns the number of kept records, writing up to `capacity`.
uint32_t bpo_contacts(const bp_world_config *config, const bp_bubble *bubbles, uint32_t n, const bp_segment *segments,
                      uint32_t ns, const uint32_t *row_offsets, const uint32_t *entries, bp_contact *out, uint32_t capacity) {
    using P = bpo::V2<float>;
    const float sepTol = config->separation_tolerance, dt = config->time_step;
    uint32_t kept = 0;
    for (uint32_t kind = 0; kind < 2; ++kind) {
        for (uint32_t i = 0; i < n; ++i) {
            const bp_bubble &a = bubbles[i];
            const P pa = bpo::v2(a.center[0], a.center[1]);
            const float ra = a.radius;
            for (uint32_t e = row_offsets[i]; e < row_offsets[i + 1]; ++e) {
                const uint32_t entryKind = entries[e] >> 30, j = entries[e] & ((1u << 30) - 1u);
                if (entryKind != kind || (kind == 0 && j <= i)) continue;
                P normal, point;
                float penetration, allowedSide = 0.0f, force;
                uint32_t feature = 0;
                int64_t second;
                if (kind == 0) {
                    const bp_bubble &b = bubbles[j];
                    const P delta = bpo::v2(b.center[0], b.center[1]) - pa;
                    const float dist = bpo::length(delta);
                    normal = dist > 0.0f ? delta / dist : bpo::v2(1.0f, 0.0f);
                    point = pa + normal * ra;
                    penetration = ra + b.radius - dist;
                    second = b.id;
                } else {
                    if (j >= ns) continue;
                    const bp_segment &sg = segments[j];
                    const P sa = bpo::v2(sg.a[0], sg.a[1]), sb = bpo::v2(sg.b[0], sg.b[1]);
                    const P edge = sb - sa;
                    const float lengthSquared = bpo::dot(edge, edge);
                    const float t = lengthSquared > 0.0f ? bpo::vclamp(bpo::dot(pa - sa, edge) / lengthSquared, 0.0f, 1.0f) : 0.0f;
                    point = sa + t * edge;
                    P delta = pa - point;
                    float dist = bpo::length(delta);
                    const float edgeLength = bpo::length(edge);
                    const P base = edgeLength > 0.0f ? bpo::v2(-edge.y, edge.x) / edgeLength : bpo::v2(0.0f, 1.0f);
                    const float sign = (((uint64_t)a.id ^ (uint64_t)sg.id) & 1u) == 0u ? 1.0f : -1.0f;
                    const bool sided = (sg.flags & BP_SEGMENT_ALLOWED_SIDE) != 0u;
                    normal = sided ? base * sg.allowed_side : (dist > 0.0f ? delta / dist : base * sign);
                    allowedSide = sided ? sg.allowed_side : 0.0f;
                    penetration = ra - dist;
                    if (sg.flags & BP_SEGMENT_POLYGON_EDGE) {
                        const P outward = base * sg.outward_sign;
                        const float along = bpo::dot(pa - sa, sb - sa);
                        const bool convexStart = (sg.flags & BP_SEGMENT_CONVEX_START) != 0u;
                        const P previous = bpo::v2(sg.previous_start[0], sg.previous_start[1]);
                        if (convexStart && along <= 0.0f && bpo::past_segment_end(pa, previous, sa)) feature = 2;
                        else feature = bpo::dot(pa - sa, outward) >= 0.0f && along > 0.0f && !bpo::past_segment_end(pa, sa, sb) ? 1u : 0u;
                        if (feature == 1) {
                            normal = outward;
                            const float depth = bpo::dot(pa - sa, outward);
                            point = pa - outward * depth;
                            penetration = ra - depth;
                        } else if (feature == 2) {
                            point = sa;
                            delta = pa - point;
                            dist = bpo::length(delta);
                            normal = dist > 0.0f ? delta / dist : outward;
                            penetration = ra - dist;
                        } else {
                            normal = outward;
                            point = sa;
                            penetration = -FLT_MAX;
                        }
                    }
                    second = sg.id;
                }
                if (!(penetration > -sepTol)) continue;
                if (kind == 0) {
                    const bp_bubble &b = bubbles[j];
                    force = bpo::pair_contact_law<float>(a.contact_stiffness, ra, b.contact_stiffness, b.radius,
                                                         ra + b.radius - penetration).force;
                } else {
                    force = bpo::wall_contact_law<float>(a.contact_stiffness, ra, ra - penetration).force;
                }
                if (kept < capacity) {
                    bp_contact &c = out[kept];
                    uint64_t h = 14695981039346656037ull;
                    h = (h ^ (uint64_t)kind) * 1099511628211ull;
                    h = (h ^ (uint64_t)a.id) * 1099511628211ull;
                    h = (h ^ (uint64_t)second) * 1099511628211ull;
                    c.id = h;
                    c.first = a.id;
                    c.second = second;
                    c.kind = kind;
                    c.feature = feature;
                    c.normal[0] = normal.x; c.normal[1] = normal.y;
                    c.point[0] = point.x; c.point[1] = point.y;
                    c.penetration = penetration;
                    c.allowed_side = allowedSide;
                    c.impulse = force * dt * dt;
                    c.index_b = j;
                }
                ++kept;
            }
        }
    }
    return kept;
}

// End-of-step bubble-body contact records (nativeBubbleBodyContacts, NativeForceKernels.metal-source:
// 333-358, read back by NativeRigidPolygonState.swift:390-402): per candidate (body, bubble) of the
// body rows, the obstacle at the body's end-of-step pose (fraction 1) and the bubble's final centre;
// kept iff distance < radius and the record's normal is not zero. Record: normal = -obstacle normal
// (from the bubble to the body), point = centre - obstacle normal * distance, impulse = the wall
// law's force at that distance times dt^2; penetration carries radius - distance. The candidates
// are the body entries of the bubbles' rows of the last grid build (entries (2 << 30) | body), taken
// body by body with the bubbles ascending, the order of the reference's body rows.
uint32_t bpo_body_contacts(const bp_world_config *config, const bp_bubble *bubbles, uint32_t n, const bp_body *bodies,
                           uint32_t nb, const float *vertices, uint32_t nv, const uint32_t *triangles, uint32_t nt,
                           const uint32_t *row_offsets, const uint32_t *entries, bp_contact *out, uint32_t capacity) {
    bpo::World<float> w;
    bpo::load(w, config, bubbles, n, nullptr, 0, bodies, nb, vertices, nv, triangles, nt, nullptr, 0, nullptr, 0, nullptr, 0,
              nullptr, 0, nullptr, 0);
    const float dt = config->time_step;
    uint32_t kept = 0;
    for (uint32_t k = 0; k < nb; ++k) {
        for (uint32_t i = 0; i < n; ++i) {
            bool candidate = false;
            for (uint32_t e = row_offsets[i]; e < row_offsets[i + 1]; ++e)
                if ((entries[e] >> 30) == 2u && (entries[e] & ((1u << 30) - 1u)) == k) candidate = true;
            if (!candidate) continue;
            const auto &a = w.b[i];
            const auto ob = bpo::body_obstacle(w.bodies[k], 1.0f, w.verts, w.tris, a.pos);
            if (!(ob.distanceValue < a.radius)) continue;
            const bpo::V2<float> normal = -ob.normal;
            if (normal.x == 0.0f && normal.y == 0.0f) continue;
            if (kept < capacity) {
                const bpo::V2<float> point = a.pos - ob.normal * ob.distanceValue;
                const float force = bpo::wall_contact_law<float>(a.Kc, a.radius, ob.distanceValue).force;
                bp_contact &c = out[kept];
                c = bp_contact{};
                uint64_t h = 14695981039346656037ull;
                h = (h ^ (uint64_t)BP_CONTACT_BODY) * 1099511628211ull;
                h = (h ^ (uint64_t)a.id) * 1099511628211ull;
                h = (h ^ (uint64_t)bodies[k].id) * 1099511628211ull;
                c.id = h;
                c.first = a.id;
                c.second = (int64_t)bodies[k].id;
                c.kind = BP_CONTACT_BODY;
                c.normal[0] = normal.x; c.normal[1] = normal.y;
                c.point[0] = point.x; c.point[1] = point.y;
                c.penetration = a.radius - ob.distanceValue;
                c.impulse = force * dt * dt;
                c.index_b = k;
            }
            ++kept;
        }
    }
    return kept;
}

// Candidate rows of one grid build with a given cell width and margins (the kernel
// nativeEmitCandidates, Geometry:300-340, and its hashed twin, Force:1010-1035).
// rows[n * limit * 2] receives (kind, other index) per slot, row_counts[2n] the bubble and
// segment counts, overflow[2] the dropped total and the widest demand.
void bpo_candidate_rows(const bp_bubble *bubbles, const float *margins, uint32_t n, const bp_segment *segments, uint32_t ns,
                        float cell_width, float separation_tolerance, uint32_t limit, uint32_t hashed, uint32_t buckets,
                        uint32_t *rows, uint32_t *row_counts, uint32_t *overflow) {
    bpo::World<float> w;
    bp_world_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    cfg.time_step = 1.0f / 60.0f;
    cfg.separation_tolerance = separation_tolerance;
    cfg.candidate_limit = limit;
    cfg.grid_bucket_count = buckets;
    bpo::load(w, &cfg, bubbles, n, segments, ns, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0);
    w.limit = limit;
    w.cfg.hashedGrid = hashed;
    w.margin.assign(margins, margins + n);
    w.cellWidth = cell_width;
    std::vector<std::vector<bpo::Candidate>> r(n);
    std::vector<uint32_t> counts(2 * n, 0);
    if (hashed) w.emit_candidates_hashed(r, counts);
    else w.emit_candidates_sorted(r, counts);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < r[i].size(); ++k) {
            rows[2 * ((size_t)i * limit + k)] = r[i][k].kind;
            rows[2 * ((size_t)i * limit + k) + 1] = r[i][k].ib;
        }
    std::memcpy(row_counts, counts.data(), counts.size() * sizeof(uint32_t));
    overflow[0] = w.overflow[0];
    overflow[1] = w.overflow[1];
}

// Planted mistake for the tests that show a pin catches it (0 switches every plant off).
// Contour and render preparation of a world's published state (draw-only):
// nativeAccumulateContactPressure (PostSolve:280-337) with nativeContourRayLimit (:268-277),
// 13 passes of nativeGenerateContours (:341-355; NativeMetalContourSmoothing) and
// nativePrepareRenderData (:382-441). A bubble's pair chords come from its row's pair entries
// (the end-of-step contact records of the last grid build) that overlap it. `stride` is the
// contour stride of the world (at least 64); points[2 * (i * stride + s)] receives the samples and
// shapes[i] the bubble's count, largest deficit, effective radius and bounds.
// Planted mistake 6: the pair chord at the radius instead of r (1 - eps).
void bpo_contours(const bp_contour_config *config, const bp_bubble *bubbles, uint32_t n, const bp_segment *segments, uint32_t ns,
                  const uint32_t *row_offsets, const uint32_t *entries, uint32_t stride, float *points, bp_bubble_contour *shapes) {
    using P = bpo::V2<float>;
    const float maxEdge = config->max_segment_length, tensionParam = config->surface_tension,
                clearance = config->position_tolerance;
    const float pi = 3.14159265358979323846f;
    auto ray = [](P center, P direction, P a, P b, float radius, float clear) {   // PostSolve:268-277
        const P edge = b - a, offset = a - center;
        const float denominator = bpo::cross(direction, edge);
        if (std::fabs(denominator) < 1e-6f) return radius;
        const float distance = bpo::cross(offset, edge) / denominator;
        const float fraction = bpo::cross(offset, direction) / denominator;
        return distance >= 0 && fraction >= 0 && fraction <= 1 ? bpo::vmin(radius, bpo::vmax(0.0f, distance - clear)) : radius;
    };
    std::vector<float> x, y, next;
    for (uint32_t i = 0; i < n; ++i) {
        const bp_bubble &bubble = bubbles[i];
        const P pos = bpo::v2(bubble.center[0], bubble.center[1]);
        const float r = bubble.radius;
        const uint32_t count = std::min(stride, std::max(64u, (uint32_t)std::ceil((r / maxEdge) * (4 * pi))));
        x.assign(count, 0.0f);
        y.assign(count, 0.0f);
        for (uint32_t sample = 0; sample < count; ++sample) {
            const float angle = 2 * pi * float(sample) / float(count);
            const P direction = bpo::v2(bpm::cosf_det(angle), bpm::sinf_det(angle));
            float limit = r;
            for (uint32_t j = 0; j < ns; ++j) {
                const bp_segment &sg = segments[j];
                if (!((bubble.category & sg.mask) != 0 && (sg.category & bubble.mask) != 0)) continue;
                limit = ray(pos, direction, bpo::v2(sg.a[0], sg.a[1]), bpo::v2(sg.b[0], sg.b[1]), limit, clearance);
            }
            for (uint32_t e = row_offsets[i]; e < row_offsets[i + 1]; ++e) {
                if ((entries[e] >> 30) != 0) continue;
                const uint32_t j = entries[e] & ((1u << 30) - 1u);
                const bp_bubble &other = bubbles[j];
                const P delta = bpo::v2(other.center[0], other.center[1]) - pos;
                const float distance = bpo::length(delta);
                const float otherRadius = other.radius;
                if (distance <= FLT_EPSILON || distance >= r + otherRadius) continue;
                const P normal = delta / distance;
                const bool lowIsSelf = i < j;
                const bp_bubble &first = lowIsSelf ? bubble : other, &second = lowIsSelf ? other : bubble;
                const bpo::Law<float> law = bpo::pair_contact_law<float>(first.contact_stiffness, first.radius,
                                                                         second.contact_stiffness, second.radius, distance);
                const float epsilon = lowIsSelf ? law.epsA : law.epsB;
                float planeDistance = bpo::vclamp(r * (1 - epsilon), 0.0f, r);
                if (bpo::g_plant == 6) planeDistance = r;
                const float halfLength = std::sqrt(bpo::vmax(0.0f, r * r - planeDistance * planeDistance));
                const P tangent = bpo::v2(-normal.y, normal.x);
                const P point = pos + normal * planeDistance;
                limit = ray(pos, direction, point - tangent * halfLength, point + tangent * halfLength, limit, clearance);
            }
            const float deficit = bpo::vclamp(1 - limit / r, 0.0f, 1.0f);
            x[sample] = deficit;
            y[sample] = deficit;
        }
        const float tension = bpo::vmin(0.8f, tensionParam / (tensionParam + 10));
        next.assign(count, 0.0f);
        for (int pass = 0; pass < 13; ++pass) {
            for (uint32_t sample = 0; sample < count; ++sample) {
                const float neighbor = (x[(sample + count - 1) % count] + x[(sample + 1) % count]) * 0.5f;
                const float mixed = x[sample] + (neighbor - x[sample]) * tension;
                next[sample] = bpo::vclamp(bpo::vmax(y[sample], mixed), 0.0f, 1.0f);
            }
            x.swap(next);
        }
        float maximum = 0;
        uint32_t anchor = 0;
        for (uint32_t sample = 0; sample < count; ++sample)
            if (x[sample] > maximum) { maximum = x[sample]; anchor = sample; }
        const float slope = maxEdge / (2 * r);
        float previous = maximum;
        for (uint32_t step = 1; step < count; ++step) {
            const uint32_t sample = (anchor + step) % count;
            const float value = bpo::vmax(x[sample], previous - slope);
            x[sample] = value;
            previous = value;
        }
        previous = maximum;
        for (uint32_t step = 1; step < count; ++step) {
            const uint32_t sample = (anchor + count - step) % count;
            const float value = bpo::vmax(x[sample], previous - slope);
            x[sample] = value;
            previous = value;
        }
        float effectiveRadius = r;
        P lo = pos, hi = pos;
        for (uint32_t sample = 0; sample < count; ++sample) {
            const float angle = 2 * pi * float(sample) / float(count);
            const P direction = bpo::v2(bpm::cosf_det(angle), bpm::sinf_det(angle));
            const P point = pos + direction * r * (1 - x[sample]);
            points[2 * ((size_t)i * stride + sample)] = point.x;
            points[2 * ((size_t)i * stride + sample) + 1] = point.y;
            lo = bpo::vmin2(lo, point);
            hi = bpo::vmax2(hi, point);
            effectiveRadius = bpo::vmin(effectiveRadius, bpo::length(point - pos));
            maximum = bpo::vmax(maximum, x[sample]);
        }
        bp_bubble_contour &out = shapes[i];
        out.count = count;
        out.maximum = maximum;
        out.effective_radius = effectiveRadius;
        out.reserved = 0;
        out.bounds[0] = lo.x;
        out.bounds[1] = lo.y;
        out.bounds[2] = hi.x;
        out.bounds[3] = hi.y;
    }
}

void bpo_set_plant(int code) { bpo::g_plant = code; }
void bpo_set_threads(int threads) { bpo::g_threads = threads < 1 ? 1 : threads; }

// Contact-law entry points (the harness kernels testEpsilon / testWallLaw / testPairLaw of
// Tests/BubblePhysicsNativeMetalTests/NativeMetalContactLawTests.swift:15-44).
void bpo_epsilon_f32(const float *q, float *out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) out[i] = bpo::epsilon_from_force_ratio<float>(q[i]);
}
void bpo_wall_law_f32(const float *in4, float *out4, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        bpo::Law<float> l = bpo::wall_contact_law<float>(in4[4 * i], in4[4 * i + 1], in4[4 * i + 2]);
        out4[4 * i] = l.force; out4[4 * i + 1] = l.epsA; out4[4 * i + 2] = l.stiffness; out4[4 * i + 3] = l.epsB;
    }
}
void bpo_pair_law_f32(const float *in4, const float *distances, float *out4, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        bpo::Law<float> l = bpo::pair_contact_law<float>(in4[4 * i], in4[4 * i + 1], in4[4 * i + 2], in4[4 * i + 3], distances[i]);
        out4[4 * i] = l.force; out4[4 * i + 1] = l.epsA; out4[4 * i + 2] = l.epsB; out4[4 * i + 3] = l.stiffness;
    }
}
void bpo_half_impulse_factor_f32(const float *gamma, const float *half, float *out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) out[i] = bpo::World<float>::half_impulse_factor(gamma[i], half[i]);
}
void bpo_detmath_f32(int which, const float *x, float *out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        switch (which) {
            case 0: out[i] = bpm::sinf_det(x[i]); break;
            case 1: out[i] = bpm::cosf_det(x[i]); break;
            case 2: out[i] = bpm::expf_det(x[i]); break;
            default: out[i] = bpm::logf_det(x[i]); break;
        }
    }
}
uint32_t bpo_abi_version(void) { return BP_ABI_VERSION; }

// Non-contact acceleration a0 of every bubble at its centre (nativeAccelerationWithoutContacts,
// Force:450-494), FP32, written to out[2n].
void bpo_noncontact_acceleration(const bp_world_config *config, const bp_bubble *bubbles, uint32_t n, const bp_motor *motors,
                                 uint32_t nm, const bp_band *bands, uint32_t nbands, const bp_field *fields, uint32_t nf,
                                 const float *field_vertices, uint32_t nfv, const int64_t *field_ids, uint32_t nfi, float *out) {
    bpo::World<float> w;
    bpo::load(w, config, bubbles, n, nullptr, 0, nullptr, 0, nullptr, 0, nullptr, 0, motors, nm, bands, nbands, fields, nf,
              field_vertices, nfv, field_ids, nfi);
    for (uint32_t i = 0; i < n; ++i) {
        bpo::V2<float> a = w.non_contact_acceleration(w.b[i], w.b[i].pos);
        out[2 * i] = a.x; out[2 * i + 1] = a.y;
    }
}

}  // extern "C"
