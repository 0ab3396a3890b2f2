// bp_device.h -- device-side records and per-bubble physics of the world step (gfx950).
//
// The functions here are the per-object arithmetic of the BubblePhysics 4.0.0 step, in the
// order the reference evaluates it, so that the FP32 results round identically to the CPU
// oracle. Citations are to the reference repository:
//   Force    = Sources/BubblePhysicsNativeMetal/Shaders/NativeForceKernels.metal-source
//   Geometry = Sources/BubblePhysicsNativeMetal/Shaders/NativeGeometryKernels.metal-source
// This translation unit is compiled with -ffp-contract=off: every + - * / sqrt rounds once,
// and the only fused multiply-add is the explicit one of past_segment_end.
#ifndef BP_DEVICE_H
#define BP_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bp_detmath.h"

namespace bpd {

#define BPD __device__ __forceinline__

// Force field (NativeForceField, NativeForceKernels.metal-source:11-17); vertex and filter
// offsets are relative to the world's field-vertex and field-id arrays.
struct DevField {
    float4 pose;              // position.xy, rotation, falloff
    float4 strength;          // directional force.xy, radial/vortex strength, radius
    uint4 shape;              // kind, vertex offset, vertex count (0 = circle), 0
    uint4 filter;             // id offset, id count, filter present, 0
};

// ---------------------------------------------------------------------------------------
// Arena: structure-of-arrays device storage of every world of a batch.
// Per-bubble dynamic state exists twice: side `pub` of a world is its published state,
// the other side is written by the step and becomes published only when the step succeeds.
// ---------------------------------------------------------------------------------------
struct Arena {
    // Bubble state, two sides.
    float2 *pos[2];
    float2 *start[2];
    float2 *vel[2];
    float *rot[2];
    float *omega[2];
    uint32_t *region[2];
    // Bubble properties.
    int64_t *id;
    float *mass, *invm, *radius, *Kc, *gamma;
    uint32_t *flags;          // BP_BUBBLE_* bits
    uint32_t *cat, *mask;
    int32_t *motor, *band;    // per-bubble record index in the world's motor/band arrays, -1 none
    // Step scratch, per bubble.
    float2 *acc, *a0, *ref;
    float *margin;
    uint32_t *rowStart, *rowLen, *grid;
    float *omega2, *bound;    // per-bubble reduction inputs (debug / cooperative path)
    // What neighbours read, packed per bubble for the global-state passes (one 16-byte load each):
    // nbrA = (x, y, r, K_c) of the working state, nbrB = (effective velocity, |a|, 0), and the
    // effective inverse mass of the step.
    float4 *nbrA, *nbrB;
    float *invmEff;
    // The same bubbles in the order of the last counting sort by cell (slot order of `grid`), so a
    // scan of a cell reads consecutive records: (x, y, r, margin) and (index, category, mask,
    // flags: 1 = finite step start, 2 = removed by the region, 4 = movable).
    float4 *cellA;
    uint4 *cellB;
    uint32_t *bodyHits;       // per body: bubbles in its candidate row at the last grid build
    uint32_t *bodyCut;        // per body: index of the k-th bubble of its row (a row cut keeps those below)
    // Adjacency entries: (kind << 30) | neighbour, CSR per world.
    uint32_t *adj;
    // Per adjacency slot, the contact geometry the last force pass computed for that row entry
    // (normal, distance and, for a body, its surface speed bound; for a segment whether the centre
    // is in its region), read back by the distance pass that follows at the same state.
    float4 *adjGeo;
    // Segments.
    float4 *segAB;            // a.xy, b.xy
    float4 *segRegion;        // previous_start.xy, outward sign, allowed side
    uint32_t *segFlags, *segCat, *segMask;
    int64_t *segId, *segOwner;
    // Bodies: pose (x, y, angle, previous angle) and previous centre, two sides.
    float4 *bodyPose[2];
    float2 *bodyPrev[2];
    float4 *bodyMotion;       // vx, vy, omega, damping (0)
    float4 *bodyShape;        // bounding radius, core radius, half length, unused
    float2 *bodyLimits;       // max linear, max angular speed
    uint4 *bodyRanges;        // vertex offset/count, triangle offset/count (world-relative)
    uint4 *bodyKind;          // shape, flags, category, mask
    uint64_t *bodyId;
    float2 *verts;
    uint4 *tris;
    float4 *motors;           // force.xy, torque, inertia
    float4 *bands;            // target.xy, stiffness, maximum force
    DevField *fields;
    float2 *fieldVerts;       // local vertices of polygon field ranges
    uint64_t *fieldIds;       // filter ids, ascending within each field's range
    float *history;           // 1024 h values per world
    // Cooperative (multi-workgroup) worlds only.
    uint32_t *coopCells;      // hashed-grid buckets of each cooperative world
    uint32_t *coopScratch;    // workgroup partials and counters of each cooperative world
    uint2 *bvhChild;          // body LBVH: children of internal node i (bit 31 = leaf), 2 slots per body
    float4 *bvhBox;           // boxes: internal nodes at [0, nb-1), leaves at [nb, 2 nb)
    uint32_t *bvhLeaf;        // body index of each sorted leaf
};

// Layout of a cooperative world's scratch (u32 words) for a grid of up to kCoopMaxGroups workgroups.
constexpr uint32_t kCoopMaxGroups = 2048;
constexpr uint32_t kCoopCounters = 32;
constexpr uint32_t kCoopScratchWords = 8 * kCoopMaxGroups + kCoopCounters;
enum CoopCounter : uint32_t {
    kCoDropped = 0, kCoMaxDemand, kCoPairs, kCoOverflowRows, kCoSafety, kCoRemovals, kCoWraps, kCoViolator,
    kCoBvhBuilt, kCoBodySafety, kCoEntries, kCoLongRows, kCoCutDropped, kCoCutWidest, kCoCutEntries
};


// Per-world constants and the published side. Written by the host, read by the kernel;
// `pub` and `failed` are updated by the kernel at the end of each step.
struct WorldDesc {
    uint32_t bubbleOffset, bubbleCount;
    uint32_t segmentOffset, segmentCount;
    uint32_t bodyOffset, bodyCount;
    uint32_t vertexOffset, triangleOffset;
    uint32_t vertexCount, triangleCount;
    uint32_t motorOffset, motorCount;
    uint32_t bandOffset, bandCount;
    uint32_t fieldOffset, fieldCount, fieldVertexOffset, fieldIdOffset;
    uint32_t coopCellOffset, coopBuckets, coopScratchOffset;   // cooperative worlds
    uint32_t adjOffset, adjCapacity;
    uint32_t historyOffset;
    uint32_t pub;
    float dt, contactTol, sepTol;
    float region[4];
    float globalForce[2];
    uint32_t applyRegion, limit, split, minimumMarginOnly, buckets;
    uint32_t assertEverySubstep, injectionEnabled, injectionAfter;
    int64_t injectionObject;
    float injectionPosition[2];
    uint32_t failed;          // kWorldFailed: failed a step, sticky, not advanced again; kWorldHeld: a
                              // step removed bubbles from the region, the world waits until they are
                              // deleted (kWorldRunnable = 0 otherwise)
    uint32_t exportAdjacency; // 1 = leave the rows of the last grid build in rowStart/rowLen/adj
    uint32_t bubbleCapacity;  // bubble slots of the world's range (transactions never grow past it)
    // Serving mirror in pinned host memory (device addresses; null outside a server): after every
    // published step the kernel writes the world's records (bp_bubble) and its bodies' poses there,
    // and a replacement transaction reads the world's new records and their motor and band indices
    // (2 x mirrorCapacity) from there.
    void *mirror;
    float4 *bodyMirror;
    int32_t *mirrorIdx;
    uint32_t mirrorCapacity;
    uint32_t reserved0;
};

enum : uint32_t { kWorldRunnable = 0u, kWorldFailed = 1u, kWorldHeld = 2u };

// One copy of a staged blob: `bytes` (a multiple of 4) from blob offset `src` to device address `dst`
// (with no blob, `src` is itself an address).
struct CopyOp {
    uint64_t src;
    void *dst;
    uint32_t bytes, pad;
};

// One world of a many-world gather (its records out of the arena) or scatter (into it).
struct WorldMove {
    const WorldDesc *desc;      // gather: the world's descriptor on the device
    void *records;              // bp_bubble records; gather: destination; scatter: source
    const int32_t *indices;     // scatter: each record's motor index, then each record's band index
    uint32_t offset, count;     // scatter: first slot of the world's range and records in it
};

// A serving transaction on one world, applied on the device before the world's step
// (bp_request of include/bubblephysics.h; `buckets` is the grid bucket count for the bubble
// count after the transaction, computed by the host, which tracks the count).
struct TxRecord {
    uint32_t world, kind, index, buckets;
    float impulse[2], direction[2];
};
// TxRecord.kind of a replacement: the world's records become the `index` records of its mirror
// (an add, merge or split the host staged on the mirror of the published state).
constexpr uint32_t kTxReplace = 100u;

// A world held after a step that removed bubbles from the region (kWorldHeld): its removed
// bubbles are deleted on the device before it steps again. `buckets` is the grid bucket count for
// the bubble count that remains; the ids of the removed bubbles go to removedIds[removedOffset...].
struct DropRecord {
    uint32_t world, buckets, removedOffset, reserved;
};

// Per-world result of the last step (bp_frame_info fields).
struct WorldStatus {
    int32_t status;
    uint32_t substeps, rebuilds, candidatePairs, candidateOverflow, maxCandidates;
    uint32_t removals, wraps;
    float time[4];
    uint64_t limiterObject;
    uint32_t limiterTerm;
    uint32_t violationKind;
    int64_t violationSegment;
    uint64_t violationObjects[2];
    float violationDetail;
    uint32_t adjacencyEntries;
    uint32_t groupWidth;
    uint32_t executed;        // 1 when the kernel ran this world's step
    uint32_t ticks;           // device wall-clock ticks (100 MHz) the workgroup spent on the step
    uint32_t seq;             // set to 1 last, with a system-scope release, when the launch asks for it
    uint32_t txResult;        // BP_REJECT_* of a serving transaction applied on the device (0 = applied)
};

// One end-of-step contact record (bp_contact without its id, which the host hashes from the key).
struct ContactRecord {
    int64_t first, second;    // bubble A id; bubble B id or segment id
    uint32_t kind, feature;   // 0 pair, 1 segment; polygon edge feature
    float normal[2], point[2];
    float penetration, allowedSide, impulse;
    uint32_t indexB;          // bubble B index or segment index
};

// Marks a world's result record complete (the thread that wrote the record calls it). With
// `notify`, the release store at system scope makes the record visible to the host while the
// launch is still running (the record then lives in pinned host memory, polled by the server).
BPD void post_status(WorldStatus *st, uint32_t notify) {
    if (notify) __hip_atomic_store(&st->seq, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------------------
// float2 helpers (componentwise, one rounding per operation).
// ---------------------------------------------------------------------------------------
BPD float2 f2(float x, float y) { return make_float2(x, y); }
BPD float2 add(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
BPD float2 sub(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
BPD float2 neg(float2 a) { return f2(-a.x, -a.y); }
BPD float2 mul(float2 a, float s) { return f2(a.x * s, a.y * s); }
BPD float2 mul(float s, float2 a) { return f2(s * a.x, s * a.y); }
BPD float2 dvs(float2 a, float s) { return f2(a.x / s, a.y / s); }
BPD float2 subs(float2 a, float s) { return f2(a.x - s, a.y - s); }
BPD float2 adds(float2 a, float s) { return f2(a.x + s, a.y + s); }
BPD float dot2(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
BPD float len2(float2 a) { return __builtin_sqrtf(a.x * a.x + a.y * a.y); }
BPD float dist2(float2 a, float2 b) { return len2(sub(a, b)); }
BPD float cross2(float2 a, float2 b) { return a.x * b.y - a.y * b.x; }
BPD float2 min2(float2 a, float2 b) { return f2(fminf(a.x, b.x), fminf(a.y, b.y)); }
BPD float2 max2(float2 a, float2 b) { return f2(fmaxf(a.x, b.x), fmaxf(a.y, b.y)); }
BPD bool finite2(float2 a) { return isfinite(a.x) && isfinite(a.y); }
BPD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
BPD float inff() { return __builtin_inff(); }

constexpr float kFloatMin = 1.17549435e-38f;   // FLT_MIN
constexpr uint32_t kFrameLimit = 1024;         // NativeMetalSubstep.frameLimit
constexpr uint32_t kMaxBuckets = 8192;         // hashed-grid buckets held in LDS
constexpr uint32_t kKindShift = 30;
constexpr uint32_t kIndexMask = (1u << 30) - 1;

// Bubble flag bits (bp_bubble.flags) and the region result.
constexpr uint32_t kEnabled = 1u, kStatic = 2u, kWrap = 4u, kForceDriven = 8u;

BPD bool filter_accepts(uint32_t ac, uint32_t am, uint32_t bc, uint32_t bm) {   // Geometry:130-132
    return (ac & bm) != 0 && (bc & am) != 0;
}

// ---------------------------------------------------------------------------------------
// Contact law, Force:50-102.
// ---------------------------------------------------------------------------------------
struct Law { float force, epsA, epsB, stiffness; };

BPD float epsilon_from_force_ratio(float q) {
    if (!(q > 0.0f)) return 0.0f;
    return 2.0f * q / ((2.0f * q + 1.0f) + __builtin_sqrtf(4.0f * q + 1.0f));
}
BPD float rest_from_force_ratio(float q) {
    return 2.0f / (1.0f + __builtin_sqrtf(fmaxf(4.0f * q, 0.0f) + 1.0f));
}
BPD Law wall_law(float K, float r, float distanceValue) {
    Law law = {0.0f, 0.0f, 0.0f, 0.0f};
    float d = fmaxf(distanceValue, r * 1e-6f);
    if (!(K > 0.0f) || !(r > 0.0f) || !(d < r)) return law;
    float rest = d / r;
    law.epsA = (r - d) / r;
    law.force = K * (r - d) * r / (d * d);
    law.stiffness = K * (1.0f + law.epsA) / (r * rest * rest * rest);
    return law;
}
BPD Law pair_law(float Ka, float ra, float Kb, float rb, float distanceValue) {
    Law law = {0.0f, 0.0f, 0.0f, 0.0f};
    float R = ra + rb;
    float d = fmaxf(distanceValue, R * 1e-6f);
    float overlap = R - d;
    if (!(Ka > 0.0f) || !(Kb > 0.0f) || !(overlap > 0.0f)) return law;
    float lowest = fminf(Ka, Kb) * overlap * R / (d * d);
    float force = lowest;
    if (Ka != Kb) {
        for (int step = 0; step < 12; ++step) {
            float qa = force / Ka, qb = force / Kb;
            float ea = epsilon_from_force_ratio(qa), eb = epsilon_from_force_ratio(qb);
            float sa = rest_from_force_ratio(qa), sb = rest_from_force_ratio(qb);
            float residual = ra * ea + rb * eb - overlap;
            float slope = ra * sa * sa * sa / ((1.0f + ea) * Ka) + rb * sb * sb * sb / ((1.0f + eb) * Kb);
            if (!(slope > kFloatMin)) break;
            force = fmaxf(force - residual / slope, lowest);
        }
    }
    // Equal stiffnesses give equal force ratios, so b's strain, rest factor and compliance
    // denominator are a's, bit for bit: computed once.
    const float qa = force / Ka;
    const float ea = epsilon_from_force_ratio(qa), sa = rest_from_force_ratio(qa);
    const float da = (1.0f + ea) * Ka;
    float eb = ea, sb = sa, db = da;
    if (Ka != Kb) {
        const float qb = force / Kb;
        eb = epsilon_from_force_ratio(qb); sb = rest_from_force_ratio(qb);
        db = (1.0f + eb) * Kb;
    }
    float compliance = ra * sa * sa * sa / da + rb * sb * sb * sb / db;
    law.force = force; law.epsA = ea; law.epsB = eb;
    law.stiffness = compliance > kFloatMin ? 1.0f / compliance : 0.0f;
    return law;
}

// pair_law for two bubbles of equal stiffness (Ka == Kb), without branches: the same operations
// in the same order as pair_law's closed form (no Newton steps are taken for equal stiffnesses),
// so the result is the same bit for bit; the zero law is selected at the end.
BPD Law pair_law_same(float K, float ra, float rb, float distanceValue) {
    const float R = ra + rb;
    const float d = fmaxf(distanceValue, R * 1e-6f);
    const float overlap = R - d;
    const bool none = !(K > 0.0f) || !(overlap > 0.0f);
    const float force = fminf(K, K) * overlap * R / (d * d);
    const float qa = force / K;
    const float ea = epsilon_from_force_ratio(qa), sa = rest_from_force_ratio(qa);
    const float da = (1.0f + ea) * K;
    const float compliance = ra * sa * sa * sa / da + rb * sa * sa * sa / da;
    Law law;
    law.force = none ? 0.0f : force;
    law.epsA = none ? 0.0f : ea;
    law.epsB = law.epsA;
    law.stiffness = none ? 0.0f : (compliance > kFloatMin ? 1.0f / compliance : 0.0f);
    return law;
}

// ---------------------------------------------------------------------------------------
// Segments, Geometry:436-440, 665-668; Force:112-140.
// ---------------------------------------------------------------------------------------
BPD float2 closest_point(float2 p, float2 a, float2 b) {
    float2 edge = sub(b, a);
    float square = dot2(edge, edge);
    float t = square > kFloatMin ? clampf(dot2(sub(p, a), edge) / square, 0.0f, 1.0f) : 0.0f;
    return add(a, mul(edge, t));
}
BPD bool past_segment_end(float2 center, float2 from, float2 to) {
    float2 edge = sub(to, from), rel = sub(center, from);
    return __builtin_fmaf(rel.x, edge.x, rel.y * edge.y) >= __builtin_fmaf(edge.x, edge.x, edge.y * edge.y);
}

struct Seg {
    float2 a, b, prevStart;
    float2 base;          // unit left normal of b - a, (0, 1) for a degenerate segment
    float outward;
    uint32_t flags;
};
constexpr uint32_t kSegPolygon = 2u, kSegConvexStart = 4u;

BPD float2 segment_base(float2 a, float2 b) {
    float2 edge = sub(b, a);
    float edgeLength = len2(edge);
    return edgeLength > 0.0f ? dvs(f2(-edge.y, edge.x), edgeLength) : f2(0.0f, 1.0f);
}

// Contact region of a segment or one-sided polygon edge (Force:112-140).
BPD bool wall_region(const Seg &s, float2 center, float2 &normal, float &distanceValue) {
    float2 start = s.a, end = s.b, edge = sub(end, start);
    if (s.flags & kSegPolygon) {
        float2 outward = mul(s.base, s.outward);
        float along = dot2(sub(center, start), edge);
        bool convexStart = (s.flags & kSegConvexStart) != 0;
        if (convexStart && along <= 0.0f && past_segment_end(center, s.prevStart, start)) {
            float2 delta = sub(center, start);
            distanceValue = len2(delta);
            normal = distanceValue > 0.0f ? dvs(delta, distanceValue) : outward;
        } else if (along > 0.0f && !past_segment_end(center, start, end) && dot2(sub(center, start), outward) >= 0.0f) {
            normal = outward;
            distanceValue = dot2(sub(center, start), outward);
        } else {
            return false;
        }
    } else {
        float2 delta = sub(center, closest_point(center, s.a, s.b));
        distanceValue = len2(delta);
        normal = distanceValue > 0.0f ? dvs(delta, distanceValue) : s.base;
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Rigid bodies as bubble obstacles, Geometry:134-143, 467-490; Force:151-210.
// ---------------------------------------------------------------------------------------
struct BodyView {
    float4 pose;      // x, y, angle, previous angle
    float2 prev;
    float4 motion;    // vx, vy, omega, damping
    float4 shape;     // bounding radius, core radius, half length
    uint4 ranges;     // vertex offset/count, triangle offset/count
    uint4 kind;       // shape, flags, category, mask
};
constexpr uint32_t kBodyStatic = 1u, kBodyKinematic = 2u;

BPD bool body_is_round(const BodyView &b) { return b.kind.x != 0u; }
BPD bool body_immovable(const BodyView &b) { return (b.kind.y & (kBodyStatic | kBodyKinematic)) != 0u; }

// A body at one step fraction (Force:151-160): interpolated pose with the rotation evaluated
// once, so that every vertex and every inside test of the body uses the same cos and sin.
struct BodyFrame {
    float2 at;            // interpolated position
    float c, s;           // cos, sin of the interpolated angle (body to world)
    float cn, sn;         // cos, sin of its negation (world to body)
    float core;           // core radius of a round body
    float4 motion;        // vx, vy, omega (clamped), damping
    uint4 ranges;         // vertex offset/count, triangle offset/count
    bool round;
};

// Interpolated position and the rotation terms (cos, sin, cos of -angle, sin of -angle).
BPD void frame_pose(float4 pose, float2 prev, float fraction, float2 &at, float4 &cs) {
    at = f2(prev.x + (pose.x - prev.x) * fraction, prev.y + (pose.y - prev.y) * fraction);
    float angle = pose.w + (pose.z - pose.w) * fraction;
    cs = make_float4(bpm::cosf_det(angle), bpm::sinf_det(angle), bpm::cosf_det(-angle), bpm::sinf_det(-angle));
}

BPD BodyFrame body_frame(const BodyView &b, float fraction) {
    BodyFrame f;
    float4 cs;
    frame_pose(b.pose, b.prev, fraction, f.at, cs);
    f.c = cs.x; f.s = cs.y; f.cn = cs.z; f.sn = cs.w;
    f.core = b.shape.y;
    f.motion = b.motion;
    f.ranges = b.ranges;
    f.round = body_is_round(b);
    return f;
}

// Triangle corner indices from the global uint4 records or from 10-bit packed LDS words.
BPD uint3 triangle_of(const uint4 *t, uint32_t i) { uint4 v = t[i]; return make_uint3(v.x, v.y, v.z); }
BPD uint3 triangle_of(const uint32_t *t, uint32_t i) { uint32_t v = t[i]; return make_uint3(v & 1023u, (v >> 10) & 1023u, v >> 20); }

template <class V>
BPD float2 frame_vertex(const BodyFrame &f, V verts, uint32_t i) {
    float2 p = verts[f.ranges.x + i];
    return add(f.at, f2(p.x * f.c - p.y * f.s, p.x * f.s + p.y * f.c));
}

// Centre inside the body (Geometry:134-143, 467-490): round core distance, else the union of
// the triangles in body coordinates.
template <int EB = 4, class V, class T>
BPD bool frame_inside(float2 p, const BodyFrame &f, V verts, T tris) {
    if (f.round) {
        float2 a = frame_vertex(f, verts, 0);
        float2 c = f.ranges.y > 1 ? frame_vertex(f, verts, 1) : a;
        return dist2(p, closest_point(p, a, c)) <= f.core;
    }
    float2 q = sub(p, f.at);
    float2 local = f2(q.x * f.cn - q.y * f.sn, q.x * f.sn + q.y * f.cn);
    // EB triangles at a time (independent tests); the union is the same in any order.
    const uint32_t count = f.ranges.w;
    for (uint32_t i = 0; i < count; i += EB) {
        bool hit = false;
#pragma unroll
        for (int k = 0; k < EB; ++k) {
            const uint32_t ti = i + k < count ? i + k : i;
            uint3 t = triangle_of(tris, f.ranges.z + ti);
            float2 a = verts[f.ranges.x + t.x], bb = verts[f.ranges.x + t.y], c = verts[f.ranges.x + t.z];
            hit = hit || (cross2(sub(bb, a), sub(local, a)) >= 0.0f && cross2(sub(c, bb), sub(local, bb)) >= 0.0f &&
                          cross2(sub(a, c), sub(local, c)) >= 0.0f);
        }
        if (hit) return true;
    }
    return false;
}

struct Obstacle { float2 normal; float distanceValue; float speedBound; };

// Body as an obstacle of a bubble centre (Force:162-210): normal, signed distance to the
// surface and the speed bound of the nearest surface point.
template <int EB = 4, class V, class T>
BPD Obstacle frame_obstacle(const BodyFrame &body, V verts, T tris, float2 center) {
    Obstacle result;
    float2 point, normal;
    float distanceValue;
    if (body.round) {
        float2 a = frame_vertex(body, verts, 0);
        float2 c = body.ranges.y > 1 ? frame_vertex(body, verts, 1) : a;
        point = closest_point(center, a, c);
        float2 delta = sub(center, point);
        float away = len2(delta);
        float2 axis = sub(c, a);
        float2 perp = f2(-axis.y, axis.x);
        float2 fallback = len2(axis) > 1e-6f ? dvs(perp, len2(perp)) : f2(0.0f, 1.0f);
        normal = away > 1e-7f ? dvs(delta, away) : fallback;
        distanceValue = away - body.core;
        point = add(point, mul(normal, body.core));
    } else {
        float best = inff();
        float2 nearest = f2(0.0f, 0.0f), outward = f2(0.0f, 1.0f);
        const uint32_t count = body.ranges.y;
        uint32_t bestEdge = 0;
        bool found = false;
        float2 p = count > 0 ? frame_vertex(body, verts, 0) : f2(0.0f, 0.0f);
        // EB edges at a time: their closest points are independent and computed together; the
        // nearest is then taken in edge order with the strict comparison, as one edge at a time.
        for (uint32_t e = 0; e < count; e += EB) {
            float2 v[EB + 1];
            v[0] = p;
#pragma unroll
            for (int k = 1; k <= EB; ++k) v[k] = frame_vertex(body, verts, e + k < count ? e + k : 0u);
            float2 cp[EB];
            float cd[EB];
#pragma unroll
            for (int k = 0; k < EB; ++k) {
                cp[k] = closest_point(center, v[k], v[k + 1]);
                cd[k] = dist2(center, cp[k]);
            }
#pragma unroll
            for (int k = 0; k < EB; ++k)
                if (e + k < count && cd[k] < best) { best = cd[k]; nearest = cp[k]; bestEdge = e + k; found = true; }
            p = v[EB];
        }
        if (found) {
            const float2 a = frame_vertex(body, verts, bestEdge);
            const float2 b = frame_vertex(body, verts, bestEdge + 1 == count ? 0u : bestEdge + 1);
            const float2 edge = sub(b, a);
            const float size = len2(edge);
            outward = size > kFloatMin ? dvs(f2(edge.y, -edge.x), size) : f2(0.0f, 1.0f);
        }
        bool inside = frame_inside<EB>(center, body, verts, tris);
        normal = inside ? outward : (best > 1e-7f ? dvs(sub(center, nearest), best) : outward);
        distanceValue = inside ? -best : best;
        point = nearest;
    }
    float2 arm = sub(point, body.at);
    result.normal = normal;
    result.distanceValue = distanceValue;
    result.speedBound = len2(f2(body.motion.x, body.motion.y)) + fabsf(body.motion.z) * len2(arm);
    return result;
}

BPD Obstacle body_obstacle(const BodyView &source, float fraction, const float2 *verts, const uint4 *tris, float2 center) {
    return frame_obstacle(body_frame(source, fraction), verts, tris, center);
}

// ---------------------------------------------------------------------------------------
// Force fields, Force:389-421.
// ---------------------------------------------------------------------------------------
// Force of one field at a point: range test in the field's frame (circle or polygon by ray
// crossings), then the directional force turned by the field's rotation, or the radial or vortex
// force with the (1 - d/R)^falloff profile.
BPD float2 field_force(const DevField &f, const float2 *verts, float2 point) {
    float2 delta = sub(point, f2(f.pose.x, f.pose.y));
    float c = bpm::cosf_det(f.pose.z), s = bpm::sinf_det(f.pose.z);
    float2 local = f2(c * delta.x + s * delta.y, -s * delta.x + c * delta.y);
    float distanceValue = len2(local), radius = f.strength.w;
    if (f.shape.z == 0u) {
        if (!(distanceValue <= radius)) return f2(0.0f, 0.0f);
    } else {
        bool inside = false;
        const uint32_t start = f.shape.y, count = f.shape.z;
        for (uint32_t i = 0; i < count; ++i) {
            float2 a = verts[start + i], b = verts[start + (i + 1 == count ? 0u : i + 1)];
            if ((a.y > local.y) != (b.y > local.y) && local.x < (b.x - a.x) * (local.y - a.y) / (b.y - a.y) + a.x) inside = !inside;
        }
        if (!inside) return f2(0.0f, 0.0f);
    }
    if (f.shape.x == 0u) return f2(c * f.strength.x - s * f.strength.y, s * f.strength.x + c * f.strength.y);
    if (!(distanceValue > 0.0f)) return f2(0.0f, 0.0f);
    float profile = f.pose.w == 0.0f ? 1.0f : bpm::powf_det(fmaxf(0.0f, 1.0f - distanceValue / radius), f.pose.w);
    float scale = f.strength.z * profile / distanceValue;
    return f.shape.x == 2u ? mul(f2(-delta.y, delta.x), scale) : mul(delta, scale);
}

// Whether a field acts on a bubble: no filter acts on every bubble, a filter only on its listed
// ids (binary search over the ascending list); an empty filter acts on none.
BPD bool field_affects(const DevField &f, const uint64_t *ids, uint64_t id) {
    if (f.filter.z == 0u) return true;
    uint32_t low = f.filter.x, end = f.filter.x + f.filter.y, high = end;
    while (low < high) {
        uint32_t mid = low + (high - low) / 2;
        if (ids[mid] < id) low = mid + 1; else high = mid;
    }
    return low < end && ids[low] == id;
}

// Exact half impulse with drag, Force:754-758.
BPD float half_impulse_factor(float gamma, float halfStep) {
    float x = gamma * halfStep;
    if (x < 0.05f)
        return halfStep * (1.0f - x * 0.5f + x * x * (1.0f / 6.0f) - x * x * x * (1.0f / 24.0f) + x * x * x * x * (1.0f / 120.0f));
    return (1.0f - bpm::expf_det(-x)) / gamma;
}

}  // namespace bpd

#endif  // BP_DEVICE_H
