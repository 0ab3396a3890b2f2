// bp_rigid.hip -- rigid solver for worlds without bubbles: one workgroup per world, every step of
// a launch inside the kernel.
//
// A step is the reference's rigid-body frame (NativeMetalSolver.encodeStepPrefix,
// NativeMetalSolver.swift:766-799, and nativeRigidFinish at the end gate, :917-921):
//   prediction with the swept vertex test against static edges (PostSolve:29-112);
//   body-pair candidate rows (Geometry:350-415, body-pair mode);
//   body-pair time of impact and rollback (PostSolve:210-271);
//   `iterations` rounds of: pair contacts per row slot (Contact:382-433), their sum per body
//   (Contact:435-449), joints (Contact:474-568) and the wall guard (PostSolve:116-178);
//   speed clamps and the region rule (PostSolve:180-206).
// Each of those is a dispatch in the reference whose threads write only their own body, row slot
// or joint; here each is a loop over the workgroup's lanes followed by a workgroup barrier, so the
// passes see the same inputs and give the same bits. Worlds of at most kRigidLdsBodies bodies run
// one wavefront with their bodies in LDS; larger worlds run 256 lanes with their bodies in global
// memory at the world's working side. Rows, corrections and wall records stay in global memory,
// where the host reads the step's contact records.
//
// Publication: the kernel copies the published side into the working bodies, steps there, and on
// success makes them the published side (NativeMetalSolver.swift:1179-1201: a failed step
// publishes nothing). A failed world is not stepped again until new input.
#include <hip/hip_runtime.h>

#include "bp_device.h"
#include "bp_rigid.h"

namespace bpr {

using namespace bpd;

// ---------------------------------------------------------------------------------------
// Geometry helpers (Geometry:134-158, 436-490, 546-622; Contact:246-378).
// ---------------------------------------------------------------------------------------
BPD float2 rot(float2 p, float angle) {
    float c = bpm::cosf_det(angle), s = bpm::sinf_det(angle);
    return f2(p.x * c - p.y * s, p.x * s + p.y * c);
}
BPD float2 unit(float2 v) { return dvs(v, len2(v)); }
BPD float mixs(float x, float y, float t) { return x + (y - x) * t; }
BPD bool finite4(const float *v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]) && isfinite(v[3]); }
BPD bool is_round(const bp_rigid_body &b) { return b.shape[0] > 0.5f; }
BPD bool immovable(const bp_rigid_body &b) { return b.flags[0] != 0u || b.filter[2] != 0u; }
BPD bool accepts(const uint32_t *a, const uint32_t *b) { return (a[0] & b[1]) != 0u && (b[0] & a[1]) != 0u; }
BPD float2 xy(const float *v) { return f2(v[0], v[1]); }

BPD float2 clamp_speed(float2 v, float limit) {
    float speed = len2(v);
    return limit > 0.0f && speed > limit ? mul(v, limit / speed) : v;
}
BPD float clamp_spin(float w, float limit) { return limit > 0.0f ? clampf(w, -limit, limit) : w; }

// A body's pose with the cosine and sine of its angle evaluated once: every vertex transform
// of one evaluation reuses them (the same values rot() would compute each time).
struct Xf {
    float2 p;
    float c, s;
    float a;      // the angle itself (inverse rotations take cos and sin of -a)
};
BPD Xf xf_at(float x, float y, float angle) { return Xf{f2(x, y), bpm::cosf_det(angle), bpm::sinf_det(angle), angle}; }
BPD Xf xf_of(const bp_rigid_body &b) { return xf_at(b.pose[0], b.pose[1], b.pose[2]); }
BPD float2 apply(const Xf &x, float2 v) { return add(x.p, f2(v.x * x.c - v.y * x.s, v.x * x.s + v.y * x.c)); }

// The world's geometry and joint tables, offsets already applied.
struct View {
    const float2 *verts;
    const uint4 *tris;
    const bp_segment *segs;
    uint32_t segCount;
    const uint4 *jranges;   // null: no joint tables
    const uint64_t *partners;
    BPD float2 vertex(const bp_rigid_body &b, const Xf &x, uint32_t i) const { return apply(x, verts[b.ranges[0] + i]); }
    BPD bool excluded(uint32_t object, uint64_t partner) const {
        if (!jranges) return false;
        uint4 r = jranges[object];
        for (uint32_t i = 0; i < r.w; ++i)
            if (partners[r.z + i] == partner) return true;
        return false;
    }
};

// Pruning. A body's vertices lie within its reach of its pose (the farthest vertex plus the shape
// radius), so geometry farther than the reach from every pose a test visits can give no hit, no
// overlap and no crossing there. The tests below skip such geometry; the slack is far above the
// rounding of the exact tests at the coordinates involved, so a skipped test is one whose exact
// evaluation returns "nothing": results are unchanged.
BPD float body_reach(const View &w, const bp_rigid_body &b) {
    float r = 0.0f;
    for (uint32_t k = 0; k < b.ranges[1]; ++k) r = fmaxf(r, len2(w.verts[b.ranges[0] + k]));
    return r + b.shape[1];
}
BPD float prune_slack(float reach, float magnitude) { return 1e-2f + 1e-4f * reach + 1e-5f * magnitude; }
BPD float magnitude2(float2 p) { return fmaxf(fabsf(p.x), fabsf(p.y)); }
// Axis-aligned boxes (lo.x, lo.y, hi.x, hi.y) that do not meet.
BPD bool boxes_apart(float4 a, float4 b) { return a.z < b.x || b.z < a.x || a.w < b.y || b.w < a.y; }

BPD float seg_distance(float2 a, float2 b, float2 c, float2 d, float2 &pa, float2 &pb) {
    float2 r = sub(b, a), s = sub(d, c);
    float den = cross2(r, s);
    if (fabsf(den) > 1e-12f) {
        float2 qp = sub(c, a);
        float t = cross2(qp, s) / den, u = cross2(qp, r) / den;
        if (t >= 0.0f && t <= 1.0f && u >= 0.0f && u <= 1.0f) {
            pa = add(a, mul(t, r));
            pb = pa;
            return 0.0f;
        }
    }
    float best = inff();
    float2 cp = closest_point(a, c, d);
    float v = dist2(a, cp);
    if (v < best) { best = v; pa = a; pb = cp; }
    cp = closest_point(b, c, d);
    v = dist2(b, cp);
    if (v < best) { best = v; pa = b; pb = cp; }
    cp = closest_point(c, a, b);
    v = dist2(c, cp);
    if (v < best) { best = v; pa = cp; pb = c; }
    cp = closest_point(d, a, b);
    v = dist2(d, cp);
    if (v < best) { best = v; pa = cp; pb = d; }
    return best;
}

struct Core {
    float2 a, b;
    float radius;
};
BPD Core core(const View &w, const bp_rigid_body &b, const Xf &x) {
    Core c;
    c.a = w.vertex(b, x, 0);
    c.b = b.ranges[1] > 1u ? w.vertex(b, x, 1) : c.a;
    c.radius = b.shape[1];
    return c;
}

// Inside test with the body's inverse rotation given (cos and sin of -angle), for callers that
// test many points against one body.
BPD bool inside_with(const View &w, float2 p, const bp_rigid_body &b, const Xf &x, float ic, float is) {
    if (is_round(b)) {
        Core c = core(w, b, x);
        return dist2(p, closest_point(p, c.a, c.b)) <= c.radius;
    }
    const float2 d = sub(p, x.p);
    const float2 local = f2(d.x * ic - d.y * is, d.x * is + d.y * ic);
    for (uint32_t i = 0; i < b.ranges[3]; ++i) {
        uint4 t = w.tris[b.ranges[2] + i];
        float2 a = w.verts[b.ranges[0] + t.x], bb = w.verts[b.ranges[0] + t.y], c = w.verts[b.ranges[0] + t.z];
        if (cross2(sub(bb, a), sub(local, a)) >= 0.0f && cross2(sub(c, bb), sub(local, bb)) >= 0.0f &&
            cross2(sub(a, c), sub(local, c)) >= 0.0f)
            return true;
    }
    return false;
}
BPD bool inside(const View &w, float2 p, const bp_rigid_body &b, const Xf &x) {
    if (is_round(b)) return inside_with(w, p, b, x, 0.0f, 0.0f);
    return inside_with(w, p, b, x, bpm::cosf_det(-x.a), bpm::sinf_det(-x.a));
}

BPD void round_round(const Core &A, const Core &B, float2 fallback, float2 &normal, float &depth, float2 &point) {
    float2 pa, pb;
    float separation = seg_distance(A.a, A.b, B.a, B.b, pa, pb);
    if (separation > 1e-4f) {
        normal = dvs(sub(pb, pa), separation);
        depth = A.radius + B.radius - separation;
        point = mul(0.5f, add(add(pa, mul(normal, A.radius)), sub(pb, mul(normal, B.radius))));
        return;
    }
    // The candidate axes are the normals of the non-degenerate cores, A's first.
    float2 ea = sub(A.b, A.a), eb = sub(B.b, B.a);
    float la = len2(ea), lb = len2(eb);
    const bool useA = la > 1e-6f, useB = lb > 1e-6f;
    const float2 axisA = useA ? dvs(f2(-ea.y, ea.x), la) : f2(0.0f, 0.0f);
    const float2 axisB = useB ? dvs(f2(-eb.y, eb.x), lb) : f2(0.0f, 0.0f);
    float2 axis = fallback;
    float overlap = 0.0f, orientation = 1.0f, best = inff();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (!(i == 0 ? useA : useB)) continue;
        float2 c = i == 0 ? axisA : axisB;
        float amin = fminf(dot2(A.a, c), dot2(A.b, c)), amax = fmaxf(dot2(A.a, c), dot2(A.b, c));
        float bmin = fminf(dot2(B.a, c), dot2(B.b, c)), bmax = fmaxf(dot2(B.a, c), dot2(B.b, c));
        float plus = amax - bmin, minus = bmax - amin;
        float value = fminf(plus, minus);
        if (value < best) {
            best = value;
            axis = c;
            overlap = value;
            if (fabsf(plus - minus) < 1e-6f) orientation = dot2(c, fallback) >= 0.0f ? 1.0f : -1.0f;
            else orientation = plus < minus ? 1.0f : -1.0f;
        }
    }
    normal = mul(orientation, axis);
    depth = overlap + A.radius + B.radius;
    point = mul(0.5f, add(pa, pb));
}

// A segment moving linearly from (pa, pb) to (ca, cb) during the step.
struct Sweep {
    float2 pa, pb, ca, cb;
};
BPD bool crosses(float2 start, float2 end, const Sweep &s) {
    float2 e0 = sub(s.pb, s.pa);
    float2 de = sub(sub(s.cb, s.ca), e0);
    float2 r0 = sub(start, s.pa);
    float2 dr = sub(sub(end, start), sub(s.ca, s.pa));
    float a = cross2(de, dr), b = cross2(e0, dr) + cross2(de, r0);
    float c = cross2(e0, r0);
    if (a == 0.0f && b == 0.0f && c == 0.0f) {
        float q0 = dot2(r0, sub(r0, e0));
        if (q0 <= 0.0f) return false;
        float qa = dot2(dr, sub(dr, de));
        float qb = dot2(dr, sub(r0, e0)) + dot2(r0, sub(dr, de));
        float minimum = fminf(q0, qa + qb + q0);
        if (qa > 0.0f) {
            float t = clampf(-qb / (2.0f * qa), 0.0f, 1.0f);
            minimum = fminf(minimum, (qa * t + qb) * t + q0);
        }
        return minimum <= 0.0f;
    }
    float roots[2] = {-1.0f, -1.0f};
    if (fabsf(a) <= 1.1920929e-7f * fmaxf(fabsf(b), fabsf(c))) {
        if (fabsf(b) > kFloatMin) roots[0] = -c / b;
    } else {
        float d = b * b - 4.0f * a * c;
        if (d > 0.0f) {
            float q = -0.5f * (b + copysignf(__builtin_sqrtf(d), b));
            roots[0] = q / a;
            roots[1] = fabsf(q) > kFloatMin ? c / q : -1.0f;
        }
    }
    for (int k = 0; k < 2; ++k) {
        float t = roots[k];
        if (!(t > 0.0f && t < 1.0f)) continue;
        float2 edge = add(e0, mul(de, t)), rel = add(r0, mul(dr, t));
        float square = dot2(edge, edge), along = dot2(rel, edge);
        if (square > kFloatMin && along >= 0.0f && along <= square) return true;
    }
    return false;
}

// The unit normals of a triangle's edges (the axes of Contact's triangle test); false when an edge
// is degenerate, which refuses every pair with that triangle.
BPD bool tri_axes(const float2 *t, float2 *axis) {
    bool ok = true;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const float2 d = sub(t[(e + 1) % 3], t[e]);
        const float m = len2(d);
        ok = ok && m > kFloatMin;
        axis[e] = m > kFloatMin ? dvs(f2(-d.y, d.x), m) : f2(0.0f, 0.0f);
    }
    return ok;
}

// Separating-axis test of two triangles over a's edge normals, then b's, with the axes given.
BPD bool sat_triangles(const float2 *a, const float2 *axA, const float2 *b, const float2 *axB, float2 &normal, float &depth,
                       float2 &point) {
    depth = inff();
#pragma unroll
    for (int shape = 0; shape < 2; ++shape)
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const float2 axis = shape == 0 ? axA[e] : axB[e];
            float amin = inff(), amax = -inff(), bmin = inff(), bmax = -inff();
            for (int k = 0; k < 3; ++k) {
                float pa = dot2(a[k], axis), pb = dot2(b[k], axis);
                amin = fminf(amin, pa);
                amax = fmaxf(amax, pa);
                bmin = fminf(bmin, pb);
                bmax = fmaxf(bmax, pb);
            }
            if (amax < bmin || bmax < amin) return false;
            float plus = amax - bmin, minus = bmax - amin;
            float overlap = fminf(plus, minus);
            if (overlap < depth) {
                depth = overlap;
                normal = plus < minus ? axis : neg(axis);
            }
        }
    float maxA = -inff(), minB = inff();
    float2 pa = f2(0.0f, 0.0f), pb = f2(0.0f, 0.0f);
    uint32_t na = 0, nb = 0;
    for (int k = 0; k < 3; ++k) {
        maxA = fmaxf(maxA, dot2(a[k], normal));
        minB = fminf(minB, dot2(b[k], normal));
    }
    for (int k = 0; k < 3; ++k) {
        if (dot2(a[k], normal) >= maxA - 1e-5f) { pa = add(pa, a[k]); ++na; }
        if (dot2(b[k], normal) <= minB + 1e-5f) { pb = add(pb, b[k]); ++nb; }
    }
    point = mul(add(dvs(pa, float(na)), dvs(pb, float(nb))), 0.5f);
    return true;
}

// Normal from the round body R to the polygon P.
BPD void round_polygon(const View &w, const Core &R, const bp_rigid_body &P, const Xf &xp, float2 &normal, float &depth,
                       float2 &point) {
    const uint32_t n = P.ranges[1];
    float insideDepth = -1.0f;
    float2 insideOut = f2(0.0f, 0.0f), insidePoint = f2(0.0f, 0.0f);
    uint32_t ends = (R.b.x != R.a.x || R.b.y != R.a.y) ? 2u : 1u, endsInside = 0;
    for (uint32_t k = 0; k < ends; ++k) {
        float2 end = k == 0 ? R.a : R.b;
        if (!inside(w, end, P, xp)) continue;
        ++endsInside;
        float nearest = inff();
        float2 nearestPoint = f2(0.0f, 0.0f), nearestOut = f2(0.0f, 0.0f);
        for (uint32_t e = 0; e < n; ++e) {
            float2 p = w.vertex(P, xp, e), q = w.vertex(P, xp, (e + 1) % n);
            float2 c = closest_point(end, p, q);
            float d = dist2(end, c);
            if (d < nearest) {
                nearest = d;
                nearestPoint = c;
                float2 edge = sub(q, p);
                float size = len2(edge);
                nearestOut = size > kFloatMin ? dvs(f2(edge.y, -edge.x), size) : f2(0.0f, 1.0f);
            }
        }
        if (nearest > insideDepth) {
            insideDepth = nearest;
            insidePoint = nearestPoint;
            insideOut = nearestOut;
        }
    }
    if (endsInside == ends) {
        normal = neg(insideOut);
        depth = R.radius + insideDepth;
        point = insidePoint;
        return;
    }
    float best = inff();
    float2 pc = f2(0.0f, 0.0f), pe = f2(0.0f, 0.0f);
    for (uint32_t e = 0; e < n; ++e) {
        float2 p = w.vertex(P, xp, e), q = w.vertex(P, xp, (e + 1) % n);
        float2 ca, ce;
        float d = seg_distance(R.a, R.b, p, q, ca, ce);
        if (d < best) {
            best = d;
            pc = ca;
            pe = ce;
        }
    }
    if (best > 1e-4f) {
        normal = dvs(sub(pe, pc), best);
        depth = R.radius - best;
        point = pe;
        return;
    }
    float minimum = inff();
    float2 minOut = f2(0.0f, 1.0f), minPoint = pe;
    for (uint32_t e = 0; e < n; ++e) {
        float2 p = w.vertex(P, xp, e), q = w.vertex(P, xp, (e + 1) % n);
        float2 edge = sub(q, p);
        float size = len2(edge);
        if (size <= kFloatMin) continue;
        float2 out = dvs(f2(edge.y, -edge.x), size);
        float da = dot2(sub(R.a, p), out), db = dot2(sub(R.b, p), out);
        float d = R.radius - fminf(da, db);
        if (d < minimum) {
            minimum = d;
            minOut = out;
            minPoint = sub(da <= db ? R.a : R.b, mul(out, fminf(da, db)));
        }
    }
    normal = neg(minOut);
    depth = isfinite(minimum) ? minimum : R.radius;
    point = minPoint;
}

BPD bool core_crosses(const View &w, const Core &R, const bp_rigid_body &P, const Xf &xp) {
    if (inside(w, R.a, P, xp) || inside(w, R.b, P, xp)) return true;
    for (uint32_t e = 0; e < P.ranges[1]; ++e) {
        float2 p = w.vertex(P, xp, e), q = w.vertex(P, xp, (e + 1) % P.ranges[1]);
        float2 ca, ce;
        if (seg_distance(R.a, R.b, p, q, ca, ce) <= 1e-4f) return true;
    }
    return false;
}

// Normal from a to b, with the bodies at the poses xa and xb (their own or interpolated ones).
BPD bool pair_contact_at(const View &w, const bp_rigid_body &a, const Xf &xa, const bp_rigid_body &b, const Xf &xb,
                         float2 &n, float &depth, float2 &point) {
    const bool ra = is_round(a), rb = is_round(b);
    if (ra && rb) {
        round_round(core(w, a, xa), core(w, b, xb), f2(1.0f, 0.0f), n, depth, point);
        return true;
    }
    if (ra != rb) {
        // The poses are selected by value: a select between references would put both in memory.
        const Xf xr = ra ? xa : xb, xp = ra ? xb : xa;
        float2 normal;
        round_polygon(w, core(w, ra ? a : b, xr), ra ? b : a, xp, normal, depth, point);
        n = ra ? normal : neg(normal);
        return true;
    }
    // Both polygons: the triangle pairs, with each corner transformed where it is used (the same
    // operations as a transform of every vertex first, without a per-lane array in memory).
    bool found = false;
    // A triangle's axes are the same for every triangle it meets: a's are taken once per triangle.
    for (uint32_t i = 0; i < a.ranges[3]; ++i) {
        uint4 ta = w.tris[a.ranges[2] + i];
        float2 av[3] = {w.vertex(a, xa, ta.x), w.vertex(a, xa, ta.y), w.vertex(a, xa, ta.z)};
        float2 axA[3];
        if (!tri_axes(av, axA)) continue;
        for (uint32_t j = 0; j < b.ranges[3]; ++j) {
            uint4 tb = w.tris[b.ranges[2] + j];
            float2 bv[3] = {w.vertex(b, xb, tb.x), w.vertex(b, xb, tb.y), w.vertex(b, xb, tb.z)};
            float2 axB[3];
            if (!tri_axes(bv, axB)) continue;
            float2 normal, p;
            float d;
            if (sat_triangles(av, axA, bv, axB, normal, d, p) && (!found || d > depth)) {
                found = true;
                depth = d;
                n = normal;
                point = p;
            }
        }
    }
    return found;
}

BPD bool pair_contact(const View &w, const bp_rigid_body &a, const bp_rigid_body &b, float2 &n, float &depth, float2 &point) {
    return pair_contact_at(w, a, xf_of(a), b, xf_of(b), n, depth, point);
}

struct P3 {
    float x, y, a;
};
BPD P3 lerp3(P3 s, P3 e, float t) { return P3{mixs(s.x, e.x, t), mixs(s.y, e.y, t), mixs(s.a, e.a, t)}; }

// Swept vertices of body b from pose p to q against the world's static edges (PostSolve:29-62).
BPD bool swept_hit(const View &w, const bp_rigid_body &b, float reach, P3 p, P3 q, uint32_t entity, float2 &point,
                   float2 &normal, int64_t &owner) {
    const Xf xp = xf_at(p.x, p.y, p.a), xq = xf_at(q.x, q.y, q.a);
    // Every swept vertex and edge stays within `reach` of the centre's path from p to q.
    const float2 c0 = f2(p.x, p.y), c1 = f2(q.x, q.y);
    const float pathMagnitude = fmaxf(magnitude2(c0), magnitude2(c1));
    for (uint32_t j = 0; j < w.segCount; ++j) {
        const bp_segment &s = w.segs[j];
        {
            const float2 a = f2(s.a[0], s.a[1]), e = f2(s.b[0], s.b[1]);
            const float r = reach + prune_slack(reach, fmaxf(pathMagnitude, fmaxf(magnitude2(a), magnitude2(e))));
            const float4 path = make_float4(fminf(c0.x, c1.x) - r, fminf(c0.y, c1.y) - r, fmaxf(c0.x, c1.x) + r, fmaxf(c0.y, c1.y) + r);
            if (boxes_apart(path, make_float4(fminf(a.x, e.x), fminf(a.y, e.y), fmaxf(a.x, e.x), fmaxf(a.y, e.y)))) continue;
        }
        const bool owned = (s.flags & BP_SEGMENT_HAS_OWNER) != 0u;
        if (owned && (uint64_t)s.owner_id == b.id) continue;
        if (owned && w.excluded(entity, (uint64_t)s.owner_id)) continue;
        const uint32_t sf[2] = {s.category, s.mask};
        if (!accepts(b.filter, sf)) continue;
        const float2 sa = f2(s.a[0], s.a[1]), sb = f2(s.b[0], s.b[1]);
        const Sweep still{sa, sb, sa, sb};
        for (uint32_t k = 0; k < b.ranges[1]; ++k) {
            float2 v = w.verts[b.ranges[0] + k];
            float2 v0 = apply(xp, v), v1 = apply(xq, v);
            bool hit = crosses(v0, v1, still);
            if (!hit) {
                float2 next = w.verts[b.ranges[0] + (k + 1) % b.ranges[1]];
                Sweep edge{v0, apply(xp, next), v1, apply(xq, next)};
                hit = crosses(sa, sa, edge) || crosses(sb, sb, edge);
            }
            if (hit) {
                point = closest_point(v1, sa, sb);
                float2 e = sub(sb, sa);
                normal = unit(f2(e.y, -e.x));
                if (dot2(sub(v0, sa), normal) < 0.0f) normal = neg(normal);
                owner = s.owner_id;
                return true;
            }
        }
    }
    return false;
}

// `reach` is the sum of both bodies' reaches: farther apart, the pair has no contact of positive
// depth at t.
BPD bool overlap_at(const View &w, const bp_rigid_body &a, const bp_rigid_body &b, float reach, float t) {
    const float2 pa = f2(mixs(a.previous[0], a.pose[0], t), mixs(a.previous[1], a.pose[1], t));
    const float2 pb = f2(mixs(b.previous[0], b.pose[0], t), mixs(b.previous[1], b.pose[1], t));
    const float limit = reach + prune_slack(reach, fmaxf(magnitude2(pa), magnitude2(pb)));
    const float2 gap = sub(pa, pb);
    if (dot2(gap, gap) > limit * limit) return false;
    const Xf xa = xf_at(pa.x, pa.y, mixs(a.pose[3], a.pose[2], t));
    const Xf xb = xf_at(pb.x, pb.y, mixs(b.pose[3], b.pose[2], t));
    float2 n, point;
    float depth;
    return pair_contact_at(w, a, xa, b, xb, n, depth, point) && depth > 1e-3f;
}

// The travel that decides whether a pair's impact time is searched (PostSolve:220-226).
BPD float pair_travel(const bp_rigid_body &a, const bp_rigid_body &b) {
    return len2(sub(sub(xy(a.pose), xy(a.previous)), sub(xy(b.pose), xy(b.previous)))) +
           fabsf(a.pose[2] - a.pose[3]) * a.physical[2] + fabsf(b.pose[2] - b.pose[3]) * b.physical[2];
}

constexpr uint32_t kNoTwin = 0xffffffffu;
constexpr uint32_t kGuardWalls = 32;

BPD bool key_less(const RigidCand &a, const RigidCand &b) {
    if (a.kind != b.kind) return a.kind < b.kind;
    if (a.first != b.first) return a.first < b.first;
    return a.second < b.second;
}

// One body-pair row entry of a contact iteration (Contact:382-433): the correction `r` (its w lanes
// carried over from `prev` when nothing is found) and the normal correction lambda it adds to the
// entry's impulse sum (0 = none).
BPD void evaluate_entry(const View &w, const bp_rigid_body *B, const RigidCand &cand, const RigidCorr &prev, RigidCorr &r,
                        bool &added, float &lambdaOut) {
    r = RigidCorr{};
    r.first[3] = prev.first[3];
    r.second[3] = prev.second[3];
    r.vfirst[3] = prev.vfirst[3];
    r.vsecond[3] = prev.vsecond[3];
    added = false;
    lambdaOut = 0.0f;
    const uint32_t ai = cand.low, bi = cand.high;
    if (cand.kind != 3u || ai >= bi) return;
    const bp_rigid_body &a = B[ai], &b = B[bi];
    if (immovable(a) && immovable(b)) return;
    float2 gap = sub(xy(a.pose), xy(b.pose));
    float radius = a.physical[2] + b.physical[2];
    if (dot2(gap, gap) > radius * radius) return;
    float2 nrm, point;
    float depth;
    if (!pair_contact(w, a, b, nrm, depth, point)) return;
    const float wa = a.physical[0], wb = b.physical[0], ia = a.physical[1], ib = b.physical[1];
    const float2 ra = sub(point, xy(a.pose)), rb = sub(point, xy(b.pose));
    const float2 va = f2(a.motion[0], a.motion[1]), vb = f2(b.motion[0], b.motion[1]);
    const float spinA = a.motion[2], spinB = b.motion[2];
    const float ca = cross2(ra, nrm), cb = cross2(rb, nrm);
    const float denom = wa + wb + ia * ca * ca + ib * cb * cb;
    if (depth < -0.001f || denom <= kFloatMin) return;
    const float lambda = fmaxf(0.0f, depth) / denom * 0.5f;
    added = true;
    lambdaOut = lambda;
    float2 m = mul(mul(neg(nrm), wa), lambda);
    r.first[0] = m.x;
    r.first[1] = m.y;
    r.first[2] = -ca * ia * lambda;
    r.first[3] = nrm.x;
    m = mul(mul(nrm, wb), lambda);
    r.second[0] = m.x;
    r.second[1] = m.y;
    r.second[2] = cb * ib * lambda;
    r.second[3] = nrm.y;
    const float2 dv = sub(vb, va);
    const float closing = dot2(dv, nrm) + spinB * cb - spinA * ca;
    const float friction = __builtin_sqrtf(fmaxf(0.0f, a.physical[3] * b.physical[3]));
    const float restitution = fmaxf(__uint_as_float(a.flags[3]), __uint_as_float(b.flags[3]));
    const float impulse = fmaxf(0.0f, -closing) * (1.0f + restitution) / denom * 0.5f;
    const float2 tangent = f2(-nrm.y, nrm.x);
    const float tangentSpeed = dot2(dv, tangent) + spinB * cross2(rb, tangent) - spinA * cross2(ra, tangent);
    const float tangentImpulse = -tangentSpeed * friction / denom * 0.5f;
    m = sub(mul(mul(neg(nrm), wa), impulse), mul(mul(tangent, wa), tangentImpulse));
    r.vfirst[0] = m.x;
    r.vfirst[1] = m.y;
    r.vfirst[2] = -ca * ia * impulse - cross2(ra, tangent) * ia * tangentImpulse;
    r.vfirst[3] = point.x;
    m = add(mul(mul(nrm, wb), impulse), mul(mul(tangent, wb), tangentImpulse));
    r.vsecond[0] = m.x;
    r.vsecond[1] = m.y;
    r.vsecond[2] = cb * ib * impulse + cross2(rb, tangent) * ib * tangentImpulse;
    r.vsecond[3] = point.y;
}

// ---------------------------------------------------------------------------------------
// Joints (Contact:452-548): one lane per joint computes both ends' corrections from the same
// state; the joint's accumulated lambdas live in its record for the step.
// ---------------------------------------------------------------------------------------
struct JEnd {
    float2 point, arm, vel;
    float im, ii, angle, omega;
};
BPD JEnd joint_end(const bp_rigid_body *B, uint32_t entity, float2 anchor) {
    JEnd e;
    e.point = anchor;
    e.arm = f2(0.0f, 0.0f);
    e.vel = f2(0.0f, 0.0f);
    e.im = e.ii = e.angle = e.omega = 0.0f;
    if (entity == BP_JOINT_WORLD_END) return e;
    const bp_rigid_body &b = B[entity];
    // Anchors are in the geometry's frame; the centre of mass sits at the local centroid.
    e.arm = rot(sub(anchor, f2(b.previous[2], b.previous[3])), b.pose[2]);
    e.point = add(xy(b.pose), e.arm);
    e.vel = f2(b.motion[0], b.motion[1]);
    e.omega = b.motion[2];
    e.angle = b.pose[2];
    e.im = b.physical[0];
    e.ii = b.physical[1];
    return e;
}
BPD float2 perp(float2 v) { return f2(-v.y, v.x); }

__device__ void rigid_solve_joints(const bp_rigid_body *B, bp_joint *joints, RigidCorr *jcorr, const RigidDesc &D,
                                   uint32_t tid, uint32_t width, const View &) {
    const float dt = fmaxf(D.dt, 1e-6f);
    const uint32_t iterations = max(D.iterations, 1u);
    for (uint32_t j = tid; j < D.jointCount; j += width) {
        bp_joint jt = joints[j];
        const JEnd a = joint_end(B, jt.ends[0], f2(jt.anchors[0], jt.anchors[1]));
        const JEnd b = joint_end(B, jt.ends[1], f2(jt.anchors[2], jt.anchors[3]));
        float2 moveA = f2(0.0f, 0.0f), moveB = f2(0.0f, 0.0f), speedA = f2(0.0f, 0.0f), speedB = f2(0.0f, 0.0f);
        float turnA = 0.0f, turnB = 0.0f, spinA = 0.0f, spinB = 0.0f;
        const float2 pA = perp(a.arm), pB = perp(b.arm);
        const uint32_t kind = jt.ends[2];
        if (kind == 0u) {
            const float2 delta = sub(b.point, a.point);
            const float length = len2(delta);
            const float2 nrm = length > 1e-6f ? dvs(delta, length) : f2(1.0f, 0.0f);
            const float ca = cross2(a.arm, nrm), cb = cross2(b.arm, nrm);
            const float wsum = a.im + b.im + a.ii * ca * ca + b.ii * cb * cb;
            if (wsum > kFloatMin) {
                const float alpha = jt.parameters[1] / (dt * dt);
                const float lambda = (-(length - jt.parameters[0]) - alpha * jt.state[0]) / (wsum + alpha);
                jt.state[0] += lambda;
                moveA = mul(mul(neg(nrm), a.im), lambda);
                turnA = -ca * a.ii * lambda;
                moveB = mul(mul(nrm, b.im), lambda);
                turnB = cb * b.ii * lambda;
                if (jt.parameters[2] > 0.0f) {
                    const float relative = dot2(sub(add(b.vel, mul(b.omega, pB)), add(a.vel, mul(a.omega, pA))), nrm);
                    const float impulse =
                        -relative * (1.0f - bpm::expf_det(-jt.parameters[2] * dt / float(iterations))) / wsum;
                    speedA = sub(speedA, mul(mul(nrm, a.im), impulse));
                    spinA = spinA - ca * a.ii * impulse;
                    speedB = add(speedB, mul(mul(nrm, b.im), impulse));
                    spinB = spinB + cb * b.ii * impulse;
                }
            }
        } else {
            const float weight = a.im + b.im;
            const float k11 = weight + a.ii * pA.x * pA.x + b.ii * pB.x * pB.x;
            const float k12 = a.ii * pA.x * pA.y + b.ii * pB.x * pB.y;
            const float k22 = weight + a.ii * pA.y * pA.y + b.ii * pB.y * pB.y;
            const float det = k11 * k22 - k12 * k12;
            if (det > kFloatMin) {
                const float2 gap = sub(b.point, a.point);
                const float2 impulse = dvs(f2(-(k22 * gap.x - k12 * gap.y), -(-k12 * gap.x + k11 * gap.y)), det);
                moveA = mul(neg(impulse), a.im);
                turnA = -a.ii * dot2(pA, impulse);
                moveB = mul(impulse, b.im);
                turnB = b.ii * dot2(pB, impulse);
            }
            const float inertia = a.ii + b.ii;
            if (inertia > kFloatMin) {
                const float angle = (b.angle + turnB) - (a.angle + turnA) - jt.parameters[3];
                float violation = 0.0f;
                if (kind == 2u) violation = angle;
                else if (jt.ends[3] & 1u)
                    violation = angle < jt.limits[0] ? angle - jt.limits[0] : (angle > jt.limits[1] ? angle - jt.limits[1] : 0.0f);
                if (violation != 0.0f) {
                    const float lambda = -violation / inertia;
                    turnA = turnA - a.ii * lambda;
                    turnB = turnB + b.ii * lambda;
                }
                if (kind == 1u && (jt.ends[3] & 2u)) {
                    const float relative = (b.omega + turnB / dt) - (a.omega + turnA / dt) - jt.limits[2];
                    const float limit = jt.limits[3] * dt;
                    const float total = clampf(jt.state[1] - relative / inertia, -limit, limit);
                    const float lambda = total - jt.state[1];
                    jt.state[1] = total;
                    spinA = spinA - a.ii * lambda;
                    spinB = spinB + b.ii * lambda;
                }
            }
        }
        RigidCorr r;
        r.first[0] = moveA.x; r.first[1] = moveA.y; r.first[2] = turnA; r.first[3] = 0.0f;
        r.second[0] = moveB.x; r.second[1] = moveB.y; r.second[2] = turnB; r.second[3] = 0.0f;
        float2 v = add(dvs(moveA, dt), speedA);
        r.vfirst[0] = v.x; r.vfirst[1] = v.y; r.vfirst[2] = turnA / dt + spinA; r.vfirst[3] = 0.0f;
        v = add(dvs(moveB, dt), speedB);
        r.vsecond[0] = v.x; r.vsecond[1] = v.y; r.vsecond[2] = turnB / dt + spinB; r.vsecond[3] = 0.0f;
        jcorr[j] = r;
        joints[j] = jt;
    }
}

// ---------------------------------------------------------------------------------------
// Workgroup helpers.
// ---------------------------------------------------------------------------------------
template <int BLOCK>
__device__ uint32_t block_prefix(uint32_t *scan, const uint32_t *counts, uint32_t *starts, uint32_t n) {
    // Exclusive prefix of counts[0..n) into starts; returns the total. scan holds BLOCK words.
    const uint32_t tid = threadIdx.x;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += BLOCK) {
        uint32_t v = base + tid < n ? counts[base + tid] : 0u;
        scan[tid] = v;
        __syncthreads();
        for (uint32_t off = 1; off < (uint32_t)BLOCK; off <<= 1) {
            uint32_t add = tid >= off ? scan[tid - off] : 0u;
            __syncthreads();
            scan[tid] += add;
            __syncthreads();
        }
        if (base + tid < n) starts[base + tid] = carry + scan[tid] - v;
        uint32_t total = scan[BLOCK - 1];
        __syncthreads();
        carry += total;
    }
    return carry;
}

// ---------------------------------------------------------------------------------------
// The step kernel.
// ---------------------------------------------------------------------------------------
// BLOCK lanes per world; CAP > 0 keeps the world's bodies (at most CAP) in LDS for the step and
// writes them to the unpublished side when the step succeeds, CAP = 0 steps them in place on the
// unpublished side in global memory. `order` lists the worlds of the launch.
template <int BLOCK, int CAP>
__global__ __launch_bounds__(BLOCK) void rigid_step_kernel(RigidArena A, RigidDesc *descs, const uint32_t *order,
                                                           bp_rigid_frame_info *infos, uint32_t worldCount, uint32_t steps) {
    __shared__ uint32_t sScan[BLOCK];
    __shared__ bp_rigid_body sBodies[CAP > 0 ? CAP : 1];
    __shared__ uint32_t sSafety, sOverflow, sMaxPer, sPairs, sRemoved, sWrapped, sTotal;
    __shared__ uint32_t sKinds[4];
    // The static polygons the wall guard tests, in body order, with their rotation terms and boxes
    // at the current iteration (worlds with more walls test every body as walls in place).
    __shared__ uint32_t sWalls, sWallIndex[kGuardWalls];
    __shared__ float4 sWallRot[kGuardWalls], sWallBox[kGuardWalls];
    const uint32_t wi = order[blockIdx.x];
    const uint32_t tid = threadIdx.x;
    RigidDesc &D = descs[wi];
    const uint32_t n = D.bodyCount;
    View w;
    w.verts = A.verts + D.vertexOffset;
    w.tris = A.tris + D.triangleOffset;
    w.segs = A.segs + D.segmentOffset;
    w.segCount = D.segmentCount;
    w.jranges = D.hasJointTables ? A.jranges + D.bodyOffset : nullptr;
    w.partners = A.partners + D.partnerOffset;
    const uint32_t stride = D.stride, limit = D.k;
    RigidCand *rows = A.rows + D.rowOffset;
    RigidCorr *corr = A.corr + D.rowOffset;
    uint32_t *slots = A.slots + D.rowOffset;
    uint32_t *twin = A.twin + D.rowOffset;
    float *toi = A.toi + D.rowOffset;
    uint32_t *toiSteps = A.toiSteps + D.rowOffset, *toiStart = A.toiStart + D.rowOffset, *toiFirst = A.toiFirst + D.rowOffset;
    uint32_t *rowCount = A.rowCount + D.bodyOffset;
    uint32_t *rowStart = A.rowStart + D.bodyOffset;
    RigidWall *walls = A.walls + D.bodyOffset;
    bp_joint *joints = A.joints + D.jointOffset;
    RigidCorr *jcorr = A.jcorr + D.jointOffset;
    const uint2 *refs = A.refs + D.refOffset;
    const float dt = D.dt;

    for (uint32_t step = 0; step < steps; ++step) {
        bp_rigid_frame_info *info = infos + (size_t)step * worldCount + wi;
        if (D.failed) {
            if (tid == 0) {
                bp_rigid_frame_info r = {};
                r.status = BP_ERR_UNAVAILABLE;
                *info = r;
            }
            continue;
        }
        bp_rigid_body *pub = A.bodies[D.side] + D.bodyOffset;
        bp_rigid_body *out = A.bodies[D.side ^ 1u] + D.bodyOffset;
        bp_rigid_body *B;
        if constexpr (CAP > 0) B = sBodies;
        else B = out;
        if (tid == 0) {
            sSafety = 0;
            sOverflow = 0;
            sMaxPer = 0;
            sRemoved = 0;
            sWrapped = 0;
        }
        for (uint32_t i = tid; i < n; i += BLOCK) {
            bp_rigid_body b = pub[i];
            // Runtime velocity update of role .dynamic before the frame (MotionControllers.swift:57-77).
            if (D.prelude && !b.flags[0] && !b.filter[2]) {
                float factor = bpm::expf_det(-b.motion[3] * dt);
                b.motion[0] = b.motion[0] * factor;
                b.motion[1] = b.motion[1] * factor;
                b.motion[0] = b.motion[0] + D.force[0] / b.mass * dt;
                b.motion[1] = b.motion[1] + D.force[1] / b.mass * dt;
            }
            B[i] = b;
            walls[i] = RigidWall{};
        }
        for (uint32_t j = tid; j < D.jointCount; j += BLOCK) {
            bp_joint jt = A.jointInput[D.jointOffset + j];
            jt.state[0] = jt.state[1] = jt.state[2] = jt.state[3] = 0.0f;
            joints[j] = jt;
        }
        __syncthreads();

        if (tid == 0) {
            uint32_t count = 0;
            for (uint32_t j = 0; j < n; ++j)
                if (B[j].flags[0] && !is_round(B[j])) {
                    if (count < kGuardWalls) sWallIndex[count] = j;
                    ++count;
                }
            sWalls = count;
        }
        // Prediction (PostSolve:64-112).
        for (uint32_t i = tid; i < n; i += BLOCK) {
            bp_rigid_body b = B[i];
            b.previous[0] = b.pose[0];
            b.previous[1] = b.pose[1];
            b.pose[3] = b.pose[2];
            b.flags[2] = 0;
            if (b.flags[0]) {
                b.motion[0] = b.motion[1] = b.motion[2] = 0.0f;
                B[i] = b;
                continue;
            }
            b.motion[2] *= bpm::expf_det(-b.motion[3] * dt);
            float2 v = clamp_speed(f2(b.motion[0], b.motion[1]), b.limits[0]);
            b.motion[0] = v.x;
            b.motion[1] = v.y;
            b.motion[2] = clamp_spin(b.motion[2], b.limits[1]);
            const P3 start{b.pose[0], b.pose[1], b.pose[2]};
            b.pose[0] += b.motion[0] * dt;
            b.pose[1] += b.motion[1] * dt;
            b.pose[2] += b.motion[2] * dt;
            if (!finite4(b.pose) || !finite4(b.motion)) {
                atomicOr(&sSafety, 1u);
                B[i] = b;
                continue;
            }
            if (b.filter[2]) {
                B[i] = b;
                continue;
            }
            float turns = ceilf(fabsf(b.pose[2] - start.a) / 0.05f);
            if (turns > 128.0f) {
                atomicOr(&sSafety, 2u | 16u);
                B[i] = b;
                continue;
            }
            const uint32_t samples = max(1u, (uint32_t)turns);
            const P3 end{b.pose[0], b.pose[1], b.pose[2]};
            const float reach = body_reach(w, b);
            bool crossed = false;
            float2 hitPoint = f2(0.0f, 0.0f), hitNormal = f2(0.0f, 0.0f);
            int64_t hitOwner = 0;
            P3 stop = start;
            for (uint32_t s = 0; s < samples && !crossed; ++s) {
                P3 p = lerp3(start, end, float(s) / float(samples));
                P3 q = lerp3(start, end, float(s + 1) / float(samples));
                crossed = swept_hit(w, b, reach, p, q, i, hitPoint, hitNormal, hitOwner);
                if (!crossed) continue;
                P3 last = p, hit = q;
                for (int k = 0; k < 12; ++k) {
                    P3 mid{0.5f * (last.x + hit.x), 0.5f * (last.y + hit.y), 0.5f * (last.a + hit.a)};
                    float2 pt, nm;
                    int64_t ow;
                    if (swept_hit(w, b, reach, p, mid, i, pt, nm, ow)) hit = mid;
                    else last = mid;
                }
                stop = last;
            }
            if (crossed) {
                float closing = fmaxf(0.0f, -dot2(f2(b.motion[0], b.motion[1]), hitNormal));
                float stopped = b.physical[0] > 0.0f ? closing / b.physical[0] * dt : 0.0f;
                b.pose[0] = stop.x;
                b.pose[1] = stop.y;
                b.pose[2] = stop.a;
                b.motion[0] = b.motion[1] = b.motion[2] = 0.0f;
                for (uint32_t j = 0; j < n; ++j)
                    if (j != i && B[j].flags[0] && B[j].id == (uint64_t)hitOwner) {
                        RigidWall r;
                        r.body = (int64_t)b.id;
                        r.owner = hitOwner;
                        r.geometry[0] = hitNormal.x;
                        r.geometry[1] = hitNormal.y;
                        r.geometry[2] = hitPoint.x;
                        r.geometry[3] = hitPoint.y;
                        r.extra[0] = stopped;
                        r.extra[1] = r.extra[2] = r.extra[3] = 0.0f;
                        walls[i] = r;
                        break;
                    }
            }
            B[i] = b;
        }
        __syncthreads();

        // Body-pair candidate rows (Geometry:350-415), one lane per row.
        for (uint32_t i = tid; i < n; i += BLOCK) {
            RigidCand *row = rows + (size_t)i * stride;
            uint32_t count = 0, dropped = 0;
            for (uint32_t o = 0; o < n; ++o) {
                if (o == i) continue;
                const uint32_t low = min(i, o), high = max(i, o);
                const bp_rigid_body &a = B[low], &c = B[high];
                if (immovable(a) && immovable(c)) continue;
                if (!accepts(a.filter, c.filter)) continue;
                if (w.excluded(i, B[o].id)) continue;
                float2 start = sub(xy(a.previous), xy(c.previous)), end = sub(xy(a.pose), xy(c.pose));
                float2 travel = sub(end, start);
                float tt = dot2(travel, travel);
                float t = tt > kFloatMin ? clampf(-dot2(start, travel) / tt, 0.0f, 1.0f) : 0.0f;
                float radius = 1.25f * (a.physical[2] + c.physical[2]);
                float2 closest = add(start, mul(t, travel));
                if (!(dot2(closest, closest) <= radius * radius)) continue;
                RigidCand cand;
                cand.first = (int64_t)a.id;
                cand.second = (int64_t)c.id;
                cand.kind = 3u;
                cand.low = low;
                cand.high = high;
                cand.impulse = 0.0f;
                uint32_t at;
                if (count < limit) at = count++;
                else {
                    ++dropped;
                    if (!key_less(cand, row[limit - 1])) continue;
                    at = limit - 1;
                }
                while (at > 0 && key_less(cand, row[at - 1])) {
                    row[at] = row[at - 1];
                    --at;
                }
                row[at] = cand;
            }
            rowCount[i] = count;
            if (dropped) atomicAdd(&sOverflow, dropped);
            atomicMax(&sMaxPer, count + dropped);
        }
        __syncthreads();
        {
            uint32_t total = block_prefix<BLOCK>(sScan, rowCount, rowStart, n);
            if (tid == 0) {
                sTotal = total;
                sPairs = total;
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < n; i += BLOCK) {
            const uint32_t s0 = rowStart[i], c = rowCount[i];
            for (uint32_t s = 0; s < c; ++s) slots[s0 + s] = i * stride + s;
        }
        __syncthreads();

        // Each used entry's twin: the same pair in the other body's row (rows are sorted by key). The
        // entry's correction starts the step at zero (only used entries are ever read).
        const uint32_t total = sTotal;
        for (uint32_t e = tid; e < total; e += BLOCK) {
            const uint32_t index = slots[e];
            corr[index] = RigidCorr{};
            const uint32_t owner = index / stride;
            const RigidCand c = rows[index];
            const uint32_t other = owner == c.low ? c.high : c.low;
            const RigidCand *orow = rows + (size_t)other * stride;
            uint32_t lo = 0, hi = rowCount[other];
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (key_less(orow[mid], c)) lo = mid + 1;
                else hi = mid;
            }
            const bool found = lo < rowCount[other] && orow[lo].kind == c.kind && orow[lo].first == c.first &&
                               orow[lo].second == c.second;
            twin[e] = found ? rowStart[other] + lo : kNoTwin;
        }
        __syncthreads();
        // Time of impact of each pair (PostSolve:220-238), once per pair; then each movable body takes
        // the earliest over its row (PostSolve:243-257) and rolls back to it (PostSolve:260-271). The
        // reference searches a pair's slices in order and bisects after the first that overlaps; here
        // (a) one lane per pair tests the start and counts the slices, (b) one lane per slice tests
        // that slice's poses and each pair keeps its earliest overlapping slice, (c) one lane per pair
        // bisects between that slice and the one before. The tests are the same pure function of the
        // same inputs, and the earliest overlapping slice is the one where the search in order stops.
        for (uint32_t e = tid; e < total; e += BLOCK) {
            const uint32_t index = slots[e], pair = twin[e];
            const RigidCand c = rows[index];
            uint32_t slices = 0;
            if (!(pair != kNoTwin && index / stride != c.low)) {
                const bp_rigid_body &a = B[c.low], &b = B[c.high];
                if (pair_travel(a, b) > 0.5f) {
                    const float reach = body_reach(w, a) + body_reach(w, b);
                    if (!overlap_at(w, a, b, reach, 0.0f)) {
                        const float count = ceilf(pair_travel(a, b) / 0.5f);
                        slices = count >= 64.0f ? 64u : max(2u, (uint32_t)count);
                        toi[e] = reach;                  // kept for (b) and (c), then replaced
                        toiFirst[e] = kNoTwin;
                    }
                }
                if (slices == 0) {
                    toi[e] = 2.0f;
                    if (pair != kNoTwin) toi[pair] = 2.0f;
                }
            }
            toiSteps[e] = slices;
        }
        __syncthreads();
        const uint32_t sliceTotal = block_prefix<BLOCK>(sScan, toiSteps, toiStart, total);
        __syncthreads();
        for (uint32_t item = tid; item < sliceTotal; item += BLOCK) {
            uint32_t lo = 0, hi = total;             // the last entry whose slices start at or before item
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (toiStart[mid] <= item) lo = mid + 1;
                else hi = mid;
            }
            const uint32_t e = lo - 1, step = item - toiStart[e] + 1, count = toiSteps[e];
            const RigidCand c = rows[slots[e]];
            if (overlap_at(w, B[c.low], B[c.high], toi[e], float(step) / float(count))) atomicMin(&toiFirst[e], step);
        }
        __syncthreads();
        // (c) The pairs with an overlapping slice, listed (in toiStart, free after (b)); the others
        // have no impact in the step.
        if (tid == 0) sKinds[0] = 0u;
        __syncthreads();
        for (uint32_t e = tid; e < total; e += BLOCK) {
            if (toiSteps[e] == 0u) continue;
            if (toiFirst[e] != kNoTwin) {
                toiStart[atomicAdd(&sKinds[0], 1u)] = e;
            } else {
                toi[e] = 2.0f;
                if (twin[e] != kNoTwin) toi[twin[e]] = 2.0f;
            }
        }
        __syncthreads();
        // The bisection's 12 halvings, two at a time by a group of four lanes: the midpoint and the
        // two midpoints that can follow it are tested together, and the group takes the path the
        // sequential bisection takes (the same midpoints, so the same bits).
        {
            const uint32_t hits = sKinds[0], member = tid & 3u;
            for (uint32_t base = 0; base < hits; base += BLOCK / 4) {
                const uint32_t h = base + tid / 4u;
                const bool active = h < hits;
                const uint32_t e = active ? toiStart[h] : 0u;
                float low = 0.0f, high = 0.0f, reach = 0.0f;
                uint32_t lowBody = 0, highBody = 0;
                if (active) {
                    const uint32_t count = toiSteps[e], first = toiFirst[e];
                    const RigidCand c = rows[slots[e]];
                    lowBody = c.low;
                    highBody = c.high;
                    reach = toi[e];
                    low = first == 1u ? 0.0f : float(first - 1u) / float(count);
                    high = float(first) / float(count);
                }
                for (int round = 0; round < 6; ++round) {
                    const float mid = 0.5f * (low + high), left = 0.5f * (low + mid), right = 0.5f * (mid + high);
                    const float t = member == 0u ? mid : (member == 1u ? left : right);
                    const bool hit = active && member < 3u && overlap_at(w, B[lowBody], B[highBody], reach, t);
                    const int lane = (int)(threadIdx.x & 63u) & ~3;
                    const bool atMid = __shfl((int)hit, lane, 64) != 0, atLeft = __shfl((int)hit, lane + 1, 64) != 0,
                               atRight = __shfl((int)hit, lane + 2, 64) != 0;
                    if (atMid) {
                        high = mid;
                        if (atLeft) high = left;
                        else low = left;
                    } else {
                        low = mid;
                        if (atRight) high = right;
                        else low = right;
                    }
                }
                if (active && member == 0u) {
                    toi[e] = low;
                    if (twin[e] != kNoTwin) toi[twin[e]] = low;
                }
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < n; i += BLOCK) {
            float earliest = 2.0f;
            if (!immovable(B[i]))
                for (uint32_t s = 0; s < rowCount[i]; ++s) earliest = fminf(earliest, toi[rowStart[i] + s]);
            B[i].limits[2] = earliest <= 1.0f ? 1.0f + earliest : 0.0f;
        }
        __syncthreads();
        for (uint32_t i = tid; i < n; i += BLOCK) {
            bp_rigid_body &body = B[i];
            if (body.limits[2] >= 1.0f) {
                float t = body.limits[2] - 1.0f;
                body.pose[0] = mixs(body.previous[0], body.pose[0], t);
                body.pose[1] = mixs(body.previous[1], body.pose[1], t);
                body.pose[2] = mixs(body.pose[3], body.pose[2], t);
            }
            body.limits[2] = 0.0f;
        }
        __syncthreads();

        // The contact rounds' work list (in toiStart, free after the impact times): the entries that
        // are evaluated, grouped by the pair's shapes (two round, mixed, two polygons) so that the
        // lanes of a wavefront run the same contact code. Entries are independent of each other, so
        // their order changes no result.
        if (tid < 4) sKinds[tid] = 0u;
        __syncthreads();
        for (uint32_t e = tid; e < total; e += BLOCK) {
            const uint32_t index = slots[e], pair = twin[e];
            const RigidCand c = rows[index];
            uint32_t tag = 0xffffffffu;
            if (!(pair != kNoTwin && index / stride != c.low)) {
                const uint32_t kind = (uint32_t)is_round(B[c.low]) + (uint32_t)is_round(B[c.high]);
                tag = (kind << 24) | atomicAdd(&sKinds[kind], 1u);
            }
            toiSteps[e] = tag;
        }
        __syncthreads();
        const uint32_t work = sKinds[0] + sKinds[1] + sKinds[2];
        for (uint32_t e = tid; e < total; e += BLOCK) {
            const uint32_t tag = toiSteps[e];
            if (tag == 0xffffffffu) continue;
            const uint32_t kind = tag >> 24;
            const uint32_t base = kind == 0u ? 0u : (kind == 1u ? sKinds[0] : sKinds[0] + sKinds[1]);
            toiStart[base + (tag & 0xffffffu)] = e;
        }
        __syncthreads();

        for (uint32_t it = 0; it < D.iterations; ++it) {
            // Pair contacts (Contact:382-433). A pair sits in the rows of both its bodies, where the
            // reference evaluates it twice on the same inputs; here the lower body's entry is evaluated
            // once and written to both entries (an entry without a twin, when a row cut or one-sided
            // joint tables leave the pair in one row only, is evaluated on its own).
            for (uint32_t k = tid; k < work; k += BLOCK) {
                const uint32_t e = toiStart[k];
                const uint32_t index = slots[e], pair = twin[e];
                const RigidCand cand = rows[index];
                RigidCorr r;
                bool added;
                float lambda;
                evaluate_entry(w, B, cand, corr[index], r, added, lambda);
                corr[index] = r;
                if (added) rows[index].impulse = cand.impulse + lambda;
                if (pair != kNoTwin) {
                    const uint32_t other = slots[pair];
                    corr[other] = r;
                    if (added) rows[other].impulse = rows[other].impulse + lambda;
                }
            }
            __syncthreads();
            // Each body sums its row in slot order (Contact:435-449).
            for (uint32_t i = tid; i < n; i += BLOCK) {
                float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
                for (uint32_t s = 0; s < rowCount[i]; ++s) {
                    const size_t at = (size_t)i * stride + s;
                    const RigidCorr &c = corr[at];
                    const bool first = rows[at].kind == 3u && rows[at].low == i;
                    const float *dp = first ? c.first : c.second;
                    const float *vp = first ? c.vfirst : c.vsecond;
                    d0 += dp[0];
                    d1 += dp[1];
                    d2 += dp[2];
                    v0 += vp[0];
                    v1 += vp[1];
                    v2 += vp[2];
                }
                bp_rigid_body &b = B[i];
                b.pose[0] += d0;
                b.pose[1] += d1;
                b.pose[2] += d2;
                b.motion[0] += v0;
                b.motion[1] += v1;
                b.motion[2] += v2;
            }
            __syncthreads();
            if (D.jointCount) {
                rigid_solve_joints(B, joints, jcorr, D, tid, BLOCK, w);
                __syncthreads();
                for (uint32_t i = tid; i < n; i += BLOCK) {
                    const uint4 range = w.jranges[i];
                    if (range.y == 0u) continue;
                    float d[3] = {0.0f, 0.0f, 0.0f}, v[3] = {0.0f, 0.0f, 0.0f};
                    for (uint32_t k = 0; k < range.y; ++k) {
                        const uint2 ref = refs[range.x + k];
                        const RigidCorr &c = jcorr[ref.x];
                        const float *dp = ref.y ? c.second : c.first;
                        const float *vp = ref.y ? c.vsecond : c.vfirst;
                        for (int q = 0; q < 3; ++q) {
                            d[q] += dp[q];
                            v[q] += vp[q];
                        }
                    }
                    const float relax = 1.0f / float(range.y);
                    bp_rigid_body &b = B[i];
                    for (int q = 0; q < 3; ++q) {
                        b.pose[q] += d[q] * relax;
                        b.motion[q] += v[q] * relax;
                    }
                }
                __syncthreads();
            }
            // Wall guard of the dynamic bodies against static polygons (PostSolve:116-178).
            const uint32_t wallCount = sWalls;
            const bool listed = wallCount <= kGuardWalls;
            if (listed) {
                for (uint32_t q = tid; q < wallCount; q += BLOCK) {
                    const bp_rigid_body &wall = B[sWallIndex[q]];
                    const Xf xw = xf_of(wall);
                    float2 lo = f2(inff(), inff()), hi = f2(-inff(), -inff());
                    for (uint32_t e = 0; e < wall.ranges[1]; ++e) {
                        const float2 v = w.vertex(wall, xw, e);
                        lo = min2(lo, v);
                        hi = max2(hi, v);
                    }
                    sWallRot[q] = make_float4(xw.c, xw.s, bpm::cosf_det(-wall.pose[2]), bpm::sinf_det(-wall.pose[2]));
                    sWallBox[q] = make_float4(lo.x, lo.y, hi.x, hi.y);
                }
                __syncthreads();
            }
            for (uint32_t i = tid; i < n; i += BLOCK) {
                if (B[i].flags[0] || B[i].filter[2]) continue;
                bp_rigid_body b = B[i];
                // The guard moves a body, never turns it: the angle's rotation terms hold for the pass.
                const Xf now = xf_of(b);
                const Xf then = xf_at(b.previous[0], b.previous[1], b.pose[3]);
                // The body's vertices at both poses (the current one as earlier walls left it) lie in
                // the box below; the guard's tests reach at most 1e-5 beyond them.
                const float reach = body_reach(w, b);
                const float2 c0 = xy(b.previous);
                for (uint32_t q = 0; q < (listed ? wallCount : n); ++q) {
                    const uint32_t j = listed ? sWallIndex[q] : q;
                    const bp_rigid_body &wall = B[j];
                    if (!listed && (!wall.flags[0] || is_round(wall))) continue;
                    if (!accepts(b.filter, wall.filter)) continue;
                    if (w.excluded(i, wall.id)) continue;
                    Xf xw;
                    float2 lo, hi;
                    if (listed) {
                        const float4 rt = sWallRot[q], bx = sWallBox[q];
                        xw = Xf{xy(wall.pose), rt.x, rt.y, wall.pose[2]};
                        lo = f2(bx.x, bx.y);
                        hi = f2(bx.z, bx.w);
                    } else {
                        xw = xf_of(wall);
                        lo = f2(inff(), inff());
                        hi = f2(-inff(), -inff());
                        for (uint32_t e = 0; e < wall.ranges[1]; ++e) {
                            const float2 v = w.vertex(wall, xw, e);
                            lo = min2(lo, v);
                            hi = max2(hi, v);
                        }
                    }
                    {
                        const float2 c1 = xy(b.pose);
                        const float bodyMagnitude = fmaxf(magnitude2(c0), magnitude2(c1));
                        const float r = reach + prune_slack(reach, fmaxf(bodyMagnitude, fmaxf(magnitude2(lo), magnitude2(hi))));
                        const float4 box = make_float4(fminf(c0.x, c1.x) - r, fminf(c0.y, c1.y) - r, fmaxf(c0.x, c1.x) + r,
                                                       fmaxf(c0.y, c1.y) + r);
                        if (boxes_apart(box, make_float4(lo.x, lo.y, hi.x, hi.y))) continue;
                    }
                    const float wic = listed ? sWallRot[q].z : bpm::cosf_det(-wall.pose[2]);
                    const float wis = listed ? sWallRot[q].w : bpm::sinf_det(-wall.pose[2]);
                    const uint32_t edges = wall.ranges[1];
                    for (uint32_t k = 0; k < b.ranges[1]; ++k) {
                        const float2 p = apply(Xf{xy(b.pose), now.c, now.s, now.a}, w.verts[b.ranges[0] + k]);
                        const float2 start = apply(then, w.verts[b.ranges[0] + k]);
                        {
                            // A vertex whose path stays away from the wall's box is inside no part of the
                            // wall and meets none of its edges.
                            const float r = prune_slack(0.0f, fmaxf(fmaxf(magnitude2(start), magnitude2(p)),
                                                                    fmaxf(magnitude2(lo), magnitude2(hi))));
                            const float4 path = make_float4(fminf(start.x, p.x) - r, fminf(start.y, p.y) - r,
                                                            fmaxf(start.x, p.x) + r, fmaxf(start.y, p.y) + r);
                            if (boxes_apart(path, make_float4(lo.x, lo.y, hi.x, hi.y))) continue;
                        }
                        const bool in = inside_with(w, p, wall, xw, wic, wis);
                        bool startedIn = inside_with(w, start, wall, xw, wic, wis);
                        if (startedIn)
                            for (uint32_t e = 0; e < edges; ++e) {
                                float2 ea = w.vertex(wall, xw, e), eb = w.vertex(wall, xw, (e + 1) % edges);
                                if (dist2(start, closest_point(start, ea, eb)) < 1e-5f) {
                                    startedIn = false;
                                    break;
                                }
                            }
                        float nearest = inff(), depth = 0.0f;
                        float2 normal = f2(0.0f, 0.0f);
                        // A vertex outside the wall acts only through an edge that its path from start
                        // to p comes near (a crossing, or a start on the edge).
                        const float4 path = make_float4(fminf(start.x, p.x), fminf(start.y, p.y), fmaxf(start.x, p.x),
                                                        fmaxf(start.y, p.y));
                        const float pathMagnitude = fmaxf(magnitude2(start), magnitude2(p));
                        for (uint32_t e = 0; e < edges; ++e) {
                            const float2 ea = w.vertex(wall, xw, e), eb = w.vertex(wall, xw, (e + 1) % edges);
                            if (!in) {
                                const float r = prune_slack(0.0f, fmaxf(pathMagnitude, fmaxf(magnitude2(ea), magnitude2(eb))));
                                const float4 grown = make_float4(path.x - r, path.y - r, path.z + r, path.w + r);
                                if (boxes_apart(grown, make_float4(fminf(ea.x, eb.x), fminf(ea.y, eb.y), fmaxf(ea.x, eb.x),
                                                                   fmaxf(ea.y, eb.y))))
                                    continue;
                            }
                            const Sweep edge{ea, eb, ea, eb};
                            const float2 dir = sub(eb, ea), out = unit(f2(dir.y, -dir.x));
                            const float signedStart = dot2(sub(start, ea), out);
                            const bool onBoundary = dist2(start, closest_point(start, ea, eb)) < 1e-5f;
                            const bool crossing = signedStart >= -1e-5f &&
                                                  (crosses(start, p, edge) || (onBoundary && dot2(sub(p, ea), out) < 0.0f));
                            if (!in && !crossing) continue;
                            if (in && !startedIn && signedStart < -1e-5f) continue;
                            const float2 closest = closest_point(p, ea, eb);
                            const float d = (crossing || !startedIn) ? dist2(start, closest_point(start, ea, eb)) : dist2(p, closest);
                            if (d < nearest) {
                                nearest = d;
                                normal = out;
                                depth = fmaxf(0.0f, -dot2(sub(p, closest), out));
                            }
                        }
                        if (isfinite(nearest) && depth > 0.0f) {
                            if (is_round(b) && core_crosses(w, core(w, b, then), wall, xw)) break;
                            const float2 shift = mul(normal, depth + 1e-5f);
                            b.pose[0] += shift.x;
                            b.pose[1] += shift.y;
                            const float2 r = sub(p, xy(b.pose));
                            const float lever = cross2(r, normal);
                            const float speed = dot2(f2(b.motion[0], b.motion[1]), normal) + b.motion[2] * lever;
                            const float denom = b.physical[0] + b.physical[1] * lever * lever;
                            const float impulse = fmaxf(0.0f, -speed) / fmaxf(denom, kFloatMin);
                            const float2 dv = mul(mul(normal, b.physical[0]), impulse);
                            b.motion[0] += dv.x;
                            b.motion[1] += dv.y;
                            b.motion[2] += lever * b.physical[1] * impulse;
                        }
                    }
                }
                B[i] = b;
            }
            __syncthreads();
        }

        // Speed clamps and the region rule (PostSolve:180-206).
        for (uint32_t i = tid; i < n; i += BLOCK) {
            bp_rigid_body b = B[i];
            if (!finite4(b.pose) || !finite4(b.previous) || !finite4(b.motion)) {
                atomicOr(&sSafety, 1u);
                continue;
            }
            if (b.flags[0]) continue;
            float2 v = clamp_speed(f2(b.motion[0], b.motion[1]), b.limits[0]);
            b.motion[0] = v.x;
            b.motion[1] = v.y;
            b.motion[2] = clamp_spin(b.motion[2], b.limits[1]);
            float2 lo = f2(inff(), inff()), hi = f2(-inff(), -inff());
            const Xf xb = xf_of(b);
            for (uint32_t k = 0; k < b.ranges[1]; ++k) {
                float2 p = w.vertex(b, xb, k);
                lo = min2(lo, p);
                hi = max2(hi, p);
            }
            lo = sub(lo, f2(b.shape[1], b.shape[1]));
            hi = add(hi, f2(b.shape[1], b.shape[1]));
            const float *rg = D.region;
            const float2 size = f2(rg[2] - rg[0], rg[3] - rg[1]);
            float2 shift = f2(0.0f, 0.0f);
            if (hi.x < rg[0] || hi.y < rg[1] || lo.x > rg[2] || lo.y > rg[3]) {
                if (!b.flags[1]) {
                    b.flags[2] = 1;
                    atomicAdd(&sRemoved, 1u);
                } else {
                    if (hi.x < rg[0]) shift.x = ceilf((rg[0] - hi.x) / size.x) * size.x;
                    if (lo.x > rg[2]) shift.x = -ceilf((lo.x - rg[2]) / size.x) * size.x;
                    if (hi.y < rg[1]) shift.y = ceilf((rg[1] - hi.y) / size.y) * size.y;
                    if (lo.y > rg[3]) shift.y = -ceilf((lo.y - rg[3]) / size.y) * size.y;
                    b.pose[0] += shift.x;
                    b.pose[1] += shift.y;
                    b.previous[0] += shift.x;
                    b.previous[1] += shift.y;
                    b.flags[2] = 2;
                    atomicAdd(&sWrapped, 1u);
                }
            }
            B[i] = b;
        }
        __syncthreads();
        const uint32_t safety = sSafety;
        if constexpr (CAP > 0) {
            if (!(safety & 3u))
                for (uint32_t i = tid; i < n; i += BLOCK) out[i] = B[i];
            __syncthreads();
        }
        if (tid == 0) {
            bp_rigid_frame_info r = {};
            r.status = (safety & 1u) ? BP_ERR_NON_FINITE : ((safety & 2u) ? BP_ERR_GUARD_FAILURE : BP_OK);
            r.guard_cause = (!(safety & 1u) && (safety & 2u) && (safety & 16u)) ? BP_GUARD_RIGID_ROTATION_BUDGET : BP_GUARD_NONE;
            r.candidate_pairs = sPairs;
            r.candidate_overflow = sOverflow;
            r.max_candidates_per_object = sMaxPer;
            r.removed = sRemoved;
            r.wrapped = sWrapped;
            *info = r;
            if (safety & 3u) {
                D.failed = 1u;
                D.lastOk = 0u;
            } else {
                D.side ^= 1u;
                D.lastOk = 1u;
            }
        }
        __syncthreads();
    }
}

}  // namespace bpr

extern "C" uint32_t bp_rigid_lds_capacity() { return bpr::kRigidLdsBodies; }

extern "C" hipError_t bp_launch_rigid_step(const bpr::RigidArena *arena, bpr::RigidDesc *descs, const uint32_t *order,
                                           uint32_t count, int small, bp_rigid_frame_info *infos, uint32_t worlds,
                                           uint32_t steps, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    if (small)
        hipLaunchKernelGGL((bpr::rigid_step_kernel<64, bpr::kRigidLdsBodies>), dim3(count), dim3(64), 0, stream, *arena, descs,
                           order, infos, worlds, steps);
    else
        hipLaunchKernelGGL((bpr::rigid_step_kernel<256, 0>), dim3(count), dim3(256), 0, stream, *arena, descs, order, infos,
                           worlds, steps);
    return hipGetLastError();
}
