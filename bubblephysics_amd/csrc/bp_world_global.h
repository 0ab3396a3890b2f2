// bp_world_global.h -- a world whose state lives in global memory: its view, its adjacency rows
// and the substep passes, shared by the single-workgroup kernel (bp_world_kernel.hip) and the
// cooperative multi-workgroup kernel (bp_world_coop.hip).
//
// The arithmetic is that of the LDS-resident kernel and of the CPU oracle, operation for
// operation, so every path gives the same bits.
#ifndef BP_WORLD_GLOBAL_H
#define BP_WORLD_GLOBAL_H

#include "bp_common.h"

namespace bpd {

struct Range { uint32_t first, stride; };

// ---------------------------------------------------------------------------------------
// The world view of one workgroup (or of a cooperative grid of workgroups).
// ---------------------------------------------------------------------------------------
struct World {
    const Arena *A;
    const WorldDesc *d;
    uint32_t n, ns, nb, w;       // w: working side
    uint32_t o;                  // bubble offset
    float2 *pos, *start, *vel;
    float *rot, *omega;
    uint32_t *region;
    const int64_t *id;
    const float *invmRaw, *radius, *Kc, *gamma, *mass;
    const uint32_t *flags, *cat, *mask;
    const int32_t *motor, *band;
    float2 *acc, *a0, *ref;
    float *margin;
    uint32_t *rowStart, *rowLen, *grid, *adj;
    float4 *nbrA, *nbrB;         // packed neighbour reads (Arena::nbrA, nbrB)
    float *invmEff;
    float4 *cellA;               // cell-sorted records (Arena::cellA, cellB)
    uint4 *cellB;
    const float4 *motors, *bands;
    const float2 *verts;
    const uint4 *tris;
    float cellWidth;
    float2 globalForce;
    float sepTol;

    BPD bool is_static(uint32_t i) const {
        return !(invmRaw[i] > 0.0f) || ((flags[i] & kEnabled) && (flags[i] & kStatic));
    }
    BPD float inv_mass(uint32_t i) const { return is_static(i) ? 0.0f : invmRaw[i]; }
    BPD float2 eff_vel(uint32_t i) const { return is_static(i) ? f2(0.0f, 0.0f) : vel[i]; }
    // The packed records of bubble i: A after its centre moved, B after its velocity and
    // acceleration changed, and the step's effective inverse mass at the start of the step.
    BPD void publish_a(uint32_t i) const { const float2 p = pos[i]; nbrA[i] = make_float4(p.x, p.y, radius[i], Kc[i]); }
    BPD void publish_b(uint32_t i) const {
        const float2 v = eff_vel(i);
        nbrB[i] = make_float4(v.x, v.y, len2(acc[i]), 0.0f);
    }
    BPD void publish_step(uint32_t i) const { invmEff[i] = inv_mass(i); publish_a(i); }
    // Bubble i's record at its slot of the cell-sorted order (written where the counting sort
    // places it in `grid`).
    BPD void publish_cell(uint32_t slot, uint32_t i) const {
        const float2 p = pos[i];
        cellA[slot] = make_float4(p.x, p.y, radius[i], margin[i]);
        cellB[slot] = make_uint4(i, cat[i], mask[i],
                                 (finite2(start[i]) ? 1u : 0u) | (region[i] == BP_REGION_REMOVED ? 2u : 0u) |
                                     (inv_mass(i) > 0.0f ? 4u : 0u));
    }

    BPD Seg segment(uint32_t j) const {
        Seg s;
        float4 ab = A->segAB[d->segmentOffset + j];
        float4 rg = A->segRegion[d->segmentOffset + j];
        s.a = f2(ab.x, ab.y); s.b = f2(ab.z, ab.w);
        s.prevStart = f2(rg.x, rg.y);
        s.outward = rg.z;
        s.flags = A->segFlags[d->segmentOffset + j];
        s.base = segment_base(s.a, s.b);
        return s;
    }
    BPD BodyView body(uint32_t k) const {
        BodyView b;
        uint32_t g = d->bodyOffset + k;
        b.pose = A->bodyPose[w][g];
        b.prev = A->bodyPrev[w][g];
        b.motion = A->bodyMotion[g];
        b.shape = A->bodyShape[g];
        b.ranges = A->bodyRanges[g];
        b.kind = A->bodyKind[g];
        return b;
    }

    // Non-contact acceleration a0 at `p`, Force:450-470 (global force, band, motor).
    Geo g;                       // geometry and force records of the world

    BPD float2 non_contact(uint32_t i, float2 p) const { return g.non_contact(flags[i], inv_mass(i), id + i, band[i], motor[i], p); }

    BPD int64_t cell_x(float2 p) const { return (int64_t)floorf(p.x / cellWidth); }
    BPD int64_t cell_y(float2 p) const { return (int64_t)floorf(p.y / cellWidth); }

    // Pair predicate of Geometry:262-273, evaluated with the lower index as the row owner.
    BPD bool pair_candidate(uint32_t lo, uint32_t hi) const {
        if (!filter_accepts(cat[lo], mask[lo], cat[hi], mask[hi])) return false;
        float reach = radius[lo] + radius[hi] + margin[lo] + margin[hi] + sepTol;
        float2 a = pos[lo], b = pos[hi];
        return fabsf(a.x - b.x) <= reach && fabsf(a.y - b.y) <= reach;
    }
    // Segment predicate, Geometry:279-290.
    BPD bool segment_candidate(uint32_t i, uint32_t j) const {
        uint32_t g = d->segmentOffset + j;
        if (!filter_accepts(cat[i], mask[i], A->segCat[g], A->segMask[g])) return false;
        float4 ab = A->segAB[g];
        float reach = radius[i] + margin[i] + sepTol;
        float2 lower = subs(min2(f2(ab.x, ab.y), f2(ab.z, ab.w)), reach);
        float2 upper = adds(max2(f2(ab.x, ab.y), f2(ab.z, ab.w)), reach);
        float2 p = pos[i];
        return p.x >= lower.x && p.y >= lower.y && p.x <= upper.x && p.y <= upper.y;
    }
    // Body predicate, Geometry:361-379.
    BPD bool body_candidate(uint32_t i, uint32_t k) const {
        uint32_t g = d->bodyOffset + k;
        uint4 kind = A->bodyKind[g];
        bool isStatic = (kind.y & kBodyStatic) != 0;
        bool obstacle = isStatic ? (kind.x != 0u) : ((kind.y & kBodyKinematic) != 0u);
        if (!obstacle) return false;
        if (!filter_accepts(cat[i], mask[i], kind.z, kind.w)) return false;
        float4 pose = A->bodyPose[w][g];
        float2 prev = A->bodyPrev[w][g];
        float br = A->bodyShape[g].x;
        float2 bodyLow = subs(min2(f2(pose.x, pose.y), prev), br);
        float2 bodyHigh = adds(max2(f2(pose.x, pose.y), prev), br);
        float reach = radius[i] + fmaxf(margin[i], 0.0f);
        float2 p = pos[i];
        float2 lo = subs(p, reach), hi = adds(p, reach);
        return !(hi.x < bodyLow.x || hi.y < bodyLow.y || bodyHigh.x < lo.x || bodyHigh.y < lo.y);
    }
};

// Bodies of bubble i by brute force over the world's bodies, in ascending body index.
struct AllBodies {
    template <class Emit> BPD void operator()(const World &W, uint32_t i, Emit emit) const {
        for (uint32_t k = 0; k < W.nb; ++k)
            if (W.body_candidate(i, k)) emit((2u << kKindShift) | k);
    }
};

// Enumerates the adjacency row of bubble i: pairs (both directions) from the 3x3 cells of
// the hashed grid, then segments, then bodies (found by `bodies`). `emit(entry)` receives packed
// entries; `forward` counts the reference's candidate-row entries (pairs j > i and segments).
template <class Emit, class Bodies = AllBodies>
BPD void enumerate_row(const World &W, const uint32_t *ends, uint32_t buckets, uint32_t i, Emit emit, uint32_t &forward,
                       Bodies bodies = Bodies()) {
    forward = 0;
    float2 p = W.pos[i];
    if (!finite2(p)) {                                        // not in the grid: no pair, no segment row;
        bodies(W, i, emit);                                   // bodies compare NaN like the reference
        return;
    }
    bool ownRow = finite2(W.start[i]);                        // Geometry:305-307
    const float ri = W.radius[i], mi = W.margin[i];
    const uint32_t ci = W.cat[i], ki = W.mask[i];
    int64_t cx = W.cell_x(p), cy = W.cell_y(p);
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
            int64_t tx = cx + dx, ty = cy + dy;
            uint32_t b = bucket_of(tx, ty, buckets);
            uint32_t lo = b == 0 ? 0u : ends[b - 1], hi = ends[b];
            for (uint32_t k = lo; k < hi; ++k) {
                // The cell-sorted record of slot k: W.pair_candidate's operands, read in one go.
                const uint4 cb = W.cellB[k];
                const uint32_t j = cb.x;
                if (j == i) continue;
                const float4 ca = W.cellA[k];
                const float2 q = f2(ca.x, ca.y);
                if (W.cell_x(q) != tx || W.cell_y(q) != ty) continue;
                if (j > i) {
                    if (!ownRow) continue;
                    if (!filter_accepts(ci, ki, cb.y, cb.z)) continue;               // pair_candidate(i, j)
                    const float reach = ri + ca.z + mi + ca.w + W.sepTol;
                    if (!(fabsf(p.x - q.x) <= reach && fabsf(p.y - q.y) <= reach)) continue;
                    ++forward;
                } else {
                    if (!(cb.w & 1u)) continue;                                      // step start not finite
                    if (!filter_accepts(cb.y, cb.z, ci, ki)) continue;               // pair_candidate(j, i)
                    const float reach = ca.z + ri + ca.w + mi + W.sepTol;
                    if (!(fabsf(q.x - p.x) <= reach && fabsf(q.y - p.y) <= reach)) continue;
                }
                emit(j);
            }
        }
    if (ownRow)
        for (uint32_t j = 0; j < W.ns; ++j)
            if (W.segment_candidate(i, j)) { ++forward; emit((1u << kKindShift) | j); }
    bodies(W, i, emit);
}

// The view of a world's working side (`wside`), over the arena.
BPD World make_world(const Arena *A, const WorldDesc *d, uint32_t wside) {
    World W;
    W.A = A; W.d = d; W.n = d->bubbleCount; W.ns = d->segmentCount; W.nb = d->bodyCount; W.w = wside;
    const uint32_t o = d->bubbleOffset;
    W.o = o;
    W.pos = A->pos[wside] + o; W.start = A->start[wside] + o; W.vel = A->vel[wside] + o;
    W.rot = A->rot[wside] + o; W.omega = A->omega[wside] + o; W.region = A->region[wside] + o;
    W.id = A->id + o; W.invmRaw = A->invm + o; W.radius = A->radius + o; W.Kc = A->Kc + o; W.gamma = A->gamma + o;
    W.mass = A->mass + o; W.flags = A->flags + o; W.cat = A->cat + o; W.mask = A->mask + o;
    W.motor = A->motor + o; W.band = A->band + o;
    W.acc = A->acc + o; W.a0 = A->a0 + o; W.ref = A->ref + o; W.margin = A->margin + o;
    W.rowStart = A->rowStart + o; W.rowLen = A->rowLen + o; W.grid = A->grid + o; W.adj = A->adj + d->adjOffset;
    W.nbrA = A->nbrA + o; W.nbrB = A->nbrB + o; W.invmEff = A->invmEff + o;
    W.cellA = A->cellA + o; W.cellB = A->cellB + o;
    W.motors = A->motors + d->motorOffset; W.bands = A->bands + d->bandOffset;
    W.verts = A->verts + d->vertexOffset; W.tris = A->tris + d->triangleOffset;
    W.cellWidth = 0.0f;
    W.globalForce = f2(d->globalForce[0], d->globalForce[1]);
    W.g = make_geo(A, d, wside);
    W.sepTol = d->sepTol;
    return W;
}

// ---------------------------------------------------------------------------------------
// Substep passes over the bubbles r.first, r.first + r.stride, ...: one workgroup covers a
// world with (threadIdx.x, BLOCK); a cooperative grid with (global thread, grid threads). The
// reductions end at the workgroup; a cooperative caller combines the workgroup results.
// ---------------------------------------------------------------------------------------

// F: contact forces, phase 1 (or phase 2 after a grid build), omega^2 and the margin bound.
template <int BLOCK>
BPD void pass_forces(World &W, Red<BLOCK> &red, Range r, const Ctl &c, uint32_t phase, float &w2Max, uint32_t &w2Owner,
                     float &mbMin, uint32_t &mbOwner) {
    const uint32_t n = W.n;
    float fraction = c.t0 > 0.0f ? clampf(c.t1 / c.t0, 0.0f, 1.0f) : 1.0f;
    float h = c.h, halfStep = 0.5f * h;
    float bestW = -1.0f, bestM = inff();
    uint32_t ownW = 0, ownM = 0;
    for (uint32_t i = r.first; i < n; i += r.stride) {
        float2 p = W.pos[i];
        float invm = W.inv_mass(i);
        float ri = W.radius[i], K = W.Kc[i];
        float2 total = f2(0.0f, 0.0f);
        float w2 = 0.0f;
        if (invm > 0.0f && W.band[i] >= 0) w2 = w2 + W.bands[W.band[i]].z * invm;
        const uint32_t *row = W.adj + W.rowStart[i];
        uint32_t len = W.rowLen[i];
        for (uint32_t e = 0; e < len; ++e) {
            uint32_t entry = row[e], kind = entry >> kKindShift, j = entry & kIndexMask;
            if (kind == 0u) {
                const float4 q = W.nbrA[j];
                const float invmJ = W.invmEff[j];
                const bool low = i < j;
                const float2 delta = low ? sub(f2(q.x, q.y), p) : sub(p, f2(q.x, q.y));
                float dd = len2(delta);
                float2 normal = dd > 0.0f ? dvs(delta, dd) : f2(1.0f, 0.0f);
                Law law = low ? pair_law(K, ri, q.w, q.z, dd) : pair_law(q.w, q.z, K, ri, dd);
                total = add(total, mul(low ? -law.force : law.force, normal));
                if (invm > 0.0f) w2 = w2 + law.stiffness * (invm + __builtin_sqrtf(invm * invmJ));
            } else if (kind == 1u) {
                float2 normal; float dd;
                if (!wall_region(W.segment(j), p, normal, dd) || !(dd < ri)) continue;
                Law law = wall_law(K, ri, dd);
                total = add(total, mul(law.force, normal));
                if (invm > 0.0f) w2 = w2 + law.stiffness * invm;
            } else {
                Obstacle ob = body_obstacle(W.body(j), fraction, W.verts, W.tris, p);
                Law law = wall_law(K, ri, ob.distanceValue);
                total = add(total, mul(law.force, ob.normal));
                if (invm > 0.0f) w2 = w2 + law.stiffness * invm;
            }
        }
        float2 v = W.vel[i];
        float2 a;
        if (phase == 2) {
            a = add(mul(total, invm), W.a0[i]);
            W.acc[i] = a;
        } else {
            a = W.acc[i];
            if (invm > 0.0f) {
                float gamma = (W.flags[i] & kForceDriven) ? fmaxf(W.gamma[i], 0.0f) : 0.0f;
                float decay = bpm::expf_det(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
                a = add(mul(total, invm), W.non_contact(i, p));
                v = add(mul(v, decay), mul(a, factor));
                W.acc[i] = a;
                W.vel[i] = v;
            }
        }
        W.nbrB[i] = make_float4(invm > 0.0f ? v.x : 0.0f, invm > 0.0f ? v.y : 0.0f, len2(a), 0.0f);
        // omega^2 (Force:510-541) and margin bound (Force:604-614) for the next reduction.
        if (w2 > bestW) { bestW = w2; ownW = i; }
        float mb;
        if (!(invm > 0.0f)) mb = inff();
        else {
            float left = W.margin[i] - len2(sub(p, W.ref[i]));
            if (isnan(left)) mb = inff();
            else if (left <= 0.0f) mb = 0.0f;
            else {
                float speed = len2(v), acl = len2(a);
                float den = speed + __builtin_sqrtf(speed * speed + 2.0f * acl * left);
                mb = den > 0.0f ? 2.0f * left / den : inff();
            }
        }
        if (mb < bestM) { bestM = mb; ownM = i; }
    }
    block_maxmin<BLOCK>(red, bestW, ownW, bestM, ownM);
    w2Max = bestW; w2Owner = ownW; mbMin = bestM; mbOwner = ownM;
}

// omega^2 and margin bound alone (after a diagnostic state injection moved a centre).
template <int BLOCK>
BPD void pass_bounds(World &W, Red<BLOCK> &red, Range r, const Ctl &c, float &w2Max, uint32_t &w2Owner, float &mbMin,
                     uint32_t &mbOwner) {
    const uint32_t n = W.n;
    float fraction = c.t0 > 0.0f ? clampf(c.t1 / c.t0, 0.0f, 1.0f) : 1.0f;
    float bestW = -1.0f, bestM = inff();
    uint32_t ownW = 0, ownM = 0;
    for (uint32_t i = r.first; i < n; i += r.stride) {
        float2 p = W.pos[i];
        float invm = W.inv_mass(i);
        float ri = W.radius[i], K = W.Kc[i];
        float w2 = 0.0f;
        if (invm > 0.0f) {
            if (W.band[i] >= 0) w2 = w2 + W.bands[W.band[i]].z * invm;
            const uint32_t *row = W.adj + W.rowStart[i];
            uint32_t len = W.rowLen[i];
            for (uint32_t e = 0; e < len; ++e) {
                uint32_t entry = row[e], kind = entry >> kKindShift, j = entry & kIndexMask;
                if (kind == 0u) {
                    const float4 q = W.nbrA[j];
                    const bool low = i < j;
                    float dd = len2(low ? sub(f2(q.x, q.y), p) : sub(p, f2(q.x, q.y)));
                    Law law = low ? pair_law(K, ri, q.w, q.z, dd) : pair_law(q.w, q.z, K, ri, dd);
                    w2 = w2 + law.stiffness * (invm + __builtin_sqrtf(invm * W.invmEff[j]));
                } else if (kind == 1u) {
                    float2 normal; float dd;
                    if (!wall_region(W.segment(j), p, normal, dd) || !(dd < ri)) continue;
                    w2 = w2 + wall_law(K, ri, dd).stiffness * invm;
                } else {
                    Obstacle ob = body_obstacle(W.body(j), fraction, W.verts, W.tris, p);
                    w2 = w2 + wall_law(K, ri, ob.distanceValue).stiffness * invm;
                }
            }
        }
        if (w2 > bestW) { bestW = w2; ownW = i; }
        float mb;
        if (!(invm > 0.0f)) mb = inff();
        else {
            float left = W.margin[i] - len2(sub(p, W.ref[i]));
            if (isnan(left)) mb = inff();
            else if (left <= 0.0f) mb = 0.0f;
            else {
                float speed = len2(W.vel[i]), acl = len2(W.acc[i]);
                float den = speed + __builtin_sqrtf(speed * speed + 2.0f * acl * left);
                mb = den > 0.0f ? 2.0f * left / den : inff();
            }
        }
        if (mb < bestM) { bestM = mb; ownM = i; }
    }
    block_maxmin<BLOCK>(red, bestW, ownW, bestM, ownM);
    w2Max = bestW; w2Owner = ownW; mbMin = bestM; mbOwner = ownM;
}

// H: the d/4 bound (Force:550-598).
template <int BLOCK>
BPD void pass_distance(World &W, Red<BLOCK> &red, Range r, const Ctl &c, float hMax, float &dbMin, uint32_t &dbOwner) {
    const uint32_t n = W.n;
    float fraction = c.t0 > 0.0f ? clampf(c.t1 / c.t0, 0.0f, 1.0f) : 1.0f;
    float bestD = inff();
    uint32_t ownD = 0;
    for (uint32_t i = r.first; i < n; i += r.stride) {
        if (!(W.inv_mass(i) > 0.0f)) continue;                // inf never replaces the initial best
        float2 p = W.pos[i], v = W.vel[i];
        float ri = W.radius[i];
        float accI = len2(W.acc[i]);
        float best = inff();
        const uint32_t *row = W.adj + W.rowStart[i];
        uint32_t len = W.rowLen[i];
        for (uint32_t e = 0; e < len; ++e) {
            uint32_t entry = row[e], kind = entry >> kKindShift, j = entry & kIndexMask;
            float dd, approach, accSum, reach, bodySpeed = 0.0f;
            if (kind == 0u) {
                const float4 q = W.nbrA[j], qb = W.nbrB[j];
                float2 delta = sub(f2(q.x, q.y), p);
                dd = len2(delta);
                float2 normal = dd > 0.0f ? dvs(delta, dd) : f2(1.0f, 0.0f);
                approach = dot2(sub(v, f2(qb.x, qb.y)), normal);
                accSum = accI + qb.z;
                reach = ri + q.z;
            } else if (kind == 1u) {
                float2 normal;
                if (!wall_region(W.segment(j), p, normal, dd)) continue;
                approach = -dot2(v, normal);
                accSum = accI;
                reach = ri;
            } else {
                Obstacle ob = body_obstacle(W.body(j), fraction, W.verts, W.tris, p);
                dd = ob.distanceValue;
                approach = -dot2(v, ob.normal);
                bodySpeed = ob.speedBound;
                accSum = accI;
                reach = ri;
            }
            bool active = dd < reach;
            float dc = fmaxf(dd, reach * 1e-6f);
            float speed;
            if (active) speed = fabsf(approach) + bodySpeed;
            else {
                if (approach + bodySpeed + accSum * hMax * 0.5f <= 0.0f) continue;
                speed = approach + bodySpeed;
            }
            float s = __builtin_sqrtf(speed * speed + accSum * dc * 0.5f);
            float den = s + speed;
            if (!(den > 0.0f)) continue;
            best = fminf(best, dc / (2.0f * den));
        }
        if (best < bestD) { bestD = best; ownD = i; }
    }
    block_min<BLOCK>(red, bestD, ownD);
    dbMin = bestD; dbOwner = ownD;
}

// D: kick and drift, phase 0 (Force:775-791).
template <int BLOCK>
BPD void pass_drift(World &W, Range r, const Ctl &c) {
    const uint32_t n = W.n;
    float h = c.h, halfStep = 0.5f * h;
    for (uint32_t i = r.first; i < n; i += r.stride) {
        float invm = W.inv_mass(i);
        if (!(invm > 0.0f)) continue;
        bool driven = (W.flags[i] & kForceDriven) != 0;
        float gamma = driven ? fmaxf(W.gamma[i], 0.0f) : 0.0f;
        float decay = bpm::expf_det(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
        float2 v = add(mul(W.vel[i], decay), mul(W.acc[i], factor));
        float om = W.omega[i];
        float angular = om * bpm::expf_det(-gamma * h);
        int32_t mi = W.motor[i];
        if (driven && mi >= 0) {
            float4 md = W.motors[mi];
            if (md.w > 0.0f) angular = angular + md.z / md.w * half_impulse_factor(gamma, h);
        }
        W.pos[i] = add(W.pos[i], mul(v, h));
        W.vel[i] = v;
        W.publish_a(i);
        W.rot[i] = W.rot[i] + om * h;
        W.omega[i] = angular;
    }
}

}  // namespace bpd

#endif  // BP_WORLD_GLOBAL_H
