// bp_world_lds.h -- LDS-resident world step: one workgroup, one world of up to BLOCK*PER bubbles.
//
// The MI355X form of the reference's single-dispatch kernel (NativeForceKernels.metal-source:
// 1236-1307), for worlds whose neighbour-visible state fits the 160 KB of LDS of a CU:
//   * each lane owns PER bubbles (index lane + s*BLOCK) and keeps their position, velocity,
//     acceleration, margin, x0, rotation and flags in registers for the whole step;
//   * what neighbours read lives in LDS as structure-of-arrays: centre, effective velocity,
//     |a| (the margin during a grid build), radius, K_c and effective inverse mass;
//   * the hashed grid (buckets and a u16 index list) and the adjacency rows (u16 entries,
//     kind << 14 | index, ordered by (kind, neighbour)) are built in LDS; rows that do not fit
//     the LDS budget of the configuration spill to the world's global adjacency area;
//   * segments, bodies, polygon vertices and triangles are staged in LDS once per step; each
//     body's pose at the current step fraction (position and rotation terms) is evaluated once
//     per substep, not once per contact;
//   * a substep is three passes (kick+drift; forces+kick+omega^2+margin bound; d/4 bound)
//     separated by workgroup barriers. Each row is walked in three runs (pairs, segments,
//     bodies), so a wavefront's lanes never wait on another contact kind's arithmetic; global
//     memory is touched only by grid builds and the final write-back of the state.
// Every quantity is computed with the same operations, in the same order, as the global-state
// kernel and the CPU oracle, so all three give the same bits.
#ifndef BP_WORLD_LDS_H
#define BP_WORLD_LDS_H

#include "bp_common.h"

namespace bpd {

constexpr uint32_t kLdsKindShift = 14;
constexpr uint32_t kLdsIndexMask = (1u << 14) - 1;
constexpr uint32_t kLdsIndexLimit = 1u << 14;

template <int BLOCK, int PER, int ROWCAP, int SEGCAP, int BODYCAP, int VERTCAP, int TRICAP>
struct LdsWorld {
    static constexpr uint32_t N = BLOCK * PER;
    static constexpr uint32_t NB = N < 64 ? 64 : (N > 2048 ? 2048 : N);
    float2 pos[N];
    float2 vel[N];     // effective velocity (zero for static bubbles)
    float aux[N];      // margin during a grid build, |a| during the substeps
    float r[N];
    float K[N];
    float invm[N];     // effective inverse mass (zero for static bubbles)
    uint32_t buckets[NB];
    uint16_t grid[N];
    uint16_t rows[ROWCAP > 0 ? ROWCAP : 2];
    // Segments (static within a step) with their unit normals.
    float4 segAB[SEGCAP];
    float4 segRegion[SEGCAP];     // previous edge start, unit normal
    float segOutward[SEGCAP];
    uint32_t segFlags[SEGCAP], segCat[SEGCAP], segMask[SEGCAP];
    // Bodies after the step's prediction, and their frame at the current step fraction.
    float4 bodyPose[BODYCAP];     // x, y, angle, previous angle
    float4 bodyExtent[BODYCAP];   // previous x, previous y, bounding radius, core radius
    float4 bodyMotion[BODYCAP];
    uint4 bodyRanges[BODYCAP];
    uint4 bodyKind[BODYCAP];
    float2 frameAt[BODYCAP];
    float4 frameCs[BODYCAP];
    float2 verts[VERTCAP];
    uint32_t tris[TRICAP];        // corner indices, 10 bits each
    Red<BLOCK> red;
};

struct RowRef {
    const uint16_t *lds;
    const uint32_t *global;
    bool inLds;
    BPD void get(uint32_t k, uint32_t &kind, uint32_t &j) const {
        if (inLds) { uint32_t e = lds[k]; kind = e >> kLdsKindShift; j = e & kLdsIndexMask; }
        else { uint32_t e = global[k]; kind = e >> kKindShift; j = e & kIndexMask; }
    }
    BPD uint32_t index(uint32_t k) const { return inLds ? (uint32_t)lds[k] & kLdsIndexMask : global[k] & kIndexMask; }
};

constexpr uint32_t kHasMotor = 1u << 31, kHasTorque = 1u << 30, kHasBand = 1u << 29;   // register-only flag bits


// All the steps of world `wi` by one workgroup, with the world's LDS state in S.
template <int BLOCK, int PER, int ROWCAP, int SEGCAP, int BODYCAP, int VERTCAP, int TRICAP>
__device__ __attribute__((always_inline)) void lds_world_steps(LdsWorld<BLOCK, PER, ROWCAP, SEGCAP, BODYCAP, VERTCAP, TRICAP> &S,
                                                               const Arena *__restrict__ A, WorldDesc *descs,
                                                               WorldStatus *stats, const uint32_t wi, uint32_t steps,
                                                               uint32_t statsStride, uint32_t notify) {
    using L = LdsWorld<BLOCK, PER, ROWCAP, SEGCAP, BODYCAP, VERTCAP, TRICAP>;
    // Row entries, polygon edges and triangles evaluated side by side: QB at a time where a lane
    // has 256 registers (workgroups of at most 512 threads), one at a time in the 1,024-thread
    // configurations, whose 128 registers per lane leave no room for more in flight.
    constexpr int QB = BLOCK <= 512 ? 4 : 1;
    WorldDesc *d = descs + wi;
    WorldStatus *st = stats + wi;
    const uint32_t lane = threadIdx.x;
    if (d->failed) {
        if (lane == 0) { st->executed = 0; post_status(st, notify); }
        return;
    }
    if (d->segmentCount > SEGCAP || d->bodyCount > BODYCAP || d->vertexCount > VERTCAP || d->triangleCount > TRICAP ||
        d->bubbleCount > L::N) {
        // The host never routes such a world here; refuse it rather than overrun LDS.
        if (lane == 0) { st->status = BP_ERR_COMMAND_FAILED; st->substeps = 0; st->executed = 1; post_status(st, notify); }
        return;
    }
    uint32_t pubNow = d->pub;                                 // published side, tracked in registers
    for (uint32_t stepIndex = 0; stepIndex < steps; ++stepIndex) {
        st = stats + (size_t)stepIndex * statsStride + wi;
        const uint64_t ticks0 = wall_clock64();
        const uint32_t pub = pubNow, ws = 1u - pub;
        const uint32_t n = d->bubbleCount, o = d->bubbleOffset;
        const Geo g = make_geo(A, d, ws);
#define ids (A->id + o)
#define catG (A->cat + o)
#define maskG (A->mask + o)
#define startG (A->start[ws] + o)
        float *history = A->history + d->historyOffset;
        const float sepTol = d->sepTol;
        const uint32_t limit = d->limit;
        const uint32_t buckets = min(d->buckets, L::NB);

        // ---- 1. Begin step (Force:285-293): own state into registers, neighbour fields into LDS.
        // Drag, motor and band records stay in memory and are read, for the bubbles whose flags
        // say they have them, where they are used: registers hold only what every bubble needs.
        // The row bounds of a bubble are its start, its length and, packed in one word, the
        // lengths of its pair block and of its pair and segment blocks (rows are below 2^16).
        float2 p[PER], v[PER], a[PER], ref[PER];
        float rot[PER], om[PER], mg[PER];
        uint32_t fl[PER], rs[PER], rl[PER], rpq[PER];
        bool valid[PER];
    #pragma unroll
        for (int s = 0; s < PER; ++s) {
            uint32_t i = lane + s * BLOCK;
            valid[s] = i < n;
            p[s] = v[s] = a[s] = ref[s] = f2(0.0f, 0.0f);
            rot[s] = om[s] = mg[s] = 0.0f;
            fl[s] = 0; rs[s] = rl[s] = rpq[s] = 0;
            if (!valid[s]) continue;
            uint32_t gi = o + i;
            p[s] = A->pos[pub][gi];
            v[s] = A->vel[pub][gi];
            rot[s] = A->rot[pub][gi];
            om[s] = A->omega[pub][gi];
            fl[s] = A->flags[gi];
            float invmRaw = A->invm[gi];
            bool isStatic = !(invmRaw > 0.0f) || ((fl[s] & kEnabled) && (fl[s] & kStatic));
            if ((fl[s] & kEnabled) && (fl[s] & kStatic)) { v[s] = f2(0.0f, 0.0f); om[s] = 0.0f; }
            startG[i] = p[s];
            S.pos[i] = p[s];
            S.vel[i] = isStatic ? f2(0.0f, 0.0f) : v[s];
            S.r[i] = A->radius[gi];
            S.K[i] = A->Kc[gi];
            S.invm[i] = isStatic ? 0.0f : invmRaw;
            // Force records (Force:423-438, resolved to indices on the host) and drag, per step.
            const int32_t mi = A->motor[gi];
            const bool driven = (fl[s] & kForceDriven) != 0;
            fl[s] &= ~(kHasMotor | kHasTorque | kHasBand);
            if (A->band[gi] >= 0) fl[s] |= kHasBand;
            if (mi >= 0) {
                const float4 md = g.motor(mi);
                fl[s] |= kHasMotor;
                if (driven && md.w > 0.0f) fl[s] |= kHasTorque;
            }
        }
        // World-uniform shortcuts of the candidate tests: when every bubble has a nonzero category
        // and a full mask, every pair passes the filter test; when every start is finite, every
        // neighbour's row exists. Either lets candidate enumeration skip the per-candidate global
        // loads of categories, masks and starts (the starts do not change until the end of step).
        // Likewise, when no force-driven bubble has drag (its drag as the step applies it,
        // max(gamma, 0), is +0), the drag is +0 for every bubble and is not read per substep.
        uint32_t localAccept = 1u, localStarts = 1u, localNoDrag = 1u;
    #pragma unroll
        for (int s = 0; s < PER; ++s) {
            if (!valid[s]) continue;
            const uint32_t gi = o + lane + s * BLOCK;
            if (A->cat[gi] == 0u || A->mask[gi] != 0xffffffffu) localAccept = 0u;
            if (!finite2(p[s])) localStarts = 0u;
            if ((fl[s] & kForceDriven) && __float_as_uint(fmaxf(A->gamma[gi], 0.0f)) != 0u) localNoDrag = 0u;
        }
        const bool allAccept = __syncthreads_and((int)localAccept) != 0;
        const bool allStarts = __syncthreads_and((int)localStarts) != 0;
        const bool allNoDrag = __syncthreads_and((int)localNoDrag) != 0;
        uint32_t safety = predict_bodies<BLOCK>(A, d, pub, ws);
        // Geometry into LDS. Bodies are read back by the lane that predicted them.
        for (uint32_t k = lane; k < g.nb; k += BLOCK) {
            const BodyView b = g.body(k);
            S.bodyPose[k] = b.pose;
            S.bodyExtent[k] = make_float4(b.prev.x, b.prev.y, b.shape.x, b.shape.y);
            S.bodyMotion[k] = b.motion;
            S.bodyRanges[k] = b.ranges;
            S.bodyKind[k] = b.kind;
        }
        for (uint32_t j = lane; j < g.ns; j += BLOCK) {
            const Seg sg = g.segment(j);
            S.segAB[j] = make_float4(sg.a.x, sg.a.y, sg.b.x, sg.b.y);
            S.segRegion[j] = make_float4(sg.prevStart.x, sg.prevStart.y, sg.base.x, sg.base.y);
            S.segOutward[j] = sg.outward;
            S.segFlags[j] = sg.flags;
            S.segCat[j] = g.segCat(j);
            S.segMask[j] = g.segMask(j);
        }
        for (uint32_t k = lane; k < d->vertexCount; k += BLOCK) S.verts[k] = A->verts[d->vertexOffset + k];
        for (uint32_t k = lane; k < d->triangleCount; k += BLOCK) {
            const uint4 t = A->tris[d->triangleOffset + k];
            S.tris[k] = t.x | (t.y << 10) | (t.z << 20);
        }
        __syncthreads();

        Ctl c;
        c.t0 = d->dt; c.t1 = n == 0 ? d->dt : 0.0f; c.h = 0.0f; c.split = (float)(d->split < 1 ? 1u : d->split);
        c.c0 = c.c1 = c.rebuild = c.limit = 0; c.limOwner = kNoOwner; c.limTerm = BP_TERM_REMAINING;
        uint32_t ov0 = 0, ov1 = 0, ov2 = 0, ov3 = 0, pairs = 0, entriesTotal = 0;
        int32_t status = BP_OK;
        bool claimed = false;
        float cellWidth = 0.0f;
        RowRef rows;
        rows.lds = S.rows;
        rows.global = A->adj + d->adjOffset;
        float4 *const geoG = A->adjGeo + d->adjOffset;
        rows.inLds = true;
        const bool ldsRowsAllowed = ROWCAP > 0 && g.ns < kLdsIndexLimit && g.nb < kLdsIndexLimit;

        auto segment_at = [&](uint32_t j) __attribute__((always_inline)) -> Seg {
            Seg sg;
            const float4 ab = S.segAB[j], rg = S.segRegion[j];
            sg.a = f2(ab.x, ab.y); sg.b = f2(ab.z, ab.w);
            sg.prevStart = f2(rg.x, rg.y);
            sg.base = f2(rg.z, rg.w);
            sg.outward = S.segOutward[j];
            sg.flags = S.segFlags[j];
            return sg;
        };
        // Body frames at the step fraction of the current pass, evaluated once per fraction.
        float framesAt = -1.0f;
        auto refresh_frames = [&](float fraction) __attribute__((always_inline)) {
            if (g.nb == 0 || fraction == framesAt) return;
            __syncthreads();
            for (uint32_t k = lane; k < g.nb; k += BLOCK) {
                const float4 ex = S.bodyExtent[k];
                float2 at;
                float4 cs;
                frame_pose(S.bodyPose[k], f2(ex.x, ex.y), fraction, at, cs);
                S.frameAt[k] = at;
                S.frameCs[k] = cs;
            }
            __syncthreads();
            framesAt = fraction;
        };
        auto frame_at = [&](uint32_t k) __attribute__((always_inline)) -> BodyFrame {
            BodyFrame f;
            const float4 cs = S.frameCs[k];
            f.at = S.frameAt[k];
            f.c = cs.x; f.s = cs.y; f.cn = cs.z; f.sn = cs.w;
            f.core = S.bodyExtent[k].w;
            f.motion = S.bodyMotion[k];
            f.ranges = S.bodyRanges[k];
            f.round = S.bodyKind[k].x != 0u;
            return f;
        };
        // The per-bubble records read on demand (drag only for force-driven bubbles, as the step
        // applies it; the motor's torque term is (torque / inertia), computed as at the start).
        auto gamma_of = [&](int s) __attribute__((always_inline)) -> float {
            if (allNoDrag) return 0.0f;
            return (fl[s] & kForceDriven) ? fmaxf(A->gamma[o + lane + s * BLOCK], 0.0f) : 0.0f;
        };
        auto motor_of = [&](int s) __attribute__((always_inline)) -> float4 { return g.motor(A->motor[o + lane + s * BLOCK]); };
        auto band_of = [&](int s) __attribute__((always_inline)) -> int32_t {
            return (fl[s] & kHasBand) ? A->band[o + lane + s * BLOCK] : -1;
        };
        auto a0_of = [&](int s, float invm, float2 pi) __attribute__((always_inline)) -> float2 {
            if (!(fl[s] & kForceDriven) || !(invm > 0.0f)) return f2(0.0f, 0.0f);
            const bool motor = (fl[s] & kHasMotor) != 0u;
            const float4 md = motor ? motor_of(s) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            return g.non_contact_with(fl[s], invm, ids + lane + s * BLOCK, band_of(s), motor, f2(md.x, md.y), pi);
        };
        // Candidate tests against the staged geometry (Geometry:279-290, 361-379).
        auto segment_candidate = [&](float2 pc, float radius, float margin, uint32_t cat, uint32_t mask, uint32_t j)
            __attribute__((always_inline)) -> bool {
            if (!filter_accepts(cat, mask, S.segCat[j], S.segMask[j])) return false;
            const float4 ab = S.segAB[j];
            const float reach = radius + margin + sepTol;
            const float2 lower = subs(min2(f2(ab.x, ab.y), f2(ab.z, ab.w)), reach);
            const float2 upper = adds(max2(f2(ab.x, ab.y), f2(ab.z, ab.w)), reach);
            return pc.x >= lower.x && pc.y >= lower.y && pc.x <= upper.x && pc.y <= upper.y;
        };
        auto body_candidate = [&](float2 pc, float radius, float margin, uint32_t cat, uint32_t mask, uint32_t k)
            __attribute__((always_inline)) -> bool {
            const uint4 kind = S.bodyKind[k];
            const bool isStatic = (kind.y & kBodyStatic) != 0;
            const bool obstacle = isStatic ? (kind.x != 0u) : ((kind.y & kBodyKinematic) != 0u);
            if (!obstacle) return false;
            if (!filter_accepts(cat, mask, kind.z, kind.w)) return false;
            const float4 pose = S.bodyPose[k], ex = S.bodyExtent[k];
            const float2 prev = f2(ex.x, ex.y);
            const float br = ex.z;
            const float2 bodyLow = subs(min2(f2(pose.x, pose.y), prev), br);
            const float2 bodyHigh = adds(max2(f2(pose.x, pose.y), prev), br);
            const float reach = radius + fmaxf(margin, 0.0f);
            const float2 lo = subs(pc, reach), hi = adds(pc, reach);
            return !(hi.x < bodyLow.x || hi.y < bodyLow.y || bodyHigh.x < lo.x || bodyHigh.y < lo.y);
        };

        // Hashed grid of the current centres in LDS (Force:1066-1105).
        auto build_hash = [&]() __attribute__((always_inline)) {
            for (uint32_t b = lane; b < buckets; b += BLOCK) S.buckets[b] = 0;
            __syncthreads();
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s] || !finite2(p[s])) continue;
                atomicAdd(&S.buckets[bucket_of(cell_coord(p[s].x, cellWidth), cell_coord(p[s].y, cellWidth), buckets)], 1u);
            }
            __syncthreads();
            bucket_prefix<BLOCK>(S.red, S.buckets, buckets);
            __syncthreads();
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s] || !finite2(p[s])) continue;
                uint32_t slot = atomicAdd(&S.buckets[bucket_of(cell_coord(p[s].x, cellWidth), cell_coord(p[s].y, cellWidth), buckets)], 1u);
                S.grid[slot] = (uint16_t)(lane + s * BLOCK);
            }
            __syncthreads();
        };

        // Adjacency row of bubble i (pairs both ways from the 3x3 cells, segments, bodies).
        auto enumerate = [&](const uint32_t i, const float2 pi, const float mi, auto emit, uint32_t &forward) __attribute__((always_inline)) {
            forward = 0;
            const float ri = S.r[i];
            const uint32_t ci = catG[i], ki = maskG[i];
            if (finite2(pi)) {
                const bool ownRow = allStarts || finite2(startG[i]);
                const int64_t cx = cell_coord(pi.x, cellWidth), cy = cell_coord(pi.y, cellWidth);
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int64_t tx = cx + dx, ty = cy + dy;
                        const uint32_t b = bucket_of(tx, ty, buckets);
                        const uint32_t lo = b == 0 ? 0u : S.buckets[b - 1], hi = S.buckets[b];
                        // Entries in groups of four: the group's index loads, then its centre,
                        // radius and margin loads, are issued together, so one LDS round trip
                        // serves four entries. The entries are still tested in bucket order.
                        for (uint32_t k0 = lo; k0 < hi; k0 += 4) {
                            uint32_t jj[4];
                            float2 qq[4];
                            float rr[4], mm[4];
    #pragma unroll
                            for (int u = 0; u < 4; ++u) jj[u] = S.grid[min(k0 + u, hi - 1)];
    #pragma unroll
                            for (int u = 0; u < 4; ++u) { qq[u] = S.pos[jj[u]]; rr[u] = S.r[jj[u]]; mm[u] = S.aux[jj[u]]; }
    #pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                if (k0 + u >= hi) break;
                                const uint32_t j = jj[u];
                                if (j == i) continue;
                                const float2 q = qq[u];
                                // The reach test first: it is cheaper than the cell test and most
                                // entries fail it (the conditions are independent, so the result
                                // is the same in either order). The reach sums the lower index's
                                // radius and margin first (r_l + r_h is exact in either order,
                                // the margins are not), and |x_l - x_h| = |x_i - x_j| exactly.
                                const bool iLow = i < j;
                                const float ml = iLow ? mi : mm[u], mh = iLow ? mm[u] : mi;
                                const float reach = ri + rr[u] + ml + mh + sepTol;
                                if (!(fabsf(pi.x - q.x) <= reach && fabsf(pi.y - q.y) <= reach)) continue;
                                if (cell_coord(q.x, cellWidth) != tx || cell_coord(q.y, cellWidth) != ty) continue;
                                if (!allAccept) {
                                    const uint32_t l = min(i, j), h = max(i, j);
                                    if (!filter_accepts(catG[l], maskG[l], catG[h], maskG[h])) continue;
                                }
                                if (j > i) {
                                    if (!ownRow) continue;
                                    ++forward;
                                } else if (!allStarts && !finite2(startG[j])) {
                                    continue;
                                }
                                emit(j);
                            }
                        }
                    }
                if (ownRow)
                    for (uint32_t j = 0; j < g.ns; ++j)
                        if (segment_candidate(pi, ri, mi, ci, ki, j)) { ++forward; emit((1u << 30) | j); }
            }
            for (uint32_t k = 0; k < g.nb; ++k)
                if (body_candidate(pi, ri, mi, ci, ki, k)) emit((2u << 30) | k);
        };

        // Rows cut to the reference's k candidates after a build that dropped some (forward_cut,
        // mark_body_cuts, cut_row), with the counters of the reference's truncated rows
        // (Geometry:291-297, 410-414; Contact:190-195). The grid slots hold the forward cuts
        // (the contact assertion hashes the grid again); the margins are still in S.aux.
        auto cut_rows = [&](uint32_t *globalRows) __attribute__((always_inline)) {
            const uint32_t *hits = g.bodyHits();
            uint32_t *bodyCut = A->bodyCut + d->bodyOffset;
            // One lane per body: the index of the k-th bubble of its row (a scan of the bubbles in
            // index order that stops there), so each body entry is then decided by one compare.
            for (uint32_t k = lane; k < g.nb; k += BLOCK)
                bodyCut[k] = body_cut_index(n, limit, hits[k], [&](uint32_t a) {
                    return body_candidate(S.pos[a], S.r[a], S.aux[a], catG[a], maskG[a], k);
                });
            __syncthreads();
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK, pc = rpq[s] & 0xffffu, segEnd = rpq[s] >> 16;
                uint32_t fp, cut;
                if (rows.inLds) {
                    uint16_t *row = S.rows + rs[s];
                    cut = forward_cut(row, pc, i, limit, fp);
                    mark_body_cuts(row, segEnd, rl[s], i, bodyCut);
                } else {
                    uint32_t *row = globalRows + rs[s];
                    cut = forward_cut(row, pc, i, limit, fp);
                    mark_body_cuts(row, segEnd, rl[s], i, bodyCut);
                }
                S.grid[i] = cut == kNoCut ? (uint16_t)0xffffu : (uint16_t)cut;
            }
            __syncthreads();
            auto cutOf = [&](uint32_t l) { const uint32_t v = S.grid[l]; return v == 0xffffu ? kNoCut : v; };
            uint32_t dropped = 0, widest = 0, kept = 0;
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK, pc = rpq[s] & 0xffffu, segEnd = rpq[s] >> 16;
                uint32_t fp;
                RowCut r;
                if (rows.inLds) {
                    uint16_t *row = S.rows + rs[s];
                    forward_cut(row, pc, i, limit, fp);
                    r = cut_row(row, pc, segEnd, rl[s], i, limit, fp, cutOf);
                } else {
                    uint32_t *row = globalRows + rs[s];
                    forward_cut(row, pc, i, limit, fp);
                    r = cut_row(row, pc, segEnd, rl[s], i, limit, fp, cutOf);
                }
                rl[s] = r.len;
                rpq[s] = r.pairEnd | (r.segEnd << 16);
                const uint32_t forward = fp + (segEnd - pc);
                dropped += (forward > limit ? forward - limit : 0u) + (r.demand > limit ? r.demand - limit : 0u);
                widest = max(widest, max(forward, r.demand));
                kept += r.len;
            }
            for (uint32_t k = lane; k < g.nb; k += BLOCK) {
                const uint32_t h = hits[k];
                dropped += h > limit ? h - limit : 0u;
                widest = max(widest, h);
            }
            atomicAdd(&S.red.counters[4], dropped);
            atomicMax(&S.red.counters[5], widest);
            atomicAdd(&S.red.counters[6], kept);
            __syncthreads();
            ov0 += S.red.counters[4];
            ov1 = max(ov1, S.red.counters[5]);
            entriesTotal = S.red.counters[6];
            __syncthreads();
        };

        // Grid build (Force:1040-1134) with CSR rows. Returns false when the rows do not fit.
        auto build = [&]() __attribute__((always_inline)) -> bool {
            float remaining = fmaxf(c.t0 - c.t1, 0.0f);
            float width = sepTol;
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                uint32_t i = lane + s * BLOCK;
                float2 a0 = a0_of(s, S.invm[i], p[s]);
                float travel = len2(S.vel[i]) * remaining + 0.5f * len2(a0) * remaining * remaining;
                mg[s] = d->minimumMarginOnly ? 0.25f * S.r[i] : fmaxf(1.05f * travel, 0.25f * S.r[i]);
                ref[s] = p[s];
                S.aux[i] = mg[s];
                width = fmaxf(width, 2.0f * (S.r[i] + mg[s]) + sepTol);
            }
            for (uint32_t k = lane; k < g.nb; k += BLOCK) g.bodyHits()[k] = 0;
            width = block_maxf<BLOCK>(S.red, width);
            cellWidth = fmaxf(width, kFloatMin);
            if (lane < 8) S.red.counters[lane] = 0;
            build_hash();
            // Pass 1: row lengths and the reference's overflow counters (Geometry:291-297, 410-414).
            // Each row is also staged in the world's global adjacency area at a fixed stride
            // (kStageStride entries per bubble, when the area holds n strides), so that pass 2
            // copies it instead of enumerating the candidates again. The block lengths of the
            // row (pairs, then segments, then bodies: the order enumeration emits them in) are
            // counted here as well.
            const bool stageOk = (uint64_t)n * kStageStride <= d->adjCapacity;
            uint32_t *const stageArea = A->adj + d->adjOffset;
            uint32_t droppedLocal = 0, maxLocal = 0, pairsLocal = 0, mine = 0, longLocal = 0;
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                uint32_t cnt = 0, forward = 0, pc = 0, sc = 0;
                uint32_t *stage = stageArea + (size_t)(lane + s * BLOCK) * kStageStride;
                enumerate(lane + s * BLOCK, p[s], mg[s], [&](uint32_t e) __attribute__((always_inline)) {
                    if (stageOk && cnt < kStageStride) stage[cnt] = e;
                    ++cnt;
                    pc += (e >> 30) == 0u;
                    sc += (e >> 30) == 1u;
                    if ((e >> 30) == 2u) atomicAdd(&g.bodyHits()[e & kIndexMask], 1u);
                }, forward);
                rl[s] = cnt;
                rpq[s] = pc | ((pc + sc) << 16);
                longLocal |= cnt > kStageStride ? 1u : 0u;
                mine += cnt;
                if (forward > limit) droppedLocal += forward - limit;
                if (cnt > limit) droppedLocal += cnt - limit;
                maxLocal = max(maxLocal, max(forward, cnt));
                pairsLocal += min(forward, limit);
            }
            __syncthreads();
            for (uint32_t k = lane; k < g.nb; k += BLOCK) {
                uint32_t hits = g.bodyHits()[k];
                maxLocal = max(maxLocal, hits);
                if (hits > limit) droppedLocal += hits - limit;
            }
            atomicAdd(&S.red.counters[0], droppedLocal);
            atomicMax(&S.red.counters[1], maxLocal);
            atomicAdd(&S.red.counters[2], pairsLocal);
            if (longLocal) atomicOr(&S.red.counters[3], 1u);
            uint32_t total;
            uint32_t base = block_exclusive_scan<BLOCK>(S.red, mine, total);
    #pragma unroll
            for (int s = 0; s < PER; ++s) { rs[s] = base; base += rl[s]; }
            entriesTotal = total;
            const uint32_t buildDropped = S.red.counters[0], buildWidest = S.red.counters[1];
            pairs = S.red.counters[2];
            rows.inLds = ldsRowsAllowed && total <= (uint32_t)ROWCAP;
            // Rows that stay in global memory stay where pass 1 staged them when every row fit
            // its stride (rows need not be contiguous: each bubble keeps its start and length).
            const bool strideRows = !rows.inLds && stageOk && S.red.counters[3] == 0u;
            if (strideRows) {
    #pragma unroll
                for (int s = 0; s < PER; ++s) rs[s] = (lane + s * BLOCK) * kStageStride;
            }
            // Rows in LDS or packed in global memory must also fit the world's adjacency slots:
            // the force pass keeps each entry's geometry in the slot of the same index.
            if (!strideRows && total > d->adjCapacity) {
                ov0 += buildDropped; ov1 = max(ov1, buildWidest);
                return false;
            }
            // Pass 2: fill the rows and order the pair block (Contact:222-238).
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK, cnt = rl[s], pc = rpq[s] & 0xffffu;
                uint32_t forward, filled = 0;
                if (rows.inLds) {
                    uint16_t *row = S.rows + rs[s];
                    if (stageOk && cnt <= kStageStride) {
                        const uint32_t *src = stageArea + (size_t)i * kStageStride;
                        for (uint32_t x = 0; x < cnt; x += 4) {
                            uint32_t ee[4];
    #pragma unroll
                            for (int u = 0; u < 4; ++u) ee[u] = src[min(x + u, cnt - 1)];
    #pragma unroll
                            for (int u = 0; u < 4; ++u)
                                if (x + u < cnt) row[x + u] = (uint16_t)(((ee[u] >> 30) << kLdsKindShift) | (ee[u] & kIndexMask));
                        }
                    } else {
                        enumerate(i, p[s], mg[s], [&](uint32_t e) __attribute__((always_inline)) {
                            row[filled++] = (uint16_t)(((e >> 30) << kLdsKindShift) | (e & kIndexMask));
                        }, forward);
                    }
                    sort_pair_block(row, pc);
                } else if (strideRows) {
                    sort_pair_block(stageArea + rs[s], pc);
                } else {
                    uint32_t *row = A->adj + d->adjOffset + rs[s];
                    enumerate(i, p[s], mg[s], [&](uint32_t e) __attribute__((always_inline)) { row[filled++] = e; }, forward);
                    sort_pair_block(row, pc);
                }
            }
            __syncthreads();
            if (buildDropped != 0u) cut_rows(stageArea);
            else { ov0 += buildDropped; ov1 = max(ov1, buildWidest); }
            return true;
        };

        // F: forces, kick (phase 1) or initial acceleration (phase 2), omega^2, margin bound.
        float w2Max = -1.0f, mbMin = inff(), dbMin = inff();
        uint32_t w2Own = 0, mbOwn = 0, dbOwn = 0;
        auto pass_forces = [&](uint32_t phase, bool kick) __attribute__((always_inline)) {
            const float fraction = step_fraction(c);
            refresh_frames(fraction);
            const float halfStep = 0.5f * c.h;
            float bestW = -1.0f, bestM = inff();
            uint32_t ownW = 0, ownM = 0;
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK;
                const float2 pi = p[s];
                const float invm = S.invm[i], ri = S.r[i], Ki = S.K[i];
                float2 total = f2(0.0f, 0.0f);
                float w2 = 0.0f;
                if (invm > 0.0f && (fl[s] & kHasBand)) w2 = w2 + g.band(band_of(s)).z * invm;
                // The row is (pairs, segments, bodies): the reference's summation order (Force:216-255).
                const uint32_t start = rs[s], pairEnd = start + (rpq[s] & 0xffffu), segEnd = start + (rpq[s] >> 16), end = start + rl[s];
                // QB pair entries at a time: their geometry and (for equal stiffnesses) their laws are
                // independent and evaluated together; the sums then take them in row order.
                for (uint32_t e = start; e < pairEnd; e += QB) {
                    uint32_t jq[QB];
                    bool okq[QB];
                    float2 nq[QB];
                    float ddq[QB], Klo[QB], Khi[QB], rlo[QB], rhi[QB];
                    bool same = true;
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        okq[q] = e + q < pairEnd;
                        const uint32_t j = rows.index(okq[q] ? e + q : start);
                        jq[q] = j;
                        const uint32_t lo = min(i, j), hi = max(i, j);
                        const float2 delta = sub(S.pos[hi], S.pos[lo]);
                        ddq[q] = len2(delta);
                        nq[q] = ddq[q] > 0.0f ? dvs(delta, ddq[q]) : f2(1.0f, 0.0f);
                        if (okq[q]) geoG[e + q] = make_float4(nq[q].x, nq[q].y, ddq[q], 0.0f);
                        Klo[q] = S.K[lo]; Khi[q] = S.K[hi]; rlo[q] = S.r[lo]; rhi[q] = S.r[hi];
                        same = same && (!okq[q] || Klo[q] == Khi[q]);
                    }
                    Law lq[QB];
                    if (same) {
#pragma unroll
                        for (int q = 0; q < QB; ++q) lq[q] = pair_law_same(Klo[q], rlo[q], rhi[q], ddq[q]);
                    } else {
                        for (int q = 0; q < QB; ++q)
                            lq[q] = okq[q] ? pair_law(Klo[q], rlo[q], Khi[q], rhi[q], ddq[q]) : Law{0.0f, 0.0f, 0.0f, 0.0f};
                    }
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        if (!okq[q]) continue;
                        total = add(total, mul(i < jq[q] ? -lq[q].force : lq[q].force, nq[q]));
                        if (invm > 0.0f) w2 = w2 + lq[q].stiffness * (invm + __builtin_sqrtf(invm * S.invm[jq[q]]));
                    }
                }
                for (uint32_t e = pairEnd; e < segEnd; ++e) {
                    float2 normal = f2(0.0f, 0.0f); float dd = 0.0f;
                    const bool inRegion = wall_region(segment_at(rows.index(e)), pi, normal, dd);
                    geoG[e] = make_float4(normal.x, normal.y, dd, inRegion ? 1.0f : 0.0f);
                    if (!inRegion || !(dd < ri)) continue;
                    const Law law = wall_law(Ki, ri, dd);
                    total = add(total, mul(law.force, normal));
                    if (invm > 0.0f) w2 = w2 + law.stiffness * invm;
                }
                for (uint32_t e = segEnd; e < end; ++e) {
                    const Obstacle ob = frame_obstacle<QB>(frame_at(rows.index(e)), S.verts, S.tris, pi);
                    geoG[e] = make_float4(ob.normal.x, ob.normal.y, ob.distanceValue, ob.speedBound);
                    const Law law = wall_law(Ki, ri, ob.distanceValue);
                    total = add(total, mul(law.force, ob.normal));
                    if (invm > 0.0f) w2 = w2 + law.stiffness * invm;
                }
                if (kick) {
                    if (phase == 2) {
                        a[s] = add(mul(total, invm), a0_of(s, invm, pi));
                    } else if (invm > 0.0f) {
                        const float gamma = gamma_of(s);
                        const float decay = bpm::expf_det(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
                        a[s] = add(mul(total, invm), a0_of(s, invm, pi));
                        v[s] = add(mul(v[s], decay), mul(a[s], factor));
                    }
                    S.vel[i] = invm > 0.0f ? v[s] : f2(0.0f, 0.0f);
                    S.aux[i] = len2(a[s]);
                }
                if (w2 > bestW) { bestW = w2; ownW = i; }
                float mb;
                if (!(invm > 0.0f)) mb = inff();
                else {
                    const float left = mg[s] - len2(sub(pi, ref[s]));
                    if (isnan(left)) mb = inff();
                    else if (left <= 0.0f) mb = 0.0f;
                    else {
                        const float speed = len2(v[s]), acl = len2(a[s]);
                        const float den = speed + __builtin_sqrtf(speed * speed + 2.0f * acl * left);
                        mb = den > 0.0f ? 2.0f * left / den : inff();
                    }
                }
                if (mb < bestM) { bestM = mb; ownM = i; }
            }
            block_maxmin<BLOCK>(S.red, bestW, ownW, bestM, ownM);
            w2Max = bestW; w2Own = ownW; mbMin = bestM; mbOwn = ownM;
        };

        // H: d/4 bound (Force:550-598).
        auto pass_distance = [&](float hMax) __attribute__((always_inline)) {
            const float fraction = step_fraction(c);
            refresh_frames(fraction);
            float bestD = inff();
            uint32_t ownD = 0;
            // One candidate's bound (Force:563-597); returns +inf for an exempt or degenerate one.
            // Written without branches (the exempt and degenerate cases are selected at the end), so
            // that the bounds of several entries can be evaluated side by side.
            auto bound = [&](float dd, float approach, float accSum, float reach, float bodySpeed) __attribute__((always_inline)) -> float {
                const bool active = dd < reach;
                const float dc = fmaxf(dd, reach * 1e-6f);
                const bool exempt = !active && approach + bodySpeed + accSum * hMax * 0.5f <= 0.0f;
                const float speed = active ? fabsf(approach) + bodySpeed : approach + bodySpeed;
                const float sq = __builtin_sqrtf(speed * speed + accSum * dc * 0.5f);
                const float den = sq + speed;
                return (exempt || !(den > 0.0f)) ? inff() : dc / (2.0f * den);
            };
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK;
                if (!(S.invm[i] > 0.0f)) continue;
                const float2 vi = v[s];
                const float ri = S.r[i], accI = S.aux[i];
                float best = inff();
                const uint32_t start = rs[s], pairEnd = start + (rpq[s] & 0xffffu), segEnd = start + (rpq[s] >> 16), end = start + rl[s];
                // QB pair entries at a time (independent bounds; the minimum does not depend on order).
                for (uint32_t e = start; e < pairEnd; e += QB) {
                    float bq[QB];
#pragma unroll
                    for (int q = 0; q < QB; ++q) {
                        const bool ok = e + q < pairEnd;
                        const uint32_t j = rows.index(ok ? e + q : start);
                        // The force pass stored the pair's geometry from its lower to its higher
                        // index: the same bits as (x_j - x_i) / d for i < j, their negation (exact)
                        // for i > j; (1, 0) at d = 0 either way.
                        const float4 gq = geoG[ok ? e + q : start];
                        const float dd = gq.z;
                        const float2 normal = !(dd > 0.0f) ? f2(1.0f, 0.0f) : (i < j ? f2(gq.x, gq.y) : f2(-gq.x, -gq.y));
                        const float b = bound(dd, dot2(sub(vi, S.vel[j]), normal), accI + S.aux[j], ri + S.r[j], 0.0f);
                        bq[q] = ok ? b : inff();
                    }
#pragma unroll
                    for (int q = 0; q < QB; ++q) best = fminf(best, bq[q]);
                }
                for (uint32_t e = pairEnd; e < segEnd; ++e) {
                    const float4 gs = geoG[e];                    // wall_region of the force pass
                    if (gs.w == 0.0f) continue;
                    best = fminf(best, bound(gs.z, -dot2(vi, f2(gs.x, gs.y)), accI, ri, 0.0f));
                }
                for (uint32_t e = segEnd; e < end; ++e) {
                    const float4 gb = geoG[e];                    // frame_obstacle of the force pass
                    best = fminf(best, bound(gb.z, -dot2(vi, f2(gb.x, gb.y)), accI, ri, gb.w));
                }
                if (best < bestD) { bestD = best; ownD = i; }
            }
            block_min<BLOCK>(S.red, bestD, ownD);
            dbMin = bestD; dbOwn = ownD;
        };

        // D: kick and drift, phase 0 (Force:775-791).
        auto pass_drift = [&]() __attribute__((always_inline)) {
            const float h = c.h, halfStep = 0.5f * h;
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK;
                if (!(S.invm[i] > 0.0f)) continue;
                const float gamma = gamma_of(s);
                const float decay = bpm::expf_det(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
                const float2 vn = add(mul(v[s], decay), mul(a[s], factor));
                float angular = om[s] * bpm::expf_det(-gamma * h);
                if (fl[s] & kHasTorque) {
                    const float4 md = motor_of(s);
                    angular = angular + (md.z / md.w) * half_impulse_factor(gamma, h);
                }
                p[s] = add(p[s], mul(vn, h));
                v[s] = vn;
                rot[s] = rot[s] + om[s] * h;
                om[s] = angular;
                S.pos[i] = p[s];
            }
        };

        // Body and one-sided edge assertions of a movable bubble (Force:889-917) against the
        // staged geometry, in the reference's order; the frames are those of the current fraction.
        auto assert_staged = [&](float2 center, float2 start, uint32_t cat, uint32_t mask, Violation &out)
            __attribute__((always_inline)) -> bool {
            for (uint32_t k = 0; k < g.nb; ++k) {
                const uint4 kind = S.bodyKind[k];
                if (!(kind.y & (kBodyStatic | kBodyKinematic)) || !filter_accepts(cat, mask, kind.z, kind.w)) continue;
                // A centre outside the body's bounding circle (with a relative slack far above the
                // rounding of the inside test) is not inside it: the triangle test is skipped.
                const float2 q = sub(center, S.frameAt[k]);
                const float reach = S.bodyExtent[k].z * 1.0001f + 1e-3f;
                if (q.x * q.x + q.y * q.y > reach * reach) continue;
                if (frame_inside<QB>(center, frame_at(k), S.verts, S.tris)) {
                    out.kind = BP_VIOLATION_INSIDE_BODY; out.other = g.bodyId(k); out.segment = 0; out.detail = 0.0f;
                    return true;
                }
            }
            for (uint32_t j = 0; j < g.ns; ++j) {
                if (!(S.segFlags[j] & kSegPolygon) || !filter_accepts(cat, mask, S.segCat[j], S.segMask[j])) continue;
                const float4 ab = S.segAB[j];
                const float2 a = f2(ab.x, ab.y), edge = sub(f2(ab.z, ab.w), a);
                const float lengthSquared = dot2(edge, edge);
                if (!(lengthSquared > 0.0f)) continue;
                // The staged unit left normal is (-e.y, e.x) / sqrt(e . e), the same operations.
                const float4 rg = S.segRegion[j];
                const float2 outward = mul(f2(rg.z, rg.w), S.segOutward[j]);
                const float before = dot2(sub(start, a), outward), after = dot2(sub(center, a), outward);
                if (!(before > 0.0f) || !(after <= 0.0f)) continue;
                const float2 crossing = add(start, mul(sub(center, start), before / (before - after)));
                const float along = dot2(sub(crossing, a), edge) / lengthSquared;
                if (along < 0.0f || along > 1.0f) continue;
                out.kind = BP_VIOLATION_EDGE_SIDE; out.segment = g.segId(j); out.other = (uint64_t)g.segOwner(j); out.detail = after;
                return true;
            }
            return false;
        };

        // Contact assertion (Force:865-918): lowest violating bubble writes the record.
        auto assert_contacts = [&]() __attribute__((always_inline)) -> bool {
            if (!(cellWidth > 0.0f)) cellWidth = 1.0f;
            __syncthreads();
            build_hash();
            const float fraction = step_fraction(c);
            refresh_frames(fraction);
            uint32_t firstIndex = 0xffffffffu;
            Violation mine = {0u, 0u, 0, 0.0f};
            const uint32_t *regionG = A->region[ws] + o;
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s] || firstIndex != 0xffffffffu) continue;
                const uint32_t i = lane + s * BLOCK;
                const float2 center = p[s];
                if (!finite2(center) || regionG[i] == BP_REGION_REMOVED) continue;
                const bool movable = S.invm[i] > 0.0f;
                uint32_t bestJ = 0xffffffffu;
                float bestD = 0.0f;
                const int64_t cx = cell_coord(center.x, cellWidth), cy = cell_coord(center.y, cellWidth);
                int x0, x1, y0, y1;
                zero_distance_span(center.x, cellWidth, cx, x0, x1);
                zero_distance_span(center.y, cellWidth, cy, y0, y1);
                for (int dx = x0; dx <= x1; ++dx)
                    for (int dy = y0; dy <= y1; ++dy) {
                        const int64_t tx = cx + dx, ty = cy + dy;
                        const uint32_t b = bucket_of(tx, ty, buckets);
                        const uint32_t lo = b == 0 ? 0u : S.buckets[b - 1], hi = S.buckets[b];
                        for (uint32_t k = lo; k < hi; ++k) {
                            const uint32_t j = S.grid[k];
                            if (j <= i || j >= bestJ) continue;
                            const float2 q = S.pos[j];
                            const float dd = len2(sub(q, center));
                            if (!(dd <= 0.0f)) continue;
                            if (cell_coord(q.x, cellWidth) != tx || cell_coord(q.y, cellWidth) != ty) continue;
                            if (regionG[j] == BP_REGION_REMOVED || !filter_accepts(catG[i], maskG[i], catG[j], maskG[j])) continue;
                            if (!movable && !(S.invm[j] > 0.0f)) continue;
                            bestJ = j; bestD = dd;
                        }
                    }
                if (bestJ != 0xffffffffu) {
                    firstIndex = i; mine.kind = BP_VIOLATION_PAIR; mine.other = (uint64_t)ids[bestJ]; mine.detail = bestD;
                    continue;
                }
                if (!movable) continue;
                if (assert_staged(center, startG[i], catG[i], maskG[i], mine)) firstIndex = i;
            }
            const uint32_t winner = block_min_u32<BLOCK>(S.red, firstIndex);
            if (winner == 0xffffffffu) return false;
            if (firstIndex == winner) {
                st->violationKind = mine.kind;
                st->violationSegment = mine.segment;
                st->violationObjects[0] = (uint64_t)ids[winner];
                st->violationObjects[1] = mine.other;
                st->violationDetail = mine.detail;
            }
            return true;
        };

        // ---- 2-3. Grid build and the substep loop (Force:1257-1304).
        const bool injection = d->injectionEnabled != 0, assertEvery = d->assertEverySubstep != 0;
        if (n > 0) {
            bool ok = build();
            if (!ok) status = BP_ERR_CAPACITY_OVERFLOW;
            if (status == BP_OK) {
                pass_forces(2u, true);
                for (uint32_t iteration = 0; iteration < 2 * kFrameLimit + 2; ++iteration) {
                    const float remaining = c.t0 - c.t1;
                    const bool over = !(remaining > 0.0f), stopped = c.rebuild != 0 || c.limit != 0;
                    const bool limited = c.c1 >= kFrameLimit;
                    if (over || stopped || limited) {
                        if (limited && !over && !stopped) c.limit = 1;
                        c.h = 0.0f;
                    } else {
                        const float stiffnessBound = w2Max > 0.0f ? 1.0f / __builtin_sqrtf(w2Max) : inff();
                        pass_distance(fminf(remaining, stiffnessBound));
                        reduce_h(c, stiffnessBound, w2Own, dbMin, dbOwn, mbMin, mbOwn, history, lane);
                    }
                    const bool finished = !(c.t0 - c.t1 > 0.0f);
                    if (!(c.h > 0.0f)) {
                        if (c.limit || finished || !c.rebuild) break;
                        c.c0 = 0; c.rebuild = 0;
                        ov2 = max(ov2, ov0); ov0 = 0; ov3 += 1;
                        __syncthreads();
                        if (!build()) { status = BP_ERR_CAPACITY_OVERFLOW; break; }
                        pass_forces(2u, true);
                        continue;
                    }
                    pass_drift();
                    __syncthreads();
                    pass_forces(1u, true);
                    if (injection && c.c1 == d->injectionAfter + 1) {
    #pragma unroll
                        for (int s = 0; s < PER; ++s) {
                            const uint32_t i = lane + s * BLOCK;
                            if (valid[s] && ids[i] == d->injectionObject) {
                                p[s] = f2(d->injectionPosition[0], d->injectionPosition[1]);
                                S.pos[i] = p[s];
                            }
                        }
                        __syncthreads();
                        pass_forces(1u, false);
                    }
                    if (assertEvery && !claimed) {
                        if (assert_contacts()) claimed = true;
                        __syncthreads();
                    }
                }
                ov2 = max(ov2, ov0);
            }
        }

        // ---- 4. Gated end of step: region, finite check, assertion (NativeMetalSolver.swift:837-945).
        const bool finished = c.limit == 0 && !(c.t0 - c.t1 > 0.0f);
        uint32_t region[PER];
    #pragma unroll
        for (int s = 0; s < PER; ++s) region[s] = BP_REGION_UNCHANGED;
        if (status == BP_OK && finished) {
            const float2 lo = f2(d->region[0], d->region[1]), hi = f2(d->region[2], d->region[3]);
    #pragma unroll
            for (int s = 0; s < PER; ++s) {
                if (!valid[s]) continue;
                const uint32_t i = lane + s * BLOCK;
                if (d->applyRegion) {
                    float2 startI = startG[i];
                    region[s] = apply_region(p[s], startI, S.r[i], fl[s], lo, hi, safety);
                    if (region[s] == BP_REGION_WRAPPED) { startG[i] = startI; S.pos[i] = p[s]; }
                }
                const uint32_t gi = o + i;
                A->region[ws][gi] = region[s];
                if (!finite2(p[s]) || !finite2(startG[i]) || !finite2(v[s]) || !isfinite(rot[s]) || !isfinite(om[s]) ||
                    !isfinite(A->mass[gi]) || !isfinite(A->invm[gi]) || !isfinite(S.r[i]))
                    safety |= 1u;
            }
            for (uint32_t j = lane; j < g.ns; j += BLOCK) {
                const Seg sg = g.segment(j);
                const float2 e = sub(sg.b, sg.a);
                if (!finite2(sg.a) || !finite2(sg.b) || !finite2(e) || !isfinite(dot2(e, e))) safety |= 1u;
            }
            __syncthreads();
            if (!claimed && n > 0 && assert_contacts()) claimed = true;
        }

        // ---- Write-back of the working side, counters and publication. A served world's mirror
        // takes the same records; it is read only after a published step (a failed world takes no
        // further requests, and the server reads the mirror of a world whose last step succeeded).
        uint32_t rmLocal = 0, wrLocal = 0;
        const bool mirrored = d->mirror != nullptr && n <= d->mirrorCapacity;
    #pragma unroll
        for (int s = 0; s < PER; ++s) {
            if (!valid[s]) continue;
            const uint32_t gi = o + lane + s * BLOCK;
            A->pos[ws][gi] = p[s];
            A->vel[ws][gi] = v[s];
            A->rot[ws][gi] = rot[s];
            A->omega[ws][gi] = om[s];
            if (mirrored) mirror_record(A, d, ws, lane + s * BLOCK, p[s], v[s], rot[s], om[s]);
            if (!(status == BP_OK && finished)) A->region[ws][gi] = BP_REGION_UNCHANGED;
            rmLocal += region[s] == BP_REGION_REMOVED;
            wrLocal += region[s] == BP_REGION_WRAPPED;
            if (d->exportAdjacency) {
                A->rowStart[gi] = rs[s];
                A->rowLen[gi] = rl[s];
                if (rows.inLds)
                    for (uint32_t e = 0; e < rl[s]; ++e) {
                        uint32_t kind, j;
                        rows.get(rs[s] + e, kind, j);
                        A->adj[d->adjOffset + rs[s] + e] = (kind << kKindShift) | j;
                    }
            }
        }
        if (mirrored) mirror_bodies(A, d, ws, lane, BLOCK);
        __syncthreads();
        if (lane < 8) S.red.counters[lane] = 0;
        __syncthreads();
        if (safety) atomicOr(&S.red.counters[5], safety);
        atomicAdd(&S.red.counters[6], rmLocal);
        atomicAdd(&S.red.counters[7], wrLocal);
        __syncthreads();
        status = final_status(status, c, S.red.counters[5], claimed, finished, n);
        if (lane == 0) {
            st->status = status;
            st->substeps = c.c1;
            st->rebuilds = ov3;
            st->candidatePairs = pairs;
            st->candidateOverflow = ov2;
            st->maxCandidates = ov1;
            st->removals = S.red.counters[6];
            st->wraps = S.red.counters[7];
            st->time[0] = c.t0; st->time[1] = c.t1; st->time[2] = c.h; st->time[3] = c.split;
            st->limiterObject = limiter_object(c, ids);
            st->limiterTerm = c.limTerm;
            if (!claimed) {
                st->violationKind = 0; st->violationSegment = 0; st->violationObjects[0] = 0; st->violationObjects[1] = 0;
                st->violationDetail = 0.0f;
            }
            st->adjacencyEntries = entriesTotal;
            st->groupWidth = BLOCK;
            st->executed = 1;
            st->ticks = (uint32_t)(wall_clock64() - ticks0);
            if (status == BP_OK) {
                d->pub = ws;
                if (S.red.counters[6] != 0) d->failed = kWorldHeld;
            } else if (status != BP_ERR_CAPACITY_OVERFLOW) {
                d->failed = kWorldFailed;
            }
            post_status(st, notify);
        }
        // A failed world stops, and so does one whose step removed bubbles from the region: the
        // removed bubbles leave the world before its next step. Later steps stay unexecuted.
        if (status != BP_OK || S.red.counters[6] != 0) break;
        pubNow = ws;
        __syncthreads();                                      // LDS is reused by the next step
    }
#undef ids
#undef catG
#undef maskG
#undef startG
}

// The worlds order[0..count) of one configuration. Without `cursor` each workgroup steps the
// world of its own index. With `cursor` (a zeroed device word) the launch has fewer workgroups
// than worlds, sized by the host to the CUs the class may hold: each workgroup steps the world of
// its index, then takes the next unclaimed index from the cursor, so the worlds run in the
// order's longest-first sequence on a fixed set of resident workgroups. Each such workgroup
// counts itself in `placed` when it starts (the host's gate for the other classes waits on it).
template <int BLOCK, int PER, int ROWCAP, int SEGCAP, int BODYCAP, int VERTCAP, int TRICAP>
__global__ __launch_bounds__(BLOCK) void world_lds_kernel(const Arena *__restrict__ A, WorldDesc *descs, WorldStatus *stats,
                                                          const uint32_t *order, uint32_t count, uint32_t steps,
                                                          uint32_t statsStride, uint32_t notify, uint32_t *cursor,
                                                          uint32_t *placed) {
    __shared__ LdsWorld<BLOCK, PER, ROWCAP, SEGCAP, BODYCAP, VERTCAP, TRICAP> S;
    __shared__ uint32_t nextSlot;
    if (placed && threadIdx.x == 0) atomicAdd(placed, 1u);
    uint32_t slot = blockIdx.x;
    while (slot < count) {
        lds_world_steps<BLOCK, PER, ROWCAP, SEGCAP, BODYCAP, VERTCAP, TRICAP>(S, A, descs, stats, order[slot], steps,
                                                                              statsStride, notify);
        if (!cursor) break;
        __syncthreads();                                      // the next world reuses LDS
        if (threadIdx.x == 0) nextSlot = gridDim.x + atomicAdd(cursor, 1u);
        __syncthreads();
        slot = nextSlot;
    }
}

}  // namespace bpd

#endif  // BP_WORLD_LDS_H
