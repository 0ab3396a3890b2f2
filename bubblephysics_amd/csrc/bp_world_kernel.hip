// bp_world_kernel.hip -- one workgroup advances one world by one step (gfx950).
//
// MI355X design of the reference's single-dispatch world kernel
// (NativeForceKernels.metal-source:1236-1307, nativeSubstepLoopSingleGroup) and, for every world
// size, of its multi-dispatch equivalent (NativeMetalSolver.swift:801-833, NativeMetalSubstepLoop
// .swift:95-120). A launch holds one workgroup per world, so one launch advances a whole batch
// of independent worlds; a plain world is a batch of one.
//
// Per step and world:
//   1. begin step and kinematic-body prediction (Force:285-293, PostSolve:66-82);
//   2. grid build: a0, margin and x0, cell width, a counting-sorted hashed grid in LDS, then a
//      CSR adjacency (rows sorted by (kind, neighbour)) in two passes without atomics on rows;
//   3. substep loop with three passes per substep instead of the reference's four:
//        D: kick and drift (phase 0),
//        F: contact forces + kick (phase 1) + omega^2 + margin bound, fused because the
//           reduction of the next substep reads exactly the positions the forces used,
//        H: d/4 bound with h_max, then h (Force:625-736) evaluated redundantly by every lane;
//   4. gated end of step: region, finite check, contact assertion, status and publication.
// Reductions keep the reference's tie rule (equal values -> lower bubble index), and every
// per-bubble sum runs in adjacency-row order, so the result is independent of the block size.
#include "bp_world_global.h"

namespace bpd {

template <int BLOCK> struct Shared {
    uint32_t buckets[kMaxBuckets];
    Red<BLOCK> red;
};

struct RowStats { uint32_t dropped, maxDemand, pairs, overflowRows; };

// Cuts the sorted rows to the reference's k candidates (forward_cut, mark_body_cuts, cut_row)
// after a build that dropped candidates, and replaces the build's counters by those of the
// reference's truncated rows (Geometry:291-297, 410-414; Contact:190-195). The grid slots hold
// the forward cuts (the contact assertion builds its own grid).
template <int BLOCK>
BPD void cut_rows(World &W, Shared<BLOCK> &sh, uint32_t limit, RowStats &stats, uint32_t &entries) {
    const uint32_t n = W.n;
    const uint32_t *hits = W.A->bodyHits + W.d->bodyOffset;
    uint32_t *bodyCut = W.A->bodyCut + W.d->bodyOffset;
    for (uint32_t k = threadIdx.x; k < W.nb; k += BLOCK)        // one lane per body
        bodyCut[k] = body_cut_index(n, limit, hits[k], [&](uint32_t a) { return W.body_candidate(a, k); });
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        uint32_t *row = W.adj + W.rowStart[i];
        const uint32_t len = W.rowLen[i];
        uint32_t pairEnd, segEnd, forwardPairs;
        row_blocks(row, len, pairEnd, segEnd);
        W.grid[i] = forward_cut(row, pairEnd, i, limit, forwardPairs);
        mark_body_cuts(row, segEnd, len, i, bodyCut);
    }
    __syncthreads();
    uint32_t dropped = 0, widest = 0, kept = 0;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        uint32_t *row = W.adj + W.rowStart[i];
        const uint32_t len = W.rowLen[i];
        uint32_t pairEnd, segEnd, forwardPairs;
        row_blocks(row, len, pairEnd, segEnd);
        forward_cut(row, pairEnd, i, limit, forwardPairs);
        const RowCut r = cut_row(row, pairEnd, segEnd, len, i, limit, forwardPairs, [&](uint32_t l) { return W.grid[l]; });
        W.rowLen[i] = r.len;
        const uint32_t forward = forwardPairs + (segEnd - pairEnd);
        dropped += (forward > limit ? forward - limit : 0u) + (r.demand > limit ? r.demand - limit : 0u);
        widest = max(widest, max(forward, r.demand));
        kept += r.len;
    }
    for (uint32_t k = threadIdx.x; k < W.nb; k += BLOCK) {
        const uint32_t h = hits[k];
        dropped += h > limit ? h - limit : 0u;
        widest = max(widest, h);
    }
    atomicAdd(&sh.red.counters[5], dropped);
    atomicMax(&sh.red.counters[6], widest);
    atomicAdd(&sh.red.counters[7], kept);
    __syncthreads();
    stats.dropped = sh.red.counters[5];
    stats.maxDemand = sh.red.counters[6];
    entries = sh.red.counters[7];
    __syncthreads();
}

// Grid build of one workgroup: Force:1040-1134 with a CSR adjacency. Returns false when the
// adjacency does not fit the world's capacity.
template <int BLOCK>
BPD bool build_grid(World &W, Shared<BLOCK> &sh, float remaining, uint32_t limit, RowStats &stats, uint32_t &entries) {
    const uint32_t n = W.n;
    const WorldDesc &d = *W.d;
    // a0, margin and x0 (Force:1046-1049, 492-499).
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        float2 p = W.pos[i];
        float2 a = W.non_contact(i, p);
        W.a0[i] = a;
        float t = fmaxf(remaining, 0.0f);
        float travel = len2(W.eff_vel(i)) * t + 0.5f * len2(a) * t * t;
        W.margin[i] = d.minimumMarginOnly ? 0.25f * W.radius[i] : fmaxf(1.05f * travel, 0.25f * W.radius[i]);
        W.ref[i] = p;
    }
    __syncthreads();
    // Cell width (Force:1052-1064).
    float width = W.sepTol;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) width = fmaxf(width, 2.0f * (W.radius[i] + W.margin[i]) + W.sepTol);
    width = block_maxf<BLOCK>(sh.red, width);
    W.cellWidth = fmaxf(width, kFloatMin);
    for (uint32_t k = threadIdx.x; k < W.nb; k += BLOCK) W.A->bodyHits[d.bodyOffset + k] = 0;
    // Hashed grid: count, exclusive prefix, scatter (Force:1066-1105).
    uint32_t buckets = d.buckets;
    for (uint32_t b = threadIdx.x; b < buckets; b += BLOCK) sh.buckets[b] = 0;
    if (threadIdx.x < 8) sh.red.counters[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        float2 p = W.pos[i];
        if (!finite2(p)) continue;
        atomicAdd(&sh.buckets[bucket_of(W.cell_x(p), W.cell_y(p), buckets)], 1u);
    }
    __syncthreads();
    {
        uint32_t perLane = (buckets + BLOCK - 1) / BLOCK;
        uint32_t first = min(threadIdx.x * perLane, buckets), last = min(first + perLane, buckets);
        uint32_t sum = 0;
        for (uint32_t b = first; b < last; ++b) sum += sh.buckets[b];
        uint32_t total;
        uint32_t running = block_exclusive_scan<BLOCK>(sh.red, sum, total);
        __syncthreads();
        for (uint32_t b = first; b < last; ++b) { uint32_t own = sh.buckets[b]; sh.buckets[b] = running; running += own; }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        float2 p = W.pos[i];
        if (!finite2(p)) continue;
        uint32_t slot = atomicAdd(&sh.buckets[bucket_of(W.cell_x(p), W.cell_y(p), buckets)], 1u);
        W.grid[slot] = i;
        W.publish_cell(slot, i);
    }
    __syncthreads();
    // Pass 1: row lengths and the reference's overflow counters. Each row is also staged at a
    // fixed stride of kStageStride entries per bubble in the world's adjacency area (when the
    // area holds n strides); when every row fits its stride, the rows stay there and pass 2 only
    // orders their pair blocks instead of enumerating the candidates again.
    const uint32_t *ends = sh.buckets;
    uint32_t droppedLocal = 0, maxLocal = 0, pairsLocal = 0, overflowLocal = 0, longLocal = 0;
    uint32_t *bodyHits = W.A->bodyHits + d.bodyOffset;
    const bool stageOk = (uint64_t)n * kStageStride <= d.adjCapacity;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        uint32_t count = 0, forward = 0;
        uint32_t *stage = W.adj + (size_t)i * kStageStride;
        auto tally = [&](uint32_t e) {
            if (stageOk && count < kStageStride) stage[count] = e;
            ++count;
            if ((e >> kKindShift) == 2u) atomicAdd(&bodyHits[e & kIndexMask], 1u);
        };
        enumerate_row(W, ends, buckets, i, tally, forward);
        W.rowLen[i] = count;
        longLocal |= count > kStageStride ? 1u : 0u;
        if (forward > limit) { droppedLocal += forward - limit; overflowLocal = 1; }
        maxLocal = max(maxLocal, max(forward, count));
        if (count > limit) { droppedLocal += count - limit; overflowLocal = 1; }
        pairsLocal += min(forward, limit);
    }
    __syncthreads();
    // Body rows of the reference (Geometry:410-414): their length enters the same counters.
    for (uint32_t k = threadIdx.x; k < W.nb; k += BLOCK) {
        uint32_t hits = bodyHits[k];
        maxLocal = max(maxLocal, hits);
        if (hits > limit) { droppedLocal += hits - limit; overflowLocal = 1; }
    }
    atomicAdd(&sh.red.counters[0], droppedLocal);
    atomicMax(&sh.red.counters[1], maxLocal);
    atomicAdd(&sh.red.counters[2], pairsLocal);
    atomicOr(&sh.red.counters[3], overflowLocal);
    if (longLocal) atomicOr(&sh.red.counters[4], 1u);
    __syncthreads();
    // Row starts: tiled exclusive scan of the row lengths in bubble order.
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += BLOCK) {
        uint32_t i = base + threadIdx.x;
        uint32_t len = i < n ? W.rowLen[i] : 0u;
        uint32_t total;
        uint32_t offset = block_exclusive_scan<BLOCK>(sh.red, len, total);
        if (i < n) W.rowStart[i] = carry + offset;
        carry += total;
    }
    entries = carry;
    stats.dropped = sh.red.counters[0];
    stats.maxDemand = sh.red.counters[1];
    stats.pairs = sh.red.counters[2];
    stats.overflowRows = sh.red.counters[3];
    const bool strideRows = stageOk && sh.red.counters[4] == 0u;
    if (strideRows) {
        // The staged rows are the rows: each bubble's start is its stride (rows need only a start
        // and a length), and only the pair block is out of order (hash-bucket order).
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
            W.rowStart[i] = i * kStageStride;
            uint32_t *row = W.adj + (size_t)i * kStageStride;
            const uint32_t count = W.rowLen[i];
            uint32_t pairs = 0;
            for (uint32_t x = 0; x < count; ++x) pairs += (row[x] >> kKindShift) == 0u ? 1u : 0u;
            sort_pair_block(row, pairs);
        }
        __syncthreads();
        if (stats.dropped != 0u) cut_rows<BLOCK>(W, sh, limit, stats, entries);
        return true;
    }
    if (carry > d.adjCapacity) return false;
    __syncthreads();
    // Pass 2: fill rows, then order each row by (kind, neighbour) (Contact:222-238).
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        uint32_t *row = W.adj + W.rowStart[i];
        uint32_t count = 0, forward = 0;
        auto put = [&](uint32_t e) { row[count++] = e; };
        enumerate_row(W, ends, buckets, i, put, forward);
        // Only the pair block can be out of order (hash-bucket order); segments and bodies follow in order.
        uint32_t pairs = 0;
        while (pairs < count && (row[pairs] >> kKindShift) == 0u) ++pairs;
        sort_pair_block(row, pairs);
    }
    __syncthreads();
    if (stats.dropped != 0u) cut_rows<BLOCK>(W, sh, limit, stats, entries);
    return true;
}

// Contact assertion (Force:865-918): the lowest-index bubble with a violation is recorded.
// Pairs are searched in the 3x3 cells of a grid of the current centres (coincident centres
// share a cell), bodies and one-sided edges by direct loops.

// Returns true (uniformly) when some bubble violates; the lowest violating bubble writes the
// record (NativeContactAssertion) into the world status.
template <int BLOCK>
BPD bool assert_contacts(World &W, Shared<BLOCK> &sh, const Ctl &c, WorldStatus *st) {
    const uint32_t n = W.n;
    const uint32_t buckets = W.d->buckets;
    if (!(W.cellWidth > 0.0f)) W.cellWidth = 1.0f;
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < buckets; b += BLOCK) sh.buckets[b] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        float2 p = W.pos[i];
        if (!finite2(p)) continue;
        atomicAdd(&sh.buckets[bucket_of(W.cell_x(p), W.cell_y(p), buckets)], 1u);
    }
    __syncthreads();
    {
        uint32_t perLane = (buckets + BLOCK - 1) / BLOCK;
        uint32_t first = min(threadIdx.x * perLane, buckets), last = min(first + perLane, buckets);
        uint32_t sum = 0;
        for (uint32_t b = first; b < last; ++b) sum += sh.buckets[b];
        uint32_t total;
        uint32_t running = block_exclusive_scan<BLOCK>(sh.red, sum, total);
        __syncthreads();
        for (uint32_t b = first; b < last; ++b) { uint32_t own = sh.buckets[b]; sh.buckets[b] = running; running += own; }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        float2 p = W.pos[i];
        if (!finite2(p)) continue;
        uint32_t slot = atomicAdd(&sh.buckets[bucket_of(W.cell_x(p), W.cell_y(p), buckets)], 1u);
        W.grid[slot] = i;
        W.publish_cell(slot, i);
    }
    __syncthreads();
    float fraction = c.t0 > 0.0f ? clampf(c.t1 / c.t0, 0.0f, 1.0f) : 1.0f;
    uint32_t firstIndex = 0xffffffffu;
    Violation mine = {0u, 0u, 0, 0.0f};
    for (uint32_t i = threadIdx.x; i < n && firstIndex == 0xffffffffu; i += BLOCK) {
        float2 center = W.pos[i], start = W.start[i];
        if (!finite2(center) || W.region[i] == BP_REGION_REMOVED) continue;
        bool movable = W.inv_mass(i) > 0.0f;
        const uint32_t ci = W.cat[i], ki = W.mask[i];
        uint32_t bestJ = 0xffffffffu;
        float bestD = 0.0f;
        int64_t cx = W.cell_x(center), cy = W.cell_y(center);
        int x0, x1, y0, y1;
        zero_distance_span(center.x, W.cellWidth, cx, x0, x1);
        zero_distance_span(center.y, W.cellWidth, cy, y0, y1);
        for (int dx = x0; dx <= x1; ++dx)
            for (int dy = y0; dy <= y1; ++dy) {
                int64_t tx = cx + dx, ty = cy + dy;
                uint32_t b = bucket_of(tx, ty, buckets);
                uint32_t lo = b == 0 ? 0u : sh.buckets[b - 1], hi = sh.buckets[b];
                for (uint32_t k = lo; k < hi; ++k) {
                    const uint4 cb = W.cellB[k];
                    uint32_t j = cb.x;
                    if (j <= i || j >= bestJ) continue;
                    const float4 ca = W.cellA[k];
                    float2 q = f2(ca.x, ca.y);
                    if (W.cell_x(q) != tx || W.cell_y(q) != ty) continue;
                    if ((cb.w & 2u) || !filter_accepts(ci, ki, cb.y, cb.z)) continue;
                    if (!movable && !(cb.w & 4u)) continue;
                    float dd = len2(sub(q, center));
                    if (dd <= 0.0f) { bestJ = j; bestD = dd; }
                }
            }
        if (bestJ != 0xffffffffu) {
            firstIndex = i; mine.kind = BP_VIOLATION_PAIR; mine.other = (uint64_t)W.id[bestJ]; mine.detail = bestD;
            break;
        }
        if (!movable) continue;
        bool found = false;
        for (uint32_t k = 0; k < W.nb && !found; ++k) {
            BodyView body = W.body(k);
            if (!body_immovable(body) || !filter_accepts(W.cat[i], W.mask[i], body.kind.z, body.kind.w)) continue;
            if (frame_inside(center, body_frame(body, fraction), W.verts, W.tris)) {
                found = true;
                mine.kind = BP_VIOLATION_INSIDE_BODY; mine.other = W.A->bodyId[W.d->bodyOffset + k]; mine.detail = 0.0f;
            }
        }
        for (uint32_t j = 0; j < W.ns && !found; ++j) {
            Seg s = W.segment(j);
            uint32_t g = W.d->segmentOffset + j;
            if (!(s.flags & kSegPolygon) || !filter_accepts(W.cat[i], W.mask[i], W.A->segCat[g], W.A->segMask[g])) continue;
            float2 edge = sub(s.b, s.a);
            float lengthSquared = dot2(edge, edge);
            if (!(lengthSquared > 0.0f)) continue;
            float2 outward = mul(dvs(f2(-edge.y, edge.x), __builtin_sqrtf(lengthSquared)), s.outward);
            float before = dot2(sub(start, s.a), outward), after = dot2(sub(center, s.a), outward);
            if (!(before > 0.0f) || !(after <= 0.0f)) continue;
            float2 crossing = add(start, mul(sub(center, start), before / (before - after)));
            float along = dot2(sub(crossing, s.a), edge) / lengthSquared;
            if (along < 0.0f || along > 1.0f) continue;
            found = true;
            mine.kind = BP_VIOLATION_EDGE_SIDE; mine.segment = W.A->segId[g];
            mine.other = (uint64_t)W.A->segOwner[g]; mine.detail = after;
        }
        if (found) { firstIndex = i; break; }
    }
    uint32_t winner = firstIndex;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) winner = min(winner, (uint32_t)__shfl_xor((int)winner, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh.red.scan[threadIdx.x >> 6] = winner;
    __syncthreads();
    winner = 0xffffffffu;
    for (int w = 0; w < BLOCK / 64; ++w) winner = min(winner, sh.red.scan[w]);
    if (winner == 0xffffffffu) return false;
    if (firstIndex == winner) {
        st->violationKind = mine.kind;
        st->violationSegment = mine.segment;
        st->violationObjects[0] = (uint64_t)W.id[winner];
        st->violationObjects[1] = mine.other;
        st->violationDetail = mine.detail;
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// The step kernel.
// ---------------------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void world_step_kernel(const Arena *__restrict__ A, WorldDesc *descs, WorldStatus *stats,
                                                           const uint32_t *order, uint32_t count, uint32_t notify) {
    __shared__ Shared<BLOCK> sh;
    if (blockIdx.x >= count) return;
    const uint32_t wi = order[blockIdx.x];
    WorldDesc *d = descs + wi;
    WorldStatus *st = stats + wi;
    const uint64_t ticks0 = wall_clock64();
    if (d->failed) {
        if (threadIdx.x == 0) { st->executed = 0; post_status(st, notify); }
        return;
    }
    const uint32_t pub = d->pub, wside = 1u - pub;
    World W = make_world(A, d, wside);
    const uint32_t o = d->bubbleOffset;
    const uint32_t n = W.n;
    float *history = A->history + d->historyOffset;

    // 1. Begin step (Force:285-293) on a working copy of the published state.
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
        float2 p = A->pos[pub][o + i];
        float2 v = A->vel[pub][o + i];
        float om = A->omega[pub][o + i];
        uint32_t f = A->flags[o + i];
        if ((f & kEnabled) && (f & kStatic)) { v = f2(0.0f, 0.0f); om = 0.0f; }
        W.pos[i] = p; W.start[i] = p; W.vel[i] = v; W.rot[i] = A->rot[pub][o + i]; W.omega[i] = om;
        W.region[i] = BP_REGION_UNCHANGED;
        W.publish_step(i);
    }
    // Kinematic and static bodies (PostSolve:66-82 for the bodies a bubble world may hold).
    uint32_t safety = 0;
    for (uint32_t k = threadIdx.x; k < W.nb; k += BLOCK) {
        uint32_t g = d->bodyOffset + k;
        float4 pose = A->bodyPose[pub][g];
        float4 motion = A->bodyMotion[g];
        uint4 kind = A->bodyKind[g];
        float2 prev = f2(pose.x, pose.y);
        pose.w = pose.z;
        if (kind.y & kBodyStatic) { motion.x = motion.y = motion.z = 0.0f; }
        else {
            float dt = d->dt;
            motion.z = motion.z * bpm::expf_det(-motion.w * dt);
            float2 lin = f2(motion.x, motion.y);
            float2 lim = A->bodyLimits[g];
            float speed = len2(lin);
            if (lim.x > 0.0f && speed > lim.x) lin = mul(lin, lim.x / speed);
            motion.x = lin.x; motion.y = lin.y;
            if (lim.y > 0.0f) motion.z = clampf(motion.z, -lim.y, lim.y);
            pose.x = pose.x + motion.x * dt;
            pose.y = pose.y + motion.y * dt;
            pose.z = pose.z + motion.z * dt;
            if (!isfinite(pose.x) || !isfinite(pose.y) || !isfinite(pose.z) || !isfinite(pose.w) || !isfinite(motion.x) ||
                !isfinite(motion.y) || !isfinite(motion.z))
                safety |= 1u;
        }
        A->bodyPose[wside][g] = pose;
        A->bodyPrev[wside][g] = prev;
        A->bodyMotion[g] = motion;
    }
    __syncthreads();

    const uint32_t limit = d->limit;
    Ctl c;
    c.t0 = d->dt; c.t1 = n == 0 ? d->dt : 0.0f; c.h = 0.0f; c.split = (float)(d->split < 1 ? 1u : d->split);
    c.c0 = c.c1 = c.rebuild = c.limit = 0; c.limOwner = kNoOwner; c.limTerm = BP_TERM_REMAINING;
    uint32_t ov0 = 0, ov1 = 0, ov2 = 0, ov3 = 0, pairs = 0, entriesTotal = 0;
    int32_t status = BP_OK;
    bool claimed = false;                 // sticky contact-assertion record of the frame

    if (n > 0) {
        float w2Max, mbMin, dbMin;
        uint32_t w2Own, mbOwn, dbOwn;
        RowStats rs;
        bool ok = build_grid<BLOCK>(W, sh, c.t0 - c.t1, limit, rs, entriesTotal);
        ov0 += rs.dropped; ov1 = max(ov1, rs.maxDemand); pairs = rs.pairs;
        if (!ok) status = BP_ERR_CAPACITY_OVERFLOW;
        if (status == BP_OK) {
            c.h = 0.0f;
            pass_forces<BLOCK>(W, sh.red, Range{threadIdx.x, BLOCK}, c, 2u, w2Max, w2Own, mbMin, mbOwn);
            for (uint32_t iteration = 0; iteration < 2 * kFrameLimit + 2; ++iteration) {
                // Reduction of h (Force:625-736).
                float remaining = c.t0 - c.t1;
                bool over = !(remaining > 0.0f), stopped = c.rebuild != 0 || c.limit != 0;
                bool limited = c.c1 >= kFrameLimit;
                if (over || stopped || limited) {
                    if (limited && !over && !stopped) c.limit = 1;
                    c.h = 0.0f;
                } else {
                    float stiffnessBound = w2Max > 0.0f ? 1.0f / __builtin_sqrtf(w2Max) : inff();
                    float hMax = fminf(remaining, stiffnessBound);
                    pass_distance<BLOCK>(W, sh.red, Range{threadIdx.x, BLOCK}, c, hMax, dbMin, dbOwn);
                    reduce_h(c, stiffnessBound, w2Own, dbMin, dbOwn, mbMin, mbOwn, history, threadIdx.x);
                }
                bool finished = !(c.t0 - c.t1 > 0.0f);
                if (!(c.h > 0.0f)) {
                    if (c.limit || finished || !c.rebuild) break;
                    c.c0 = 0; c.rebuild = 0;
                    ov2 = max(ov2, ov0); ov0 = 0; ov3 += 1;
                    __syncthreads();
                    ok = build_grid<BLOCK>(W, sh, c.t0 - c.t1, limit, rs, entriesTotal);
                    ov0 += rs.dropped; ov1 = max(ov1, rs.maxDemand); pairs = rs.pairs;
                    if (!ok) { status = BP_ERR_CAPACITY_OVERFLOW; break; }
                    pass_forces<BLOCK>(W, sh.red, Range{threadIdx.x, BLOCK}, c, 2u, w2Max, w2Own, mbMin, mbOwn);
                    continue;
                }
                pass_drift<BLOCK>(W, Range{threadIdx.x, BLOCK}, c);
                __syncthreads();
                pass_forces<BLOCK>(W, sh.red, Range{threadIdx.x, BLOCK}, c, 1u, w2Max, w2Own, mbMin, mbOwn);
                if (d->injectionEnabled && c.c1 == d->injectionAfter + 1) {
                    for (uint32_t i = threadIdx.x; i < n; i += BLOCK)
                        if (W.id[i] == d->injectionObject) {
                            W.pos[i] = f2(d->injectionPosition[0], d->injectionPosition[1]);
                            W.publish_a(i);
                        }
                    __syncthreads();
                    pass_bounds<BLOCK>(W, sh.red, Range{threadIdx.x, BLOCK}, c, w2Max, w2Own, mbMin, mbOwn);
                }
                if (d->assertEverySubstep && !claimed) {
                    if (assert_contacts<BLOCK>(W, sh, c, st)) claimed = true;
                    __syncthreads();
                }
            }
            ov2 = max(ov2, ov0);
        }
    }
    // 4. Gated end of step (NativeMetalSolver.swift:837-945; Force:822-828).
    bool finished = c.limit == 0 && !(c.t0 - c.t1 > 0.0f);
    uint32_t removals = 0, wraps = 0;
    if (status == BP_OK && finished) {
        if (d->applyRegion) {                                   // PostSolve:536-566
            float2 lo = f2(d->region[0], d->region[1]), hi = f2(d->region[2], d->region[3]);
            for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {
                uint32_t f = W.flags[i];
                if (!(f & kEnabled) || (f & kStatic)) continue;
                float2 p = W.pos[i];
                float r = W.radius[i];
                float2 nearest = f2(clampf(p.x, lo.x, hi.x), clampf(p.y, lo.y, hi.y));
                if (len2(sub(p, nearest)) <= r) continue;
                if (!(f & kWrap)) { W.region[i] = BP_REGION_REMOVED; continue; }
                float2 span = adds(sub(hi, lo), 2.0f * r);
                float delta[2] = {0.0f, 0.0f};
                float pp[2] = {p.x, p.y}, rl[2] = {lo.x, lo.y}, rh[2] = {hi.x, hi.y}, sp[2] = {span.x, span.y};
                for (int axis = 0; axis < 2; ++axis) {
                    if (pp[axis] > rh[axis]) delta[axis] = -sp[axis] * fmaxf(1.0f, ceilf((pp[axis] - r - rh[axis]) / sp[axis]));
                    else if (pp[axis] < rl[axis]) delta[axis] = sp[axis] * fmaxf(1.0f, ceilf((rl[axis] - r - pp[axis]) / sp[axis]));
                }
                float2 dl = f2(delta[0], delta[1]);
                float2 wrapped = add(p, dl);
                float2 nearestWrapped = f2(clampf(wrapped.x, lo.x, hi.x), clampf(wrapped.y, lo.y, hi.y));
                if (len2(sub(wrapped, nearestWrapped)) > r) dl = add(dl, sub(nearestWrapped, wrapped));
                float2 np = add(p, dl), ns = add(W.start[i], dl);
                if (!finite2(np) || !finite2(ns)) { safety |= 1u; continue; }
                W.pos[i] = np; W.start[i] = ns;
                W.region[i] = BP_REGION_WRAPPED;
            }
        }
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) {     // PostSolve:5-27
            if (!finite2(W.pos[i]) || !finite2(W.start[i]) || !finite2(W.vel[i]) || !isfinite(W.rot[i]) ||
                !isfinite(W.omega[i]) || !isfinite(W.mass[i]) || !isfinite(W.invmRaw[i]) || !isfinite(W.radius[i]))
                safety |= 1u;
        }
        for (uint32_t j = threadIdx.x; j < W.ns; j += BLOCK) {
            Seg s = W.segment(j);
            float2 e = sub(s.b, s.a);
            if (!finite2(s.a) || !finite2(s.b) || !finite2(e) || !isfinite(dot2(e, e))) safety |= 1u;
        }
        __syncthreads();
        if (!claimed && n > 0 && assert_contacts<BLOCK>(W, sh, c, st)) claimed = true;
    }
    // Safety bits of every lane.
    __syncthreads();
    if (threadIdx.x == 0) sh.red.counters[5] = 0;
    __syncthreads();
    if (safety) atomicOr(&sh.red.counters[5], safety);
    uint32_t rmLocal = 0, wrLocal = 0;
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK) { rmLocal += W.region[i] == BP_REGION_REMOVED; wrLocal += W.region[i] == BP_REGION_WRAPPED; }
    if (threadIdx.x == 0) { sh.red.counters[6] = 0; sh.red.counters[7] = 0; }
    __syncthreads();
    atomicAdd(&sh.red.counters[6], rmLocal);
    atomicAdd(&sh.red.counters[7], wrLocal);
    __syncthreads();
    removals = sh.red.counters[6]; wraps = sh.red.counters[7];
    uint32_t safetyAll = sh.red.counters[5];
    // Error precedence (NativeMetalSolver.swift:1026-1040) and the host finite check (:1067).
    if (status == BP_OK) {
        if (c.limit) status = BP_ERR_SUBSTEP_LIMIT_EXCEEDED;
        else if (safetyAll & 1u) status = BP_ERR_NON_FINITE;
        else if (claimed) status = BP_ERR_CONTACT_VIOLATION;
        else if (n > 0 && !finished) status = BP_ERR_COMMAND_FAILED;   // loop ended without finishing
    }
    if (d->mirror && status == BP_OK && n <= d->mirrorCapacity) {
        for (uint32_t i = threadIdx.x; i < n; i += BLOCK) mirror_record(A, d, wside, i, W.pos[i], W.vel[i], W.rot[i], W.omega[i]);
        mirror_bodies(A, d, wside, threadIdx.x, BLOCK);
    }
    if (threadIdx.x == 0) {
        st->status = status;
        st->substeps = c.c1;
        st->rebuilds = ov3;
        st->candidatePairs = pairs;
        st->candidateOverflow = ov2;
        st->maxCandidates = ov1;
        st->removals = removals;
        st->wraps = wraps;
        st->time[0] = c.t0; st->time[1] = c.t1; st->time[2] = c.h; st->time[3] = c.split;
        st->limiterObject = limiter_object(c, W.id);
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
            d->pub = wside;
            if (removals != 0) d->failed = kWorldHeld;    // removed bubbles leave before the next step
        } else if (status != BP_ERR_CAPACITY_OVERFLOW) {
            d->failed = kWorldFailed;
        }
        post_status(st, notify);
    }
}

}  // namespace bpd

// Host launchers (used by bp_api.cpp).
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_world_step(int block, const bpd::Arena *arena, bpd::WorldDesc *descs, bpd::WorldStatus *stats,
                                           const uint32_t *order, uint32_t count, uint32_t notify, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    switch (block) {
        case 256: hipLaunchKernelGGL(bpd::world_step_kernel<256>, dim3(count), dim3(256), 0, stream, arena, descs, stats, order, count, notify); break;
        case 512: hipLaunchKernelGGL(bpd::world_step_kernel<512>, dim3(count), dim3(512), 0, stream, arena, descs, stats, order, count, notify); break;
        case 1024: hipLaunchKernelGGL(bpd::world_step_kernel<1024>, dim3(count), dim3(1024), 0, stream, arena, descs, stats, order, count, notify); break;
        default: hipLaunchKernelGGL(bpd::world_step_kernel<64>, dim3(count), dim3(64), 0, stream, arena, descs, stats, order, count, notify); break;
    }
    return hipGetLastError();
}
