// bp_world_lds.hip -- LDS-resident world step: one workgroup, one world of up to BLOCK*PER bubbles.
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
//   * a substep is three passes (kick+drift; forces+kick+omega^2+margin bound; d/4 bound)
//     separated by workgroup barriers, with global memory touched only for segments, bodies and
//     force records (small, cached) and by the final write-back of the state.
// Every quantity is computed with the same operations, in the same order, as the global-state
// kernel and the CPU oracle, so all three give the same bits.
#include "bp_common.h"

namespace bpd {

constexpr uint32_t kLdsKindShift = 14;
constexpr uint32_t kLdsIndexMask = (1u << 14) - 1;
constexpr uint32_t kLdsIndexLimit = 1u << 14;

template <int BLOCK, int PER, int ROWCAP>
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
};

template <int BLOCK, int PER, int ROWCAP>
__global__ __launch_bounds__(BLOCK) void world_lds_kernel(const Arena *__restrict__ A, WorldDesc *descs, WorldStatus *stats,
                                                          const uint32_t *order, uint32_t count) {
    using L = LdsWorld<BLOCK, PER, ROWCAP>;
    __shared__ L S;
    if (blockIdx.x >= count) return;
    const uint32_t wi = order[blockIdx.x];
    WorldDesc *d = descs + wi;
    WorldStatus *st = stats + wi;
    const uint32_t lane = threadIdx.x;
    if (d->failed) {
        if (lane == 0) st->executed = 0;
        return;
    }
    const uint32_t pub = d->pub, ws = 1u - pub;
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
    float2 p[PER], v[PER], a[PER], ref[PER];
    float rot[PER], om[PER], mg[PER];
    uint32_t fl[PER], rs[PER], rl[PER];
    bool valid[PER];
#pragma unroll
    for (int s = 0; s < PER; ++s) {
        uint32_t i = lane + s * BLOCK;
        valid[s] = i < n;
        p[s] = v[s] = a[s] = ref[s] = f2(0.0f, 0.0f);
        rot[s] = om[s] = mg[s] = 0.0f;
        fl[s] = 0; rs[s] = rl[s] = 0;
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
    }
    uint32_t safety = predict_bodies<BLOCK>(A, d, pub, ws);
    __syncthreads();

    Ctl c;
    c.t0 = d->dt; c.t1 = n == 0 ? d->dt : 0.0f; c.h = 0.0f; c.split = (float)(d->split < 1 ? 1u : d->split);
    c.c0 = c.c1 = c.rebuild = c.limit = 0; c.limObj = 0; c.limTerm = BP_TERM_REMAINING;
    uint32_t ov0 = 0, ov1 = 0, ov2 = 0, ov3 = 0, pairs = 0, entriesTotal = 0;
    int32_t status = BP_OK;
    bool claimed = false;
    float cellWidth = 0.0f;
    RowRef rows;
    rows.lds = S.rows;
    rows.global = A->adj + d->adjOffset;
    rows.inLds = true;
    const bool ldsRowsAllowed = ROWCAP > 0 && g.ns < kLdsIndexLimit && g.nb < kLdsIndexLimit;

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
            const bool ownRow = finite2(startG[i]);
            const int64_t cx = cell_coord(pi.x, cellWidth), cy = cell_coord(pi.y, cellWidth);
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    const int64_t tx = cx + dx, ty = cy + dy;
                    const uint32_t b = bucket_of(tx, ty, buckets);
                    const uint32_t lo = b == 0 ? 0u : S.buckets[b - 1], hi = S.buckets[b];
                    for (uint32_t k = lo; k < hi; ++k) {
                        const uint32_t j = S.grid[k];
                        if (j == i) continue;
                        const float2 q = S.pos[j];
                        if (cell_coord(q.x, cellWidth) != tx || cell_coord(q.y, cellWidth) != ty) continue;
                        const uint32_t l = min(i, j), h = max(i, j);
                        const float2 pl = l == i ? pi : q, ph = l == i ? q : pi;
                        const float ml = l == i ? mi : S.aux[j], mh = l == i ? S.aux[j] : mi;
                        const float reach = S.r[l] + S.r[h] + ml + mh + sepTol;
                        if (!(fabsf(pl.x - ph.x) <= reach && fabsf(pl.y - ph.y) <= reach)) continue;
                        if (!filter_accepts(catG[l], maskG[l], catG[h], maskG[h])) continue;
                        if (j > i) {
                            if (!ownRow) continue;
                            ++forward;
                        } else if (!finite2(startG[j])) {
                            continue;
                        }
                        emit(j);
                    }
                }
            if (ownRow)
                for (uint32_t j = 0; j < g.ns; ++j)
                    if (g.segment_candidate(pi, ri, mi, ci, ki, j)) { ++forward; emit((1u << 30) | j); }
        }
        for (uint32_t k = 0; k < g.nb; ++k)
            if (g.body_candidate(pi, ri, mi, ci, ki, k)) emit((2u << 30) | k);
    };

    // Grid build (Force:1040-1134) with CSR rows. Returns false when the rows do not fit.
    auto build = [&]() __attribute__((always_inline)) -> bool {
        float remaining = fmaxf(c.t0 - c.t1, 0.0f);
        float width = sepTol;
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            if (!valid[s]) continue;
            uint32_t i = lane + s * BLOCK;
            float2 a0 = g.non_contact(fl[s], S.invm[i], A->band[o + i], A->motor[o + i], p[s]);
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
        uint32_t droppedLocal = 0, maxLocal = 0, pairsLocal = 0, mine = 0;
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            if (!valid[s]) continue;
            uint32_t cnt = 0, forward = 0;
            enumerate(lane + s * BLOCK, p[s], mg[s], [&](uint32_t e) __attribute__((always_inline)) {
                ++cnt;
                if ((e >> 30) == 2u) atomicAdd(&g.bodyHits()[e & kIndexMask], 1u);
            }, forward);
            rl[s] = cnt;
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
        uint32_t total;
        uint32_t base = block_exclusive_scan<BLOCK>(S.red, mine, total);
#pragma unroll
        for (int s = 0; s < PER; ++s) { rs[s] = base; base += rl[s]; }
        entriesTotal = total;
        ov0 += S.red.counters[0];
        ov1 = max(ov1, S.red.counters[1]);
        pairs = S.red.counters[2];
        rows.inLds = ldsRowsAllowed && total <= (uint32_t)ROWCAP;
        if (!rows.inLds && total > d->adjCapacity) return false;
        // Pass 2: fill the rows and order the pair block (Contact:222-238).
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            if (!valid[s]) continue;
            uint32_t forward, cnt = 0;
            if (rows.inLds) {
                uint16_t *row = S.rows + rs[s];
                enumerate(lane + s * BLOCK, p[s], mg[s], [&](uint32_t e) __attribute__((always_inline)) {
                    row[cnt++] = (uint16_t)(((e >> 30) << kLdsKindShift) | (e & kIndexMask));
                }, forward);
                uint32_t pc = 0;
                while (pc < cnt && (row[pc] >> kLdsKindShift) == 0u) ++pc;
                for (uint32_t x = 1; x < pc; ++x) {
                    uint16_t val = row[x];
                    uint32_t loc = x;
                    while (loc > 0 && val < row[loc - 1]) { row[loc] = row[loc - 1]; --loc; }
                    row[loc] = val;
                }
            } else {
                uint32_t *row = A->adj + d->adjOffset + rs[s];
                enumerate(lane + s * BLOCK, p[s], mg[s], [&](uint32_t e) __attribute__((always_inline)) { row[cnt++] = e; }, forward);
                uint32_t pc = 0;
                while (pc < cnt && (row[pc] >> kKindShift) == 0u) ++pc;
                for (uint32_t x = 1; x < pc; ++x) {
                    uint32_t val = row[x];
                    uint32_t loc = x;
                    while (loc > 0 && val < row[loc - 1]) { row[loc] = row[loc - 1]; --loc; }
                    row[loc] = val;
                }
            }
        }
        __syncthreads();
        return true;
    };

    // F: forces, kick (phase 1) or initial acceleration (phase 2), omega^2, margin bound.
    float w2Max = -1.0f, mbMin = inff(), dbMin = inff();
    uint32_t w2Own = 0, mbOwn = 0, dbOwn = 0;
    auto pass_forces = [&](uint32_t phase, bool kick) __attribute__((always_inline)) {
        const float fraction = step_fraction(c);
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
            const int32_t bi = A->band[o + i];
            if (invm > 0.0f && bi >= 0) w2 = w2 + g.band(bi).z * invm;
            const uint32_t start = rs[s], len = rl[s];
            for (uint32_t e = 0; e < len; ++e) {
                uint32_t kind, j;
                rows.get(start + e, kind, j);
                if (kind == 0u) {
                    const uint32_t lo = min(i, j), hi = max(i, j);
                    const float2 delta = sub(S.pos[hi], S.pos[lo]);
                    const float dd = len2(delta);
                    const float2 normal = dd > 0.0f ? dvs(delta, dd) : f2(1.0f, 0.0f);
                    const Law law = pair_law(S.K[lo], S.r[lo], S.K[hi], S.r[hi], dd);
                    total = add(total, mul(i == lo ? -law.force : law.force, normal));
                    if (invm > 0.0f) w2 = w2 + law.stiffness * (invm + __builtin_sqrtf(invm * S.invm[j]));
                } else if (kind == 1u) {
                    float2 normal; float dd;
                    if (!wall_region(g.segment(j), pi, normal, dd) || !(dd < ri)) continue;
                    const Law law = wall_law(Ki, ri, dd);
                    total = add(total, mul(law.force, normal));
                    if (invm > 0.0f) w2 = w2 + law.stiffness * invm;
                } else {
                    const Obstacle ob = body_obstacle(g.body(j), fraction, g.verts(), g.tris(), pi);
                    const Law law = wall_law(Ki, ri, ob.distanceValue);
                    total = add(total, mul(law.force, ob.normal));
                    if (invm > 0.0f) w2 = w2 + law.stiffness * invm;
                }
            }
            if (kick) {
                if (phase == 2) {
                    a[s] = add(mul(total, invm), g.non_contact(fl[s], invm, bi, A->motor[o + i], pi));
                } else if (invm > 0.0f) {
                    const float gamma = (fl[s] & kForceDriven) ? fmaxf(A->gamma[o + i], 0.0f) : 0.0f;
                    const float decay = bpm::expf_det(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
                    a[s] = add(mul(total, invm), g.non_contact(fl[s], invm, bi, A->motor[o + i], pi));
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
        float bestD = inff();
        uint32_t ownD = 0;
#pragma unroll
        for (int s = 0; s < PER; ++s) {
            if (!valid[s]) continue;
            const uint32_t i = lane + s * BLOCK;
            if (!(S.invm[i] > 0.0f)) continue;
            const float2 pi = p[s], vi = v[s];
            const float ri = S.r[i], accI = S.aux[i];
            float best = inff();
            const uint32_t start = rs[s], len = rl[s];
            for (uint32_t e = 0; e < len; ++e) {
                uint32_t kind, j;
                rows.get(start + e, kind, j);
                float dd, approach, accSum, reach, bodySpeed = 0.0f;
                if (kind == 0u) {
                    const float2 delta = sub(S.pos[j], pi);
                    dd = len2(delta);
                    const float2 normal = dd > 0.0f ? dvs(delta, dd) : f2(1.0f, 0.0f);
                    approach = dot2(sub(vi, S.vel[j]), normal);
                    accSum = accI + S.aux[j];
                    reach = ri + S.r[j];
                } else if (kind == 1u) {
                    float2 normal;
                    if (!wall_region(g.segment(j), pi, normal, dd)) continue;
                    approach = -dot2(vi, normal);
                    accSum = accI;
                    reach = ri;
                } else {
                    const Obstacle ob = body_obstacle(g.body(j), fraction, g.verts(), g.tris(), pi);
                    dd = ob.distanceValue;
                    approach = -dot2(vi, ob.normal);
                    bodySpeed = ob.speedBound;
                    accSum = accI;
                    reach = ri;
                }
                const bool active = dd < reach;
                const float dc = fmaxf(dd, reach * 1e-6f);
                float speed;
                if (active) speed = fabsf(approach) + bodySpeed;
                else {
                    if (approach + bodySpeed + accSum * hMax * 0.5f <= 0.0f) continue;
                    speed = approach + bodySpeed;
                }
                const float sq = __builtin_sqrtf(speed * speed + accSum * dc * 0.5f);
                const float den = sq + speed;
                if (!(den > 0.0f)) continue;
                best = fminf(best, dc / (2.0f * den));
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
            const bool driven = (fl[s] & kForceDriven) != 0;
            const float gamma = driven ? fmaxf(A->gamma[o + i], 0.0f) : 0.0f;
            const float decay = bpm::expf_det(-gamma * halfStep), factor = half_impulse_factor(gamma, halfStep);
            const float2 vn = add(mul(v[s], decay), mul(a[s], factor));
            float angular = om[s] * bpm::expf_det(-gamma * h);
            const int32_t mi = A->motor[o + i];
            if (driven && mi >= 0) {
                const float4 md = g.motor(mi);
                if (md.w > 0.0f) angular = angular + md.z / md.w * half_impulse_factor(gamma, h);
            }
            p[s] = add(p[s], mul(vn, h));
            v[s] = vn;
            rot[s] = rot[s] + om[s] * h;
            om[s] = angular;
            S.pos[i] = p[s];
        }
    };

    // Contact assertion (Force:865-918): lowest violating bubble writes the record.
    auto assert_contacts = [&]() __attribute__((always_inline)) -> bool {
        if (!(cellWidth > 0.0f)) cellWidth = 1.0f;
        __syncthreads();
        build_hash();
        const float fraction = step_fraction(c);
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
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy) {
                    const int64_t tx = cx + dx, ty = cy + dy;
                    const uint32_t b = bucket_of(tx, ty, buckets);
                    const uint32_t lo = b == 0 ? 0u : S.buckets[b - 1], hi = S.buckets[b];
                    for (uint32_t k = lo; k < hi; ++k) {
                        const uint32_t j = S.grid[k];
                        if (j <= i || j >= bestJ) continue;
                        const float2 q = S.pos[j];
                        if (cell_coord(q.x, cellWidth) != tx || cell_coord(q.y, cellWidth) != ty) continue;
                        const float dd = len2(sub(q, center));
                        if (!(dd <= 0.0f)) continue;
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
            if (assert_body_and_edges(g, center, startG[i], catG[i], maskG[i], fraction, mine)) firstIndex = i;
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
                    reduce_h(c, stiffnessBound, w2Own, dbMin, dbOwn, mbMin, mbOwn, ids, history, lane);
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
                if (d->injectionEnabled && c.c1 == d->injectionAfter + 1) {
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
                if (d->assertEverySubstep && !claimed) {
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

    // ---- Write-back of the working side, counters and publication.
    uint32_t rmLocal = 0, wrLocal = 0;
#pragma unroll
    for (int s = 0; s < PER; ++s) {
        if (!valid[s]) continue;
        const uint32_t gi = o + lane + s * BLOCK;
        A->pos[ws][gi] = p[s];
        A->vel[ws][gi] = v[s];
        A->rot[ws][gi] = rot[s];
        A->omega[ws][gi] = om[s];
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
        st->limiterObject = c.limObj;
        st->limiterTerm = c.limTerm;
        if (!claimed) {
            st->violationKind = 0; st->violationSegment = 0; st->violationObjects[0] = 0; st->violationObjects[1] = 0;
            st->violationDetail = 0.0f;
        }
        st->adjacencyEntries = entriesTotal;
        st->groupWidth = BLOCK;
        st->executed = 1;
        if (status == BP_OK) d->pub = ws;
        else if (status != BP_ERR_CAPACITY_OVERFLOW) d->failed = 1;
    }
#undef ids
#undef catG
#undef maskG
#undef startG
}

}  // namespace bpd

// Launch configurations of the LDS kernel: (threads, bubbles per thread, LDS row capacity).
#define BP_LDS_CONFIGS(X)      \
    X(0, 64, 1, 1024)          \
    X(1, 128, 1, 2048)         \
    X(2, 256, 1, 4096)         \
    X(3, 512, 1, 8192)         \
    X(4, 1024, 1, 16384)       \
    X(5, 256, 2, 8192)         \
    X(6, 512, 2, 16384)        \
    X(7, 1024, 2, 24576)       \
    X(8, 256, 4, 16384)        \
    X(9, 1024, 4, 0)

extern "C" __attribute__((visibility("hidden"))) int bp_lds_config_count() { return 10; }

extern "C" __attribute__((visibility("hidden"))) void bp_lds_config(int id, uint32_t *threads, uint32_t *per, uint32_t *rowcap) {
#define BP_LDS_DESCRIBE(ID, B, P, R) \
    if (id == ID) { *threads = B; *per = P; *rowcap = R; return; }
    BP_LDS_CONFIGS(BP_LDS_DESCRIBE)
#undef BP_LDS_DESCRIBE
    *threads = 0; *per = 0; *rowcap = 0;
}

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_world_lds(int id, const bpd::Arena *arena, bpd::WorldDesc *descs,
                                                                           bpd::WorldStatus *stats, const uint32_t *order,
                                                                           uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
#define BP_LDS_LAUNCH(ID, B, P, R)                                                                                       \
    if (id == ID) {                                                                                                      \
        hipLaunchKernelGGL((bpd::world_lds_kernel<B, P, R>), dim3(count), dim3(B), 0, stream, arena, descs, stats, order, count); \
        return hipGetLastError();                                                                                        \
    }
    BP_LDS_CONFIGS(BP_LDS_LAUNCH)
#undef BP_LDS_LAUNCH
    return hipErrorInvalidValue;
}
