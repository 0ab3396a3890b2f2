// bp_world_coop.hip -- one large world advanced by a cooperative grid of workgroups (gfx950).
//
// The reference splits a world above one workgroup into many dispatches per substep with a host
// loop between them (NativeMetalSolver.swift:801-833, NativeMetalSubstepLoop.swift:195-273). Here
// the whole step is one cooperative launch: every lane of the grid owns bubbles (grid-stride),
// the state lives in global memory (structure of arrays), and grid barriers separate the phases.
// A substep costs three barriers:
//   H: d/4 bound, workgroup minima -> grid barrier -> every workgroup combines the minima;
//   D: kick and drift -> grid barrier;
//   F: forces, kick, omega^2, margin bound, workgroup extrema -> grid barrier -> combine.
// The h rule then runs redundantly in every lane on identical values, so the control flow is
// uniform across the grid and every workgroup reaches every barrier.
//
// The grid build is a counting sort of the centres into hashed buckets in global memory
// (histogram, grid-wide exclusive scan, scatter), then CSR rows in two passes (lengths,
// grid-wide scan, fill). Bodies are found through a linear BVH over their sweep boxes
// (body_bvh_kernel, built once per step before the cooperative launch); the exact candidate
// predicate (Geometry:361-379) is applied at the leaves, so the rows equal the brute-force rows.
#include <hip/hip_cooperative_groups.h>

#include "bp_world_global.h"

namespace cg = cooperative_groups;

namespace bpd {

// ---------------------------------------------------------------------------------------
// Linear BVH of the bodies of one world.
// ---------------------------------------------------------------------------------------
constexpr uint32_t kBvhMaxBodies = 8192;
constexpr uint32_t kBvhLeaf = 0x80000000u;
constexpr int kBvhThreads = 1024;
constexpr int kBvhStack = 72;          // a Karras tree over unique 64-bit keys is at most 64 deep

BPD uint32_t spread16(uint32_t x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Sweep box of body k for the step, enlarged so that it also holds the body's outline at every
// step fraction (the largest vertex distance, plus the core radius of a round body). The
// enlargement only makes the tree conservative; every leaf is tested with the exact predicate.
BPD float4 body_sweep_box(const Arena *A, const WorldDesc *d, uint32_t side, uint32_t k) {
    const uint32_t g = d->bodyOffset + k;
    const float4 pose = A->bodyPose[side][g];
    const float2 prev = A->bodyPrev[side][g];
    const float4 shape = A->bodyShape[g];
    const uint4 ranges = A->bodyRanges[g];
    const uint4 kind = A->bodyKind[g];
    float outline = 0.0f;
    for (uint32_t v = 0; v < ranges.y; ++v) outline = fmaxf(outline, len2(A->verts[d->vertexOffset + ranges.x + v]));
    if (kind.x != 0u) outline = outline + shape.y;
    const float reach = fmaxf(shape.x, outline * 1.0001f + 0.01f);
    const float2 lo = subs(min2(f2(pose.x, pose.y), prev), reach), hi = adds(max2(f2(pose.x, pose.y), prev), reach);
    return make_float4(lo.x, lo.y, hi.x, hi.y);
}

BPD int prefix_length(const uint64_t *keys, uint32_t count, int64_t i, int64_t j) {
    if (j < 0 || j >= (int64_t)count) return -1;
    return __builtin_clzll(keys[i] ^ keys[j]);
}

// One workgroup per cooperative world: predicts the bodies for the step (PostSolve:66-82) and
// builds the LBVH of their sweep boxes (Karras 2012: Morton order, then each internal node
// from the longest common prefix of its key range).
__global__ __launch_bounds__(kBvhThreads) void body_bvh_kernel(const Arena *__restrict__ A, WorldDesc *descs, const uint32_t *order) {
    __shared__ uint64_t keys[kBvhMaxBodies];
    __shared__ uint16_t parentNode[kBvhMaxBodies], parentLeaf[kBvhMaxBodies];
    __shared__ uint32_t arrivals[kBvhMaxBodies];
    __shared__ Red<kBvhThreads> red;
    __shared__ uint32_t flag;
    WorldDesc *d = descs + order[blockIdx.x];
    if (d->failed) return;
    const uint32_t pub = d->pub, ws = 1u - pub;
    uint32_t *counters = A->coopScratch + d->coopScratchOffset + 8 * kCoopMaxGroups;
    if (threadIdx.x == 0) flag = 0;
    __syncthreads();
    const uint32_t safety = predict_bodies<kBvhThreads>(A, d, pub, ws);
    if (safety) atomicOr(&flag, safety);
    const uint32_t nb = d->bodyCount;
    const uint32_t base = 2 * d->bodyOffset;
    // Leaf boxes by body index, from the poses this lane has just written.
    float2 lo = f2(inff(), inff()), hi = f2(-inff(), -inff());
    for (uint32_t k = threadIdx.x; k < nb; k += kBvhThreads) {
        const float4 box = body_sweep_box(A, d, ws, k);
        A->bvhBox[base + nb + k] = box;
        const float2 c = f2(0.5f * (box.x + box.z), 0.5f * (box.y + box.w));
        lo = min2(lo, c);
        hi = max2(hi, c);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        counters[kCoBodySafety] = flag;
        counters[kCoBvhBuilt] = 0;
    }
    if (nb < 2 || nb > kBvhMaxBodies) return;              // brute force (or a single leaf)
    lo.x = -block_maxf<kBvhThreads>(red, -lo.x);
    lo.y = -block_maxf<kBvhThreads>(red, -lo.y);
    hi.x = block_maxf<kBvhThreads>(red, hi.x);
    hi.y = block_maxf<kBvhThreads>(red, hi.y);
    const float2 span = f2(fmaxf(hi.x - lo.x, 1e-30f), fmaxf(hi.y - lo.y, 1e-30f));
    uint32_t padded = 1;
    while (padded < nb) padded <<= 1;
    for (uint32_t k = threadIdx.x; k < padded; k += kBvhThreads) {
        if (k >= nb) { keys[k] = ~0ull; continue; }
        const float4 box = A->bvhBox[base + nb + k];
        const float2 c = f2(0.5f * (box.x + box.z), 0.5f * (box.y + box.w));
        float qx = (c.x - lo.x) / span.x * 65535.0f, qy = (c.y - lo.y) / span.y * 65535.0f;
        qx = isfinite(qx) ? fminf(fmaxf(qx, 0.0f), 65535.0f) : 0.0f;
        qy = isfinite(qy) ? fminf(fmaxf(qy, 0.0f), 65535.0f) : 0.0f;
        const uint32_t morton = spread16((uint32_t)qx) | (spread16((uint32_t)qy) << 1);
        keys[k] = ((uint64_t)morton << 32) | k;
    }
    __syncthreads();
    // Bitonic sort of the padded key array in LDS.
    for (uint32_t size = 2; size <= padded; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < padded / 2; t += kBvhThreads) {
                const uint32_t i = 2 * t - (t & (stride - 1));
                const uint32_t j = i + stride;
                const bool up = (i & size) == 0;
                const uint64_t a = keys[i], b = keys[j];
                if ((a > b) == up) { keys[i] = b; keys[j] = a; }
            }
            __syncthreads();
        }
    // Internal nodes: children and key ranges.
    for (uint32_t i = threadIdx.x; i + 1 < nb; i += kBvhThreads) {
        const int64_t ii = i;
        const int dir = prefix_length(keys, nb, ii, ii + 1) - prefix_length(keys, nb, ii, ii - 1) > 0 ? 1 : -1;
        const int minPrefix = prefix_length(keys, nb, ii, ii - dir);
        int64_t lmax = 2;
        while (prefix_length(keys, nb, ii, ii + lmax * dir) > minPrefix) lmax <<= 1;
        int64_t l = 0;
        for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
            if (prefix_length(keys, nb, ii, ii + (l + t) * dir) > minPrefix) l += t;
        const int64_t j = ii + l * dir;
        const int nodePrefix = prefix_length(keys, nb, ii, j);
        int64_t s = 0;
        for (int64_t div = 2;; div <<= 1) {
            const int64_t t = (l + div - 1) / div;
            if (prefix_length(keys, nb, ii, ii + (s + t) * dir) > nodePrefix) s += t;
            if (t <= 1) break;
        }
        const int64_t split = ii + s * dir + (dir < 0 ? -1 : 0);
        const int64_t first = min(ii, j), last = max(ii, j);
        const uint32_t left = first == split ? (kBvhLeaf | (uint32_t)(keys[split] & 0xffffffffu)) : (uint32_t)split;
        const uint32_t right = last == split + 1 ? (kBvhLeaf | (uint32_t)(keys[split + 1] & 0xffffffffu)) : (uint32_t)(split + 1);
        A->bvhChild[base + i] = make_uint2(left, right);
        if (left & kBvhLeaf) parentLeaf[left & ~kBvhLeaf] = (uint16_t)i;
        else parentNode[left] = (uint16_t)i;
        if (right & kBvhLeaf) parentLeaf[right & ~kBvhLeaf] = (uint16_t)i;
        else parentNode[right] = (uint16_t)i;
        arrivals[i] = 0u;
    }
    __syncthreads();
    // Node boxes bottom-up: each leaf climbs towards the root; at every node the first child to
    // arrive stops and the second writes the union of both children's boxes (exact min and max;
    // a leaf with a non-finite box makes every box above it the whole plane, so it overlaps every
    // query), so each box is written once, after both of its children.
    for (uint32_t k = threadIdx.x; k < nb; k += kBvhThreads) {
        uint32_t node = parentLeaf[k];
        for (;;) {
            __threadfence();                                  // this lane's box before its arrival
            if (atomicAdd(&arrivals[node], 1u) == 0u) break;
            __threadfence();                                  // the sibling's box after its arrival
            const uint2 ch = A->bvhChild[base + node];
            float4 box = make_float4(inff(), inff(), -inff(), -inff());
            const uint32_t kids[2] = {ch.x, ch.y};
            for (int c = 0; c < 2; ++c) {
                const bool leaf = (kids[c] & kBvhLeaf) != 0u;
                const float4 b = leaf ? A->bvhBox[base + nb + (kids[c] & ~kBvhLeaf)] : A->bvhBox[base + kids[c]];
                if (leaf && (isnan(b.x) || isnan(b.y) || isnan(b.z) || isnan(b.w))) {
                    box = make_float4(-inff(), -inff(), inff(), inff());
                    break;
                }
                box = make_float4(fminf(box.x, b.x), fminf(box.y, b.y), fmaxf(box.z, b.z), fmaxf(box.w, b.w));
            }
            A->bvhBox[base + node] = box;
            if (node == 0u) break;
            node = parentNode[node];
        }
    }
    for (uint32_t q = threadIdx.x; q < nb; q += kBvhThreads) A->bvhLeaf[d->bodyOffset + q] = (uint32_t)(keys[q] & 0xffffffffu);
    __syncthreads();
    if (threadIdx.x == 0) counters[kCoBvhBuilt] = 1;
}

BPD bool box_overlaps(float4 box, float2 lo, float2 hi) {       // the comparison of Geometry:375-378
    return !(hi.x < box.x || hi.y < box.y || box.z < lo.x || box.w < lo.y);
}

// Visits every body whose tree box overlaps [lo, hi] (a superset of the exact candidates).
template <class Visit>
BPD void bvh_query(const Arena *A, uint32_t base, uint32_t nb, float2 lo, float2 hi, Visit visit) {
    uint32_t stack[kBvhStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const uint2 ch = A->bvhChild[base + stack[--sp]];
        const uint32_t kids[2] = {ch.x, ch.y};
        for (int c = 0; c < 2; ++c) {
            const uint32_t node = kids[c];
            if (node & kBvhLeaf) {
                const uint32_t k = node & ~kBvhLeaf;
                if (box_overlaps(A->bvhBox[base + nb + k], lo, hi)) visit(k);
            } else if (box_overlaps(A->bvhBox[base + node], lo, hi) && sp < kBvhStack) {
                stack[sp++] = node;
            }
        }
    }
}

// Body rows through the tree when it was built, else brute force.
struct TreeBodies {
    const Arena *A;
    uint32_t base, nb;
    bool built;
    template <class Emit> BPD void operator()(const World &W, uint32_t i, Emit emit) const {
        if (!built) { AllBodies()(W, i, emit); return; }
        const float reach = W.radius[i] + fmaxf(W.margin[i], 0.0f);
        const float2 p = W.pos[i];
        bvh_query(A, base, nb, subs(p, reach), adds(p, reach), [&](uint32_t k) {
            if (W.body_candidate(i, k)) emit((2u << kKindShift) | k);
        });
    }
};

// ---------------------------------------------------------------------------------------
// Grid-wide primitives. Every workgroup combines the per-workgroup partials itself, in the same
// order, so every lane of the grid holds the same result.
// ---------------------------------------------------------------------------------------
struct CoopScratch {
    float *width, *a, *b, *d;      // cell width; omega^2 max and margin min; d/4 min (separate arrays, so that
    uint32_t *ao, *bo, *dO;        // a reduction never overwrites partials another workgroup is still reading)
    uint32_t *scan, *counters;
};

BPD CoopScratch coop_scratch(const Arena *A, const WorldDesc *d) {
    uint32_t *s = A->coopScratch + d->coopScratchOffset;
    CoopScratch c;
    c.width = reinterpret_cast<float *>(s);
    c.a = reinterpret_cast<float *>(s + kCoopMaxGroups);
    c.ao = s + 2 * kCoopMaxGroups;
    c.b = reinterpret_cast<float *>(s + 3 * kCoopMaxGroups);
    c.bo = s + 4 * kCoopMaxGroups;
    c.d = reinterpret_cast<float *>(s + 5 * kCoopMaxGroups);
    c.dO = s + 6 * kCoopMaxGroups;
    c.scan = s + 7 * kCoopMaxGroups;
    c.counters = s + 8 * kCoopMaxGroups;
    return c;
}

template <int BLOCK>
BPD uint32_t block_sum_u32(Red<BLOCK> &r, uint32_t v) {
    uint32_t total;
    (void)block_exclusive_scan<BLOCK>(r, v, total);
    return total;
}
template <int BLOCK>
BPD uint32_t block_max_u32(Red<BLOCK> &r, uint32_t v) { return ~block_min_u32<BLOCK>(r, ~v); }

// Max of (a, owner) and min of (b, owner) over the grid (ties to the lower index).
template <int BLOCK>
BPD void grid_maxmin(cg::grid_group &grid, Red<BLOCK> &red, const CoopScratch &s, float &a, uint32_t &ao, float &b, uint32_t &bo) {
    if (threadIdx.x == 0) { s.a[blockIdx.x] = a; s.ao[blockIdx.x] = ao; s.b[blockIdx.x] = b; s.bo[blockIdx.x] = bo; }
    grid.sync();
    a = -1.0f; ao = 0; b = inff(); bo = 0;
    for (uint32_t w = threadIdx.x; w < gridDim.x; w += BLOCK) {
        const float va = s.a[w], vb = s.b[w];
        const uint32_t oa = s.ao[w], ob = s.bo[w];
        if (va > a || (va == a && oa < ao)) { a = va; ao = oa; }
        if (vb < b || (vb == b && ob < bo)) { b = vb; bo = ob; }
    }
    block_maxmin<BLOCK>(red, a, ao, b, bo);
}
template <int BLOCK>
BPD void grid_min(cg::grid_group &grid, Red<BLOCK> &red, const CoopScratch &s, float &b, uint32_t &bo) {
    if (threadIdx.x == 0) { s.d[blockIdx.x] = b; s.dO[blockIdx.x] = bo; }
    grid.sync();
    b = inff(); bo = 0;
    for (uint32_t w = threadIdx.x; w < gridDim.x; w += BLOCK) {
        const float vb = s.d[w];
        const uint32_t ob = s.dO[w];
        if (vb < b || (vb == b && ob < bo)) { b = vb; bo = ob; }
    }
    block_min<BLOCK>(red, b, bo);
}

// Exclusive prefix sum of src[0, count) into dst (may alias src); returns the total. Each
// workgroup owns a contiguous chunk, each lane a contiguous part of it.
template <int BLOCK>
BPD uint32_t grid_scan(cg::grid_group &grid, Red<BLOCK> &red, const CoopScratch &s, const uint32_t *src, uint32_t *dst,
                       uint32_t count) {
    const uint32_t G = gridDim.x;
    const uint32_t chunk = (count + G - 1) / G;
    const uint32_t first = min(blockIdx.x * chunk, count), last = min(first + chunk, count);
    const uint32_t per = (last - first + BLOCK - 1) / BLOCK;
    const uint32_t a = min(first + threadIdx.x * per, last), b = min(a + per, last);
    uint32_t sum = 0;
    for (uint32_t k = a; k < b; ++k) sum += src[k];
    uint32_t groupTotal;
    const uint32_t offset = block_exclusive_scan<BLOCK>(red, sum, groupTotal);
    if (threadIdx.x == 0) s.scan[blockIdx.x] = groupTotal;
    grid.sync();
    uint32_t before = 0, all = 0;
    for (uint32_t w = threadIdx.x; w < G; w += BLOCK) {
        const uint32_t v = s.scan[w];
        all += v;
        if (w < blockIdx.x) before += v;
    }
    before = block_sum_u32<BLOCK>(red, before);
    all = block_sum_u32<BLOCK>(red, all);
    uint32_t run = before + offset;
    for (uint32_t k = a; k < b; ++k) { const uint32_t v = src[k]; dst[k] = run; run += v; }
    return all;
}

// Counting sort of the centres into the world's hashed buckets: after it, bucket b holds
// grid[ends[b-1], ends[b]) (ends[-1] = 0). Ends with a grid barrier.
template <int BLOCK>
BPD void coop_hash(cg::grid_group &grid, World &W, Red<BLOCK> &red, const CoopScratch &s, Range R, uint32_t *cells, uint32_t B) {
    for (uint32_t i = R.first; i < W.n; i += R.stride) {
        const float2 p = W.pos[i];
        if (!finite2(p)) continue;
        atomicAdd(&cells[bucket_of(W.cell_x(p), W.cell_y(p), B)], 1u);
    }
    grid.sync();
    (void)grid_scan<BLOCK>(grid, red, s, cells, cells, B);
    grid.sync();
    for (uint32_t i = R.first; i < W.n; i += R.stride) {
        const float2 p = W.pos[i];
        if (!finite2(p)) continue;
        const uint32_t slot = atomicAdd(&cells[bucket_of(W.cell_x(p), W.cell_y(p), B)], 1u);
        W.grid[slot] = i;
        W.publish_cell(slot, i);
    }
    grid.sync();
}

BPD void sort_run(uint32_t *row, uint32_t from, uint32_t to) {        // insertion sort of row[from, to)
    for (uint32_t a = from + 1; a < to; ++a) {
        const uint32_t v = row[a];
        uint32_t loc = a;
        while (loc > from && v < row[loc - 1]) { row[loc] = row[loc - 1]; --loc; }
        row[loc] = v;
    }
}

// Cuts the sorted rows to the reference's k candidates (forward_cut, mark_body_cuts, cut_row)
// after a build that dropped candidates, and replaces the build's counters by those of the
// reference's truncated rows (Geometry:291-297, 410-414; Contact:190-195). The grid slots hold
// the forward cuts (the contact assertion hashes the grid again).
template <int BLOCK>
BPD void coop_cut_rows(cg::grid_group &grid, World &W, Red<BLOCK> &red, const CoopScratch &s, Range R, uint32_t limit,
                       uint32_t &dropped, uint32_t &maxDemand, uint32_t &entries) {
    const uint32_t n = W.n;
    const uint32_t *hits = W.A->bodyHits + W.d->bodyOffset;
    uint32_t *bodyCut = W.A->bodyCut + W.d->bodyOffset;
    for (uint32_t k = R.first; k < W.nb; k += R.stride)          // one lane per body
        bodyCut[k] = body_cut_index(n, limit, hits[k], [&](uint32_t a) { return W.body_candidate(a, k); });
    grid.sync();
    for (uint32_t i = R.first; i < n; i += R.stride) {
        uint32_t *row = W.adj + W.rowStart[i];
        const uint32_t len = W.rowLen[i];
        uint32_t pairEnd, segEnd, forwardPairs;
        row_blocks(row, len, pairEnd, segEnd);
        W.grid[i] = forward_cut(row, pairEnd, i, limit, forwardPairs);
        mark_body_cuts(row, segEnd, len, i, bodyCut);
    }
    grid.sync();
    uint32_t droppedLocal = 0, widest = 0, kept = 0;
    for (uint32_t i = R.first; i < n; i += R.stride) {
        uint32_t *row = W.adj + W.rowStart[i];
        const uint32_t len = W.rowLen[i];
        uint32_t pairEnd, segEnd, forwardPairs;
        row_blocks(row, len, pairEnd, segEnd);
        forward_cut(row, pairEnd, i, limit, forwardPairs);
        const RowCut r = cut_row(row, pairEnd, segEnd, len, i, limit, forwardPairs, [&](uint32_t l) { return W.grid[l]; });
        W.rowLen[i] = r.len;
        const uint32_t forward = forwardPairs + (segEnd - pairEnd);
        droppedLocal += (forward > limit ? forward - limit : 0u) + (r.demand > limit ? r.demand - limit : 0u);
        widest = max(widest, max(forward, r.demand));
        kept += r.len;
    }
    for (uint32_t k = R.first; k < W.nb; k += R.stride) {
        const uint32_t h = hits[k];
        droppedLocal += h > limit ? h - limit : 0u;
        widest = max(widest, h);
    }
    droppedLocal = block_sum_u32<BLOCK>(red, droppedLocal);
    kept = block_sum_u32<BLOCK>(red, kept);
    widest = block_max_u32<BLOCK>(red, widest);
    if (threadIdx.x == 0) {
        atomicAdd(&s.counters[kCoCutDropped], droppedLocal);
        atomicMax(&s.counters[kCoCutWidest], widest);
        atomicAdd(&s.counters[kCoCutEntries], kept);
    }
    grid.sync();
    dropped = s.counters[kCoCutDropped];
    maxDemand = s.counters[kCoCutWidest];
    entries = s.counters[kCoCutEntries];
}

// Grid build of a cooperative world (Force:1040-1134 and Geometry:300-426 with CSR rows).
template <int BLOCK>
BPD bool coop_build(cg::grid_group &grid, World &W, Red<BLOCK> &red, const CoopScratch &s, Range R, uint32_t *cells,
                    uint32_t B, const TreeBodies &bodies, float remaining, uint32_t limit, uint32_t &dropped,
                    uint32_t &maxDemand, uint32_t &pairs, uint32_t &entries) {
    const uint32_t n = W.n;
    const WorldDesc &d = *W.d;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    float width = W.sepTol;
    for (uint32_t i = R.first; i < n; i += R.stride) {              // Force:1046-1049, 492-499
        const float2 p = W.pos[i];
        const float2 a = W.non_contact(i, p);
        W.a0[i] = a;
        const float t = fmaxf(remaining, 0.0f);
        const float travel = len2(W.eff_vel(i)) * t + 0.5f * len2(a) * t * t;
        const float m = d.minimumMarginOnly ? 0.25f * W.radius[i] : fmaxf(1.05f * travel, 0.25f * W.radius[i]);
        W.margin[i] = m;
        W.ref[i] = p;
        width = fmaxf(width, 2.0f * (W.radius[i] + m) + W.sepTol);
    }
    width = block_maxf<BLOCK>(red, width);
    if (threadIdx.x == 0) s.width[blockIdx.x] = width;
    for (uint32_t b = gtid; b < B; b += R.stride) cells[b] = 0;
    uint32_t *bodyHits = W.A->bodyHits + d.bodyOffset;
    for (uint32_t k = gtid; k < W.nb; k += R.stride) bodyHits[k] = 0;
    if (gtid == 0) {
        for (uint32_t k = kCoDropped; k <= kCoOverflowRows; ++k) s.counters[k] = 0;
        for (uint32_t k = kCoLongRows; k <= kCoCutEntries; ++k) s.counters[k] = 0;
    }
    grid.sync();
    width = W.sepTol;
    for (uint32_t w = threadIdx.x; w < gridDim.x; w += BLOCK) width = fmaxf(width, s.width[w]);
    width = block_maxf<BLOCK>(red, width);
    W.cellWidth = fmaxf(width, kFloatMin);
    coop_hash<BLOCK>(grid, W, red, s, R, cells, B);
    // Row lengths and the reference's overflow counters (Geometry:291-297, 410-414). Each row
    // is also staged at a fixed stride of kStageStride entries per bubble in the world's adjacency
    // area (when the area holds n strides); when every row fits, the rows stay there and the fill
    // pass only orders their pair and body runs instead of enumerating the candidates again.
    const bool stageOk = (uint64_t)n * kStageStride <= d.adjCapacity;
    uint32_t droppedLocal = 0, maxLocal = 0, pairsLocal = 0, longLocal = 0;
    for (uint32_t i = R.first; i < n; i += R.stride) {
        uint32_t count = 0, forward = 0;
        uint32_t *stage = W.adj + (size_t)i * kStageStride;
        auto tally = [&](uint32_t e) {
            if (stageOk && count < kStageStride) stage[count] = e;
            ++count;
            if ((e >> kKindShift) == 2u) atomicAdd(&bodyHits[e & kIndexMask], 1u);
        };
        enumerate_row(W, cells, B, i, tally, forward, bodies);
        W.rowLen[i] = count;
        longLocal |= count > kStageStride ? 1u : 0u;
        if (forward > limit) droppedLocal += forward - limit;
        if (count > limit) droppedLocal += count - limit;
        maxLocal = max(maxLocal, max(forward, count));
        pairsLocal += min(forward, limit);
    }
    grid.sync();
    for (uint32_t k = gtid; k < W.nb; k += R.stride) {
        const uint32_t hits = bodyHits[k];
        maxLocal = max(maxLocal, hits);
        if (hits > limit) droppedLocal += hits - limit;
    }
    droppedLocal = block_sum_u32<BLOCK>(red, droppedLocal);
    pairsLocal = block_sum_u32<BLOCK>(red, pairsLocal);
    maxLocal = block_max_u32<BLOCK>(red, maxLocal);
    if (threadIdx.x == 0) {
        atomicAdd(&s.counters[kCoDropped], droppedLocal);
        atomicAdd(&s.counters[kCoPairs], pairsLocal);
        atomicMax(&s.counters[kCoMaxDemand], maxLocal);
    }
    if (longLocal) atomicOr(&s.counters[kCoLongRows], 1u);
    entries = grid_scan<BLOCK>(grid, red, s, W.rowLen, W.rowStart, n);
    grid.sync();
    dropped = s.counters[kCoDropped];
    maxDemand = s.counters[kCoMaxDemand];
    pairs = s.counters[kCoPairs];
    if (stageOk && s.counters[kCoLongRows] == 0u) {
        // The staged rows are the rows (a start and a length per bubble).
        for (uint32_t i = R.first; i < n; i += R.stride) {
            W.rowStart[i] = i * kStageStride;
            uint32_t *row = W.adj + (size_t)i * kStageStride;
            const uint32_t count = W.rowLen[i];
            uint32_t pairEnd = 0, segs = 0;
            for (uint32_t x = 0; x < count; ++x) {
                const uint32_t kind = row[x] >> kKindShift;
                pairEnd += kind == 0u ? 1u : 0u;
                segs += kind == 1u ? 1u : 0u;
            }
            sort_pair_block(row, pairEnd);
            sort_run(row, pairEnd + segs, count);
        }
        grid.sync();
        if (dropped != 0u) coop_cut_rows<BLOCK>(grid, W, red, s, R, limit, dropped, maxDemand, entries);
        return true;
    }
    if (entries > d.adjCapacity) return false;
    // Fill the rows, then order the pair run and the body run (Contact:222-238).
    for (uint32_t i = R.first; i < n; i += R.stride) {
        uint32_t *row = W.adj + W.rowStart[i];
        uint32_t count = 0, forward = 0;
        auto put = [&](uint32_t e) { row[count++] = e; };
        enumerate_row(W, cells, B, i, put, forward, bodies);
        uint32_t pairEnd = 0;
        while (pairEnd < count && (row[pairEnd] >> kKindShift) == 0u) ++pairEnd;
        uint32_t segEnd = pairEnd;
        while (segEnd < count && (row[segEnd] >> kKindShift) == 1u) ++segEnd;
        sort_run(row, 0, pairEnd);
        sort_run(row, segEnd, count);
    }
    grid.sync();
    if (dropped != 0u) coop_cut_rows<BLOCK>(grid, W, red, s, R, limit, dropped, maxDemand, entries);
    return true;
}

// Contact assertion over the grid (Force:841-927): the lowest violating bubble wins.
template <int BLOCK>
BPD bool coop_assert(cg::grid_group &grid, World &W, Red<BLOCK> &red, const CoopScratch &s, Range R, uint32_t *cells,
                     uint32_t B, const TreeBodies &tree, const Ctl &c, WorldStatus *st) {
    const uint32_t n = W.n;
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    if (!(W.cellWidth > 0.0f)) W.cellWidth = 1.0f;
    for (uint32_t b = gtid; b < B; b += R.stride) cells[b] = 0;
    if (gtid == 0) s.counters[kCoViolator] = 0xffffffffu;
    grid.sync();
    coop_hash<BLOCK>(grid, W, red, s, R, cells, B);
    const float fraction = step_fraction(c);
    uint32_t firstIndex = 0xffffffffu;
    Violation mine = {0u, 0u, 0, 0.0f};
    for (uint32_t i = R.first; i < n && firstIndex == 0xffffffffu; i += R.stride) {
        const float2 center = W.pos[i], start = W.start[i];
        if (!finite2(center) || W.region[i] == BP_REGION_REMOVED) continue;
        const bool movable = W.inv_mass(i) > 0.0f;
        const uint32_t ci = W.cat[i], ki = W.mask[i];
        uint32_t bestJ = 0xffffffffu;
        float bestD = 0.0f;
        const int64_t cx = W.cell_x(center), cy = W.cell_y(center);
        int x0, x1, y0, y1;
        zero_distance_span(center.x, W.cellWidth, cx, x0, x1);
        zero_distance_span(center.y, W.cellWidth, cy, y0, y1);
        for (int dx = x0; dx <= x1; ++dx)
            for (int dy = y0; dy <= y1; ++dy) {
                const int64_t tx = cx + dx, ty = cy + dy;
                const uint32_t b = bucket_of(tx, ty, B);
                const uint32_t lo = b == 0 ? 0u : cells[b - 1], hi = cells[b];
                for (uint32_t k = lo; k < hi; ++k) {
                    const uint4 cb = W.cellB[k];
                    const uint32_t j = cb.x;
                    if (j <= i || j >= bestJ) continue;
                    const float4 ca = W.cellA[k];
                    const float2 q = f2(ca.x, ca.y);
                    if (W.cell_x(q) != tx || W.cell_y(q) != ty) continue;
                    if ((cb.w & 2u) || !filter_accepts(ci, ki, cb.y, cb.z)) continue;
                    if (!movable && !(cb.w & 4u)) continue;
                    const float dd = len2(sub(q, center));
                    if (dd <= 0.0f) { bestJ = j; bestD = dd; }
                }
            }
        if (bestJ != 0xffffffffu) {
            firstIndex = i; mine.kind = BP_VIOLATION_PAIR; mine.other = (uint64_t)W.id[bestJ]; mine.detail = bestD;
            break;
        }
        if (!movable) continue;
        // Inside an immovable body: the lowest such body, found through the tree when built.
        uint32_t body = 0xffffffffu;
        auto inside = [&](uint32_t k) {
            if (k >= body) return;
            const BodyView bv = W.body(k);
            if (!body_immovable(bv) || !filter_accepts(W.cat[i], W.mask[i], bv.kind.z, bv.kind.w)) return;
            if (frame_inside(center, body_frame(bv, fraction), W.verts, W.tris)) body = k;
        };
        if (tree.built) bvh_query(W.A, tree.base, tree.nb, center, center, inside);
        else for (uint32_t k = 0; k < W.nb && body == 0xffffffffu; ++k) inside(k);
        if (body != 0xffffffffu) {
            firstIndex = i; mine.kind = BP_VIOLATION_INSIDE_BODY; mine.other = W.A->bodyId[W.d->bodyOffset + body];
            mine.detail = 0.0f;
            break;
        }
        Violation edge = {0u, 0u, 0, 0.0f};
        Geo geo = W.g;
        geo.nb = 0;                                           // bodies were checked above
        if (assert_body_and_edges(geo, center, start, W.cat[i], W.mask[i], fraction, edge)) {
            firstIndex = i; mine = edge;
            break;
        }
    }
    if (firstIndex != 0xffffffffu) atomicMin(&s.counters[kCoViolator], firstIndex);
    grid.sync();
    const uint32_t winner = s.counters[kCoViolator];
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
// The cooperative step kernel: one world, gridDim.x workgroups of BLOCK lanes.
// ---------------------------------------------------------------------------------------
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void world_coop_kernel(const Arena *__restrict__ A, WorldDesc *descs, WorldStatus *stats,
                                                           uint32_t wi) {
    cg::grid_group grid = cg::this_grid();
    __shared__ Red<BLOCK> red;
    WorldDesc *d = descs + wi;
    WorldStatus *st = stats + wi;
    const uint64_t ticks0 = wall_clock64();
    const uint32_t gtid = blockIdx.x * BLOCK + threadIdx.x;
    const Range R{gtid, gridDim.x * BLOCK};
    if (d->failed) {                                          // uniform: every workgroup returns
        if (gtid == 0) st->executed = 0;
        return;
    }
    const uint32_t pub = d->pub, wside = 1u - pub;
    World W = make_world(A, d, wside);
    const uint32_t n = W.n, o = d->bubbleOffset;
    const CoopScratch s = coop_scratch(A, d);
    uint32_t *cells = A->coopCells + d->coopCellOffset;
    const uint32_t B = d->coopBuckets;
    float *history = A->history + d->historyOffset;

    // 1. Begin step (Force:285-293); the bodies were predicted by body_bvh_kernel.
    for (uint32_t i = R.first; i < n; i += R.stride) {
        float2 p = A->pos[pub][o + i];
        float2 v = A->vel[pub][o + i];
        float om = A->omega[pub][o + i];
        const uint32_t f = A->flags[o + i];
        if ((f & kEnabled) && (f & kStatic)) { v = f2(0.0f, 0.0f); om = 0.0f; }
        W.pos[i] = p; W.start[i] = p; W.vel[i] = v; W.rot[i] = A->rot[pub][o + i]; W.omega[i] = om;
        W.region[i] = BP_REGION_UNCHANGED;
        W.publish_step(i);
    }
    if (gtid == 0) {
        s.counters[kCoSafety] = 0; s.counters[kCoRemovals] = 0; s.counters[kCoWraps] = 0;
    }
    grid.sync();
    uint32_t safety = gtid == 0 ? s.counters[kCoBodySafety] : 0u;
    TreeBodies tree;
    tree.A = A; tree.base = 2 * d->bodyOffset; tree.nb = W.nb; tree.built = s.counters[kCoBvhBuilt] != 0;

    const uint32_t limit = d->limit;
    Ctl c;
    c.t0 = d->dt; c.t1 = n == 0 ? d->dt : 0.0f; c.h = 0.0f; c.split = (float)(d->split < 1 ? 1u : d->split);
    c.c0 = c.c1 = c.rebuild = c.limit = 0; c.limOwner = kNoOwner; c.limTerm = BP_TERM_REMAINING;
    uint32_t ov1 = 0, ov2 = 0, ov3 = 0, pairs = 0, entriesTotal = 0, dropped = 0, ov0 = 0;
    int32_t status = BP_OK;
    bool claimed = false;

    if (n > 0) {
        float w2Max, mbMin, dbMin;
        uint32_t w2Own, mbOwn, dbOwn, maxDemand;
        bool ok = coop_build<BLOCK>(grid, W, red, s, R, cells, B, tree, c.t0 - c.t1, limit, dropped, maxDemand, pairs, entriesTotal);
        ov0 += dropped; ov1 = max(ov1, maxDemand);
        if (!ok) status = BP_ERR_CAPACITY_OVERFLOW;
        if (status == BP_OK) {
            pass_forces<BLOCK>(W, red, R, c, 2u, w2Max, w2Own, mbMin, mbOwn);
            grid_maxmin<BLOCK>(grid, red, s, w2Max, w2Own, mbMin, mbOwn);
            for (uint32_t iteration = 0; iteration < 2 * kFrameLimit + 2; ++iteration) {
                const float remaining = c.t0 - c.t1;
                const bool over = !(remaining > 0.0f), stopped = c.rebuild != 0 || c.limit != 0;
                const bool limited = c.c1 >= kFrameLimit;
                if (over || stopped || limited) {
                    if (limited && !over && !stopped) c.limit = 1;
                    c.h = 0.0f;
                } else {
                    const float stiffnessBound = w2Max > 0.0f ? 1.0f / __builtin_sqrtf(w2Max) : inff();
                    pass_distance<BLOCK>(W, red, R, c, fminf(remaining, stiffnessBound), dbMin, dbOwn);
                    grid_min<BLOCK>(grid, red, s, dbMin, dbOwn);
                    reduce_h(c, stiffnessBound, w2Own, dbMin, dbOwn, mbMin, mbOwn, history, gtid);
                }
                const bool finished = !(c.t0 - c.t1 > 0.0f);
                if (!(c.h > 0.0f)) {
                    if (c.limit || finished || !c.rebuild) break;
                    c.c0 = 0; c.rebuild = 0;
                    ov2 = max(ov2, ov0); ov0 = 0; ov3 += 1;
                    ok = coop_build<BLOCK>(grid, W, red, s, R, cells, B, tree, c.t0 - c.t1, limit, dropped, maxDemand, pairs,
                                           entriesTotal);
                    ov0 += dropped; ov1 = max(ov1, maxDemand);
                    if (!ok) { status = BP_ERR_CAPACITY_OVERFLOW; break; }
                    pass_forces<BLOCK>(W, red, R, c, 2u, w2Max, w2Own, mbMin, mbOwn);
                    grid_maxmin<BLOCK>(grid, red, s, w2Max, w2Own, mbMin, mbOwn);
                    continue;
                }
                pass_drift<BLOCK>(W, R, c);
                grid.sync();
                pass_forces<BLOCK>(W, red, R, c, 1u, w2Max, w2Own, mbMin, mbOwn);
                grid_maxmin<BLOCK>(grid, red, s, w2Max, w2Own, mbMin, mbOwn);
                if (d->injectionEnabled && c.c1 == d->injectionAfter + 1) {
                    for (uint32_t i = R.first; i < n; i += R.stride)
                        if (W.id[i] == d->injectionObject) {
                            W.pos[i] = f2(d->injectionPosition[0], d->injectionPosition[1]);
                            W.publish_a(i);
                        }
                    grid.sync();
                    pass_bounds<BLOCK>(W, red, R, c, w2Max, w2Own, mbMin, mbOwn);
                    grid_maxmin<BLOCK>(grid, red, s, w2Max, w2Own, mbMin, mbOwn);
                }
                if (d->assertEverySubstep && !claimed) {
                    if (coop_assert<BLOCK>(grid, W, red, s, R, cells, B, tree, c, st)) claimed = true;
                    grid.sync();
                }
            }
            ov2 = max(ov2, ov0);
        }
    }
    // 4. Gated end of step (NativeMetalSolver.swift:837-945).
    const bool finished = c.limit == 0 && !(c.t0 - c.t1 > 0.0f);
    uint32_t rmLocal = 0, wrLocal = 0;
    if (status == BP_OK && finished) {
        const float2 lo = f2(d->region[0], d->region[1]), hi = f2(d->region[2], d->region[3]);
        for (uint32_t i = R.first; i < n; i += R.stride) {
            uint32_t region = BP_REGION_UNCHANGED;
            if (d->applyRegion) {
                float2 p = W.pos[i], start = W.start[i];
                region = apply_region(p, start, W.radius[i], W.flags[i], lo, hi, safety);
                if (region == BP_REGION_WRAPPED) { W.pos[i] = p; W.start[i] = start; }
            }
            W.region[i] = region;
            rmLocal += region == BP_REGION_REMOVED;
            wrLocal += region == BP_REGION_WRAPPED;
            if (!finite2(W.pos[i]) || !finite2(W.start[i]) || !finite2(W.vel[i]) || !isfinite(W.rot[i]) ||
                !isfinite(W.omega[i]) || !isfinite(W.mass[i]) || !isfinite(W.invmRaw[i]) || !isfinite(W.radius[i]))
                safety |= 1u;
        }
        for (uint32_t j = gtid; j < W.ns; j += R.stride) {
            const Seg sg = W.segment(j);
            const float2 e = sub(sg.b, sg.a);
            if (!finite2(sg.a) || !finite2(sg.b) || !finite2(e) || !isfinite(dot2(e, e))) safety |= 1u;
        }
        grid.sync();
        if (!claimed && n > 0 && coop_assert<BLOCK>(grid, W, red, s, R, cells, B, tree, c, st)) claimed = true;
    }
    safety = block_max_u32<BLOCK>(red, safety & 1u);          // bit 0 is the only safety bit
    rmLocal = block_sum_u32<BLOCK>(red, rmLocal);
    wrLocal = block_sum_u32<BLOCK>(red, wrLocal);
    if (threadIdx.x == 0) {
        if (safety) atomicOr(&s.counters[kCoSafety], safety);
        atomicAdd(&s.counters[kCoRemovals], rmLocal);
        atomicAdd(&s.counters[kCoWraps], wrLocal);
    }
    grid.sync();
    status = final_status(status, c, s.counters[kCoSafety], claimed, finished, n);
    if (gtid == 0) {
        st->status = status;
        st->substeps = c.c1;
        st->rebuilds = ov3;
        st->candidatePairs = pairs;
        st->candidateOverflow = ov2;
        st->maxCandidates = ov1;
        st->removals = s.counters[kCoRemovals];
        st->wraps = s.counters[kCoWraps];
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
            if (s.counters[kCoRemovals] != 0) d->failed = kWorldHeld;   // removed bubbles leave before the next step
        } else if (status != BP_ERR_CAPACITY_OVERFLOW) {
            d->failed = kWorldFailed;
        }
    }
}

}  // namespace bpd

constexpr int kCoopBlock = 256;

// Largest cooperative grid the device holds at once for the step kernel (0 when unavailable).
extern "C" __attribute__((visibility("hidden"))) uint32_t bp_coop_max_groups() {
    int device = 0, cus = 0, perCu = 0, coop = 0;
    if (hipGetDevice(&device) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess || !coop) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, bpd::world_coop_kernel<kCoopBlock>, kCoopBlock, 0) != hipSuccess)
        return 0;
    uint32_t groups = (uint32_t)cus * (uint32_t)perCu;
    return groups < bpd::kCoopMaxGroups ? groups : bpd::kCoopMaxGroups;
}

extern "C" __attribute__((visibility("hidden"))) uint32_t bp_coop_block() { return kCoopBlock; }

// One cooperative world: the body tree of the step, then the cooperative step kernel.
// `orderEntry` points at the world's index in device memory; `world` is the same index.
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_world_coop(const bpd::Arena *arena, bpd::WorldDesc *descs,
                                                                             bpd::WorldStatus *stats, const uint32_t *orderEntry,
                                                                             uint32_t world, uint32_t groups, hipStream_t stream) {
    hipLaunchKernelGGL(bpd::body_bvh_kernel, dim3(1), dim3(bpd::kBvhThreads), 0, stream, arena, descs, orderEntry);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    void *args[] = {(void *)&arena, (void *)&descs, (void *)&stats, (void *)&world};
    return hipLaunchCooperativeKernel((const void *)bpd::world_coop_kernel<kCoopBlock>, dim3(groups), dim3(kCoopBlock), args, 0,
                                      stream);
}
