// bp_common.h -- workgroup primitives and per-world geometry shared by the world-step kernels.
//
// Reductions keep the tie rule of nativeReduceSubstepInGroup (NativeForceKernels.metal-source:
// 646-698): equal values go to the lower bubble index, so a reduction's result does not depend on
// the workgroup width or on which lane held which bubble.
#ifndef BP_COMMON_H
#define BP_COMMON_H

#include "bp_device.h"
#include "../../include/bubblephysics.h"

namespace bpd {

constexpr uint32_t kNoOwner = 0xffffffffu;

// NativeSubstepControl, held in registers and uniform across the workgroup. The limiter is kept
// as the limiting bubble's index and turned into its id once, at the end of the step.
struct Ctl {
    float t0, t1, h, split;
    uint32_t c0, c1, rebuild, limit;
    uint32_t limOwner;
    uint32_t limTerm;
};

template <int BLOCK> struct Red {
    float v[BLOCK / 64];
    uint32_t o[BLOCK / 64];
    float v2[BLOCK / 64];
    uint32_t o2[BLOCK / 64];
    uint32_t scan[BLOCK / 64];
    uint32_t counters[8];
};

BPD void wave_max_owner(float &v, uint32_t &o) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float ov = __shfl_xor(v, off, 64);
        uint32_t oo = __shfl_xor(o, off, 64);
        if (ov > v || (ov == v && oo < o)) { v = ov; o = oo; }
    }
}
BPD void wave_min_owner(float &v, uint32_t &o) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float ov = __shfl_xor(v, off, 64);
        uint32_t oo = __shfl_xor(o, off, 64);
        if (ov < v || (ov == v && oo < o)) { v = ov; o = oo; }
    }
}

// Max of (a, owner) and min of (b, owner) over the workgroup, broadcast to every lane.
template <int BLOCK>
BPD void block_maxmin(Red<BLOCK> &r, float &a, uint32_t &ao, float &b, uint32_t &bo) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_max_owner(a, ao);
    wave_min_owner(b, bo);
    if (BLOCK == 64) return;
    __syncthreads();
    if (lane == 0) { r.v[wave] = a; r.o[wave] = ao; r.v2[wave] = b; r.o2[wave] = bo; }
    __syncthreads();
    a = r.v[0]; ao = r.o[0]; b = r.v2[0]; bo = r.o2[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) {
        float v = r.v[w]; uint32_t o = r.o[w];
        if (v > a || (v == a && o < ao)) { a = v; ao = o; }
        v = r.v2[w]; o = r.o2[w];
        if (v < b || (v == b && o < bo)) { b = v; bo = o; }
    }
}
template <int BLOCK>
BPD void block_min(Red<BLOCK> &r, float &b, uint32_t &bo) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_min_owner(b, bo);
    if (BLOCK == 64) return;
    __syncthreads();
    if (lane == 0) { r.v2[wave] = b; r.o2[wave] = bo; }
    __syncthreads();
    b = r.v2[0]; bo = r.o2[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) {
        float v = r.v2[w]; uint32_t o = r.o2[w];
        if (v < b || (v == b && o < bo)) { b = v; bo = o; }
    }
}
template <int BLOCK>
BPD float block_maxf(Red<BLOCK> &r, float v) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    if (BLOCK == 64) return v;
    __syncthreads();
    if (lane == 0) r.v[wave] = v;
    __syncthreads();
    float m = r.v[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) m = fmaxf(m, r.v[w]);
    return m;
}
template <int BLOCK>
BPD uint32_t block_min_u32(Red<BLOCK> &r, uint32_t v) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off, 64));
    if (BLOCK == 64) return v;
    __syncthreads();
    if (lane == 0) r.scan[wave] = v;
    __syncthreads();
    uint32_t m = r.scan[0];
#pragma unroll
    for (int w = 1; w < BLOCK / 64; ++w) m = min(m, r.scan[w]);
    return m;
}
// Exclusive prefix sum of one value per lane; returns the lane's offset and the total.
template <int BLOCK>
BPD uint32_t block_exclusive_scan(Red<BLOCK> &r, uint32_t value, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inclusive = value;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t other = __shfl_up(inclusive, off, 64);
        if (lane >= off) inclusive += other;
    }
    if (BLOCK == 64) {
        total = __shfl(inclusive, 63, 64);
        return inclusive - value;
    }
    __syncthreads();
    if (lane == 63) r.scan[wave] = inclusive;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) {
        uint32_t s = r.scan[w];
        if (w < wave) before += s;
        total += s;
    }
    return before + inclusive - value;
}

// Exclusive prefix over a bucket array in LDS, each lane owning a contiguous chunk (Force:1080-1094).
template <int BLOCK>
BPD void bucket_prefix(Red<BLOCK> &r, uint32_t *buckets, uint32_t count) {
    uint32_t perLane = (count + BLOCK - 1) / BLOCK;
    uint32_t first = min((uint32_t)threadIdx.x * perLane, count), last = min(first + perLane, count);
    uint32_t sum = 0;
    for (uint32_t b = first; b < last; ++b) sum += buckets[b];
    uint32_t total;
    uint32_t running = block_exclusive_scan<BLOCK>(r, sum, total);
    __syncthreads();
    for (uint32_t b = first; b < last; ++b) { uint32_t own = buckets[b]; buckets[b] = running; running += own; }
}

// The modulus is a mask when the bucket count is a power of two (the automatic counts are).
BPD uint32_t bucket_of(int64_t cx, int64_t cy, uint32_t buckets) {      // Force:1000-1002
    const uint32_t h = (uint32_t)cx * 73856093u ^ (uint32_t)cy * 19349663u;
    return (buckets & (buckets - 1u)) == 0u ? (h & (buckets - 1u)) : h % buckets;
}
// floor(x / w) as int64; cells within the int32 range take the one-instruction conversion,
// which gives the same integer.
BPD int64_t cell_coord(float x, float width) {
    const float f = floorf(x / width);
    return fabsf(f) < 2147483648.0f ? (int64_t)(int32_t)f : (int64_t)f;
}

constexpr uint32_t kStageStride = 64;      // staged row entries per bubble in a world's global adjacency area

// Sorts the first `count` of N register values ascending with a bitonic network (every index is a
// compile-time constant once unrolled, so the values stay in registers). Slots past `count`
// hold the largest value and stay at the end.
template <int N>
BPD void bitonic_sort(uint32_t (&v)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t lo = min(v[i], v[l]), hi = max(v[i], v[l]);
                    if ((i & k) == 0) { v[i] = lo; v[l] = hi; } else { v[i] = hi; v[l] = lo; }
                }
            }
}
template <int N, typename T>
BPD void sort_small_block(T *row, uint32_t count) {
    uint32_t v[N];
#pragma unroll
    for (int u = 0; u < N; ++u) v[u] = (uint32_t)u < count ? (uint32_t)row[u] : 0xffffffffu;
    bitonic_sort<N>(v);
#pragma unroll
    for (int u = 0; u < N; ++u)
        if ((uint32_t)u < count) row[u] = (T)v[u];
}
// Orders the pair block of an adjacency row (entries of kind 0, so the value is the neighbour
// index) ascending: blocks of up to 16 entries are loaded together, sorted in registers and
// stored together; longer blocks are insertion-sorted in place.
template <typename T>
BPD void sort_pair_block(T *row, uint32_t count) {
    if (count < 2) return;
    if (count <= 8) { sort_small_block<8>(row, count); return; }
    if (count <= 16) { sort_small_block<16>(row, count); return; }
    for (uint32_t x = 1; x < count; ++x) {
        const T val = row[x];
        uint32_t loc = x;
        while (loc > 0 && val < row[loc - 1]) { row[loc] = row[loc - 1]; --loc; }
        row[loc] = val;
    }
}

// ---------------------------------------------------------------------------------------
// Rows cut to the reference's k candidates (Geometry:230-251 and 276-298, 350-415;
// Contact:190-238).
//
// A build here lists every candidate of a bubble in one row, both directions of each pair,
// sorted by (kind, neighbour). The reference keeps less when a row overflows:
//   - the candidate row of bubble i holds the k smallest keys of its pairs j > i, then its
//     segments: its k smallest forward neighbours, and k - min(forward pairs, k) segments;
//   - the candidate row of a body holds its k bubbles of smallest index;
//   - the adjacency of bubble i takes the surviving candidates in arrival order (pairs (a, i)
//     by ascending a, then the pairs (i, j) by ascending j, segments, bodies: the row's own
//     order) and keeps the first k.
// Bubble ids ascend with the index, so key order is index order. When a build drops anything,
// the rows are cut to exactly that in two passes over the bubbles with a barrier between:
//   1. forward_cut gives each bubble the largest forward neighbour its candidate row keeps
//      (kNoCut when it keeps every one), stored where every bubble can read it, and
//      mark_body_cuts marks, in the bubble's own row, the bodies that do not keep it;
//   2. cut_row compacts each bubble's own row in place: pair (l, h) stays iff h <= cut(l).
// Nothing of this runs for a build that drops nothing.
// ---------------------------------------------------------------------------------------
constexpr uint32_t kNoCut = 0xffffffffu;

template <typename T> struct RowCodec;
template <> struct RowCodec<uint16_t> {        // LDS rows: kind << 14 | index
    static constexpr uint32_t kMask = (1u << 14) - 1;
    static constexpr uint16_t kDrop = 0xffff;  // kind 3: a marked entry
};
template <> struct RowCodec<uint32_t> {        // global rows: kind << 30 | index
    static constexpr uint32_t kMask = (1u << 30) - 1;
    static constexpr uint32_t kDrop = 0xffffffffu;
};

// Forward pair count of bubble i (its pairs j > i) and the largest neighbour its candidate row
// keeps, from the sorted pair run row[0, pairEnd).
template <typename T>
BPD uint32_t forward_cut(const T *row, uint32_t pairEnd, uint32_t i, uint32_t limit, uint32_t &forwardPairs) {
    uint32_t lo = 0, hi = pairEnd;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (((uint32_t)row[mid] & RowCodec<T>::kMask) < i) lo = mid + 1; else hi = mid;
    }
    forwardPairs = pairEnd - lo;
    return forwardPairs > limit ? ((uint32_t)row[lo + limit - 1] & RowCodec<T>::kMask) : kNoCut;
}

// Index of the limit-th bubble (in index order) of body k's full row, or kNoCut when the row
// holds at most `limit` bubbles: a body's k-entry row keeps exactly the bubbles of lower index.
// `hits` is the row's length, `candidate(a)` the candidate predicate of bubble a.
template <class Candidate>
BPD uint32_t body_cut_index(uint32_t n, uint32_t limit, uint32_t hits, Candidate candidate) {
    if (hits <= limit) return kNoCut;
    uint32_t count = 0;
    for (uint32_t a = 0; a < n; ++a)
        if (candidate(a) && ++count == limit) return a;
    return kNoCut;
}

// Marks the entries of the body run row[segEnd, end) of bubble i whose body's row does not keep
// it: bodyCut[k] (body_cut_index) is below i, so k bubbles of lower index fill the body's row.
template <typename T>
BPD void mark_body_cuts(T *row, uint32_t segEnd, uint32_t end, uint32_t i, const uint32_t *bodyCut) {
    for (uint32_t e = segEnd; e < end; ++e) {
        const uint32_t k = (uint32_t)row[e] & RowCodec<T>::kMask;
        if (bodyCut[k] != kNoCut && bodyCut[k] < i) row[e] = RowCodec<T>::kDrop;
    }
}

// Block ends of a global row ordered (pairs, segments, bodies).
BPD void row_blocks(const uint32_t *row, uint32_t len, uint32_t &pairEnd, uint32_t &segEnd) {
    pairEnd = 0;
    while (pairEnd < len && (row[pairEnd] >> kKindShift) == 0u) ++pairEnd;
    segEnd = pairEnd;
    while (segEnd < len && (row[segEnd] >> kKindShift) == 1u) ++segEnd;
}

struct RowCut {
    uint32_t len, pairEnd, segEnd;   // the cut row and its block ends
    uint32_t demand;                 // adjacency entries the reference's candidates ask for
};

// Compacts the row of bubble i in place. `cutOf(l)` is bubble l's forward cut.
template <typename T, class CutOf>
BPD RowCut cut_row(T *row, uint32_t pairEnd, uint32_t segEnd, uint32_t end, uint32_t i, uint32_t limit,
                   uint32_t forwardPairs, CutOf cutOf) {
    RowCut r{0u, 0u, 0u, 0u};
    const uint32_t own = cutOf(i);
    for (uint32_t e = 0; e < pairEnd; ++e) {
        const T v = row[e];
        const uint32_t j = (uint32_t)v & RowCodec<T>::kMask;
        if (j < i ? i <= cutOf(j) : j <= own) {
            if (r.len < limit) row[r.len++] = v;
            ++r.demand;
        }
    }
    r.pairEnd = r.len;
    const uint32_t segKeep = min(segEnd - pairEnd, limit - min(forwardPairs, limit));
    for (uint32_t e = pairEnd; e < pairEnd + segKeep; ++e) {
        const T v = row[e];
        if (r.len < limit) row[r.len++] = v;
        ++r.demand;
    }
    r.segEnd = r.len;
    for (uint32_t e = segEnd; e < end; ++e) {
        const T v = row[e];
        if (v == RowCodec<T>::kDrop) continue;
        if (r.len < limit) row[r.len++] = v;
        ++r.demand;
    }
    return r;
}

// The cell offsets along one axis (within -1..1 of the centre's cell `c`) that can hold another
// centre at FP32 distance 0 from a centre with coordinate `v`: the contact assertion's pair test
// (Force:880-881). sqrt(dx^2 + dy^2) is 0 only when both squares flush to zero, so
// |dx|, |dy| < 2.7e-23, and floor(x / w) is monotonic in x, so the cells of v - 1e-20 and
// v + 1e-20 bound the cell of every such centre. Away from a cell boundary this is the own cell.
BPD void zero_distance_span(float v, float width, int64_t c, int &lo, int &hi) {
    lo = cell_coord(v - 1e-20f, width) < c ? -1 : 0;
    hi = cell_coord(v + 1e-20f, width) > c ? 1 : 0;
}

// ---------------------------------------------------------------------------------------
// Geometry, force records and configuration of one world (working side of the body poses).
// ---------------------------------------------------------------------------------------
struct Geo {
    const Arena *A;
    uint32_t ns, nb;
    uint32_t so, bo, vo, to, mo, bao;   // segment, body, vertex, triangle, motor, band offsets
    uint32_t fo, nf, fvo, fio;          // field offset and count, field-vertex and field-id offsets
    uint32_t side;                      // working side of the body poses
    float2 globalForce;
    float sepTol;

    BPD const float2 *verts() const { return A->verts + vo; }
    BPD const uint4 *tris() const { return A->tris + to; }
    BPD uint32_t segCat(uint32_t j) const { return A->segCat[so + j]; }
    BPD uint32_t segMask(uint32_t j) const { return A->segMask[so + j]; }
    BPD int64_t segId(uint32_t j) const { return A->segId[so + j]; }
    BPD int64_t segOwner(uint32_t j) const { return A->segOwner[so + j]; }
    BPD uint64_t bodyId(uint32_t k) const { return A->bodyId[bo + k]; }
    BPD uint32_t *bodyHits() const { return A->bodyHits + bo; }
    BPD float4 band(int32_t k) const { return A->bands[bao + k]; }
    BPD float4 motor(int32_t k) const { return A->motors[mo + k]; }

    BPD Seg segment(uint32_t j) const {
        Seg s;
        float4 ab = A->segAB[so + j], rg = A->segRegion[so + j];
        s.a = f2(ab.x, ab.y); s.b = f2(ab.z, ab.w);
        s.prevStart = f2(rg.x, rg.y);
        s.outward = rg.z;
        s.flags = A->segFlags[so + j];
        s.base = segment_base(s.a, s.b);
        return s;
    }
    BPD BodyView body(uint32_t k) const {
        BodyView b;
        uint32_t g = bo + k;
        b.pose = A->bodyPose[side][g]; b.prev = A->bodyPrev[side][g]; b.motion = A->bodyMotion[g];
        b.shape = A->bodyShape[g]; b.ranges = A->bodyRanges[g]; b.kind = A->bodyKind[g];
        return b;
    }
    // Segment candidate of a bubble, Geometry:279-290.
    BPD bool segment_candidate(float2 p, float radius, float margin, uint32_t cat, uint32_t mask, uint32_t j) const {
        if (!filter_accepts(cat, mask, segCat(j), segMask(j))) return false;
        float4 ab = A->segAB[so + j];
        float reach = radius + margin + sepTol;
        float2 lower = subs(min2(f2(ab.x, ab.y), f2(ab.z, ab.w)), reach);
        float2 upper = adds(max2(f2(ab.x, ab.y), f2(ab.z, ab.w)), reach);
        return p.x >= lower.x && p.y >= lower.y && p.x <= upper.x && p.y <= upper.y;
    }
    // Bubble-body candidate, Geometry:361-379 (NaN centres compare like the reference).
    BPD bool body_candidate(float2 p, float radius, float margin, uint32_t cat, uint32_t mask, uint32_t k) const {
        uint32_t g = bo + k;
        uint4 kind = A->bodyKind[g];
        bool isStatic = (kind.y & kBodyStatic) != 0;
        bool obstacle = isStatic ? (kind.x != 0u) : ((kind.y & kBodyKinematic) != 0u);
        if (!obstacle) return false;
        if (!filter_accepts(cat, mask, kind.z, kind.w)) return false;
        float4 pose = A->bodyPose[side][g];
        float2 prev = A->bodyPrev[side][g];
        float br = A->bodyShape[g].x;
        float2 bodyLow = subs(min2(f2(pose.x, pose.y), prev), br);
        float2 bodyHigh = adds(max2(f2(pose.x, pose.y), prev), br);
        float reach = radius + fmaxf(margin, 0.0f);
        float2 lo = subs(p, reach), hi = adds(p, reach);
        return !(hi.x < bodyLow.x || hi.y < bodyLow.y || bodyHigh.x < lo.x || bodyHigh.y < lo.y);
    }
    // Non-contact acceleration a0 (Force:450-470): global force, band, motor.
    // idp points at the bubble's id; it is read only when the world has force fields.
    BPD float2 non_contact_with(uint32_t flags, float invm, const int64_t *idp, int32_t bandIndex, bool hasMotor, float2 motorForce,
                                float2 p) const {
        if (!(flags & kForceDriven) || !(invm > 0.0f)) return f2(0.0f, 0.0f);
        float2 force = globalForce;
        if (nf > 0) {
            const uint64_t id = (uint64_t)*idp;
            for (uint32_t k = 0; k < nf; ++k) {
                const DevField f = A->fields[fo + k];
                if (field_affects(f, A->fieldIds + fio, id)) force = add(force, field_force(f, A->fieldVerts + fvo, p));
            }
        }
        if (bandIndex >= 0) {
            float4 bd = band(bandIndex);
            float2 bf = mul(sub(f2(bd.x, bd.y), p), bd.z);
            float lim = bd.w, magnitude = len2(bf);
            if (lim > 0.0f && magnitude > lim) bf = mul(bf, lim / magnitude);
            force = add(force, bf);
        }
        if (hasMotor) force = add(force, motorForce);
        return mul(force, invm);
    }
    BPD float2 non_contact(uint32_t flags, float invm, const int64_t *idp, int32_t bandIndex, int32_t motorIndex, float2 p) const {
        float2 mf = f2(0.0f, 0.0f);
        if (motorIndex >= 0) { float4 md = motor(motorIndex); mf = f2(md.x, md.y); }
        return non_contact_with(flags, invm, idp, bandIndex, motorIndex >= 0, mf, p);
    }
};

BPD Geo make_geo(const Arena *A, const WorldDesc *d, uint32_t side) {
    Geo g;
    g.A = A;
    g.ns = d->segmentCount; g.nb = d->bodyCount;
    g.so = d->segmentOffset; g.bo = d->bodyOffset; g.vo = d->vertexOffset; g.to = d->triangleOffset;
    g.mo = d->motorOffset; g.bao = d->bandOffset;
    g.fo = d->fieldOffset; g.nf = d->fieldCount; g.fvo = d->fieldVertexOffset; g.fio = d->fieldIdOffset;
    g.side = side;
    g.globalForce = f2(d->globalForce[0], d->globalForce[1]);
    g.sepTol = d->sepTol;
    return g;
}

// Kinematic and static bodies at the start of a step (PostSolve:66-82 for the bodies a bubble
// world may hold): previous <- pose, kinematic pose advanced by its clamped velocity.
template <int BLOCK>
BPD uint32_t predict_bodies(const Arena *A, const WorldDesc *d, uint32_t pub, uint32_t wside) {
    uint32_t safety = 0;
    for (uint32_t k = threadIdx.x; k < d->bodyCount; k += BLOCK) {
        uint32_t g = d->bodyOffset + k;
        float4 pose = A->bodyPose[pub][g];
        float4 motion = A->bodyMotion[g];
        uint4 kind = A->bodyKind[g];
        float2 prev = f2(pose.x, pose.y);
        pose.w = pose.z;
        if (kind.y & kBodyStatic) {
            motion.x = motion.y = motion.z = 0.0f;
        } else {
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
    return safety;
}

// Reduction of h for the whole world once the three terms are known (Force:633-736).
// Every lane evaluates it with the same inputs, so the control state stays uniform.
BPD void reduce_h(Ctl &c, float stiffnessBound, uint32_t w2Owner, float dbMin, uint32_t dbOwner, float mbMin,
                  uint32_t mbOwner, float *history, uint32_t lane) {
    float remaining = c.t0 - c.t1;
    if (mbMin <= 0.0f) {
        c.rebuild = 1; c.h = 0.0f;
        c.limOwner = mbOwner; c.limTerm = BP_TERM_MARGIN;
        return;
    }
    float split = fmaxf(c.split, 1.0f);
    float h = remaining;
    uint32_t term = BP_TERM_REMAINING;
    uint32_t object = kNoOwner;
    if (split > 1.0f && c.t0 / split < h) h = c.t0 / split;
    if (stiffnessBound / split < h) { h = stiffnessBound / split; term = BP_TERM_STIFFNESS; object = w2Owner; }
    if (dbMin / split < h) { h = dbMin / split; term = BP_TERM_DISTANCE; object = dbOwner; }
    if (mbMin < h) { h = mbMin; term = BP_TERM_MARGIN; object = mbOwner; c.rebuild = 1; }
    h = fmaxf(h, c.t0 * 1e-7f);
    uint32_t frameSubstep = c.c1;
    if (h >= remaining || remaining - h <= c.t0 * 1e-5f) { h = remaining; c.t1 = c.t0; }
    else c.t1 = c.t1 + h;
    c.h = h;
    c.c0 += 1;
    c.c1 = frameSubstep + 1;
    c.limOwner = object; c.limTerm = term;
    if (lane == 0) history[frameSubstep] = h;
}

// Id of the limiting bubble (substepControl.limiter.x), 0 when no bubble limits the step.
BPD uint64_t limiter_object(const Ctl &c, const int64_t *ids) {
    return c.limOwner == kNoOwner ? 0u : (uint64_t)ids[c.limOwner];
}

// Region removal or wrap of one bubble (PostSolve:536-566). Returns the region result and
// updates p and start; sets safety bit 1 (and leaves the record) when the wrap is not finite.
BPD uint32_t apply_region(float2 &p, float2 &start, float r, uint32_t flags, float2 lo, float2 hi, uint32_t &safety) {
    if (!(flags & kEnabled) || (flags & kStatic)) return BP_REGION_UNCHANGED;
    float2 nearest = f2(clampf(p.x, lo.x, hi.x), clampf(p.y, lo.y, hi.y));
    if (len2(sub(p, nearest)) <= r) return BP_REGION_UNCHANGED;
    if (!(flags & kWrap)) return BP_REGION_REMOVED;
    float2 span = adds(sub(hi, lo), 2.0f * r);
    float dx = 0.0f, dy = 0.0f;
    if (p.x > hi.x) dx = -span.x * fmaxf(1.0f, ceilf((p.x - r - hi.x) / span.x));
    else if (p.x < lo.x) dx = span.x * fmaxf(1.0f, ceilf((lo.x - r - p.x) / span.x));
    if (p.y > hi.y) dy = -span.y * fmaxf(1.0f, ceilf((p.y - r - hi.y) / span.y));
    else if (p.y < lo.y) dy = span.y * fmaxf(1.0f, ceilf((lo.y - r - p.y) / span.y));
    float2 dl = f2(dx, dy);
    float2 wrapped = add(p, dl);
    float2 nearestWrapped = f2(clampf(wrapped.x, lo.x, hi.x), clampf(wrapped.y, lo.y, hi.y));
    if (len2(sub(wrapped, nearestWrapped)) > r) dl = add(dl, sub(nearestWrapped, wrapped));
    float2 np = add(p, dl), ns = add(start, dl);
    if (!finite2(np) || !finite2(ns)) { safety |= 1u; return BP_REGION_UNCHANGED; }
    p = np;
    start = ns;
    return BP_REGION_WRAPPED;
}

struct Violation { uint32_t kind; uint64_t other; int64_t segment; float detail; };

// Body and one-sided edge assertions of a movable bubble (Force:889-917), in the reference's order.
BPD bool assert_body_and_edges(const Geo &g, float2 center, float2 start, uint32_t cat, uint32_t mask, float fraction,
                               Violation &out) {
    for (uint32_t k = 0; k < g.nb; ++k) {
        BodyView body = g.body(k);
        if (!body_immovable(body) || !filter_accepts(cat, mask, body.kind.z, body.kind.w)) continue;
        if (frame_inside(center, body_frame(body, fraction), g.verts(), g.tris())) {
            out.kind = BP_VIOLATION_INSIDE_BODY; out.other = g.bodyId(k); out.segment = 0; out.detail = 0.0f;
            return true;
        }
    }
    for (uint32_t j = 0; j < g.ns; ++j) {
        Seg s = g.segment(j);
        if (!(s.flags & kSegPolygon) || !filter_accepts(cat, mask, g.segCat(j), g.segMask(j))) continue;
        float2 edge = sub(s.b, s.a);
        float lengthSquared = dot2(edge, edge);
        if (!(lengthSquared > 0.0f)) continue;
        float2 outward = mul(dvs(f2(-edge.y, edge.x), __builtin_sqrtf(lengthSquared)), s.outward);
        float before = dot2(sub(start, s.a), outward), after = dot2(sub(center, s.a), outward);
        if (!(before > 0.0f) || !(after <= 0.0f)) continue;
        float2 crossing = add(start, mul(sub(center, start), before / (before - after)));
        float along = dot2(sub(crossing, s.a), edge) / lengthSquared;
        if (along < 0.0f || along > 1.0f) continue;
        out.kind = BP_VIOLATION_EDGE_SIDE; out.segment = g.segId(j); out.other = (uint64_t)g.segOwner(j); out.detail = after;
        return true;
    }
    return false;
}

BPD float step_fraction(const Ctl &c) { return c.t0 > 0.0f ? clampf(c.t1 / c.t0, 0.0f, 1.0f) : 1.0f; }

// Status word of a finished step (error precedence, NativeMetalSolver.swift:1026-1040).
BPD int32_t final_status(int32_t status, const Ctl &c, uint32_t safetyAll, bool claimed, bool finished, uint32_t n) {
    if (status != BP_OK) return status;
    if (c.limit) return BP_ERR_SUBSTEP_LIMIT_EXCEEDED;
    if (safetyAll & 1u) return BP_ERR_NON_FINITE;
    if (claimed) return BP_ERR_CONTACT_VIOLATION;
    if (n > 0 && !finished) return BP_ERR_COMMAND_FAILED;
    return BP_OK;
}

}  // namespace bpd

namespace bpd {

// Record i of a served world's published state into its mirror in pinned host memory (the server
// reads it once the launch has completed, for the host staging of add, merge and split).
BPD void mirror_record(const Arena *A, const WorldDesc *d, uint32_t side, uint32_t i, float2 pos, float2 vel, float rot,
                       float om) {
    const uint32_t g = d->bubbleOffset + i;
    bp_bubble r;
    r.id = A->id[g];
    r.center[0] = pos.x; r.center[1] = pos.y;
    const float2 st = A->start[side][g];
    r.previous_center[0] = st.x; r.previous_center[1] = st.y;
    r.velocity[0] = vel.x; r.velocity[1] = vel.y;
    r.rotation = rot;
    r.angular_velocity = om;
    r.mass = A->mass[g];
    r.inverse_mass = A->invm[g];
    r.radius = A->radius[g];
    r.contact_stiffness = A->Kc[g];
    r.damping = A->gamma[g];
    r.flags = A->flags[g];
    r.category = A->cat[g];
    r.mask = A->mask[g];
    static_cast<bp_bubble *>(d->mirror)[i] = r;
}

// The bodies' published poses into the mirror (lanes `lane`, `lane + width`, ...).
BPD void mirror_bodies(const Arena *A, const WorldDesc *d, uint32_t side, uint32_t lane, uint32_t width) {
    for (uint32_t k = lane; k < d->bodyCount; k += width) d->bodyMirror[k] = A->bodyPose[side][d->bodyOffset + k];
}

}  // namespace bpd

#endif  // BP_COMMON_H
