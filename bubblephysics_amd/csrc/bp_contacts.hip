// bp_contacts.hip -- end-of-step contact records of a world's published state
// (bp_world_contacts, bp_batch_contacts).
//
// The reference builds them after the region pass from the candidate list of the last grid
// build and the final centres (NativeMetalSolver.swift:846-871):
//   * one record per candidate: a bubble pair (the lower index first) or a bubble and a segment
//     (nativeBuildContacts, NativeGeometryKernels.metal-source:683-728, with the polygon edge
//     feature of :670-681);
//   * kept iff penetration > -separation tolerance (nativePrepareContactFlags,
//     NativeContactKernels.metal-source:146-153), compacted in candidate order (:155-160);
//   * the contact law's force at the record's geometry times dt^2 (nativeContactImpulses,
//     NativeForceKernels.metal-source:261-280).
// The candidate list is ordered as the reference compacts it (NativeGeometryKernels.metal-source:
// 625-651): every bubble's pair candidates (partner index ascending), then every bubble's segment
// candidates (segment ascending, which is id order). Here the candidates are read from the
// exported adjacency rows of the last build: the pair entries above the bubble's own index and the
// segment entries. Body entries feed the bubble-body records (nativeBubbleBodyContacts), a list of
// their own (bp_world_body_contacts).
//
// One workgroup per world: a counting pass writes each list slot's kept-record count, a block scan
// in chunks gives the offsets, and a writing pass recomputes and stores the kept records.
#include "bp_world_global.h"

#include <mutex>

namespace bpd {

namespace {

constexpr int kContactBlock = 256;
constexpr uint32_t kSegAllowedSide = 16u;      // BP_SEGMENT_ALLOWED_SIDE

struct Contact {
    float2 normal, point;
    float penetration, allowedSide;
    uint32_t feature;
};

// nativeBuildContacts, pair branch (Geometry:690-695).
__device__ Contact pair_contact(float2 pa, float ra, float2 pb, float rb) {
    Contact c;
    const float2 delta = sub(pb, pa);
    const float distanceValue = len2(delta);
    c.normal = distanceValue > 0.0f ? dvs(delta, distanceValue) : f2(1.0f, 0.0f);
    c.point = add(pa, mul(c.normal, ra));
    c.penetration = ra + rb - distanceValue;
    c.allowedSide = 0.0f;
    c.feature = 0u;
    return c;
}

// nativePolygonEdgeFeature (Geometry:674-681).
__device__ uint32_t polygon_edge_feature(float2 center, float2 start, float2 end, float2 previousStart, uint32_t flags,
                                         float2 outward) {
    const float along = dot2(sub(center, start), sub(end, start));
    const bool convexStart = (flags & kSegConvexStart) != 0u;
    if (convexStart && along <= 0.0f && past_segment_end(center, previousStart, start)) return 2u;
    const bool interior = along > 0.0f && !past_segment_end(center, start, end);
    return dot2(sub(center, start), outward) >= 0.0f && interior ? 1u : 0u;
}

// nativeBuildContacts, segment branch (Geometry:696-724).
__device__ Contact segment_contact(float2 pa, float ra, int64_t idA, const World &W, uint32_t j) {
    const Arena *A = W.A;
    const uint32_t g = W.d->segmentOffset + j;
    const float4 ab = A->segAB[g], rg = A->segRegion[g];
    const uint32_t flags = A->segFlags[g];
    const float2 sa = f2(ab.x, ab.y), sb = f2(ab.z, ab.w);
    Contact c;
    const float2 edge = sub(sb, sa);
    const float lengthSquared = dot2(edge, edge);
    const float t = lengthSquared > 0.0f ? clampf(dot2(sub(pa, sa), edge) / lengthSquared, 0.0f, 1.0f) : 0.0f;
    c.point = add(sa, mul(t, edge));
    float2 delta = sub(pa, c.point);
    float distanceValue = len2(delta);
    const float edgeLength = len2(edge);
    const float2 base = edgeLength > 0.0f ? dvs(f2(-edge.y, edge.x), edgeLength) : f2(0.0f, 1.0f);
    const float sign = (((uint64_t)idA ^ (uint64_t)A->segId[g]) & 1u) == 0u ? 1.0f : -1.0f;
    const bool sided = (flags & kSegAllowedSide) != 0u;
    c.normal = sided ? mul(base, rg.w) : (distanceValue > 0.0f ? dvs(delta, distanceValue) : mul(base, sign));
    c.allowedSide = sided ? rg.w : 0.0f;
    c.penetration = ra - distanceValue;
    c.feature = 0u;
    if (flags & kSegPolygon) {
        const float2 outward = mul(base, rg.z);
        c.feature = polygon_edge_feature(pa, sa, sb, f2(rg.x, rg.y), flags, outward);
        if (c.feature == 1u) {
            c.normal = outward;
            const float depth = dot2(sub(pa, sa), outward);
            c.point = sub(pa, mul(outward, depth));
            c.penetration = ra - depth;
        } else if (c.feature == 2u) {
            c.point = sa;
            delta = sub(pa, c.point);
            distanceValue = len2(delta);
            c.normal = distanceValue > 0.0f ? dvs(delta, distanceValue) : outward;
            c.penetration = ra - distanceValue;
        } else {
            c.normal = outward;
            c.point = sa;
            c.penetration = -3.402823466e+38f;
        }
    }
    return c;
}

// nativeBubbleBodyContacts (NativeForceKernels.metal-source:337-358): bubble i against its body
// candidates at the bodies' end-of-step poses (fraction 1); a record iff distance < radius and the
// normal is not zero. Records of one bubble come in its row's order (body ascending); the host
// orders them body by body.
template <bool WRITE>
__device__ uint32_t walk_bodies(const World &W, uint32_t i, float dt, const uint32_t *offsets, ContactRecord *out,
                                uint32_t capacity) {
    const float2 pa = W.pos[i];
    const float ra = W.radius[i], Ka = W.Kc[i];
    const uint32_t *row = W.adj + W.rowStart[i];
    const uint32_t len = W.rowLen[i];
    uint32_t kept = 0, base = 0;
    if (WRITE) base = i == 0 ? 0u : offsets[i - 1];
    for (uint32_t e = 0; e < len; ++e) {
        const uint32_t entry = row[e], kind = entry >> kKindShift, k = entry & kIndexMask;
        if (kind != 2u) continue;
        const Obstacle ob = body_obstacle(W.body(k), 1.0f, W.verts, W.tris, pa);
        if (!(ob.distanceValue < ra)) continue;
        const float2 normal = f2(-ob.normal.x, -ob.normal.y);
        if (normal.x == 0.0f && normal.y == 0.0f) continue;
        if (WRITE && base + kept < capacity) {
            const float2 point = sub(pa, mul(ob.normal, ob.distanceValue));
            ContactRecord r;
            r.first = W.id[i];
            r.second = (int64_t)W.A->bodyId[W.d->bodyOffset + k];
            r.kind = 2u;
            r.feature = 0u;
            r.normal[0] = normal.x; r.normal[1] = normal.y;
            r.point[0] = point.x; r.point[1] = point.y;
            r.penetration = ra - ob.distanceValue;
            r.allowedSide = 0.0f;
            r.impulse = wall_law(Ka, ra, ob.distanceValue).force * dt * dt;
            r.indexB = k;
            out[base + kept] = r;
        }
        ++kept;
    }
    return kept;
}

// The list slot of bubble i's pair candidates is i, of its segment candidates n + i.
template <bool WRITE>
__device__ uint32_t walk_slot(const World &W, uint32_t slot, float sepTol, float dt, const uint32_t *offsets,
                              ContactRecord *out, uint32_t capacity) {
    const uint32_t n = W.n;
    const bool segments = slot >= n;
    const uint32_t i = segments ? slot - n : slot;
    const float2 pa = W.pos[i];
    const float ra = W.radius[i], Ka = W.Kc[i];
    const uint32_t *row = W.adj + W.rowStart[i];
    const uint32_t len = W.rowLen[i];
    uint32_t kept = 0, base = 0;
    if (WRITE) base = slot == 0 ? 0u : offsets[slot - 1];
    for (uint32_t e = 0; e < len; ++e) {
        const uint32_t entry = row[e], kind = entry >> kKindShift, j = entry & kIndexMask;
        if (kind != (segments ? 1u : 0u) || (kind == 0u && j <= i)) continue;
        const Contact c = segments ? segment_contact(pa, ra, W.id[i], W, j) : pair_contact(pa, ra, W.pos[j], W.radius[j]);
        if (!(c.penetration > -sepTol)) continue;
        if (WRITE && base + kept < capacity) {
            ContactRecord r;
            r.first = W.id[i];
            r.second = segments ? W.A->segId[W.d->segmentOffset + j] : W.id[j];
            r.kind = segments ? 1u : 0u;
            r.feature = c.feature;
            r.normal[0] = c.normal.x; r.normal[1] = c.normal.y;
            r.point[0] = c.point.x; r.point[1] = c.point.y;
            r.penetration = c.penetration;
            r.allowedSide = c.allowedSide;
            // nativeContactImpulses: the law at the distance the record implies.
            float force;
            if (!segments) force = pair_law(Ka, ra, W.Kc[j], W.radius[j], W.radius[i] + W.radius[j] - c.penetration).force;
            else force = wall_law(Ka, ra, ra - c.penetration).force;
            r.impulse = force * dt * dt;
            r.indexB = j;
            out[base + kept] = r;
        }
        ++kept;
    }
    return kept;
}

// BODIES = false: pair and segment records (2 n slots); true: bubble-body records (n slots).
template <bool BODIES>
__global__ __launch_bounds__(kContactBlock) void contacts_kernel(const Arena *__restrict__ A, const WorldDesc *descs,
                                                                 uint32_t world, uint32_t *offsets, ContactRecord *out,
                                                                 uint32_t capacity, uint32_t *total) {
    const WorldDesc *d = descs + world;
    const World W = make_world(A, d, d->pub);
    const uint32_t slots = BODIES ? W.n : 2 * W.n;
    const float sepTol = d->sepTol, dt = d->dt;
    for (uint32_t s = threadIdx.x; s < slots; s += kContactBlock)
        offsets[s] = BODIES ? walk_bodies<false>(W, s, dt, nullptr, nullptr, 0)
                            : walk_slot<false>(W, s, sepTol, dt, nullptr, nullptr, 0);
    __syncthreads();
    // Inclusive scan over the slots in chunks of the workgroup width (nativeScanInGroup).
    __shared__ uint32_t part[kContactBlock];
    uint32_t carry = 0;
    for (uint32_t chunk = 0; chunk < slots; chunk += kContactBlock) {
        const uint32_t s = chunk + threadIdx.x;
        uint32_t v = s < slots ? offsets[s] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t dist = 1; dist < kContactBlock; dist <<= 1) {
            const uint32_t add = threadIdx.x >= dist ? part[threadIdx.x - dist] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (s < slots) offsets[s] = part[threadIdx.x] + carry;
        carry += part[kContactBlock - 1];
        __syncthreads();
    }
    for (uint32_t s = threadIdx.x; s < slots; s += kContactBlock) {
        if (BODIES) walk_bodies<true>(W, s, dt, offsets, out, capacity);
        else walk_slot<true>(W, s, sepTol, dt, offsets, out, capacity);
    }
    if (threadIdx.x == 0) *total = carry;
}

}  // namespace

}  // namespace bpd

// Contact records of one world into `out` (device, `capacity` records); `offsets` holds 2 n
// device words; `total` receives the number of kept records (those past `capacity` are counted,
// not written).
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_contacts(const bpd::Arena *arena, const bpd::WorldDesc *descs,
                                                                           uint32_t world, uint32_t *offsets, void *out,
                                                                           uint32_t capacity, uint32_t *total,
                                                                           hipStream_t stream) {
    hipLaunchKernelGGL(bpd::contacts_kernel<false>, dim3(1), dim3(bpd::kContactBlock), 0, stream, arena, descs, world, offsets,
                       static_cast<bpd::ContactRecord *>(out), capacity, total);
    return hipGetLastError();
}

// Bubble-body records of one world, as bp_launch_contacts with n offset words.
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_body_contacts(const bpd::Arena *arena,
                                                                                const bpd::WorldDesc *descs, uint32_t world,
                                                                                uint32_t *offsets, void *out,
                                                                                uint32_t capacity, uint32_t *total,
                                                                                hipStream_t stream) {
    hipLaunchKernelGGL(bpd::contacts_kernel<true>, dim3(1), dim3(bpd::kContactBlock), 0, stream, arena, descs, world, offsets,
                       static_cast<bpd::ContactRecord *>(out), capacity, total);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Scratch of the entry point that allocates nothing (bp_world_contacts_noalloc): one fixed
// device array per GPU for the slot offsets (2 n words) and the records, instead of a buffer
// allocated per call. Every call on a GPU shares it and it keeps the last call's contents; calls
// that overlap on one GPU race.
// ---------------------------------------------------------------------------------------
namespace bpd {
constexpr uint32_t kNoAllocMaxBubbles = 32768;                 // worlds up to this many bubbles
constexpr uint32_t kNoAllocMaxRecords = 65536;                 // up to this many records per call
constexpr size_t kNoAllocOffsetBytes = (size_t)2 * kNoAllocMaxBubbles * sizeof(uint32_t);   // 256 KiB
constexpr size_t kNoAllocRecordBytes = (size_t)kNoAllocMaxRecords * sizeof(ContactRecord);  // 3.5 MiB
static_assert(kNoAllocOffsetBytes % 256 == 0 && kNoAllocRecordBytes % 256 == 0, "scratch sections stay 256-aligned");
}  // namespace bpd

static __device__ __align__(256) unsigned char scratch[bpd::kNoAllocOffsetBytes + bpd::kNoAllocRecordBytes + 256];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchPtr[kMaxDevices] = {};                          // device address of `scratch`, per GPU
}  // namespace

// Device address of this GPU's scratch array, looked up once per device with that device current.
// The array's name is a host-side stand-in: only this address may be given to kernels and copies.
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_contacts_scratch(void **out, uint32_t *maxBubbles,
                                                                            uint32_t *maxRecords) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[device]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[device] = p;
    }
    *out = g_scratchPtr[device];
    *maxBubbles = bpd::kNoAllocMaxBubbles;
    *maxRecords = bpd::kNoAllocMaxRecords;
    return hipSuccess;
}
