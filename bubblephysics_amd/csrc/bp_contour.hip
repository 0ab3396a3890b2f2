// bp_contour.hip -- contour and render preparation of a world's published state
// (bp_world_contours, bp_batch_contours).
//
// The reference runs three kernels at the end of every frame (NativeMetalSolver.swift:884-903):
// nativeAccumulateContactPressure (one thread per contour sample: the ray of the sample's angle
// cut by the segments and by the flat chords of the overlapping pair partners,
// NativePostSolveKernels.metal-source:280-337), 13 Jacobi smoothing passes
// (nativeGenerateContoursFused, :358-386) and nativePrepareRenderData (:388-441: the slope-bounded
// majorant from the deepest sample, points, bounds, largest deficit and effective radius).
// Here the three are one kernel with one wavefront per bubble: its lanes take the samples, the
// deficits stay in LDS through the passes, lane 0 runs the two serial majorant sweeps, and the
// bounds and extrema are wave reductions (min and max are exact, so the order does not matter).
#include "bp_world_global.h"

namespace bpd {

namespace {

constexpr int kContourLanes = 64;
constexpr uint32_t kContourCap = 2048;     // samples held in LDS per bubble
constexpr float kPi = 3.14159265358979323846f;

// nativeContourRayLimit (PostSolve:268-277).
BPD float ray_limit(float2 center, float2 direction, float2 a, float2 b, float radius, float clearance) {
    const float2 edge = sub(b, a), offset = sub(a, center);
    const float den = cross2(direction, edge);
    if (fabsf(den) < 1e-6f) return radius;
    const float dist = cross2(offset, edge) / den;
    const float fraction = cross2(offset, direction) / den;
    return dist >= 0.0f && fraction >= 0.0f && fraction <= 1.0f ? fminf(radius, fmaxf(0.0f, dist - clearance)) : radius;
}

__device__ float wave_min(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ float wave_max(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__global__ __launch_bounds__(kContourLanes) void contour_kernel(const Arena *__restrict__ A, const WorldDesc *descs,
                                                               uint32_t world, bp_contour_config cfg, uint32_t stride,
                                                               float2 *points, bp_bubble_contour *shapes) {
    __shared__ float level[2][kContourCap];
    __shared__ float floorLevel[kContourCap];
    const WorldDesc *d = descs + world;
    const World W = make_world(A, d, d->pub);
    const uint32_t i = blockIdx.x, lane = threadIdx.x;
    if (i >= W.n) return;
    const float2 pos = W.pos[i];
    const float r = W.radius[i], Ki = W.Kc[i];
    const uint32_t cat = W.cat[i], msk = W.mask[i];
    const uint32_t count = min(stride, max(64u, (uint32_t)ceilf((r / cfg.max_segment_length) * (4.0f * kPi))));
    const uint32_t *row = W.adj + W.rowStart[i];
    const uint32_t len = W.rowLen[i];

    // Ray limits (PostSolve:280-337).
    for (uint32_t s = lane; s < count; s += kContourLanes) {
        const float angle = 2.0f * kPi * float(s) / float(count);
        const float2 dir = f2(bpm::cosf_det(angle), bpm::sinf_det(angle));
        float limit = r;
        for (uint32_t j = 0; j < W.ns; ++j) {
            const uint32_t g = d->segmentOffset + j;
            if (!filter_accepts(cat, msk, A->segCat[g], A->segMask[g])) continue;
            const float4 ab = A->segAB[g];
            limit = ray_limit(pos, dir, f2(ab.x, ab.y), f2(ab.z, ab.w), limit, cfg.position_tolerance);
        }
        for (uint32_t e = 0; e < len; ++e) {
            const uint32_t entry = row[e];
            if ((entry >> kKindShift) != 0u) continue;
            const uint32_t j = entry & kIndexMask;
            const float2 delta = sub(W.pos[j], pos);
            const float dist = len2(delta);
            const float rj = W.radius[j];
            if (dist <= 1.1920929e-7f || dist >= r + rj) continue;
            const float2 nrm = dvs(delta, dist);
            const bool lowIsSelf = i < j;
            const Law law = lowIsSelf ? pair_law(Ki, r, W.Kc[j], rj, dist) : pair_law(W.Kc[j], rj, Ki, r, dist);
            const float eps = lowIsSelf ? law.epsA : law.epsB;
            const float plane = clampf(r * (1.0f - eps), 0.0f, r);
            const float half = __builtin_sqrtf(fmaxf(0.0f, r * r - plane * plane));
            const float2 tangent = f2(-nrm.y, nrm.x);
            const float2 point = add(pos, mul(nrm, plane));
            limit = ray_limit(pos, dir, sub(point, mul(tangent, half)), add(point, mul(tangent, half)), limit,
                              cfg.position_tolerance);
        }
        const float deficit = clampf(1.0f - limit / r, 0.0f, 1.0f);
        level[0][s] = deficit;
        floorLevel[s] = deficit;
    }
    __syncthreads();

    // 13 Jacobi passes (PostSolve:358-386).
    const float tension = fminf(0.8f, cfg.surface_tension / (cfg.surface_tension + 10.0f));
    int cur = 0;
    for (int pass = 0; pass < 13; ++pass) {
        for (uint32_t s = lane; s < count; s += kContourLanes) {
            const float x = level[cur][s];
            const float neighbor = (level[cur][(s + count - 1) % count] + level[cur][(s + 1) % count]) * 0.5f;
            const float mixed = x + (neighbor - x) * tension;
            level[cur ^ 1][s] = clampf(fmaxf(floorLevel[s], mixed), 0.0f, 1.0f);
        }
        cur ^= 1;
        __syncthreads();
    }
    float *x = level[cur];

    // Slope-bounded majorant from the deepest sample (PostSolve:392-417), one lane.
    __shared__ float sMaximum;
    if (lane == 0) {
        float maximum = 0.0f;
        uint32_t anchor = 0;
        for (uint32_t s = 0; s < count; ++s)
            if (x[s] > maximum) {
                maximum = x[s];
                anchor = s;
            }
        const float slope = cfg.max_segment_length / (2.0f * r);
        float previous = maximum;
        for (uint32_t step = 1; step < count; ++step) {
            const uint32_t s = (anchor + step) % count;
            const float value = fmaxf(x[s], previous - slope);
            x[s] = value;
            previous = value;
        }
        previous = maximum;
        for (uint32_t step = 1; step < count; ++step) {
            const uint32_t s = (anchor + count - step) % count;
            const float value = fmaxf(x[s], previous - slope);
            x[s] = value;
            previous = value;
        }
        sMaximum = maximum;
    }
    __syncthreads();

    // Points, bounds and extrema (PostSolve:418-441).
    float2 lo = pos, hi = pos;
    float nearest = r, deepest = sMaximum;
    float2 *out = points + (size_t)i * stride;
    for (uint32_t s = lane; s < count; s += kContourLanes) {
        const float angle = 2.0f * kPi * float(s) / float(count);
        const float2 dir = f2(bpm::cosf_det(angle), bpm::sinf_det(angle));
        const float2 point = add(pos, mul(mul(dir, r), 1.0f - x[s]));
        out[s] = point;
        lo = min2(lo, point);
        hi = max2(hi, point);
        nearest = fminf(nearest, len2(sub(point, pos)));
        deepest = fmaxf(deepest, x[s]);
    }
    lo.x = wave_min(lo.x);
    lo.y = wave_min(lo.y);
    hi.x = wave_max(hi.x);
    hi.y = wave_max(hi.y);
    nearest = wave_min(nearest);
    deepest = wave_max(deepest);
    if (lane == 0) {
        bp_bubble_contour c;
        c.count = count;
        c.maximum = deepest;
        c.effective_radius = nearest;
        c.reserved = 0.0f;
        c.bounds[0] = lo.x;
        c.bounds[1] = lo.y;
        c.bounds[2] = hi.x;
        c.bounds[3] = hi.y;
        shapes[i] = c;
    }
}

}  // namespace

}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) uint32_t bp_contour_capacity() { return bpd::kContourCap; }

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_contours(const bpd::Arena *arena,
                                                                               const bpd::WorldDesc *descs, uint32_t world,
                                                                               uint32_t bubbles, bp_contour_config cfg,
                                                                               uint32_t stride, float2 *points,
                                                                               bp_bubble_contour *shapes, hipStream_t stream) {
    if (bubbles == 0) return hipSuccess;
    hipLaunchKernelGGL(bpd::contour_kernel, dim3(bubbles), dim3(bpd::kContourLanes), 0, stream, arena, descs, world, cfg,
                       stride, points, shapes);
    return hipGetLastError();
}
