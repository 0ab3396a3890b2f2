// bp_audit.hip -- FP64 audit of a world's published state (bp_world_audit, bp_batch_audit).
//
// One workgroup per world. Each lane sums, in FP64, the momentum, angular momentum and kinetic
// energy of its movable bubbles and the elastic energy of the contacts in their adjacency rows
// (the rows of the last grid build, which hold every contact of the published state: a grid is
// rebuilt before any bubble leaves its margin). A contact's distance is the FP32 one the step
// uses (pair centre distance in FP64 from the FP32 centres; wall regions and body obstacles at
// the end of the step in FP32); its energy is FP64:
//   U(eps; K, r) = K r (eps / (1 - eps) + ln(1 - eps))          ContactLawReference.swift:26-31
// per deflected bubble, with a pair of different materials split by
//   r_a eps(F/K_a) + r_b eps(F/K_b) = overlap                    ContactLawReference.swift:68-82
// solved for F by bisection. The workgroup then reduces its lanes in a fixed tree.
#include "bp_world_global.h"

namespace bpd {

namespace {

__device__ double audit_epsilon(double q) { return q > 0.0 ? 2.0 * q / ((2.0 * q + 1.0) + sqrt(4.0 * q + 1.0)) : 0.0; }

__device__ double audit_spring(double eps, double K, double r) {
    return eps > 0.0 ? K * r * (eps / (1.0 - eps) + log1p(-eps)) : 0.0;
}

// Wall and body contacts: the wall law's deflection (Force:60-69) with d floored at r 1e-6.
__device__ double audit_wall(double K, double r, double d, uint32_t &contacts) {
    const double dd = fmax(d, r * 1e-6);
    if (!(K > 0.0) || !(r > 0.0) || !(dd < r)) return 0.0;
    ++contacts;
    return audit_spring((r - dd) / r, K, r);
}

__device__ double audit_pair(double Ka, double ra, double Kb, double rb, double d, uint32_t &contacts) {
    const double R = ra + rb, dd = fmax(d, R * 1e-6), overlap = R - dd;
    if (!(Ka > 0.0) || !(Kb > 0.0) || !(overlap > 0.0)) return 0.0;
    ++contacts;
    if (Ka == Kb) return audit_spring(overlap / R, Ka, R);
    double lo = 0.0, hi = fmax(Ka, Kb);
    while (ra * audit_epsilon(hi / Ka) + rb * audit_epsilon(hi / Kb) < overlap) hi *= 2.0;
    for (int k = 0; k < 100; ++k) {
        const double mid = 0.5 * (lo + hi);
        if (ra * audit_epsilon(mid / Ka) + rb * audit_epsilon(mid / Kb) < overlap) lo = mid; else hi = mid;
    }
    const double f = 0.5 * (lo + hi);
    return audit_spring(audit_epsilon(f / Ka), Ka, ra) + audit_spring(audit_epsilon(f / Kb), Kb, rb);
}

constexpr int kAuditBlock = 256;

}  // namespace

struct AuditRecord {        // bp_audit
    double momentum[2], angular, kinetic, contact;
    uint32_t contacts, bubbles;
};

__global__ __launch_bounds__(kAuditBlock) void audit_kernel(const Arena *__restrict__ A, const WorldDesc *descs,
                                                            AuditRecord *out) {
    const WorldDesc *d = descs + blockIdx.x;
    const uint32_t pub = d->pub;
    World W = make_world(A, d, pub);
    const bool rows = d->exportAdjacency != 0;
    double px = 0.0, py = 0.0, ang = 0.0, kin = 0.0, pot = 0.0;
    uint32_t contacts = 0;
    for (uint32_t i = threadIdx.x; i < W.n; i += kAuditBlock) {
        const float2 p = W.pos[i];
        if (!W.is_static(i)) {
            const double m = W.mass[i], vx = W.vel[i].x, vy = W.vel[i].y;
            px += m * vx;
            py += m * vy;
            ang += m * ((double)p.x * vy - (double)p.y * vx);
            kin += 0.5 * m * (vx * vx + vy * vy);
        }
        if (!rows) continue;
        const float ri = W.radius[i], K = W.Kc[i];
        const uint32_t *row = W.adj + W.rowStart[i];
        const uint32_t len = W.rowLen[i];
        for (uint32_t e = 0; e < len; ++e) {
            const uint32_t entry = row[e], kind = entry >> kKindShift, j = entry & kIndexMask;
            if (kind == 0u) {
                if (j <= i) continue;                        // each pair once, from its lower end
                const double dx = (double)W.pos[j].x - (double)p.x, dy = (double)W.pos[j].y - (double)p.y;
                pot += audit_pair(K, ri, W.Kc[j], W.radius[j], sqrt(dx * dx + dy * dy), contacts);
            } else if (kind == 1u) {
                float2 normal; float dd;
                if (!wall_region(W.segment(j), p, normal, dd)) continue;
                pot += audit_wall(K, ri, dd, contacts);
            } else {
                const Obstacle ob = body_obstacle(W.body(j), 1.0f, W.verts, W.tris, p);
                pot += audit_wall(K, ri, ob.distanceValue, contacts);
            }
        }
    }
    __shared__ double red[5][kAuditBlock];
    __shared__ uint32_t redc[kAuditBlock];
    red[0][threadIdx.x] = px; red[1][threadIdx.x] = py; red[2][threadIdx.x] = ang; red[3][threadIdx.x] = kin;
    red[4][threadIdx.x] = pot; redc[threadIdx.x] = contacts;
    __syncthreads();
    for (int s = kAuditBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
            redc[threadIdx.x] += redc[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        AuditRecord r;
        r.momentum[0] = red[0][0]; r.momentum[1] = red[1][0]; r.angular = red[2][0]; r.kinetic = red[3][0];
        r.contact = rows ? red[4][0] : __builtin_nan("");
        r.contacts = rows ? redc[0] : 0u;
        r.bubbles = W.n;
        out[blockIdx.x] = r;
    }
}

}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_audit(const bpd::Arena *arena, const bpd::WorldDesc *descs,
                                                                        uint32_t worlds, void *out, hipStream_t stream) {
    if (worlds == 0) return hipSuccess;
    hipLaunchKernelGGL(bpd::audit_kernel, dim3(worlds), dim3(bpd::kAuditBlock), 0, stream, arena, descs,
                       static_cast<bpd::AuditRecord *>(out));
    return hipGetLastError();
}
