// bp_transfer.hip -- one world's bubbles between host records and the structure-of-arrays arena
// in one copy each way (serving transactions read a world, change it and write it back).
#include "bp_device.h"
#include "../../include/bubblephysics.h"

namespace bpd {

// State of the published side of one world into bp_bubble records (dynamic fields only; the
// host keeps the properties it wrote).
__global__ void gather_world_kernel(const Arena *__restrict__ A, const WorldDesc *d, bp_bubble *out) {
    const uint32_t n = d->bubbleCount, o = d->bubbleOffset, p = d->pub;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = o + i;
        const float2 c = A->pos[p][g], s = A->start[p][g], v = A->vel[p][g];
        bp_bubble &b = out[i];
        b.center[0] = c.x; b.center[1] = c.y;
        b.previous_center[0] = s.x; b.previous_center[1] = s.y;
        b.velocity[0] = v.x; b.velocity[1] = v.y;
        b.rotation = A->rot[p][g];
        b.angular_velocity = A->omega[p][g];
    }
}

// Records (and their motor and band indices) into both sides of the world's range.
__global__ void scatter_world_kernel(const Arena *__restrict__ A, uint32_t offset, uint32_t count, const bp_bubble *in,
                                     const int32_t *motorIdx, const int32_t *bandIdx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const bp_bubble b = in[i];
        const uint32_t g = offset + i;
        for (int side = 0; side < 2; ++side) {
            A->pos[side][g] = make_float2(b.center[0], b.center[1]);
            A->start[side][g] = make_float2(b.previous_center[0], b.previous_center[1]);
            A->vel[side][g] = make_float2(b.velocity[0], b.velocity[1]);
            A->rot[side][g] = b.rotation;
            A->omega[side][g] = b.angular_velocity;
            A->region[side][g] = 0u;
        }
        A->id[g] = b.id;
        A->mass[g] = b.mass; A->invm[g] = b.inverse_mass; A->radius[g] = b.radius; A->Kc[g] = b.contact_stiffness;
        A->gamma[g] = b.damping; A->flags[g] = b.flags; A->cat[g] = b.category; A->mask[g] = b.mask;
        A->motor[g] = motorIdx[i];
        A->band[g] = bandIdx[i];
    }
}

}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_gather_world(const bpd::Arena *arena, const bpd::WorldDesc *desc,
                                                                               bp_bubble *out, uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(bpd::gather_world_kernel, dim3(blocks), dim3(256), 0, stream, arena, desc, out);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_scatter_world(const bpd::Arena *arena, uint32_t offset, uint32_t count,
                                                                                const bp_bubble *in, const int32_t *motorIdx,
                                                                                const int32_t *bandIdx, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(bpd::scatter_world_kernel, dim3(blocks), dim3(256), 0, stream, arena, offset, count, in, motorIdx, bandIdx);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Serving: transactions applied on the device, and full records of one world.
// ---------------------------------------------------------------------------------------
namespace bpd {

// Copies bubble slot `src` to slot `dst` of a world (published side and properties).
__device__ void copy_slot(const Arena *__restrict__ A, uint32_t p, uint32_t src, uint32_t dst) {
    A->pos[p][dst] = A->pos[p][src];
    A->start[p][dst] = A->start[p][src];
    A->vel[p][dst] = A->vel[p][src];
    A->rot[p][dst] = A->rot[p][src];
    A->omega[p][dst] = A->omega[p][src];
    A->region[p][dst] = A->region[p][src];
    A->id[dst] = A->id[src];
    A->mass[dst] = A->mass[src]; A->invm[dst] = A->invm[src]; A->radius[dst] = A->radius[src];
    A->Kc[dst] = A->Kc[src]; A->gamma[dst] = A->gamma[src]; A->flags[dst] = A->flags[src];
    A->cat[dst] = A->cat[src]; A->mask[dst] = A->mask[src];
    A->motor[dst] = A->motor[src]; A->band[dst] = A->band[src];
}

// Moves elements [from + 1, n) of one array to [from, n - 1), a workgroup-wide chunk at a time:
// a chunk is read whole before it is written, and no chunk writes what a later chunk reads.
template <class T> __device__ void shift_down(T *base, uint32_t from, uint32_t n) {
    for (uint32_t c = from + 1; c < n; c += blockDim.x) {
        const uint32_t s = c + threadIdx.x;
        const bool ok = s < n;
        T v{};
        if (ok) v = base[s];
        __syncthreads();
        if (ok) base[s - 1] = v;
        __syncthreads();
    }
}

// Removes slot `from` of a world's range [o, o + n): every later bubble moves down one slot.
__device__ void remove_slot(const Arena *__restrict__ A, uint32_t p, uint32_t o, uint32_t from, uint32_t n) {
    const uint32_t f = o + from, e = o + n;
    shift_down(A->pos[p], f, e); shift_down(A->start[p], f, e); shift_down(A->vel[p], f, e);
    shift_down(A->rot[p], f, e); shift_down(A->omega[p], f, e); shift_down(A->region[p], f, e);
    shift_down(A->id, f, e); shift_down(A->mass, f, e); shift_down(A->invm, f, e); shift_down(A->radius, f, e);
    shift_down(A->Kc, f, e); shift_down(A->gamma, f, e); shift_down(A->flags, f, e); shift_down(A->cat, f, e);
    shift_down(A->mask, f, e); shift_down(A->motor, f, e); shift_down(A->band, f, e);
}

// One workgroup per transaction (each names a different world). The arithmetic is FP32, one
// rounding per operation, in the order of the host-side statement of the same transactions:
//   impulse: v_i = float(double(v_i) + double(J) / double(m_i)) for a movable bubble (the
//            reference's FP64 form, TransactionStager.swift:312-324);
//   add:     a copy of bubble i with a new id (largest + 1) at c_i + (2 r_i) * direction;
//   remove:  bubble i leaves;
//   merge:   bubbles a < b (i and its next index, or previous for the last) become one bubble in
//            slot a: mass m_a + m_b, radius sqrt(r_a^2 + r_b^2), centre and velocity the
//            mass-weighted means;
//   split:   bubble i becomes two halves (mass / 2, radius / sqrt 2) at c_i +- r_half * direction;
//            the first keeps the id and slot, the second takes a new id (largest + 1).
// The index is `index % n`. Structural changes keep ids sorted, so no reordering is needed.
// Like a host write of the world's bubbles, a transaction clears the world's failed mark.
__global__ void transaction_kernel(const Arena *__restrict__ A, WorldDesc *descs, const TxRecord *txs) {
    const TxRecord tx = txs[blockIdx.x];
    WorldDesc *d = descs + tx.world;
    const uint32_t n = d->bubbleCount, o = d->bubbleOffset, p = d->pub;
    if (n == 0 || tx.kind == BP_TX_NONE) return;
    // The host refuses growth past the world's slots before it launches; this keeps the kernel
    // inside the range whatever it is given.
    if ((tx.kind == BP_TX_ADD || tx.kind == BP_TX_SPLIT) && n + 1 > d->bubbleCapacity) return;
    const uint32_t i = tx.index % n;
    uint32_t newCount = n;
    if (tx.kind == BP_TX_IMPULSE) {
        if (threadIdx.x == 0) {
            const uint32_t g = o + i;
            const float invm = A->invm[g];
            if (invm > 0.0f) {
                const double m = (double)A->mass[g];
                const float2 v = A->vel[p][g];
                A->vel[p][g] = make_float2((float)((double)v.x + (double)tx.impulse[0] / m),
                                           (float)((double)v.y + (double)tx.impulse[1] / m));
            }
        }
    } else if (tx.kind == BP_TX_ADD) {
        if (threadIdx.x == 0) {
            const uint32_t g = o + i, t = o + n;
            copy_slot(A, p, g, t);
            A->id[t] = A->id[o + n - 1] + 1;
            const float offset = 2.0f * A->radius[g];
            const float2 c = A->pos[p][g];
            const float2 nc = make_float2(c.x + offset * tx.direction[0], c.y + offset * tx.direction[1]);
            A->pos[p][t] = nc; A->start[p][t] = nc; A->region[p][t] = 0u;
            A->motor[t] = -1; A->band[t] = -1;      // the new id has no motor or band record
        }
        newCount = n + 1;
    } else if (tx.kind == BP_TX_REMOVE) {
        if (n > 1) { remove_slot(A, p, o, i, n); newCount = n - 1; }
    } else if (tx.kind == BP_TX_MERGE) {
        if (n > 1) {
            const uint32_t k = i + 1 < n ? i + 1 : i - 1;
            const uint32_t a = min(i, k), b = max(i, k), ga = o + a, gb = o + b;
            __syncthreads();
            if (threadIdx.x == 0) {
                const float ma = A->mass[ga], mb = A->mass[gb];
                const float m = ma + mb;
                const float ra = A->radius[ga], rb = A->radius[gb];
                A->radius[ga] = __builtin_sqrtf(ra * ra + rb * rb);
                A->mass[ga] = m;
                if (m > 0.0f) {                  // a massless pair keeps a's centre and velocity
                    const float2 ca = A->pos[p][ga], cb = A->pos[p][gb], va = A->vel[p][ga], vb = A->vel[p][gb];
                    const float2 c = make_float2((ca.x * ma + cb.x * mb) / m, (ca.y * ma + cb.y * mb) / m);
                    A->invm[ga] = 1.0f / m;
                    A->pos[p][ga] = c;
                    A->vel[p][ga] = make_float2((va.x * ma + vb.x * mb) / m, (va.y * ma + vb.y * mb) / m);
                } else {
                    A->invm[ga] = 0.0f;
                }
                A->start[p][ga] = A->pos[p][ga]; A->region[p][ga] = 0u;
            }
            __syncthreads();
            remove_slot(A, p, o, b, n);
            newCount = n - 1;
        }
    } else if (tx.kind == BP_TX_SPLIT) {
        if (threadIdx.x == 0) {
            const uint32_t g = o + i, t = o + n;
            const float r = A->radius[g] * 0.70710678f;
            const float m = A->mass[g] * 0.5f;
            const float invm = m > 0.0f ? 1.0f / m : 0.0f;
            const float2 c = A->pos[p][g];
            const float dx = tx.direction[0] * r, dy = tx.direction[1] * r;
            copy_slot(A, p, g, t);
            const float2 c1 = make_float2(c.x + dx, c.y + dy), c2 = make_float2(c.x - dx, c.y - dy);
            A->radius[g] = r; A->mass[g] = m; A->invm[g] = invm;
            A->pos[p][g] = c1; A->start[p][g] = c1; A->region[p][g] = 0u;
            A->radius[t] = r; A->mass[t] = m; A->invm[t] = invm;
            A->pos[p][t] = c2; A->start[p][t] = c2; A->region[p][t] = 0u;
            A->id[t] = A->id[o + n - 1] + 1;
            A->motor[t] = -1; A->band[t] = -1;
        }
        newCount = n + 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d->bubbleCount = newCount;
        d->buckets = tx.buckets;
        d->failed = 0;              // new input: a world that failed a step is stepped again
    }
}

// Complete records (state and properties) of one world's published side.
__global__ void gather_full_world_kernel(const Arena *__restrict__ A, const WorldDesc *d, bp_bubble *out) {
    const uint32_t n = d->bubbleCount, o = d->bubbleOffset, p = d->pub;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = o + i;
        const float2 c = A->pos[p][g], s = A->start[p][g], v = A->vel[p][g];
        bp_bubble b;
        b.id = A->id[g];
        b.center[0] = c.x; b.center[1] = c.y;
        b.previous_center[0] = s.x; b.previous_center[1] = s.y;
        b.velocity[0] = v.x; b.velocity[1] = v.y;
        b.rotation = A->rot[p][g];
        b.angular_velocity = A->omega[p][g];
        b.mass = A->mass[g]; b.inverse_mass = A->invm[g]; b.radius = A->radius[g];
        b.contact_stiffness = A->Kc[g]; b.damping = A->gamma[g]; b.flags = A->flags[g];
        b.category = A->cat[g]; b.mask = A->mask[g];
        out[i] = b;
    }
}

}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_transactions(const bpd::Arena *arena, bpd::WorldDesc *descs,
                                                                               const bpd::TxRecord *txs, uint32_t count,
                                                                               hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(bpd::transaction_kernel, dim3(count), dim3(256), 0, stream, arena, descs, txs);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_gather_full_world(const bpd::Arena *arena,
                                                                                    const bpd::WorldDesc *desc, bp_bubble *out,
                                                                                    uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(bpd::gather_full_world_kernel, dim3(blocks), dim3(256), 0, stream, arena, desc, out);
    return hipGetLastError();
}

// Fill of a u32 array by a kernel: graph captures record kernels, not memset nodes.
namespace bpd {
__global__ void fill_u32_kernel(uint32_t *p, size_t words, uint32_t value) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x) p[i] = value;
}
}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_fill_u32(uint32_t *p, size_t words, uint32_t value,
                                                                           hipStream_t stream) {
    if (words == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(bpd::fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, p, words, value);
    return hipGetLastError();
}
