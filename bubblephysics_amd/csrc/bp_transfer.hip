// bp_transfer.hip -- one world's bubbles between host records and the structure-of-arrays arena
// in one copy each way (serving transactions read a world, change it and write it back).
#include "bp_device.h"
#include "../../include/bubblephysics.h"

namespace bpd {

// State of the published side of one world into bp_bubble records: the dynamic fields only (the
// host keeps the properties it wrote), or, with FULL, every field (a world leaving the device,
// whose properties transactions may have changed).
template <bool FULL>
__device__ void gather_range(const Arena *__restrict__ A, const WorldDesc *d, bp_bubble *out, uint32_t first, uint32_t stride) {
    const uint32_t n = d->bubbleCount, o = d->bubbleOffset, p = d->pub;
    for (uint32_t i = first; i < n; i += stride) {
        const uint32_t g = o + i;
        const float2 c = A->pos[p][g], s = A->start[p][g], v = A->vel[p][g];
        bp_bubble &b = out[i];
        b.center[0] = c.x; b.center[1] = c.y;
        b.previous_center[0] = s.x; b.previous_center[1] = s.y;
        b.velocity[0] = v.x; b.velocity[1] = v.y;
        b.rotation = A->rot[p][g];
        b.angular_velocity = A->omega[p][g];
        if (FULL) {
            b.id = A->id[g];
            b.mass = A->mass[g]; b.inverse_mass = A->invm[g]; b.radius = A->radius[g]; b.contact_stiffness = A->Kc[g];
            b.damping = A->gamma[g]; b.flags = A->flags[g]; b.category = A->cat[g]; b.mask = A->mask[g];
        }
    }
}

template <bool FULL>
__global__ void gather_world_kernel(const Arena *__restrict__ A, const WorldDesc *d, bp_bubble *out) {
    gather_range<FULL>(A, d, out, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Every field of several worlds' records, one row of workgroups per world.
__global__ void gather_worlds_full_kernel(const Arena *__restrict__ A, const WorldMove *moves) {
    const WorldMove m = moves[blockIdx.y];
    gather_range<true>(A, m.desc, static_cast<bp_bubble *>(m.records), blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Records (and their motor and band indices) into both sides of the world's range.
__device__ void scatter_range(const Arena *__restrict__ A, uint32_t offset, uint32_t count, const bp_bubble *in,
                              const int32_t *motorIdx, const int32_t *bandIdx, bool clearAcc, uint32_t first, uint32_t stride) {
    for (uint32_t i = first; i < count; i += stride) {
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
        if (clearAcc) A->acc[g] = make_float2(0.f, 0.f);
    }
}

__global__ void scatter_world_kernel(const Arena *__restrict__ A, uint32_t offset, uint32_t count, const bp_bubble *in,
                                     const int32_t *motorIdx, const int32_t *bandIdx) {
    scatter_range(A, offset, count, in, motorIdx, bandIdx, false, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Several worlds' records into their ranges, one row of workgroups per world; the accelerations
// of the range start at zero (a world moved in has no previous force evaluation there).
__global__ void scatter_worlds_kernel(const Arena *__restrict__ A, const WorldMove *moves) {
    const WorldMove m = moves[blockIdx.y];
    scatter_range(A, m.offset, m.count, static_cast<const bp_bubble *>(m.records), m.indices, m.indices + m.count, true,
                  blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Copies of one staged blob to many destinations (a world moved into another world's ranges):
// one workgroup per copy, 4-byte words (every record array has 4-byte elements or wider).
__global__ void multi_copy_kernel(const char *blob, const CopyOp *ops) {
    const CopyOp op = ops[blockIdx.x];
    const uint32_t *src = blob ? reinterpret_cast<const uint32_t *>(blob + op.src) : reinterpret_cast<const uint32_t *>(op.src);
    uint32_t *dst = reinterpret_cast<uint32_t *>(op.dst);
    for (uint32_t i = threadIdx.x; i < op.bytes / 4u; i += blockDim.x) dst[i] = src[i];
}

}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_gather_world(const bpd::Arena *arena, const bpd::WorldDesc *desc,
                                                                               bp_bubble *out, uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(bpd::gather_world_kernel<false>, dim3(blocks), dim3(256), 0, stream, arena, desc, out);
    return hipGetLastError();
}

// Every field of the world's records (count: at least the world's bubble count on the device, the
// grid's size; `out` holds that many).
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_gather_world_full(const bpd::Arena *arena,
                                                                                    const bpd::WorldDesc *desc, bp_bubble *out,
                                                                                    uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + 255) / 256;
    hipLaunchKernelGGL(bpd::gather_world_kernel<true>, dim3(blocks), dim3(256), 0, stream, arena, desc, out);
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

// One workgroup per transaction (each names a different world). Only the transactions that
// need no host staging run here; add, merge and split are staged on the host with the reference's
// placement rules (bp_stage_add, bp_stage_merge, bp_stage_split) and written back before the launch.
//   impulse: applyImpulse (TransactionStager.swift:312-324): only a .dynamic or .motorDriven bubble
//            (force-driven without a band drive) takes it, else incompatibleMotionRole;
//            v = float(double(v) + double(J) / double(m)), refused (invalidValue) when not finite;
//   remove:  bubble i leaves (not the last bubble of a world).
// The index is `index % n`. A world that failed a step takes no transaction: its failure is final.
// The outcome (BP_REJECT_*, 0 = applied) goes to the world's result record.
__global__ void transaction_kernel(const Arena *__restrict__ A, WorldDesc *descs, const TxRecord *txs, WorldStatus *stats) {
    const TxRecord tx = txs[blockIdx.x];
    WorldDesc *d = descs + tx.world;
    const uint32_t n = d->bubbleCount, o = d->bubbleOffset, p = d->pub;
    if (n == 0 || tx.kind == BP_TX_NONE || d->failed == kWorldFailed) return;
    const uint32_t i = tx.index % n;
    uint32_t newCount = n, outcome = BP_REJECT_NONE;
    if (tx.kind == BP_TX_IMPULSE) {
        if (threadIdx.x == 0) {
            const uint32_t g = o + i;
            const uint32_t fl = A->flags[g];
            const bool movable = (fl & BP_BUBBLE_FORCE_DRIVEN) != 0u && (fl & BP_BUBBLE_STATIC) == 0u && A->band[g] < 0;
            if (!movable) {
                outcome = BP_REJECT_INCOMPATIBLE_MOTION_ROLE;
            } else {
                const double m = (double)A->mass[g];
                const float2 v = A->vel[p][g];
                const float2 nv = make_float2((float)((double)v.x + (double)tx.impulse[0] / m),
                                              (float)((double)v.y + (double)tx.impulse[1] / m));
                if (isfinite(nv.x) && isfinite(nv.y)) A->vel[p][g] = nv;
                else outcome = BP_REJECT_INVALID_VALUE;
            }
        }
    } else if (tx.kind == BP_TX_REMOVE) {
        if (n > 1) { remove_slot(A, p, o, i, n); newCount = n - 1; }
    } else if (tx.kind == kTxReplace && d->mirror && tx.index <= d->mirrorCapacity && tx.index <= d->bubbleCapacity) {
        // The host staged an add, merge or split on the mirror: its records replace the world's,
        // on both sides, with their motor and band indices.
        const bp_bubble *in = static_cast<const bp_bubble *>(d->mirror);
        const uint32_t count = tx.index, cap = d->mirrorCapacity;
        for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) {
            const bp_bubble b = in[k];
            const uint32_t g = o + k;
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
            A->motor[g] = d->mirrorIdx[k];
            A->band[g] = d->mirrorIdx[cap + k];
        }
        newCount = count;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d->bubbleCount = newCount;
        d->buckets = tx.buckets;
        stats[tx.world].txResult = outcome;
    }
}

// Deletes the bubbles that a held world's last step removed from the region (region result
// BP_REGION_REMOVED on the published side), keeping the order of the others, and makes the world
// runnable again: the public world deletes such objects before its next frame
// (NativePublicWorld.swift:264-267). One workgroup per world. The range is compacted in place a
// chunk of blockDim bubbles at a time: a chunk is read whole before it is written, and its
// bubbles only move down, below the next chunk.
__global__ void drop_removed_kernel(const Arena *__restrict__ A, WorldDesc *descs, const DropRecord *recs,
                                    int64_t *removedIds) {
    __shared__ uint32_t scanKeep[256], scanDrop[256];
    const DropRecord r = recs[blockIdx.x];
    WorldDesc *d = descs + r.world;
    if (d->failed != kWorldHeld) return;
    const uint32_t n = d->bubbleCount, o = d->bubbleOffset, p = d->pub, t = threadIdx.x;
    uint32_t kept = 0, dropped = 0;
    for (uint32_t c = 0; c < n; c += blockDim.x) {
        const uint32_t i = c + t;
        const bool ok = i < n;
        const uint32_t g = o + i;
        const bool keep = ok && A->region[p][g] != BP_REGION_REMOVED;
        float2 pos{}, start{}, vel{};
        float rot = 0.0f, om = 0.0f, mass = 0.0f, invm = 0.0f, radius = 0.0f, Kc = 0.0f, gamma = 0.0f;
        uint32_t region = 0, flags = 0, cat = 0, mask = 0;
        int32_t motor = -1, band = -1;
        int64_t id = 0;
        if (ok) {
            pos = A->pos[p][g]; start = A->start[p][g]; vel = A->vel[p][g]; rot = A->rot[p][g]; om = A->omega[p][g];
            region = A->region[p][g]; id = A->id[g]; mass = A->mass[g]; invm = A->invm[g]; radius = A->radius[g];
            Kc = A->Kc[g]; gamma = A->gamma[g]; flags = A->flags[g]; cat = A->cat[g]; mask = A->mask[g];
            motor = A->motor[g]; band = A->band[g];
        }
        // Inclusive scans of the keep and drop marks over the chunk (blockDim <= 256).
        scanKeep[t] = keep ? 1u : 0u;
        scanDrop[t] = (ok && !keep) ? 1u : 0u;
        __syncthreads();
        for (uint32_t step = 1; step < blockDim.x; step <<= 1) {
            const uint32_t ak = t >= step ? scanKeep[t - step] : 0u, ad = t >= step ? scanDrop[t - step] : 0u;
            __syncthreads();
            scanKeep[t] += ak; scanDrop[t] += ad;
            __syncthreads();
        }
        if (keep) {
            const uint32_t q = o + kept + scanKeep[t] - 1;
            A->pos[p][q] = pos; A->start[p][q] = start; A->vel[p][q] = vel; A->rot[p][q] = rot; A->omega[p][q] = om;
            A->region[p][q] = region; A->id[q] = id; A->mass[q] = mass; A->invm[q] = invm; A->radius[q] = radius;
            A->Kc[q] = Kc; A->gamma[q] = gamma; A->flags[q] = flags; A->cat[q] = cat; A->mask[q] = mask;
            A->motor[q] = motor; A->band[q] = band;
        } else if (ok && removedIds) {
            removedIds[r.removedOffset + dropped + scanDrop[t] - 1] = id;
        }
        const uint32_t chunkKept = scanKeep[blockDim.x - 1], chunkDropped = scanDrop[blockDim.x - 1];
        __syncthreads();
        kept += chunkKept;
        dropped += chunkDropped;
    }
    if (t == 0) {
        d->bubbleCount = kept;
        d->buckets = r.buckets;
        d->failed = kWorldRunnable;
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
                                                                               bpd::WorldStatus *stats, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(bpd::transaction_kernel, dim3(count), dim3(256), 0, stream, arena, descs, txs, stats);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_drop_removed(const bpd::Arena *arena, bpd::WorldDesc *descs,
                                                                               const bpd::DropRecord *recs, uint32_t count,
                                                                               int64_t *removedIds, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(bpd::drop_removed_kernel, dim3(count), dim3(256), 0, stream, arena, descs, recs, removedIds);
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

// Gate of the launches that follow it on a stream: one lane waits until `target` pooled
// workgroups have started (they count themselves in *placed), or `maxTicks` ticks of the
// 100 MHz wall clock have passed, whichever comes first, so the wait always ends.
namespace bpd {
__global__ void gate_kernel(const uint32_t *placed, uint32_t target, uint32_t maxTicks) {
    if (threadIdx.x != 0) return;
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(placed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target &&
           wall_clock64() - t0 < (uint64_t)maxTicks)
        __builtin_amdgcn_s_sleep(4);
}
}  // namespace bpd

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_gate(const uint32_t *placed, uint32_t target,
                                                                       uint32_t maxTicks, hipStream_t stream) {
    hipLaunchKernelGGL(bpd::gate_kernel, dim3(1), dim3(64), 0, stream, placed, target, maxTicks);
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

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_multi_copy(const char *blob, const bpd::CopyOp *ops,
                                                                             uint32_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(bpd::multi_copy_kernel, dim3(count), dim3(256), 0, stream, blob, ops);
    return hipGetLastError();
}

// Many-world gather and scatter: `moves` (device-readable) lists `count` worlds of at most
// `maxRecords` records each.
extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_gather_worlds_full(const bpd::Arena *arena,
                                                                                     const bpd::WorldMove *moves, uint32_t count,
                                                                                     uint32_t maxRecords, hipStream_t stream) {
    if (count == 0 || maxRecords == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((maxRecords + 255) / 256, 64);
    hipLaunchKernelGGL(bpd::gather_worlds_full_kernel, dim3(blocks, count), dim3(256), 0, stream, arena, moves);
    return hipGetLastError();
}

extern "C" __attribute__((visibility("hidden"))) hipError_t bp_launch_scatter_worlds(const bpd::Arena *arena,
                                                                                 const bpd::WorldMove *moves, uint32_t count,
                                                                                 uint32_t maxRecords, hipStream_t stream) {
    if (count == 0 || maxRecords == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((maxRecords + 255) / 256, 64);
    hipLaunchKernelGGL(bpd::scatter_worlds_kernel, dim3(blocks, count), dim3(256), 0, stream, arena, moves);
    return hipGetLastError();
}
