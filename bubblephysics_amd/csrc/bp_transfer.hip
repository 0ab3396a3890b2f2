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
