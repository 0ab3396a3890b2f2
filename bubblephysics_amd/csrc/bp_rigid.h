// bp_rigid.h -- internal: device layout of the rigid solver's batch (bp_rigid.hip kernels,
// bp_rigid_api.cpp host side).
#ifndef BP_RIGID_H
#define BP_RIGID_H
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bubblephysics.h"

namespace bpr {

// A body-pair candidate of a body's row: NativeCandidate (NativeGeometryKernels.metal-source:27)
// with kind 3, the pair's lower and higher body index and, in place of metadata.w, the sum of the
// iterations' normal corrections. 32 bytes.
struct RigidCand {
    int64_t first, second;
    uint32_t kind, low, high;
    float impulse;
};

// NativeRigidCorrection (NativeGeometryKernels.metal-source:51): position and velocity
// corrections of both ends; the w lanes carry the normal and the point. 64 bytes.
struct RigidCorr {
    float first[4], second[4], vfirst[4], vsecond[4];
};

// NativeRigidWall (NativeGeometryKernels.metal-source:54): the last sweep a body stopped at
// against a static body's edge. 48 bytes.
struct RigidWall {
    int64_t body, owner;
    float geometry[4];
    float extra[4];
};

// One world of the batch: ranges of the packed arrays and the world's controls.
struct RigidDesc {
    uint32_t bodyOffset, bodyCount, vertexOffset, triangleOffset;
    uint32_t segmentOffset, segmentCount, jointOffset, jointCount;
    uint32_t refOffset, partnerOffset, hasJointTables, rowOffset;
    uint32_t stride, k, iterations, prelude;
    float dt, force[2], pad0;
    float region[4];
    uint32_t side, failed, lastOk, pad1;
};

struct RigidArena {
    bp_rigid_body *bodies[2];     // published / working sides, indexed by bodyOffset
    const float2 *verts;
    const uint4 *tris;
    const bp_segment *segs;
    const bp_joint *jointInput;   // the world's joint records as given (state ignored)
    bp_joint *joints;             // the step's records (state = accumulated lambdas)
    RigidCorr *jcorr;
    const uint2 *refs;
    const uint4 *jranges;         // indexed by bodyOffset
    const uint64_t *partners;
    RigidCand *rows;              // indexed by rowOffset + body * stride + slot
    RigidCorr *corr;
    uint32_t *slots;              // used row slots of the step, compacted
    uint32_t *twin;               // per used slot: the slot of the same pair in the other body's row
    float *toi;                   // per used slot: the pair's time of impact
    uint32_t *toiSteps;           // per used slot: slices of the pair's impact-time search (0: none)
    uint32_t *toiStart;           // per used slot: first slice's index in the world's slice list
    uint32_t *toiFirst;           // per used slot: first slice whose pose overlaps
    uint32_t *rowCount;           // indexed by bodyOffset
    uint32_t *rowStart;
    RigidWall *walls;
};

}  // namespace bpr

namespace bpr {
// Worlds of at most this many bodies run one wavefront per world with their bodies in LDS;
// larger worlds run 256 lanes per world with their bodies in global memory.
constexpr uint32_t kRigidLdsBodies = 96;
}

// Launches the worlds listed in order[count]: small = 1 for worlds of at most kRigidLdsBodies bodies.
extern "C" hipError_t bp_launch_rigid_step(const bpr::RigidArena *arena, bpr::RigidDesc *descs, const uint32_t *order,
                                           uint32_t count, int small, bp_rigid_frame_info *infos, uint32_t worlds,
                                           uint32_t steps, hipStream_t stream);

#endif  // BP_RIGID_H
