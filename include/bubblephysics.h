/*
 * bubblephysics.h -- C ABI of the MI355X-native world step of BubblePhysics 4.0.0.
 *
 * This header is the drop-in boundary for the reference's solver seam
 *   NativeMetalFrameExecuting.executeFrame(input: NativeSolverInput, scenarioStep:) -> NativeSolverFrame
 *   (Sources/BubblePhysicsNativeMetal/NativeSolverInput.swift:255-264),
 * implemented in the reference by NativeMetalSolver.executeSnapshot
 *   (Sources/BubblePhysicsNativeMetal/NativeMetalSolver.swift:380-1218).
 *
 * Plain C types only: fixed-width integers, float arrays and opaque handles.
 * The library owns every device allocation; input arrays are borrowed for the
 * duration of a call; output arrays are allocated by the caller.
 *
 * Record layouts are shared by the HIP library (libbubblephysics_hip.so) and
 * by the CPU oracle used in the tests.
 */
#ifndef BUBBLEPHYSICS_H
#define BUBBLEPHYSICS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BP_ABI_VERSION 1u

/* ------------------------------------------------------------------------ */
/* Status codes.                                                              */
/* One code per case of NativeMetalSolverError (NativeMetalSolver.swift:5-7)  */
/* and NativeMetalWorldError (NativeSolverInput.swift:266-272).               */
/* ------------------------------------------------------------------------ */
typedef int32_t bp_status;
enum {
    BP_OK = 0,
    BP_ERR_INVALID_ARGUMENT = 1,       /* malformed call (null handle, size mismatch)        */
    BP_ERR_RESOURCE_UNAVAILABLE = 2,   /* NativeMetalSolverError.resourceUnavailable (no GPU) */
    BP_ERR_ALLOCATION_FAILED = 3,      /* NativeMetalSolverError.allocationFailed             */
    BP_ERR_COMMAND_SETUP_FAILED = 4,   /* NativeMetalSolverError.commandSetupFailed           */
    BP_ERR_COMMAND_FAILED = 5,         /* NativeMetalSolverError.commandFailed                */
    BP_ERR_INVALID_SNAPSHOT = 6,       /* NativeMetalSolverError.invalidSnapshot              */
    BP_ERR_CAPACITY_OVERFLOW = 7,      /* NativeMetalSolverError.capacityOverflow             */
    BP_ERR_FRAME_IN_FLIGHT = 8,        /* NativeMetalWorldError.frameInFlight                 */
    BP_ERR_NON_FINITE = 9,             /* NativeMetalWorldError.nonFinite                     */
    BP_ERR_GUARD_FAILURE = 10,         /* NativeMetalWorldError.guardFailure                  */
    BP_ERR_CONTACT_VIOLATION = 11,     /* NativeMetalWorldError.contactViolation              */
    BP_ERR_SUBSTEP_LIMIT_EXCEEDED = 12,/* NativeMetalWorldError.substepLimitExceeded          */
    BP_ERR_UNAVAILABLE = 13,           /* NativeMetalWorldError.unavailable                   */
    BP_ERR_INVALID_FRAME = 14          /* NativeMetalWorldError.invalidFrame: the step's read-back
                                          record is inconsistent (claims success for an unfinished
                                          step or more than 1024 substeps); nothing is published  */
};

/* Term that fixed h in the last substep: NativeMetalSubstep.Term
 * (NativeForceRecords.swift:104-106). */
enum { BP_TERM_REMAINING = 0, BP_TERM_STIFFNESS = 1, BP_TERM_DISTANCE = 2, BP_TERM_MARGIN = 3, BP_TERM_NONE = 0xffffffffu };

/* Kind of contact violation: NativeMetalContactAssertion.Kind (NativeForceRecords.swift:44-46). */
enum { BP_VIOLATION_NONE = 0, BP_VIOLATION_PAIR = 1, BP_VIOLATION_INSIDE_BODY = 2, BP_VIOLATION_EDGE_SIDE = 3 };

/* Execution path (NativeMetalExecutionPath, NativeMetalBackend.swift:60-65), extended with the
 * MI355X forms of this library. */
enum {
    BP_PATH_WORKGROUP = 0,        /* whole world in one workgroup (reference: singleDispatch)     */
    BP_PATH_COOPERATIVE = 1,      /* world spread over many workgroups (reference: multipleDispatches) */
    BP_PATH_NONE = 2              /* world without bubbles: no substep loop                       */
};

/* ------------------------------------------------------------------------ */
/* Records.                                                                   */
/* ------------------------------------------------------------------------ */

/* bp_bubble.flags: publicFlags.xyz and the force-driven marker contact.w of NativeBubble
 * (NativeGeometryKernels.metal-source:4-17; NativePublicWorld.swift:203-212). */
enum {
    BP_BUBBLE_ENABLED = 1u,       /* publicFlags.x */
    BP_BUBBLE_STATIC = 2u,        /* publicFlags.y: motion role .static            */
    BP_BUBBLE_WRAP = 4u,          /* publicFlags.z: region exit policy .wrap        */
    BP_BUBBLE_FORCE_DRIVEN = 8u   /* contact.w: roles dynamic, motorDriven, playerDriven */
};

/* Region result written by the step (publicFlags.w, NativePostSolveKernels.metal-source:536-566). */
enum { BP_REGION_UNCHANGED = 0, BP_REGION_REMOVED = 1, BP_REGION_WRAPPED = 2 };

/* One bubble: NativeSolverBubble (NativeSolverInput.swift:18-57) with the public properties the
 * step reads (NativePublicProperties, NativePublicWorld.swift:68-94). 72 bytes. */
typedef struct bp_bubble {
    int64_t id;                 /* ObjectID raw value; a world's bubbles are sorted by id        */
    float center[2];            /* position.xy                                                    */
    float previous_center[2];   /* position.zw: output only, centre at the start of the step      */
    float velocity[2];          /* motion.xy                                                      */
    float rotation;             /* motion.z                                                       */
    float angular_velocity;     /* motion.w                                                       */
    float mass;                 /* physical.x                                                     */
    float inverse_mass;         /* physical.y; 0 = static                                         */
    float radius;               /* physical.z (targetRadius)                                      */
    float contact_stiffness;    /* contact.y (K_c); <= 0 means no contact force                   */
    float damping;              /* material.x (gamma, 1/s)                                        */
    uint32_t flags;             /* BP_BUBBLE_*                                                    */
    uint32_t category;          /* filter.x                                                       */
    uint32_t mask;              /* filter.y                                                       */
} bp_bubble;

/* bp_segment.flags (NativeMetalSegment.init, NativeMetalBufferLayout.swift:64-78). */
enum {
    BP_SEGMENT_HAS_OWNER = 1u,    /* flags.y */
    BP_SEGMENT_POLYGON_EDGE = 2u, /* flags.w bit 0: one-sided static polygon edge with regions */
    BP_SEGMENT_CONVEX_START = 4u, /* flags.w bit 1 */
    BP_SEGMENT_CONVEX_END = 8u,   /* flags.w bit 2 (informative) */
    BP_SEGMENT_ALLOWED_SIDE = 16u /* flags.z: allowed_side is set */
};

/* One static segment: NativeSolverSegment (NativeSolverInput.swift:80-114) packed as
 * NativeMetalSegment. A world's segments are sorted by id. 64 bytes. */
typedef struct bp_segment {
    int64_t id;
    int64_t owner_id;
    float a[2];                 /* current.xy */
    float b[2];                 /* current.zw */
    float previous_start[2];    /* region.xy: start of the previous polygon edge */
    float outward_sign;         /* region.z */
    float allowed_side;         /* collision.x */
    uint32_t flags;             /* BP_SEGMENT_* */
    uint32_t category;
    uint32_t mask;
    uint32_t reserved;
} bp_segment;

/* bp_body.shape (NativeRigidBody.shape.x, NativeRigidPolygonState.swift:34-36). */
enum { BP_SHAPE_POLYGON = 0, BP_SHAPE_CIRCLE = 1, BP_SHAPE_CAPSULE = 2 };
/* bp_body.flags. */
enum { BP_BODY_STATIC = 1u, BP_BODY_KINEMATIC = 2u };

/* A rigid body as a bubble obstacle: the fields of NativeRigidBody
 * (NativeRigidPolygonState.swift:24-40) that the bubble step reads. Poses are centre-of-mass
 * poses; vertices are local (centroid-relative) and shared in one array per world.
 * Dynamic bodies never share a world with bubbles (NativeCompleteStepper.swift:71-88) and are
 * rejected. 96 bytes. */
typedef struct bp_body {
    uint64_t id;
    float pose[3];              /* x, y, angle at the end of the last step             */
    float previous[3];          /* output: x, y, angle at the start of the last step   */
    float velocity[3];          /* vx, vy, angular velocity (kinematic drive)          */
    float bounding_radius;      /* physical.z                                          */
    float core_radius;          /* shape.y                                             */
    float half_length;          /* shape.z                                             */
    float max_linear_speed;     /* limits.x, 0 = none                                   */
    float max_angular_speed;    /* limits.y, 0 = none                                   */
    uint32_t shape;             /* BP_SHAPE_*                                           */
    uint32_t flags;             /* BP_BODY_*                                            */
    uint32_t category;
    uint32_t mask;
    uint32_t vertex_offset;     /* ranges.x */
    uint32_t vertex_count;      /* ranges.y */
    uint32_t triangle_offset;   /* ranges.z */
    uint32_t triangle_count;    /* ranges.w */
} bp_body;

/* Motor drive of a .motorDriven bubble: NativeMetalMotor (NativeForceRecords.swift:29-32). */
typedef struct bp_motor {
    int64_t bubble_id;
    float force[2];
    float torque;
    float inertia;
} bp_motor;

/* Rubber band of a .playerDriven bubble: NativeMetalBandDrive (NativeForceRecords.swift:9-12). */
typedef struct bp_band {
    int64_t bubble_id;
    float target[2];
    float stiffness;
    float maximum_force;        /* 0 = no limit */
} bp_band;

/* Force field: NativeForceField (NativeForceKernels.metal-source:11-17), built as in
 * NativeForceInput.appendField (NativeForceRecords.swift:212-239). Fields act in ascending id
 * order on force-driven bubbles inside their range (a circle of `radius`, or a polygon given by
 * `vertex_count` local vertices in the world's field-vertex array). */
enum { BP_FIELD_DIRECTIONAL = 0u, BP_FIELD_RADIAL = 1u, BP_FIELD_VORTEX = 2u };
typedef struct bp_field {
    int64_t id;
    float position[2];          /* pose.xy                                              */
    float rotation;             /* pose.z                                               */
    float falloff;              /* pose.w: exponent of (1 - d/R), 0 = constant          */
    float direction[2];         /* strength.xy: force of a directional field (local)     */
    float strength;             /* strength.z: radial or vortex strength                 */
    float radius;               /* strength.w: circle radius or polygon fade radius      */
    uint32_t kind;              /* BP_FIELD_*                                           */
    uint32_t vertex_offset;     /* shape.y, into the world's field vertices             */
    uint32_t vertex_count;      /* shape.z, 0 = circle                                  */
    uint32_t filter_offset;     /* filter.x, into the world's filter ids (ascending)     */
    uint32_t filter_count;      /* filter.y                                             */
    uint32_t has_filter;        /* filter.z: 1 = acts only on the listed ids            */
} bp_field;

/* Step configuration: NativeSolverConfiguration (NativeSolverInput.swift:116-162) plus the
 * package switches of NativeMetalConfiguration (NativeMetalBackend.swift:2-34). */
typedef struct bp_world_config {
    float time_step;              /* dt                                                    */
    float contact_tolerance;      /* thresholds.z (0.001)                                  */
    float separation_tolerance;   /* thresholds.w (0.002)                                  */
    float region[4];              /* min.x, min.y, max.x, max.y                            */
    float global_force[2];        /* NativeMetalForceControl.global                        */
    uint32_t apply_region;        /* 1 = public path: region removal/wrap at step end     */
    uint32_t candidate_limit;     /* k per object, 1..128 (default 128)                    */
    uint32_t substep_split;       /* substepSplitMultiplier, >= 1                          */
    uint32_t minimum_margin_only; /* 1 = margin r/4 only                                   */
    uint32_t grid_bucket_count;   /* hashed-grid buckets, 1..2048 (default 2048)           */
    uint32_t assert_every_substep;/* 1 = contact assertion after each substep              */
    uint32_t injection_enabled;   /* injectedSubstepState present                          */
    uint32_t injection_after_substep;
    int64_t injection_object;
    float injection_position[2];
    uint32_t force_cooperative;   /* 1 = multi-workgroup path even for small worlds        */
    uint32_t export_adjacency;    /* 1 = keep the rows of the last grid build readable after a
                                     step (adjacency reads, contact energy of the audit); a plain
                                     bp_world always keeps them                               */
} bp_world_config;

/* Details of a failure: NativeMetalFailureDetail (NativeSolverInput.swift:280-290). */
typedef struct bp_failure_detail {
    uint64_t object_ids[2];
    int32_t violation_kind;       /* BP_VIOLATION_*                                  */
    int32_t edge_segment_id_low;  /* low 32 bits of the segment id for edge side, else -1 */
    float distance;               /* NaN when absent                                  */
    uint32_t limiting_term;       /* BP_TERM_*                                         */
} bp_failure_detail;

/* Per-step result: telemetry of NativeMetalFrameTelemetry (NativeMetalBackend.swift:72-122)
 * and the overflow counters of NativeSolverFrame (NativeSolverInput.swift:247-250). */
typedef struct bp_frame_info {
    bp_status status;
    uint32_t substep_count;          /* substepControl.counts.y                         */
    uint32_t grid_rebuild_count;     /* rebuilds forced by the margin term              */
    uint32_t candidate_pair_count;   /* contactControl.counts.x of the last grid build  */
    uint32_t candidate_overflow;     /* dropped candidates, max over grid builds        */
    uint32_t max_candidates_per_object;
    uint32_t execution_path;         /* BP_PATH_*                                       */
    uint32_t group_width;            /* threads of the workgroup that ran the loop      */
    uint32_t region_removals;
    uint32_t region_wraps;
    float gpu_milliseconds;
    float simulated_time;            /* substepControl.time.y at the end                */
    uint64_t limiter_object;         /* substepControl.limiter.x                        */
    uint32_t limiter_term;           /* substepControl.limiter.y                        */
    uint32_t adjacency_entries;      /* entries of the adjacency of the last grid build */
    uint32_t kernel_class;           /* kernel configuration that ran the step (bp_batch classes) */
    uint32_t device_ticks;           /* device wall-clock ticks (100 MHz) of the world's step, batched kernels */
    bp_failure_detail detail;
} bp_frame_info;

/* ------------------------------------------------------------------------ */
/* Library.                                                                   */
/* ------------------------------------------------------------------------ */
uint32_t bp_abi_version(void);
/* sizeof of bp_bubble, bp_segment, bp_body, bp_motor, bp_band, bp_world_config, bp_frame_info,
 * bp_batch_world_desc, bp_field, in that order (the role of NativeMetalShaderABITests.swift:61-381). */
void bp_abi_record_sizes(uint32_t *out9);
/* Fills a configuration with the reference defaults (NativeSolverConfiguration.init,
 * NativeSolverInput.swift:128-143; NativeMetalConfiguration, NativeMetalBackend.swift:2-34). */
void bp_world_config_default(bp_world_config *config);
const char *bp_status_string(bp_status status);
/* Last error text of the calling thread (device error strings, argument checks). */
const char *bp_last_error(void);

/* ------------------------------------------------------------------------ */
/* Plain world: one world per handle (NativeMetalSolver, NativeMetalSolver.swift:255-378).     */
/* ------------------------------------------------------------------------ */
typedef struct bp_world bp_world;

bp_status bp_world_create(const bp_world_config *config, bp_world **out_world);
void bp_world_destroy(bp_world *world);
bp_status bp_world_set_config(bp_world *world, const bp_world_config *config);
/* Full snapshot of the bubbles (replaces the published state; ids sorted ascending). */
bp_status bp_world_set_bubbles(bp_world *world, const bp_bubble *bubbles, uint32_t count);
/* Number of the last published snapshot: it changes with every full upload, every deletion by
 * bp_world_remove_bubbles and every successful step (a failed step publishes nothing). */
uint64_t bp_world_publication(const bp_world *world);
/* Delta upload: bubbleChanges of NativeMetalSnapshot (NativeMetalBufferLayout.swift:209-289),
 * taken against the snapshot `base_publication` names. A delta against any other snapshot than the
 * last published one returns BP_ERR_INVALID_SNAPSHOT and changes nothing (the previous snapshot
 * must be the last published one, NativeMetalSolver.swift:404-419). */
bp_status bp_world_update_bubbles(bp_world *world, uint64_t base_publication, const uint32_t *indices,
                                  const bp_bubble *bubbles, uint32_t count);
bp_status bp_world_set_segments(bp_world *world, const bp_segment *segments, uint32_t count);
bp_status bp_world_set_bodies(bp_world *world, const bp_body *bodies, uint32_t count,
                              const float *vertices, uint32_t vertex_count,
                              const uint32_t *triangles, uint32_t triangle_count);
bp_status bp_world_set_forces(bp_world *world, const bp_motor *motors, uint32_t motor_count,
                              const bp_band *bands, uint32_t band_count);
/* Force fields with their polygon vertices (float2 pairs) and filter ids (NativeForceInput
 * fields, fieldVertices and fieldObjectIDs, NativeForceRecords.swift:177-183). */
bp_status bp_world_set_fields(bp_world *world, const bp_field *fields, uint32_t field_count, const float *vertices,
                              uint32_t vertex_count, const int64_t *filter_ids, uint32_t filter_count);
/* One step of dt (executeFrame). On any error the published state is unchanged. */
bp_status bp_world_step(bp_world *world, int32_t scenario_step, bp_frame_info *out_info);
bp_status bp_world_read_bubbles(bp_world *world, bp_bubble *out, uint32_t count);
/* Region results of the last published step: NativeSolverFrame.regionRemovals and .regionWraps
 * (NativeSolverInput.swift:243-244; nativeApplyPublicRegion, NativePostSolveKernels.metal-source:536-566),
 * bubble ids ascending. As in the reference's frame, removed bubbles are still in the published
 * state; the public world deletes them before its next frame (NativePublicWorld.swift:264-267),
 * here with bp_world_remove_bubbles. *out_removed / *out_wrapped receive the counts; an array
 * smaller than its count returns BP_ERR_CAPACITY_OVERFLOW (null arrays ask for counts only). */
bp_status bp_world_region_results(bp_world *world, int64_t *removed, uint32_t removed_capacity, uint32_t *out_removed,
                                  int64_t *wrapped, uint32_t wrapped_capacity, uint32_t *out_wrapped);
/* Deletes bubbles by id from the published state, with their motor and band records: the next
 * snapshot of a public world that dropped them (NativePublicWorld.swift:264-267). An id the world
 * does not hold returns BP_ERR_INVALID_SNAPSHOT and changes nothing. */
bp_status bp_world_remove_bubbles(bp_world *world, const int64_t *ids, uint32_t count);
bp_status bp_world_read_bodies(bp_world *world, bp_body *out, uint32_t count);
uint32_t bp_world_bubble_count(const bp_world *world);
/* Adjacency of the last grid build (test seam, like NativeMetalSolver.candidateKeys):
 * row_offsets[count+1], entries packed as (kind << 30) | neighbour index. */
bp_status bp_world_read_adjacency(bp_world *world, uint32_t *row_offsets, uint32_t offsets_count,
                                  uint32_t *entries, uint32_t entries_capacity, uint32_t *out_entry_count);
/* h of each substep of the last step (stepHistory), up to 1024 values. */
bp_status bp_world_read_step_history(bp_world *world, float *out, uint32_t capacity, uint32_t *out_count);
/* FP64 audit of a world's published state: momentum, angular momentum and kinetic energy of
 * the movable bubbles, and the elastic energy of the contacts in the rows of the last grid build,
 * U = K r (eps / (1 - eps) + ln(1 - eps)) per deflected bubble (ContactLawReference.swift:26-31;
 * a pair of different materials splits its overlap as r_a eps(F/K_a) + r_b eps(F/K_b) = overlap,
 * :68-82). Sums are FP64 on the device. No reference counterpart: the reference audits energy and
 * momentum only in its tests (NativeMetalSubstepKernelTests.swift:957-978). */
typedef struct bp_audit {
    double momentum[2];         /* sum of m v                                                   */
    double angular_momentum;    /* sum of m (x cross v), about the origin                        */
    double kinetic_energy;      /* sum of m |v|^2 / 2                                            */
    double contact_energy;      /* NaN when the world's rows are not kept (export_adjacency = 0) */
    uint32_t contacts;          /* contacts with a deflection in contact_energy                  */
    uint32_t bubbles;
} bp_audit;
bp_status bp_world_audit(bp_world *world, bp_audit *out);

/* End-of-step contact records of a world's published state: NativeSolverContact
 * (NativeSolverInput.swift:188-200) as the reference fills it from its contact cache
 * (NativeMetalSolver.swift:1041-1060). Built from the candidate list of the last grid build at the
 * final centres (nativeBuildContacts, NativeGeometryKernels.metal-source:683-728), kept iff
 * penetration > -separation tolerance (NativeContactKernels.metal-source:146-160), in candidate
 * order: every bubble's pairs (the lower index first, partner ascending), then every bubble's
 * segments; `impulse` is the contact law's force at the record's geometry times dt^2
 * (nativeContactImpulses, NativeForceKernels.metal-source:261-280). The world's rows must be kept
 * (export_adjacency, always on for bp_world) and its last step must have succeeded. 64 bytes. */
enum { BP_CONTACT_PAIR = 0, BP_CONTACT_SEGMENT = 1, BP_CONTACT_BODY = 2, BP_CONTACT_RIGID = 3 };
typedef struct bp_contact {
    uint64_t id;                /* FNV-1a of (kind, first, second), NativeMetalSolver.swift:1047-1052 */
    int64_t first;              /* bubble A id                                                     */
    int64_t second;             /* bubble B id, the segment id, or the body id                     */
    uint32_t kind;              /* BP_CONTACT_*                                                    */
    uint32_t feature;           /* polygon edge: 1 edge interior, 2 convex start vertex; else 0   */
    float normal[2];
    float point[2];             /* pointQ                                                          */
    float penetration;          /* body records: radius - distance                                */
    float allowed_side;         /* 0: the segment has no allowed side (the reference's nil)        */
    float impulse;              /* accumulatedCompression: F dt^2                                  */
    uint32_t index_b;           /* bubble B index, the segment index, or the body index            */
} bp_contact;
/* Writes up to `capacity` records; *out_count receives the number of records the step has
 * (more than `capacity` means the list was cut, as the reference cuts at its contact capacity). */
bp_status bp_world_contacts(bp_world *world, bp_contact *out, uint32_t capacity, uint32_t *out_count);

/* End-of-step bubble-body contact records (nativeBubbleBodyContacts,
 * NativeForceKernels.metal-source:333-358, read back as the rigid records of
 * NativeRigidPolygonState.swift:390-402): per bubble-body candidate of the last grid build (the body
 * entries of the bubbles' rows), the body's obstacle at its end-of-step pose against the bubble's
 * final centre; kept iff distance < radius (and a non-zero normal). kind BP_CONTACT_BODY, first = the
 * bubble id, second = the body id, normal from the bubble to the body, point = centre - obstacle
 * normal * distance on the body's surface, penetration = radius - distance, impulse = the wall law's
 * force at that distance times dt^2, index_b = the body index. Ordered body by body, bubbles
 * ascending, as the reference's body rows. Rows of a bubble that overflowed its row capacity hold
 * only the entries the cut kept, where the reference's body rows keep every candidate. Writes up to
 * `capacity` records; *out_count receives the number the step has. */
bp_status bp_world_body_contacts(bp_world *world, bp_contact *out, uint32_t capacity, uint32_t *out_count);

/* bp_world_contacts without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Its working memory is a fixed array of about 3.8 MiB in the library, one per
 * GPU, held until the process ends: worlds of at most 32,768 bubbles and at most 65,536 records per
 * call (larger return BP_ERR_CAPACITY_OVERFLOW before any device work). Every call on a GPU shares
 * that array, so calls that overlap on one GPU race. The world must already be on the device (a
 * step since its last upload), else BP_ERR_UNAVAILABLE. Results equal bp_world_contacts. */
bp_status bp_world_contacts_noalloc(bp_world *world, bp_contact *out, uint32_t capacity, uint32_t *out_count);

/* Contour and render preparation of a world's published state: the draw-only contour the
 * reference builds at the end of every frame (NativeMetalSolver.swift:884-903). Per bubble, rays at
 * `count` angles are cut by every segment the bubble's filter accepts and by the flat chord each
 * overlapping pair partner leaves at r (1 - eps) from the centre, eps the contact law's deflection
 * of this bubble (nativeAccumulateContactPressure and nativeContourRayLimit,
 * NativePostSolveKernels.metal-source:268-337); the deficits are smoothed by 13 Jacobi passes that
 * never go below the cut (nativeGenerateContours, :341-386), closed by the smallest slope-bounded
 * majorant from the deepest sample, and turned into points, bounds, the largest deficit and the
 * effective radius (nativePrepareRenderData, :388-441). count = min(stride, max(64, ceil(4 pi r /
 * max_segment_length))). Pair partners are the overlapping pair entries of the bubble's row of the
 * last grid build (the end-of-step contact records). The contour never feeds back into the step. */
typedef struct bp_contour_config {
    float max_segment_length;   /* maxContourSegmentLength (default 8)                        */
    float surface_tension;      /* contourSurfaceTension (default 20)                          */
    float position_tolerance;   /* positionTolerance (default 1e-4): the rays' clearance       */
    uint32_t reserved;
} bp_contour_config;
typedef struct bp_bubble_contour {
    uint32_t count;             /* points of the bubble's contour                              */
    float maximum;              /* deformation.x: the largest deficit (fraction of the radius) */
    float effective_radius;     /* deformation.y: the nearest contour point's distance          */
    float reserved;
    float bounds[4];            /* min.x, min.y, max.x, max.y of the contour and the centre    */
} bp_bubble_contour;
/* The contour stride of a world's published bubbles: max over bubbles of
 * max(64, ceil(4 pi r / max_segment_length)) (NativeMetalSolver.swift:407-411). */
bp_status bp_world_contour_stride(bp_world *world, const bp_contour_config *config, uint32_t *out_stride);
/* points holds 2 * stride floats per bubble (bubble i's samples from 2 * i * stride); stride must be
 * at least the world's contour stride and at most 2048; shapes holds one record per bubble. The
 * world's rows must be kept (export_adjacency) and its last step must have succeeded. */
bp_status bp_world_contours(bp_world *world, const bp_contour_config *config, uint32_t stride, float *points,
                            bp_bubble_contour *shapes, uint32_t bubble_count);

/* Contact transitions of one step: the contactBegan / contactEnded events of
 * NativePublicWorld.swift:340-391 derived from the step's contact records (bp_world_contacts) and
 * the previous step's set of contact pairs. Keys are (bubble A, bubble B) or (bubble A, segment
 * owner), ordered by raw ObjectID; only records with penetration > 0 or impulse > 0 whose key has an
 * object of interest count. Began events (key order) carry the penetration-weighted mean point,
 * the normalised normal sum from the key's first object to its second and the impulse F dt;
 * ended events follow in key order. `history` holds pairs as (first, second) u64 couples: the
 * previous set on input (history_count pairs), the current set on output (*out_history_count; more
 * than history_capacity returns BP_ERR_CAPACITY_OVERFLOW). Host-only: no device work. Bubble-body
 * records (BP_CONTACT_BODY, bp_world_body_contacts) enter the sums as the reference's rigid records
 * (NativePublicWorld.swift:377-380): key (bubble, body), weight 1, their normal, impulse / dt, and no
 * penetration test; pass them after the pair and segment records, the reference's order. Rigid
 * solver records (BP_CONTACT_RIGID, bp_rigid_batch_contacts) enter the same way. */
enum { BP_EVENT_CONTACT_BEGAN = 1, BP_EVENT_CONTACT_ENDED = 2 };
typedef struct bp_contact_event {
    uint32_t kind;              /* BP_EVENT_CONTACT_*                                             */
    uint32_t reserved;
    uint64_t first, second;     /* ObjectID raw values, first <= second                          */
    float point[2];             /* began: penetration-weighted mean point                         */
    float normal[2];            /* began: unit normal from first to second (0 when the sum is 0)  */
    float impulse;              /* began: sum of F dt (record impulse / dt)                       */
    float reserved2;
} bp_contact_event;
bp_status bp_contact_transitions(const bp_contact *contacts, uint32_t count, const bp_segment *segments,
                                 uint32_t segment_count, const uint64_t *interested, uint32_t interested_count,
                                 float step_duration, uint64_t *history, uint32_t history_count,
                                 uint32_t history_capacity, uint32_t *out_history_count,
                                 bp_contact_event *events, uint32_t event_capacity, uint32_t *out_event_count);

/* Structural staging of merge and split on the host: TransactionStager.merge / .split
 * (TransactionStager.swift:630-735) with StructuralOperations (StructuralOperations.swift:10-151)
 * and the placement rules of PlacementGeometry (PlacementGeometry.swift:42-159). Host-only: the
 * caller passes the world's published bubbles (sorted by id), its bodies at their current poses
 * with their local vertices, and the simulation region, and receives the world's new bubble list
 * (sorted by id) to write back. A refused transaction returns BP_ERR_INVALID_ARGUMENT with the
 * reason in `why` (TransactionRejectionReason); the world is unchanged. */
enum {
    BP_PLACEMENT_OK = 0,
    BP_PLACEMENT_OUTSIDE_REGION = 1,       /* PlacementIssue.outsideRegion    */
    BP_PLACEMENT_INSIDE_RIGID_BODY = 2,    /* PlacementIssue.insideRigidBody  */
    BP_PLACEMENT_COINCIDENT_CENTER = 3     /* PlacementIssue.coincidentCenter */
};
enum {
    BP_REJECT_NONE = 0,
    BP_REJECT_INVALID_STRUCTURAL_OPERATION = 1,
    BP_REJECT_INVALID_PLACEMENT = 2,       /* issue and object_id say where and for which result */
    BP_REJECT_DUPLICATE_ID = 3,
    BP_REJECT_UNKNOWN_OBJECT = 4,
    BP_REJECT_INCOMPATIBLE_MOTION_ROLE = 5, /* incompatibleMotionRole: an impulse on a bubble that is
                                               neither .dynamic nor .motorDriven                */
    BP_REJECT_INVALID_VALUE = 6            /* invalidValue: a non-finite impulse or velocity after it */
};
typedef struct bp_stage_rejection {
    uint32_t reason;            /* BP_REJECT_*    */
    uint32_t issue;             /* BP_PLACEMENT_* */
    int64_t object_id;
} bp_stage_rejection;
typedef struct bp_stage_scene {
    const bp_bubble *bubbles;   /* sorted by id */
    uint32_t bubble_count;
    uint32_t body_count;
    const bp_body *bodies;      /* pose = the body's current position and rotation */
    const float *vertices;      /* local polygon vertices (x, y), indexed by vertex_offset */
    uint32_t vertex_count;
    float region[4];            /* simulationRegion: min.x, min.y, max.x, max.y (inclusive) */
} bp_stage_scene;
/* Merge: `sources` (two or more distinct ids) become `result` (its id, filter, material and flags;
 * its mass when result_mass_given, else the sources' total). Centre: the mean contact point of
 * the source pairs, or the first candidate placement that passes. */
bp_status bp_stage_merge(const bp_stage_scene *scene, const int64_t *sources, uint32_t source_count,
                         const bp_bubble *result, int32_t result_mass_given, bp_bubble *out,
                         uint32_t out_capacity, uint32_t *out_count, bp_stage_rejection *why);
/* Add: TransactionStager.add (TransactionStager.swift:406-421): refused when the id is taken
 * (BP_REJECT_DUPLICATE_ID) or when the bubble's centre fails the placement rules: outside the
 * region, inside a rigid body it collides with, or on the centre of a bubble it collides with
 * (BP_REJECT_INVALID_PLACEMENT with the issue; PlacementGeometry.swift:42-94). */
bp_status bp_stage_add(const bp_stage_scene *scene, const bp_bubble *bubble, bp_bubble *out, uint32_t out_capacity,
                       uint32_t *out_count, bp_stage_rejection *why);
/* Split: `source` becomes `first` and `second` with area shares summing to 1 (within 1e-6); the
 * parts' masses and radii follow the shares. */
bp_status bp_stage_split(const bp_stage_scene *scene, int64_t source, const bp_bubble *first, float first_share,
                         const bp_bubble *second, float second_share, bp_bubble *out, uint32_t out_capacity,
                         uint32_t *out_count, bp_stage_rejection *why);

/* Test switch mirroring the injected NativeMetalCommandExecuting of the reference
 * (NativeMetalSolverLifecycleTests.swift:292-308): the next step fails with BP_ERR_COMMAND_FAILED
 * before touching the device. */
bp_status bp_world_debug_fail_next_command(bp_world *world);

/* ------------------------------------------------------------------------ */
/* Host geometry preparation.                                                */
/* ------------------------------------------------------------------------ */
/* Static polygon -> one-sided solver segments with edge regions
 * (NativeStaticPolygonEdges.segments, NativeStaticPolygonEdges.swift:16-37). Polygons of zero
 * area give plain two-sided segments. Writes *out_count segments (ids first_id, first_id+1, ...). */
bp_status bp_static_polygon_edges(const float *points, uint32_t count, int64_t first_id, int64_t owner_id, uint32_t has_owner,
                                  bp_segment *out, uint32_t capacity, uint32_t *out_count);
/* Polygon rigid shape -> centroid, centroid-relative counter-clockwise vertices, ear-clipped
 * triangles (a, b, c, 0) and bounding radius (NativeRigidGeometry.init,
 * NativeRigidPolygonState.swift:89-133). out_vertices holds 2*count floats, out_triangles 4*count. */
bp_status bp_polygon_body_geometry(const float *local_points, uint32_t count, float *out_centroid, float *out_vertices,
                                   uint32_t *out_vertex_count, uint32_t *out_triangles, uint32_t *out_triangle_count,
                                   float *out_radius);

/* Random sequential placement of non-overlapping discs in a rectangle (host spawning helper,
 * in the role of PlacementGeometry, Sources/BubblePhysicsRuntime/PlacementGeometry.swift:42-50).
 * region = (min.x, min.y, max.x, max.y); obstacles are (x, y, r) triples to keep clear of; discs
 * that find no place after `attempts` tries get NaN centres. */
bp_status bp_place_discs(const float *radii, uint32_t count, const float *region, float inset, float gap,
                         const float *obstacles, uint32_t obstacle_count, uint64_t seed, uint32_t attempts,
                         float *out_xy, uint32_t *out_placed);

/* ------------------------------------------------------------------------ */
/* Batch: many independent worlds advanced by one launch per step.             */
/* No reference counterpart; built on the same kernels as bp_world.            */
/* ------------------------------------------------------------------------ */
typedef struct bp_batch bp_batch;

typedef struct bp_batch_world_desc {
    bp_world_config config;
    const bp_bubble *bubbles;   uint32_t bubble_count;
    const bp_segment *segments; uint32_t segment_count;
    const bp_body *bodies;      uint32_t body_count;
    const float *vertices;      uint32_t vertex_count;     /* float2 pairs */
    const uint32_t *triangles;  uint32_t triangle_count;   /* uint4 records (a, b, c, 0) */
    const bp_motor *motors;     uint32_t motor_count;
    const bp_band *bands;       uint32_t band_count;
    const bp_field *fields;     uint32_t field_count;
    const float *field_vertices; uint32_t field_vertex_count;   /* float2 pairs */
    const int64_t *field_filter_ids; uint32_t field_filter_count;
} bp_batch_world_desc;

bp_status bp_batch_create(bp_batch **out_batch);
void bp_batch_destroy(bp_batch *batch);
/* Adds a world; returns its index. Worlds are packed on the next step. */
bp_status bp_batch_add_world(bp_batch *batch, const bp_batch_world_desc *desc, uint32_t *out_index);
uint32_t bp_batch_world_count(const bp_batch *batch);
/* Bubbles of one world (transactions applied by a bp_server may have changed it). */
uint32_t bp_batch_bubble_count(const bp_batch *batch, uint32_t world);
/* Advances every world by `steps` steps; infos[world_count * steps] (step-major) may be null.
 * A failed world stops at its failing step and keeps its last published state. Bubbles a step
 * removes from the region leave their world before its next step, as the public world deletes
 * them (NativePublicWorld.swift:264-267). */
bp_status bp_batch_step(bp_batch *batch, uint32_t steps, bp_frame_info *infos);
/* Region results of a world of a batch: the ids removed during the last stepping call (already
 * deleted from the world) and the ids wrapped in its last published step, ascending; arrays as
 * in bp_world_region_results. */
bp_status bp_batch_region_results(bp_batch *batch, uint32_t world, int64_t *removed, uint32_t removed_capacity,
                                  uint32_t *out_removed, int64_t *wrapped, uint32_t wrapped_capacity, uint32_t *out_wrapped);
bp_status bp_batch_read_bubbles(bp_batch *batch, uint32_t world, bp_bubble *out, uint32_t count);
/* Published bubbles of every world in one read: offsets[world_count + 1] receives each world's
 * first record (a null `out` asks for the offsets only). */
bp_status bp_batch_read_all_bubbles(bp_batch *batch, bp_bubble *out, uint32_t capacity, uint32_t *offsets,
                                    uint32_t offsets_count);
/* Bodies of one world at their published poses (count = the world's body count). */
bp_status bp_batch_read_bodies(bp_batch *batch, uint32_t world, bp_body *out, uint32_t count);
/* Captures `steps` steps of every resident world (the launches of every kernel class on the
 * batch's streams, forked from and joined back to one stream) into a hipGraph, replayed by
 * bp_batch_replay with one graph launch. A change of the batch's layout (new worlds, residency,
 * a world moving to another kernel class) drops the graph; replay then returns
 * BP_ERR_INVALID_ARGUMENT until the next capture. */
bp_status bp_batch_capture(bp_batch *batch, uint32_t steps);
/* One replay of the captured steps; infos[world_count * steps] (step-major) may be null. */
bp_status bp_batch_replay(bp_batch *batch, bp_frame_info *infos);
/* Device time of the last bp_batch_step (HIP events around the launches). */
float bp_batch_last_gpu_milliseconds(const bp_batch *batch);
/* Kernel configurations ("classes") a world can run in: LDS-resident configurations for small
 * worlds and the global-state kernel. For each class of the last bp_batch_step: device time of its
 * launches (HIP events on its stream), worlds in it, threads per workgroup and bubble capacity.
 * Arrays hold bp_batch_class_count() entries; any may be null. */
uint32_t bp_batch_class_count(void);
bp_status bp_batch_last_class_timing(const bp_batch *batch, float *milliseconds, uint32_t *worlds, uint32_t *threads,
                                     uint32_t *capacity);
/* FP64 audit (bp_world_audit) of every world: out[world_count]. */
bp_status bp_batch_audit(bp_batch *batch, bp_audit *out, uint32_t count);
/* End-of-step contact records (bp_world_contacts) of one world of a batch. */
bp_status bp_batch_contacts(bp_batch *batch, uint32_t world, bp_contact *out, uint32_t capacity, uint32_t *out_count);
/* End-of-step bubble-body contact records (bp_world_body_contacts) of one world of a batch. */
bp_status bp_batch_body_contacts(bp_batch *batch, uint32_t world, bp_contact *out, uint32_t capacity, uint32_t *out_count);
/* Contours (bp_world_contours) of one world of a batch. */
bp_status bp_batch_contours(bp_batch *batch, uint32_t world, const bp_contour_config *config, uint32_t stride, float *points,
                            bp_bubble_contour *shapes, uint32_t bubble_count);
/* Autotuner: for each bin of world sizes (bin b: 2^b to 2^(b+1) - 1 bubbles), times one step of
 * the bin's worlds in every kernel configuration that holds them all (the LDS-resident
 * configurations and the one-workgroup global-state kernel: workgroup shape and the LDS-to-HBM
 * cutover), then the best configuration with half and twice the automatic grid bucket count.
 * Every trial starts from the batch's published state, which the call leaves unchanged; the
 * batch adopts the choices (the next step repacks). Results do not depend on the choices.
 * No reference counterpart (its single-dispatch path has a fixed width, NativeMetalBackend.swift:17-21). */
typedef struct bp_tune_bin {
    uint32_t min_bubbles, max_bubbles;
    uint32_t worlds;               /* worlds of the batch in the bin                              */
    int32_t best_class;            /* chosen kernel configuration, -1 = none                       */
    int32_t best_bucket_shift;     /* grid buckets: -1 half, 0 automatic, +1 twice the automatic count */
    float best_rate;               /* bubble-substeps/s of the bin's trial step with the choices   */
    float class_rate[12];          /* per kernel configuration; 0 = not tried (does not hold the bin) */
    float bucket_rate[2];          /* best configuration with half and with twice the buckets      */
} bp_tune_bin;
bp_status bp_batch_autotune(bp_batch *batch, uint32_t repeats, bp_tune_bin *out, uint32_t capacity, uint32_t *out_count);
/* Applies a table of bp_batch_autotune (for instance one cached for this GPU) without trials. */
bp_status bp_batch_apply_tuning(bp_batch *batch, const bp_tune_bin *bins, uint32_t count);
/* Pins a world to a class (-1 = automatic); used by the autotuner. */
bp_status bp_batch_set_world_class(bp_batch *batch, uint32_t world, int32_t cls);
/* Workgroup pools of bp_batch_step (on by default). No reference counterpart. Once a step of every
 * world has measured their device times, each LDS kernel class is launched with a share of the
 * CUs in proportion to its measured work instead of one workgroup per world: its workgroups take
 * the class's worlds longest first, the next one when they finish one. Results are the same bit
 * for bit either way; only the placement of worlds on CUs changes. 0 turns pools off. */
bp_status bp_batch_set_pooling(bp_batch *batch, uint32_t enabled);

/* ------------------------------------------------------------------------ */
/* Serving: continuous batching over resident worlds. No reference counterpart. */
/* ------------------------------------------------------------------------ */
/* Advances only the listed worlds (distinct, resident) by `steps` steps; infos[steps * count]
 * (step-major, list order) may be null. */
bp_status bp_batch_step_worlds(bp_batch *batch, const uint32_t *worlds, uint32_t count, uint32_t steps,
                               bp_frame_info *infos);
/* Non-blocking form: queues the launches and returns. bp_batch_poll returns BP_ERR_FRAME_IN_FLIGHT
 * while they run, then BP_OK with infos filled as by bp_batch_step_worlds. One submission at a
 * time; other batch calls return BP_ERR_FRAME_IN_FLIGHT until it is collected. */
bp_status bp_batch_submit_worlds(bp_batch *batch, const uint32_t *worlds, uint32_t count, uint32_t steps);
bp_status bp_batch_poll(bp_batch *batch, bp_frame_info *infos);
/* Replaces a world's bubbles (ids strictly increasing; the count may change): a transaction
 * applied between steps. Within the slots the world was packed with it is written in place. */
bp_status bp_batch_write_bubbles(bp_batch *batch, uint32_t world, const bp_bubble *bubbles, uint32_t count);
/* Residency: a non-resident world lives in host memory only and takes no device memory after the
 * next step repacks the batch; it cannot be stepped until it is resident again. */
bp_status bp_batch_set_resident(bp_batch *batch, uint32_t world, uint32_t resident);
/* Device bytes a world holds (or would hold) while resident. */
bp_status bp_batch_world_footprint(const bp_batch *batch, uint32_t world, uint64_t *bytes);

/* ------------------------------------------------------------------------ */
/* Server: a native request loop over the worlds of a batch. No reference     */
/* counterpart: the reference steps one world per call from its game loop     */
/* (BubbleWorld.submit / advance, BubbleWorld.swift:252-283).                  */
/* A dispatcher thread takes step requests as they arrive and launches, per   */
/* kernel class, the waiting worlds that are not already on the device;       */
/* launches of different classes and consecutive launches of one class run    */
/* concurrently on several streams. Every world keeps its state on the device; */
/* worlds are admitted against a device-memory budget and the least recently  */
/* used idle worlds are evicted to host memory when an admission needs room.  */
/* ------------------------------------------------------------------------ */
typedef struct bp_server bp_server;

/* Transaction carried by a request, applied before the step with the reference's rules
 * (TransactionStager). The bubble it names is `index % bubble count`; new ids are the largest id + 1.
 * A refused transaction leaves the world unchanged, the step still runs, and the result's
 * transaction_result gives the reason (BP_REJECT_*), as the public world reports a
 * transactionRejected event and advances. */
enum {
    BP_TX_NONE = 0,
    BP_TX_IMPULSE = 1,  /* applyImpulse (TransactionStager.swift:312-324), on the device:
                           v = float(double(v) + double(J) / double(m)) for a .dynamic or .motorDriven
                           bubble; other roles: BP_REJECT_INCOMPATIBLE_MOTION_ROLE; a non-finite impulse
                           or result: BP_REJECT_INVALID_VALUE                                           */
    BP_TX_ADD = 2,      /* add (:406-413) of a bubble with bubble i's properties at c_i + 2 r_i * direction,
                           at rest, staged by bp_stage_add (placement rules)                           */
    BP_TX_REMOVE = 3,   /* remove (:429-436) of bubble i (not the last bubble of a world), on the device  */
    BP_TX_MERGE = 4,    /* mergeBubbles (:630-674) of bubble i and its index neighbour (i + 1, or i - 1
                           for the last) into a bubble with the lower id and the first source's
                           properties, staged by bp_stage_merge                                         */
    BP_TX_SPLIT = 5     /* splitBubble (:675-735) of bubble i into halves (shares 1/2): the first keeps
                           the id, the second takes a new one; staged by bp_stage_split                */
};

typedef struct bp_request {
    double arrival;             /* bp_server_run_trace: seconds after the start of the trace  */
    uint64_t tag;               /* caller's id, returned with the result                       */
    uint32_t world;
    uint32_t transaction;       /* BP_TX_*                                                     */
    uint32_t index;             /* bubble of the transaction: index % bubble count             */
    float impulse[2];           /* BP_TX_IMPULSE                                               */
    float direction[2];         /* BP_TX_ADD, BP_TX_SPLIT: unit vector                         */
    uint32_t reserved;
} bp_request;

typedef struct bp_server_result {
    uint64_t tag;
    uint32_t world;
    bp_status status;           /* status of the step; BP_ERR_CAPACITY_OVERFLOW when a transaction
                                   needs more bubble slots than the world was packed with (no step);
                                   BP_ERR_UNAVAILABLE for a world that failed an earlier step: a
                                   world failure is final (BubbleWorld.swift:296-299, 424-430), so
                                   such a request runs neither its transaction nor a step          */
    uint32_t substeps;
    uint32_t bubbles;           /* bubbles of the world in the step                            */
    double arrival;             /* seconds on the server clock (bp_server_clock)               */
    double completed;           /* seconds on the server clock when the result reached the host */
    uint32_t launch_worlds;     /* worlds in the launch that carried the step                  */
    uint32_t kernel_class;
    float device_milliseconds;  /* time the world's workgroup spent on the step                */
    uint32_t transaction_result;/* BP_REJECT_* of the request's transaction (0: applied or none) */
} bp_server_result;

typedef struct bp_server_config {
    uint64_t memory_budget;         /* device bytes of resident worlds; 0 = no limit           */
    uint32_t streams;               /* HIP streams in all (default 32, at least 3): one for reads,
                                       one for host staging of add, merge and split, the rest
                                       for launches (one in flight on each, the batch's own
                                       streams first); the process should have as many hardware
                                       queues (GPU_MAX_HW_QUEUES)                               */
    uint32_t max_inflight_per_class;/* launches of one kernel class in flight at once (default 32) */
    uint32_t max_worlds_per_launch; /* default 1024                                             */
    uint32_t big_world;             /* worlds above this many bubbles run in the one-workgroup
                                       global-state kernel (default 4096)                       */
} bp_server_config;

typedef struct bp_server_stats {
    uint64_t submitted, completed, launches, launched_worlds, transactions, rejected;
    uint64_t evictions, admissions, repacks;
    uint64_t resident_bytes;
    uint64_t admission_microseconds;  /* dispatcher time spent moving worlds in and out of residency */
} bp_server_stats;

/* sizeof of bp_request, bp_server_result, bp_server_config, bp_server_stats, in that order. */
void bp_server_record_sizes(uint32_t *out4);
void bp_server_config_default(bp_server_config *config);
/* Starts a server over every world of `batch`. While it lives the batch's own entry points
 * return BP_ERR_FRAME_IN_FLIGHT. */
bp_status bp_server_create(bp_batch *batch, const bp_server_config *config, bp_server **out_server);
/* Finishes the launches in flight (requests not yet launched are dropped), stops the
 * dispatcher and hands the batch back with every world's newest state. */
void bp_server_destroy(bp_server *server);
/* Queues one step request (non-blocking, thread-safe); its arrival time is the call time. */
bp_status bp_server_submit(bp_server *server, const bp_request *request);
/* Results completed since the last poll (at most `capacity`). */
bp_status bp_server_poll(bp_server *server, bp_server_result *out, uint32_t capacity, uint32_t *out_count);
/* Blocks until every submitted request has a result. */
bp_status bp_server_wait_idle(bp_server *server);
/* Replays a trace: submits trace[i] at its arrival time (seconds after the call) and returns
 * when all have completed; results[i] belongs to trace[i] (and is not returned by poll). */
bp_status bp_server_run_trace(bp_server *server, const bp_request *trace, uint32_t count, bp_server_result *results);
/* A world's bubbles (state and properties) while the server is idle (after wait_idle). */
bp_status bp_server_read_bubbles(bp_server *server, uint32_t world, bp_bubble *out, uint32_t capacity, uint32_t *out_count);
bp_status bp_server_get_stats(bp_server *server, bp_server_stats *out);
/* Seconds since the server started (the clock of bp_server_result). */
double bp_server_clock(const bp_server *server);

/* ------------------------------------------------------------------------ */
/* Rigid solver for worlds without bubbles.                                   */
/* The rigid-body step the reference runs once per frame for every world with */
/* rigid bodies (NativeMetalSolver.encodeStepPrefix, NativeMetalSolver.swift: */
/* 766-799, then nativeRigidFinish at the end gate, :917-921): prediction with */
/* the swept test against static edges, body-pair candidate rows, body-pair   */
/* CCD, 8 iterations of contacts, joints and the wall guard, and the region   */
/* rule. In a world with bubbles its bodies are static or kinematic and the   */
/* bubble step above covers them; these entry points are for the worlds whose */
/* bodies move under contacts and joints. One workgroup per world.            */
/* ------------------------------------------------------------------------ */

/* One rigid body: NativeRigidBody (NativeRigidPolygonState.swift:24-41;
 * NativeGeometryKernels.metal-source:35-41), field for field. 160 bytes. A world's bodies are in
 * ascending id order. Vertices are centroid-relative (bp_polygon_body_geometry); a circle has one
 * vertex (0, 0), a capsule two (-half_length, 0) and (half_length, 0). */
typedef struct bp_rigid_body {
    uint64_t id;                /* identity.x: ObjectID                                                 */
    float mass;                 /* identity.y (unused by the reference's kernels): the body's mass, read
                                   only by the motion prelude of bp_rigid_config                        */
    uint32_t reserved;
    float pose[4];              /* centre of mass x, y, angle; w: angle at the start of the last step    */
    float previous[4];          /* centre of mass at the start of the last step x, y; local centroid x, y */
    float motion[4];            /* vx, vy, angular velocity, angular damping (1/s; 0 for kinematic)      */
    float physical[4];          /* inverse mass, inverse moment (0 = static or kinematic), bounding radius, friction */
    uint32_t ranges[4];         /* vertex offset, vertex count, triangle offset, triangle count          */
    uint32_t flags[4];          /* x static, y wrap at the region, z region result (BP_REGION_*), w restitution (float bits) */
    uint32_t filter[4];         /* category, mask, kinematic (1), layer                                  */
    float shape[4];             /* kind (BP_SHAPE_*), core radius, half length, 0                          */
    float limits[4];            /* maximum linear and angular speed (0 = none), CCD scratch, 0           */
} bp_rigid_body;

/* One joint: NativeJointRecord (NativeJointState.swift:6-14). Ends are body indices of the world;
 * BP_JOINT_WORLD_END is a world point. 96 bytes. */
#define BP_JOINT_WORLD_END 0xffffffffu
enum { BP_JOINT_DISTANCE = 0, BP_JOINT_HINGE = 1, BP_JOINT_WELD = 2 };
enum { BP_JOINT_HAS_LIMITS = 1u, BP_JOINT_HAS_MOTOR = 2u };
typedef struct bp_joint {
    uint64_t id;                /* JointID                                                               */
    uint64_t reserved;
    uint32_t ends[4];           /* end 0, end 1, kind (BP_JOINT_*), flags (BP_JOINT_HAS_*)               */
    float anchors[4];           /* anchor of end 0 and end 1: body-local (geometry origin) or world point */
    float parameters[4];        /* length, softness (compliance), damping, reference angle               */
    float limits[4];            /* lower and upper angle, motor target angular velocity, maximum torque   */
    float state[4];             /* accumulated lambda of the step: distance, motor (zeroed every step)    */
} bp_joint;

/* A joint as the public model states it (Joint; SceneBuilders.distanceJoint / hingeJoint /
 * weldJoint), the input of bp_rigid_joint_tables. */
typedef struct bp_joint_desc {
    uint64_t id;
    uint64_t body_a;
    uint64_t body_b;            /* ignored unless has_body_b                                             */
    uint32_t has_body_b;        /* 0: end b is the world point anchor_b                                  */
    uint32_t kind;              /* BP_JOINT_*                                                            */
    float anchor_a[2];
    float anchor_b[2];
    float length, softness, damping;   /* distance                                                      */
    float reference_angle;             /* hinge and weld                                                */
    uint32_t flags;                    /* hinge: BP_JOINT_HAS_LIMITS, BP_JOINT_HAS_MOTOR                */
    uint32_t collide_connected;
    float lower, upper;                /* hinge limits                                                  */
    float motor_speed, max_torque;     /* hinge motor                                                   */
} bp_joint_desc;

/* Joint tables of one world (NativeJointState.prepare, NativeJointState.swift:73-141, for a world
 * without bubbles): records sorted by JointID with a world end moved to end 0; per body its
 * joint references (joint index, side) and its sorted, distinct non-colliding partners (ids).
 * Joints naming an unknown body are skipped, as the reference skips them.
 * ranges[4 * body_count] = (reference start, reference count, partner start, partner count);
 * references holds 2 * joint_count pairs at most, partners 2 * joint_count ids at most. */
bp_status bp_rigid_joint_tables(const bp_joint_desc *joints, uint32_t joint_count, const uint64_t *body_ids,
                                uint32_t body_count, bp_joint *records, uint32_t *out_record_count,
                                uint32_t *references, uint32_t *out_reference_count, uint32_t *ranges,
                                uint64_t *partners, uint32_t *out_partner_count);

typedef struct bp_rigid_config {
    float time_step;            /* dt                                                                     */
    float region[4];            /* simulation region: min.x, min.y, max.x, max.y                          */
    uint32_t candidate_limit;   /* k per body, 1..128 (default 128)                                       */
    uint32_t iterations;        /* contact iterations per step (NativeMetalSolver.rigidSolverIterationCount = 8) */
    uint32_t motion_prelude;    /* 1: before each step, every dynamic body (not static, not kinematic) takes
                                   the runtime's velocity update of MotionControllers.prepare
                                   (MotionControllers.swift:57-77, role .dynamic, no fields):
                                   v = v * exp(-damping dt), then v += global_force / mass * dt; damping is
                                   motion[3]. It lets one launch carry a world with a global force through
                                   several steps, as the public world's frames would.                    */
    float global_force[2];      /* WorldConfiguration.globalForce (a force, not an acceleration)          */
} bp_rigid_config;

/* Cause of BP_ERR_GUARD_FAILURE (NativeMetalGuardCause, NativeSolverInput.swift:292-311). */
enum { BP_GUARD_NONE = 0, BP_GUARD_RIGID_ROTATION_BUDGET = 1 };

typedef struct bp_rigid_frame_info {
    bp_status status;           /* BP_OK, BP_ERR_NON_FINITE or BP_ERR_GUARD_FAILURE (reference precedence) */
    uint32_t guard_cause;       /* BP_GUARD_*                                                             */
    uint32_t candidate_pairs;   /* body-pair entries over all rows                                        */
    uint32_t candidate_overflow;/* candidates dropped by full rows                                        */
    uint32_t max_candidates_per_object;
    uint32_t removed;           /* bodies the region rule removed (flags[2] == BP_REGION_REMOVED)         */
    uint32_t wrapped;
    uint32_t reserved;
} bp_rigid_frame_info;

typedef struct bp_rigid_batch bp_rigid_batch;

/* One world of a rigid batch. Triangles are uint4 records (a, b, c, 0) with vertex numbers
 * local to their body; segments are the static polygon edges (bp_static_polygon_edges with the
 * static body as owner), which the prediction's swept test reads. Joint arrays are the output of
 * bp_rigid_joint_tables (all null / zero for a world without joints). */
typedef struct bp_rigid_world_desc {
    bp_rigid_config config;
    const bp_rigid_body *bodies;  uint32_t body_count;
    const float *vertices;        uint32_t vertex_count;    /* float2 pairs */
    const uint32_t *triangles;    uint32_t triangle_count;  /* uint4 records */
    const bp_segment *segments;   uint32_t segment_count;
    const bp_joint *joints;       uint32_t joint_count;
    const uint32_t *joint_references; uint32_t joint_reference_count;   /* (joint, side) pairs */
    const uint32_t *joint_ranges;                                       /* uint4 per body      */
    const uint64_t *joint_partners; uint32_t joint_partner_count;
} bp_rigid_world_desc;

bp_status bp_rigid_batch_create(bp_rigid_batch **out_batch);
void bp_rigid_batch_destroy(bp_rigid_batch *batch);
/* Adds a world (validated: sorted ids, ranges inside the arrays, a triangle's corners inside its
 * body, known shapes, finite records); returns its index. Worlds are packed on the next step. */
bp_status bp_rigid_batch_add_world(bp_rigid_batch *batch, const bp_rigid_world_desc *desc, uint32_t *out_index);
uint32_t bp_rigid_batch_world_count(const bp_rigid_batch *batch);
/* Replaces a world (bodies, geometry, segments, joints, configuration): the next frame's input, as
 * the public world re-stages its bodies every frame (NativeRigidPolygonState.prepare,
 * NativeRigidPolygonState.swift:193-301); for instance after dropping the bodies the region rule
 * removed. A failed world steps again after new input. */
bp_status bp_rigid_batch_set_world(bp_rigid_batch *batch, uint32_t world, const bp_rigid_world_desc *desc);
/* Advances every world by `steps` steps in one launch (one workgroup per world, all steps inside
 * the launch); infos[world_count * steps] (step-major) may be null. A world whose step fails keeps
 * its last published bodies and is not stepped again until new input. */
bp_status bp_rigid_batch_step(bp_rigid_batch *batch, uint32_t steps, bp_rigid_frame_info *infos);
bp_status bp_rigid_batch_read_bodies(bp_rigid_batch *batch, uint32_t world, bp_rigid_body *out, uint32_t count);
/* Joint records of a world after its last step (state = the step's accumulated lambdas). */
bp_status bp_rigid_batch_read_joints(bp_rigid_batch *batch, uint32_t world, bp_joint *out, uint32_t count);
/* Rigid contact records of a world's last published step (NativeRigidPolygonState.contactRecords,
 * NativeRigidPolygonState.swift:367-428): each body pair once, from the row of its lower body
 * index, with the last iteration's normal (from first to second) and point and the sum of the
 * iterations' normal corrections; then the prediction's stopped sweeps against static bodies that
 * the rows did not report. kind BP_CONTACT_RIGID, first / second = body ids, index_b = the second
 * body's index, penetration 0. Writes up to `capacity`; *out_count receives the number of records. */
bp_status bp_rigid_batch_contacts(bp_rigid_batch *batch, uint32_t world, bp_contact *out, uint32_t capacity,
                                  uint32_t *out_count);
/* Device time of the last bp_rigid_batch_step (HIP events around the launch). */
float bp_rigid_batch_last_gpu_milliseconds(const bp_rigid_batch *batch);

#ifdef __cplusplus
}
#endif
#endif /* BUBBLEPHYSICS_H */
