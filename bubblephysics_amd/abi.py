"""ctypes mirror of include/bubblephysics.h and the loader of the HIP library.

The library is built in-tree (``bubblephysics_amd/lib/libbubblephysics_hip.so``) by
``__graft_entry__.build()``. There is no CPU fallback: a missing library raises
:class:`LibraryMissing`, and a machine without a GPU gets ``BP_ERR_RESOURCE_UNAVAILABLE``
from the create calls.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "libbubblephysics_hip.so"

# Status codes (bp_status).
BP_OK = 0
BP_ERR_INVALID_ARGUMENT = 1
BP_ERR_RESOURCE_UNAVAILABLE = 2
BP_ERR_ALLOCATION_FAILED = 3
BP_ERR_COMMAND_SETUP_FAILED = 4
BP_ERR_COMMAND_FAILED = 5
BP_ERR_INVALID_SNAPSHOT = 6
BP_ERR_CAPACITY_OVERFLOW = 7
BP_ERR_FRAME_IN_FLIGHT = 8
BP_ERR_NON_FINITE = 9
BP_ERR_GUARD_FAILURE = 10
BP_ERR_CONTACT_VIOLATION = 11
BP_ERR_SUBSTEP_LIMIT_EXCEEDED = 12
BP_ERR_UNAVAILABLE = 13
BP_ERR_INVALID_FRAME = 14

STATUS_NAMES = {
    0: "ok", 1: "invalidArgument", 2: "resourceUnavailable", 3: "allocationFailed", 4: "commandSetupFailed",
    5: "commandFailed", 6: "invalidSnapshot", 7: "capacityOverflow", 8: "frameInFlight", 9: "nonFinite",
    10: "guardFailure", 11: "contactViolation", 12: "substepLimitExceeded", 13: "unavailable",
}

BP_TERM_REMAINING, BP_TERM_STIFFNESS, BP_TERM_DISTANCE, BP_TERM_MARGIN = 0, 1, 2, 3
BP_TERM_NONE = 0xFFFFFFFF
BP_VIOLATION_NONE, BP_VIOLATION_PAIR, BP_VIOLATION_INSIDE_BODY, BP_VIOLATION_EDGE_SIDE = 0, 1, 2, 3

BP_BUBBLE_ENABLED, BP_BUBBLE_STATIC, BP_BUBBLE_WRAP, BP_BUBBLE_FORCE_DRIVEN = 1, 2, 4, 8
BP_REGION_UNCHANGED, BP_REGION_REMOVED, BP_REGION_WRAPPED = 0, 1, 2
BP_SEGMENT_HAS_OWNER, BP_SEGMENT_POLYGON_EDGE, BP_SEGMENT_CONVEX_START, BP_SEGMENT_CONVEX_END = 1, 2, 4, 8
BP_SEGMENT_ALLOWED_SIDE = 16
BP_SHAPE_POLYGON, BP_SHAPE_CIRCLE, BP_SHAPE_CAPSULE = 0, 1, 2
BP_BODY_STATIC, BP_BODY_KINEMATIC = 1, 2
BP_FIELD_DIRECTIONAL, BP_FIELD_RADIAL, BP_FIELD_VORTEX = 0, 1, 2
BP_PATH_WORKGROUP, BP_PATH_COOPERATIVE, BP_PATH_NONE = 0, 1, 2


class Bubble(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int64), ("center", ctypes.c_float * 2), ("previous_center", ctypes.c_float * 2),
        ("velocity", ctypes.c_float * 2), ("rotation", ctypes.c_float), ("angular_velocity", ctypes.c_float),
        ("mass", ctypes.c_float), ("inverse_mass", ctypes.c_float), ("radius", ctypes.c_float),
        ("contact_stiffness", ctypes.c_float), ("damping", ctypes.c_float), ("flags", ctypes.c_uint32),
        ("category", ctypes.c_uint32), ("mask", ctypes.c_uint32),
    ]


class Segment(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int64), ("owner_id", ctypes.c_int64), ("a", ctypes.c_float * 2), ("b", ctypes.c_float * 2),
        ("previous_start", ctypes.c_float * 2), ("outward_sign", ctypes.c_float), ("allowed_side", ctypes.c_float),
        ("flags", ctypes.c_uint32), ("category", ctypes.c_uint32), ("mask", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
    ]


class Body(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64), ("pose", ctypes.c_float * 3), ("previous", ctypes.c_float * 3),
        ("velocity", ctypes.c_float * 3), ("bounding_radius", ctypes.c_float), ("core_radius", ctypes.c_float),
        ("half_length", ctypes.c_float), ("max_linear_speed", ctypes.c_float), ("max_angular_speed", ctypes.c_float),
        ("shape", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("category", ctypes.c_uint32), ("mask", ctypes.c_uint32),
        ("vertex_offset", ctypes.c_uint32), ("vertex_count", ctypes.c_uint32), ("triangle_offset", ctypes.c_uint32),
        ("triangle_count", ctypes.c_uint32),
    ]


class Motor(ctypes.Structure):
    _fields_ = [("bubble_id", ctypes.c_int64), ("force", ctypes.c_float * 2), ("torque", ctypes.c_float),
                ("inertia", ctypes.c_float)]


class Band(ctypes.Structure):
    _fields_ = [("bubble_id", ctypes.c_int64), ("target", ctypes.c_float * 2), ("stiffness", ctypes.c_float),
                ("maximum_force", ctypes.c_float)]


class Field(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("position", ctypes.c_float * 2), ("rotation", ctypes.c_float),
                ("falloff", ctypes.c_float), ("direction", ctypes.c_float * 2), ("strength", ctypes.c_float),
                ("radius", ctypes.c_float), ("kind", ctypes.c_uint32), ("vertex_offset", ctypes.c_uint32),
                ("vertex_count", ctypes.c_uint32), ("filter_offset", ctypes.c_uint32), ("filter_count", ctypes.c_uint32),
                ("has_filter", ctypes.c_uint32)]


class WorldConfig(ctypes.Structure):
    _fields_ = [
        ("time_step", ctypes.c_float), ("contact_tolerance", ctypes.c_float), ("separation_tolerance", ctypes.c_float),
        ("region", ctypes.c_float * 4), ("global_force", ctypes.c_float * 2), ("apply_region", ctypes.c_uint32),
        ("candidate_limit", ctypes.c_uint32), ("substep_split", ctypes.c_uint32), ("minimum_margin_only", ctypes.c_uint32),
        ("grid_bucket_count", ctypes.c_uint32), ("assert_every_substep", ctypes.c_uint32),
        ("injection_enabled", ctypes.c_uint32), ("injection_after_substep", ctypes.c_uint32),
        ("injection_object", ctypes.c_int64), ("injection_position", ctypes.c_float * 2),
        ("force_cooperative", ctypes.c_uint32), ("export_adjacency", ctypes.c_uint32),
    ]


class FailureDetail(ctypes.Structure):
    _fields_ = [("object_ids", ctypes.c_uint64 * 2), ("violation_kind", ctypes.c_int32),
                ("edge_segment_id_low", ctypes.c_int32), ("distance", ctypes.c_float), ("limiting_term", ctypes.c_uint32)]


class FrameInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("substep_count", ctypes.c_uint32), ("grid_rebuild_count", ctypes.c_uint32),
        ("candidate_pair_count", ctypes.c_uint32), ("candidate_overflow", ctypes.c_uint32),
        ("max_candidates_per_object", ctypes.c_uint32), ("execution_path", ctypes.c_uint32),
        ("group_width", ctypes.c_uint32), ("region_removals", ctypes.c_uint32), ("region_wraps", ctypes.c_uint32),
        ("gpu_milliseconds", ctypes.c_float), ("simulated_time", ctypes.c_float), ("limiter_object", ctypes.c_uint64),
        ("limiter_term", ctypes.c_uint32), ("adjacency_entries", ctypes.c_uint32), ("kernel_class", ctypes.c_uint32),
        ("device_ticks", ctypes.c_uint32), ("detail", FailureDetail),
    ]


class BatchWorldDesc(ctypes.Structure):
    _fields_ = [
        ("config", WorldConfig),
        ("bubbles", ctypes.c_void_p), ("bubble_count", ctypes.c_uint32),
        ("segments", ctypes.c_void_p), ("segment_count", ctypes.c_uint32),
        ("bodies", ctypes.c_void_p), ("body_count", ctypes.c_uint32),
        ("vertices", ctypes.c_void_p), ("vertex_count", ctypes.c_uint32),
        ("triangles", ctypes.c_void_p), ("triangle_count", ctypes.c_uint32),
        ("motors", ctypes.c_void_p), ("motor_count", ctypes.c_uint32),
        ("bands", ctypes.c_void_p), ("band_count", ctypes.c_uint32),
        ("fields", ctypes.c_void_p), ("field_count", ctypes.c_uint32),
        ("field_vertices", ctypes.c_void_p), ("field_vertex_count", ctypes.c_uint32),
        ("field_filter_ids", ctypes.c_void_p), ("field_filter_count", ctypes.c_uint32),
    ]


class TuneBin(ctypes.Structure):
    _fields_ = [
        ("min_bubbles", ctypes.c_uint32), ("max_bubbles", ctypes.c_uint32), ("worlds", ctypes.c_uint32),
        ("best_class", ctypes.c_int32), ("best_bucket_shift", ctypes.c_int32), ("best_rate", ctypes.c_float),
        ("class_rate", ctypes.c_float * 12), ("bucket_rate", ctypes.c_float * 2),
    ]


class Audit(ctypes.Structure):
    _fields_ = [
        ("momentum", ctypes.c_double * 2), ("angular_momentum", ctypes.c_double), ("kinetic_energy", ctypes.c_double),
        ("contact_energy", ctypes.c_double), ("contacts", ctypes.c_uint32), ("bubbles", ctypes.c_uint32),
    ]


# Serving (bp_server).
BP_TX_NONE, BP_TX_IMPULSE, BP_TX_ADD, BP_TX_REMOVE, BP_TX_MERGE, BP_TX_SPLIT = 0, 1, 2, 3, 4, 5


class Request(ctypes.Structure):
    _fields_ = [
        ("arrival", ctypes.c_double), ("tag", ctypes.c_uint64), ("world", ctypes.c_uint32),
        ("transaction", ctypes.c_uint32), ("index", ctypes.c_uint32), ("impulse", ctypes.c_float * 2),
        ("direction", ctypes.c_float * 2), ("reserved", ctypes.c_uint32),
    ]


class ServerResult(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_uint64), ("world", ctypes.c_uint32), ("status", ctypes.c_int32), ("substeps", ctypes.c_uint32),
        ("bubbles", ctypes.c_uint32), ("arrival", ctypes.c_double), ("completed", ctypes.c_double),
        ("launch_worlds", ctypes.c_uint32), ("kernel_class", ctypes.c_uint32), ("device_milliseconds", ctypes.c_float),
        ("transaction_result", ctypes.c_uint32),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("memory_budget", ctypes.c_uint64), ("streams", ctypes.c_uint32), ("max_inflight_per_class", ctypes.c_uint32),
        ("max_worlds_per_launch", ctypes.c_uint32), ("big_world", ctypes.c_uint32),
    ]


class Contact(ctypes.Structure):          # bp_contact
    _fields_ = [("id", ctypes.c_uint64), ("first", ctypes.c_int64), ("second", ctypes.c_int64), ("kind", ctypes.c_uint32),
                ("feature", ctypes.c_uint32), ("normal", ctypes.c_float * 2), ("point", ctypes.c_float * 2),
                ("penetration", ctypes.c_float), ("allowed_side", ctypes.c_float), ("impulse", ctypes.c_float),
                ("index_b", ctypes.c_uint32)]


class ContactEvent(ctypes.Structure):     # bp_contact_event
    _fields_ = [("kind", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("first", ctypes.c_uint64),
                ("second", ctypes.c_uint64), ("point", ctypes.c_float * 2), ("normal", ctypes.c_float * 2),
                ("impulse", ctypes.c_float), ("reserved2", ctypes.c_float)]


class ServerStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "submitted", "completed", "launches", "launched_worlds", "transactions", "rejected", "evictions", "admissions",
        "repacks", "resident_bytes", "admission_microseconds")]


class StageRejection(ctypes.Structure):
    _fields_ = [("reason", ctypes.c_uint32), ("issue", ctypes.c_uint32), ("object_id", ctypes.c_int64)]


class StageScene(ctypes.Structure):
    _fields_ = [("bubbles", ctypes.c_void_p), ("bubble_count", ctypes.c_uint32), ("body_count", ctypes.c_uint32),
                ("bodies", ctypes.c_void_p), ("vertices", ctypes.c_void_p), ("vertex_count", ctypes.c_uint32),
                ("region", ctypes.c_float * 4)]



# Rigid solver for worlds without bubbles (bp_rigid_*).
class RigidBody(ctypes.Structure):      # bp_rigid_body, NativeRigidBody field for field
    _fields_ = [("id", ctypes.c_uint64), ("mass", ctypes.c_float), ("reserved", ctypes.c_uint32),
                ("pose", ctypes.c_float * 4), ("previous", ctypes.c_float * 4), ("motion", ctypes.c_float * 4),
                ("physical", ctypes.c_float * 4), ("ranges", ctypes.c_uint32 * 4), ("flags", ctypes.c_uint32 * 4),
                ("filter", ctypes.c_uint32 * 4), ("shape", ctypes.c_float * 4), ("limits", ctypes.c_float * 4)]


class Joint(ctypes.Structure):          # bp_joint, NativeJointRecord
    _fields_ = [("id", ctypes.c_uint64), ("reserved", ctypes.c_uint64), ("ends", ctypes.c_uint32 * 4),
                ("anchors", ctypes.c_float * 4), ("parameters", ctypes.c_float * 4), ("limits", ctypes.c_float * 4),
                ("state", ctypes.c_float * 4)]


class JointDesc(ctypes.Structure):      # bp_joint_desc
    _fields_ = [("id", ctypes.c_uint64), ("body_a", ctypes.c_uint64), ("body_b", ctypes.c_uint64),
                ("has_body_b", ctypes.c_uint32), ("kind", ctypes.c_uint32), ("anchor_a", ctypes.c_float * 2),
                ("anchor_b", ctypes.c_float * 2), ("length", ctypes.c_float), ("softness", ctypes.c_float),
                ("damping", ctypes.c_float), ("reference_angle", ctypes.c_float), ("flags", ctypes.c_uint32),
                ("collide_connected", ctypes.c_uint32), ("lower", ctypes.c_float), ("upper", ctypes.c_float),
                ("motor_speed", ctypes.c_float), ("max_torque", ctypes.c_float)]


class RigidConfig(ctypes.Structure):    # bp_rigid_config
    _fields_ = [("time_step", ctypes.c_float), ("region", ctypes.c_float * 4), ("candidate_limit", ctypes.c_uint32),
                ("iterations", ctypes.c_uint32), ("motion_prelude", ctypes.c_uint32), ("global_force", ctypes.c_float * 2)]


class RigidFrameInfo(ctypes.Structure):  # bp_rigid_frame_info
    _fields_ = [("status", ctypes.c_int32), ("guard_cause", ctypes.c_uint32), ("candidate_pairs", ctypes.c_uint32),
                ("candidate_overflow", ctypes.c_uint32), ("max_candidates_per_object", ctypes.c_uint32),
                ("removed", ctypes.c_uint32), ("wrapped", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class RigidWorldDesc(ctypes.Structure):  # bp_rigid_world_desc
    _fields_ = [("config", RigidConfig), ("bodies", ctypes.c_void_p), ("body_count", ctypes.c_uint32),
                ("vertices", ctypes.c_void_p), ("vertex_count", ctypes.c_uint32),
                ("triangles", ctypes.c_void_p), ("triangle_count", ctypes.c_uint32),
                ("segments", ctypes.c_void_p), ("segment_count", ctypes.c_uint32),
                ("joints", ctypes.c_void_p), ("joint_count", ctypes.c_uint32),
                ("joint_references", ctypes.c_void_p), ("joint_reference_count", ctypes.c_uint32),
                ("joint_ranges", ctypes.c_void_p),
                ("joint_partners", ctypes.c_void_p), ("joint_partner_count", ctypes.c_uint32)]



class ContourConfig(ctypes.Structure):    # bp_contour_config
    _fields_ = [("max_segment_length", ctypes.c_float), ("surface_tension", ctypes.c_float),
                ("position_tolerance", ctypes.c_float), ("reserved", ctypes.c_uint32)]


class BubbleContour(ctypes.Structure):    # bp_bubble_contour
    _fields_ = [("count", ctypes.c_uint32), ("maximum", ctypes.c_float), ("effective_radius", ctypes.c_float),
                ("reserved", ctypes.c_float), ("bounds", ctypes.c_float * 4)]


def contour_config(max_segment_length=8.0, surface_tension=20.0, position_tolerance=1e-4) -> ContourConfig:
    """NativeSolverConfiguration's contour defaults (NativeSolverInput.swift:128-132)."""
    c = ContourConfig()
    c.max_segment_length, c.surface_tension, c.position_tolerance = max_segment_length, surface_tension, position_tolerance
    return c

BP_JOINT_WORLD_END = 0xFFFFFFFF
BP_JOINT_DISTANCE, BP_JOINT_HINGE, BP_JOINT_WELD = 0, 1, 2
BP_JOINT_HAS_LIMITS, BP_JOINT_HAS_MOTOR = 1, 2
BP_GUARD_NONE, BP_GUARD_RIGID_ROTATION_BUDGET = 0, 1
BP_CONTACT_RIGID = 3
RIGID_BODY_DTYPE = np.dtype(RigidBody)
JOINT_DTYPE = np.dtype(Joint)
JOINT_DESC_DTYPE = np.dtype(JointDesc)
RIGID_FRAME_DTYPE = np.dtype(RigidFrameInfo)
CONTOUR_DTYPE = np.dtype(BubbleContour)

BP_PLACEMENT_OK, BP_PLACEMENT_OUTSIDE_REGION, BP_PLACEMENT_INSIDE_RIGID_BODY, BP_PLACEMENT_COINCIDENT_CENTER = 0, 1, 2, 3
(BP_REJECT_NONE, BP_REJECT_INVALID_STRUCTURAL_OPERATION, BP_REJECT_INVALID_PLACEMENT, BP_REJECT_DUPLICATE_ID,
 BP_REJECT_UNKNOWN_OBJECT, BP_REJECT_INCOMPATIBLE_MOTION_ROLE, BP_REJECT_INVALID_VALUE) = 0, 1, 2, 3, 4, 5, 6


# numpy views of the record arrays (same layout as the C structures).
BUBBLE_DTYPE = np.dtype(Bubble)
SEGMENT_DTYPE = np.dtype(Segment)
BODY_DTYPE = np.dtype(Body)
MOTOR_DTYPE = np.dtype(Motor)
BAND_DTYPE = np.dtype(Band)
FIELD_DTYPE = np.dtype(Field)
FRAME_DTYPE = np.dtype(FrameInfo)
AUDIT_DTYPE = np.dtype(Audit)
CONTACT_DTYPE = np.dtype(Contact)
EVENT_DTYPE = np.dtype(ContactEvent)
BP_EVENT_CONTACT_BEGAN, BP_EVENT_CONTACT_ENDED = 1, 2
BP_CONTACT_PAIR, BP_CONTACT_SEGMENT, BP_CONTACT_BODY = 0, 1, 2
TUNE_DTYPE = np.dtype(TuneBin)
REQUEST_DTYPE = np.dtype(Request)
RESULT_DTYPE = np.dtype(ServerResult)

# Every exported function of the C ABI with its argument and result types.
_P = ctypes.c_void_p
_U32, _I32, _I64, _U64, _F = ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float
SIGNATURES = {
    "bp_abi_version": ([], _U32),
    "bp_abi_record_sizes": ([_P], None),
    "bp_world_config_default": ([_P], None),
    "bp_status_string": ([_I32], ctypes.c_char_p),
    "bp_last_error": ([], ctypes.c_char_p),
    "bp_world_create": ([_P, _P], _I32),
    "bp_world_destroy": ([_P], None),
    "bp_world_set_config": ([_P, _P], _I32),
    "bp_world_set_bubbles": ([_P, _P, _U32], _I32),
    "bp_world_update_bubbles": ([_P, _U64, _P, _P, _U32], _I32),
    "bp_world_publication": ([_P], _U64),
    "bp_world_region_results": ([_P, _P, _U32, _P, _P, _U32, _P], _I32),
    "bp_world_remove_bubbles": ([_P, _P, _U32], _I32),
    "bp_batch_region_results": ([_P, _U32, _P, _U32, _P, _P, _U32, _P], _I32),
    "bp_world_set_segments": ([_P, _P, _U32], _I32),
    "bp_world_set_bodies": ([_P, _P, _U32, _P, _U32, _P, _U32], _I32),
    "bp_world_set_forces": ([_P, _P, _U32, _P, _U32], _I32),
    "bp_world_set_fields": ([_P, _P, _U32, _P, _U32, _P, _U32], _I32),
    "bp_world_step": ([_P, _I32, _P], _I32),
    "bp_world_read_bubbles": ([_P, _P, _U32], _I32),
    "bp_world_read_bodies": ([_P, _P, _U32], _I32),
    "bp_world_bubble_count": ([_P], _U32),
    "bp_world_read_adjacency": ([_P, _P, _U32, _P, _U32, _P], _I32),
    "bp_world_read_step_history": ([_P, _P, _U32, _P], _I32),
    "bp_world_debug_fail_next_command": ([_P], _I32),
    "bp_world_audit": ([_P, _P], _I32),
    "bp_batch_audit": ([_P, _P, _U32], _I32),
    "bp_world_contacts": ([_P, _P, _U32, _P], _I32),
    "bp_world_contacts_noalloc": ([_P, _P, _U32, _P], _I32),
    "bp_batch_contacts": ([_P, _U32, _P, _U32, _P], _I32),
    "bp_world_body_contacts": ([_P, _P, _U32, _P], _I32),
    "bp_batch_body_contacts": ([_P, _U32, _P, _U32, _P], _I32),
    "bp_batch_set_pooling": ([_P, _U32], _I32),
    "bp_contact_transitions": ([_P, _U32, _P, _U32, _P, _U32, ctypes.c_float, _P, _U32, _U32, _P, _P, _U32, _P], _I32),
    "bp_stage_merge": ([_P, _P, _U32, _P, ctypes.c_int32, _P, _U32, _P, _P], _I32),
    "bp_stage_add": ([_P, _P, _P, _U32, _P, _P], _I32),
    "bp_stage_split": ([_P, ctypes.c_int64, _P, ctypes.c_float, _P, ctypes.c_float, _P, _U32, _P, _P], _I32),
    "bp_static_polygon_edges": ([_P, _U32, _I64, _I64, _U32, _P, _U32, _P], _I32),
    "bp_polygon_body_geometry": ([_P, _U32, _P, _P, _P, _P, _P, _P], _I32),
    "bp_place_discs": ([_P, _U32, _P, _F, _F, _P, _U32, _U64, _U32, _P, _P], _I32),
    "bp_batch_create": ([_P], _I32),
    "bp_batch_destroy": ([_P], None),
    "bp_batch_add_world": ([_P, _P, _P], _I32),
    "bp_batch_world_count": ([_P], _U32),
    "bp_batch_bubble_count": ([_P, _U32], _U32),
    "bp_batch_step": ([_P, _U32, _P], _I32),
    "bp_batch_read_bubbles": ([_P, _U32, _P, _U32], _I32),
    "bp_batch_read_bodies": ([_P, _U32, _P, _U32], _I32),
    "bp_batch_read_all_bubbles": ([_P, _P, _U32, _P, _U32], _I32),
    "bp_batch_last_gpu_milliseconds": ([_P], _F),
    "bp_batch_last_class_timing": ([_P, _P, _P, _P, _P], _I32),
    "bp_batch_class_count": ([], _U32),
    "bp_batch_step_worlds": ([_P, _P, _U32, _U32, _P], _I32),
    "bp_batch_submit_worlds": ([_P, _P, _U32, _U32], _I32),
    "bp_batch_poll": ([_P, _P], _I32),
    "bp_batch_write_bubbles": ([_P, _U32, _P, _U32], _I32),
    "bp_batch_set_resident": ([_P, _U32, _U32], _I32),
    "bp_batch_world_footprint": ([_P, _U32, _P], _I32),
    "bp_batch_set_world_class": ([_P, _U32, _I32], _I32),
    "bp_batch_capture": ([_P, _U32], _I32),
    "bp_batch_replay": ([_P, _P], _I32),
    "bp_batch_autotune": ([_P, _U32, _P, _U32, _P], _I32),
    "bp_batch_apply_tuning": ([_P, _P, _U32], _I32),
    "bp_world_contour_stride": ([_P, _P, _P], _I32),
    "bp_world_contours": ([_P, _P, _U32, _P, _P, _U32], _I32),
    "bp_batch_contours": ([_P, _U32, _P, _U32, _P, _P, _U32], _I32),
    "bp_rigid_joint_tables": ([_P, _U32, _P, _U32, _P, _P, _P, _P, _P, _P, _P], _I32),
    "bp_rigid_batch_create": ([_P], _I32),
    "bp_rigid_batch_destroy": ([_P], None),
    "bp_rigid_batch_add_world": ([_P, _P, _P], _I32),
    "bp_rigid_batch_world_count": ([_P], _U32),
    "bp_rigid_batch_set_world": ([_P, _U32, _P], _I32),
    "bp_rigid_batch_step": ([_P, _U32, _P], _I32),
    "bp_rigid_batch_read_bodies": ([_P, _U32, _P, _U32], _I32),
    "bp_rigid_batch_read_joints": ([_P, _U32, _P, _U32], _I32),
    "bp_rigid_batch_contacts": ([_P, _U32, _P, _U32, _P], _I32),
    "bp_rigid_batch_last_gpu_milliseconds": ([_P], _F),
    "bp_server_record_sizes": ([_P], None),
    "bp_server_config_default": ([_P], None),
    "bp_server_create": ([_P, _P, _P], _I32),
    "bp_server_destroy": ([_P], None),
    "bp_server_submit": ([_P, _P], _I32),
    "bp_server_poll": ([_P, _P, _U32, _P], _I32),
    "bp_server_wait_idle": ([_P], _I32),
    "bp_server_run_trace": ([_P, _P, _U32, _P], _I32),
    "bp_server_read_bubbles": ([_P, _U32, _P, _U32, _P], _I32),
    "bp_server_get_stats": ([_P, _P], _I32),
    "bp_server_clock": ([_P], ctypes.c_double),
}


class LibraryMissing(RuntimeError):
    """The HIP library is not built; run __graft_entry__.build()."""


class BubblePhysicsError(RuntimeError):
    def __init__(self, status: int, where: str, text: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)} ({status}) {text}".strip())


_lib = None


def load_library(path: os.PathLike | None = None) -> ctypes.CDLL:
    """Loads the HIP library once; raises LibraryMissing when it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise LibraryMissing(f"{p} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.bp_abi_version() != 1:
        raise RuntimeError("bubblephysics ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def ptr(array: np.ndarray | None) -> int | None:
    """Address of a C-contiguous numpy array (None for empty or missing arrays)."""
    if array is None or array.size == 0:
        return None
    assert array.flags["C_CONTIGUOUS"], "arrays passed to the C ABI must be contiguous"
    return array.ctypes.data


def check(status: int, where: str) -> None:
    if status != BP_OK:
        lib = load_library()
        raise BubblePhysicsError(status, where, lib.bp_last_error().decode(errors="replace"))


def default_config() -> WorldConfig:
    cfg = WorldConfig()
    load_library().bp_world_config_default(ctypes.byref(cfg))
    return cfg


def default_config_values() -> WorldConfig:
    """Reference defaults without loading the library (NativeSolverInput.swift:128-143)."""
    cfg = WorldConfig()
    cfg.time_step = 1.0 / 60.0
    cfg.contact_tolerance = 0.001
    cfg.separation_tolerance = 0.002
    cfg.candidate_limit = 128
    cfg.substep_split = 1
    cfg.grid_bucket_count = 0
    return cfg
