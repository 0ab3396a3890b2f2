"""ctypes wrapper of the CPU oracle (oracle/bp_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the CPU-baseline
leg of bench.py, as the checker of the HIP library. Nothing in bubblephysics_amd imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "lib" / "libbp_oracle.so"


class OracleFrame(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32), ("substeps", ctypes.c_uint32), ("rebuilds", ctypes.c_uint32),
        ("candidate_pairs", ctypes.c_uint32), ("candidate_overflow", ctypes.c_uint32), ("max_candidates", ctypes.c_uint32),
        ("removals", ctypes.c_uint32), ("wraps", ctypes.c_uint32), ("elapsed", ctypes.c_double),
        ("limiter_object", ctypes.c_uint64), ("limiter_term", ctypes.c_uint32), ("violation_kind", ctypes.c_uint32),
        ("violation_segment", ctypes.c_int64), ("violation_objects", ctypes.c_uint64 * 2), ("violation_detail", ctypes.c_double),
    ]


FRAME_DTYPE = np.dtype(OracleFrame)
_lib = None


def build() -> Path:
    """Builds the oracle library with its Makefile (g++, strict IEEE flags)."""
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        P, U32, I32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
        lib.bpo_world_run.argtypes = [P, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32,
                                      U32, I32, U32, U32, P, P, P, U32, P, P, P, P, P, U32, P]
        lib.bpo_set_plant.argtypes = [I32]
        lib.bpo_noncontact_acceleration.argtypes = [P, P, U32, P, U32, P, U32, P, U32, P, U32, P, U32, P]
        lib.bpo_world_run.restype = ctypes.c_int
        lib.bpo_epsilon_f32.argtypes = [P, P, U32]
        lib.bpo_wall_law_f32.argtypes = [P, P, U32]
        lib.bpo_pair_law_f32.argtypes = [P, P, P, U32]
        lib.bpo_half_impulse_factor_f32.argtypes = [P, P, P, U32]
        lib.bpo_detmath_f32.argtypes = [ctypes.c_int, P, P, U32]
        lib.bpo_audit.argtypes = [P, P, U32, P, U32, P, U32, P, U32, P, U32, P]
        lib.bpo_contacts.argtypes = [P, P, U32, P, U32, P, P, P, U32]
        lib.bpo_contacts.restype = U32
        lib.bpo_body_contacts.argtypes = [P, P, U32, P, U32, P, U32, P, U32, P, P, P, U32]
        lib.bpo_body_contacts.restype = U32
        _lib = lib
    return _lib


def _p(a):
    return None if a is None or a.size == 0 else a.ctypes.data


def set_threads(threads: int):
    """Host threads of the oracle's per-bubble phases (1 = sequential); the results do not change."""
    load().bpo_set_threads(int(threads))


def set_plant(code: int):
    """Switches on a planted mistake of the oracle (0 = none); only for tests that show a pin fails."""
    load().bpo_set_plant(int(code))


def run(scene, steps: int | None = None, *, precision: int = 32, assert_mode: int = 1, grid: str = "auto",
        adjacency: bool = False, history: bool = False, region: bool = False, capture: bool = False):
    """Runs `steps` steps of a scene in the oracle; returns (bubbles, bodies, frames, extras).

    The scene is not modified. grid: "auto" (hashed for <= 1024 bubbles like the reference's
    single-dispatch path, sorted otherwise), "hashed" or "sorted". region: extras["region"] holds
    each bubble's region result of the last step (BP_REGION_*). capture: extras["substeps"] holds
    the (x, y, vx, vy) of every bubble after each substep of the last step, shape (substeps, n, 4).
    """
    lib = load()
    steps = scene.steps if steps is None else steps
    bubbles = scene.bubbles.copy()
    bodies = scene.bodies.copy()
    frames = np.zeros(max(steps, 1), FRAME_DTYPE)
    n = len(bubbles)
    row_offsets = entries = hist = None
    if adjacency:
        row_offsets = np.zeros(n + 1, np.uint32)
        entries = np.zeros(max(n * 256, 16), np.uint32)
    terms = objects = None
    if history:
        hist = np.zeros(1024, np.float32)
        terms = np.zeros(1024, np.uint32)
        objects = np.zeros(1024, np.uint64)
    verts = np.ascontiguousarray(scene.vertices, np.float32)
    tris = np.ascontiguousarray(scene.triangles, np.uint32)
    fverts = np.ascontiguousarray(scene.field_vertices, np.float32).reshape(-1, 2)
    fids = np.ascontiguousarray(scene.field_filter_ids, np.int64)
    override = {"auto": 0, "hashed": 1, "sorted": 2}[grid]
    region_out = np.zeros(max(n, 1), np.uint32) if region else None
    cap = np.zeros(4 * max(n, 1) * 1024, np.float32) if capture else None
    cap_count = ctypes.c_uint32(0)
    executed = lib.bpo_world_run(ctypes.byref(scene.config), _p(bubbles), n, _p(scene.segments), len(scene.segments),
                                 _p(bodies), len(bodies), _p(verts), len(verts), _p(tris), len(tris), _p(scene.motors),
                                 len(scene.motors), _p(scene.bands), len(scene.bands), _p(scene.fields), len(scene.fields),
                                 _p(fverts), len(fverts), _p(fids), len(fids), steps, precision, assert_mode,
                                 override, _p(frames), _p(row_offsets), _p(entries), 0 if entries is None else len(entries),
                                 _p(hist), _p(terms), _p(objects), _p(region_out), _p(cap), 0 if cap is None else len(cap),
                                 ctypes.byref(cap_count))
    extras = {"executed": executed}
    if region:
        extras["region"] = region_out[:n].copy()
    if capture:
        extras["substeps"] = cap[: 4 * cap_count.value].reshape(-1, max(n, 1), 4).copy() if n else np.zeros((0, 0, 4))
    if adjacency:
        extras["row_offsets"] = row_offsets
        extras["entries"] = entries[: row_offsets[-1]].copy()
    if history:
        # h, limiting term and limiting bubble id of every substep of the last executed step.
        count = int(frames[executed - 1]["substeps"]) if executed else 0
        extras["history"] = hist
        extras["terms"] = terms[:count].copy()
        extras["objects"] = objects[:count].copy()
    return bubbles, bodies, frames[:executed], extras


AUDIT_FIELDS = [("momentum", np.float64, (2,)), ("angular_momentum", np.float64), ("kinetic_energy", np.float64),
                ("contact_energy", np.float64), ("contacts", np.uint32), ("bubbles", np.uint32)]
AUDIT_DTYPE = np.dtype(AUDIT_FIELDS)


def audit(scene, bubbles=None, bodies=None):
    """FP64 audit by brute force of a world state (the scene's records, or the given bubbles and
    bodies on the scene's geometry): momentum, angular momentum, kinetic and contact energy."""
    lib = load()
    b = np.ascontiguousarray(scene.bubbles if bubbles is None else bubbles)
    bod = np.ascontiguousarray(scene.bodies if bodies is None else bodies)
    verts = np.ascontiguousarray(scene.vertices, np.float32)
    tris = np.ascontiguousarray(scene.triangles, np.uint32)
    out = np.zeros(1, AUDIT_DTYPE)
    lib.bpo_audit(ctypes.byref(scene.config), _p(b), len(b), _p(scene.segments), len(scene.segments), _p(bod), len(bod),
                  _p(verts), len(verts), _p(tris), len(tris), _p(out))
    return out[0]


CONTACT_FIELDS = [("id", np.uint64), ("first", np.int64), ("second", np.int64), ("kind", np.uint32), ("feature", np.uint32),
                  ("normal", np.float32, (2,)), ("point", np.float32, (2,)), ("penetration", np.float32),
                  ("allowed_side", np.float32), ("impulse", np.float32), ("index_b", np.uint32)]
CONTACT_DTYPE = np.dtype(CONTACT_FIELDS)


def contacts(scene, bubbles, row_offsets, entries, capacity=None):
    """End-of-step contact records (bp_contact) of a world state, from the rows of its last grid
    build (run(..., adjacency=True) extras) on the scene's configuration and segments."""
    lib = load()
    b = np.ascontiguousarray(bubbles)
    offs = np.ascontiguousarray(row_offsets, np.uint32)
    ents = np.ascontiguousarray(entries, np.uint32)
    cap = int(offs[-1]) if capacity is None else capacity
    out = np.zeros(max(cap, 1), CONTACT_DTYPE)
    k = lib.bpo_contacts(ctypes.byref(scene.config), _p(b), len(b), _p(scene.segments), len(scene.segments), _p(offs),
                         _p(ents), _p(out), cap)
    return out[: min(k, cap)], int(k)


def body_contacts(scene, bubbles, bodies, row_offsets, entries, capacity=None):
    """End-of-step bubble-body contact records (bp_contact, kind 2) of a world state: the body
    entries of its last build's rows against the bodies at their end-of-step poses."""
    lib = load()
    b = np.ascontiguousarray(bubbles)
    bd = np.ascontiguousarray(bodies)
    verts = np.ascontiguousarray(scene.vertices, np.float32).reshape(-1, 2)
    tris = np.ascontiguousarray(scene.triangles, np.uint32).reshape(-1, 4)
    offs = np.ascontiguousarray(row_offsets, np.uint32)
    ents = np.ascontiguousarray(entries, np.uint32)
    cap = int(offs[-1]) if capacity is None else capacity
    out = np.zeros(max(cap, 1), CONTACT_DTYPE)
    k = lib.bpo_body_contacts(ctypes.byref(scene.config), _p(b), len(b), _p(bd), len(bd), _p(verts), len(verts), _p(tris),
                              len(tris), _p(offs), _p(ents), _p(out), cap)
    return out[: min(k, cap)], int(k)


def noncontact_acceleration(scene):
    """a0 of every bubble of a scene at its centre (FP32), shape (n, 2)."""
    n = len(scene.bubbles)
    out = np.zeros((n, 2), np.float32)
    fverts = np.ascontiguousarray(scene.field_vertices, np.float32).reshape(-1, 2)
    fids = np.ascontiguousarray(scene.field_filter_ids, np.int64)
    load().bpo_noncontact_acceleration(ctypes.byref(scene.config), _p(scene.bubbles), n, _p(scene.motors), len(scene.motors),
                                       _p(scene.bands), len(scene.bands), _p(scene.fields), len(scene.fields), _p(fverts),
                                       len(fverts), _p(fids), len(fids), _p(out))
    return out


def epsilon(q):
    q = np.ascontiguousarray(q, np.float32)
    out = np.empty_like(q)
    load().bpo_epsilon_f32(_p(q), _p(out), len(q))
    return out


def wall_law(inputs4):
    x = np.ascontiguousarray(inputs4, np.float32).reshape(-1, 4)
    out = np.empty_like(x)
    load().bpo_wall_law_f32(_p(x), _p(out), len(x))
    return out


def pair_law(inputs4, distances):
    x = np.ascontiguousarray(inputs4, np.float32).reshape(-1, 4)
    d = np.ascontiguousarray(distances, np.float32)
    out = np.empty_like(x)
    load().bpo_pair_law_f32(_p(x), _p(d), _p(out), len(x))
    return out


def half_impulse_factor(gamma, half):
    g = np.ascontiguousarray(gamma, np.float32)
    h = np.ascontiguousarray(half, np.float32)
    out = np.empty_like(g)
    load().bpo_half_impulse_factor_f32(_p(g), _p(h), _p(out), len(g))
    return out


def detmath(which: str, x):
    code = {"sin": 0, "cos": 1, "exp": 2, "log": 3}[which]
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    load().bpo_detmath_f32(code, _p(x), _p(out), len(x))
    return out


def candidate_rows(bubbles, margins, segments, cell_width, separation_tolerance, limit, *, hashed=False, buckets=2048):
    """Candidate rows of one grid build: (rows[n, limit, 2] of (kind, other), counts[n, 2], overflow[2])."""
    lib = load()
    if not hasattr(lib, "_rows_ready"):
        P, U32, F = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float
        lib.bpo_candidate_rows.argtypes = [P, P, U32, P, U32, F, F, U32, U32, U32, P, P, P]
        lib._rows_ready = True
    n = len(bubbles)
    b = np.ascontiguousarray(bubbles)
    m = np.ascontiguousarray(margins, np.float32)
    s = np.ascontiguousarray(segments)
    rows = np.full((max(n, 1), limit, 2), 0xFFFFFFFF, np.uint32)
    counts = np.zeros((max(n, 1), 2), np.uint32)
    overflow = np.zeros(2, np.uint32)
    lib.bpo_candidate_rows(_p(b), _p(m), n, _p(s), len(s), cell_width, separation_tolerance, limit, 1 if hashed else 0,
                           buckets, _p(rows), _p(counts), _p(overflow))
    return rows[:n], counts[:n], overflow


def run_rigid(scene, steps: int, *, plant: int = 0, contact_capacity: int = 4096):
    """Runs `steps` steps of a rigid world (bubblephysics_amd.rigid.RigidScene) in the oracle's
    restatement of the reference's rigid-body frame (oracle/bp_oracle_rigid.cpp).

    Returns (bodies, joints, infos[ran], contacts of the last published step). A failed step
    publishes nothing and ends the run."""
    from bubblephysics_amd import abi
    lib = load()
    if not hasattr(lib, "_rigid_ready"):
        P, U32 = ctypes.c_void_p, ctypes.c_uint32
        lib.bpo_rigid_run.argtypes = [P, P, U32, P, P, P, U32, P, U32, P, P, P, U32, P, P, U32, P]
        lib.bpo_rigid_run.restype = U32
        lib.bpo_rigid_set_plant.argtypes = [ctypes.c_int]
        lib._rigid_ready = True
    bodies = np.ascontiguousarray(scene.bodies.copy())
    joints = np.ascontiguousarray(scene.joints.copy())
    infos = np.zeros(steps, abi.RIGID_FRAME_DTYPE)
    contacts = np.zeros(contact_capacity, abi.CONTACT_DTYPE)
    count = ctypes.c_uint32()
    verts = np.ascontiguousarray(scene.vertices, dtype=np.float32)
    tris = np.ascontiguousarray(scene.triangles, dtype=np.uint32)
    segs = np.ascontiguousarray(scene.segments)
    refs = np.ascontiguousarray(scene.references, dtype=np.uint32)
    ranges = None if scene.ranges is None or len(joints) == 0 else np.ascontiguousarray(scene.ranges, dtype=np.uint32)
    partners = np.ascontiguousarray(scene.partners, dtype=np.uint64)
    lib.bpo_rigid_set_plant(int(plant))
    try:
        ran = lib.bpo_rigid_run(ctypes.byref(scene.config), _p(bodies), len(bodies), _p(verts), _p(tris), _p(segs), len(segs),
                                _p(joints), len(joints), _p(refs), _p(ranges), _p(partners), steps, _p(infos),
                                _p(contacts), contact_capacity, ctypes.byref(count))
    finally:
        lib.bpo_rigid_set_plant(0)
    return bodies, joints, infos[:ran], contacts[:min(count.value, contact_capacity)].copy()


def contour_stride(bubbles, config) -> int:
    """max over bubbles of max(64, ceil(4 pi r / max_segment_length)) in double (NativeMetalSolver.swift:407-411)."""
    import math
    best = 64.0
    for r in bubbles["radius"]:
        best = max(best, math.ceil(4.0 * math.pi * float(r) / float(np.float32(config.max_segment_length))))
    return int(best)


def contours(bubbles, segments, row_offsets, entries, config, stride=None):
    """Draw-only contours of a world state from the rows of its last grid build (oracle restatement
    of nativeAccumulateContactPressure, nativeGenerateContours and nativePrepareRenderData):
    (points [n, stride, 2], shapes [n])."""
    from bubblephysics_amd import abi
    lib = load()
    if not hasattr(lib, "_contours_ready"):
        P, U32 = ctypes.c_void_p, ctypes.c_uint32
        lib.bpo_contours.argtypes = [P, P, U32, P, U32, P, P, U32, P, P]
        lib._contours_ready = True
    stride = stride or contour_stride(bubbles, config)
    b = np.ascontiguousarray(bubbles)
    n = len(b)
    points = np.zeros((n, stride, 2), np.float32)
    shapes = np.zeros(n, abi.CONTOUR_DTYPE)
    segs = np.ascontiguousarray(segments)
    lib.bpo_contours(ctypes.byref(config), _p(b), n, _p(segs), len(segs), _p(np.ascontiguousarray(row_offsets, np.uint32)),
                     _p(np.ascontiguousarray(entries, np.uint32)), stride, _p(points), _p(shapes))
    return points, shapes
