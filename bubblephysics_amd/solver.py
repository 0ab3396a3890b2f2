"""Solver-level host interface: the reference's NativeMetalFrameExecuting seam over the C ABI.

Mirrors NativeMetalSolver.executeFrame(input:scenarioStep:) -> NativeSolverFrame
(Sources/BubblePhysicsNativeMetal/NativeSolverInput.swift:255-264): the solver keeps the world on
the device, publishes a step only when it succeeds, and reports failures with the reference's
error cases (NativeMetalWorldError) and failure detail (NativeMetalFailureDetail).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import abi


@dataclass
class FailureDetail:
    object_ids: tuple
    violation_kind: int
    edge_segment_id: int | None
    distance: float | None
    limiting_term: int | None


@dataclass
class SolverFrame:
    """NativeSolverFrame (NativeSolverInput.swift:224-253) for the fields the step produces."""
    status: int
    bubbles: np.ndarray
    substep_count: int
    grid_rebuild_count: int
    candidate_pair_count: int
    contact_candidate_overflow_count: int
    maximum_contact_candidates_per_object: int
    group_width: int
    gpu_milliseconds: float
    region_removals: set
    region_wraps: set
    limiter_object: int
    limiter_term: int
    detail: FailureDetail | None
    execution_path: int = 0          # BP_PATH_WORKGROUP or BP_PATH_COOPERATIVE

    @property
    def ok(self) -> bool:
        return self.status == abi.BP_OK


def _detail(info: abi.FrameInfo) -> FailureDetail | None:
    if info.status not in (abi.BP_ERR_CONTACT_VIOLATION, abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED):
        return None
    d = info.detail
    dist = None if np.isnan(d.distance) else float(d.distance)
    return FailureDetail(tuple(int(x) for x in d.object_ids), int(d.violation_kind),
                         None if d.edge_segment_id_low < 0 else int(d.edge_segment_id_low), dist,
                         None if d.limiting_term == abi.BP_TERM_NONE else int(d.limiting_term))


class NativeWorld:
    """One world on the GPU (bp_world). Raises BubblePhysicsError for call errors; step results
    come back as SolverFrame with the status of the step."""

    def __init__(self, scene=None, config: abi.WorldConfig | None = None):
        self._lib = abi.load_library()
        self._handle = ctypes.c_void_p()
        cfg = config if config is not None else (scene.config if scene is not None else abi.default_config())
        abi.check(self._lib.bp_world_create(ctypes.byref(cfg), ctypes.byref(self._handle)), "bp_world_create")
        self._count = 0
        self._bodies = 0
        if scene is not None:
            self.load(scene)

    def close(self):
        if self._handle:
            self._lib.bp_world_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def load(self, scene):
        self.set_config(scene.config)
        self.set_bubbles(scene.bubbles)
        self.set_segments(scene.segments)
        self.set_bodies(scene.bodies, scene.vertices, scene.triangles)
        self.set_forces(scene.motors, scene.bands)
        if len(scene.fields):
            self.set_fields(scene.fields, scene.field_vertices, scene.field_filter_ids)

    def set_config(self, cfg: abi.WorldConfig):
        abi.check(self._lib.bp_world_set_config(self._handle, ctypes.byref(cfg)), "bp_world_set_config")

    def set_bubbles(self, bubbles: np.ndarray):
        b = np.ascontiguousarray(bubbles, abi.BUBBLE_DTYPE)
        abi.check(self._lib.bp_world_set_bubbles(self._handle, abi.ptr(b), len(b)), "bp_world_set_bubbles")
        self._count = len(b)

    @property
    def publication(self) -> int:
        """Number of the last published snapshot (the base a delta upload must name)."""
        return int(self._lib.bp_world_publication(self._handle))

    def update_bubbles(self, indices, bubbles: np.ndarray, base_publication: int | None = None):
        """Delta upload against the snapshot `base_publication` (default: the last published one)."""
        idx = np.ascontiguousarray(indices, np.uint32)
        b = np.ascontiguousarray(bubbles, abi.BUBBLE_DTYPE)
        base = self.publication if base_publication is None else int(base_publication)
        abi.check(self._lib.bp_world_update_bubbles(self._handle, base, abi.ptr(idx), abi.ptr(b), len(b)),
                  "bp_world_update_bubbles")

    def region_results(self):
        """(removed ids, wrapped ids) of the last published step, ascending."""
        return _region_results(lambda *a: self._lib.bp_world_region_results(self._handle, *a), "bp_world_region_results")

    def remove_bubbles(self, ids):
        """Deletes bubbles by id, as the public world deletes the ones the region removed."""
        arr = np.ascontiguousarray(ids, np.int64)
        abi.check(self._lib.bp_world_remove_bubbles(self._handle, abi.ptr(arr), len(arr)), "bp_world_remove_bubbles")
        self._count -= len(arr)

    def set_segments(self, segments: np.ndarray):
        s = np.ascontiguousarray(segments, abi.SEGMENT_DTYPE)
        abi.check(self._lib.bp_world_set_segments(self._handle, abi.ptr(s), len(s)), "bp_world_set_segments")

    def set_bodies(self, bodies: np.ndarray, vertices: np.ndarray, triangles: np.ndarray):
        b = np.ascontiguousarray(bodies, abi.BODY_DTYPE)
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 2)
        t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 4)
        abi.check(self._lib.bp_world_set_bodies(self._handle, abi.ptr(b), len(b), abi.ptr(v), len(v), abi.ptr(t), len(t)),
                  "bp_world_set_bodies")
        self._bodies = len(b)

    def set_forces(self, motors: np.ndarray, bands: np.ndarray):
        m = np.ascontiguousarray(motors, abi.MOTOR_DTYPE)
        b = np.ascontiguousarray(bands, abi.BAND_DTYPE)
        abi.check(self._lib.bp_world_set_forces(self._handle, abi.ptr(m), len(m), abi.ptr(b), len(b)), "bp_world_set_forces")

    def set_fields(self, fields: np.ndarray, vertices: np.ndarray, filter_ids: np.ndarray):
        f = np.ascontiguousarray(fields, abi.FIELD_DTYPE)
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 2)
        ids = np.ascontiguousarray(filter_ids, np.int64)
        abi.check(self._lib.bp_world_set_fields(self._handle, abi.ptr(f), len(f), abi.ptr(v), len(v), abi.ptr(ids), len(ids)),
                  "bp_world_set_fields")

    def step(self, scenario_step: int = 0) -> SolverFrame:
        info = abi.FrameInfo()
        status = self._lib.bp_world_step(self._handle, scenario_step, ctypes.byref(info))
        if status in (abi.BP_ERR_INVALID_ARGUMENT, abi.BP_ERR_ALLOCATION_FAILED, abi.BP_ERR_RESOURCE_UNAVAILABLE):
            abi.check(status, "bp_world_step")
        return self._frame(info)

    def _frame(self, info: abi.FrameInfo) -> SolverFrame:
        bubbles = self.read_bubbles()
        removed, wrapped = (set(), set())
        if info.status == abi.BP_OK and (info.region_removals or info.region_wraps):
            r, w = self.region_results()
            removed, wrapped = set(int(x) for x in r), set(int(x) for x in w)
        return SolverFrame(
            status=int(info.status), bubbles=bubbles, substep_count=int(info.substep_count),
            grid_rebuild_count=int(info.grid_rebuild_count), candidate_pair_count=int(info.candidate_pair_count),
            contact_candidate_overflow_count=int(info.candidate_overflow),
            maximum_contact_candidates_per_object=int(info.max_candidates_per_object), group_width=int(info.group_width),
            gpu_milliseconds=float(info.gpu_milliseconds), region_removals=removed, region_wraps=wrapped,
            limiter_object=int(info.limiter_object), limiter_term=int(info.limiter_term), detail=_detail(info),
            execution_path=int(info.execution_path))

    def read_bubbles(self) -> np.ndarray:
        out = np.zeros(self._count, abi.BUBBLE_DTYPE)
        abi.check(self._lib.bp_world_read_bubbles(self._handle, abi.ptr(out), self._count), "bp_world_read_bubbles")
        return out

    def read_bodies(self) -> np.ndarray:
        out = np.zeros(self._bodies, abi.BODY_DTYPE)
        abi.check(self._lib.bp_world_read_bodies(self._handle, abi.ptr(out), self._bodies), "bp_world_read_bodies")
        return out

    def read_adjacency(self):
        """(row offsets, entries packed as kind << 30 | neighbour) of the last grid build."""
        offsets = np.zeros(self._count + 1, np.uint32)
        total = ctypes.c_uint32(0)
        abi.check(self._lib.bp_world_read_adjacency(self._handle, abi.ptr(offsets), len(offsets), None, 0, ctypes.byref(total)),
                  "bp_world_read_adjacency")
        entries = np.zeros(max(total.value, 1), np.uint32)
        abi.check(self._lib.bp_world_read_adjacency(self._handle, abi.ptr(offsets), len(offsets), abi.ptr(entries), len(entries),
                                                    ctypes.byref(total)), "bp_world_read_adjacency")
        return offsets, entries[: total.value]

    def audit(self) -> np.ndarray:
        """FP64 audit (bp_audit record) of the published state."""
        out = np.zeros(1, abi.AUDIT_DTYPE)
        abi.check(self._lib.bp_world_audit(self._handle, abi.ptr(out)), "bp_world_audit")
        return out[0]

    def contacts(self, capacity: int | None = None) -> np.ndarray:
        """End-of-step contact records (bp_contact) of the published state."""
        return _read_contacts(lambda out, cap, count: self._lib.bp_world_contacts(self._handle, out, cap, count),
                              "bp_world_contacts", capacity)

    def contour_stride(self, config: abi.ContourConfig | None = None) -> int:
        cfg = config or abi.contour_config()
        out = ctypes.c_uint32()
        abi.check(self._lib.bp_world_contour_stride(self._handle, ctypes.byref(cfg), ctypes.byref(out)), "bp_world_contour_stride")
        return out.value

    def contours(self, config: abi.ContourConfig | None = None, stride: int | None = None):
        """Draw-only contours of the published state (bp_world_contours): (points [n, stride, 2],
        shapes [n])."""
        cfg = config or abi.contour_config()
        stride = stride or self.contour_stride(cfg)
        n = self._count
        points = np.zeros((n, stride, 2), np.float32)
        shapes = np.zeros(n, abi.CONTOUR_DTYPE)
        abi.check(self._lib.bp_world_contours(self._handle, ctypes.byref(cfg), stride, abi.ptr(points), abi.ptr(shapes), n),
                  "bp_world_contours")
        return points, shapes

    def body_contacts(self, capacity: int | None = None) -> np.ndarray:
        """End-of-step bubble-body contact records (bp_contact, kind BP_CONTACT_BODY)."""
        return _read_contacts(lambda out, cap, count: self._lib.bp_world_body_contacts(self._handle, out, cap, count),
                              "bp_world_body_contacts", capacity)

    def read_step_history(self) -> np.ndarray:
        out = np.zeros(1024, np.float32)
        count = ctypes.c_uint32(0)
        abi.check(self._lib.bp_world_read_step_history(self._handle, abi.ptr(out), 1024, ctypes.byref(count)),
                  "bp_world_read_step_history")
        return out[: count.value]

    def fail_next_command(self):
        abi.check(self._lib.bp_world_debug_fail_next_command(self._handle), "bp_world_debug_fail_next_command")


def _region_results(call, name: str):
    """Runs a region-results call: counts first, then the ids."""
    nr, nw = ctypes.c_uint32(0), ctypes.c_uint32(0)
    abi.check(call(None, 0, ctypes.byref(nr), None, 0, ctypes.byref(nw)), name)
    removed = np.zeros(max(nr.value, 1), np.int64)
    wrapped = np.zeros(max(nw.value, 1), np.int64)
    abi.check(call(abi.ptr(removed), len(removed), ctypes.byref(nr), abi.ptr(wrapped), len(wrapped), ctypes.byref(nw)), name)
    return removed[: nr.value], wrapped[: nw.value]


def _read_contacts(call, name: str, capacity: int | None) -> np.ndarray:
    """Runs a contact-record call; without a capacity, asks for the count first."""
    count = ctypes.c_uint32(0)
    if capacity is None:
        abi.check(call(None, 0, ctypes.byref(count)), name)
        capacity = count.value
    out = np.zeros(max(capacity, 1), abi.CONTACT_DTYPE)
    abi.check(call(abi.ptr(out), capacity, ctypes.byref(count)), name)
    return out[: min(count.value, capacity)]


def contact_transitions(contacts: np.ndarray, segments: np.ndarray, interested, step_duration: float,
                        history: np.ndarray | None = None):
    """Contact began/ended events of one step (bp_contact_transitions) from its contact records and
    the previous step's pair set; returns (events, new pair set as an (n, 2) uint64 array)."""
    lib = abi.load_library()
    c = np.ascontiguousarray(contacts, abi.CONTACT_DTYPE)
    seg = np.ascontiguousarray(segments, abi.SEGMENT_DTYPE)
    ids = np.ascontiguousarray(np.asarray(list(interested), np.uint64))
    prev = np.zeros((0, 2), np.uint64) if history is None else np.ascontiguousarray(history, np.uint64).reshape(-1, 2)
    cap = len(c) + len(prev) + 1
    hist = np.zeros((cap, 2), np.uint64)
    hist[: len(prev)] = prev
    events = np.zeros(cap, abi.EVENT_DTYPE)
    hc, ec = ctypes.c_uint32(0), ctypes.c_uint32(0)
    abi.check(lib.bp_contact_transitions(abi.ptr(c), len(c), abi.ptr(seg), len(seg), abi.ptr(ids), len(ids),
                                         ctypes.c_float(step_duration), abi.ptr(hist), len(prev), cap, ctypes.byref(hc),
                                         abi.ptr(events), cap, ctypes.byref(ec)), "bp_contact_transitions")
    return events[: ec.value].copy(), hist[: hc.value].copy()


class StageRejected(abi.BubblePhysicsError):
    """A merge or split the staging refused (TransactionRejectionReason): reason, issue, object id."""

    def __init__(self, why: abi.StageRejection, where: str):
        super().__init__(abi.BP_ERR_INVALID_ARGUMENT, where, f"reason {why.reason} issue {why.issue} object {why.object_id}")
        self.reason, self.issue, self.object_id = int(why.reason), int(why.issue), int(why.object_id)


def _stage(fn, name, bubbles, bodies, vertices, region, extra_out, *args):
    lib = abi.load_library()
    b = np.ascontiguousarray(bubbles, abi.BUBBLE_DTYPE)
    bod = np.ascontiguousarray(np.zeros(0, abi.BODY_DTYPE) if bodies is None else bodies, abi.BODY_DTYPE)
    v = np.ascontiguousarray(np.zeros((0, 2), np.float32) if vertices is None else vertices, np.float32).reshape(-1, 2)
    sc = abi.StageScene(abi.ptr(b), len(b), len(bod), abi.ptr(bod), abi.ptr(v), len(v), (ctypes.c_float * 4)(*region))
    out = np.zeros(len(b) + extra_out, abi.BUBBLE_DTYPE)
    count, why = ctypes.c_uint32(0), abi.StageRejection()
    status = getattr(lib, fn)(ctypes.byref(sc), *args, abi.ptr(out), len(out), ctypes.byref(count), ctypes.byref(why))
    if status != abi.BP_OK and why.reason != abi.BP_REJECT_NONE:
        raise StageRejected(why, name)
    abi.check(status, name)
    return out[: count.value].copy()


def stage_merge(bubbles, sources, result, *, result_mass=None, bodies=None, vertices=None, region=(-1e30, -1e30, 1e30, 1e30)):
    """World bubbles after merging `sources` into `result` (a BUBBLE_DTYPE record: id, filter,
    material, flags) -- bp_stage_merge, TransactionStager.merge. Raises StageRejected."""
    res = np.zeros(1, abi.BUBBLE_DTYPE)
    res[0] = result
    if result_mass is not None:
        res[0]["mass"] = result_mass
    ids = np.ascontiguousarray(np.asarray(list(sources), np.int64))
    return _stage("bp_stage_merge", "bp_stage_merge", bubbles, bodies, vertices, region, 1, abi.ptr(ids), len(ids),
                  abi.ptr(res), 1 if result_mass is not None else 0)


def stage_add(bubbles, bubble, *, bodies=None, vertices=None, region=(-1e30, -1e30, 1e30, 1e30)):
    """World bubbles after adding `bubble` (a BUBBLE_DTYPE record) -- bp_stage_add,
    TransactionStager.add with the placement rules. Raises StageRejected."""
    rec = np.zeros(1, abi.BUBBLE_DTYPE)
    rec[0] = bubble
    return _stage("bp_stage_add", "bp_stage_add", bubbles, bodies, vertices, region, 1, abi.ptr(rec))


def stage_split(bubbles, source, first, first_share, second, second_share, *, bodies=None, vertices=None,
                region=(-1e30, -1e30, 1e30, 1e30)):
    """World bubbles after splitting `source` into `first` and `second` with area shares --
    bp_stage_split, TransactionStager.split. Raises StageRejected."""
    parts = np.zeros(2, abi.BUBBLE_DTYPE)
    parts[0], parts[1] = first, second
    return _stage("bp_stage_split", "bp_stage_split", bubbles, bodies, vertices, region, 2, ctypes.c_int64(int(source)),
                  abi.ptr(parts[0:1]), ctypes.c_float(first_share), abi.ptr(parts[1:2]), ctypes.c_float(second_share))


class Batch:
    """Many independent worlds advanced together (bp_batch)."""

    def __init__(self, scenes=()):
        self._lib = abi.load_library()
        self._handle = ctypes.c_void_p()
        abi.check(self._lib.bp_batch_create(ctypes.byref(self._handle)), "bp_batch_create")
        self._counts = []
        for s in scenes:
            self.add(s)

    def add(self, scene) -> int:
        index = ctypes.c_uint32(0)
        d = scene.desc()
        abi.check(self._lib.bp_batch_add_world(self._handle, ctypes.byref(d), ctypes.byref(index)), "bp_batch_add_world")
        self._counts.append(len(scene.bubbles))
        return index.value

    def step(self, steps: int = 1) -> np.ndarray:
        """Runs `steps` steps; returns FrameInfo records shaped (steps, worlds)."""
        infos = np.zeros(steps * len(self._counts), abi.FRAME_DTYPE)
        abi.check(self._lib.bp_batch_step(self._handle, steps, abi.ptr(infos)), "bp_batch_step")
        return infos.reshape(steps, len(self._counts))

    def region_results(self, world: int):
        """(ids removed during the last stepping call, ids wrapped in the last published step)."""
        return _region_results(lambda *a: self._lib.bp_batch_region_results(self._handle, world, *a),
                               "bp_batch_region_results")

    def read_all_bubbles(self):
        """(records of every world concatenated, offsets[worlds + 1]) in one read."""
        offsets = np.zeros(len(self._counts) + 1, np.uint32)
        abi.check(self._lib.bp_batch_read_all_bubbles(self._handle, None, 0, abi.ptr(offsets), len(offsets)),
                  "bp_batch_read_all_bubbles")
        out = np.zeros(max(int(offsets[-1]), 1), abi.BUBBLE_DTYPE)
        abi.check(self._lib.bp_batch_read_all_bubbles(self._handle, abi.ptr(out), len(out), abi.ptr(offsets), len(offsets)),
                  "bp_batch_read_all_bubbles")
        return out[: int(offsets[-1])], offsets

    def read_bodies(self, world: int, count: int) -> np.ndarray:
        """Bodies of one world at their published poses."""
        out = np.zeros(count, abi.BODY_DTYPE)
        abi.check(self._lib.bp_batch_read_bodies(self._handle, world, abi.ptr(out), count), "bp_batch_read_bodies")
        return out

    def read_bubbles(self, world: int) -> np.ndarray:
        self._counts[world] = int(self._lib.bp_batch_bubble_count(self._handle, world))
        out = np.zeros(self._counts[world], abi.BUBBLE_DTYPE)
        abi.check(self._lib.bp_batch_read_bubbles(self._handle, world, abi.ptr(out), len(out)), "bp_batch_read_bubbles")
        return out

    def audit(self) -> np.ndarray:
        """FP64 audit (bp_audit records) of every world's published state."""
        out = np.zeros(len(self._counts), abi.AUDIT_DTYPE)
        abi.check(self._lib.bp_batch_audit(self._handle, abi.ptr(out), len(out)), "bp_batch_audit")
        return out

    def contacts(self, world: int, capacity: int | None = None) -> np.ndarray:
        """End-of-step contact records (bp_contact) of one world's published state."""
        return _read_contacts(lambda out, cap, count: self._lib.bp_batch_contacts(self._handle, world, out, cap, count),
                              "bp_batch_contacts", capacity)

    def body_contacts(self, world: int, capacity: int | None = None) -> np.ndarray:
        """End-of-step bubble-body contact records (bp_contact, kind BP_CONTACT_BODY) of one world."""
        return _read_contacts(lambda out, cap, count: self._lib.bp_batch_body_contacts(self._handle, world, out, cap, count),
                              "bp_batch_body_contacts", capacity)

    def contours(self, world: int, stride: int, config: abi.ContourConfig | None = None):
        """Draw-only contours of one world (bp_batch_contours): (points [n, stride, 2], shapes [n])."""
        cfg = config or abi.contour_config()
        n = self._counts[world]
        points = np.zeros((n, stride, 2), np.float32)
        shapes = np.zeros(n, abi.CONTOUR_DTYPE)
        abi.check(self._lib.bp_batch_contours(self._handle, world, ctypes.byref(cfg), stride, abi.ptr(points), abi.ptr(shapes),
                                              n), "bp_batch_contours")
        return points, shapes

    def capture(self, steps: int = 1):
        """Captures `steps` steps of every world into a hipGraph (bp_batch_capture)."""
        abi.check(self._lib.bp_batch_capture(self._handle, steps), "bp_batch_capture")
        self._graph_steps = steps

    def replay(self) -> np.ndarray:
        """One replay of the captured steps; FrameInfo records shaped (steps, worlds)."""
        steps = self._graph_steps
        infos = np.zeros(steps * len(self._counts), abi.FRAME_DTYPE)
        abi.check(self._lib.bp_batch_replay(self._handle, abi.ptr(infos)), "bp_batch_replay")
        return infos.reshape(steps, len(self._counts))

    def last_class_timing(self):
        """Per kernel class: (device milliseconds, worlds, threads per workgroup, bubble capacity)."""
        k = self._lib.bp_batch_class_count()
        ms = np.zeros(k, np.float32)
        worlds, threads, cap = (np.zeros(k, np.uint32) for _ in range(3))
        abi.check(self._lib.bp_batch_last_class_timing(self._handle, abi.ptr(ms), abi.ptr(worlds), abi.ptr(threads),
                                                       abi.ptr(cap)), "bp_batch_last_class_timing")
        return ms, worlds, threads, cap

    def set_world_class(self, world: int, cls: int):
        abi.check(self._lib.bp_batch_set_world_class(self._handle, world, cls), "bp_batch_set_world_class")

    def set_pooling(self, enabled: bool):
        """Workgroup pools of the batched launch (bp_batch_set_pooling); results do not change."""
        abi.check(self._lib.bp_batch_set_pooling(self._handle, 1 if enabled else 0), "bp_batch_set_pooling")

    # ---- serving -------------------------------------------------------------------------
    def step_worlds(self, worlds, steps: int = 1) -> np.ndarray:
        """Advances only the listed worlds; returns FrameInfo records shaped (steps, len(worlds))."""
        w = np.ascontiguousarray(worlds, np.uint32)
        infos = np.zeros(steps * len(w), abi.FRAME_DTYPE)
        abi.check(self._lib.bp_batch_step_worlds(self._handle, abi.ptr(w), len(w), steps, abi.ptr(infos)),
                  "bp_batch_step_worlds")
        return infos.reshape(steps, len(w))

    def submit_worlds(self, worlds, steps: int = 1):
        """Queues a step of the listed worlds and returns at once (collect it with poll())."""
        self._pending = np.ascontiguousarray(worlds, np.uint32)
        self._pending_steps = steps
        abi.check(self._lib.bp_batch_submit_worlds(self._handle, abi.ptr(self._pending), len(self._pending), steps),
                  "bp_batch_submit_worlds")

    def poll(self):
        """None while the submitted step runs, then its FrameInfo records (steps, len(worlds))."""
        infos = np.zeros(self._pending_steps * len(self._pending), abi.FRAME_DTYPE)
        status = self._lib.bp_batch_poll(self._handle, abi.ptr(infos))
        if status == abi.BP_ERR_FRAME_IN_FLIGHT:
            return None
        abi.check(status, "bp_batch_poll")
        return infos.reshape(self._pending_steps, len(self._pending))

    def write_bubbles(self, world: int, bubbles: np.ndarray):
        """Replaces a world's bubbles between steps (a transaction)."""
        b = np.ascontiguousarray(bubbles, abi.BUBBLE_DTYPE)
        abi.check(self._lib.bp_batch_write_bubbles(self._handle, world, abi.ptr(b), len(b)), "bp_batch_write_bubbles")
        self._counts[world] = len(b)

    def set_resident(self, world: int, resident: bool):
        abi.check(self._lib.bp_batch_set_resident(self._handle, world, 1 if resident else 0), "bp_batch_set_resident")

    def footprint(self, world: int) -> int:
        out = ctypes.c_uint64(0)
        abi.check(self._lib.bp_batch_world_footprint(self._handle, world, ctypes.byref(out)), "bp_batch_world_footprint")
        return int(out.value)

    @property
    def last_gpu_milliseconds(self) -> float:
        return float(self._lib.bp_batch_last_gpu_milliseconds(self._handle))

    def close(self):
        if self._handle:
            self._lib.bp_batch_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
