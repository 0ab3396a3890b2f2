"""Small worlds of the reference's substep-kernel tests, built as whole worlds.

The reference drives single kernels on hand-made buffers (the Rig of
Tests/BubblePhysicsNativeMetalTests/NativeMetalSubstepKernelTests.swift:15-226). Here the same
scenes are complete worlds that the oracle and the HIP library step end to end: bubbles with the
Rig's defaults (mass 0.36, radius 12, K_c 229, no drag, role "force"), the wall of the Rig (the
static-polygon edges of the square (0,0)-(200,200), :234-239, whose bottom edge lies on y = 0
with outward normal (0, -1)), and force records. The broadphase finds the candidate rows that
the Rig lists by hand; bubbles that the Rig keeps apart by giving them no pair rows are kept
apart here with collision filters.
"""
from __future__ import annotations

import numpy as np

from bubblephysics_amd import abi, scenes

SQUARE = [(0.0, 0.0), (200.0, 0.0), (200.0, 200.0), (0.0, 200.0)]
DT = np.float32(1.0 / 60.0)


def body(id, center, velocity=(0.0, 0.0), mass=0.36, radius=12.0, stiffness=229.0, damping=0.0, role="force",
         category=1, mask=0xFFFFFFFF):
    return dict(id=id, center=center, velocity=velocity, mass=mass, radius=radius, stiffness=stiffness, damping=damping,
                role=role, category=category, mask=mask)


def wall_segments():
    """Static-polygon edges of the Rig's square, first id 1, owner 100 (:237)."""
    return scenes.static_polygon_edges(SQUARE, 1, 100)


def field(id, kind, *, position=(0.0, 0.0), rotation=0.0, falloff=0.0, direction=(0.0, 0.0), strength=0.0, radius=None,
          polygon=None, filter_ids=None):
    """A force field as NativeForceInput.appendField builds it (NativeForceRecords.swift:212-239): a
    circle of `radius`, or a polygon whose fade radius is its farthest vertex from the centre."""
    return dict(id=id, kind=kind, position=position, rotation=rotation, falloff=falloff, direction=direction,
                strength=strength, radius=radius, polygon=polygon, filter_ids=filter_ids)


def world(bodies, *, walls=False, global_force=(0.0, 0.0), dt=1.0 / 60.0, split=1, motors=(), bands=(), fields=(),
          minimum_margin_only=0, region=(-1e5, -1e5, 1e5, 1e5)) -> scenes.Scene:
    """A world of Rig bodies. motors: (bubble id, fx, fy, torque, inertia); bands: (bubble id, tx, ty, K, cap);
    fields: records of field()."""
    cfg = scenes.make_config(region, dt=dt, apply_region=0)
    cfg.global_force[:] = [float(v) for v in global_force]
    cfg.substep_split = int(split)
    cfg.minimum_margin_only = int(minimum_margin_only)
    n = len(bodies)
    b = np.zeros(n, abi.BUBBLE_DTYPE)
    for k, s in enumerate(sorted(bodies, key=lambda x: x["id"])):
        b[k]["id"] = s["id"]
        b[k]["center"] = s["center"]
        b[k]["previous_center"] = s["center"]
        b[k]["velocity"] = s["velocity"]
        b[k]["mass"] = np.float32(s["mass"])
        b[k]["inverse_mass"] = 0.0 if s["role"] == "static" else np.float32(1.0) / np.float32(s["mass"])
        b[k]["radius"] = s["radius"]
        b[k]["contact_stiffness"] = s["stiffness"]
        b[k]["damping"] = s["damping"]
        flags = abi.BP_BUBBLE_ENABLED
        if s["role"] == "static":
            flags |= abi.BP_BUBBLE_STATIC
        if s["role"] == "force":
            flags |= abi.BP_BUBBLE_FORCE_DRIVEN
        b[k]["flags"] = flags
        b[k]["category"] = s["category"]
        b[k]["mask"] = s["mask"]
    m = np.zeros(len(motors), abi.MOTOR_DTYPE)
    for k, (bid, fx, fy, torque, inertia) in enumerate(motors):
        m[k]["bubble_id"] = bid
        m[k]["force"] = (fx, fy)
        m[k]["torque"] = torque
        m[k]["inertia"] = inertia
    bd = np.zeros(len(bands), abi.BAND_DTYPE)
    for k, (bid, tx, ty, stiff, cap) in enumerate(bands):
        bd[k]["bubble_id"] = bid
        bd[k]["target"] = (tx, ty)
        bd[k]["stiffness"] = stiff
        bd[k]["maximum_force"] = cap
    f = np.zeros(len(fields), abi.FIELD_DTYPE)
    verts, ids = [], []
    for k, spec in enumerate(fields):
        f[k]["id"] = spec["id"]
        f[k]["kind"] = spec["kind"]
        f[k]["position"] = spec["position"]
        f[k]["rotation"] = spec["rotation"]
        f[k]["falloff"] = spec["falloff"]
        f[k]["direction"] = spec["direction"]
        f[k]["strength"] = spec["strength"]
        if spec["polygon"] is not None:
            pts = np.asarray(spec["polygon"], np.float32)
            f[k]["vertex_offset"], f[k]["vertex_count"] = len(verts), len(pts)
            verts.extend(pts.tolist())
            f[k]["radius"] = float(np.max(np.sqrt((pts.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32))))
        else:
            f[k]["radius"] = spec["radius"]
        if spec["filter_ids"] is not None:
            f[k]["has_filter"] = 1
            f[k]["filter_offset"], f[k]["filter_count"] = len(ids), len(spec["filter_ids"])
            ids.extend(sorted(spec["filter_ids"]))
    scene = scenes.Scene("kernel-test", cfg, b, motors=m, bands=bd, fields=f,
                         field_vertices=np.asarray(verts, np.float32).reshape(-1, 2), field_filter_ids=np.asarray(ids, np.int64))
    if walls:
        scene.segments = wall_segments()
    return scene


def relative(value, expected):
    return abs(float(value) - float(expected)) / max(abs(float(expected)), np.finfo(np.float64).tiny)


PARITY_SCENES = (["wall-stiffness-2.0", "wall-stiffness-0.4", "wall-stiffness-3.0"] + [f"winner-{w}" for w in (0, 31, 64, 100, 129)]
                 + ["pair-gershgorin", "pair-series", "approach", "turning", "ties", "bounce-1000", "pair-bounce", "margin-only",
                    "drag-motor", "band-oscillator", "fields"])


def parity_scene(name):
    """(scene, steps) of one of PARITY_SCENES."""
    for n, scene, steps in parity_scenes():
        if n == name:
            return scene, steps
    raise KeyError(name)


# The scenes below are shared by the oracle tests and the GPU parity test: (name, scene, steps).
def parity_scenes():
    out = []
    for d, mass, k in [(2.0, 0.36, 229.0), (0.4, 38.0, 229.0), (3.0, 1.0, 22900.0)]:
        out.append((f"wall-stiffness-{d}", world([body(7, (100.0, -d), mass=mass, stiffness=k)], walls=True), 3))
    for winner in (0, 31, 64, 100, 129):
        bodies = [body(i + 1, (100.0, -2.0), mass=0.36 if i == winner else 38.0, category=2, mask=1) for i in range(130)]
        out.append((f"winner-{winner}", world(bodies, walls=True), 2))
    out.append(("pair-gershgorin", world([body(5, (0.0, 0.0), mass=1.0), body(9, (3.0, 0.0), mass=4.0)]), 3))
    out.append(("pair-series", world([body(3, (0.0, 0.0), mass=1.0), body(4, (6.0, 0.0), mass=1.0, radius=20.0,
                                                                          stiffness=22900.0)]), 3))
    out.append(("approach", world([body(21, (100.0, -14.0), velocity=(0.0, 1000.0))], walls=True), 3))
    out.append(("turning", world([body(2, (100.0, -14.0), velocity=(0.0, -100.0), mass=1.0)], walls=True,
                                 global_force=(0.0, 1e6)), 2))
    out.append(("ties", world([body(31, (60.0, -14.0), velocity=(0.0, 1000.0)), body(32, (140.0, -14.0), velocity=(0.0, 1000.0)),
                               body(33, (100.0, -14.0), velocity=(0.0, 100.0))], walls=True), 3))
    out.append(("bounce-1000", world([body(1, (100.0, -80.0), velocity=(0.0, 1000.0))], walls=True), 12))
    out.append(("pair-bounce", world([body(1, (0.0, 0.0), velocity=(600.0, 0.0)), body(2, (60.0, 0.0), velocity=(-400.0, 0.0),
                                                                                       mass=1.5)]), 8))
    out.append(("margin-only", world([body(1, (0.0, 0.0), velocity=(3000.0, 0.0))], minimum_margin_only=1), 2))
    out.append(("drag-motor", world([body(1, (0.0, 0.0), velocity=(10.0, -4.0), mass=2.0, damping=2.0),
                                     body(2, (50.0, 0.0), mass=2.0, role="path")],
                                    motors=[(1, 0.0, 0.0, 4.5, 9.0), (2, 0.0, 0.0, 4.5, 9.0)], global_force=(3.0, 0.0)), 4))
    out.append(("band-oscillator", world([body(1, (0.0, 0.0), mass=1.0)], bands=[(1, 50.0, 0.0, 40000.0, 0.0)], split=4), 2))
    out.append(("fields", world([body(1, (10.0, 20.0), mass=0.5), body(2, (40.0, 20.0), mass=2.0),
                                 body(3, (10.0, 60.0), role="path"), body(4, (-30.0, -30.0), role="static"),
                                 body(5, (400.0, 0.0), mass=1.0), body(6, (-20.0, 35.0), mass=1.0)],
                                fields=reference_fields(), global_force=(1.0, -2.0)), 4))
    return out


BOX = [(-50.0, -50.0), (50.0, -50.0), (50.0, 50.0), (-50.0, 50.0)]


def reference_fields():
    """The five fields of NativeMetalSubstepKernelTests.swift:263-270."""
    from bubblephysics_amd.abi import BP_FIELD_DIRECTIONAL as D, BP_FIELD_RADIAL as R, BP_FIELD_VORTEX as V
    return [field(10, D, direction=(30.0, -40.0), radius=100.0, position=(5.0, -5.0), rotation=0.5),
            field(11, R, strength=-50.0, falloff=2.0, radius=80.0),
            field(12, V, strength=10.0, falloff=1.0, polygon=BOX, position=(5.0, 5.0), rotation=0.3),
            field(13, R, strength=7.0, falloff=0.0, radius=200.0, filter_ids=[2]),
            field(14, D, direction=(1000.0, 1000.0), radius=300.0, filter_ids=[])]
