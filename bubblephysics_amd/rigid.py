"""Rigid solver for worlds without bubbles: the host mirror of the reference's rigid-body frame.

``RigidObject`` is a public rigid body (RigidBodyDescriptor with its material);
``records_for`` stages a list of them into solver records the way NativeRigidPolygonState.prepare
does (NativeRigidPolygonState.swift:193-301): centre-of-mass poses from the geometry's centroid,
the kinematic centre-of-mass velocity, inverse mass and moment, the material's friction,
restitution and speed limits, the filter and the shape. ``public_state`` maps records back to
object positions as NativeRigidPolygonState.completed does (:326-352). Static polygons also become
one-sided solver segments (NativePublicWorld.swift:218-238), which the prediction's swept test
reads. ``RigidBatch`` drives ``bp_rigid_batch_*`` of the HIP library (no CPU fallback).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

ROLES = ("static", "dynamic", "kinematic")


def _f32(x) -> np.float32:
    return np.float32(x)


def _rot(p, angle):
    """Rotation in FP32 (host mapping of the public world; the solver never calls it)."""
    c, s = np.cos(_f32(angle), dtype=np.float32), np.sin(_f32(angle), dtype=np.float32)
    x, y = _f32(p[0]), _f32(p[1])
    return (np.float32(x * c - y * s), np.float32(x * s + y * c))


@dataclass
class RigidObject:
    """A rigid body as the public model states it (RigidBodyDescriptor + PhysicsMaterial)."""
    id: int
    shape: str = "polygon"               # polygon | circle | capsule
    vertices: list | None = None          # polygon: local vertices (geometry frame)
    radius: float = 0.0                   # circle and capsule
    half_length: float = 0.0              # capsule (axis along local x)
    position: tuple = (0.0, 0.0)
    rotation: float = 0.0
    velocity: tuple = (0.0, 0.0)
    angular_velocity: float = 0.0
    mass: float = 1.0
    moment: float | None = None           # None: from the shape (RigidShape.momentOfInertia)
    role: str = "static"
    wrap: bool = False                    # regionExitPolicy .wrap
    damping: float = 0.0
    friction: float = 0.0
    restitution: float = 0.0
    max_linear_speed: float = 0.0
    max_angular_speed: float = 0.0
    category: int = 1
    mask: int = 0xFFFFFFFF
    layer: int = 0
    previous_position: tuple | None = None


def box(id, center, half, role="static", velocity=(0.0, 0.0), rotation=0.0, **kw) -> RigidObject:
    """SceneBuilders.box: a 2·half box of mass 1 and moment (w² + h²) / 12."""
    w, h = 2.0 * half[0], 2.0 * half[1]
    verts = [(-half[0], -half[1]), (half[0], -half[1]), (half[0], half[1]), (-half[0], half[1])]
    kw.setdefault("moment", (w * w + h * h) / 12.0)
    return RigidObject(id=id, shape="polygon", vertices=verts, position=tuple(center), rotation=rotation,
                       velocity=tuple(velocity), role=role, **kw)


def circle(id, center, radius, role="dynamic", velocity=(0.0, 0.0), **kw) -> RigidObject:
    return RigidObject(id=id, shape="circle", radius=radius, position=tuple(center), velocity=tuple(velocity),
                       role=role, **kw)


def capsule(id, center, half_length, radius, role="dynamic", velocity=(0.0, 0.0), rotation=0.0, **kw) -> RigidObject:
    return RigidObject(id=id, shape="capsule", radius=radius, half_length=half_length, position=tuple(center),
                       rotation=rotation, velocity=tuple(velocity), role=role, **kw)


def shape_moment(o: RigidObject) -> float:
    """RigidShape.momentOfInertia(mass:) in Double, rounded to Float (Objects.swift:181-220)."""
    m = float(o.mass)
    if o.shape == "circle":
        r = float(np.float32(o.radius))
        return float(np.float32(m * r * r / 2))
    if o.shape == "capsule":
        h, r = float(np.float32(o.half_length)), float(np.float32(o.radius))
        rect_area, disc_area = 4 * h * r, math.pi * r * r
        rect_mass = m * rect_area / (rect_area + disc_area)
        disc_mass = m - rect_mass
        rect = rect_mass * (4 * h * h + 4 * r * r) / 12
        caps = disc_mass * (r * r / 2 + h * h + 8 * h * r / (3 * math.pi))
        return float(np.float32(rect + caps))
    vs = [(np.float32(x), np.float32(y)) for x, y in o.vertices]
    ox, oy = vs[0]
    pts = [(float(np.float32(x - ox)), float(np.float32(y - oy))) for x, y in vs]
    area = second = 0.0
    cx = cy = 0.0
    for i in range(len(pts)):
        a, b = pts[i], pts[(i + 1) % len(pts)]
        cr = a[0] * b[1] - a[1] * b[0]
        area += cr
        cx += (a[0] + b[0]) * cr
        cy += (a[1] + b[1]) * cr
        second += cr * (a[0] * a[0] + a[0] * b[0] + b[0] * b[0] + a[1] * a[1] + a[1] * b[1] + b[1] * b[1])
    area /= 2
    cx /= 6 * area
    cy /= 6 * area
    per_mass = (second / 12) / area
    return float(np.float32(m * (per_mass - (cx * cx + cy * cy))))


def body_geometry(o: RigidObject):
    """NativeRigidGeometry.init (NativeRigidPolygonState.swift:65-133): (vertices, triangles,
    centroid, radius, kind, core radius, half length)."""
    if o.shape == "circle":
        return [(0.0, 0.0)], [], (0.0, 0.0), float(np.float32(o.radius)), abi.BP_SHAPE_CIRCLE, o.radius, 0.0
    if o.shape == "capsule":
        h, r = np.float32(o.half_length), np.float32(o.radius)
        return [(-h, 0.0), (h, 0.0)], [], (0.0, 0.0), float(np.float32(h + r)), abi.BP_SHAPE_CAPSULE, r, h
    lib = abi.load_library()
    pts = np.asarray(o.vertices, dtype=np.float32).reshape(-1, 2)
    n = len(pts)
    centroid = np.zeros(2, np.float32)
    verts = np.zeros((n, 2), np.float32)
    tris = np.zeros((n, 4), np.uint32)
    nv, nt, radius = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_float()
    abi.check(lib.bp_polygon_body_geometry(abi.ptr(pts), n, abi.ptr(centroid), abi.ptr(verts), ctypes.byref(nv),
                                           abi.ptr(tris), ctypes.byref(nt), ctypes.byref(radius)), "bp_polygon_body_geometry")
    return ([tuple(v) for v in verts[:nv.value]], [tuple(t) for t in tris[:nt.value]], (centroid[0], centroid[1]),
            radius.value, abi.BP_SHAPE_POLYGON, 0.0, 0.0)


@dataclass
class RigidScene:
    """Solver input of one rigid world."""
    bodies: np.ndarray
    vertices: np.ndarray          # (n, 2) float32
    triangles: np.ndarray         # (n, 4) uint32
    segments: np.ndarray
    centroids: np.ndarray         # (bodies, 2) float32, the local centroid of each body
    config: abi.RigidConfig
    joints: np.ndarray = field(default_factory=lambda: np.zeros(0, abi.JOINT_DTYPE))
    references: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    ranges: np.ndarray | None = None
    partners: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint64))


def rigid_config(dt=1.0 / 60.0, region=(0.0, 0.0, 600.0, 300.0), candidate_limit=128, iterations=8, gravity=None):
    cfg = abi.RigidConfig()
    cfg.time_step = dt
    cfg.region[:] = region
    cfg.candidate_limit = candidate_limit
    cfg.iterations = iterations
    if gravity is not None:
        cfg.motion_prelude = 1
        cfg.global_force[:] = gravity
    return cfg


def records_for(objects, config: abi.RigidConfig, joints=(), segment_first_id: int = 1) -> RigidScene:
    """Stages public rigid objects (sorted by id here) into solver records as
    NativeRigidPolygonState.prepare does, with the static polygons' one-sided edges."""
    lib = abi.load_library()
    objs = sorted(objects, key=lambda o: o.id)
    dt = np.float32(config.time_step)
    bodies = np.zeros(len(objs), abi.RIGID_BODY_DTYPE)
    verts, tris, cents, segs = [], [], [], []
    next_segment = segment_first_id
    for i, o in enumerate(objs):
        assert o.role in ROLES, o.role
        gv, gt, centroid, radius, kind, core, half = body_geometry(o)
        cx, cy = np.float32(centroid[0]), np.float32(centroid[1])
        offset = _rot((cx, cy), o.rotation)
        static, kinematic = o.role == "static", o.role == "kinematic"
        vx, vy = (np.float32(0), np.float32(0)) if static else (np.float32(o.velocity[0]), np.float32(o.velocity[1]))
        if kinematic:
            nxt = np.float32(np.float32(o.rotation) + np.float32(o.angular_velocity) * dt)
            rotated = _rot((cx, cy), nxt)
            vx = np.float32(vx + np.float32(rotated[0] - offset[0]) / dt)
            vy = np.float32(vy + np.float32(rotated[1] - offset[1]) / dt)
        prev = o.previous_position if o.previous_position is not None else o.position
        b = bodies[i]
        b["id"] = o.id
        b["mass"] = o.mass
        px, py = np.float32(o.position[0]), np.float32(o.position[1])
        b["pose"] = (px + offset[0], py + offset[1], o.rotation, o.rotation)
        b["previous"] = (np.float32(prev[0]) + offset[0], np.float32(prev[1]) + offset[1], cx, cy)
        b["motion"] = (vx, vy, 0.0 if static else o.angular_velocity, 0.0 if kinematic else o.damping)
        moment = o.moment if o.moment is not None else shape_moment(o)
        movable = not (static or kinematic)
        b["physical"] = (np.float32(1) / np.float32(o.mass) if movable else 0.0,
                         np.float32(1) / np.float32(moment) if movable else 0.0, radius, o.friction)
        b["ranges"] = (len(verts), len(gv), len(tris), len(gt))
        b["flags"] = (1 if static else 0, 1 if o.wrap else 0, 0, int(np.float32(o.restitution).view(np.uint32)))
        b["filter"] = (o.category, o.mask, 1 if kinematic else 0, (o.layer & 0xFFFFFFFF) ^ 0x80000000)
        b["shape"] = (kind, core, half, 0.0)
        b["limits"] = (o.max_linear_speed, o.max_angular_speed, 0.0, 0.0)
        verts.extend(gv)
        tris.extend(gt)
        cents.append((cx, cy))
        if static and o.shape == "polygon":
            c, s = np.cos(np.float32(o.rotation), dtype=np.float32), np.sin(np.float32(o.rotation), dtype=np.float32)
            pts = np.array([(np.float32(x) * c - np.float32(y) * s + px, np.float32(x) * s + np.float32(y) * c + py)
                            for x, y in o.vertices], dtype=np.float32)
            out = np.zeros(len(pts) * 2 + 2, abi.SEGMENT_DTYPE)
            count = ctypes.c_uint32()
            abi.check(lib.bp_static_polygon_edges(abi.ptr(pts), len(pts), next_segment, o.id, 1, abi.ptr(out), len(out),
                                                  ctypes.byref(count)), "bp_static_polygon_edges")
            edges = out[:count.value].copy()
            edges["category"] = o.category
            edges["mask"] = o.mask
            segs.append(edges)
            next_segment += count.value
    scene = RigidScene(bodies=bodies, vertices=np.asarray(verts, np.float32).reshape(-1, 2),
                       triangles=np.asarray(tris, np.uint32).reshape(-1, 4),
                       segments=np.concatenate(segs) if segs else np.zeros(0, abi.SEGMENT_DTYPE),
                       centroids=np.asarray(cents, np.float32).reshape(-1, 2), config=config)
    if joints:
        scene.joints, scene.references, scene.ranges, scene.partners = joint_tables(joints, bodies["id"])
    return scene


def joint_desc(id, body_a, body_b=None, kind=abi.BP_JOINT_DISTANCE, anchor_a=(0.0, 0.0), anchor_b=(0.0, 0.0),
               length=0.0, softness=0.0, damping=0.0, reference_angle=0.0, limits=None, motor=None,
               collide_connected=False):
    """A public joint (Joint; SceneBuilders.distanceJoint / hingeJoint / weldJoint)."""
    d = np.zeros(1, abi.JOINT_DESC_DTYPE)[0]
    d["id"], d["body_a"] = id, body_a
    d["has_body_b"] = 0 if body_b is None else 1
    d["body_b"] = 0 if body_b is None else body_b
    d["kind"] = kind
    d["anchor_a"], d["anchor_b"] = anchor_a, anchor_b
    d["length"], d["softness"], d["damping"], d["reference_angle"] = length, softness, damping, reference_angle
    flags = 0
    if limits is not None:
        flags |= abi.BP_JOINT_HAS_LIMITS
        d["lower"], d["upper"] = limits
    if motor is not None:
        flags |= abi.BP_JOINT_HAS_MOTOR
        d["motor_speed"], d["max_torque"] = motor
    d["flags"] = flags
    d["collide_connected"] = 1 if collide_connected else 0
    return d


def joint_tables(joints, body_ids):
    """bp_rigid_joint_tables: (records, references, ranges, partners) of a world's joints."""
    lib = abi.load_library()
    descs = np.array(list(joints), dtype=abi.JOINT_DESC_DTYPE)
    ids = np.ascontiguousarray(body_ids, dtype=np.uint64)
    nj, nb = len(descs), len(ids)
    records = np.zeros(max(nj, 1), abi.JOINT_DTYPE)
    refs = np.zeros(max(2 * nj, 1) * 2, np.uint32)
    ranges = np.zeros(max(nb, 1) * 4, np.uint32)
    partners = np.zeros(max(2 * nj, 1), np.uint64)
    rc, fc, pc = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    abi.check(lib.bp_rigid_joint_tables(abi.ptr(descs), nj, abi.ptr(ids), nb, abi.ptr(records), ctypes.byref(rc),
                                        abi.ptr(refs), ctypes.byref(fc), abi.ptr(ranges), abi.ptr(partners),
                                        ctypes.byref(pc)), "bp_rigid_joint_tables")
    return records[:rc.value].copy(), refs[:2 * fc.value].copy(), ranges[:4 * nb].copy(), partners[:pc.value].copy()


def public_state(bodies: np.ndarray):
    """Object positions from solver records (NativeRigidPolygonState.completed): position =
    pose - rotate(centroid, angle); returns (position (n, 2), rotation, velocity, angular velocity)."""
    pos = np.zeros((len(bodies), 2), np.float32)
    for i, b in enumerate(bodies):
        off = _rot((b["previous"][2], b["previous"][3]), b["pose"][2])
        pos[i] = (np.float32(b["pose"][0]) - off[0], np.float32(b["pose"][1]) - off[1])
    return pos, bodies["pose"][:, 2].copy(), bodies["motion"][:, :2].copy(), bodies["motion"][:, 2].copy()


def crowd(seed, bodies=48, statics=6, joints=True, gravity=(0.0, -300.0)):
    """A seeded crowd of boxes, circles, capsules and convex polygons in a walled 400 x 500 box with a
    static circle and capsule, a kinematic bar and five joints (hinge with limits, soft damped
    distance to the world, weld, colliding distance, motor hinge to the world): the rigid solver's
    bench and parity workload. Returns (objects, config, joints)."""
    rng = np.random.default_rng(seed)
    objs = [box(1, (200, 5), (200, 5), role="static"), box(2, (5, 200), (5, 200), role="static"),
            box(3, (395, 200), (5, 200), role="static"),
            RigidObject(id=4, shape="circle", radius=12.0, position=(200.0, 120.0), role="static"),
            RigidObject(id=5, shape="capsule", radius=6.0, half_length=20.0, position=(120.0, 200.0), rotation=0.4,
                        role="static"),
            box(6, (300, 90), (30, 4), role="kinematic", velocity=(-20, 0))][:statics]
    kinds = ("box", "circle", "capsule", "poly")
    for k in range(bodies):
        x, y = 30 + (k % 8) * 45 + rng.uniform(-4, 4), 160 + (k // 8) * 40 + rng.uniform(-4, 4)
        v = (float(rng.normal(0, 40)), float(rng.normal(0, 40)))
        kind = kinds[k % 4]
        common = dict(velocity=v, angular_velocity=float(rng.normal(0, 2)), friction=float(rng.uniform(0, 0.6)),
                      restitution=float(rng.uniform(0, 0.4)), damping=float(rng.uniform(0, 0.5)), mass=float(rng.uniform(0.5, 2)))
        if kind == "box":
            objs.append(box(10 + k, (x, y), (float(rng.uniform(5, 12)), float(rng.uniform(4, 10))), role="dynamic",
                            rotation=float(rng.uniform(0, 3)), **common))
        elif kind == "circle":
            objs.append(circle(10 + k, (x, y), float(rng.uniform(5, 12)), **common))
        elif kind == "capsule":
            objs.append(capsule(10 + k, (x, y), float(rng.uniform(4, 10)), float(rng.uniform(3, 7)),
                                rotation=float(rng.uniform(0, 3)), **common))
        else:
            m = int(rng.integers(3, 8))
            r = float(rng.uniform(6, 13))
            pts = [(r * math.cos(2 * math.pi * i / m), r * math.sin(2 * math.pi * i / m)) for i in range(m)]
            objs.append(RigidObject(id=10 + k, shape="polygon", vertices=pts, position=(x, y), role="dynamic",
                                    rotation=float(rng.uniform(0, 3)), **common))
    js = []
    if joints and bodies >= 9:
        js = [joint_desc(1, 10, 11, kind=abi.BP_JOINT_HINGE, anchor_a=(8, 0), anchor_b=(-8, 0), limits=(-0.3, 0.3)),
              joint_desc(2, 12, None, anchor_b=(130.0, 330.0), length=60.0, softness=0.001, damping=1.0),
              joint_desc(3, 14, 15, kind=abi.BP_JOINT_WELD, anchor_a=(10, 0), anchor_b=(-10, 0)),
              joint_desc(4, 16, 17, length=45.0, collide_connected=True),
              joint_desc(5, 18, None, kind=abi.BP_JOINT_HINGE, anchor_b=(240.0, 300.0), motor=(1.5, 5_000))]
    return objs, rigid_config(region=(0, 0, 400, 500), gravity=gravity), js


class RigidBatch:
    """Many rigid worlds advanced by one launch per call (bp_rigid_batch_*)."""

    def __init__(self, scenes=()):
        self.lib = abi.load_library()
        h = ctypes.c_void_p()
        abi.check(self.lib.bp_rigid_batch_create(ctypes.byref(h)), "bp_rigid_batch_create")
        self.h = h
        self.counts: list[tuple[int, int]] = []
        for s in scenes:
            self.add(s)

    def close(self):
        if getattr(self, "h", None):
            self.lib.bp_rigid_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _desc(scene: RigidScene):
        keep = []

        def p(a):
            if a is None or a.size == 0:
                return None
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a.ctypes.data

        d = abi.RigidWorldDesc()
        d.config = scene.config
        d.bodies, d.body_count = p(scene.bodies), len(scene.bodies)
        d.vertices, d.vertex_count = p(scene.vertices), len(scene.vertices)
        d.triangles, d.triangle_count = p(scene.triangles), len(scene.triangles)
        d.segments, d.segment_count = p(scene.segments), len(scene.segments)
        d.joints, d.joint_count = p(scene.joints), len(scene.joints)
        d.joint_references, d.joint_reference_count = p(scene.references), len(scene.references) // 2
        d.joint_ranges = p(scene.ranges) if scene.ranges is not None and len(scene.joints) else None
        d.joint_partners, d.joint_partner_count = p(scene.partners), len(scene.partners)
        return d, keep

    def add(self, scene: RigidScene) -> int:
        d, keep = self._desc(scene)
        idx = ctypes.c_uint32()
        abi.check(self.lib.bp_rigid_batch_add_world(self.h, ctypes.byref(d), ctypes.byref(idx)), "bp_rigid_batch_add_world")
        self.counts.append((len(scene.bodies), len(scene.joints)))
        return idx.value

    def set_world(self, world: int, scene: RigidScene):
        d, keep = self._desc(scene)
        abi.check(self.lib.bp_rigid_batch_set_world(self.h, world, ctypes.byref(d)), "bp_rigid_batch_set_world")
        self.counts[world] = (len(scene.bodies), len(scene.joints))

    def step(self, steps: int = 1) -> np.ndarray:
        """Advances every world; returns the frame records [steps, worlds]."""
        infos = np.zeros((steps, len(self.counts)), abi.RIGID_FRAME_DTYPE)
        abi.check(self.lib.bp_rigid_batch_step(self.h, steps, abi.ptr(infos)), "bp_rigid_batch_step")
        return infos

    def read_bodies(self, world: int) -> np.ndarray:
        out = np.zeros(self.counts[world][0], abi.RIGID_BODY_DTYPE)
        abi.check(self.lib.bp_rigid_batch_read_bodies(self.h, world, abi.ptr(out), len(out)), "bp_rigid_batch_read_bodies")
        return out

    def read_joints(self, world: int) -> np.ndarray:
        out = np.zeros(self.counts[world][1], abi.JOINT_DTYPE)
        abi.check(self.lib.bp_rigid_batch_read_joints(self.h, world, abi.ptr(out), len(out)), "bp_rigid_batch_read_joints")
        return out

    def contacts(self, world: int, capacity: int = 4096) -> np.ndarray:
        out = np.zeros(capacity, abi.CONTACT_DTYPE)
        n = ctypes.c_uint32()
        abi.check(self.lib.bp_rigid_batch_contacts(self.h, world, abi.ptr(out), capacity, ctypes.byref(n)),
                  "bp_rigid_batch_contacts")
        return out[:min(n.value, capacity)].copy()

    def last_gpu_ms(self) -> float:
        return float(self.lib.bp_rigid_batch_last_gpu_milliseconds(self.h))
