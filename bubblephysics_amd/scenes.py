"""Synthetic worlds of the benchmark configurations and of the reference's test scenes.

Every world is plain data (numpy record arrays in the layout of include/bubblephysics.h):
bubbles sorted by id, one-sided wall segments, rigid-body obstacles, motor and band records.
Inputs are seeded (numpy PCG64, seed 20261201 + configuration number) so that the CPU oracle
and the HIP library step identical worlds.

Choices the reference leaves open (kept identical for GPU and CPU):
  * material: K_c = 229, gamma = 0 (the game value of PublicWorldStressBenchmark.swift:684-685);
  * mass m = 0.02 r^2 (PublicWorldStressBenchmark.swift:765 without its floor of 1);
  * gravity g: a .motorDriven bubble with motor force m*g (Model/Motion.swift:38-45), so the
    step computes (m*g)*(1/m) in FP32 exactly as the reference's force path does;
  * walls: four 1-unit static rectangles on the region border (PublicWorldStressBenchmark
    .swift:770-780), which become 16 one-sided segments;
  * rigid bodies: kinematic with constant velocity (dynamic bodies never share a world with
    bubbles, NativeCompleteStepper.swift:71-88).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

BASE_SEED = 20261201
CONTACT_STIFFNESS = 229.0
GRAVITY = -980.0


@dataclass
class Scene:
    """One world: configuration, records and the number of steps it is run for."""
    name: str
    config: abi.WorldConfig
    bubbles: np.ndarray
    segments: np.ndarray = field(default_factory=lambda: np.zeros(0, abi.SEGMENT_DTYPE))
    bodies: np.ndarray = field(default_factory=lambda: np.zeros(0, abi.BODY_DTYPE))
    vertices: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))
    triangles: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.uint32))
    motors: np.ndarray = field(default_factory=lambda: np.zeros(0, abi.MOTOR_DTYPE))
    bands: np.ndarray = field(default_factory=lambda: np.zeros(0, abi.BAND_DTYPE))
    fields: np.ndarray = field(default_factory=lambda: np.zeros(0, abi.FIELD_DTYPE))
    field_vertices: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.float32))
    field_filter_ids: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    steps: int = 1
    placed_without_overlap: int = 0

    def copy(self) -> "Scene":
        c = abi.WorldConfig()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self.config), ctypes.sizeof(c))
        return Scene(self.name, c, self.bubbles.copy(), self.segments.copy(), self.bodies.copy(), self.vertices.copy(),
                     self.triangles.copy(), self.motors.copy(), self.bands.copy(), self.fields.copy(), self.field_vertices.copy(),
                     self.field_filter_ids.copy(), self.steps, self.placed_without_overlap)

    def desc(self) -> abi.BatchWorldDesc:
        """C descriptor pointing at this scene's arrays (the scene must outlive the call)."""
        d = abi.BatchWorldDesc()
        d.config = self.config
        for name in ("bubbles", "segments", "bodies", "vertices", "triangles", "motors", "bands", "fields", "field_vertices",
                     "field_filter_ids"):
            arr = getattr(self, name)
            if not arr.flags["C_CONTIGUOUS"]:
                arr = np.ascontiguousarray(arr)
                setattr(self, name, arr)
            setattr(d, name, abi.ptr(arr) if len(arr) else None)
        d.bubble_count = len(self.bubbles)
        d.segment_count = len(self.segments)
        d.body_count = len(self.bodies)
        d.vertex_count = len(self.vertices)
        d.triangle_count = len(self.triangles)
        d.motor_count = len(self.motors)
        d.band_count = len(self.bands)
        d.field_count = len(self.fields)
        d.field_vertex_count = len(self.field_vertices)
        d.field_filter_count = len(self.field_filter_ids)
        return d


def make_config(region, dt=1.0 / 60.0, apply_region=1) -> abi.WorldConfig:
    cfg = abi.default_config_values()
    cfg.time_step = dt
    cfg.region[:] = [float(v) for v in region]
    cfg.apply_region = apply_region
    return cfg


def make_bubbles(centers, radii, velocities, *, first_id=1, stiffness=CONTACT_STIFFNESS, damping=0.0,
                 flags=abi.BP_BUBBLE_ENABLED | abi.BP_BUBBLE_FORCE_DRIVEN) -> np.ndarray:
    n = len(radii)
    b = np.zeros(n, abi.BUBBLE_DTYPE)
    r = np.asarray(radii, np.float32)
    b["id"] = np.arange(first_id, first_id + n, dtype=np.int64)
    b["center"] = np.asarray(centers, np.float32).reshape(n, 2)
    b["previous_center"] = b["center"]
    b["velocity"] = np.asarray(velocities, np.float32).reshape(n, 2)
    mass = (np.float32(0.02) * r * r).astype(np.float32)
    b["mass"] = mass
    b["inverse_mass"] = (np.float32(1.0) / mass).astype(np.float32)
    b["radius"] = r
    b["contact_stiffness"] = np.float32(stiffness)
    b["damping"] = np.float32(damping)
    b["flags"] = flags
    b["category"] = 1
    b["mask"] = 0xFFFFFFFF
    return b


def gravity_motors(bubbles: np.ndarray, g: float = GRAVITY) -> np.ndarray:
    """Motor records m*g for every bubble (NativeForceInput.init, NativeForceRecords.swift:198-202)."""
    m = np.zeros(len(bubbles), abi.MOTOR_DTYPE)
    m["bubble_id"] = bubbles["id"]
    mass = bubbles["mass"].astype(np.float32)
    m["force"][:, 0] = np.float32(0.0)
    m["force"][:, 1] = (mass * np.float32(g)).astype(np.float32)
    m["torque"] = np.float32(0.0)
    r = bubbles["radius"].astype(np.float32)
    m["inertia"] = (np.float32(0.5) * mass * r * r).astype(np.float32)
    return m


def static_polygon_edges(points, first_id: int, owner_id: int) -> np.ndarray:
    lib = abi.load_library()
    pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 2))
    out = np.zeros(len(pts) + 1, abi.SEGMENT_DTYPE)
    count = ctypes.c_uint32(0)
    abi.check(lib.bp_static_polygon_edges(abi.ptr(pts), len(pts), first_id, owner_id, 1, abi.ptr(out), len(out),
                                          ctypes.byref(count)), "bp_static_polygon_edges")
    return out[: count.value].copy()


def polygon_geometry(local_points):
    """(centroid, centroid-relative vertices, triangles, bounding radius) of a polygon shape."""
    lib = abi.load_library()
    pts = np.ascontiguousarray(np.asarray(local_points, np.float32).reshape(-1, 2))
    n = len(pts)
    centroid = np.zeros(2, np.float32)
    verts = np.zeros((n, 2), np.float32)
    tris = np.zeros((n, 4), np.uint32)
    vc, tc, radius = ctypes.c_uint32(0), ctypes.c_uint32(0), ctypes.c_float(0)
    abi.check(lib.bp_polygon_body_geometry(abi.ptr(pts), n, abi.ptr(centroid), abi.ptr(verts), ctypes.byref(vc), abi.ptr(tris),
                                           ctypes.byref(tc), ctypes.byref(radius)), "bp_polygon_body_geometry")
    return centroid, verts[: vc.value].copy(), tris[: tc.value].copy(), float(radius.value)


class BodyBuilder:
    """Accumulates rigid bodies with their shared vertex and triangle arrays."""

    def __init__(self):
        self.records = []
        self.vertices = []
        self.triangles = []

    def _add(self, body_id, pose, velocity, shape, flags, verts, tris, bounding, core=0.0, half=0.0):
        rec = np.zeros(1, abi.BODY_DTYPE)[0]
        rec["id"] = body_id
        rec["pose"] = np.asarray(pose, np.float32)
        rec["previous"] = np.asarray(pose, np.float32)
        rec["velocity"] = np.asarray(velocity, np.float32)
        rec["bounding_radius"] = bounding
        rec["core_radius"] = core
        rec["half_length"] = half
        rec["shape"] = shape
        rec["flags"] = flags
        rec["category"] = 1
        rec["mask"] = 0xFFFFFFFF
        rec["vertex_offset"] = sum(len(v) for v in self.vertices)
        rec["vertex_count"] = len(verts)
        rec["triangle_offset"] = sum(len(t) for t in self.triangles)
        rec["triangle_count"] = len(tris)
        self.records.append(rec)
        self.vertices.append(np.asarray(verts, np.float32).reshape(-1, 2))
        self.triangles.append(np.asarray(tris, np.uint32).reshape(-1, 4))

    def polygon(self, body_id, center, angle, velocity, local_points, flags=abi.BP_BODY_KINEMATIC):
        _, verts, tris, radius = polygon_geometry(local_points)
        self._add(body_id, (center[0], center[1], angle), (velocity[0], velocity[1], 0.0), abi.BP_SHAPE_POLYGON, flags,
                  verts, tris, radius)

    def static_polygon_world(self, body_id, world_points):
        """Static polygon given in world coordinates: its pose is the centroid, angle 0."""
        centroid, verts, tris, radius = polygon_geometry(world_points)
        self._add(body_id, (centroid[0], centroid[1], 0.0), (0.0, 0.0, 0.0), abi.BP_SHAPE_POLYGON, abi.BP_BODY_STATIC,
                  verts, tris, radius)

    def circle(self, body_id, center, velocity, radius, flags=abi.BP_BODY_KINEMATIC, angle=0.0):
        self._add(body_id, (center[0], center[1], angle), (velocity[0], velocity[1], 0.0), abi.BP_SHAPE_CIRCLE, flags,
                  [(0.0, 0.0)], np.zeros((0, 4), np.uint32), radius, core=radius)

    def capsule(self, body_id, center, angle, velocity, half_length, radius, flags=abi.BP_BODY_KINEMATIC):
        self._add(body_id, (center[0], center[1], angle), (velocity[0], velocity[1], 0.0), abi.BP_SHAPE_CAPSULE, flags,
                  [(-half_length, 0.0), (half_length, 0.0)], np.zeros((0, 4), np.uint32), half_length + radius,
                  core=radius, half=half_length)

    def build(self):
        bodies = np.array(self.records, dtype=abi.BODY_DTYPE) if self.records else np.zeros(0, abi.BODY_DTYPE)
        order = np.argsort(bodies["id"], kind="stable") if len(bodies) else np.zeros(0, np.int64)
        bodies = bodies[order]
        verts = np.concatenate(self.vertices) if self.vertices else np.zeros((0, 2), np.float32)
        tris = np.concatenate(self.triangles) if self.triangles else np.zeros((0, 4), np.uint32)
        return bodies, np.ascontiguousarray(verts, np.float32), np.ascontiguousarray(tris, np.uint32)


def boundary_walls(width: float, height: float, builder: BodyBuilder, first_body_id: int, first_segment_id: int = 1):
    """Four 1-unit static rectangles on the region border -> 16 one-sided segments."""
    walls = [
        [(0, 0), (1, 0), (1, height), (0, height)],
        [(width - 1, 0), (width, 0), (width, height), (width - 1, height)],
        [(0, 0), (width, 0), (width, 1), (0, 1)],
        [(0, height - 1), (width, height - 1), (width, height), (0, height)],
    ]
    segments = []
    next_id = first_segment_id
    for k, pts in enumerate(walls):
        seg = static_polygon_edges(pts, next_id, first_body_id + k)
        next_id += len(seg)
        segments.append(seg)
        builder.static_polygon_world(first_body_id + k, pts)
    return np.concatenate(segments)


def place(radii, region, rng, *, inset=1.0, gap=0.0, attempts=2000, obstacles=None):
    """Random non-overlapping placement (largest radius first); returns centres (NaN = unplaced)."""
    lib = abi.load_library()
    radii = np.asarray(radii, np.float32)
    order = np.argsort(-radii, kind="stable")
    r_sorted = np.ascontiguousarray(radii[order])
    reg = np.asarray(region, np.float32)
    obs = np.zeros((0, 3), np.float32) if obstacles is None else np.ascontiguousarray(np.asarray(obstacles, np.float32))
    out = np.zeros((len(radii), 2), np.float32)
    placed = ctypes.c_uint32(0)
    seed = int(rng.integers(0, 2**63 - 1))
    abi.check(lib.bp_place_discs(abi.ptr(r_sorted), len(r_sorted), abi.ptr(reg), float(inset), float(gap), abi.ptr(obs),
                                 len(obs), seed, attempts, abi.ptr(out), ctypes.byref(placed)), "bp_place_discs")
    centres = np.empty_like(out)
    centres[order] = out
    return centres, placed.value


def _assemble(name, cfg, centres, radii, velocities, *, motors=False, builder=None, segments=None, steps=1, placed=0,
              damping=0.0):
    flags = abi.BP_BUBBLE_ENABLED | abi.BP_BUBBLE_FORCE_DRIVEN
    b = make_bubbles(centres, radii, velocities, damping=damping, flags=flags)
    scene = Scene(name, cfg, b, steps=steps, placed_without_overlap=placed)
    if motors:
        scene.motors = gravity_motors(b)
    if segments is not None:
        scene.segments = segments
    if builder is not None:
        scene.bodies, scene.vertices, scene.triangles = builder.build()
    return scene


def config1(n: int = 256, steps: int = 120, seed_offset: int = 1) -> Scene:
    """Plain world of 256 bubbles with walls and gravity (region [0,375]x[0,700])."""
    rng = np.random.Generator(np.random.PCG64(BASE_SEED + seed_offset))
    width, height = 375.0, 700.0
    radii = rng.uniform(8.0, 24.0, n).astype(np.float32)
    region = (0.0, 0.0, width, height)
    centres, placed = place(radii, region, rng, attempts=20000)
    missing = np.isnan(centres[:, 0])
    if missing.any():
        # The area cannot hold every bubble without overlap: the rest goes in with d >= (r_a + r_b) / 2.
        ok = ~missing
        obs = np.column_stack([centres[ok], radii[ok]]).astype(np.float32)
        extra, _ = place(radii[missing], region, rng, gap=-8.0, attempts=20000, obstacles=obs)
        centres[missing] = extra
        still = np.isnan(centres[:, 0])
        if still.any():
            # Last resort: a jittered lattice in the top band.
            k = int(still.sum())
            xs = np.linspace(30.0, width - 30.0, max(k, 1))
            centres[still, 0] = xs[:k]
            centres[still, 1] = height - 40.0 - rng.uniform(0, 10, k)
    velocities = rng.normal(0.0, 30.0, (n, 2)).astype(np.float32)
    builder = BodyBuilder()
    segments = boundary_walls(width, height, builder, first_body_id=100_001)
    cfg = make_config(region)
    return _assemble("config1-plain-256", cfg, centres, radii, velocities, motors=True, builder=builder,
                     segments=segments, steps=steps, placed=placed)


def _random_bodies(rng, builder, count_poly, count_circle, count_capsule, region, first_id, speed=50.0):
    """Kinematic polygons, circles and capsules placed without overlap; returns (x, y, r) obstacles."""
    width, height = region[2] - region[0], region[3] - region[1]
    shapes = []
    for _ in range(count_poly):
        k = int(rng.integers(3, 9))
        rad = float(rng.uniform(10.0, 40.0))
        angles = np.sort(rng.uniform(0.0, 2.0 * math.pi, k))
        while np.diff(np.concatenate([angles, angles[:1] + 2 * math.pi])).max() >= math.pi:
            angles = np.sort(rng.uniform(0.0, 2.0 * math.pi, k))
        pts = np.column_stack([rad * np.cos(angles), rad * np.sin(angles)]).astype(np.float32)
        # The body sits on its centroid, so its reach is the farthest vertex from the centroid.
        shapes.append(("polygon", polygon_geometry(pts)[3], pts))
    for _ in range(count_circle):
        shapes.append(("circle", float(rng.uniform(8.0, 30.0)), None))
    for _ in range(count_capsule):
        half, r = float(rng.uniform(10.0, 40.0)), float(rng.uniform(4.0, 12.0))
        shapes.append(("capsule", half + r, (half, r)))
    bounds = np.array([s[1] for s in shapes], np.float32)
    centres, _ = place(bounds, region, rng, inset=2.0, gap=4.0, attempts=5000)
    obstacles = []
    for k, (kind, bound, data) in enumerate(shapes):
        c = centres[k]
        if np.isnan(c[0]):
            continue
        v = rng.normal(0.0, speed, 2).astype(np.float32)
        angle = float(rng.uniform(0.0, 2.0 * math.pi))
        bid = first_id + k
        if kind == "polygon":
            builder.polygon(bid, c, angle, v, data)
        elif kind == "circle":
            builder.circle(bid, c, v, bound)
        else:
            builder.capsule(bid, c, angle, v, data[0], data[1])
        obstacles.append((c[0], c[1], bound))
    return np.asarray(obstacles, np.float32).reshape(-1, 3)


def config2(steps: int = 300, seed_offset: int = 2) -> Scene:
    """1024 bubbles (log-uniform radii 4..32) and 64 kinematic bodies in [0,1024]x[0,2048]."""
    rng = np.random.Generator(np.random.PCG64(BASE_SEED + seed_offset))
    width, height = 1024.0, 2048.0
    region = (0.0, 0.0, width, height)
    builder = BodyBuilder()
    segments = boundary_walls(width, height, builder, first_body_id=200_001)
    obstacles = _random_bodies(rng, builder, 24, 20, 20, region, first_id=100_001)
    n = 1024
    radii = np.exp(rng.uniform(math.log(4.0), math.log(32.0), n)).astype(np.float32)
    centres, placed = place(radii, region, rng, attempts=5000, obstacles=obstacles)
    keep = ~np.isnan(centres[:, 0])
    velocities = rng.normal(0.0, 50.0, (n, 2)).astype(np.float32)
    cfg = make_config(region)
    return _assemble("config2-1024-bodies", cfg, centres[keep], radii[keep], velocities[keep], builder=builder,
                     segments=segments, steps=steps, placed=int(keep.sum()))


def config3(steps: int = 100, seed_offset: int = 3, n: int = 262144, bodies: int = 4096, side: float = 16384.0,
            speed: float = 50.0) -> Scene:
    """Large world: 262,144 bubbles (log-uniform radii 2..24) and 4,096 kinematic bodies in 16384^2,
    bubble and body velocities N(0, speed)."""
    rng = np.random.Generator(np.random.PCG64(BASE_SEED + seed_offset))
    region = (0.0, 0.0, side, side)
    builder = BodyBuilder()
    segments = boundary_walls(side, side, builder, first_body_id=10_000_001)
    third = bodies // 3
    obstacles = _random_bodies(rng, builder, bodies - 2 * third, third, third, region, first_id=1_000_001, speed=speed)
    radii = np.exp(rng.uniform(math.log(2.0), math.log(24.0), n)).astype(np.float32)
    centres, placed = place(radii, region, rng, attempts=200, obstacles=obstacles)
    keep = ~np.isnan(centres[:, 0])
    velocities = rng.normal(0.0, speed, (n, 2)).astype(np.float32)
    cfg = make_config(region)
    return _assemble("config3-large", cfg, centres[keep], radii[keep], velocities[keep], builder=builder,
                     segments=segments, steps=steps, placed=int(keep.sum()))


def batch_world(rng, n: int, *, bodies_max: int = 16, radius=(6.0, 20.0), packing=(0.4, 0.7), name="batch", steps=200,
                gravity=False) -> Scene:
    """One world of the batched configuration: square region sized for a packing fraction."""
    radii = rng.uniform(radius[0], radius[1], n).astype(np.float32)
    phi = float(rng.uniform(*packing))
    area = float(np.sum(np.pi * radii.astype(np.float64) ** 2)) / phi
    side = float(math.ceil(math.sqrt(area))) + 2.0
    region = (0.0, 0.0, side, side)
    builder = BodyBuilder()
    segments = boundary_walls(side, side, builder, first_body_id=2_000_000_001)
    nb = int(rng.integers(0, bodies_max + 1))
    obstacles = np.zeros((0, 3), np.float32)
    if nb:
        npoly = nb // 3 + nb % 3
        obstacles = _random_bodies(rng, builder, npoly, nb // 3, nb // 3, region, first_id=1_000_000_001)
    centres, placed = place(radii, region, rng, attempts=300, obstacles=obstacles)
    missing = np.isnan(centres[:, 0])
    if missing.any():
        ok = ~missing
        obs = np.concatenate([obstacles, np.column_stack([centres[ok], radii[ok]])]).astype(np.float32)
        extra, _ = place(radii[missing], region, rng, gap=-0.5 * float(radius[0]) * 2, attempts=2000, obstacles=obs)
        centres[missing] = extra
    keep = ~np.isnan(centres[:, 0])
    velocities = rng.normal(0.0, 30.0, (n, 2)).astype(np.float32)
    cfg = make_config(region)
    return _assemble(name, cfg, centres[keep], radii[keep], velocities[keep], motors=gravity, builder=builder,
                     segments=segments, steps=steps, placed=int((~missing).sum()))


def config4(worlds: int = 4096, steps: int = 200, seed_offset: int = 4, size=(16, 4096)) -> list[Scene]:
    """4,096 independent worlds with bubble counts log-uniform in [16, 4096]."""
    rng = np.random.Generator(np.random.PCG64(BASE_SEED + seed_offset))
    counts = np.exp(rng.uniform(math.log(size[0]), math.log(size[1]), worlds)).astype(np.int64)
    return [batch_world(rng, int(c), name=f"config4-world{k}", steps=steps) for k, c in enumerate(counts)]


def pair_scene(distance: float, ra=24.0, rb=24.0, ka=229.0, kb=229.0, velocity=(0.0, 0.0), dt=1.0 / 60.0) -> Scene:
    """Two bubbles on the x axis (the pair scenes of NativeMetalContactLawTests.swift:399-441)."""
    cfg = make_config((-1e4, -1e4, 1e4, 1e4), dt=dt, apply_region=0)
    b = make_bubbles([(0.0, 0.0), (distance, 0.0)], [ra, rb], [velocity, (0.0, 0.0)])
    b["contact_stiffness"] = [ka, kb]
    b["mass"] = np.float32(0.36)
    b["inverse_mass"] = np.float32(1.0) / np.float32(0.36)
    return Scene("pair", cfg, b)
