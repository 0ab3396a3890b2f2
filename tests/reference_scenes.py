"""World-level scenes of the reference's runtime tests, as plain worlds for the oracle and the GPU.

The reference builds these worlds through its public API (ForceScene,
Tests/BubblePhysicsRuntimeTests/ForceSceneSupport.swift:10-150; CrowdScene,
CrowdSceneSupport.swift:36-80; the worlds of RegionPolicyTests.swift and StaticPolygonEdgeTests.swift)
and steps them with BubbleWorld.advance. Here the same objects are mapped the way
NativePublicWorld.step maps them onto the solver input (NativePublicWorld.swift:162-258):

  * bubbles sorted by id; role .static -> inverse mass 0 and the static flag; roles .dynamic and
    .motorDriven are force-driven (NativePublicWorld.swift:88-93); .wrap sets the wrap flag;
  * a .motorDriven bubble gets a motor record with inertia m r^2 / 2 (NativeForceRecords.swift:198-202);
  * static polygons become one-sided edge segments (NativeStaticPolygonEdges), segment ids counting
    from 1 over the static polygons in id order, each segment carrying its polygon's filter;
  * static circles and capsules and kinematic bodies of any shape become body obstacles;
  * the global force is the configuration's global force, the region the simulation region,
    and the region pass runs at the end of every step (the public path).

`Runner` steps a world one frame at a time the way BubbleWorld does: it may change motor drives
before a frame, and it deletes the bubbles a frame removed from the region before the next frame
(NativePublicWorld.swift:264-267). `OracleRunner` steps the oracle, `GpuRunner` the HIP library.
"""
from __future__ import annotations

import math

import numpy as np

from bubblephysics_amd import abi, scenes

FORCE_SCENE_STIFFNESS = 229.0
ALL = 0xFFFFFFFF


def game_mass(n: float) -> float:
    """ForceScene.gameMass: m = (n / 2)^(1 / 2.25) (ForceSceneSupport.swift:27)."""
    return float(np.float32((n / 2.0) ** (1.0 / 2.25)))


class World:
    """Objects of one world in the reference's vocabulary, mapped onto solver records by build()."""

    def __init__(self, *, half: float | None = 2000.0, region=None, dt: float = 1.0 / 60.0,
                 stiffness: float = FORCE_SCENE_STIFFNESS, damping: float = 0.0, apply_region: int = 1):
        self.region = tuple(region) if region is not None else (-half, -half, half, half)
        self.dt = dt
        self.stiffness = stiffness
        self.damping = damping
        self.apply_region = apply_region
        self.bubbles = []
        self.static_polygons = []          # (id, world points, category, mask)
        self.bodies = scenes.BodyBuilder()
        self.global_force = (0.0, 0.0)
        self.motors = {}                   # bubble id -> (fx, fy, torque)
        self.native = {}                   # configuration switches (NativeMetalConfiguration)

    # -- objects -------------------------------------------------------------------------------
    def bubble(self, id, at, *, radius=24.0, mass=1.0, velocity=(0.0, 0.0), role="dynamic", category=1, mask=ALL,
               wrap=False, stiffness=None, damping=None):
        self.bubbles.append(dict(id=id, at=at, radius=radius, mass=mass, velocity=velocity, role=role,
                                 category=category, mask=mask, wrap=wrap,
                                 stiffness=self.stiffness if stiffness is None else stiffness,
                                 damping=self.damping if damping is None else damping))
        return self

    def motor(self, id, force, torque=0.0):
        """setMotorDrive: the bubble becomes .motorDriven with this drive."""
        self.motors[id] = (float(force[0]), float(force[1]), float(torque))
        for b in self.bubbles:
            if b["id"] == id:
                b["role"] = "motor"
        return self

    def static_polygon(self, id, points, category=1, mask=ALL):
        self.static_polygons.append((id, [tuple(map(float, p)) for p in points], category, mask))
        return self

    def wall(self, id, x, thickness=1.0, half_height=1900.0):
        """ForceScene.wall: x in [x, x + thickness], y in [-half_height, half_height] (:64-70)."""
        return self.static_polygon(id, [(x, -half_height), (x + thickness, -half_height), (x + thickness, half_height),
                                        (x, half_height)])

    def square(self, id, at, side, *, role="static", rotation=0.0, velocity=(0.0, 0.0), angular=0.0, category=1,
               mask=ALL):
        """ForceScene.square: a square of `side` centred at `at` (:73-79)."""
        h = side / 2.0
        local = [(-h, -h), (h, -h), (h, h), (-h, h)]
        if role == "static":
            c, s = math.cos(rotation), math.sin(rotation)
            pts = [(x * c - y * s + at[0], x * s + y * c + at[1]) for x, y in local]
            return self.static_polygon(id, np.asarray(pts, np.float32).tolist(), category, mask)
        self.bodies.polygon(id, at, rotation, (velocity[0], velocity[1]), local, flags=abi.BP_BODY_KINEMATIC)
        self._body_extra(id, angular, category, mask)
        return self

    def circle(self, id, at, radius, *, role="static", velocity=(0.0, 0.0), category=1, mask=ALL):
        flags = abi.BP_BODY_STATIC if role == "static" else abi.BP_BODY_KINEMATIC
        self.bodies.circle(id, at, velocity, radius, flags=flags)
        self._body_extra(id, 0.0, category, mask)
        return self

    def capsule(self, id, at, half_length, radius, *, rotation=0.0, role="static", velocity=(0.0, 0.0), category=1,
                mask=ALL):
        flags = abi.BP_BODY_STATIC if role == "static" else abi.BP_BODY_KINEMATIC
        self.bodies.capsule(id, at, rotation, velocity, half_length, radius, flags=flags)
        self._body_extra(id, 0.0, category, mask)
        return self

    def _body_extra(self, id, angular, category, mask):
        rec = self.bodies.records[-1]
        rec["velocity"][2] = np.float32(angular)
        rec["category"] = category
        rec["mask"] = mask

    def set_global_force(self, f):
        self.global_force = (float(f[0]), float(f[1]))
        return self

    # -- solver input --------------------------------------------------------------------------
    def build(self) -> scenes.Scene:
        cfg = scenes.make_config(self.region, dt=self.dt, apply_region=self.apply_region)
        cfg.global_force[:] = list(self.global_force)
        for key, value in self.native.items():
            if key == "injection_position":
                cfg.injection_position[:] = list(value)
            else:
                setattr(cfg, key, value)
        specs = sorted(self.bubbles, key=lambda b: b["id"])
        b = np.zeros(len(specs), abi.BUBBLE_DTYPE)
        for k, s in enumerate(specs):
            b[k]["id"] = s["id"]
            b[k]["center"] = s["at"]
            b[k]["previous_center"] = s["at"]
            b[k]["velocity"] = s["velocity"]
            b[k]["mass"] = np.float32(s["mass"])
            b[k]["inverse_mass"] = 0.0 if s["role"] == "static" else np.float32(1.0) / np.float32(s["mass"])
            b[k]["radius"] = s["radius"]
            b[k]["contact_stiffness"] = s["stiffness"]
            b[k]["damping"] = s["damping"]
            flags = abi.BP_BUBBLE_ENABLED
            if s["role"] == "static":
                flags |= abi.BP_BUBBLE_STATIC
            if s["role"] in ("dynamic", "motor"):
                flags |= abi.BP_BUBBLE_FORCE_DRIVEN
            if s["wrap"]:
                flags |= abi.BP_BUBBLE_WRAP
            b[k]["flags"] = flags
            b[k]["category"] = s["category"]
            b[k]["mask"] = s["mask"]
        scene = scenes.Scene("reference-scene", cfg, b)
        segs, next_id = [], 1
        for pid, pts, cat, mask in sorted(self.static_polygons, key=lambda p: p[0]):
            seg = scenes.static_polygon_edges(pts, next_id, pid)
            seg["category"] = cat
            seg["mask"] = mask
            next_id += len(seg)
            segs.append(seg)
        if segs:
            scene.segments = np.concatenate(segs)
        if self.bodies.records:
            scene.bodies, scene.vertices, scene.triangles = self.bodies.build()
        scene.motors = motor_records(b, self.motors)
        return scene


def motor_records(bubbles: np.ndarray, motors: dict) -> np.ndarray:
    """Motor records of the .motorDriven bubbles, ids ascending (NativeForceRecords.swift:198-202)."""
    out = np.zeros(len(motors), abi.MOTOR_DTYPE)
    for k, bid in enumerate(sorted(motors)):
        row = np.nonzero(bubbles["id"] == bid)[0]
        fx, fy, torque = motors[bid]
        out[k]["bubble_id"] = bid
        out[k]["force"] = (fx, fy)
        out[k]["torque"] = torque
        if len(row):
            m, r = np.float32(bubbles[row[0]]["mass"]), np.float32(bubbles[row[0]]["radius"])
            out[k]["inertia"] = np.float32(0.5) * m * r * r
    return out


def injecting(world: World, object_id: int, position) -> World:
    """NativeMetalConfiguration with assertEverySubstep and a state injected after substep 0 of every
    frame (ContactAssertionTests.swift:11-16)."""
    world.native.update(assert_every_substep=1, injection_enabled=1, injection_after_substep=0,
                        injection_object=int(object_id), injection_position=(float(position[0]), float(position[1])))
    return world


class Runner:
    """Frame-by-frame stepping with the public world's bookkeeping. Subclasses implement _step."""

    def __init__(self, scene: scenes.Scene):
        self.scene = scene.copy()
        self.motors = {}
        for m in self.scene.motors:
            self.motors[int(m["bubble_id"])] = (float(m["force"][0]), float(m["force"][1]), float(m["torque"]))
        self.frames = []

    def set_motor(self, bubble_id, force, torque=0.0):
        self.motors[int(bubble_id)] = (float(force[0]), float(force[1]), float(torque))

    def position(self, bubble_id):
        row = np.nonzero(self.scene.bubbles["id"] == bubble_id)[0]
        return None if not len(row) else self.scene.bubbles[row[0]]["center"].astype(np.float64)

    def velocity(self, bubble_id):
        row = np.nonzero(self.scene.bubbles["id"] == bubble_id)[0]
        return None if not len(row) else self.scene.bubbles[row[0]]["velocity"].astype(np.float64)

    def body_pose(self, body_id):
        row = np.nonzero(self.scene.bodies["id"] == body_id)[0]
        return self.scene.bodies[row[0]]["pose"].astype(np.float64)

    def step(self):
        """One frame; returns a dict with the status, the region results (ids) and, when captured,
        the per-substep states. After a failed frame the world keeps its last published state."""
        self.scene.motors = motor_records(self.scene.bubbles, self.motors)
        result = self._step()
        self.frames.append(result)
        if result["status"] == abi.BP_OK:
            self.scene.bubbles = result["bubbles"]
            self.scene.bodies = result["bodies"]
            removed = result["removed"]
            if len(removed):
                keep = ~np.isin(self.scene.bubbles["id"], removed)
                self.scene.bubbles = self.scene.bubbles[keep].copy()
                for bid in removed:
                    self.motors.pop(int(bid), None)
        return result

    def run(self, frames, until=None, before=None):
        """Steps until `frames` frames, the first failing frame, or until(result) is true."""
        out = []
        for k in range(frames):
            if before is not None:
                before(k, self)
            r = self.step()
            out.append(r)
            if r["status"] != abi.BP_OK or (until is not None and until(r)):
                break
        return out


class OracleRunner(Runner):
    def __init__(self, scene, oracle, *, capture=False):
        super().__init__(scene)
        self.oracle = oracle
        self.capture = capture

    def _step(self):
        cb, cbod, cf, ex = self.oracle.run(self.scene, 1, region=True, capture=self.capture)
        f = cf[0]
        region = ex["region"]
        return {"status": int(f["status"]), "frame": f, "bubbles": cb, "bodies": cbod,
                "removed": cb["id"][region == abi.BP_REGION_REMOVED].copy(),
                "wrapped": cb["id"][region == abi.BP_REGION_WRAPPED].copy(),
                "substeps": ex.get("substeps"), "ids": cb["id"].copy(),
                "position": {int(i): cb[k]["center"].astype(np.float64) for k, i in enumerate(cb["id"])},
                "velocity": {int(i): cb[k]["velocity"].astype(np.float64) for k, i in enumerate(cb["id"])}}


class GpuRunner(Runner):
    """The HIP library through the C ABI (bp_world); the world is uploaded again only when the
    public world deletes bubbles or changes a drive."""

    def __init__(self, scene):
        from bubblephysics_amd import solver
        super().__init__(scene)
        self.world = solver.NativeWorld(self.scene)
        self._uploaded_motors = dict(self.motors)

    def _step(self):
        if self.motors != self._uploaded_motors:
            self.world.set_forces(motor_records(self.scene.bubbles, self.motors), self.scene.bands)
            self._uploaded_motors = dict(self.motors)
        frame = self.world.step(len(self.frames))
        bubbles = frame.bubbles.copy()
        bodies = self.world.read_bodies() if len(self.scene.bodies) else self.scene.bodies.copy()
        removed = np.array(sorted(frame.region_removals), np.int64)
        wrapped = np.array(sorted(frame.region_wraps), np.int64)
        if frame.status == abi.BP_OK and len(removed):
            self.world.remove_bubbles(removed)
            self._uploaded_motors = {k: v for k, v in self.motors.items() if k not in set(removed.tolist())}
        return {"status": int(frame.status), "frame": frame, "bubbles": bubbles, "bodies": bodies, "removed": removed,
                "wrapped": wrapped, "substeps": None, "ids": bubbles["id"].copy(),
                "position": {int(i): bubbles[k]["center"].astype(np.float64) for k, i in enumerate(bubbles["id"])},
                "velocity": {int(i): bubbles[k]["velocity"].astype(np.float64) for k, i in enumerate(bubbles["id"])}}


class BatchRunner(Runner):
    """The HIP library's batch path (bp_batch) with the world pinned to one kernel class; the batch
    deletes region-removed bubbles itself before the next step."""

    def __init__(self, scene, kernel_class=None):
        from bubblephysics_amd import solver
        super().__init__(scene)
        self.batch = solver.Batch([self.scene])
        if kernel_class is not None:
            self.batch.set_world_class(0, kernel_class)

    def _step(self):
        info = self.batch.step(1)[0, 0]
        bubbles = self.batch.read_bubbles(0)
        bodies = self.batch.read_bodies(0, len(self.scene.bodies)) if len(self.scene.bodies) else self.scene.bodies.copy()
        removed, wrapped = self.batch.region_results(0)
        return {"status": int(info["status"]), "frame": info, "bubbles": bubbles, "bodies": bodies, "removed": removed,
                "wrapped": wrapped, "substeps": None, "ids": bubbles["id"].copy(), "deleted": True,
                "position": {int(i): bubbles[k]["center"].astype(np.float64) for k, i in enumerate(bubbles["id"])},
                "velocity": {int(i): bubbles[k]["velocity"].astype(np.float64) for k, i in enumerate(bubbles["id"])}}

    def step(self):
        result = self._step()
        self.frames.append(result)
        if result["status"] == abi.BP_OK:
            self.scene.bubbles = result["bubbles"]
            self.scene.bodies = result["bodies"]
        return result


def frame_summary(result):
    """(status, substeps, violation kind, violation objects) of a frame of any runner."""
    f = result["frame"]
    if isinstance(f, np.void):
        if "violation_kind" in f.dtype.names:                # oracle frame
            kind = int(f["violation_kind"]) if int(f["status"]) == abi.BP_ERR_CONTACT_VIOLATION else 0
            objs = tuple(sorted(int(x) for x in f["violation_objects"] if x)) if kind else ()
            return int(f["status"]), int(f["substeps"]), kind, objs
        d = f["detail"]                                       # bp_frame_info record
        kind = int(d["violation_kind"]) if int(f["status"]) == abi.BP_ERR_CONTACT_VIOLATION else 0
        objs = tuple(sorted(int(x) for x in d["object_ids"] if x)) if kind else ()
        return int(f["status"]), int(f["substep_count"]), kind, objs
    kind = f.detail.violation_kind if f.detail is not None and f.status == abi.BP_ERR_CONTACT_VIOLATION else 0
    objs = tuple(sorted(int(x) for x in f.detail.object_ids if x)) if kind else ()
    return int(f.status), int(f.substep_count), int(kind), objs
