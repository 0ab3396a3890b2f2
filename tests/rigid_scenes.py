"""Rigid worlds without bubbles, restated from the reference's rigid-body tests.

Each case builds the objects of one test of the reference (paths relative to the reference
repository) with the scene builders of Tests/BubblePhysicsMetalRenderTests/Support/
PhysicsSceneBuilders.swift (box: mass 1, moment (w^2 + h^2) / 12; circle and capsule: mass 1,
moment from the shape). A world's global force acts through the runtime's velocity update before
each step (the motion prelude of bp_rigid_config), as BubbleWorld.advance applies it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from bubblephysics_amd import abi, rigid

STEP = 1.0 / 60.0
SPEED = 720.0  # RigidCCDTests: 12 units per step at 1/60 s
J = rigid.joint_desc


@dataclass
class Case:
    name: str
    objects: list
    config: abi.RigidConfig
    steps: int
    joints: list = field(default_factory=list)
    interested: tuple = ()

    def scene(self) -> rigid.RigidScene:
        return rigid.records_for(self.objects, self.config, joints=self.joints)


def cfg(region=(0.0, 0.0, 600.0, 300.0), gravity=None, dt=STEP, iterations=8, candidate_limit=128):
    return rigid.rigid_config(dt=dt, region=region, gravity=gravity, iterations=iterations, candidate_limit=candidate_limit)


# ---- RigidCCDTests.swift ------------------------------------------------------------------
def ccd_floor():            # :19-35
    return Case("ccd-floor", [rigid.box(2, (300, 100), (80, 2), role="static"),
                              rigid.box(1, (300, 164), (10, 2), role="dynamic", velocity=(0, -SPEED))],
                cfg((0, 0, 600, 600), gravity=(0, -400)), 120, interested=(1,))


def ccd_head_on():          # :37-55
    return Case("ccd-head-on", [rigid.box(1, (100, 300), (2, 20), role="dynamic", velocity=(SPEED, 0)),
                                rigid.box(2, (220, 300), (2, 20), role="dynamic", velocity=(-SPEED, 0))],
                cfg((0, 0, 600, 600)), 60, interested=(1,))


def ccd_thin_wall():        # :57-68
    return Case("ccd-thin-wall", [rigid.box(2, (300, 300), (2, 60), role="static"),
                                  rigid.box(1, (120, 300), (2, 20), role="dynamic", velocity=(SPEED, 0))],
                cfg((0, 0, 600, 600)), 60)


def ccd_kinematic():        # :70-85
    return Case("ccd-kinematic", [rigid.box(2, (300, 300), (2, 60), role="kinematic"),
                                  rigid.box(1, (120, 300), (2, 20), role="dynamic", velocity=(SPEED, 0))],
                cfg((0, 0, 600, 600)), 60)


def ccd_round():            # :87-110
    return Case("ccd-round", [rigid.box(2, (300, 300), (2, 60), role="dynamic"),
                              rigid.circle(1, (120, 300), 2, velocity=(SPEED, 0)),
                              rigid.capsule(3, (100, 100), 5, 2, velocity=(SPEED, 0)),
                              rigid.capsule(4, (220, 100), 5, 2, velocity=(-SPEED, 0))],
                cfg((0, 0, 600, 600)), 60)


# ---- RigidShapeContactTests.swift -----------------------------------------------------------
HALF_WIDTH = {"box": 10.0, "circle": 9.0, "capsule": 14.0}


def shape(kind, id, at, role, velocity=(0.0, 0.0), taller=False):
    if kind == "box":
        return rigid.box(id, at, (10, 9 if taller else 7), role=role, velocity=velocity)
    if kind == "circle":
        return rigid.circle(id, at, 9, role=role, velocity=velocity)
    return rigid.capsule(id, at, 8, 6, role=role, velocity=velocity)


def shape_against_static(mover, target):      # :47-84
    tx = 150.0
    sx = tx - HALF_WIDTH[target] - HALF_WIDTH[mover] - 12.4
    return Case(f"static-{mover}-{target}", [shape(target, 2, (tx, 100), "static", taller=True),
                                             shape(mover, 1, (sx, 100), "dynamic", velocity=(60, 0))],
                cfg(), 90, interested=(1,))


def shape_separation(mover, target):          # :88-112
    tx = 150.0
    sx = tx - HALF_WIDTH[target] - HALF_WIDTH[mover] - 12
    return Case(f"separate-{mover}-{target}", [shape(target, 2, (tx, 100), "dynamic"),
                                               shape(mover, 1, (sx, 100), "dynamic", velocity=(60, 0))],
                cfg(), 60)


RIGID_SHAPES = ("box", "circle", "capsule")
SEPARATION_PAIRS = [(a, b) for i, a in enumerate(RIGID_SHAPES) for b in RIGID_SHAPES[i:]]


# ---- JointWorldTests.swift ------------------------------------------------------------------
JOINT_REGION = (-1000.0, -1000.0, 2000.0, 2000.0)
G = 400.0
PIVOT = (300.0, 400.0)


def pendulum():             # :28-48
    length, theta = 100.0, 0.1
    start = (PIVOT[0] + length * math.sin(theta), PIVOT[1] - length * math.cos(theta))
    return Case("pendulum", [rigid.circle(1, start, 5)], cfg(JOINT_REGION, gravity=(0, -G)), 420,
                joints=[J(1, 1, None, anchor_b=PIVOT, length=length)])


def rigid_distance():       # :51-64
    return Case("rigid-distance", [rigid.circle(1, (PIVOT[0] + 100, PIVOT[1]), 5)], cfg(JOINT_REGION, gravity=(0, -G)), 240,
                joints=[J(1, 1, None, anchor_b=PIVOT, length=100.0)])


def soft_distance():        # :67-85
    return Case("soft-distance", [rigid.circle(1, (PIVOT[0] + 130, PIVOT[1]), 5)], cfg(JOINT_REGION), 150,
                joints=[J(1, 1, None, anchor_b=PIVOT, length=100.0, softness=0.01, damping=2)])


HINGE_PIVOT = (200.0, 300.0)


def hinge_limits():         # :90-110
    return Case("hinge-limits", [rigid.box(1, (HINGE_PIVOT[0] + 20, HINGE_PIVOT[1]), (20, 2), role="dynamic")],
                cfg(JOINT_REGION, gravity=(0, -G)), 240,
                joints=[J(1, 1, None, kind=abi.BP_JOINT_HINGE, anchor_a=(-20, 0), anchor_b=HINGE_PIVOT, limits=(-0.5, 0.5))])


def hinge_motor():          # :113-123
    return Case("hinge-motor", [rigid.box(1, HINGE_PIVOT, (20, 4), role="dynamic")], cfg(JOINT_REGION), 90,
                joints=[J(1, 1, None, kind=abi.BP_JOINT_HINGE, anchor_b=HINGE_PIVOT, motor=(2, 100_000))])


def hinge_motor_between(movable_is_first):    # :126-143
    a, b = (2, 1) if movable_is_first else (1, 2)
    return Case(f"hinge-motor-{'first' if movable_is_first else 'second'}",
                [rigid.box(1, HINGE_PIVOT, (30, 30), role="static"), rigid.box(2, HINGE_PIVOT, (20, 4), role="dynamic")],
                cfg(JOINT_REGION), 60, joints=[J(1, a, b, kind=abi.BP_JOINT_HINGE, motor=(2, 100_000))])


def hinge_motor_load(max_torque):             # :146-163
    return Case(f"hinge-load-{int(max_torque)}", [rigid.box(1, (HINGE_PIVOT[0] + 30, HINGE_PIVOT[1]), (30, 2), role="dynamic")],
                cfg(JOINT_REGION, gravity=(0, -G)), 60,
                joints=[J(1, 1, None, kind=abi.BP_JOINT_HINGE, anchor_a=(-30, 0), anchor_b=HINGE_PIVOT,
                          motor=(1, max_torque))])


def weld():                 # :166-189 (the impulse [0, 150] on body 2 before the first step)
    objs = [rigid.box(1, (200, 300), (20, 5), role="dynamic"), rigid.box(2, (240, 300), (20, 5), role="dynamic",
                                                                          velocity=(0.0, 150.0))]
    return Case("weld", objs, cfg((-5000, -5000, 5000, 5000), gravity=(0, -G)), 180,
                joints=[J(1, 1, 2, kind=abi.BP_JOINT_WELD, anchor_a=(20, 0), anchor_b=(-20, 0))])


def connected(collide):     # :222-256
    return Case(f"connected-{'on' if collide else 'off'}",
                [rigid.box(1, (200, 200), (20, 20), role="dynamic"), rigid.box(2, (220, 200), (20, 20), role="dynamic")],
                cfg(JOINT_REGION), 90, interested=(1,),
                joints=[J(1, 1, 2, kind=abi.BP_JOINT_HINGE, anchor_a=(10, 0), anchor_b=(-10, 0), collide_connected=collide)])


def joint_determinism():    # :260-284
    pivot = (300.0, 500.0)
    objs = [rigid.box(1, (200, 300), (20, 5), role="dynamic"), rigid.box(2, (240, 300), (20, 5), role="dynamic"),
            rigid.circle(3, (pivot[0] + 80, pivot[1]), 6),
            rigid.box(5, (600, 200), (30, 3), role="static"), rigid.box(6, (600, 230), (20, 5), role="dynamic")]
    joints = [J(1, 1, 2, kind=abi.BP_JOINT_WELD, anchor_a=(20, 0), anchor_b=(-20, 0)),
              J(2, 3, None, anchor_b=pivot, length=80.0),
              J(4, 6, None, kind=abi.BP_JOINT_HINGE, anchor_a=(-20, 0), anchor_b=(580, 230), motor=(1, 50_000))]
    return Case("joint-mix", objs, cfg(JOINT_REGION, gravity=(0, -G)), 120, joints=joints)


# ---- VelocityLimitTests.swift -----------------------------------------------------------------
LIMIT_REGION = (-5000.0, -5000.0, 5000.0, 5000.0)


def speed_limits():         # :15-37
    lim = dict(max_linear_speed=50.0, max_angular_speed=3.0)
    return Case("speed-limits", [rigid.box(2, (200, 100), (10, 7), role="dynamic", **lim),
                                 rigid.circle(3, (300, 100), 9, angular_velocity=100, **lim),
                                 rigid.capsule(4, (400, 100), 8, 5, angular_velocity=-100, **lim)],
                cfg(LIMIT_REGION, gravity=(0, -100_000)), 30)


def kinematic_speed_limit():  # :40-56
    return Case("kinematic-limit", [rigid.box(1, (0, 100), (10, 7), role="kinematic", velocity=(120.0, 0.0),
                                              max_linear_speed=50.0, max_angular_speed=3.0)], cfg(LIMIT_REGION), 60)


def no_speed_limit():       # :59-67
    return Case("no-limit", [rigid.box(1, (200, 100), (10, 7), role="dynamic")], cfg(LIMIT_REGION, gravity=(0, -100_000)), 5)


# ---- NativeMetalRigidPolygonTests.swift (NativePublicWorld: no global force) -------------------
def unit_square(id, position=(0.0, 0.0), velocity=(0.0, 0.0), angular=0.0, role="dynamic"):
    return rigid.RigidObject(id=id, shape="polygon", vertices=[(-1, -1), (1, -1), (1, 1), (-1, 1)], position=position,
                             velocity=velocity, angular_velocity=angular, mass=2.0, moment=2.0, role=role)


POLY_REGION = (-100.0, -100.0, 100.0, 100.0)


def motor_polygon():        # :62-68 (dt 0.25)
    return Case("motor-polygon", [unit_square(2**63 - 1, velocity=(2, 0), angular=1)], cfg(POLY_REGION, dt=0.25), 1)


def polygon_against_wall():  # :47-60
    wall = rigid.RigidObject(id=3, shape="polygon", vertices=[(2, -10), (2.25, -10), (2.25, 10), (2, 10)], role="static")
    return Case("polygon-wall", [unit_square(2**63 - 1, velocity=(8, 0), angular=0.5), wall], cfg(POLY_REGION), 60)


def polygons_swap():        # :94-107
    right = rigid.RigidObject(id=3, shape="polygon", vertices=[(-1, -1), (1, -1), (1, 1), (-1, 1)], position=(3, 0),
                              velocity=(-360, 0), mass=2.0, moment=2.0, role="dynamic")
    return Case("polygons-swap", [unit_square(2**63 - 1, position=(-3, 0), velocity=(360, 0)), right], cfg(POLY_REGION), 1)


# ---- region rule (nativeRigidFinish, PostSolve:180-206) --------------------------------------
def region_exit(wrap):
    return Case(f"region-{'wrap' if wrap else 'remove'}",
                [rigid.circle(1, (390, 200), 5, velocity=(1200, 0), wrap=wrap),
                 rigid.box(2, (100, 380), (4, 4), role="dynamic", velocity=(300, 2400), wrap=wrap)],
                cfg((0, 0, 400, 400)), 2)


def mixed_world(seed, bodies=48, statics=6, joints=True):
    """The seeded rigid crowd of bubblephysics_amd.rigid.crowd (the rigid bench workload) as a case."""
    objs, config, js = rigid.crowd(seed, bodies=bodies, statics=statics, joints=joints)
    return Case(f"mixed-{seed}", objs, config, 40, joints=js, interested=(10, 12))


def reference_cases():
    cases = [ccd_floor(), ccd_head_on(), ccd_thin_wall(), ccd_kinematic(), ccd_round()]
    cases += [shape_against_static(m, t) for m in RIGID_SHAPES for t in RIGID_SHAPES]
    cases += [shape_separation(m, t) for m, t in SEPARATION_PAIRS]
    cases += [pendulum(), rigid_distance(), soft_distance(), hinge_limits(), hinge_motor(), hinge_motor_between(False),
              hinge_motor_between(True), hinge_motor_load(6_000), hinge_motor_load(30_000), weld(), connected(False),
              connected(True), joint_determinism(), speed_limits(), kinematic_speed_limit(), no_speed_limit(),
              motor_polygon(), polygon_against_wall(), polygons_swap(), region_exit(False), region_exit(True)]
    return cases


def simulate_oracle(case: Case, steps: int | None = None, plant: int = 0):
    """Steps the oracle one frame at a time; returns per-frame (bodies, contact records, info) and
    the contact events (bp_contact_transitions with the case's objects of interest)."""
    from bubblephysics_amd import solver
    from oracle import bp_oracle
    scene = case.scene()
    frames, events = [], []
    history = None
    for _ in range(case.steps if steps is None else steps):
        bodies, joints, infos, contacts = bp_oracle.run_rigid(scene, 1, plant=plant)
        assert len(infos) == 1
        if infos[0]["status"] != abi.BP_OK:
            frames.append((bodies, contacts, infos[0]))
            break
        scene.bodies = bodies
        frames.append((bodies, contacts, infos[0]))
        if case.interested or history is not None:
            ev, history = solver.contact_transitions(contacts, scene.segments, case.interested, case.config.time_step, history)
            events.extend(ev)
    return frames, events


def positions(bodies):
    return rigid.public_state(bodies)[0]
