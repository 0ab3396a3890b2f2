"""Pins of the static-polygon edges (bp_static_polygon_edges, NativeStaticPolygonEdges.swift:16-72)
and of the oracle's one-sided edge contact (wall_region, NativeForceKernels.metal-source:112-140)
by the reference's StaticPolygonEdgeTests.swift, restated on the oracle with its scenes and
tolerances: a thin wall deflects a pressed bubble by r eps (eps / (1 - eps)^2 = F / K_c) like a
thick one, a collinear vertex changes nothing, a convex vertex is one spring at any angle and a
bubble slides through it at the law distance, the winding of the vertices does not change the
trajectories, a fast bubble never ends a step inside or behind a wall, and an oblique entry over a
convex corner is pushed out by the crossed edge only.

The worlds follow the test's own helper (:21-45): region +-extent, static polygons with ids 9001,
9002, ..., bubble material K_c = 20 / 0.0108 with drag, one step accepting the scene before the
measured steps, and the motor force set before each step where the test sets it per frame.
"""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi, scenes
import contact_law_reference as law
import reference_scenes as rs

STIFFNESS = float(np.float32(20.0) / np.float32(0.0108))
DT = 1.0 / 60.0


def deflection(radius, force):
    """r eps from eps / (1 - eps)^2 = F / K_c (:16-19, the Double law)."""
    return float(np.float32(radius * law.epsilon_for_ratio(force / STIFFNESS)))


def world(extent, polygons, damping=5.0, reversed_=False):
    w = rs.World(half=extent, stiffness=STIFFNESS, damping=damping)
    for k, pts in enumerate(polygons):
        w.static_polygon(9001 + k, list(reversed(pts)) if reversed_ else pts)
    return w


def rectangle(face_x, thickness, y_min=-400.0, y_max=400.0):
    """:57-61: a wall whose face is x = face_x, counter-clockwise."""
    return [(face_x - thickness, y_min), (face_x, y_min), (face_x, y_max), (face_x - thickness, y_max)]


def crowd_mass(radius):
    return float(np.float32(radius) * np.float32(radius) / np.float32(400.0))


def runner_for(oracle, w):
    r = rs.OracleRunner(w.build(), oracle)
    assert r.step()["status"] == abi.BP_OK                 # the step that accepts the scene
    return r


def settle_against_face(runner, radius, face_x, steps):
    """:68-80: deflection r - (x - face_x) after `steps`, and its spread over the last 60."""
    tail, position = [], None
    for k in range(steps):
        out = runner.step()
        assert out["status"] == abi.BP_OK, k
        position = out["position"][1]
        if k >= steps - 60:
            tail.append(radius - (position[0] - face_x))
    return tail[-1], max(tail) - min(tail), position


@pytest.mark.parametrize("pressure", [300.0, 700.0, 1100.0, 1500.0, 2500.0])
def test_thin_wall_deflects_by_the_front_edge_only_and_like_a_thick_wall(oracle, pressure):
    # :82-105: r 148, m 1, gamma 5 against the middle of a wall 48 or 400 thick.
    radius, face_x = 148.0, -100.0
    expected = deflection(radius, pressure)
    got = []
    for thickness in (48.0, 400.0):
        w = world(1000.0, [rectangle(face_x, thickness)])
        w.bubble(1, (face_x + radius, 0.0), radius=radius, mass=1.0).motor(1, (-pressure, 0.0))
        d, spread, _ = settle_against_face(runner_for(oracle, w), radius, face_x, 1200)
        assert abs(d - expected) <= 0.02 * expected, (thickness, d, expected)
        assert spread < 0.005 * expected, (thickness, spread)
        got.append(d)
    assert abs(got[0] - got[1]) <= 0.02 * expected


def test_collinear_vertex_does_not_change_the_contact(oracle):
    # :107-127: the face given as two collinear segments of 400 behaves as one segment.
    radius, face_x, pressure = 24.0, -100.0, 600.0
    expected = deflection(radius, pressure)
    two = [(face_x - 48, -400.0), (face_x, -400.0), (face_x, 0.0), (face_x, 400.0), (face_x - 48, 400.0)]
    results = []
    for polygon in (two, rectangle(face_x, 48.0)):
        w = world(1000.0, [polygon])
        w.bubble(1, (face_x + radius, 0.0), radius=radius, mass=1.36).motor(1, (-pressure, 0.0))
        results.append(settle_against_face(runner_for(oracle, w), radius, face_x, 600))
    for d, spread, _ in results:
        assert abs(d - expected) <= 0.02 * expected
    assert results[0][1] < 0.005 * expected
    assert np.allclose(results[0][2], results[1][2], atol=1e-2)


def test_collinear_vertex_is_merged_into_one_edge():
    # NativeStaticPolygonEdges.swift:40-60: |e1 x e2| <= 1e-5 |e1||e2| with e1.e2 > 0 merges the
    # edges, so the two-segment face and the single face give the same segments.
    face_x = -100.0
    two = [(face_x - 48, -400.0), (face_x, -400.0), (face_x, 0.0), (face_x, 400.0), (face_x - 48, 400.0)]
    a = scenes.static_polygon_edges(two, 1, 9001)
    b = scenes.static_polygon_edges(rectangle(face_x, 48.0), 1, 9001)
    assert len(a) == len(b) == 4
    key = lambda s: sorted((tuple(x["a"]), tuple(x["b"]), int(x["flags"]), float(x["outward_sign"])) for x in s)
    assert key(a) == key(b)


APEX_ANGLES = (30.0, 90.0, 120.0)


def wedge(degrees, length=300.0):
    """:135-141: apex (0, 0) pointing to -x, interior angle `degrees`, clockwise."""
    half = degrees * math.pi / 360.0
    c, s = float(np.float32(length * math.cos(half))), float(np.float32(length * math.sin(half)))
    return [(0.0, 0.0), (c, s), (c, -s)]


def nearest_point(p, polygon):
    """:151-162 (Float arithmetic is not needed for a direction)."""
    best, best_d = np.array(polygon[0], np.float64), math.inf
    for i in range(len(polygon)):
        a, b = np.array(polygon[i], np.float64), np.array(polygon[(i + 1) % len(polygon)], np.float64)
        e = b - a
        t = min(max(np.dot(p - a, e) / max(np.dot(e, e), 1e-38), 0.0), 1.0)
        q = a + t * e
        if np.hypot(*(p - q)) < best_d:
            best_d, best = np.hypot(*(p - q)), q
    return best


def unit(v):
    return v / max(np.hypot(*v), 1e-38)


@pytest.mark.parametrize("degrees", APEX_ANGLES)
def test_convex_vertex_is_one_spring_for_every_angle(oracle, degrees):
    # :166-185: pressed toward the apex every frame, the distance to it settles at r - r eps.
    radius, pressure = 24.0, 600.0
    expected = deflection(radius, pressure)
    w = world(1000.0, [wedge(degrees)])
    w.bubble(1, (-radius, 0.0), radius=radius, mass=1.36)
    runner = runner_for(oracle, w)
    position, tail = np.array([-radius, 0.0]), []
    for k in range(420):
        runner.set_motor(1, unit(-position) * pressure)
        out = runner.step()
        assert out["status"] == abi.BP_OK
        position = out["position"][1]
        if k >= 360:
            tail.append(radius - np.hypot(*position))
    assert abs(tail[-1] - expected) <= 0.02 * expected, (tail[-1], expected)
    assert max(tail) - min(tail) < 0.005 * expected


@pytest.mark.parametrize("degrees", APEX_ANGLES)
@pytest.mark.parametrize("side", [1.0, -1.0])
def test_slide_through_a_convex_vertex_keeps_the_distance_from_the_edge(oracle, degrees, side):
    # :191-236: pressed to the nearest point with 600 and pushed along with 30, the bubble passes
    # the apex keeping r - r eps to the boundary within 10 % of r eps.
    radius, pressure, tangential = 24.0, 600.0, 30.0
    expected = deflection(radius, pressure)
    polygon = wedge(degrees)
    half = degrees * math.pi / 360.0
    start = np.array([math.cos(half), side * math.sin(half)])
    finish = np.array([math.cos(half), -side * math.sin(half)])
    outward = np.array([-math.sin(half), side * math.cos(half)])
    first = np.float32(40 * start + (radius - expected) * outward).astype(np.float64)
    toward0 = unit(nearest_point(first, polygon) - first)
    sense = 1.0 if np.dot(np.array([-toward0[1], toward0[0]]), -start) >= 0 else -1.0
    w = world(1000.0, [polygon], damping=4.0)
    w.bubble(1, tuple(first), radius=radius, mass=1.36)
    runner = runner_for(oracle, w)

    def push(p, slide):
        toward = unit(nearest_point(p, polygon) - p)
        perp = np.array([-toward[1], toward[0]])
        return toward * pressure + (sense * perp * tangential if slide else 0.0)

    position = first
    for _ in range(240):
        runner.set_motor(1, push(position, False))
        out = runner.step()
        assert out["status"] == abi.BP_OK
        position = out["position"][1]
    worst, reached = 0.0, False
    for _ in range(1500):
        runner.set_motor(1, push(position, True))
        out = runner.step()
        assert out["status"] == abi.BP_OK
        position = out["position"][1]
        worst = max(worst, abs(np.hypot(*(position - nearest_point(position, polygon))) - (radius - expected)))
        if np.dot(position, finish) >= 40:
            reached = True
            break
    assert reached
    assert worst <= 0.1 * expected, (worst, expected)


def winding_worlds(reversed_):
    radius, pressure = 24.0, 600.0
    thick, thin, corner, inner = (-1200.0, 1200.0), (1200.0, 1200.0), (-1200.0, -1200.0), (1200.0, -1200.0)
    arm, width = 400.0, 48.0
    polygons = [rectangle(thick[0], 400.0, thick[1] - 400.0, thick[1] + 400.0),
                rectangle(thin[0], 48.0, thin[1] - 400.0, thin[1] + 400.0),
                rectangle(corner[0], 48.0, corner[1] - 800.0, corner[1]),
                [(x + inner[0] - width, y + inner[1] - width)
                 for x, y in [(0, 0), (arm, 0), (arm, width), (width, width), (width, arm), (0, arm)]]]
    bisector = np.array([1.0, 1.0]) / math.sqrt(2.0)
    angle = 195.0 * math.pi / 180.0
    aim = np.array([math.cos(angle), math.sin(angle)])
    pushes = [(np.array(thick) + [radius, 0.0], np.array([-pressure, 0.0])),
              (np.array(thin) + [radius, 0.0], np.array([-pressure, 0.0])),
              (np.array(corner) - aim * 50, aim * pressure),
              (np.array(inner) + bisector * 100, -bisector * pressure)]
    w = world(2500.0, polygons, damping=2.0, reversed_=reversed_)
    for k, (at, force) in enumerate(pushes):
        w.bubble(k + 1, tuple(np.float32(at)), radius=radius, mass=1.36).motor(k + 1, tuple(np.float32(force)))
    return w


def test_winding_of_the_vertices_does_not_change_the_trajectories(oracle):
    # :242-289
    runners = [runner_for(oracle, winding_worlds(r)) for r in (False, True)]
    worst, slid = 0.0, None
    for k in range(600):
        outs = [r.step() for r in runners]
        assert all(o["status"] == abi.BP_OK for o in outs)
        for b in range(1, 5):
            worst = max(worst, float(np.hypot(*(outs[0]["position"][b] - outs[1]["position"][b]))))
        slid = outs[0]["position"][3]
    assert worst <= 1e-2
    assert slid[1] < -1200.0 - 300.0


@pytest.mark.parametrize("degrees", [0.0, 30.0])
def test_fast_bubble_never_ends_a_step_inside_or_behind_the_wall(oracle, degrees):
    # :295-313: r 12 at 6000 units/s, 100 units per step, more than 2 r + T.
    radius, speed = 12.0, 6000.0
    heading = degrees * math.pi / 180.0
    velocity = (float(np.float32(-math.cos(heading) * speed)), float(np.float32(math.sin(heading) * speed)))
    w = world(2500.0, [rectangle(0.0, 48.0)])
    w.bubble(1, (200.0, 0.0), radius=radius, mass=crowd_mass(radius), velocity=velocity)
    runner = runner_for(oracle, w)
    for k in range(40):
        out = runner.step()
        assert out["status"] == abi.BP_OK, k
        p = out["position"][1]
        if abs(p[1]) <= 400:
            assert p[0] >= -radius / 10, (k, p)


def test_oblique_entry_over_a_convex_corner_is_pushed_out_by_the_crossed_edge_only(oracle):
    # :320-344
    radius, damping = 12.0, 5.0
    velocity = (1200.0, -1800.0)
    w = world(2500.0, [rectangle(-100.0, 48.0)], damping=damping)
    w.bubble(1, (-220.0, 510.0), radius=radius, mass=crowd_mass(radius), velocity=velocity)
    runner = runner_for(oracle, w)
    previous_x, bounced = -math.inf, False
    for k in range(6):
        out = runner.step()
        assert out["status"] == abi.BP_OK
        p, v = out["position"][1], out["velocity"][1]
        assert not (-148 < p[0] < -100 and p[1] < 400), (k, p)
        assert p[0] > previous_x, (k, p)
        previous_x = p[0]
        bounced |= v[1] > 0
        dragged = velocity[0] * math.exp(-damping * DT * (k + 2))
        assert v[0] > 0.88 * dragged, (k, v)
    assert bounced


def test_thin_wall_pin_catches_two_sided_edges(oracle):
    # The mistake the thin-wall test guards against: polygon edges as plain two-sided segments (no
    # edge regions). The back face then pushes too, and the deflection leaves the 2 % band.
    radius, face_x, pressure = 148.0, -100.0, 2500.0
    expected = deflection(radius, pressure)
    w = world(1000.0, [rectangle(face_x, 48.0)])
    w.bubble(1, (face_x + radius, 0.0), radius=radius, mass=1.0).motor(1, (-pressure, 0.0))
    scene = w.build()
    scene.segments["flags"] &= ~np.uint32(abi.BP_SEGMENT_POLYGON_EDGE | abi.BP_SEGMENT_CONVEX_START)
    runner = rs.OracleRunner(scene, oracle)
    runner.step()
    d, _, _ = settle_against_face(runner, radius, face_x, 1200)
    assert abs(d - expected) > 0.02 * expected, (d, expected)
