"""Pins of the oracle's body obstacle, body prediction, region pass and contact assertion by the
reference's own world-level tests, restated on the oracle with the same scenes and tolerances:

  * RigidContactTests.swift:31-86 (a bubble pressed against a static circle, a static capsule, a
    standing kinematic square and its corner holds |d - d*| <= 0.5 % d* + 1e-3, d* from the Double
    law), :90-111 (filters let a bubble through a static circle), :145-183 (a spinning kinematic
    square never holds the bubble's centre and keeps its own path);
  * ContactAssertionTests.swift:35-117 (injected states: coincident centres, centre inside a static
    polygon, circle, capsule or kinematic body, the thin-wall side change, excluding filters,
    static bubbles, non-finite state), :118-166 (wraps onto free space, onto an overlapping
    bubble, into a body, onto a centre), :168-235 (the substep limit and game scenes within it);
  * RegionPolicyTests.swift:9-17, 51-60, 107-134, 258-275 (removal outside a corner, wrap across
    several region widths, diagonal corner wrap, the wrap seam, wrapped ids).

Each group also runs with a mistake planted in the oracle (a dropped term, a reversed sign, an
ignored term) to show that its pin fails then.
"""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi
import contact_law_reference as law
import reference_scenes as rs

K = rs.FORCE_SCENE_STIFFNESS


@pytest.fixture
def planted(oracle):
    """Runs a block with a planted mistake in the oracle, always switching it off again."""
    class Plant:
        def __init__(self, code):
            self.code = code

        def __enter__(self):
            oracle.set_plant(self.code)

        def __exit__(self, *exc):
            oracle.set_plant(0)
    yield Plant
    oracle.set_plant(0)


def d_star(force, radius=24.0):
    """ContactLawReference.equilibriumDistance: r (1 - eps(F / K))."""
    return radius * (1 - law.epsilon_for_ratio(force / K))


def oracle_run(oracle, world, frames, **kw):
    runner = rs.OracleRunner(world.build(), oracle, **kw)
    return runner, runner.run(frames)


# ---- RigidContactTests ---------------------------------------------------------------------------
SURFACE = {"circle": 50.0, "capsule": 20.0, "kinematic": 50.0}


def pressed(kind, force):
    """RigidContactTests.assertEquilibrium (:31-48): bubble "2" at rest at d* left of the body,
    pressed by a global force F in +x."""
    w = rs.World().bubble(1, (-(SURFACE[kind] + d_star(force)), 0.0)).set_global_force((force, 0.0))
    if kind == "circle":
        w.circle(2, (0.0, 0.0), 50.0)
    elif kind == "capsule":
        w.capsule(2, (0.0, 0.0), 50.0, 20.0, rotation=math.pi / 2)     # its flat side faces the bubble
    else:
        w.square(2, (0.0, 0.0), 100.0, role="kinematic")
    return w


def equilibrium_violations(oracle, kind, force, frames=1200):
    _, out = oracle_run(oracle, pressed(kind, force), frames)
    ds, bad = d_star(force), []
    for k, r in enumerate(out):
        if r["status"] != abi.BP_OK:
            return [("status", k, r["status"])]
        d = -r["position"][1][0] - SURFACE[kind]
        if not (d > 0 and abs(d - ds) <= 0.005 * ds + 1e-3):
            bad.append((k, d, ds))
    return bad if len(out) == frames else [("frames", len(out))]


@pytest.mark.parametrize("kind", ["circle", "capsule", "kinematic"])
@pytest.mark.parametrize("force", [100.0, 1200.0, 10000.0])
def test_bubble_pressed_against_a_body_holds_the_law_distance(oracle, kind, force):
    assert equilibrium_violations(oracle, kind, force) == []


@pytest.mark.parametrize("force", [100.0, 1200.0])
def test_bubble_pressed_into_a_kinematic_corner_holds_the_law_distance_from_the_vertex(oracle, force):
    # :53-72: the surface near a polygon corner is the arc around the vertex.
    diagonal = float(np.float32(math.sqrt(0.5)))
    ds = d_star(force)
    corner = np.array([-50.0, -50.0])
    start = corner - np.array([diagonal, diagonal]) * ds
    w = rs.World().bubble(1, tuple(np.float32(start))).square(2, (0.0, 0.0), 100.0, role="kinematic")
    w.set_global_force((force * diagonal, force * diagonal))
    _, out = oracle_run(oracle, w, 600)
    assert len(out) == 600 and all(r["status"] == abi.BP_OK for r in out)
    for k, r in enumerate(out):
        d = float(np.hypot(*(r["position"][1] - corner)))
        assert abs(d - ds) <= 0.005 * ds + 1e-3, (k, d, ds)


def test_round_core_and_polygon_edge_pins_catch_planted_mistakes(oracle, planted):
    with planted(1):        # distance to the core axis without the core radius
        assert equilibrium_violations(oracle, "circle", 1200.0, frames=120) != []
        assert equilibrium_violations(oracle, "capsule", 1200.0, frames=120) != []
    with planted(2):        # the nearest vertex taken for the nearest point of a polygon edge
        assert equilibrium_violations(oracle, "kinematic", 1200.0, frames=120) != []
    assert equilibrium_violations(oracle, "circle", 1200.0, frames=120) == []


def test_filter_lets_a_bubble_through_a_static_circle_while_the_default_bubble_bounces(oracle):
    # :90-111
    def run(category, mask):
        w = rs.World().bubble(1, (-200.0, 0.0), velocity=(300.0, 0.0), category=category, mask=mask)
        w.circle(2, (0.0, 0.0), 50.0)
        _, out = oracle_run(oracle, w, 90)
        assert len(out) == 90 and all(r["status"] == abi.BP_OK for r in out)
        return out
    blocked = run(1, rs.ALL)
    assert all(r["position"][1][0] < -50 for r in blocked)
    assert blocked[-1]["velocity"][1][0] < 0
    passing = run(2, 2)
    assert passing[-1]["position"][1][0] > 50
    assert abs(passing[-1]["velocity"][1][0] - 300.0) <= 1e-3


def inside_square(p, center, rotation, half):
    """ForceScene.isInsideSquare: boundary counts as inside."""
    d = np.asarray(p, np.float64) - np.asarray(center, np.float64)
    c, s = math.cos(rotation), math.sin(rotation)
    local = (c * d[0] + s * d[1], -s * d[0] + c * d[1])
    return abs(local[0]) <= half and abs(local[1]) <= half


def spinning_run(oracle, frames=30):
    """:145-183: a 50-unit square at 1000 units/s spinning at 20 rad/s meets bubble "2" at rest."""
    w = rs.World().bubble(1, (360.0, 0.0)).square(2, (250.0, 0.0), 50.0, role="kinematic", velocity=(1000.0, 0.0),
                                                    angular=20.0)
    scene = w.build()
    previous = (np.array([250.0, 0.0]), 0.0)
    checked, inside = 0, []
    for k in range(frames):
        cb, cbod, cf, ex = oracle.run(scene, 1, capture=True, history=True)
        assert cf[0]["status"] == abi.BP_OK, k
        pose = cbod[0]["pose"].astype(np.float64)
        h = ex["history"][: int(cf[0]["substeps"])].astype(np.float64)
        total, elapsed = h.sum(), 0.0
        for j, sample in enumerate(ex["substeps"]):
            elapsed += h[j]
            fraction = elapsed / total if total > 0 else 1.0
            center = previous[0] + (pose[:2] - previous[0]) * fraction
            rotation = previous[1] + (pose[2] - previous[1]) * fraction
            checked += 1
            if inside_square(sample[0, :2], center, rotation, 25.0):
                inside.append((k, j))
        previous = (pose[:2].copy(), float(pose[2]))
        scene.bubbles, scene.bodies = cb, cbod
    return scene, checked, inside


def test_spinning_kinematic_body_never_holds_the_bubble_centre_and_keeps_its_path(oracle):
    scene, checked, inside = spinning_run(oracle)
    assert inside == [] and checked > 30
    assert scene.bubbles[0]["velocity"][0] > 0.5 * 1000.0
    pose = scene.bodies[0]["pose"]
    assert abs(pose[0] - (250.0 + 1000.0 * 30 / 60)) <= 2e-3 and abs(pose[1]) <= 2e-3
    assert abs(pose[2] - 20.0 * 30 / 60) <= 1e-3


def test_body_prediction_pin_catches_a_planted_mistake(oracle, planted):
    with planted(5):        # the body prediction ignores the angular velocity
        scene, _, _ = spinning_run(oracle, frames=6)
        assert abs(scene.bodies[0]["pose"][2] - 20.0 * 6 / 60) > 1e-3


# ---- ContactAssertionTests -----------------------------------------------------------------------
def violation(oracle, world, frames=5):
    """Runs until the first issue (ForceScene.run); returns (status, objects, kind, segment) of it."""
    _, out = oracle_run(oracle, world, frames)
    last = out[-1]
    f = last["frame"]
    return last["status"], {int(x) for x in f["violation_objects"] if x}, int(f["violation_kind"]), int(f["violation_segment"])


INJECTED = {
    "pair": (lambda w: w.bubble(1, (0.0, 0.0)).bubble(2, (200.0, 0.0)), (200.0, 0.0), abi.BP_VIOLATION_PAIR),
    "static polygon": (lambda w: w.bubble(1, (300.0, 0.0)).square(2, (0.0, 0.0), 100.0), (10.0, 10.0),
                       abi.BP_VIOLATION_EDGE_SIDE),
    "static circle": (lambda w: w.bubble(1, (300.0, 0.0)).circle(2, (0.0, 0.0), 50.0), (0.0, 20.0),
                      abi.BP_VIOLATION_INSIDE_BODY),
    "static capsule": (lambda w: w.bubble(1, (300.0, 0.0)).capsule(2, (0.0, 0.0), 50.0, 20.0), (30.0, 5.0),
                       abi.BP_VIOLATION_INSIDE_BODY),
    "kinematic body": (lambda w: w.bubble(1, (300.0, 0.0)).square(2, (0.0, 0.0), 100.0, role="kinematic"), (10.0, 10.0),
                       abi.BP_VIOLATION_INSIDE_BODY),
    "thin wall side": (lambda w: w.bubble(1, (80.0, 0.0)).wall(2, 100.0), (150.0, 0.0), abi.BP_VIOLATION_EDGE_SIDE),
}


def injected_world(name):
    build, position, _ = INJECTED[name]
    return rs.injecting(build(rs.World()), 1, position)


@pytest.mark.parametrize("name", list(INJECTED))
def test_injected_state_fails_the_world_with_a_contact_violation(oracle, name):
    # :35-80: an injected state after substep 0 with the assertion after every substep.
    status, objects, kind, segment = violation(oracle, injected_world(name))
    assert status == abi.BP_ERR_CONTACT_VIOLATION and {1, 2} <= objects, (status, objects)
    assert kind == INJECTED[name][2]
    if kind == abi.BP_VIOLATION_EDGE_SIDE:
        assert segment > 0                                     # the diagnostics name the edge


def test_assertion_pins_catch_planted_mistakes(oracle, planted):
    with planted(4):        # the side after the move read at the step's start point
        status, _, kind, _ = violation(oracle, injected_world("thin wall side"))
        assert not (status == abi.BP_ERR_CONTACT_VIOLATION and kind == abi.BP_VIOLATION_EDGE_SIDE)


def test_excluding_filters_let_objects_pass_through_each_other(oracle):
    # :82-98
    pair = rs.injecting(rs.World().bubble(1, (0.0, 0.0), category=1, mask=1).bubble(2, (200.0, 0.0), category=2, mask=2),
                        1, (200.0, 0.0))
    _, out = oracle_run(oracle, pair, 30)
    assert len(out) == 30 and all(r["status"] == abi.BP_OK for r in out)
    body = rs.injecting(rs.World().bubble(1, (300.0, 0.0), category=2, mask=2).square(2, (0.0, 0.0), 100.0), 1, (10.0, 10.0))
    _, out = oracle_run(oracle, body, 30)
    assert len(out) == 30 and all(r["status"] == abi.BP_OK for r in out)


def test_static_bubbles_are_not_asserted_against_each_other_nor_bodies(oracle):
    # :102-116
    pair = rs.injecting(rs.World().bubble(1, (0.0, 0.0), role="static").bubble(2, (200.0, 0.0), role="static"),
                        1, (200.0, 0.0))
    _, out = oracle_run(oracle, pair, 10)
    assert len(out) == 10 and all(r["status"] == abi.BP_OK for r in out)
    body = rs.injecting(rs.World().bubble(1, (300.0, 0.0), role="static").square(2, (0.0, 0.0), 100.0), 1, (10.0, 10.0))
    _, out = oracle_run(oracle, body, 10)
    assert len(out) == 10 and all(r["status"] == abi.BP_OK for r in out)


def test_non_finite_injected_position_fails_with_non_finite_values(oracle):
    # :118-126
    w = rs.injecting(rs.World().bubble(1, (0.0, 0.0)), 1, (float("nan"), 0.0))
    status, _, _, _ = violation(oracle, w)
    assert status == abi.BP_ERR_NON_FINITE


def wrap_scene(rest=lambda w: w):
    """:131-136: a .wrap bubble leaves the 2000 x 2000 region through the top in one frame."""
    return rest(rs.World(half=1000.0).bubble(1, (0.0, 1000.0), velocity=(0.0, 3600.0), wrap=True))


def test_wrap_onto_free_space_keeps_simulating(oracle):
    runner, out = oracle_run(oracle, wrap_scene(), 30)
    assert len(out) == 30 and all(r["status"] == abi.BP_OK for r in out)
    assert out[0]["position"][1][1] < -900
    assert list(out[0]["wrapped"]) == [1]


def test_wrap_onto_an_overlapping_bubble_pushes_them_apart(oracle):
    _, out = oracle_run(oracle, wrap_scene(lambda w: w.bubble(2, (30.0, -988.0))), 30)
    assert len(out) == 30 and all(r["status"] == abi.BP_OK for r in out)
    b, v = out[3]["position"][2], out[3]["velocity"][2]
    assert b[0] > 30.05 and 0.5 < v[0] < 3.0, (b, v)


def test_wrap_into_a_rigid_body_fails_the_world(oracle):
    status, objects, _, _ = violation(oracle, wrap_scene(lambda w: w.square(2, (0.0, -988.0), 20.0)))
    assert status == abi.BP_ERR_CONTACT_VIOLATION and {1, 2} <= objects


def test_wrap_exactly_onto_another_centre_fails_the_world(oracle):
    _, out = oracle_run(oracle, wrap_scene(), 1)
    landing = out[0]["position"][1]
    assert landing[1] < -900
    status, objects, kind, _ = violation(oracle, wrap_scene(lambda w: w.bubble(2, tuple(np.float32(landing)))))
    assert status == abi.BP_ERR_CONTACT_VIOLATION and {1, 2} <= objects and kind == abi.BP_VIOLATION_PAIR


def test_kinematic_squeeze_against_a_wall_exceeds_the_substep_limit(oracle):
    # :168-191: first contact after 0.76 s (frame 46), the limit must act within 1 s more.
    w = rs.World().bubble(1, (300.0, 0.0)).wall(2, 360.0).square(3, (150.0, 0.0), 100.0, role="kinematic",
                                                                    velocity=(100.0, 0.0))
    _, out = oracle_run(oracle, w, 46 + 60 + 5)
    last = out[-1]
    assert last["status"] == abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED, (len(out), last["status"])
    assert int(last["frame"]["limiter_object"]) == 1
    assert int(last["frame"]["limiter_term"]) != abi.BP_TERM_NONE
    assert all(int(r["frame"]["substeps"]) <= 1024 for r in out[:-1])


def assert_within_the_limit(oracle, world, frames, until=None):
    _, out = oracle_run(oracle, world, frames)
    assert all(r["status"] == abi.BP_OK for r in out)
    counts = [int(r["frame"]["substeps"]) for r in out]
    assert 0 < max(counts) <= 1024
    return out


def test_game_scenes_never_exceed_the_substep_limit(oracle):
    # :196-235
    spinning = rs.World().bubble(1, (360.0, 0.0)).square(2, (250.0, 0.0), 50.0, role="kinematic", velocity=(1000.0, 0.0),
                                                          angular=20.0)
    assert len(assert_within_the_limit(oracle, spinning, 30)) == 30
    for radius, mass in ((24.0, 1.0), (64.0, rs.game_mass(128)), (148.0, rs.game_mass(7168))):
        w = rs.World().bubble(1, (360.0 - radius - 100.0, 0.0), radius=radius, mass=mass, velocity=(3000.0, 0.0)).wall(2, 360.0)
        runner = rs.OracleRunner(w.build(), oracle)
        out = runner.run(200, until=lambda r: r["velocity"][1][0] < 0 and r["position"][1][0] < 360.0 - radius)
        assert all(r["status"] == abi.BP_OK for r in out)
        assert 0 < max(int(r["frame"]["substeps"]) for r in out) <= 1024
        assert out[-1]["velocity"][1][0] < 0, radius              # it bounced back
    crowd = rs.World(damping=1.0)
    for row in range(4):
        for column in range(4):
            crowd.bubble(1 + row * 4 + column, (80.0 + column * 50.0, row * 50.0 - 75.0))
    crowd.bubble(100, (0.0, 0.0), velocity=(3000.0, 0.0))
    assert len(assert_within_the_limit(oracle, crowd, 120)) == 120


# ---- RegionPolicyTests -----------------------------------------------------------------------------
def region_world(at, velocity=(0.0, 0.0), wrap=False):
    """RegionPolicyTests.initial (:258-266): region [0, 10]^2, dt 0.25, one bubble of radius 1."""
    return rs.World(region=(0.0, 0.0, 10.0, 10.0), dt=0.25).bubble(1, at, radius=1.0, velocity=velocity, wrap=wrap)


def test_circle_outside_a_corner_is_removed(oracle):
    # :9-16: its bounds overlap the region, its collider does not.
    runner, out = oracle_run(oracle, region_world((10.8, 10.8)), 1)
    assert out[0]["status"] == abi.BP_OK and list(out[0]["removed"]) == [1]
    assert len(runner.scene.bubbles) == 0


def overshoot(oracle):
    runner, out = oracle_run(oracle, region_world((-25.0, 5.0), (-4.0, 0.0), wrap=True), 1)
    return out[0]["bubbles"][0]


def test_wrap_keeps_the_overshoot_across_several_region_widths(oracle):
    # :107-117: x = -26 wraps by three periods of 12 to 10, the start -25 to 11.
    b = overshoot(oracle)
    assert tuple(b["center"]) == (10.0, 5.0)
    assert tuple(b["previous_center"]) == (11.0, 5.0)
    assert tuple(b["velocity"]) == (-4.0, 0.0)


def test_region_pin_catches_a_planted_mistake(oracle, planted):
    with planted(3):        # the wrap period without the collider's diameter
        assert tuple(overshoot(oracle)["center"]) != (10.0, 5.0)


def test_wrapped_object_is_reported_and_a_step_without_wraps_reports_none(oracle):
    # :51-67: the wrap goes into the snapped set; a short step wraps nothing.
    _, out = oracle_run(oracle, region_world((-25.0, 5.0), (-4.0, 0.0), wrap=True), 1)
    assert list(out[0]["wrapped"]) == [1] and list(out[0]["removed"]) == []
    w = region_world((9.0, 5.0), (4.0, 0.0), wrap=True)
    w.dt = 0.001
    _, out = oracle_run(oracle, w, 1)
    assert list(out[0]["wrapped"]) == [] and list(out[0]["removed"]) == []


def test_diagonal_corner_wrap_reenters_and_does_not_oscillate(oracle):
    # :119-134
    _, out = oracle_run(oracle, region_world((10.8, 10.8), wrap=True), 2)
    first, second = out[0]["position"][1], out[1]["position"][1]
    dx = max(0.0, max(-first[0], first[0] - 10.0))
    dy = max(0.0, max(-first[1], first[1] - 10.0))
    assert math.hypot(dx, dy) <= 1.0
    assert abs(second[0] - first[0]) <= 1e-4 and abs(second[1] - first[1]) <= 1e-4
    assert tuple(out[1]["velocity"][1]) == (0.0, 0.0)


def test_wrap_keeps_the_velocity_and_makes_no_seam_contact(oracle):
    # :206-224: the opposite-edge bubble overlaps only a toroidal image of the wrapped one.
    w = region_world((9.0, 5.0), (4.0, 0.0), wrap=True).bubble(2, (0.25, 5.0), radius=1.0)
    runner, out = oracle_run(oracle, w, 3)
    assert all(r["status"] == abi.BP_OK for r in out)
    assert abs(out[0]["position"][2][0] - 0.25) <= 1e-4
    b1 = out[2]["bubbles"][out[2]["bubbles"]["id"] == 1][0]
    assert abs(b1["center"][0] - 0.0) <= 1e-4 and abs(b1["previous_center"][0] + 1.0) <= 1e-4
    assert tuple(b1["velocity"]) == (4.0, 0.0)
