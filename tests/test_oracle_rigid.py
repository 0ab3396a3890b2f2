"""Pins of the oracle's rigid solver (oracle/bp_oracle_rigid.cpp) for worlds without bubbles.

The reference's own rigid tests are restated on the oracle with their thresholds:
RigidCCDTests.swift, RigidShapeContactTests.swift, JointWorldTests.swift, VelocityLimitTests.swift
(Tests/BubblePhysicsMetalRenderTests) and NativeMetalRigidPolygonTests.swift
(Tests/BubblePhysicsNativeMetalTests). Closed forms pin what those leave open: free flight, the
angular damping factor, the first contact iteration of overlapping boxes and circles, and the
region rule's shift. Planted mistakes (bpo_rigid_set_plant) show that the checks catch a flipped
contact normal, a dropped time of impact, a joint correction of the wrong sign, a wrap without
the region size and a dropped damping factor. CPU only.
"""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi, rigid
from oracle import bp_oracle
from rigid_scenes import (HALF_WIDTH, HINGE_PIVOT, PIVOT, RIGID_SHAPES, SEPARATION_PAIRS, Case, cfg, ccd_floor,
                          ccd_head_on, ccd_kinematic, ccd_round, ccd_thin_wall, connected, hinge_limits, hinge_motor,
                          hinge_motor_between, hinge_motor_load, joint_determinism, kinematic_speed_limit, motor_polygon,
                          no_speed_limit, pendulum, polygon_against_wall, polygons_swap, positions, region_exit,
                          rigid_distance, shape_against_static, shape_separation, simulate_oracle, soft_distance,
                          speed_limits, weld)


def began(events, a, b):
    return [e for e in events if e["kind"] == abi.BP_EVENT_CONTACT_BEGAN and {int(e["first"]), int(e["second"])} == {a, b}]


def ended(events, a, b):
    return [e for e in events if e["kind"] == abi.BP_EVENT_CONTACT_ENDED and {int(e["first"]), int(e["second"])} == {a, b}]


def all_ok(frames):
    return all(info["status"] == abi.BP_OK for _, _, info in frames)


def by_id(bodies, id):
    return int(np.nonzero(bodies["id"] == np.uint64(id))[0][0])


# ---- RigidCCDTests.swift ----------------------------------------------------------------------
def ccd_floor_holds(plant=0):
    frames, events = simulate_oracle(ccd_floor(), plant=plant)
    lowest = min(positions(b)[by_id(b, 1)][1] for b, _, _ in frames)
    final = positions(frames[-1][0])[by_id(frames[-1][0], 1)][1]
    return (all_ok(frames) and len(frames) == 120 and lowest > 100 + 2 + 2 - 0.5 and abs(final - 104) <= 0.6
            and len(began(events, 1, 2)) == 1 and not ended(events, 1, 2))


def test_fast_box_does_not_tunnel_through_thin_static_floor():
    assert ccd_floor_holds()


def ccd_head_on_holds(plant=0):
    frames, events = simulate_oracle(ccd_head_on(), plant=plant)
    gaps = [positions(b)[1][0] - positions(b)[0][0] for b, _, _ in frames]
    last = positions(frames[-1][0])
    return (all_ok(frames) and min(gaps) > 4 - 0.5 and last[0][0] < last[1][0] and len(began(events, 1, 2)) == 1
            and not ended(events, 1, 2))


def test_two_fast_boxes_head_on_do_not_pass_through_each_other():
    assert ccd_head_on_holds()


def test_fast_box_does_not_tunnel_through_thin_static_wall():
    frames, _ = simulate_oracle(ccd_thin_wall())
    assert all_ok(frames)
    assert max(positions(b)[by_id(b, 1)][0] for b, _, _ in frames) < 300 - 2 - 2 + 0.5


def test_kinematic_body_is_not_pushed_back_and_stops_fast_box():
    frames, _ = simulate_oracle(ccd_kinematic())
    assert all_ok(frames)
    assert max(positions(b)[by_id(b, 1)][0] for b, _, _ in frames) < 300 - 2 - 2 + 0.5
    last = frames[-1][0]
    assert tuple(positions(last)[by_id(last, 2)]) == (300.0, 300.0)


def test_fast_circle_and_capsules_do_not_tunnel():
    frames, _ = simulate_oracle(ccd_round())
    assert all_ok(frames)
    for b, _, _ in frames:
        p = positions(b)
        assert p[by_id(b, 1)][0] <= p[by_id(b, 2)][0], "the circle may not pass the thin box"
        assert p[by_id(b, 3)][0] <= p[by_id(b, 4)][0], "the capsules may not pass each other"
    p = positions(frames[-1][0])
    assert p[by_id(frames[-1][0], 3)][0] < p[by_id(frames[-1][0], 4)][0]


# ---- RigidShapeContactTests.swift ---------------------------------------------------------------
@pytest.mark.parametrize("mover,target", [(m, t) for m in RIGID_SHAPES for t in RIGID_SHAPES])
def test_every_rigid_shape_pair_collides_against_a_static_body(mover, target):
    frames, events = simulate_oracle(shape_against_static(mover, target))
    assert all_ok(frames) and len(frames) == 90
    last = frames[-1][0]
    assert np.all(np.isfinite(last["pose"])) and np.all(np.isfinite(last["motion"]))
    hits = began(events, 1, 2)
    assert hits, "a contactBegan event"
    e = hits[0]
    assert e["impulse"] > 0
    assert abs(math.hypot(*e["normal"]) - 1) <= 1e-3
    toward = 1.0 if int(e["first"]) == 1 else -1.0
    assert e["normal"][0] * toward > 0.7, "the normal points from the first object to the second"
    expected = 150 - HALF_WIDTH[target] - HALF_WIDTH[mover]
    assert abs(positions(last)[by_id(last, 1)][0] - expected) <= 2.0, "the mover stops at the surface"
    assert tuple(positions(last)[by_id(last, 2)]) == (150.0, 100.0), "the static body stays"


def separation_holds(mover, target, plant=0):
    frames, _ = simulate_oracle(shape_separation(mover, target), plant=plant)
    last = frames[-1][0]
    pushed, moved = by_id(last, 2), by_id(last, 1)
    p = positions(last)
    return (all_ok(frames) and last["motion"][pushed][0] > 3
            and p[pushed][0] - p[moved][0] >= HALF_WIDTH[mover] + HALF_WIDTH[target] - 2.0)


@pytest.mark.parametrize("mover,target", SEPARATION_PAIRS)
def test_every_rigid_shape_pair_separates_when_dynamic(mover, target):
    assert separation_holds(mover, target)


# ---- JointWorldTests.swift --------------------------------------------------------------------
def test_pendulum_period_matches_the_formula_within_two_percent():
    frames, _ = simulate_oracle(pendulum())
    assert all_ok(frames)
    crossings, previous = [], None
    for k, (b, _, _) in enumerate(frames):
        x = positions(b)[0][0] - PIVOT[0]
        vx = float(b["motion"][0][0])
        if previous is not None and previous < 0 <= x:
            crossings.append((k + 1) / 60 - (x / max(vx, 1e-3)) / 60)
        previous = x
    assert len(crossings) >= 2
    expected = 2 * math.pi * math.sqrt(100 / 400)
    assert abs((crossings[1] - crossings[0]) - expected) <= expected * 0.02


def rigid_distance_worst(plant=0):
    frames, _ = simulate_oracle(rigid_distance(), plant=plant)
    if not all_ok(frames):
        return math.inf
    return max(abs(math.dist(positions(b)[0], PIVOT) - 100) for b, _, _ in frames)


def test_rigid_distance_keeps_length_within_one_percent():
    assert rigid_distance_worst() <= 100 * 0.01


def test_soft_damped_distance_oscillates_around_rest_length_and_decays():
    frames, _ = simulate_oracle(soft_distance())
    assert all_ok(frames) and len(frames) == 150
    d = [math.dist(positions(b)[0], PIVOT) for b, _, _ in frames]
    early = max(abs(x - 100) for x in d[:45])
    late = max(abs(x - 100) for x in d[105:])
    assert early > 15 and any(x < 100 - 3 for x in d) and late < early * 0.5 and max(d) > 100 + 15


def test_hinge_limits_keep_the_angle_within_range():
    frames, _ = simulate_oracle(hinge_limits())
    assert all_ok(frames)
    rotations = [float(b["pose"][0][2]) for b, _, _ in frames]
    assert min(rotations) >= -0.5 - 0.05 and max(rotations) <= 0.5 + 0.05 and rotations[-1] <= -0.45
    p = positions(frames[-1][0])[0]
    r = rotations[-1]
    anchor = (p[0] - 20 * math.cos(r), p[1] - 20 * math.sin(r))
    assert math.dist(anchor, HINGE_PIVOT) < 1.0


def test_hinge_motor_without_load_reaches_target_velocity():
    frames, _ = simulate_oracle(hinge_motor())
    assert all_ok(frames) and abs(float(frames[-1][0]["motion"][0][2]) - 2) <= 0.1


def test_hinge_motor_drives_the_second_body_relative_to_the_first():
    second, _ = simulate_oracle(hinge_motor_between(False))
    first, _ = simulate_oracle(hinge_motor_between(True))
    assert all_ok(second) and all_ok(first)
    assert abs(float(second[-1][0]["motion"][1][2]) - 2) <= 0.1
    assert abs(float(first[-1][0]["motion"][1][2]) + 2) <= 0.1


def test_hinge_motor_under_load_beyond_maximum_torque_cannot_reach_target():
    weak, _ = simulate_oracle(hinge_motor_load(6_000))
    strong, _ = simulate_oracle(hinge_motor_load(30_000))
    assert all_ok(weak) and all_ok(strong)
    assert float(weak[-1][0]["motion"][0][2]) < 0.5
    assert abs(float(strong[-1][0]["motion"][0][2]) - 1) <= 0.05


def test_weld_keeps_relative_pose():
    frames, _ = simulate_oracle(weld())
    assert all_ok(frames)
    worst_offset = worst_angle = 0.0
    for b, _, _ in frames:
        p, r, _, _ = rigid.public_state(b)
        c, s = math.cos(-r[0]), math.sin(-r[0])
        d = p[1] - p[0]
        local = (d[0] * c - d[1] * s, d[0] * s + d[1] * c)
        worst_offset = max(worst_offset, math.dist(local, (40, 0)))
        worst_angle = max(worst_angle, abs(float(r[1] - r[0])))
    assert worst_offset <= 0.8 and worst_angle <= 0.03
    assert abs(float(rigid.public_state(frames[-1][0])[1][0])) > 0.05, "the pair turned"


def test_connected_bodies_do_not_collide_unless_enabled():
    def run(collide):
        frames, events = simulate_oracle(connected(collide))
        assert all_ok(frames)
        b = frames[-1][0]
        p, r, _, _ = rigid.public_state(b)
        starts = [e for e in events if e["kind"] == abi.BP_EVENT_CONTACT_BEGAN]
        return math.dist(p[0], p[1]), len(starts), len(began(events, 1, 2)), abs(float(r[0])) + abs(float(r[1]))
    off = run(False)
    assert abs(off[0] - 20) <= 1.5 and off[1] == 0 and off[2] == 0 and off[3] <= 0.05
    on = run(True)
    assert on[1] > 0 and on[2] > 0 and on[3] > 0.5 and abs(on[0] - 20) <= 1.5


def test_joint_scene_is_bit_identical_across_two_runs():
    case = joint_determinism()
    a = bp_oracle.run_rigid(case.scene(), case.steps)[0]
    b = bp_oracle.run_rigid(case.scene(), case.steps)[0]
    assert a.tobytes() == b.tobytes()


# ---- VelocityLimitTests.swift -----------------------------------------------------------------
def test_no_object_exceeds_material_speed_limits_under_strong_global_force():
    frames, _ = simulate_oracle(speed_limits())
    assert all_ok(frames) and len(frames) == 30
    lin = np.max([np.hypot(b["motion"][:, 0], b["motion"][:, 1]) for b, _, _ in frames], axis=0)
    ang = np.max([np.abs(b["motion"][:, 2]) for b, _, _ in frames], axis=0)
    assert np.all(lin <= 50 * 1.001) and np.all(ang <= 3 * 1.001)
    ids = list(frames[0][0]["id"])
    assert lin[ids.index(2)] > 50 * 0.9 and ang[ids.index(3)] > 3 * 0.9


def test_kinematic_body_driven_above_the_limit_moves_at_the_limit():
    frames, _ = simulate_oracle(kinematic_speed_limit())
    assert all_ok(frames)
    assert max(float(np.hypot(*b["motion"][0][:2])) for b, _, _ in frames) <= 50 * 1.001
    expected = 50 * len(frames) / 60
    assert abs(float(positions(frames[-1][0])[0][0]) - expected) <= expected * 0.02


def test_material_without_limits_is_not_clamped():
    frames, _ = simulate_oracle(no_speed_limit())
    assert all_ok(frames) and float(np.hypot(*frames[-1][0]["motion"][0][:2])) > 1000


# ---- NativeMetalRigidPolygonTests.swift ---------------------------------------------------------
def test_motor_polygon_integrates_declared_linear_and_angular_velocity():
    frames, _ = simulate_oracle(motor_polygon())
    b = frames[-1][0]
    assert all_ok(frames)
    assert abs(float(positions(b)[0][0]) - 0.5) <= 1e-4 and abs(float(b["pose"][0][2]) - 0.25) <= 1e-4


def test_polygon_contact_keeps_finite_state_and_does_not_move_through_static_wall():
    frames, _ = simulate_oracle(polygon_against_wall())
    assert all_ok(frames)
    for b, _, _ in frames:
        i = by_id(b, 2**63 - 1)
        assert np.all(np.isfinite(b["pose"][i])) and np.all(np.isfinite(b["motion"][i]))
        assert positions(b)[i][0] < 1.1


def test_fast_dynamic_polygons_cannot_swap_positions_without_contact():
    frames, _ = simulate_oracle(polygons_swap())
    b = frames[-1][0]
    assert all_ok(frames)
    a, c = positions(b)[by_id(b, 2**63 - 1)][0], positions(b)[by_id(b, 3)][0]
    assert a < c and c - a >= 2 - 0.05


def test_removing_an_earlier_polygon_keeps_the_remaining_velocity():
    # :83-92: a world re-staged without its first body keeps the other's published motion.
    earlier = rigid.RigidObject(id=3, shape="polygon", vertices=[(-1, -1), (1, -1), (1, 1), (-1, 1)], position=(-10, 0),
                                role="static")
    moving = rigid.RigidObject(id=2**63 - 1, shape="polygon", vertices=[(-1, -1), (1, -1), (1, 1), (-1, 1)],
                               velocity=(2, 0), mass=2.0, moment=2.0, role="dynamic")
    config = cfg((-100, -100, 100, 100), dt=0.25)
    b, _, infos, _ = bp_oracle.run_rigid(rigid.records_for([moving, earlier], config), 1)
    assert infos[0]["status"] == abi.BP_OK
    p, r, v, w = rigid.public_state(b)
    i = by_id(b, 2**63 - 1)
    moving.position, moving.velocity = tuple(p[i]), tuple(v[i])
    b2, _, infos, _ = bp_oracle.run_rigid(rigid.records_for([moving], config), 1)
    assert infos[0]["status"] == abi.BP_OK
    assert abs(float(positions(b2)[0][0]) - 1) <= 1e-4 and abs(float(b2["motion"][0][0]) - 2) <= 1e-4


# ---- closed forms ----------------------------------------------------------------------------
def test_free_flight_integrates_velocity_and_spin():
    o = rigid.RigidObject(id=1, shape="circle", radius=3, position=(50, 60), velocity=(30, -12), angular_velocity=2.5,
                          role="dynamic")
    config = cfg((-1000, -1000, 1000, 1000))
    b, _, infos, _ = bp_oracle.run_rigid(rigid.records_for([o], config), 30)
    assert np.all(infos["status"] == abi.BP_OK)
    t = 30 / 60
    assert np.allclose(positions(b)[0], (50 + 30 * t, 60 - 12 * t), rtol=0, atol=1e-3)
    assert abs(float(b["pose"][0][2]) - 2.5 * t) <= 1e-4


def damping_error(plant=0):
    o = rigid.RigidObject(id=1, shape="circle", radius=3, position=(0, 0), angular_velocity=4.0, damping=1.5, role="dynamic")
    b, _, infos, _ = bp_oracle.run_rigid(rigid.records_for([o], cfg((-1000, -1000, 1000, 1000))), 20, plant=plant)
    assert np.all(infos["status"] == abi.BP_OK)
    return abs(float(b["motion"][0][2]) - 4.0 * math.exp(-1.5 * 20 / 60)) / (4.0 * math.exp(-1.5 * 20 / 60))


def test_angular_damping_follows_the_exponential():
    assert damping_error() <= 1e-5


def overlap_record(shape, plant=0):
    # Two bodies overlapping by 2 along x, one contact iteration, no motion: the pair's record has
    # the unit normal from the lower id to the higher, and lambda = depth / (w_a + w_b) / 2.
    if shape == "box":
        objs = [rigid.box(1, (100, 100), (10, 10), role="dynamic"), rigid.box(2, (118, 100), (10, 10), role="dynamic")]
    else:
        objs = [rigid.circle(1, (100, 100), 10), rigid.circle(2, (118, 100), 10)]
    case = Case("overlap", objs, cfg((0, 0, 400, 400), iterations=1), 1)
    b, _, infos, contacts = bp_oracle.run_rigid(case.scene(), 1, plant=plant)
    assert infos[0]["status"] == abi.BP_OK and len(contacts) == 1
    return b, contacts[0]


@pytest.mark.parametrize("shape", ["box", "circle"])
def test_first_contact_iteration_matches_the_closed_form(shape):
    b, rec = overlap_record(shape)
    assert (int(rec["first"]), int(rec["second"]), int(rec["kind"])) == (1, 2, abi.BP_CONTACT_RIGID)
    assert tuple(rec["normal"]) == (1.0, 0.0)
    assert abs(float(rec["impulse"]) - 2.0 / 2.0 * 0.5) <= 1e-6
    assert abs(float(rec["point"][0]) - 109.0) <= 1e-4 and abs(float(rec["point"][1]) - 100.0) <= 1e-4
    # Each body moves by lambda * w along the normal: the overlap halves.
    assert abs(float(positions(b)[1][0] - positions(b)[0][0]) - 19.0) <= 1e-4


def test_region_wrap_shifts_by_whole_region_sizes():
    # The circle leaves through x = 400 and comes back one region width to the left; the box leaves
    # through y = 400 and comes back one height below; pose and step-start pose move together.
    frames, _ = simulate_oracle(region_exit(True))
    assert all_ok(frames)
    seen = {}
    for fb, _, _ in frames:
        for i in range(len(fb)):
            if int(fb["flags"][i][2]) == abi.BP_REGION_WRAPPED:
                seen.setdefault(int(fb["id"][i]), (fb["pose"][i].copy(), fb["previous"][i].copy()))
    assert set(seen) == {1, 2}
    pose, prev = seen[1]
    assert 0 <= pose[0] <= 30 and abs(float(pose[0] - prev[0]) - 1200 / 60) <= 1e-3
    pose, prev = seen[2]
    assert pose[1] <= 60 and abs(float(pose[1] - prev[1]) - 2400 / 60) <= 1e-3


def test_region_removal_marks_the_bodies_that_left():
    frames, _ = simulate_oracle(region_exit(False))
    assert all_ok(frames)
    marks = {int(b["id"][i]): int(b["flags"][i][2]) for b, _, _ in frames for i in range(len(b))
             if int(b["flags"][i][2]) != 0}
    assert marks == {1: abi.BP_REGION_REMOVED, 2: abi.BP_REGION_REMOVED}


def test_rotation_budget_refuses_the_frame():
    o = rigid.RigidObject(id=1, shape="polygon", vertices=[(-5, -5), (5, -5), (5, 5), (-5, 5)], angular_velocity=500.0,
                          role="dynamic", mass=1.0, moment=10.0)
    scene = rigid.records_for([o], cfg((-1000, -1000, 1000, 1000)))
    before = scene.bodies.copy()
    b, _, infos, _ = bp_oracle.run_rigid(scene, 3)
    assert len(infos) == 1 and infos[0]["status"] == abi.BP_ERR_GUARD_FAILURE
    assert infos[0]["guard_cause"] == abi.BP_GUARD_RIGID_ROTATION_BUDGET
    assert b.tobytes() == before.tobytes(), "a failed frame publishes nothing"


def test_overflowing_motion_is_non_finite():
    o = rigid.RigidObject(id=1, shape="circle", radius=2, velocity=(3e38, 0), role="dynamic")
    b, _, infos, _ = bp_oracle.run_rigid(rigid.records_for([o], cfg((-1000, -1000, 1000, 1000), dt=100.0)), 1)
    assert infos[0]["status"] == abi.BP_ERR_NON_FINITE


# ---- planted mistakes ---------------------------------------------------------------------------
def test_planted_normal_flip_is_caught():
    _, rec = overlap_record("box", plant=1)
    assert tuple(rec["normal"]) != (1.0, 0.0)


def test_planted_dropped_time_of_impact_is_caught():
    assert ccd_head_on_holds() and not ccd_head_on_holds(plant=2)


def test_planted_joint_sign_is_caught():
    assert rigid_distance_worst(plant=3) > 100 * 0.01


def test_planted_wrap_without_region_size_is_caught():
    frames, _ = simulate_oracle(region_exit(True), plant=4)
    wrapped = [fb["pose"][0][0] for fb, _, _ in frames if int(fb["flags"][0][2]) == abi.BP_REGION_WRAPPED]
    assert not wrapped or not (0 <= wrapped[0] <= 30)


def test_planted_dropped_damping_is_caught():
    assert damping_error(plant=5) > 1e-3
