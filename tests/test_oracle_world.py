"""World-level pins of the oracle: the h rule, the integrator and the 1D scenes of the reference's
substep-kernel tests (Tests/BubblePhysicsNativeMetalTests/NativeMetalSubstepKernelTests.swift),
each checked against the closed form or the Double reference the reference's test uses, with the
same tolerance. The oracle's per-substep trace gives h, the limiting term and its bubble for
every substep, which the reference reads from its substep control buffer."""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi
from kernel_scenes import DT, body, relative, world
import contact_law_reference as ref

TERM = {"remaining": abi.BP_TERM_REMAINING, "stiffness": abi.BP_TERM_STIFFNESS, "distance": abi.BP_TERM_DISTANCE,
        "margin": abi.BP_TERM_MARGIN}


def wall_stiffness(distance, radius, stiffness):
    """dF/dx of the wall law by central difference of the Double force (:241-246)."""
    eps = 1 - distance / radius
    h = 1e-6 * (1 - eps)
    return (ref.force(eps + h, stiffness) - ref.force(eps - h, stiffness)) / (2 * h * radius)


def pair_stiffness(overlap, ra, ka, rb, kb):
    """dF/d(overlap) of the Double pair solution (:249-254)."""
    h = 1e-5 * min(ra, rb)
    return (ref.pair_contact(overlap + h, ra, ka, rb, kb).force - ref.pair_contact(overlap - h, ra, ka, rb, kb).force) / (2 * h)


def distance_bound(distance, speed, acceleration_sum):
    """h = d / (2 (s + v)), s = sqrt(v^2 + A d / 2) (:575-579)."""
    s = math.sqrt(speed * speed + acceleration_sum * distance / 2)
    return distance / (2 * (s + speed))


def first_substep(oracle, scene):
    _, _, frames, ex = oracle.run(scene, 1, history=True)
    return frames[0], ex


def test_no_contacts_take_the_whole_step_in_one_substep(oracle):
    # :451-465
    f, ex = first_substep(oracle, world([body(41, (0.0, 0.0), velocity=(5.0, 0.0)), body(42, (500.0, 0.0))]))
    assert f["status"] == 0 and f["substeps"] == 1 and f["rebuilds"] == 0
    assert ex["history"][0] == DT
    assert ex["terms"][0] == TERM["remaining"]


@pytest.mark.parametrize("distance,mass,stiffness", [(2.0, 0.36, 229.0), (0.4, 38.0, 229.0), (3.0, 1.0, 22900.0)])
def test_wall_contact_limits_h_to_one_over_omega(oracle, distance, mass, stiffness):
    # :489-506: omega^2 = k / m with k the derivative of the wall law.
    f, ex = first_substep(oracle, world([body(7, (100.0, -distance), mass=mass, stiffness=stiffness)], walls=True))
    expected = 1 / math.sqrt(wall_stiffness(distance, 12.0, stiffness) / mass)
    assert ex["terms"][0] == TERM["stiffness"] and ex["objects"][0] == 7
    assert relative(ex["history"][0], expected) <= 1e-3
    assert ex["history"][0] < DT


@pytest.mark.parametrize("winner", [0, 1, 2, 31, 32, 64, 95, 96, 99, 100, 126, 127, 128, 129])
def test_reduction_finds_the_winner_at_any_index(oracle, winner):
    # :508-527: 130 bubbles on the wall, the light one sets h. Bubbles do not see each other
    # (the Rig gives them wall rows only).
    bodies = [body(i + 1, (100.0, -2.0), mass=0.36 if i == winner else 38.0, category=2, mask=1) for i in range(130)]
    f, ex = first_substep(oracle, world(bodies, walls=True))
    expected = 1 / math.sqrt(wall_stiffness(2.0, 12.0, 229.0) / 0.36)
    assert ex["objects"][0] == winner + 1
    assert relative(ex["history"][0], expected) <= 1e-3


def test_pair_stiffness_term_is_gershgorin_with_the_lighter_bubble_owning_omega(oracle):
    # :528-538: omega_1^2 = k (1/1 + 1/sqrt(4)).
    f, ex = first_substep(oracle, world([body(5, (0.0, 0.0), mass=1.0), body(9, (3.0, 0.0), mass=4.0)]))
    k = pair_stiffness(21.0, 12.0, 229.0, 12.0, 229.0)
    assert ex["terms"][0] == TERM["stiffness"] and ex["objects"][0] == 5
    assert relative(ex["history"][0], 1 / math.sqrt(k * 1.5)) <= 2e-3


def test_static_neighbour_adds_no_partner_term(oracle):
    # :540-548
    f, ex = first_substep(oracle, world([body(1, (0.0, 0.0), mass=1.0), body(2, (3.0, 0.0), mass=1.0, role="static")]))
    k = pair_stiffness(21.0, 12.0, 229.0, 12.0, 229.0)
    assert ex["objects"][0] == 1
    assert relative(ex["history"][0], 1 / math.sqrt(k)) <= 2e-3


def test_pair_of_different_materials_is_two_springs_in_series(oracle):
    # :550-566: k = dF/d(overlap) of the bisection solution; equal masses 1 give omega^2 = 2k.
    f, ex = first_substep(oracle, world([body(3, (0.0, 0.0), mass=1.0),
                                         body(4, (6.0, 0.0), mass=1.0, radius=20.0, stiffness=22900.0)]))
    k = pair_stiffness(26.0, 12.0, 229.0, 20.0, 22900.0)
    assert ex["terms"][0] == TERM["stiffness"]
    assert relative(ex["history"][0], 1 / math.sqrt(2 * k)) <= 5e-3


def test_approaching_candidate_is_limited_to_a_quarter_of_the_distance(oracle):
    # :581-589
    f, ex = first_substep(oracle, world([body(21, (100.0, -14.0), velocity=(0.0, 1000.0))], walls=True))
    assert ex["terms"][0] == TERM["distance"] and ex["objects"][0] == 21
    assert relative(ex["history"][0], 14.0 / 4000) <= 1e-4


def test_receding_candidate_is_exempt_unless_it_turns_back_within_the_step(oracle):
    # :591-610
    f, ex = first_substep(oracle, world([body(1, (100.0, -14.0), velocity=(0.0, -1000.0))], walls=True))
    assert f["substeps"] == 1 and ex["history"][0] == DT and ex["terms"][0] == TERM["remaining"]
    f, ex = first_substep(oracle, world([body(2, (100.0, -14.0), velocity=(0.0, -100.0), mass=1.0)], walls=True,
                                        global_force=(0.0, 1e6)))
    assert -100 + 1e6 * (1 / 60) / 2 > 0
    assert ex["terms"][0] == TERM["distance"]
    assert relative(ex["history"][0], distance_bound(14.0, -100.0, 1e6)) <= 1e-3


def test_active_contact_bound_is_the_same_in_both_directions(oracle):
    # :612-625: d = 8 < r = 12, same h approaching and receding.
    hs = []
    for v in (1000.0, -1000.0):
        f, ex = first_substep(oracle, world([body(6, (100.0, -8.0), velocity=(0.0, v))], walls=True))
        a = ref.force(1 - 8.0 / 12, 229.0) / 0.36
        assert ex["terms"][0] == TERM["distance"]
        assert relative(ex["history"][0], distance_bound(8.0, 1000.0, a)) <= 2e-3
        hs.append(ex["history"][0])
    assert hs[0] == hs[1]


def test_pair_bound_uses_the_relative_normal_velocity(oracle):
    # :627-636: tangential velocities do not enter v_n.
    f, ex = first_substep(oracle, world([body(1, (0.0, 0.0), velocity=(400.0, 3.0), mass=1.0),
                                         body(2, (30.0, 0.0), velocity=(-600.0, -7.0), mass=1.0)]))
    assert ex["terms"][0] == TERM["distance"]
    assert relative(ex["history"][0], 30.0 / 4000) <= 1e-4


def test_equal_bounds_go_to_the_lower_index(oracle):
    # :638-646
    f, ex = first_substep(oracle, world([body(31, (60.0, -14.0), velocity=(0.0, 1000.0)),
                                         body(32, (140.0, -14.0), velocity=(0.0, 1000.0)),
                                         body(33, (100.0, -14.0), velocity=(0.0, 100.0))], walls=True))
    assert ex["objects"][0] == 31
    assert relative(ex["history"][0], 14.0 / 4000) <= 1e-4


@pytest.mark.parametrize("split,status,substeps", [(1024, abi.BP_OK, 1024), (1025, abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED, 1024)])
def test_exactly_1024_substeps_are_allowed(oracle, split, status, substeps):
    # :648-680: the limit is raised only when time is left after the 1024th substep.
    f, ex = first_substep(oracle, world([body(1, (0.0, 0.0))], split=split))
    assert f["status"] == status and f["substeps"] == substeps


def test_forced_split_divides_the_step_into_equal_parts(oracle):
    # :392-414
    results = []
    for multiplier in (1, 4, 7):
        scene = world([body(1, (0.0, 0.0), velocity=(3.0, 0.0), mass=0.5, damping=1.0)], global_force=(3.0, 0.0),
                      split=multiplier)
        b, _, frames, ex = oracle.run(scene, 1, history=True)
        h = ex["history"][: frames[0]["substeps"]].astype(np.float64)
        results.append((int(frames[0]["substeps"]), float(b["velocity"][0, 0]), float(b["center"][0, 0]), h))
    assert [r[0] for r in results] == [1, 4, 7]
    for r in results[1:]:
        assert abs(r[3].sum() - 1 / 60) <= 1e-7
        assert r[3].max() - r[3].min() <= 1e-7
        assert abs(r[1] - results[0][1]) <= 1e-4 and abs(r[2] - results[0][2]) <= 1e-4


def test_minimum_margin_only_keeps_a_quarter_radius_and_the_margin_term_rebuilds(oracle):
    # :681-698 and :700-723: margin r/4 = 3 at 3000 units/s gives h = 3/3000 and a rebuild flag.
    f, ex = first_substep(oracle, world([body(1, (0.0, 0.0), velocity=(3000.0, 0.0))], minimum_margin_only=1))
    assert ex["terms"][0] == TERM["margin"] and ex["objects"][0] == 1
    assert relative(ex["history"][0], 3.0 / 3000) <= 1e-3
    assert f["rebuilds"] >= 1 and f["status"] == 0
    # Without the switch the motion term (1.05 * 3000 / 60 = 52.5) covers the whole step.
    f, ex = first_substep(oracle, world([body(1, (0.0, 0.0), velocity=(3000.0, 0.0))]))
    assert f["substeps"] == 1 and f["rebuilds"] == 0


@pytest.mark.parametrize("gamma", [0.0, 1.0, 5.0, 40.0])
def test_exact_half_impulses_under_constant_force(oracle, gamma):
    # :793-812: v1 = v0 e^{-gh} + a (1 - e^{-gh}) / g, x1 = h * v(h/2).
    scene = world([body(1, (0.0, 0.0), velocity=(10.0, -4.0), mass=0.5, damping=gamma)], global_force=(3.0, 0.0))
    b, _, frames, _ = oracle.run(scene, 1)
    assert frames[0]["substeps"] == 1
    h, a, g = 1 / 60, 6.0, gamma
    decay = math.exp(-g * h)
    vx = 10 + a * h if g == 0 else 10 * decay + a * (1 - decay) / g
    vy = -4 * decay
    assert abs(b["velocity"][0, 0] - vx) <= 2e-5 * max(1, abs(vx))
    assert abs(b["velocity"][0, 1] - vy) <= 2e-5 * max(1, abs(vy))
    half = math.exp(-g * h / 2)
    mid = 10 + a * h / 2 if g == 0 else 10 * half + a * (1 - half) / g
    assert abs(b["center"][0, 0] - h * mid) <= 1e-5


@pytest.mark.parametrize("step", [1 / 60, 4 / 60])
def test_terminal_velocity_is_force_over_drag_at_any_step(oracle, step):
    # :814-822
    scene = world([body(1, (0.0, 0.0), mass=0.5, damping=5.0)], global_force=(3.0, 0.0), dt=step)
    b, _, frames, _ = oracle.run(scene, int(round(10 / step)))
    assert (frames["status"] == 0).all()
    assert abs(b["velocity"][0, 0] - 1.2) <= 1e-4


def test_roles_path_moves_with_its_velocity_static_stays_force_driven_feels_forces(oracle):
    # :824-837. The world step also runs nativeBeginStep (Force:285-293), which sets a static
    # bubble's velocity to zero; the kernel test checks only that integration leaves it alone.
    scene = world([body(1, (0.0, 0.0), velocity=(20.0, 0.0), damping=5.0, role="path"),
                   body(2, (50.0, 0.0), velocity=(20.0, 0.0), damping=5.0, role="static"),
                   body(3, (100.0, 0.0), velocity=(20.0, 0.0), damping=5.0)], global_force=(100.0, 0.0))
    b, _, _, _ = oracle.run(scene, 1)
    h = np.float32(1 / 60)
    assert b["velocity"][0, 0] == 20.0 and abs(b["center"][0, 0] - 20 * h) <= 1e-5
    assert b["center"][1, 0] == 50.0 and b["velocity"][1, 0] == 0.0
    assert b["velocity"][2, 0] < 100 and b["velocity"][2, 0] != 20.0


def test_rotation_advances_and_decays_with_drag(oracle):
    # :839-850
    scene = world([body(1, (0.0, 0.0), damping=3.0)])
    scene.bubbles["angular_velocity"][0] = 2.0
    b, _, _, _ = oracle.run(scene, 1)
    h = 1 / 60
    assert abs(b["rotation"][0] - 2 * h) <= 1e-6
    assert abs(b["angular_velocity"][0] - 2 * math.exp(-3 * h)) <= 1e-5


@pytest.mark.parametrize("gamma", [0.0, 2.0])
def test_motor_torque_spins_force_driven_bubbles_only(oracle, gamma):
    # :378-390: I = 9, M = 4.5: omega = M / (I g) (1 - e^{-g h}).
    scene = world([body(1, (0.0, 0.0), mass=2.0, damping=gamma), body(2, (50.0, 0.0), mass=2.0, role="path")],
                  motors=[(1, 0.0, 0.0, 4.5, 9.0), (2, 0.0, 0.0, 4.5, 9.0)])
    b, _, _, _ = oracle.run(scene, 1)
    h = 1 / 60
    expected = 0.5 * h if gamma == 0 else 0.5 / gamma * (1 - math.exp(-gamma * h))
    assert abs(b["angular_velocity"][0] - expected) <= 1e-6
    assert b["angular_velocity"][1] == 0.0


def test_motor_and_global_force_set_the_acceleration(oracle):
    # :319-343 (motor part): force (6, -8) plus global (1, 1) on mass 2, no contacts: v = a h.
    scene = world([body(3, (10.0, 20.0), mass=2.0), body(5, (500.0, 20.0), mass=1.0)],
                  motors=[(3, 6.0, -8.0, 2.0, 1.0)], global_force=(1.0, 1.0))
    b, _, frames, _ = oracle.run(scene, 1)
    h = 1 / 60
    assert frames[0]["substeps"] == 1
    assert abs(b["velocity"][0, 0] / h - 3.5) <= 1e-3 and abs(b["velocity"][0, 1] / h + 3.5) <= 1e-3
    assert abs(b["velocity"][1, 0] / h - 1.0) <= 1e-3 and abs(b["velocity"][1, 1] / h - 1.0) <= 1e-3


def test_band_follows_the_harmonic_oscillator(oracle):
    # :345-364: K = 40000, m = 1, split 4: x(t) = 50 (1 - cos wt).
    scene = world([body(1, (0.0, 0.0), mass=1.0)], bands=[(1, 50.0, 0.0, 40000.0, 0.0)], split=4)
    b, _, frames, _ = oracle.run(scene, 1)
    assert frames[0]["substeps"] >= 12
    w, t = 200.0, 1 / 60
    assert abs(b["center"][0, 0] - 50 * (1 - math.cos(w * t))) <= 0.03 * 50
    assert abs(b["velocity"][0, 0] - 50 * w * math.sin(w * t)) <= 0.05 * 50 * w


def test_band_stiffness_limits_h_through_omega(oracle):
    # :366-376: h = sqrt(m / K).
    f, ex = first_substep(oracle, world([body(1, (0.0, 0.0), mass=0.36)], bands=[(1, 0.0, 0.0, 300000.0, 0.0)]))
    assert ex["terms"][0] == TERM["stiffness"] and ex["objects"][0] == 1
    assert abs(ex["history"][0] - math.sqrt(0.36 / 300000.0)) <= 1e-5


@pytest.mark.parametrize("speed", [1000.0, 3000.0])
def test_bounce_on_the_wall_matches_the_1d_h_rule_simulation(oracle, speed):
    # :884-916: per frame within 0.3 units and 3% of the speed, rebound within 2%, peak
    # substeps within 25% of the 1D simulation.
    scene = world([body(1, (100.0, -80.0), velocity=(0.0, speed))], walls=True)
    wall = ref.WallScene(radius=12, mass=0.36, stiffness=229, damping=0, start_distance=80, velocity=-speed).run(40)
    assert not wall.penetrated
    cur, counts, rebound = scene.copy(), [], None
    for frame in range(min(40, len(wall.frames))):
        b, _, frames, _ = oracle.run(cur, 1)
        assert frames[0]["status"] == 0
        counts.append(int(frames[0]["substeps"]))
        distance = -float(b["center"][0, 1])
        assert distance > 0
        assert abs(distance - wall.frames[frame][1]) <= 0.3
        assert abs(-float(b["velocity"][0, 1]) - wall.frames[frame][2]) <= 0.03 * speed
        cur.bubbles = b
        if b["velocity"][0, 1] < 0 and distance > 12:
            rebound = -float(b["velocity"][0, 1])
            break
    assert rebound is not None
    assert abs(rebound - speed) / speed <= 0.02
    peak, ref_peak = max(counts), max(f[0] for f in wall.frames)
    assert abs(peak - ref_peak) <= 0.25 * ref_peak


def test_substep_history_sums_to_the_step(oracle):
    # :918-948: the h of a frame with many substeps sums to dt within 1e-6.
    cur = world([body(1, (100.0, -60.0), velocity=(0.0, 1000.0))], walls=True)
    deepest = 0
    for _ in range(10):
        b, _, frames, ex = oracle.run(cur, 1, history=True)
        count = int(frames[0]["substeps"])
        h = ex["history"][:count].astype(np.float64)
        assert (h > 0).all() and abs(h.sum() - 1 / 60) <= 1e-6
        deepest = max(deepest, count)
        cur.bubbles = b
    assert deepest > 10


def test_pair_bounce_conserves_momentum_and_energy(oracle):
    # :957-978: momentum within 2e-4 * |p| * 10, energy within 2% after separation.
    scene = world([body(1, (0.0, 0.0), velocity=(600.0, 0.0)), body(2, (60.0, 0.0), velocity=(-400.0, 0.0), mass=1.5)])
    p0 = 0.36 * 600 - 1.5 * 400
    e0 = 0.5 * 0.36 * 600 ** 2 + 0.5 * 1.5 * 400 ** 2
    cur, separated = scene.copy(), False
    for _ in range(12):
        b, _, frames, _ = oracle.run(cur, 1)
        assert frames[0]["status"] == 0
        v = b["velocity"][:, 0].astype(np.float64)
        assert abs(0.36 * v[0] + 1.5 * v[1] - p0) <= 2e-4 * max(1, abs(p0)) * 10
        assert b["center"][0, 0] < b["center"][1, 0]
        cur.bubbles = b
        if b["center"][1, 0] - b["center"][0, 0] > 24 and v[0] < v[1]:
            separated = True
            break
    assert separated
    e = 0.5 * 0.36 * v[0] ** 2 + 0.5 * 1.5 * v[1] ** 2
    assert abs(e - e0) / e0 <= 0.02


def test_planted_sign_mistake_in_the_distance_bound_fails_the_closed_form():
    # A bound with the speed subtracted instead of added misses the approach case by far more
    # than the 1e-4 tolerance used above, so the check cannot pass a mistaken restatement.
    wrong = 14.0 / (2 * (math.sqrt(1000.0 ** 2) - 1000.0 + 1e-30))
    assert relative(wrong, 14.0 / 4000) > 1e-4
    assert relative(distance_bound(14.0, 1000.0, 0.0), 14.0 / 4000) <= 1e-12


def _radial(strength, falloff, radius, point):
    """ContactLawReference.RadialField.force (ContactLawReference.swift:186-197)."""
    d = math.hypot(point[0], point[1])
    if not (d > 0 and d <= radius):
        return np.zeros(2)
    profile = 1.0 if falloff == 0 else max(0.0, 1 - d / radius) ** falloff
    return np.asarray(point, np.float64) * (strength * profile / d)


def test_fields_global_force_filters_roles_and_mass_set_a0(oracle):
    # NativeMetalSubstepKernelTests.swift:263-308: directional (rotated), radial with falloff,
    # vortex in a rotated box, a filtered field acting on id 2 only, an empty filter acting on none.
    from kernel_scenes import reference_fields
    bodies = [body(1, (10.0, 20.0), mass=0.5), body(2, (10.0, 20.0), mass=2.0), body(3, (10.0, 20.0), role="path"),
              body(4, (10.0, 20.0), role="static"), body(5, (400.0, 0.0), mass=1.0)]
    scene = world(bodies, fields=reference_fields(), global_force=(1.0, -2.0))
    got = oracle.noncontact_acceleration(scene).astype(np.float64)

    def expected(id, mass, point):
        force = np.array([1.0, -2.0])
        c, s = math.cos(0.5), math.sin(0.5)
        delta = np.asarray(point) - np.array([5.0, -5.0])
        local = np.array([c * delta[0] + s * delta[1], -s * delta[0] + c * delta[1]])
        if np.hypot(*local) <= 100:
            force += np.array([c * 30 - s * -40, s * 30 + c * -40])
        force += _radial(-50, 2, 80, point)
        c3, s3 = math.cos(0.3), math.sin(0.3)
        d3 = np.asarray(point) - np.array([5.0, 5.0])
        l3 = np.array([c3 * d3[0] + s3 * d3[1], -s3 * d3[0] + c3 * d3[1]])
        if abs(l3[0]) <= 50 and abs(l3[1]) <= 50 and np.hypot(*d3) > 0:
            r = 50 * math.sqrt(2)
            scale = 10 * max(0.0, 1 - np.hypot(*l3) / r) ** 1 / np.hypot(*l3)
            force += np.array([-d3[1], d3[0]]) * scale
        if id == 2:
            force += _radial(7, 0, 200, point)
        return force / mass

    for index, mass in [(0, 0.5), (1, 2.0), (4, 1.0)]:
        want = expected(int(scene.bubbles["id"][index]), mass, scene.bubbles["center"][index].astype(np.float64))
        assert np.linalg.norm(got[index] - want) / max(np.linalg.norm(want), 1e-9) <= 1e-4, index
    assert np.any(got[0] != 0)
    assert np.all(got[2] == 0) and np.all(got[3] == 0)


def test_planted_field_rotation_mistake_is_caught():
    # Turning the directional force by -0.5 instead of +0.5 moves it by far more than 1e-4.
    c, s = math.cos(0.5), math.sin(0.5)
    right = np.array([c * 30 - s * -40, s * 30 + c * -40])
    wrong = np.array([c * 30 + s * -40, -s * 30 + c * -40])
    assert np.linalg.norm(right - wrong) / np.linalg.norm(right) > 1e-4


def test_per_phase_threads_give_the_same_bits(oracle):
    # The per-phase CPU baseline runs the per-bubble phases and the candidate rows on several host
    # threads; every thread count must give the sequential run's bits (state, frames, rows).
    from bubblephysics_amd import scenes
    for scene, steps in ((scenes.config1(), 12), (scenes.config2(), 6)):
        runs = []
        for threads in (1, 4):
            oracle.set_threads(threads)
            try:
                cb, cbod, cf, ex = oracle.run(scene, steps, adjacency=True)
            finally:
                oracle.set_threads(1)
            runs.append((cb.tobytes(), cbod.tobytes(), cf.tobytes(), ex["row_offsets"].tobytes(), ex["entries"].tobytes()))
        assert runs[0] == runs[1]
