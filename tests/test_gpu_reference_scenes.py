"""GPU parity on the reference's world-level scenes (the scenes that pin the oracle's body obstacle,
body prediction, region pass and contact assertion, tests/test_oracle_reference_scenes.py): the
HIP library steps each scene frame by frame beside the oracle, in each kernel family (the
LDS-resident workgroup kernel through bp_world, the one-workgroup global-state kernel through a
batch, the cooperative grid), and every frame must agree: status, substep count, violation kind and
objects, removed and wrapped ids, and the state of every bubble and body bit for bit."""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi
import reference_scenes as rs
import test_oracle_reference_scenes as scenes_

pytestmark = pytest.mark.gpu

FAMILIES = ["workgroup", "global", "cooperative"]
STATE = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def gpu_runner(scene, family):
    if family == "workgroup":
        return rs.GpuRunner(scene)
    if family == "global":
        return rs.BatchRunner(scene, kernel_class=10)
    scene = scene.copy()
    scene.config.force_cooperative = 1
    return rs.GpuRunner(scene)


def compare(world, oracle, family, frames, label):
    scene = world.build()
    cpu = rs.OracleRunner(scene, oracle)
    gpu = gpu_runner(scene, family)
    for k in range(frames):
        print(f"[gpu-ref] {label} {family} frame {k}", flush=True) if k == 0 else None
        c, g = cpu.step(), gpu.step()
        assert rs.frame_summary(g) == rs.frame_summary(c), (label, family, k, rs.frame_summary(g), rs.frame_summary(c))
        if c["status"] != abi.BP_OK:
            return k + 1
        assert list(g["removed"]) == list(c["removed"]) and list(g["wrapped"]) == list(c["wrapped"]), (label, family, k)
        assert np.array_equal(gpu.scene.bubbles["id"], cpu.scene.bubbles["id"]), (label, family, k)
        for f in STATE:
            a = np.ascontiguousarray(gpu.scene.bubbles[f]).view(np.uint32)
            b = np.ascontiguousarray(cpu.scene.bubbles[f]).view(np.uint32)
            assert np.array_equal(a, b), (label, family, k, f, gpu.scene.bubbles[f], cpu.scene.bubbles[f])
        if len(cpu.scene.bodies):
            a = np.ascontiguousarray(gpu.scene.bodies["pose"]).view(np.uint32)
            b = np.ascontiguousarray(cpu.scene.bodies["pose"]).view(np.uint32)
            assert np.array_equal(a, b), (label, family, k, "body pose")
    return frames


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("kind", ["circle", "capsule", "kinematic"])
def test_pressed_bubble_matches_the_oracle(oracle, family, kind):
    for force in (100.0, 10000.0):
        compare(scenes_.pressed(kind, force), oracle, family, 120, f"{kind} F={force}")


@pytest.mark.parametrize("family", FAMILIES)
def test_corner_filter_and_spinning_body_match_the_oracle(oracle, family):
    diagonal = float(np.float32(math.sqrt(0.5)))
    ds = scenes_.d_star(1200.0)
    start = np.array([-50.0, -50.0]) - np.array([diagonal, diagonal]) * ds
    corner = rs.World().bubble(1, tuple(np.float32(start))).square(2, (0.0, 0.0), 100.0, role="kinematic")
    corner.set_global_force((1200.0 * diagonal, 1200.0 * diagonal))
    compare(corner, oracle, family, 120, "corner")
    for cat, mask in ((1, rs.ALL), (2, 2)):
        w = rs.World().bubble(1, (-200.0, 0.0), velocity=(300.0, 0.0), category=cat, mask=mask).circle(2, (0.0, 0.0), 50.0)
        compare(w, oracle, family, 90, f"filter {cat}")
    spinning = rs.World().bubble(1, (360.0, 0.0)).square(2, (250.0, 0.0), 50.0, role="kinematic", velocity=(1000.0, 0.0),
                                                          angular=20.0)
    compare(spinning, oracle, family, 30, "spinning")


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("name", list(scenes_.INJECTED))
def test_injected_violation_matches_the_oracle(oracle, family, name):
    ran = compare(scenes_.injected_world(name), oracle, family, 5, name)
    assert ran < 5 or name == "never"                  # the world fails within the run


@pytest.mark.parametrize("family", FAMILIES)
def test_filters_static_bubbles_and_non_finite_match_the_oracle(oracle, family):
    pair = rs.injecting(rs.World().bubble(1, (0.0, 0.0), category=1, mask=1).bubble(2, (200.0, 0.0), category=2, mask=2),
                        1, (200.0, 0.0))
    compare(pair, oracle, family, 10, "excluding pair")
    static = rs.injecting(rs.World().bubble(1, (300.0, 0.0), role="static").square(2, (0.0, 0.0), 100.0), 1, (10.0, 10.0))
    compare(static, oracle, family, 10, "static in body")
    nan = rs.injecting(rs.World().bubble(1, (0.0, 0.0)), 1, (float("nan"), 0.0))
    assert compare(nan, oracle, family, 5, "non-finite") == 1


@pytest.mark.parametrize("family", FAMILIES)
def test_wraps_match_the_oracle(oracle, family):
    compare(scenes_.wrap_scene(), oracle, family, 30, "wrap free")
    compare(scenes_.wrap_scene(lambda w: w.bubble(2, (30.0, -988.0))), oracle, family, 30, "wrap overlap")
    compare(scenes_.wrap_scene(lambda w: w.square(2, (0.0, -988.0), 20.0)), oracle, family, 5, "wrap into body")


@pytest.mark.parametrize("family", FAMILIES)
def test_region_removal_and_wrap_match_the_oracle(oracle, family):
    compare(scenes_.region_world((10.8, 10.8)), oracle, family, 2, "corner removal")
    compare(scenes_.region_world((-25.0, 5.0), (-4.0, 0.0), wrap=True), oracle, family, 3, "overshoot")
    compare(scenes_.region_world((10.8, 10.8), wrap=True), oracle, family, 3, "diagonal wrap")
    seam = scenes_.region_world((9.0, 5.0), (4.0, 0.0), wrap=True).bubble(2, (0.25, 5.0), radius=1.0)
    compare(seam, oracle, family, 4, "seam")
    # Removal of one of several bubbles, then further steps of the rest (the deletion keeps the
    # others' order and state).
    crowd = rs.World(region=(0.0, 0.0, 100.0, 100.0), dt=1.0 / 60.0)
    for k in range(12):
        crowd.bubble(k + 1, (10.0 + 7.0 * k, 50.0 + (k % 3)), radius=3.0, velocity=(0.0, 0.0) if k != 5 else (0.0, 9000.0))
    compare(crowd, oracle, family, 6, "crowd removal")


@pytest.mark.parametrize("family", ["workgroup", "global"])
def test_squeeze_to_the_substep_limit_matches_the_oracle(oracle, family):
    w = rs.World().bubble(1, (300.0, 0.0)).wall(2, 360.0).square(3, (150.0, 0.0), 100.0, role="kinematic",
                                                                    velocity=(100.0, 0.0))
    ran = compare(w, oracle, family, 46 + 60 + 5, "squeeze")
    assert ran < 46 + 60 + 5


def test_region_ids_through_the_c_abi():
    # NativeSolverFrame.regionRemovals / regionWraps (NativeSolverInput.swift:243-244): a removal
    # and a wrap in one step; the solver's frame keeps the removed bubble until the caller deletes it.
    from bubblephysics_amd import solver
    w = rs.World(region=(0.0, 0.0, 10.0, 10.0), dt=0.25)
    w.bubble(1, (10.8, 10.8), radius=1.0).bubble(2, (-25.0, 5.0), radius=1.0, velocity=(-4.0, 0.0), wrap=True)
    w.bubble(3, (5.0, 5.0), radius=1.0)
    world = solver.NativeWorld(w.build())
    frame = world.step(0)
    assert frame.status == abi.BP_OK
    assert frame.region_removals == {1} and frame.region_wraps == {2}
    assert len(frame.bubbles) == 3                         # still in the published frame
    removed, wrapped = world.region_results()
    assert list(removed) == [1] and list(wrapped) == [2]
    world.remove_bubbles(removed)
    assert list(world.read_bubbles()["id"]) == [2, 3]
    frame = world.step(1)
    assert frame.status == abi.BP_OK and frame.region_removals == set() and list(frame.bubbles["id"]) == [2, 3]


# ---- StaticPolygonEdgeTests scenes (tests/test_oracle_polygon_edges.py) --------------------------
import test_oracle_polygon_edges as edges_


def polygon_world(kind):
    if kind == "thin wall":
        w = edges_.world(1000.0, [edges_.rectangle(-100.0, 48.0)])
        return w.bubble(1, (48.0, 0.0), radius=148.0, mass=1.0).motor(1, (-2500.0, 0.0))
    if kind == "collinear":
        two = [(-148.0, -400.0), (-100.0, -400.0), (-100.0, 0.0), (-100.0, 400.0), (-148.0, 400.0)]
        return edges_.world(1000.0, [two]).bubble(1, (-76.0, 0.0), radius=24.0, mass=1.36).motor(1, (-600.0, 0.0))
    if kind == "winding":
        return edges_.winding_worlds(True)
    if kind == "fast":
        v = (float(np.float32(-math.cos(math.pi / 6) * 6000.0)), float(np.float32(math.sin(math.pi / 6) * 6000.0)))
        return edges_.world(2500.0, [edges_.rectangle(0.0, 48.0)]).bubble(1, (200.0, 0.0), radius=12.0,
                                                                         mass=edges_.crowd_mass(12.0), velocity=v)
    return edges_.world(2500.0, [edges_.rectangle(-100.0, 48.0)]).bubble(1, (-220.0, 510.0), radius=12.0,
                                                                        mass=edges_.crowd_mass(12.0),
                                                                        velocity=(1200.0, -1800.0))


@pytest.mark.parametrize("family", FAMILIES)
@pytest.mark.parametrize("kind", ["thin wall", "collinear", "winding", "fast", "oblique"])
def test_polygon_edge_scenes_match_the_oracle(oracle, family, kind):
    frames = {"thin wall": 60, "collinear": 60, "winding": 60, "fast": 40, "oblique": 6}[kind]
    compare(polygon_world(kind), oracle, family, frames, kind)


@pytest.mark.parametrize("family", ["workgroup", "cooperative"])
def test_convex_vertex_with_a_drive_set_every_frame_matches_the_oracle(oracle, family):
    # StaticPolygonEdgeTests.swift:166-185: the motor force changes before every frame.
    w = edges_.world(1000.0, [edges_.wedge(90.0)]).bubble(1, (-24.0, 0.0), radius=24.0, mass=1.36)
    scene = w.build()
    cpu, gpu = rs.OracleRunner(scene, oracle), gpu_runner(scene, family)
    for k in range(120):
        position = cpu.position(1)
        force = tuple(np.float32(edges_.unit(-position) * 600.0)) if k else (0.0, 0.0)
        if k:
            cpu.set_motor(1, force)
            gpu.set_motor(1, force)
        c, g = cpu.step(), gpu.step()
        assert rs.frame_summary(g) == rs.frame_summary(c), k
        for f in STATE:
            assert np.array_equal(np.ascontiguousarray(gpu.scene.bubbles[f]).view(np.uint32),
                                  np.ascontiguousarray(cpu.scene.bubbles[f]).view(np.uint32)), (k, f)
