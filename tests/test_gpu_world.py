"""GPU parity: the HIP world step against the CPU oracle on the same seeded worlds, through the
C ABI. Integer and index results (substep counts, rebuilds, candidate pairs, adjacency rows) must be
equal; floating-point state is compared bitwise, which holds because kernels and oracle round every
operation identically (the tolerance the step guarantees is 1e-5 relative; bitwise is stricter)."""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, solver
import kernel_scenes

pytestmark = pytest.mark.gpu

STATE_FIELDS = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_same_state(gpu, cpu, label=""):
    for f in STATE_FIELDS:
        g, c = bits(gpu[f]), bits(cpu[f])
        if not np.array_equal(g, c):
            bad = np.nonzero((g != c).reshape(len(gpu), -1).any(axis=1))[0]
            raise AssertionError(f"{label} field {f} differs at bubbles {bad[:8]}: "
                                 f"gpu {gpu[f][bad[:2]]} cpu {cpu[f][bad[:2]]}")


def run_both(scene, steps, oracle, *, check_adjacency=False):
    world = solver.NativeWorld(scene)
    cur = scene.copy()
    for k in range(steps):
        frame = world.step(k)
        cb, cbod, cf, ex = oracle.run(cur, 1, adjacency=check_adjacency)
        c = cf[0]
        assert frame.status == c["status"], f"step {k}"
        assert frame.substep_count == c["substeps"], f"step {k}"
        assert frame.grid_rebuild_count == c["rebuilds"], f"step {k}"
        assert frame.candidate_pair_count == c["candidate_pairs"], f"step {k}"
        assert frame.maximum_contact_candidates_per_object == c["max_candidates"], f"step {k}"
        if frame.status != abi.BP_OK:
            return world, frame, c
        assert frame.limiter_term == c["limiter_term"] and frame.limiter_object == c["limiter_object"], f"step {k}"
        assert_same_state(frame.bubbles, cb, f"step {k}")
        if check_adjacency:
            off, ent = world.read_adjacency()
            assert np.array_equal(off, ex["row_offsets"]), f"step {k}"
            assert np.array_equal(ent, ex["entries"]), f"step {k}"
        gb = world.read_bodies()
        assert np.array_equal(bits(gb["pose"]), bits(cbod["pose"])), f"step {k}"
        cur.bubbles, cur.bodies = cb, cbod
    return world, None, None


def test_config1_plain_world_matches_oracle_bitwise(oracle):
    run_both(scenes.config1(), 120, oracle, check_adjacency=True)


def test_config2_bodies_match_oracle_bitwise(oracle):
    run_both(scenes.config2(), 25, oracle, check_adjacency=True)


def test_large_world_cooperative_grid_matches_oracle(oracle):
    # Above the widest LDS configuration a world runs on a cooperative grid of workgroups.
    rng = np.random.Generator(np.random.PCG64(7))
    scene = scenes.batch_world(rng, 5000, bodies_max=8)
    world, _, _ = run_both(scene, 8, oracle, check_adjacency=True)
    assert world.step(8).execution_path == abi.BP_PATH_COOPERATIVE


def test_large_world_in_one_workgroup_matches_oracle(oracle):
    # The single-workgroup global-state kernel (class 10) on the same world.
    rng = np.random.Generator(np.random.PCG64(7))
    scene = scenes.batch_world(rng, 5000, bodies_max=8)
    batch = solver.Batch([scene])
    batch.set_world_class(0, 10)
    infos = batch.step(4)
    cb, _, cf, _ = oracle.run(scene, 4)
    assert (infos[:, 0]["kernel_class"] == 10).all()
    assert np.array_equal(infos[:, 0]["substep_count"], cf["substeps"])
    assert_same_state(batch.read_bubbles(0), cb, "one workgroup")


def test_small_world_forced_onto_the_cooperative_grid_matches_oracle(oracle):
    # force_cooperative: the reference's multi-dispatch path for a small world (sorted grid in
    # the oracle, cooperative grid here); results are the same bits.
    scene = scenes.config1()
    scene.config.force_cooperative = 1
    world, _, _ = run_both(scene, 12, oracle, check_adjacency=True)
    assert world.step(12).execution_path == abi.BP_PATH_COOPERATIVE


def test_cooperative_world_with_body_tree_matches_oracle(oracle):
    # 20,000 bubbles and 300 kinematic bodies: body rows come from the LBVH of the sweep boxes.
    n = 20000
    scene = scenes.config3(n=n, bodies=300, side=16384.0 * math.sqrt(n / 262144))
    run_both(scene, 3, oracle, check_adjacency=True)


def test_hash_bucket_count_does_not_change_the_step(oracle):
    # InGroupGridCollisionTests.swift:34-41: one bucket for every cell gives the same result.
    base = scenes.config1()
    one = base.copy()
    one.config.grid_bucket_count = 1
    odd = base.copy()
    odd.config.grid_bucket_count = 7
    results = []
    for s in (base, one, odd):
        w = solver.NativeWorld(s)
        for k in range(10):
            f = w.step(k)
        results.append(f.bubbles)
    assert_same_state(results[0], results[1], "1 bucket")
    assert_same_state(results[0], results[2], "7 buckets")


def test_batch_of_uneven_worlds_matches_oracle_and_plain_world(oracle):
    rng = np.random.Generator(np.random.PCG64(11))
    sizes = [16, 60, 200, 300, 600, 1000, 1500]
    worlds = [scenes.batch_world(rng, n, bodies_max=6) for n in sizes]
    batch = solver.Batch(worlds)
    steps = 6
    infos = batch.step(steps)
    for w, scene in enumerate(worlds):
        cb, _, cf, _ = oracle.run(scene, steps)
        assert np.array_equal(infos[:, w]["status"], cf["status"]), w
        assert np.array_equal(infos[:, w]["substep_count"], cf["substeps"]), w
        assert_same_state(batch.read_bubbles(w), cb, f"world {w}")
    # The same world alone gives the same bits.
    alone = solver.NativeWorld(worlds[3])
    for k in range(steps):
        alone.step(k)
    assert_same_state(alone.read_bubbles(), batch.read_bubbles(3), "alone vs batch")


def test_substep_limit_is_reported_and_nothing_is_published(oracle):
    scene = scenes.config1()
    scene.config.substep_split = 2048          # h <= dt / 2048: more than 1024 substeps per step
    world = solver.NativeWorld(scene)
    before = world.read_bubbles()
    frame = world.step(0)
    _, _, cf, _ = oracle.run(scene, 1)
    assert frame.status == abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED == cf[0]["status"]
    assert frame.substep_count == 1024
    # The detail names the object and the term of the last substep (NativeMetalSolver.swift:1027-1030).
    assert frame.detail is not None
    assert frame.detail.limiting_term == cf[0]["limiter_term"]
    assert frame.detail.object_ids[0] == cf[0]["limiter_object"]
    assert_same_state(world.read_bubbles(), before, "unpublished")


def test_non_finite_state_is_reported(oracle):
    scene = scenes.config1()
    scene.bubbles["velocity"][5] = (np.nan, 0.0)
    world = solver.NativeWorld(scene)
    frame = world.step(0)
    _, _, cf, _ = oracle.run(scene, 1)
    assert frame.status == cf[0]["status"]
    assert frame.status in (abi.BP_ERR_NON_FINITE, abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED)


def test_injected_centre_is_a_contact_violation(oracle):
    # ContactAssertionTests.swift:35-83: a centre moved onto another bubble after a substep.
    scene = scenes.config1()
    scene.config.assert_every_substep = 1
    scene.config.injection_enabled = 1
    scene.config.injection_after_substep = 0
    scene.config.injection_object = int(scene.bubbles["id"][10])
    scene.config.injection_position[:] = [float(v) for v in scene.bubbles["center"][20]]
    world = solver.NativeWorld(scene)
    frame = world.step(0)
    _, _, cf, _ = oracle.run(scene, 1)
    c = cf[0]
    assert frame.status == c["status"]
    if frame.status == abi.BP_ERR_CONTACT_VIOLATION:
        assert frame.detail.violation_kind == c["violation_kind"] == abi.BP_VIOLATION_PAIR
        assert frame.detail.object_ids == (int(c["violation_objects"][0]), int(c["violation_objects"][1]))


@pytest.mark.parametrize("cls", [None, 10])
def test_centre_injected_onto_a_static_bubble_is_a_pair_violation(oracle, cls):
    # ContactAssertionTests.swift:35-83 with the target held still, so the centres coincide
    # exactly after the injection: the pair test d <= 0 (Force:880-881) must fire, in the
    # LDS-resident kernel and in the global-state kernel alike.
    scene = scenes.config1()
    target = 20
    scene.bubbles["inverse_mass"][target] = 0.0
    scene.bubbles["flags"][target] |= abi.BP_BUBBLE_STATIC
    scene.config.assert_every_substep = 1
    scene.config.injection_enabled = 1
    scene.config.injection_after_substep = 0
    scene.config.injection_object = int(scene.bubbles["id"][10])
    scene.config.injection_position[:] = [float(v) for v in scene.bubbles["center"][target]]
    _, _, cf, _ = oracle.run(scene, 1)
    c = cf[0]
    assert c["status"] == abi.BP_ERR_CONTACT_VIOLATION and c["violation_kind"] == abi.BP_VIOLATION_PAIR
    batch = solver.Batch([scene])
    if cls is not None:
        batch.set_world_class(0, cls)
    info = batch.step(1)[0][0]
    assert info["status"] == c["status"]
    assert info["detail"]["violation_kind"] == c["violation_kind"]
    assert (int(info["detail"]["object_ids"][0]), int(info["detail"]["object_ids"][1])) == (
        int(c["violation_objects"][0]), int(c["violation_objects"][1]))
    if cls is not None:
        assert info["kernel_class"] == cls


def test_command_failure_switch_leaves_the_world_usable(oracle):
    scene = scenes.config1()
    world = solver.NativeWorld(scene)
    world.step(0)
    before = world.read_bubbles()
    world.fail_next_command()
    frame = world.step(1)
    assert frame.status == abi.BP_ERR_COMMAND_FAILED
    assert_same_state(world.read_bubbles(), before, "after command failure")
    assert world.step(2).status == abi.BP_OK


def test_step_history_sums_to_the_frame_step(oracle):
    # NativeMetalSubstepKernelTests.swift:918-948: captured h sums to dt within 1e-6.
    scene = scenes.config1()
    world = solver.NativeWorld(scene)
    for k in range(20):
        f = world.step(k)
        h = world.read_step_history()
        assert len(h) == f.substep_count
        assert abs(float(np.sum(h.astype(np.float64))) - 1 / 60) <= 1e-6


def test_pair_only_world_conserves_momentum_bitwise_in_forces(oracle):
    # Exactly opposite pair forces: sum of m v of a free two-bubble world is unchanged.
    scene = scenes.pair_scene(40.0, velocity=(30.0, 0.0))
    world = solver.NativeWorld(scene)
    for k in range(30):
        f = world.step(k)
        assert f.status == 0
    b = f.bubbles
    p = (b["mass"][:, None].astype(np.float64) * b["velocity"].astype(np.float64)).sum(axis=0)
    p0 = (scene.bubbles["mass"][:, None].astype(np.float64) * scene.bubbles["velocity"].astype(np.float64)).sum(axis=0)
    assert np.allclose(p, p0, rtol=1e-5, atol=1e-4)


def test_every_kernel_configuration_gives_the_same_bits(oracle, hip_library):
    # Autotuner safety: the LDS configurations and the global-state kernel are interchangeable.
    rng = np.random.Generator(np.random.PCG64(5))
    scene = scenes.batch_world(rng, 60, bodies_max=6)
    classes = int(hip_library.bp_batch_class_count())
    batch = solver.Batch([scene] * classes)
    for k in range(classes):
        batch.set_world_class(k, k)
    steps = 5
    infos = batch.step(steps)
    cb, _, cf, _ = oracle.run(scene, steps)
    assert sorted(set(int(x) for x in infos[0]["kernel_class"])) == list(range(classes))
    for k in range(classes):
        assert np.array_equal(infos[:, k]["substep_count"], cf["substeps"]), k
        assert np.array_equal(infos[:, k]["candidate_pair_count"], cf["candidate_pairs"]), k
        assert_same_state(batch.read_bubbles(k), cb, f"class {k}")


@pytest.mark.parametrize("name", kernel_scenes.PARITY_SCENES)
def test_reference_kernel_scenes_match_oracle_bitwise(oracle, name):
    # The scenes of the reference's substep-kernel tests, whose closed forms pin the oracle
    # (test_oracle_world.py), stepped on the GPU: same h history, limiter and state bits.
    scene, steps = kernel_scenes.parity_scene(name)
    world = solver.NativeWorld(scene)
    cur = scene.copy()
    for k in range(steps):
        frame = world.step(k)
        cb, _, cf, ex = oracle.run(cur, 1, history=True)
        c = cf[0]
        assert (frame.status, frame.substep_count, frame.grid_rebuild_count) == (c["status"], c["substeps"], c["rebuilds"]), k
        if frame.status != abi.BP_OK:
            break
        assert frame.limiter_term == c["limiter_term"] and frame.limiter_object == c["limiter_object"], k
        h = world.read_step_history()
        assert np.array_equal(bits(h), bits(ex["history"][: c["substeps"]])), k
        assert_same_state(frame.bubbles, cb, f"{name} step {k}")
        cur.bubbles = cb


def test_config3_large_world_matches_the_oracle_fixture():
    # The 262,144-bubble world of config 3 (cooperative grid, body LBVH): counters and the SHA-256
    # of the state bits after each step equal the oracle's (tests/golden/config3_oracle.json, written
    # by tests/golden/make_config3_fixture.py; the oracle takes about a minute per step here).
    import hashlib
    import json
    from pathlib import Path
    golden = json.loads((Path(__file__).resolve().parent / "golden" / "config3_oracle.json").read_text())
    scene = scenes.config3()
    world = solver.NativeWorld(scene)
    for k, want in enumerate(golden["steps"]):
        f = world.step(k)
        got = (f.status, f.substep_count, f.grid_rebuild_count, f.candidate_pair_count, f.maximum_contact_candidates_per_object)
        assert got == (want["status"], want["substeps"], want["rebuilds"], want["candidate_pairs"], want["max_candidates"]), k
        h = hashlib.sha256()
        for field in STATE_FIELDS:
            h.update(np.ascontiguousarray(f.bubbles[field]).tobytes())
        assert h.hexdigest() == want["state_sha256"], k
        assert f.execution_path == abi.BP_PATH_COOPERATIVE


def test_undersized_bounding_radius_is_refused():
    # The kernels skip a body's exact inside test for centres outside its bounding circle, so a
    # body whose bounding radius does not hold its vertices (and a round body's core radius) is
    # refused as an invalid snapshot instead of missing violations.
    import reference_scenes as rs
    for shrink in ("square", "capsule"):
        w = rs.World().bubble(1, (0.0, 200.0))
        w = w.square(2, (0.0, 0.0), 100.0, role="kinematic") if shrink == "square" else w.capsule(2, (0.0, 0.0), 40.0, 10.0)
        scene = w.build()
        scene.bodies[0]["bounding_radius"] = np.float32(0.9 * float(scene.bodies[0]["bounding_radius"]))
        frame = solver.NativeWorld(scene).step(0)
        assert frame.status == abi.BP_ERR_INVALID_SNAPSHOT, shrink
        scene.bodies[0]["bounding_radius"] = np.float32(float(scene.bodies[0]["bounding_radius"]) / 0.9 * 1.0001)
        assert solver.NativeWorld(scene).step(0).status == abi.BP_OK, shrink
