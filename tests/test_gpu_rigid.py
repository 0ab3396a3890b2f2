"""GPU parity of the rigid solver for worlds without bubbles (bp_rigid_batch_*, bp_rigid.hip).

Every case of tests/rigid_scenes.py (the reference's rigid, joint, speed-limit and polygon scenes,
restated; the oracle is pinned on them in test_oracle_rigid.py) runs on the device through the
C ABI and must equal the oracle bit for bit: the published body records, the joints' lambdas, the
frame records and the last step's contact records. Worlds run one per launch and all together in
one launch; a failing world stops without touching the others; seeded crowds with joints check
the kernel at a workload beyond the reference's scenes, with full and cut candidate rows.
"""
import numpy as np
import pytest

from bubblephysics_amd import abi, rigid
from rigid_scenes import Case, cfg, mixed_world, reference_cases

pytestmark = pytest.mark.gpu

FRAME_FIELDS = ("status", "guard_cause", "candidate_pairs", "candidate_overflow", "max_candidates_per_object", "removed",
                "wrapped")


def oracle_run(oracle, case: Case, steps=None):
    return oracle.run_rigid(case.scene(), case.steps if steps is None else steps)


def assert_world_equal(batch, world, case, expect, last_info):
    bodies, joints, infos, contacts = expect
    got = batch.read_bodies(world)
    assert got.tobytes() == bodies.tobytes(), f"{case.name}: bodies differ"
    if len(joints):
        assert batch.read_joints(world).tobytes() == joints.tobytes(), f"{case.name}: joint lambdas differ"
    for f in FRAME_FIELDS:
        assert int(last_info[f]) == int(infos[-1][f]), (case.name, f, int(last_info[f]), int(infos[-1][f]))
    if infos[-1]["status"] == abi.BP_OK:
        rec = batch.contacts(world)
        assert rec.tobytes() == contacts.tobytes(), f"{case.name}: contact records differ"


@pytest.mark.parametrize("case", reference_cases(), ids=lambda c: c.name)
def test_reference_scene_matches_the_oracle_bitwise(oracle, case):
    expect = oracle_run(oracle, case)
    print(f"rigid case {case.name}: {len(case.objects)} bodies, {case.steps} steps", flush=True)
    batch = rigid.RigidBatch([case.scene()])
    infos = batch.step(case.steps)
    assert [int(s) for s in infos["status"][:len(expect[2]), 0]] == [int(s) for s in expect[2]["status"]]
    assert_world_equal(batch, 0, case, expect, infos[len(expect[2]) - 1, 0])


def test_every_scene_in_one_launch_matches_the_oracle(oracle):
    cases = [c for c in reference_cases() if c.steps >= 60]
    steps = 60
    batch = rigid.RigidBatch([c.scene() for c in cases])
    infos = batch.step(steps)
    for w, case in enumerate(cases):
        expect = oracle_run(oracle, case, steps)
        assert_world_equal(batch, w, case, expect, infos[len(expect[2]) - 1, w])


def test_steps_one_launch_at_a_time_equal_one_launch(oracle):
    case = reference_cases()[0]
    a = rigid.RigidBatch([case.scene()])
    for _ in range(20):
        a.step(1)
    b = rigid.RigidBatch([case.scene()])
    b.step(20)
    assert a.read_bodies(0).tobytes() == b.read_bodies(0).tobytes()
    assert a.contacts(0).tobytes() == b.contacts(0).tobytes()


@pytest.mark.parametrize("seed,limit", [(1, 128), (2, 128), (3, 3)])
def test_seeded_crowds_with_joints_match_the_oracle(oracle, seed, limit):
    case = mixed_world(seed)
    case.config.candidate_limit = limit
    expect = oracle_run(oracle, case)
    print(f"rigid crowd seed {seed}, k {limit}: {len(case.objects)} bodies, {case.steps} steps", flush=True)
    batch = rigid.RigidBatch([case.scene()])
    infos = batch.step(case.steps)
    assert int(infos[-1, 0]["status"]) == int(expect[2][-1]["status"]) == abi.BP_OK
    if limit < 128:
        assert int(expect[2]["candidate_overflow"].max()) > 0, "the cut rows must drop candidates"
    assert_world_equal(batch, 0, case, expect, infos[-1, 0])


def test_many_crowds_in_one_launch(oracle):
    cases = [mixed_world(100 + s, bodies=24 + (s % 5) * 8, joints=(s % 2 == 0)) for s in range(24)]
    batch = rigid.RigidBatch([c.scene() for c in cases])
    infos = batch.step(20)
    for w, case in enumerate(cases):
        expect = oracle_run(oracle, case, 20)
        assert_world_equal(batch, w, case, expect, infos[-1, w])


def spinning_world():
    o = rigid.RigidObject(id=1, shape="polygon", vertices=[(-5, -5), (5, -5), (5, 5), (-5, 5)], angular_velocity=500.0,
                          role="dynamic", mass=1.0, moment=10.0)
    return Case("spinning", [o], cfg((-1000, -1000, 1000, 1000)), 3)


def test_a_failing_world_stops_alone_and_steps_again_after_new_input(oracle):
    good = reference_cases()[1]
    bad = spinning_world()
    batch = rigid.RigidBatch([good.scene(), bad.scene()])
    infos = batch.step(3)
    assert list(infos[:, 1]["status"]) == [abi.BP_ERR_GUARD_FAILURE, abi.BP_ERR_UNAVAILABLE, abi.BP_ERR_UNAVAILABLE]
    assert int(infos[0, 1]["guard_cause"]) == abi.BP_GUARD_RIGID_ROTATION_BUDGET
    assert batch.read_bodies(1).tobytes() == bad.scene().bodies.tobytes(), "a failed frame publishes nothing"
    with pytest.raises(abi.BubblePhysicsError):
        batch.contacts(1)
    expect = oracle_run(oracle, good, 3)
    assert_world_equal(batch, 0, good, expect, infos[-1, 0])
    calm = bad.scene()
    calm.bodies["motion"][0][2] = 1.0
    batch.set_world(1, calm)
    infos = batch.step(2)
    assert list(infos[:, 1]["status"]) == [abi.BP_OK, abi.BP_OK]
    expect = oracle.run_rigid(calm, 2)
    assert batch.read_bodies(1).tobytes() == expect[0].tobytes()
    # The other world went on from where it was.
    expect = oracle_run(oracle, good, 5)
    assert batch.read_bodies(0).tobytes() == expect[0].tobytes()


def test_piled_crowds_match_the_oracle_over_many_launches(oracle):
    # The bench workload after the crowds have fallen and piled up: resting contacts, walls pushing
    # back, pairs colliding inside a step (impact-time bisection), joints under load.
    cases = [mixed_world(200 + s) for s in range(3)]
    batch = rigid.RigidBatch([c.scene() for c in cases])
    for _ in range(4):
        infos = batch.step(40)
        assert (infos["status"] == abi.BP_OK).all()
    for w, case in enumerate(cases):
        expect = oracle_run(oracle, case, 160)
        print(f"piled crowd {w}: {len(expect[3])} contact records", flush=True)
        assert len(expect[3]) > 20, "the crowd must have piled up"
        assert_world_equal(batch, w, case, expect, infos[-1, w])


def test_a_world_beyond_the_lds_class_matches_the_oracle(oracle):
    # More bodies than the one-wavefront class holds: the 256-lane kernel with bodies in global memory.
    case = mixed_world(300, bodies=120)
    assert len(case.objects) > 96
    expect = oracle_run(oracle, case, 60)
    batch = rigid.RigidBatch([case.scene()])
    infos = batch.step(60)
    assert_world_equal(batch, 0, case, expect, infos[-1, 0])
