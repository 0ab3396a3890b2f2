"""End-of-step bubble-body contact records (bp_world_body_contacts, bp_batch_body_contacts; the
oracle's bpo_body_contacts) and the contact events they feed.

The reference writes one record per candidate of a body's bubble row (nativeBubbleBodyContacts,
NativeForceKernels.metal-source:333-358): the obstacle at the body's end-of-step pose against the
bubble's final centre, kept when distance < radius, with the normal from the bubble to the body, the
point on the body's surface and the wall law's F dt^2. The host reads them back as rigid records
(NativeRigidPolygonState.swift:390-402) and adds them to the contact-event sums with weight 1
(NativePublicWorld.swift:377-380).

CPU: the oracle's records against closed forms of the obstacle -- a circle, a capsule, a polygon
face, a polygon corner, a centre inside a polygon, a moved body taken at its end pose -- and the
impulse against the Double contact law (ContactLawReference.swift); strict distance < radius;
candidates only from the rows; body-major order; the reference's event scene
(RigidContactTests.swift:110-139) on the oracle, which a planted obstacle mistake fails.
GPU: the device records bit for bit against the oracle's on the oracle's own end state and rows,
and the event scene frame by frame."""
import math

import numpy as np
import pytest

import contact_law_reference as ref
import reference_scenes as rs
from bubblephysics_amd import abi, scenes, solver

DT = 1.0 / 60.0


def body_rows(n, pairs):
    """Rows holding the body entries `pairs` = [(bubble index, body index), ...] only."""
    offsets, entries = [0], []
    for i in range(n):
        entries += [(2 << 30) | k for (b, k) in sorted(pairs) if b == i]
        offsets.append(len(entries))
    return np.array(offsets, np.uint32), np.array(entries, np.uint32)


def records(oracle, scene, pairs=None):
    n, nb = len(scene.bubbles), len(scene.bodies)
    if pairs is None:
        pairs = [(i, k) for i in range(n) for k in range(nb)]
    offsets, entries = body_rows(n, pairs)
    out, total = oracle.body_contacts(scene, scene.bubbles, scene.bodies, offsets, entries)
    assert total == len(out)
    return out


def wall_impulse(penetration, radius, stiffness=229.0):
    return ref.force(penetration / radius, stiffness) * DT * DT


def test_circle_record_closed_form(oracle):
    w = rs.World().bubble(1, (-60.0, 0.0)).circle(2, (0.0, 0.0), 50.0).build()
    (c,) = records(oracle, w)
    assert c["kind"] == abi.BP_CONTACT_BODY and c["first"] == 1 and c["second"] == 2 and c["index_b"] == 0
    assert tuple(c["normal"]) == (1.0, 0.0)                   # from the bubble to the body
    assert tuple(c["point"]) == (-50.0, 0.0)                  # on the circle's surface
    assert c["penetration"] == 14.0                           # radius 24 - distance 10
    assert np.isclose(c["impulse"], wall_impulse(14.0, 24.0), rtol=1e-5)
    assert c["feature"] == 0 and c["allowed_side"] == 0.0
    # distance == radius is no contact (strict <), nor is a larger one.
    assert len(records(oracle, rs.World().bubble(1, (-74.0, 0.0)).circle(2, (0.0, 0.0), 50.0).build())) == 0
    assert len(records(oracle, rs.World().bubble(1, (0.0, 80.0)).circle(2, (0.0, 0.0), 50.0).build())) == 0


def test_capsule_record_closed_form(oracle):
    w = rs.World().bubble(1, (15.0, 15.0)).capsule(2, (0.0, 0.0), 40.0, 10.0).build()
    (c,) = records(oracle, w)
    assert tuple(c["normal"]) == (0.0, -1.0) and tuple(c["point"]) == (15.0, 10.0)
    assert c["penetration"] == 19.0
    # Beyond the end of the core segment the nearest core point is its end.
    w = rs.World().bubble(1, (52.0, 0.0)).capsule(2, (0.0, 0.0), 40.0, 10.0).build()
    (c,) = records(oracle, w)
    assert tuple(c["normal"]) == (-1.0, 0.0) and tuple(c["point"]) == (50.0, 0.0) and c["penetration"] == 22.0


def test_polygon_face_corner_and_inside(oracle):
    square = lambda at: rs.World().bubble(1, at).square(2, (0.0, 0.0), 100.0, role="kinematic").build()
    (c,) = records(oracle, square((0.0, 58.0)))                # above the top face
    assert tuple(c["normal"]) == (0.0, -1.0) and tuple(c["point"]) == (0.0, 50.0) and c["penetration"] == 16.0
    (c,) = records(oracle, square((56.0, 56.0)))               # outside the corner (50, 50)
    d = math.sqrt(72.0)
    assert tuple(c["point"]) == (50.0, 50.0)
    assert np.allclose(c["normal"], (-1 / math.sqrt(2), -1 / math.sqrt(2)), atol=1e-6)
    assert np.isclose(c["penetration"], 24.0 - d, atol=1e-5)
    (c,) = records(oracle, square((0.0, 45.0)))                # centre inside: distance -5 to the top face
    assert tuple(c["normal"]) == (0.0, -1.0) and tuple(c["point"]) == (0.0, 50.0) and c["penetration"] == 29.0
    assert np.isfinite(c["impulse"]) and c["impulse"] > wall_impulse(16.0, 24.0)   # past the law's range (ε > 1)


def test_moved_body_is_taken_at_its_end_pose(oracle):
    w = rs.World().bubble(1, (-60.0, 0.0)).circle(2, (0.0, 0.0), 50.0, role="kinematic").build()
    w.bodies[0]["previous"][:2] = (400.0, 0.0)                 # the step moved it from x = 400 to 0
    (c,) = records(oracle, w)
    assert tuple(c["point"]) == (-50.0, 0.0) and c["penetration"] == 14.0


def test_candidates_come_from_the_rows_and_records_go_body_by_body(oracle):
    w = (rs.World().bubble(1, (-60.0, 0.0)).bubble(2, (60.0, 0.0)).bubble(3, (0.0, 300.0))
         .circle(7, (0.0, 0.0), 50.0).circle(8, (0.0, 250.0), 40.0).build())
    out = records(oracle, w)
    assert [(int(c["second"]), int(c["first"])) for c in out] == [(7, 1), (7, 2), (8, 3)]
    # A touching pair missing from the rows gives no record.
    out = records(oracle, w, pairs=[(0, 0), (2, 1)])
    assert [(int(c["second"]), int(c["first"])) for c in out] == [(7, 1), (8, 3)]


def _events(runner_records, interested, frames):
    """Began/ended events per frame from each frame's records, via bp_contact_transitions."""
    history, out = None, []
    for k, recs in enumerate(runner_records):
        events, history = solver.contact_transitions(recs, np.zeros(0, abi.SEGMENT_DTYPE), interested, DT, history)
        out.append(events)
    return out


def _event_scene():
    """RigidContactTests.testBubbleBodyContactProducesBeganAndEndedEvents (:110-139): a bubble at
    (-120, 0) moving at 200 against a static circle of radius 50 at the origin."""
    return rs.World().bubble(1, (-120.0, 0.0), velocity=(200.0, 0.0)).circle(2, (0.0, 0.0), 50.0).build()


def _oracle_frames(oracle, scene, frames):
    runner = rs.OracleRunner(scene, oracle)
    recs = []
    for _ in range(frames):
        cb, cbod, cf, ex = oracle.run(runner.scene, 1, region=True, adjacency=True)
        assert cf[0]["status"] == abi.BP_OK
        body, _ = oracle.body_contacts(runner.scene, cb, cbod, ex["row_offsets"], ex["entries"])
        pair, _ = oracle.contacts(runner.scene, cb, ex["row_offsets"], ex["entries"])
        recs.append(np.concatenate([pair, body]))
        runner.step()
    return recs


def check_event_scene(events):
    began = [(k, e) for k, ev in enumerate(events) for e in ev if e["kind"] == abi.BP_EVENT_CONTACT_BEGAN
             and (e["first"], e["second"]) == (1, 2)]
    ended = [k for k, ev in enumerate(events) for e in ev if e["kind"] == abi.BP_EVENT_CONTACT_ENDED
             and (e["first"], e["second"]) == (1, 2)]
    assert len(began) == 1 and len(ended) == 1
    step, e = began[0]
    assert abs(e["normal"][0] - 1.0) <= 1e-3 and abs(e["normal"][1]) <= 1e-3
    assert abs(e["point"][0] + 50.0) <= 1.0
    assert e["impulse"] > 0.0
    assert ended[0] > step


def test_reference_event_scene_on_the_oracle(oracle):
    events = _events(_oracle_frames(oracle, _event_scene(), 120), [1], 120)
    check_event_scene(events)


def test_planted_round_core_mistake_fails_the_event_scene(oracle):
    # A circle's distance without its radius: the world fails or the event is wrong.
    oracle.set_plant(1)
    try:
        with pytest.raises(AssertionError):
            check_event_scene(_events(_oracle_frames(oracle, _event_scene(), 120), [1], 120))
    finally:
        oracle.set_plant(0)


def test_body_records_enter_the_sums_as_rigid_records():
    # Weight 1 and no penetration test; the key is ordered by id, the normal flipped with it.
    rec = np.zeros(2, abi.CONTACT_DTYPE)
    rec[0] = (0, 9, 4, abi.BP_CONTACT_BODY, 0, (0.6, 0.8), (3.0, 4.0), -1.0, 0.0, 0.5 * DT, 0)
    rec[1] = (0, 9, 5, abi.BP_CONTACT_BODY, 0, (1.0, 0.0), (1.0, 1.0), 2.0, 0.0, 0.0, 1)
    events, history = solver.contact_transitions(rec, np.zeros(0, abi.SEGMENT_DTYPE), [9], DT)
    assert [(int(e["first"]), int(e["second"])) for e in events] == [(4, 9), (5, 9)]
    assert np.allclose(events[0]["normal"], (-0.6, -0.8)) and tuple(events[0]["point"]) == (3.0, 4.0)
    assert np.isclose(events[0]["impulse"], 0.5) and events[1]["impulse"] == 0.0
    assert history.tolist() == [[4, 9], [5, 9]]


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [1, 6])
def test_world_body_contacts_match_the_oracle(oracle, steps):
    scene = scenes.config2()
    world = solver.NativeWorld(scene)
    for k in range(steps):
        assert world.step(k).status == abi.BP_OK
    cb, cbod, frames, ex = oracle.run(scene, steps, adjacency=True)
    assert frames[-1]["status"] == abi.BP_OK
    assert world.read_bubbles().tobytes() == cb.tobytes()
    want, total = oracle.body_contacts(scene, cb, cbod, ex["row_offsets"], ex["entries"])
    got = world.body_contacts()
    assert len(got) == total > 0
    assert got.tobytes() == want.tobytes()
    assert world.body_contacts(capacity=3).tobytes() == want[:3].tobytes()


@pytest.mark.gpu
def test_batch_body_contacts_match_the_oracle(oracle):
    rng = np.random.Generator(np.random.PCG64(47))
    ws = [scenes.batch_world(rng, n, bodies_max=12) for n in (40, 700, 3000)]
    for w in ws:
        w.config.export_adjacency = 1
    batch = solver.Batch(ws)
    infos = batch.step(3)
    checked = 0
    for k, w in enumerate(ws):
        if infos[-1][k]["status"] != abi.BP_OK or not len(w.bodies):
            continue
        cb, cbod, frames, ex = oracle.run(w, 3, adjacency=True)
        want, total = oracle.body_contacts(w, cb, cbod, ex["row_offsets"], ex["entries"])
        got = batch.body_contacts(k)
        assert len(got) == total, k
        assert got.tobytes() == want.tobytes(), k
        checked += 1
    assert checked >= 2


@pytest.mark.gpu
def test_reference_event_scene_on_the_gpu(oracle):
    scene = _event_scene()
    want = _oracle_frames(oracle, scene, 120)
    runner = rs.GpuRunner(scene)
    got = []
    for k in range(120):
        r = runner.step()
        assert r["status"] == abi.BP_OK
        got.append(np.concatenate([runner.world.contacts(), runner.world.body_contacts()]))
        assert got[-1].tobytes() == want[k].tobytes(), k
    check_event_scene(_events(got, [1], 120))
