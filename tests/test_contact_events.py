"""Contact began/ended transitions (bp_contact_transitions), the host logic of
NativePublicWorld.swift:340-391 over a step's contact records.

CPU only: the library's host entry point on hand-made records, against the rules of the
reference written out per case (weighted mean point, normal from the key's first object, impulse
F dt, interest, owners of segments, key order), and on the records of consecutive steps of a
stepped world (GPU)."""
import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, solver

DT = np.float32(1.0 / 60.0)


def record(first, second, kind=abi.BP_CONTACT_PAIR, normal=(1.0, 0.0), point=(0.0, 0.0), penetration=1.0, impulse=0.0):
    r = np.zeros(1, abi.CONTACT_DTYPE)[0]
    r["first"], r["second"], r["kind"] = first, second, kind
    r["normal"], r["point"], r["penetration"], r["impulse"] = normal, point, penetration, impulse
    return r


def segments(*owners):
    """Segments with ids 1.. and the given owners (None: no owner)."""
    s = np.zeros(len(owners), abi.SEGMENT_DTYPE)
    for k, o in enumerate(owners):
        s[k]["id"] = k + 1
        if o is not None:
            s[k]["owner_id"], s[k]["flags"] = o, abi.BP_SEGMENT_HAS_OWNER
    return s


def run(contacts, segs, interested, history=None):
    return solver.contact_transitions(np.array(contacts, abi.CONTACT_DTYPE), segs, interested, float(DT), history)


def test_began_event_weights_point_and_sums_normal_and_impulse():
    # Bubble 5 touches two edges (segments 1 and 2) of the polygon owned by 900.
    c = [record(5, 1, abi.BP_CONTACT_SEGMENT, normal=(0.0, 1.0), point=(10.0, 0.0), penetration=3.0, impulse=0.02),
         record(5, 2, abi.BP_CONTACT_SEGMENT, normal=(1.0, 0.0), point=(0.0, 10.0), penetration=1.0, impulse=0.01)]
    events, hist = run(c, segments(900, 900), [5])
    assert len(events) == 1 and hist.tolist() == [[5, 900]]
    e = events[0]
    assert e["kind"] == abi.BP_EVENT_CONTACT_BEGAN and (e["first"], e["second"]) == (5, 900)
    w1, w2 = np.float32(3.0) + np.float32(1e-4), np.float32(1.0) + np.float32(1e-4)
    px = (np.float32(10.0) * w1 + np.float32(0.0) * w2) / (w1 + w2)
    py = (np.float32(0.0) * w1 + np.float32(10.0) * w2) / (w1 + w2)
    assert np.allclose(e["point"], [px, py], rtol=1e-6)
    # Normals point from the segment to the bubble; the key runs from 5 to 900: flipped once.
    assert np.allclose(e["normal"], [-np.sqrt(0.5), -np.sqrt(0.5)], rtol=1e-6)
    assert np.isclose(e["impulse"], (0.02 + 0.01) / (1.0 / 60.0), rtol=1e-5)


def test_pair_normal_follows_the_key_order():
    # Bubble 9 (A) above bubble 4: the key is (4, 9), so the normal from 4 to 9 is the record's flipped.
    events, _ = run([record(9, 4, normal=(0.0, -1.0))], segments(), [4])
    assert (events[0]["first"], events[0]["second"]) == (4, 9)
    assert tuple(events[0]["normal"]) == (0.0, 1.0)
    events, _ = run([record(4, 9, normal=(0.0, 1.0))], segments(), [9])
    assert tuple(events[0]["normal"]) == (0.0, 1.0)


def test_interest_near_misses_and_ownerless_segments():
    c = [record(1, 2), record(3, 4, penetration=-0.001), record(3, 4, penetration=-0.001, impulse=0.5),
         record(6, 1, abi.BP_CONTACT_SEGMENT), record(7, 2, abi.BP_CONTACT_SEGMENT)]
    events, hist = run(c, segments(None, 800), [1, 3, 7])
    keys = [(int(e["first"]), int(e["second"])) for e in events]
    # (1, 2): interest in 1; (3, 4): only the record with an impulse counts; (6, seg 1): no owner;
    # (7, seg 2): owner 800.
    assert keys == [(1, 2), (3, 4), (7, 800)] and hist.tolist() == [[1, 2], [3, 4], [7, 800]]
    assert run(c, segments(None, 800), [])[0].size == 0


def test_ended_events_follow_began_events_in_key_order():
    prev = np.array([[1, 2], [2, 8], [3, 4]], np.uint64)
    events, hist = run([record(3, 4), record(5, 6)], segments(), [1, 2, 3, 5], prev)
    got = [(int(e["kind"]), int(e["first"]), int(e["second"])) for e in events]
    assert got == [(abi.BP_EVENT_CONTACT_BEGAN, 5, 6), (abi.BP_EVENT_CONTACT_ENDED, 1, 2), (abi.BP_EVENT_CONTACT_ENDED, 2, 8)]
    assert hist.tolist() == [[3, 4], [5, 6]]
    # No interest, but a previous set: every previous pair ends.
    events, hist = run([], segments(), [], prev)
    assert [int(e["kind"]) for e in events] == [abi.BP_EVENT_CONTACT_ENDED] * 3 and len(hist) == 0


@pytest.mark.gpu
def test_transitions_of_consecutive_steps():
    scene = scenes.config1()
    world = solver.NativeWorld(scene)
    ids = scene.bubbles["id"].astype(np.uint64)
    history = None
    for k in range(6):
        assert world.step(k).status == abi.BP_OK
        contacts = world.contacts()
        events, new = solver.contact_transitions(contacts, scene.segments, ids, float(scene.config.time_step), history)
        began = {(int(e["first"]), int(e["second"])) for e in events if e["kind"] == abi.BP_EVENT_CONTACT_BEGAN}
        ended = {(int(e["first"]), int(e["second"])) for e in events if e["kind"] == abi.BP_EVENT_CONTACT_ENDED}
        old = set() if history is None else {tuple(map(int, p)) for p in history}
        cur = {tuple(map(int, p)) for p in new}
        assert began == cur - old and ended == old - cur
        pairs = {(int(min(c["first"], c["second"])), int(max(c["first"], c["second"]))) for c in contacts
                 if c["kind"] == abi.BP_CONTACT_PAIR and (c["penetration"] > 0 or c["impulse"] > 0)}
        assert pairs <= cur
        history = new
    assert len(history) > 0
