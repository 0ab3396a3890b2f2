"""End-of-step contact records (bp_world_contacts, bp_batch_contacts; the oracle's bpo_contacts).

CPU: the oracle's records against closed forms of nativeBuildContacts
(NativeGeometryKernels.metal-source:683-728): pair normal, point and penetration; the closest point
of a plain segment; the edge-interior and convex-start-vertex features of a static polygon edge and
the record an edge drops (:674-681); the separation threshold of the compaction
(NativeContactKernels.metal-source:146-160); the impulse F dt^2 against the Double contact law of
ContactLawReference.swift (tests/contact_law_reference.py); the candidate order (pairs of every
bubble, then segments, Geometry:625-651); the FNV-1a contact id (NativeMetalSolver.swift:1047-1052);
and, on a stepped world, the pair records against brute force over all pairs.
GPU: the device records, field for field and bit for bit, against the oracle's on the oracle's own
end state and rows (the GPU end state equals the oracle's, tested elsewhere)."""
import numpy as np
import pytest

import contact_law_reference as ref
import kernel_scenes as ks
from bubblephysics_amd import abi, scenes, solver

DT = np.float32(1.0 / 60.0)


def all_rows(scene):
    """Rows listing every other bubble (ascending) and every segment: candidates without a broadphase."""
    n, ns = len(scene.bubbles), len(scene.segments)
    offsets, entries = [0], []
    for i in range(n):
        entries += [j for j in range(n) if j != i] + [(1 << 30) | j for j in range(ns)]
        offsets.append(len(entries))
    return np.array(offsets, np.uint32), np.array(entries, np.uint32)


def records(oracle, scene):
    offsets, entries = all_rows(scene)
    out, total = oracle.contacts(scene, scene.bubbles, offsets, entries)
    assert total == len(out)
    return out


def test_pair_record_closed_form(oracle):
    for d, ra, rb in ((20.0, 12.0, 12.0), (15.0, 9.0, 14.0), (3.0, 12.0, 12.0)):
        w = ks.world([ks.body(1, (0.0, 0.0), radius=ra), ks.body(2, (d, 0.0), radius=rb)])
        (c,) = records(oracle, w)
        assert c["kind"] == abi.BP_CONTACT_PAIR and c["first"] == 1 and c["second"] == 2 and c["index_b"] == 1
        assert tuple(c["normal"]) == (1.0, 0.0)
        assert tuple(c["point"]) == (np.float32(ra), 0.0)
        assert c["penetration"] == np.float32(ra + rb - d)
        eps = (ra + rb - d) / (ra + rb)
        assert np.isclose(c["impulse"], ref.force(eps, 229.0) * (1.0 / 60.0) ** 2, rtol=1e-5)
        assert c["allowed_side"] == 0.0 and c["feature"] == 0
    # The pair is listed once, from its lower index, whatever the direction of the normal.
    w = ks.world([ks.body(5, (10.0, 30.0)), ks.body(6, (10.0, 10.0))])
    (c,) = records(oracle, w)
    assert c["first"] == 5 and tuple(c["normal"]) == (0.0, -1.0) and tuple(c["point"]) == (10.0, 18.0)


def test_pairs_of_different_materials_use_the_series_law(oracle):
    w = ks.world([ks.body(1, (0.0, 0.0), radius=12.0, stiffness=229.0), ks.body(2, (18.0, 0.0), radius=10.0, stiffness=900.0)])
    (c,) = records(oracle, w)
    pc = ref.pair_contact(4.0, 12.0, 229.0, 10.0, 900.0)
    assert np.isclose(c["impulse"], pc.force * (1.0 / 60.0) ** 2, rtol=1e-3)


def test_separation_threshold_keeps_near_misses_only(oracle):
    # separation tolerance 0.002: a gap of 0.001 is kept with a negative penetration, 0.003 is not.
    near = records(oracle, ks.world([ks.body(1, (0.0, 0.0)), ks.body(2, (24.001, 0.0))]))
    far = records(oracle, ks.world([ks.body(1, (0.0, 0.0)), ks.body(2, (24.003, 0.0))]))
    assert len(near) == 1 and -0.002 < near[0]["penetration"] < 0 and near[0]["impulse"] == 0.0
    assert len(far) == 0


def test_polygon_edge_features(oracle):
    # The Rig's square (0,0)-(200,200): edges 1..4 from (0,0) counter-clockwise, outward normals out.
    w = ks.world([ks.body(10, (100.0, -5.0))], walls=True)
    (c,) = records(oracle, w)
    assert c["kind"] == abi.BP_CONTACT_SEGMENT and c["second"] == 1 and c["feature"] == 1
    assert tuple(c["normal"]) == (0.0, -1.0) and tuple(c["point"]) == (100.0, 0.0) and c["penetration"] == 7.0
    assert np.isclose(c["impulse"], ref.force(7.0 / 12.0, 229.0) * (1.0 / 60.0) ** 2, rtol=1e-5)
    # Outside the corner (0,0): the convex start vertex of edge 1 (feature 2); edge 4 ends there and drops it.
    w = ks.world([ks.body(10, (-5.0, -5.0))], walls=True)
    (c,) = records(oracle, w)
    assert c["second"] == 1 and c["feature"] == 2 and tuple(c["point"]) == (0.0, 0.0)
    dist = np.float32(np.sqrt(np.float32(50.0)))
    assert c["penetration"] == np.float32(12.0) - dist
    assert np.allclose(c["normal"], [-np.sqrt(0.5), -np.sqrt(0.5)], atol=1e-7)
    # Inside the square, past the edge's outward side: no record (feature 0 drops it).
    assert len(records(oracle, ks.world([ks.body(10, (100.0, 5.0))], walls=True))) == 0


def test_plain_segment_closest_point(oracle):
    seg = np.zeros(1, abi.SEGMENT_DTYPE)
    seg[0]["id"], seg[0]["a"], seg[0]["b"] = 7, (0.0, 0.0), (100.0, 0.0)
    seg[0]["category"], seg[0]["mask"] = 1, 0xFFFFFFFF
    for x, y, point in ((50.0, 8.0, (50.0, 0.0)), (-6.0, 8.0, (0.0, 0.0)), (103.0, -4.0, (100.0, 0.0))):
        w = ks.world([ks.body(3, (x, y))])
        w.segments = seg
        (c,) = records(oracle, w)
        dx, dy = np.float32(x) - np.float32(point[0]), np.float32(y) - np.float32(point[1])
        dist = np.float32(np.sqrt(np.float32(dx * dx + dy * dy)))
        assert tuple(c["point"]) == point and c["feature"] == 0 and c["second"] == 7
        assert c["penetration"] == np.float32(12.0) - dist
        assert np.allclose(c["normal"], [dx / dist, dy / dist], rtol=1e-6)
        assert np.isclose(c["impulse"], ref.force(float(c["penetration"]) / 12.0, 229.0) * (1.0 / 60.0) ** 2, rtol=1e-5)
    # An allowed side fixes the normal to the edge's left normal times the side.
    seg[0]["flags"] = abi.BP_SEGMENT_ALLOWED_SIDE
    seg[0]["allowed_side"] = -1.0
    w = ks.world([ks.body(3, (50.0, 8.0))])
    w.segments = seg
    (c,) = records(oracle, w)
    assert tuple(c["normal"]) == (0.0, -1.0) and c["allowed_side"] == -1.0


def test_candidate_order_and_contact_ids(oracle):
    # Bubbles 0-1 and 1-2 touch, bubble 0 and 2 touch the bottom edge: pairs of every bubble first.
    w = ks.world([ks.body(1, (50.0, -5.0)), ks.body(2, (68.0, -18.0)), ks.body(3, (86.0, -5.0))], walls=True)
    out = records(oracle, w)
    assert [(int(c["kind"]), int(c["first"]), int(c["second"])) for c in out] == [(0, 1, 2), (0, 2, 3), (1, 1, 1), (1, 3, 1)]
    mask = (1 << 64) - 1
    for c in out:
        h = 14695981039346656037
        for v in (int(c["kind"]), int(c["first"]) & mask, int(c["second"]) & mask):
            h = ((h ^ v) * 1099511628211) & mask
        assert int(c["id"]) == h
    assert len(set(out["id"].tolist())) == len(out)


def test_capacity_counts_every_record(oracle):
    w = ks.world([ks.body(1, (50.0, -5.0)), ks.body(2, (68.0, -18.0)), ks.body(3, (86.0, -5.0))], walls=True)
    offsets, entries = all_rows(w)
    full, total = oracle.contacts(w, w.bubbles, offsets, entries)
    cut, total2 = oracle.contacts(w, w.bubbles, offsets, entries, capacity=2)
    assert total == total2 == 4 and len(cut) == 2 and cut.tobytes() == full[:2].tobytes()


def test_pair_records_of_a_stepped_world_match_brute_force(oracle):
    scene = scenes.config1()
    cb, _, frames, extras = oracle.run(scene, 6, adjacency=True)
    assert frames[-1]["status"] == abi.BP_OK
    out, total = oracle.contacts(scene, cb, extras["row_offsets"], extras["entries"])
    got = {(int(c["first"]), int(c["second"])) for c in out if c["kind"] == abi.BP_CONTACT_PAIR}
    pos = cb["center"].astype(np.float32)
    want = set()
    tol = np.float32(scene.config.separation_tolerance)
    for i in range(len(cb)):
        d = pos[i + 1:] - pos[i]
        dist = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)).astype(np.float32)
        pen = (cb["radius"][i] + cb["radius"][i + 1:]).astype(np.float32) - dist
        for k in np.nonzero(pen > -tol)[0]:
            want.add((int(cb["id"][i]), int(cb["id"][i + 1 + k])))
    assert got == want and len(got) > 100


def _oracle_end(oracle, scene, steps):
    cb, _, frames, extras = oracle.run(scene, steps, adjacency=True)
    assert frames[-1]["status"] == abi.BP_OK
    want, total = oracle.contacts(scene, cb, extras["row_offsets"], extras["entries"])
    return cb, want, total


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("config1", 8), ("config2", 4)])
def test_world_contacts_match_the_oracle(oracle, name, steps):
    scene = getattr(scenes, name)()
    world = solver.NativeWorld(scene)
    for k in range(steps):
        assert world.step(k).status == abi.BP_OK
    cb, want, total = _oracle_end(oracle, scene, steps)
    assert world.read_bubbles().tobytes() == cb.tobytes()
    got = world.contacts()
    assert len(got) == total > 0
    assert got.tobytes() == want.tobytes()
    assert (got["kind"] == abi.BP_CONTACT_SEGMENT).any() and (got["feature"] > 0).any()
    cut = world.contacts(capacity=5)
    assert cut.tobytes() == want[:5].tobytes()


@pytest.mark.gpu
def test_batch_contacts_match_the_oracle(oracle):
    rng = np.random.Generator(np.random.PCG64(43))
    ws = [scenes.batch_world(rng, n, bodies_max=8) for n in (30, 400, 2500)]
    for w in ws:
        w.config.export_adjacency = 1
    quiet = ws[0].copy()
    quiet.config.export_adjacency = 0
    batch = solver.Batch(ws + [quiet])
    infos = batch.step(3)
    for k, w in enumerate(ws):
        if infos[-1][k]["status"] != abi.BP_OK:
            continue
        cb, want, total = _oracle_end(oracle, w, 3)
        got = batch.contacts(k)
        assert len(got) == total, k
        assert got.tobytes() == want.tobytes(), k
    with pytest.raises(abi.BubblePhysicsError):
        batch.contacts(len(ws))
