"""Structural staging on the host (bp_stage_merge, bp_stage_split): TransactionStager.merge / .split
(TransactionStager.swift:630-735), StructuralOperations.swift:10-151 and the placement rules of
PlacementGeometry.swift:42-159.

CPU only. Pinned by closed forms and conservation laws rather than by the restated formulas: the
merge point of two touching discs is the middle of their overlap; the merged radius keeps the
area, the velocity the momentum, the angular velocity the angular momentum about the new centre;
half a disc has its centroid at 4r/(3 pi); any split keeps area, mass, centre of mass, momentum
and angular momentum, and a part that has to move keeps the momenta about the parts' new centre
of mass; the placement search takes its candidates in the reference's order (points towards each
source, the sources' centres, the rings) and refuses with the first cause when none passes."""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi, solver

F = np.float32


def bubble(id_, x, y, r=10.0, m=2.0, v=(0.0, 0.0), w=0.0, rot=0.0, cat=1, mask=0xFFFFFFFF):
    b = np.zeros(1, abi.BUBBLE_DTYPE)[0]
    b["id"], b["center"], b["previous_center"], b["velocity"] = id_, (x, y), (x, y), v
    b["radius"], b["mass"], b["inverse_mass"], b["angular_velocity"], b["rotation"] = r, m, 1.0 / m, w, rot
    b["contact_stiffness"], b["flags"], b["category"], b["mask"] = 229.0, abi.BP_BUBBLE_ENABLED, cat, mask
    return b


def world(*bs):
    return np.array(sorted(bs, key=lambda b: int(b["id"])), abi.BUBBLE_DTYPE)


def result(id_, cat=1, mask=0xFFFFFFFF):
    return bubble(id_, 0.0, 0.0, cat=cat, mask=mask)


def circle_body(x, y, r, cat=1, mask=0xFFFFFFFF):
    b = np.zeros(1, abi.BODY_DTYPE)
    b[0]["id"], b[0]["pose"], b[0]["core_radius"], b[0]["bounding_radius"] = 900, (x, y, 0.0), r, r
    b[0]["shape"], b[0]["category"], b[0]["mask"] = abi.BP_SHAPE_CIRCLE, cat, mask
    return b


def by_id(bs, id_):
    (k,) = np.nonzero(bs["id"] == id_)[0]
    return bs[k]


def test_merge_of_two_touching_discs():
    a = bubble(1, 0.0, 0.0, r=12.0, m=2.0, v=(0.0, 1.0))
    b = bubble(2, 18.0, 0.0, r=8.0, m=3.0, v=(1.0, -1.0))
    keep = bubble(5, 100.0, 100.0)
    out = solver.stage_merge(world(a, b, keep), [2, 1], result(7))
    assert out["id"].tolist() == [5, 7] and out[0].tobytes() == keep.tobytes()
    m = by_id(out, 7)
    # The overlap runs from x = 10 (b's surface) to 12 (a's): its middle is the merge point.
    assert tuple(m["center"]) == (11.0, 0.0) and tuple(m["previous_center"]) == (11.0, 0.0)
    assert m["radius"] == F(math.sqrt(12.0 ** 2 + 8.0 ** 2)) and m["mass"] == 5.0 and m["inverse_mass"] == F(1.0 / 5.0)
    assert np.allclose(m["velocity"], [(2 * 0 + 3 * 1) / 5, (2 * 1 - 3 * 1) / 5], rtol=1e-6)
    # Angular momentum about (11, 0): a at dx = -11 moving +y, b at dx = 7 moving (1, -1).
    L = 2 * (-11 * 1.0) + 3 * (7 * -1.0 - 0 * 1.0)
    assert np.isclose(m["angular_velocity"], L / (0.5 * 5 * (144 + 64)), rtol=1e-6)
    assert m["rotation"] == 0.0


def test_merge_result_mass_override_keeps_the_sources_kinematics():
    a, b = bubble(1, 0.0, 0.0, v=(0.0, 1.0)), bubble(2, 20.0, 0.0, v=(0.0, -1.0))
    m = solver.stage_merge(world(a, b), [1, 2], result(1), result_mass=10.0)[0]
    assert m["id"] == 1 and m["mass"] == 10.0 and m["inverse_mass"] == F(0.1)
    assert np.allclose(m["velocity"], [0.0, 0.0]) and np.isclose(m["angular_velocity"], -40.0 / (0.5 * 4 * 200), rtol=1e-6)


def test_merge_point_moves_off_another_centre_and_out_of_a_body():
    a, b = bubble(1, 0.0, 0.0), bubble(2, 20.0, 0.0)
    third = bubble(3, 10.0, 0.0)
    # The contact point (10, 0) is bubble 3's centre: the first candidate, a quarter of the way to
    # source 1, is taken.
    m = by_id(solver.stage_merge(world(a, b, third), [1, 2], result(8)), 8)
    assert tuple(m["center"]) == (7.5, 0.0)
    # A bubble whose filter never meets the result's does not count.
    other = bubble(3, 10.0, 0.0, cat=2, mask=2)
    assert tuple(by_id(solver.stage_merge(world(a, b, other), [1, 2], result(8)), 8)["center"]) == (10.0, 0.0)
    # A rigid disc of radius 3 at (10, 0) covers (10, 0) and (7.5, 0); (5, 0) is the next candidate.
    m = by_id(solver.stage_merge(world(a, b), [1, 2], result(8), bodies=circle_body(10.0, 0.0, 3.0)), 8)
    assert tuple(m["center"]) == (5.0, 0.0)
    # A body the result's filter does not meet is no obstacle.
    m = by_id(solver.stage_merge(world(a, b), [1, 2], result(8), bodies=circle_body(10.0, 0.0, 3.0, cat=4, mask=4)), 8)
    assert tuple(m["center"]) == (10.0, 0.0)


def test_polygon_body_counts_its_boundary_as_inside():
    a, b = bubble(1, 0.0, 0.0), bubble(2, 20.0, 0.0)
    body = np.zeros(1, abi.BODY_DTYPE)
    body[0]["pose"], body[0]["shape"], body[0]["vertex_count"] = (12.0, 0.0, 0.0), abi.BP_SHAPE_POLYGON, 4
    body[0]["category"], body[0]["mask"] = 1, 0xFFFFFFFF
    square = np.array([[-2.0, -2.0], [2.0, -2.0], [2.0, 2.0], [-2.0, 2.0]], np.float32)   # x in [10, 14]
    m = by_id(solver.stage_merge(world(a, b), [1, 2], result(8), bodies=body, vertices=square), 8)
    # (10, 0) lies on the square's left edge: moved to (7.5, 0).
    assert tuple(m["center"]) == (7.5, 0.0)


def test_merge_refusals():
    a, b, c = bubble(1, 0.0, 0.0), bubble(2, 20.0, 0.0), bubble(3, 50.0, 0.0)
    w = world(a, b, c)
    cases = [([1], 9, abi.BP_REJECT_INVALID_STRUCTURAL_OPERATION), ([1, 1], 9, abi.BP_REJECT_INVALID_STRUCTURAL_OPERATION),
             ([1, 4], 9, abi.BP_REJECT_UNKNOWN_OBJECT), ([1, 2], 3, abi.BP_REJECT_DUPLICATE_ID)]
    for sources, rid, reason in cases:
        with pytest.raises(solver.StageRejected) as e:
            solver.stage_merge(w, sources, result(rid))
        assert e.value.reason == reason, sources
    # A region above every candidate (the rings around source 1 reach 8 * sqrt(200) / 4 < 29):
    # refused with the contact point's cause.
    with pytest.raises(solver.StageRejected) as e:
        solver.stage_merge(w, [1, 2], result(9), region=(-100.0, 30.0, 100.0, 100.0))
    assert (e.value.reason, e.value.issue, e.value.object_id) == (abi.BP_REJECT_INVALID_PLACEMENT, abi.BP_PLACEMENT_OUTSIDE_REGION, 9)


def conserved(src, parts):
    """Area, mass, momentum and angular momentum (about the parts' centre of mass, where the
    source's spin goes) of a split; returns the parts' centre of mass relative to the source."""
    m = parts["mass"].astype(np.float64)
    r = parts["radius"].astype(np.float64)
    c = parts["center"].astype(np.float64)
    v = parts["velocity"].astype(np.float64)
    M, R = float(src["mass"]), float(src["radius"])
    assert np.isclose((r * r).sum(), R * R, rtol=1e-6) and np.isclose(m.sum(), M, rtol=1e-6)
    assert np.allclose((v * m[:, None]).sum(0), M * src["velocity"].astype(np.float64), rtol=1e-5, atol=1e-5)
    d = c - src["center"].astype(np.float64)
    com = (d * m[:, None]).sum(0) / M
    e = d - com
    L = (m * (e[:, 0] * v[:, 1] - e[:, 1] * v[:, 0])).sum() + (0.5 * m * r * r * parts["angular_velocity"]).sum()
    assert np.isclose(L, 0.5 * M * R * R * float(src["angular_velocity"]), rtol=1e-5)
    return com


def test_half_split_closed_form():
    src = bubble(4, 3.0, -2.0, r=10.0, m=4.0, v=(1.0, 2.0), w=0.5)
    out = solver.stage_split(world(src), 4, result(4), 0.5, result(6), 0.5)
    a, b = by_id(out, 4), by_id(out, 6)
    x = 4 * 10.0 / (3 * math.pi)                    # centroid of half a disc
    assert np.allclose(a["center"], [3.0 - x, -2.0], rtol=1e-6) and np.allclose(b["center"], [3.0 + x, -2.0], rtol=1e-6)
    assert a["radius"] == b["radius"] == F(10.0 * math.sqrt(0.5)) and a["mass"] == b["mass"] == 2.0
    inertia = 0.5 * (2 * 50 + 2 * 50) + 2 * 2 * x * x
    omega = 0.5 * 4 * 100 * 0.5 / inertia
    assert np.isclose(a["angular_velocity"], omega, rtol=1e-6) and a["angular_velocity"] == b["angular_velocity"]
    assert np.allclose(a["velocity"], [1.0, 2.0 - x * omega], rtol=1e-6)
    assert np.allclose(b["velocity"], [1.0, 2.0 + x * omega], rtol=1e-6)
    assert np.allclose(conserved(src, out), [0.0, 0.0], atol=1e-5)


@pytest.mark.parametrize("share,rot", [(0.3, 0.0), (0.8, 1.1), (0.05, -2.5)])
def test_uneven_split_keeps_area_mass_and_momenta(share, rot):
    src = bubble(10, 40.0, 7.0, r=15.0, m=6.0, v=(-3.0, 0.5), w=-1.5, rot=rot)
    out = solver.stage_split(world(src, bubble(30, 500.0, 500.0)), 10, result(11), share, result(12), 1.0 - share)
    parts = np.array([by_id(out, 11), by_id(out, 12)])
    assert np.allclose(conserved(src, parts), [0.0, 0.0], atol=1e-4)
    # The parts sit on the source's axis, the smaller share farther out; the chord splits the area.
    axis = np.array([math.cos(rot), math.sin(rot)])
    d = parts["center"].astype(np.float64) - [40.0, 7.0]
    assert np.allclose(axis[0] * d[:, 1] - axis[1] * d[:, 0], 0.0, atol=1e-4) and (d @ axis)[0] < 0 < (d @ axis)[1]
    R, x1 = 15.0, abs((d @ axis)[0])
    # The segment of area share * pi R^2 has its centroid at 2/3 (R^2 - c^2)^1.5 / area.
    c = -R
    lo, hi = -R, R
    for _ in range(200):
        c = 0.5 * (lo + hi)
        if R * R * (math.asin(c / R) + math.pi / 2) + c * math.sqrt(R * R - c * c) < share * math.pi * R * R:
            lo = c
        else:
            hi = c
    assert np.isclose(x1, 2.0 / 3 * (R * R - c * c) ** 1.5 / (share * math.pi * R * R), rtol=1e-5)


def test_blocked_centroid_moves_and_momenta_stay():
    src = bubble(4, 0.0, 0.0, r=10.0, m=4.0, v=(1.0, 2.0), w=0.5)
    x = 4 * 10.0 / (3 * math.pi)
    blocker = bubble(2, -x, 0.0, r=3.0)
    out = solver.stage_split(world(src, blocker), 4, result(4), 0.5, result(6), 0.5)
    a, b = by_id(out, 4), by_id(out, 6)
    # The first part's centroid is bubble 2's centre: a quarter of the way to the source centre.
    assert np.allclose(a["center"], [F(-x) + (0 - F(-x)) * F(0.25), 0.0], rtol=1e-6)
    assert np.allclose(b["center"], [x, 0.0], rtol=1e-6)
    # The parts' centre of mass has moved off the source's centre; the momenta are kept about it.
    assert conserved(src, np.array([a, b]))[0] > 0.5


def test_split_refusals():
    src = bubble(4, 0.0, 0.0)
    w = world(src, bubble(9, 100.0, 0.0))
    for s0, s1, first, second, reason in [(0.5, 0.6, 4, 6, abi.BP_REJECT_INVALID_STRUCTURAL_OPERATION),
                                          (1.0, 0.0, 4, 6, abi.BP_REJECT_INVALID_STRUCTURAL_OPERATION),
                                          (0.5, 0.5, 4, 9, abi.BP_REJECT_DUPLICATE_ID),
                                          (0.5, 0.5, 6, 6, abi.BP_REJECT_DUPLICATE_ID)]:
        with pytest.raises(solver.StageRejected) as e:
            solver.stage_split(w, 4, result(first), s0, result(second), s1)
        assert e.value.reason == reason, (s0, s1, first, second)
    with pytest.raises(solver.StageRejected) as e:
        solver.stage_split(w, 5, result(4), 0.5, result(6), 0.5)
    assert e.value.reason == abi.BP_REJECT_UNKNOWN_OBJECT
