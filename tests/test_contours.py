"""Draw-only contours of a world's published state (bp_world_contours, bp_batch_contours).

The oracle's restatement (bpo_contours) is pinned by the reference's contour tests:
  * ContourChordTests.swift:14-56: a pair pressed together by equal and opposite forces F; each
    contour's flat chord lies at r (1 - eps) (within 2e-3 of the radii sum) and the two chords meet
    (their sum equals the centre distance within 1e-3 of the radii sum), also for different sizes
    and stiffnesses;
  * ContactContourTests.swift:52-128: a soft bubble pressed against a wall and a soft pair pressed
    together: every frame after the fifth the contour is no shallower than the geometric dent
    (0.02), every point lies within r + 1e-3, no edge is longer than the segment length, and the
    settled contour does not enter the wall or the partner by more than 0.5;
  * closed forms: a free bubble's contour is its circle (count = max(64, ceil(4 pi r / 8)) points);
  * a planted mistake (the chord at the radius instead of r (1 - eps)) fails the chord pin.
The GPU tests hold the HIP kernel to the oracle bit for bit on worlds the step tests already hold
bitwise (the contour reads only the published state and the rows of the last grid build).
"""
import math

import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, solver
from oracle import bp_oracle
import reference_scenes as rs

K = 229.0
CFG = abi.contour_config()


def eps(q):
    return 2 * q / ((2 * q + 1) + math.sqrt(4 * q + 1))


def chord_world(force, rb=24.0, kb=K):
    ra = 24.0
    rest = ra * (1 - eps(force / K)) + rb * (1 - eps(force / kb))
    half = rest / 2
    w = rs.World(half=2000)
    mb = 1.0 if rb == 24.0 else rs.game_mass(7168)
    w.bubble(1, (-half, 0), radius=ra, mass=1.0).bubble(2, (half, 0), radius=rb, mass=mb, stiffness=kb)
    w.motor(1, (force, 0)).motor(2, (-force, 0))
    return w.build()


def oracle_contours(scene, steps, plant=0):
    cb, _, frames, ex = bp_oracle.run(scene, steps, adjacency=True)
    assert all(int(f["status"]) == abi.BP_OK for f in frames)
    bp_oracle.set_plant(plant)
    try:
        pts, shapes = bp_oracle.contours(cb, scene.segments, ex["row_offsets"], ex["entries"], CFG)
    finally:
        bp_oracle.set_plant(0)
    return cb, pts, shapes


def chord(points, count, center, direction):
    return max(float(np.dot(p - center, direction)) for p in points[:count])


def chords_hold(force, rb=24.0, kb=K, plant=0):
    scene = chord_world(force, rb, kb)
    cb, pts, shapes = oracle_contours(scene, 60, plant)
    pa, pb = cb["center"][0].astype(np.float64), cb["center"][1].astype(np.float64)
    d = float(np.linalg.norm(pb - pa))
    n = (pb - pa) / d
    ca = chord(pts[0], shapes[0]["count"], pa, n)
    cbh = chord(pts[1], shapes[1]["count"], pb, -n)
    ra, radii = 24.0, 24.0 + rb
    return (ra - ca > 0.05 and rb - cbh > 0.05 and ca + cbh - d <= 1e-3 * radii and abs(ca + cbh - d) <= 1e-3 * radii
            and abs(ca - ra * (1 - eps(force / K))) <= 2e-3 * radii and abs(cbh - rb * (1 - eps(force / kb))) <= 2e-3 * radii)


@pytest.mark.parametrize("force,rb,kb", [(100, 24.0, K), (1200, 24.0, K), (100, 148.0, K), (1200, 148.0, K),
                                         (1200, 24.0, 10 * K)])
def test_pair_chords_lie_at_r_one_minus_eps_and_meet(force, rb, kb):
    assert chords_hold(force, rb, kb)


def test_planted_chord_at_the_radius_is_caught():
    assert not chords_hold(1200, plant=6)


def test_free_bubble_contour_is_its_circle():
    w = rs.World(half=2000)
    w.bubble(1, (3.0, -7.0), radius=40.0)
    scene = w.build()
    cb, pts, shapes = oracle_contours(scene, 1)
    count = int(shapes[0]["count"])
    assert count == max(64, math.ceil(4 * math.pi * 40 / 8)) == 63 + 1
    c = cb["center"][0].astype(np.float64)
    r = np.linalg.norm(pts[0][:count] - c, axis=1)
    assert np.allclose(r, 40.0, atol=1e-4) and shapes[0]["maximum"] == 0 and abs(shapes[0]["effective_radius"] - 40) < 1e-4


# ---- ContactContourTests.swift ---------------------------------------------------------------
RADIUS = 20.0
INNER = (180.0, 320.0)
THICK = 48.0


def crowd_world(stiffness):
    outer = (INNER[0] + THICK, INNER[1] + THICK)
    w = rs.World(region=(-outer[0], -outer[1], outer[0], outer[1]), stiffness=stiffness)
    rects = [((-outer[0], -outer[1]), (-INNER[0], outer[1])), ((INNER[0], -outer[1]), (outer[0], outer[1])),
             ((-INNER[0], -outer[1]), (INNER[0], -INNER[1])), ((-INNER[0], INNER[1]), (INNER[0], outer[1]))]
    for k, (lo, hi) in enumerate(rects):
        w.static_polygon(1000 + k, [(lo[0], lo[1]), (hi[0], lo[1]), (hi[0], hi[1]), (lo[0], hi[1])])
    return w


def shortfall(points, center, normal, distance):
    worst = -math.inf
    for p in points:
        off = p - center
        reach = float(np.linalg.norm(off))
        if reach <= 1e-6:
            continue
        u = off / reach
        align = float(np.dot(u, normal))
        limit = min(RADIUS, distance / align) if align > 1e-6 else RADIUS
        worst = max(worst, reach - limit)
    return worst


def invariants_hold(points, center):
    if len(points) < 3:
        return False
    for k, p in enumerate(points):
        if np.linalg.norm(p - center) > RADIUS + 1e-3 or np.linalg.norm(points[(k + 1) % len(points)] - p) > 8.0:
            return False
    return True


def step_with_contours(scene, frames):
    for _ in range(frames):
        cb, cbod, fr, ex = bp_oracle.run(scene, 1, adjacency=True, region=True)
        assert int(fr[0]["status"]) == abi.BP_OK
        scene.bubbles, scene.bodies = cb, cbod
        pts, shapes = bp_oracle.contours(cb, scene.segments, ex["row_offsets"], ex["entries"], CFG)
        yield cb, pts, shapes


def test_wall_contour_is_not_shallower_than_the_geometric_dent():
    wall_x = -INNER[0]
    w = crowd_world(RADIUS / 0.06)
    w.bubble(1, (wall_x + RADIUS, 0.0), radius=RADIUS, mass=1.0).set_global_force((-100.0, 0.0))
    scene = w.build()
    checked = 0
    for step, (cb, pts, shapes) in enumerate(step_with_contours(scene, 600)):
        if step < 5:
            continue
        c = cb["center"][0].astype(np.float64)
        contour = pts[0][: shapes[0]["count"]].astype(np.float64)
        assert invariants_hold(contour, c), step
        assert shortfall(contour, c, np.array([-1.0, 0.0]), float(c[0] - wall_x)) <= 0.02, step
        checked += 1
        if step == 599:
            assert RADIUS - min(np.linalg.norm(contour - c, axis=1)) > 0.15 * RADIUS
            assert -(contour[:, 0].min() - wall_x) <= 0.5
    assert checked > 500


def test_pair_contours_are_not_shallower_than_the_chord_and_do_not_overlap():
    w = crowd_world(RADIUS / 0.2)
    w.bubble(1, (-RADIUS, 0.0), radius=RADIUS, mass=1.0).bubble(2, (RADIUS, 0.0), radius=RADIUS, mass=1.0)
    w.motor(1, (30.0, 0.0)).motor(2, (-30.0, 0.0))
    scene = w.build()
    for step, (cb, pts, shapes) in enumerate(step_with_contours(scene, 400)):
        if step < 5:
            continue
        a, b = cb["center"][0].astype(np.float64), cb["center"][1].astype(np.float64)
        ca, cbb = pts[0][: shapes[0]["count"]].astype(np.float64), pts[1][: shapes[1]["count"]].astype(np.float64)
        assert invariants_hold(ca, a) and invariants_hold(cbb, b), step
        d = float(np.linalg.norm(b - a))
        n = (b - a) / d
        assert shortfall(ca, a, n, d / 2) <= 0.02 and shortfall(cbb, b, -n, d / 2) <= 0.02, step
        if step == 399:
            assert d < 2 * RADIUS - 5
            assert RADIUS - min(np.linalg.norm(ca - a, axis=1)) > 0.1 * RADIUS
            assert max(float(np.dot(p - a, n)) - d / 2 for p in ca) <= 0.5
            assert max(-float(np.dot(p - b, n)) - d / 2 for p in cbb) <= 0.5


# ---- GPU: bit for bit against the oracle ------------------------------------------------------
def gpu_vs_oracle(scene, steps):
    cb, pts, shapes = oracle_contours(scene, steps)
    world = solver.NativeWorld(scene)
    for k in range(steps):
        f = world.step(k)
        assert f.status == abi.BP_OK
    assert world.read_bubbles().tobytes() == cb.tobytes()
    gp, gs = world.contours(CFG, stride=pts.shape[1])
    for i in range(len(cb)):
        n = int(shapes[i]["count"])
        assert gp[i][:n].tobytes() == pts[i][:n].tobytes(), f"bubble {i}: contour points differ"
    assert gs.tobytes() == shapes.tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("force,rb", [(1200, 24.0), (100, 148.0)])
def test_gpu_chord_contours_equal_the_oracle(force, rb):
    gpu_vs_oracle(chord_world(force, rb), 60)


@pytest.mark.gpu
def test_gpu_config1_contours_equal_the_oracle():
    gpu_vs_oracle(scenes.config1(steps=10), 10)


@pytest.mark.gpu
def test_gpu_wall_press_contours_equal_the_oracle():
    w = crowd_world(RADIUS / 0.06)
    w.bubble(1, (-INNER[0] + RADIUS, 0.0), radius=RADIUS, mass=1.0).set_global_force((-100.0, 0.0))
    w.bubble(2, (0.0, 0.0), radius=RADIUS, mass=1.0).bubble(3, (30.0, 10.0), radius=15.0, mass=1.0)
    gpu_vs_oracle(w.build(), 40)


@pytest.mark.gpu
def test_gpu_batch_contours_equal_the_oracle():
    rng = np.random.Generator(np.random.PCG64(77))
    worlds = [scenes.batch_world(rng, n, bodies_max=6) for n in (40, 300, 900)]
    for w in worlds:
        w.config.export_adjacency = 1
    batch = solver.Batch(worlds)
    batch.step(5)
    for wi, scene in enumerate(worlds):
        cb, pts, shapes = oracle_contours(scene, 5)
        assert batch.read_bubbles(wi).tobytes() == cb.tobytes()
        stride = pts.shape[1]
        gp, gs = batch.contours(wi, stride, CFG)
        for i in range(len(cb)):
            n = int(shapes[i]["count"])
            assert gp[i][:n].tobytes() == pts[i][:n].tobytes(), f"world {wi} bubble {i}"
        assert gs.tobytes() == shapes.tobytes()
