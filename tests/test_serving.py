"""Serving: subset launches, asynchronous completion, device-side transactions, residency and the
native continuous-batching server. The GPU tests hold served worlds to the oracle (after the
host-side statement of each transaction) and to the same world served alone."""
import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, serving, solver

STATE = ("center", "previous_center", "velocity", "rotation", "angular_velocity")
PROPS = ("id", "mass", "inverse_mass", "radius", "contact_stiffness", "damping", "flags", "category", "mask")


def same_bits(a, b, props=False):
    if len(a) != len(b):
        return False
    fields = STATE + (PROPS if props else ("id",))
    return all(np.array_equal(np.ascontiguousarray(a[f]).view(np.uint8), np.ascontiguousarray(b[f]).view(np.uint8))
               for f in fields)


def worlds(seed, sizes, bodies_max=4):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [scenes.batch_world(rng, n, bodies_max=bodies_max) for n in sizes]


def request(world, kind, seed, arrival=0.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    return serving.Request(arrival, world, **serving.transaction_params(rng, kind))


# ---- CPU: request traces and the host-side statement of the transactions ------------------
def test_poisson_trace_is_seeded_and_has_the_rate():
    a = serving.poisson_requests(100, 2000.0, 2.0, seed=3)
    b = serving.poisson_requests(100, 2000.0, 2.0, seed=3)
    key = lambda r: (r.arrival, r.world, r.transaction, r.index, r.impulse, r.direction)
    assert [key(r) for r in a] == [key(r) for r in b]
    assert abs(len(a) - 4000) < 4 * np.sqrt(4000)
    assert all(x.arrival <= y.arrival for x, y in zip(a, a[1:]))
    frac = np.mean([r.transaction is not None for r in a])
    assert 0.07 < frac < 0.13
    for r in a:
        assert abs(np.hypot(*r.direction) - 1.0) < 1e-6
    trace = serving.to_trace(a)
    assert trace.dtype == abi.REQUEST_DTYPE and np.array_equal(trace["tag"], np.arange(len(a)))


def stage_kw(scene, bodies=None):
    """Staging context of a world: its bodies (at the given poses), vertices and region."""
    return dict(bodies=scene.bodies if bodies is None else bodies, vertices=scene.vertices, region=tuple(scene.config.region),
                bands=scene.bands)


@pytest.mark.parametrize("kind", serving.KINDS)
def test_transactions_keep_ids_sorted_and_conserve_what_they_should(kind):
    b = scenes.make_bubbles([(0, 0), (30, 0), (60, 0), (90, 5)], [10, 12, 8, 9], [(1, 2), (3, 4), (5, 6), (7, 8)])
    for seed in range(6):
        req = request(0, kind, seed)
        out, reason = serving.apply_transaction(b, req)
        assert reason == abi.BP_REJECT_NONE
        assert np.all(np.diff(out["id"]) > 0)
        mv = lambda x: (x["mass"][:, None].astype(np.float64) * x["velocity"].astype(np.float64)).sum(axis=0)
        i = req.index % len(b)
        if kind in ("merge", "split"):
            # Mass, momentum and area are kept (StructuralOperations.swift:10-151).
            assert np.isclose(out["mass"].sum(dtype=np.float64), b["mass"].sum(dtype=np.float64), rtol=1e-6)
            assert np.allclose(mv(out), mv(b), rtol=1e-5, atol=1e-4)
            assert np.isclose((out["radius"].astype(np.float64) ** 2).sum(), (b["radius"].astype(np.float64) ** 2).sum(), rtol=1e-5)
        if kind == "merge":
            k = i + 1 if i + 1 < len(b) else i - 1
            assert len(out) == len(b) - 1 and b["id"][max(i, k)] not in out["id"]
        if kind == "split":
            assert len(out) == len(b) + 1 and out["id"][-1] == b["id"].max() + 1
            # Equal shares: the two half-disc centroids sit symmetrically about the old centre,
            # 4 r / (3 pi) away (the centroid of a half disc).
            c = b["center"][i].astype(np.float64)
            d1, d2 = out["center"][i] - c, out["center"][-1] - c
            assert np.allclose(d1, -d2, atol=1e-4)
            assert np.isclose(np.hypot(*d1), 4 * b["radius"][i] / (3 * np.pi), rtol=1e-4)
        if kind == "add":
            assert len(out) == len(b) + 1 and out["id"][-1] == b["id"].max() + 1
            assert np.isclose(np.hypot(*(out["center"][-1] - b["center"][i])), 2 * b["radius"][i], rtol=1e-5)
            assert np.all(out["velocity"][-1] == 0)
        if kind == "remove":
            assert len(out) == len(b) - 1 and b["id"][i] not in out["id"]
        if kind == "impulse":
            dv = out["velocity"][i] - b["velocity"][i]
            want = np.array(req.impulse, np.float64) / np.float64(b["mass"][i])
            assert len(out) == len(b) and np.allclose(dv, want, rtol=1e-5, atol=1e-4)


def test_transactions_leave_a_single_bubble_world_whole():
    b = scenes.make_bubbles([(0, 0)], [10], [(1, 2)])
    assert len(serving.apply_transaction(b, request(0, "remove", 1))[0]) == 1
    out, reason = serving.apply_transaction(b, request(0, "merge", 1))
    assert len(out) == 1 and reason == abi.BP_REJECT_INVALID_STRUCTURAL_OPERATION


def test_impulse_follows_the_reference_role_and_finiteness_rules():
    # TransactionStager.swift:312-324: only .dynamic and .motorDriven bubbles take an impulse, the
    # impulse and the velocity after it must be finite; a refused impulse changes nothing.
    b = scenes.make_bubbles([(0, 0), (30, 0), (60, 0)], [10, 12, 8], [(1, 2), (3, 4), (5, 6)])
    b["flags"][1] = abi.BP_BUBBLE_ENABLED | abi.BP_BUBBLE_STATIC      # .static
    b["inverse_mass"][1] = 0
    b["flags"][2] = abi.BP_BUBBLE_ENABLED                              # .pathDriven / .kinematic: not force-driven
    bands = np.zeros(1, abi.BAND_DTYPE)
    bands[0]["bubble_id"] = b["id"][0]                                 # .playerDriven: force-driven with a band
    for i, reason in ((1, abi.BP_REJECT_INCOMPATIBLE_MOTION_ROLE), (2, abi.BP_REJECT_INCOMPATIBLE_MOTION_ROLE)):
        req = serving.Request(0.0, 0, "impulse", index=i, impulse=(5.0, 5.0))
        out, got = serving.apply_transaction(b, req)
        assert got == reason and np.array_equal(out, b)
    out, got = serving.apply_transaction(b, serving.Request(0.0, 0, "impulse", index=0, impulse=(5.0, 5.0)), bands=bands)
    assert got == abi.BP_REJECT_INCOMPATIBLE_MOTION_ROLE and np.array_equal(out, b)
    out, got = serving.apply_transaction(b, serving.Request(0.0, 0, "impulse", index=0, impulse=(float("nan"), 0.0)))
    assert got == abi.BP_REJECT_INVALID_VALUE and np.array_equal(out, b)
    huge = scenes.make_bubbles([(0, 0)], [10], [(3e38, 0)])
    out, got = serving.apply_transaction(huge, serving.Request(0.0, 0, "impulse", index=0, impulse=(3e38, 0.0)))
    assert got == abi.BP_REJECT_INVALID_VALUE and np.array_equal(out, huge)


def test_structural_transactions_follow_placement_and_keep_static_bubbles_static():
    # Placement (PlacementGeometry.swift:42-94): an add outside the region is refused; merge and
    # split of static bubbles keep inverse mass 0 (resultDescriptor, TransactionStager.swift:621-629).
    b = scenes.make_bubbles([(10, 10), (40, 10)], [10, 12], [(0, 0), (0, 0)])
    b["flags"] = abi.BP_BUBBLE_ENABLED | abi.BP_BUBBLE_STATIC
    b["inverse_mass"] = 0
    region = (0.0, 0.0, 100.0, 100.0)
    add = serving.Request(0.0, 0, "add", index=0, direction=(-1.0, 0.0))      # lands at x = -10
    out, got = serving.apply_transaction(b, add, region=region)
    assert got == abi.BP_REJECT_INVALID_PLACEMENT and np.array_equal(out, b)
    for kind in ("merge", "split"):
        out, got = serving.apply_transaction(b, serving.Request(0.0, 0, kind, index=0), region=region)
        assert got == abi.BP_REJECT_NONE and np.all(out["inverse_mass"] == 0), kind


# ---- GPU: batch subset launches -------------------------------------------------------------
@pytest.mark.gpu
def test_listed_worlds_advance_and_the_others_do_not(oracle):
    ws = worlds(21, [40, 300, 70, 1200, 20])
    batch = solver.Batch(ws)
    before = [batch.read_bubbles(w) for w in range(len(ws))]
    infos = batch.step_worlds([3, 1], 3)
    assert infos.shape == (3, 2)
    for j, w in enumerate([3, 1]):
        cb, _, cf, _ = oracle.run(ws[w], 3)
        assert np.array_equal(infos[:, j]["substep_count"], cf["substeps"])
        assert same_bits(batch.read_bubbles(w), cb)
    for w in (0, 2, 4):
        assert same_bits(batch.read_bubbles(w), before[w])


@pytest.mark.gpu
def test_submitted_step_is_collected_by_polling(oracle):
    ws = worlds(22, [50, 500])
    batch = solver.Batch(ws)
    batch.submit_worlds([1, 0], 2)
    infos = None
    while infos is None:
        infos = batch.poll()
    for j, w in enumerate([1, 0]):
        cb, _, cf, _ = oracle.run(ws[w], 2)
        assert np.array_equal(infos[:, j]["substep_count"], cf["substeps"])
        assert same_bits(batch.read_bubbles(w), cb)


@pytest.mark.gpu
def test_evicted_world_keeps_its_state_and_resumes(oracle):
    ws = worlds(24, [80, 400, 30])
    batch = solver.Batch(ws)
    batch.step(2)
    kept = batch.read_bubbles(1)
    batch.set_resident(1, False)
    with pytest.raises(RuntimeError):
        batch.step_worlds([1], 1)
    batch.step_worlds([0, 2], 1)
    batch.set_resident(1, True)
    assert same_bits(batch.read_bubbles(1), kept)
    cur = ws[1].copy()
    cur.bubbles, _, _, _ = oracle.run(ws[1], 2)
    batch.step_worlds([1], 1)
    cb, _, _, _ = oracle.run(cur, 1)
    assert same_bits(batch.read_bubbles(1), cb)


# ---- GPU: the native server -------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("kind", serving.KINDS)
def test_served_transaction_and_step_match_the_oracle(oracle, kind):
    # A step, then a served request carrying the transaction, then a plain served request: the
    # device transaction kernel must equal the host statement, and the steps the oracle's.
    ws = worlds(23, [60, 900])
    batch = solver.Batch(ws)
    batch.step(1)
    with serving.Server(batch) as server:
        req = request(1, kind, 5)
        server.submit(req, tag=7)
        server.wait_idle()
        r = server.poll()
        assert len(r) == 1 and r[0]["tag"] == 7 and r[0]["world"] == 1
        cur = ws[1].copy()
        cur.bubbles, cur.bodies, _, _ = oracle.run(ws[1], 1)
        cur.bubbles, reason = serving.apply_transaction(cur.bubbles, req, **stage_kw(cur))
        assert r[0]["bubbles"] == len(cur.bubbles) and r[0]["transaction_result"] == reason
        cb, _, cf, _ = oracle.run(cur, 1)
        assert r[0]["status"] == cf[0]["status"] and r[0]["substeps"] == cf[0]["substeps"]
        assert same_bits(server.read_bubbles(1), cb, props=True)
        server.submit(serving.Request(0.0, 1), tag=8)
        server.wait_idle()
        r = server.poll()
        cur.bubbles = cb
        cb, _, cf, _ = oracle.run(cur, 1)
        assert r[0]["substeps"] == cf[0]["substeps"]
        assert same_bits(server.read_bubbles(1), cb, props=True)
    # The batch takes the worlds back with their new counts.
    assert same_bits(batch.read_bubbles(1), cb, props=True)
    cur.bubbles = cb
    cb, _, cf, _ = oracle.run(cur, 1)
    assert batch.step(1)[0, 1]["substep_count"] == cf[0]["substeps"]
    assert same_bits(batch.read_bubbles(1), cb, props=True)


@pytest.mark.gpu
def test_growing_world_moves_to_a_wider_kernel_class(oracle):
    # 64 bubbles fill the narrowest configuration (64 threads x 1); an add moves the world up.
    rng = np.random.Generator(np.random.PCG64(31))
    w = scenes.batch_world(rng, 64, bodies_max=0, packing=(0.3, 0.3))
    assert len(w.bubbles) == 64
    with serving.Server([w]) as server:
        server.submit(serving.Request(0.0, 0), tag=0)
        req = request(0, "add", 3)
        server.submit(req, tag=1)
        server.wait_idle()
        r = np.sort(server.poll(), order="tag")
        assert list(r["bubbles"]) == [64, 65] and r[0]["kernel_class"] != r[1]["kernel_class"]
        cur = w.copy()
        cur.bubbles, cur.bodies, _, _ = oracle.run(w, 1)
        cur.bubbles, reason = serving.apply_transaction(cur.bubbles, req, **stage_kw(cur))
        assert reason == abi.BP_REJECT_NONE
        cb, _, _, _ = oracle.run(cur, 1)
        assert same_bits(server.read_bubbles(0), cb, props=True)


@pytest.mark.gpu
def test_transaction_beyond_the_packed_slots_is_rejected():
    ws = worlds(26, [16])
    n = len(ws[0].bubbles)
    slots = n + max(8, n // 8)
    with serving.Server(ws) as server:
        for k in range(slots - n + 24):
            server.submit(request(0, "add", 100 + k), tag=k)
        server.wait_idle()
        r = np.sort(server.poll(), order="tag")
        overflow = r["status"] == abi.BP_ERR_CAPACITY_OVERFLOW
        applied = ~overflow & (r["transaction_result"] == abi.BP_REJECT_NONE)
        # Adds that pass placement grow the world until its slots are full; later ones are refused
        # for capacity (no step), while adds that fail placement are refused with the reason.
        assert applied.sum() == slots - n and overflow.sum() > 0
        assert np.all(r["bubbles"][~overflow] <= slots)
        assert server.stats()["rejected"] == overflow.sum()
        assert len(server.read_bubbles(0)) == slots


@pytest.mark.gpu
@pytest.mark.parametrize("sizes,in_place", [([30, 200, 60, 700, 45, 120, 2500], False),
                                            ([150, 140, 160, 150, 145, 155, 150, 140], True)],
                         ids=["uneven-repacks", "even-in-place"])
def test_each_served_world_gets_the_result_it_gets_alone(sizes, in_place):
    # Continuous batching with a budget that forces evictions: every world ends in the same
    # state as when its own requests (and transactions) are replayed on it alone, with the
    # transactions applied on the host by their statement in numpy. Uneven worlds need repacking
    # rounds (the large world fits no evicted world's ranges); worlds of about one size move into
    # the ranges of the worlds they evict, with no repack.
    ws = worlds(25, sizes)
    requests = serving.poisson_requests(len(ws), 600.0, 0.4, seed=9, transaction_fraction=0.3)
    probe = solver.Batch(ws)
    budget = sum(sorted(probe.footprint(w) for w in range(len(ws)))[-3:])   # about three worlds at a time
    probe.close()
    server = serving.Server(ws, memory_budget=budget)
    stats = server.run(requests)
    counters = server.stats()
    assert counters["completed"] == len(requests) and counters["evictions"] > 0
    if in_place:
        assert counters["evictions"] > counters["repacks"], counters
    else:
        assert counters["repacks"] > 0
    final = []
    for w in range(len(ws)):
        final.append(server.read_bubbles(w))
    server.close()
    s = stats.summary()
    assert s["served"] == len(requests) and s["p99_ms"] >= s["p50_ms"] > 0
    assert np.all(stats.results["completed"] >= stats.results["arrival"])
    for w in range(len(ws)):
        served = sorted((r for r in requests if r.world == w), key=lambda r: r.done)
        alone = solver.Batch([ws[w]])
        for r in served:
            if r.status == abi.BP_ERR_CAPACITY_OVERFLOW:    # refused by the server: no transaction, no step
                continue
            if r.status == abi.BP_ERR_UNAVAILABLE:          # the world failed earlier: nothing more runs
                assert alone.step_worlds([0], 1)[0, 0]["status"] == abi.BP_ERR_UNAVAILABLE
                continue
            if r.transaction:
                kw = stage_kw(ws[w], bodies=alone.read_bodies(0, len(ws[w].bodies)))
                staged, reason = serving.apply_transaction(alone.read_bubbles(0), r, **kw)
                assert reason == r.tx_result, (w, r.transaction, reason, r.tx_result)
                if reason == abi.BP_REJECT_NONE:
                    alone.write_bubbles(0, staged)
            info = alone.step_worlds([0], 1)[0, 0]
            assert info["status"] == r.status and info["substep_count"] == r.substeps, (w, info["status"], r.status)
        assert same_bits(final[w], alone.read_bubbles(0), props=True), w


@pytest.mark.gpu
def test_structural_transactions_after_region_removals_match_alone():
    # A step that removes bubbles from the region leaves the world's serving mirror stale (the
    # server deletes the bubbles on the device afterwards): the next add, merge or split is
    # staged from a read-back of the device state instead. Results equal the world served alone.
    rng = np.random.Generator(np.random.PCG64(41))
    w = scenes.batch_world(rng, 150, bodies_max=3)
    w.segments = w.segments[:0]                          # no walls: bubbles may leave
    side = float(w.config.region[2])
    w.config.region[2] = side - 60.0                     # the right strip lies outside the region
    kinds = [None, "merge", None, "add", "split", None, "merge", "add", None, "split", "merge", None]
    reqs = []
    for k, kind in enumerate(kinds * 2):
        r = request(0, kind, 100 + k, arrival=0.002 * k)
        reqs.append(r)
    server = serving.Server([w])
    stats = server.run(reqs)
    final = server.read_bubbles(0)
    server.close()
    assert stats.summary()["served"] == len(reqs)
    alone = solver.Batch([w])
    removed_any = False
    for r in sorted(reqs, key=lambda r: r.done):
        if r.status == abi.BP_ERR_UNAVAILABLE:
            assert alone.step_worlds([0], 1)[0, 0]["status"] == abi.BP_ERR_UNAVAILABLE
            continue
        if r.transaction:
            kw = stage_kw(w, bodies=alone.read_bodies(0, len(w.bodies)))
            staged, reason = serving.apply_transaction(alone.read_bubbles(0), r, **kw)
            assert reason == r.tx_result, (r.transaction, reason, r.tx_result)
            if reason == abi.BP_REJECT_NONE:
                alone.write_bubbles(0, staged)
        info = alone.step_worlds([0], 1)[0, 0]
        removed_any |= int(info["region_removals"]) > 0
        assert info["status"] == r.status and info["substep_count"] == r.substeps
    assert removed_any
    assert same_bits(final, alone.read_bubbles(0), props=True)
