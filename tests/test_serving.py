"""Serving: subset launches, asynchronous completion, transactions, residency and continuous
batching. The GPU tests hold the served worlds to the oracle and to the same world served alone."""
import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, serving, solver

STATE = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def same_bits(a, b):
    return all(np.array_equal(np.ascontiguousarray(a[f]).view(np.uint32), np.ascontiguousarray(b[f]).view(np.uint32))
               for f in STATE) and np.array_equal(a["id"], b["id"])


def worlds(seed, sizes):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [scenes.batch_world(rng, n, bodies_max=4) for n in sizes]


# ---- CPU: request traces and transactions -------------------------------------------------
def test_poisson_trace_is_seeded_and_has_the_rate():
    a = serving.poisson_requests(100, 2000.0, 2.0, seed=3)
    b = serving.poisson_requests(100, 2000.0, 2.0, seed=3)
    assert [(r.arrival, r.world, r.transaction, r.seed) for r in a] == [(r.arrival, r.world, r.transaction, r.seed) for r in b]
    assert abs(len(a) - 4000) < 4 * np.sqrt(4000)
    assert all(x.arrival <= y.arrival for x, y in zip(a, a[1:]))
    frac = np.mean([r.transaction is not None for r in a])
    assert 0.07 < frac < 0.13


@pytest.mark.parametrize("kind", serving.KINDS)
def test_transactions_keep_ids_sorted_and_conserve_what_they_should(kind):
    b = scenes.make_bubbles([(0, 0), (30, 0), (60, 0), (90, 5)], [10, 12, 8, 9], [(1, 2), (3, 4), (5, 6), (7, 8)])
    out = serving.apply_transaction(b, kind, seed=11)
    assert np.all(np.diff(out["id"]) > 0)
    mv = lambda x: (x["mass"][:, None].astype(np.float64) * x["velocity"].astype(np.float64)).sum(axis=0)
    if kind in ("merge", "split"):
        # Mass, momentum and area are kept (StructuralOperations.swift:10-151).
        assert np.isclose(out["mass"].sum(dtype=np.float64), b["mass"].sum(dtype=np.float64), rtol=1e-6)
        assert np.allclose(mv(out), mv(b), rtol=1e-5)
        assert np.isclose((out["radius"].astype(np.float64) ** 2).sum(), (b["radius"].astype(np.float64) ** 2).sum(), rtol=1e-5)
    if kind == "merge":
        assert len(out) == len(b) - 1
    if kind == "split":
        assert len(out) == len(b) + 1
    if kind == "add":
        assert len(out) == len(b) + 1 and out["id"][-1] == b["id"].max() + 1
    if kind == "remove":
        assert len(out) == len(b) - 1
    if kind == "impulse":
        assert len(out) == len(b) and (out["velocity"] != b["velocity"]).sum() >= 1


# ---- GPU ------------------------------------------------------------------------------
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
@pytest.mark.parametrize("kind", serving.KINDS)
def test_transaction_between_steps_matches_oracle(oracle, kind):
    ws = worlds(23, [60, 900])
    batch = solver.Batch(ws)
    batch.step(1)
    cur = ws[1].copy()
    cur.bubbles, _, _, _ = oracle.run(ws[1], 1)
    changed = serving.apply_transaction(batch.read_bubbles(1), kind, seed=5)
    batch.write_bubbles(1, changed)
    cur.bubbles = changed
    infos = batch.step_worlds([1], 2)
    cb, _, cf, _ = oracle.run(cur, 2)
    assert np.array_equal(infos[:, 0]["substep_count"], cf["substeps"])
    assert same_bits(batch.read_bubbles(1), cb)


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


@pytest.mark.gpu
def test_each_served_world_gets_the_result_it_gets_alone():
    # Continuous batching with a budget that forces evictions: every world ends in the same
    # state as when its own requests (and transactions) are replayed on it alone.
    ws = worlds(25, [30, 200, 60, 700, 45, 120])
    requests = serving.poisson_requests(len(ws), 400.0, 0.25, seed=9, transaction_fraction=0.3)
    probe = solver.Batch(ws)
    budget = sum(sorted(probe.footprint(w) for w in range(len(ws)))[-3:])   # about three worlds at a time
    probe.close()
    server = serving.Server(ws, memory_budget=budget)
    order = {w: [] for w in range(len(ws))}
    stats = server.run(requests, on_done=lambda r: order[r.world].append(r))
    assert stats.served == len(requests) and stats.evictions > 0
    for w, served in order.items():
        if not served:
            continue
        alone = solver.Batch([ws[w]])
        for r in sorted(served, key=lambda r: r.done):
            if r.transaction:
                alone.write_bubbles(0, serving.apply_transaction(alone.read_bubbles(0), r.transaction, r.seed))
            alone.step_worlds([0], 1)
        server.batch.set_resident(w, True)
        assert same_bits(server.batch.read_bubbles(w), alone.read_bubbles(0)), w
