"""hipGraph replay of batch steps and the autotuner: both change time only, never results.

The graph test replays the captured step several times with eager steps and host reads between
the replays, and checks every replay against a batch stepped eagerly (graph replays of memset
nodes have misbehaved on this hardware; the capture clears its result records with a kernel)."""
import numpy as np
import pytest

from bubblephysics_amd import abi, autotune, scenes, solver

STATE = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def same_bits(a, b):
    return len(a) == len(b) and all(
        np.array_equal(np.ascontiguousarray(a[f]).view(np.uint32), np.ascontiguousarray(b[f]).view(np.uint32)) for f in STATE)


def worlds(seed, sizes, bodies_max=6):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [scenes.batch_world(rng, n, bodies_max=bodies_max) for n in sizes]


def test_tuning_table_round_trips_through_json():
    t = np.zeros(2, abi.TUNE_DTYPE)
    t[0]["min_bubbles"], t[0]["max_bubbles"], t[0]["best_class"], t[0]["best_bucket_shift"] = 16, 31, 1, -1
    t[1]["min_bubbles"], t[1]["max_bubbles"], t[1]["best_class"] = 32, 63, 3
    t[0]["class_rate"][:3] = [1.0, 2.0, 3.0]
    back = autotune.from_json(autotune.to_json(t, "gfx950"))
    assert back.tobytes() == t.tobytes()


@pytest.mark.gpu
def test_graph_replays_match_eager_steps():
    ws = worlds(51, [20, 90, 300, 1100, 2600, 5000])
    graph, eager = solver.Batch(ws), solver.Batch(ws)
    graph.capture(2)
    for rep in range(3):
        gi = graph.replay()
        ei = eager.step(2)
        assert np.array_equal(gi["substep_count"], ei["substep_count"]) and np.array_equal(gi["status"], ei["status"]), rep
        for w in range(len(ws)):
            assert same_bits(graph.read_bubbles(w), eager.read_bubbles(w)), (rep, w)
        # Eager work between replays: the graph steps from whatever state the batch holds.
        graph.step(1)
        eager.step(1)


@pytest.mark.gpu
def test_graph_is_dropped_when_the_layout_changes():
    ws = worlds(52, [40, 200])
    b = solver.Batch(ws)
    b.capture(1)
    b.replay()
    b.set_resident(1, False)
    with pytest.raises(abi.BubblePhysicsError):
        b.replay()


@pytest.mark.gpu
def test_autotune_keeps_state_and_results():
    ws = worlds(53, [16, 24, 40, 70, 150, 300, 600, 1300, 2800])
    tuned, plain = solver.Batch(ws), solver.Batch(ws)
    tuned.step(1)
    plain.step(1)
    before = [tuned.read_bubbles(w) for w in range(len(ws))]
    table = autotune.tune(tuned, repeats=1)
    # best_class -1 keeps the automatic classes (no configuration beat them by more than 2 %).
    assert len(table) >= 5 and np.all(table["best_class"] >= -1) and np.all(table["best_rate"] > 0)
    for r in table:
        assert (r["class_rate"] > 0).any()
        if r["best_class"] >= 0:
            assert r["class_rate"][r["best_class"]] > 0
    for w in range(len(ws)):
        assert same_bits(tuned.read_bubbles(w), before[w]), w
    ti, pi = tuned.step(3), plain.step(3)
    assert np.array_equal(ti["substep_count"], pi["substep_count"])
    for w in range(len(ws)):
        assert same_bits(tuned.read_bubbles(w), plain.read_bubbles(w)), w
    # The table applies to another batch without trials and changes nothing but the classes.
    again = solver.Batch(ws)
    autotune.apply(again, table)
    again.step(4)
    for w in range(len(ws)):
        assert same_bits(again.read_bubbles(w), plain.read_bubbles(w)), w
