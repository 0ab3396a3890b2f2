"""Rows that overflow the candidate limit: the reference keeps the k smallest keys of each bubble's
forward row and of each body's row, and the first k adjacency entries of each bubble
(NativeGeometryKernels.metal-source:230-251, 276-298, 350-415; NativeContactKernels.metal-source:
190-238). The GPU builds full rows and cuts them to exactly that when a build drops candidates.

CPU: the oracle's truncated rows equal its full rows (k = 128, nothing dropped) cut by that rule,
restated here over plain lists, so the rule the kernels implement is the reference's.
GPU: worlds stepped with a small candidate limit, so that rows overflow in every build (until the
contacts they drop trip the contact assertion), on every kernel family -- LDS rows, global rows of
the LDS kernel, the global-state kernel, the cooperative grid -- against the oracle, which
truncates as the reference does: state bits, substep and rebuild counts, the overflow counters,
the status and the adjacency rows."""
import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, solver
from test_gpu_world import assert_same_state, bits



def limited(scene, limit):
    scene.config.candidate_limit = limit
    return scene


def rows_of(offsets, entries):
    return [[int(e) for e in entries[offsets[i]:offsets[i + 1]]] for i in range(len(offsets) - 1)]


def cut_rows(full, k):
    """The reference's rows from full rows sorted by (kind, neighbour), entries kind << 30 | index."""
    kind, index = (lambda e: e >> 30), (lambda e: e & ((1 << 30) - 1))
    forward = [[index(e) for e in r if kind(e) == 0 and index(e) > i] for i, r in enumerate(full)]
    cut = [f[k - 1] if len(f) > k else None for f in forward]      # largest forward neighbour kept
    body_rows = {}
    for i, r in enumerate(full):
        for e in r:
            if kind(e) == 2:
                body_rows.setdefault(index(e), []).append(i)
    out = []
    for i, r in enumerate(full):
        arrived = [e for e in r if kind(e) == 0 and (cut[min(i, index(e))] is None or max(i, index(e)) <= cut[min(i, index(e))])]
        arrived += [e for e in r if kind(e) == 1][: k - min(len(forward[i]), k)]
        arrived += [e for e in r if kind(e) == 2 and body_rows[index(e)].index(i) < k]
        out.append(arrived[:k])
    return out


@pytest.mark.parametrize("make,k", [(scenes.config1, 1), (scenes.config1, 3), (scenes.config1, 6), (scenes.config2, 2),
                                    (scenes.config2, 3)])
def test_oracle_truncation_is_the_cut_of_full_rows(oracle, make, k):
    full_scene, cut_scene = make(), limited(make(), k)
    full_scene.config.candidate_limit = 128
    _, _, ff, fe = oracle.run(full_scene, 1, adjacency=True)
    _, _, cf, ce = oracle.run(cut_scene, 1, adjacency=True)
    assert ff["rebuilds"][0] == cf["rebuilds"][0] == 0          # both rows come from the step's first build
    full = rows_of(fe["row_offsets"], fe["entries"])
    assert ff["candidate_overflow"][0] == 0 and cf["candidate_overflow"][0] > 0
    assert cut_rows(full, k) == rows_of(ce["row_offsets"], ce["entries"])
    if make is scenes.config2:
        widest_body = max(np.bincount([e & 0xFFFF for r in full for e in r if e >> 30 == 2]))
        assert widest_body > k                                       # a body row is cut too


def step_world_against_oracle(scene, steps, oracle):
    world = solver.NativeWorld(scene)
    cur = scene.copy()
    dropped, paths = 0, set()
    for k in range(steps):
        frame = world.step(k)
        paths.add(frame.execution_path)
        cb, cbod, cf, ex = oracle.run(cur, 1, adjacency=True)
        c = cf[0]
        assert frame.status == c["status"], f"step {k}"
        assert frame.substep_count == c["substeps"] and frame.grid_rebuild_count == c["rebuilds"], f"step {k}"
        assert frame.candidate_pair_count == c["candidate_pairs"], f"step {k}"
        assert frame.contact_candidate_overflow_count == c["candidate_overflow"], f"step {k}"
        assert frame.maximum_contact_candidates_per_object == c["max_candidates"], f"step {k}"
        if frame.status != abi.BP_OK:       # dropped contacts let bubbles sink in: the assertion trips
            break
        assert_same_state(frame.bubbles, cb, f"step {k}")
        off, ent = world.read_adjacency()
        assert np.array_equal(off, ex["row_offsets"]), f"step {k}"
        assert np.array_equal(ent, ex["entries"]), f"step {k}"
        assert np.array_equal(bits(world.read_bodies()["pose"]), bits(cbod["pose"])), f"step {k}"
        dropped += int(c["candidate_overflow"])
        cur.bubbles, cur.bodies = cb, cbod
    return paths, dropped


@pytest.mark.gpu
@pytest.mark.parametrize("limit", [1, 3, 6])
def test_lds_rows_cut_like_the_reference(oracle, limit):
    _, dropped = step_world_against_oracle(limited(scenes.config1(), limit), 12, oracle)
    assert dropped > 0


@pytest.mark.gpu
def test_body_rows_cut_like_the_reference(oracle):
    # config2 has rigid bodies; with k = 2 their rows overflow as well as the bubbles' rows.
    scene = limited(scenes.config2(), 2)
    _, dropped = step_world_against_oracle(scene, 8, oracle)
    assert dropped > 0


@pytest.mark.gpu
def test_cooperative_rows_cut_like_the_reference(oracle):
    scene = limited(scenes.config2(), 3)
    scene.config.force_cooperative = 1
    paths, dropped = step_world_against_oracle(scene, 6, oracle)
    assert paths == {abi.BP_PATH_COOPERATIVE} and dropped > 0


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [9, 10])
def test_global_rows_cut_like_the_reference(oracle, cls):
    # Class 9 is the LDS kernel with rows in global memory; class 10 the global-state kernel.
    rng = np.random.Generator(np.random.PCG64(11))
    scene = limited(scenes.batch_world(rng, 1800, bodies_max=8), 4)
    batch = solver.Batch([scene])
    batch.set_world_class(0, cls)
    infos = batch.step(4)
    cb, _, cf, _ = oracle.run(scene, 4)
    info = infos[:, 0]
    assert (info["kernel_class"] == cls).all()
    assert (info["status"] == cf["status"]).all()
    assert np.array_equal(info["substep_count"], cf["substeps"])
    assert np.array_equal(info["candidate_overflow"], cf["candidate_overflow"])
    assert np.array_equal(info["max_candidates_per_object"], cf["max_candidates"])
    assert np.array_equal(info["candidate_pair_count"], cf["candidate_pairs"])
    assert int(cf["candidate_overflow"].sum()) > 0
    assert_same_state(batch.read_bubbles(0), cb, f"class {cls}")
