"""Workgroup pools of the batched launch (bp_batch_set_pooling).

Once a step has measured every world's device time, the classes whose workgroup fills a CU are
launched with a share of the CUs instead of one workgroup per world: each workgroup takes the
class's worlds longest first, the next one when it finishes one, and the other classes wait
behind a gate until the pools are placed. Only the placement of worlds on CUs changes, so the
results must be the same bits as one workgroup per world, and as the oracle's.
"""
import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, solver
from test_gpu_world import assert_same_state


def _worlds():
    # 600 worlds of the 512-thread and 1,024-thread classes: more than the CUs, so the pools
    # hold fewer workgroups than worlds; and small worlds behind the gate.
    rng = np.random.Generator(np.random.PCG64(31))
    sizes = [int(x) for x in rng.integers(300, 1000, 600)] + [int(x) for x in rng.integers(16, 200, 80)]
    return [scenes.batch_world(rng, n, bodies_max=4) for n in sizes]


@pytest.mark.gpu
def test_pooled_launches_give_the_same_bits(oracle):
    worlds = _worlds()
    runs = []
    for pooling in (False, True):
        batch = solver.Batch(worlds)
        batch.set_pooling(pooling)
        first = batch.step(1)                  # measures the worlds' device times
        infos = batch.step(4)                  # pooled when pooling is on
        out, offsets = batch.read_all_bubbles()
        runs.append((first, infos, out, offsets))
        batch.close()
    (f0, i0, b0, o0), (f1, i1, b1, o1) = runs
    assert np.array_equal(o0, o1)
    assert np.array_equal(i0["status"], i1["status"])
    assert np.array_equal(i0["substep_count"], i1["substep_count"])
    assert b0.tobytes() == b1.tobytes()
    # And both are the oracle's, on a few worlds of each kind.
    for w in (0, 299, 599, 600, 679):
        cb, _, cf, _ = oracle.run(worlds[w], 5)
        assert np.array_equal(np.concatenate([f1[:, w]["status"], i1[:, w]["status"]]), cf["status"]), w
        assert_same_state(b1[o1[w]:o1[w + 1]], cb, f"world {w}")


@pytest.mark.gpu
def test_pools_follow_the_pooling_switch_between_calls():
    worlds = _worlds()[:300]
    batch = solver.Batch(worlds)
    batch.step(1)
    batch.set_pooling(False)
    a = batch.step(2)
    batch.set_pooling(True)
    b = batch.step(2)
    assert (a["status"] == abi.BP_OK).any() and (b["status"] == abi.BP_OK).any()
    ref = solver.Batch(worlds)
    ref.set_pooling(False)
    ref.step(5)
    assert batch.read_all_bubbles()[0].tobytes() == ref.read_all_bubbles()[0].tobytes()
