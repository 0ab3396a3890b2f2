"""GPU parity at the measured workload: the 4,096 worlds of the batched benchmark (scenes.config4())
advanced by the HIP batch for the benchmark's 3 warm-up and 20 timed steps equal the oracle's run
of each world (tests/golden/config4_oracle.json, written by tests/golden/make_config4_fixture.py):
steps executed, status of the last one, substeps, the objects a failure names and the bits of the
final state."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from bubblephysics_amd import abi, scenes, solver

pytestmark = pytest.mark.gpu

STATE_FIELDS = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def digest(bubbles):
    h = hashlib.sha256()
    for f in STATE_FIELDS:
        h.update(np.ascontiguousarray(bubbles[f]).tobytes())
    return h.hexdigest()[:16]


def test_config4_batch_matches_the_oracle_world_by_world():
    golden = json.loads((Path(__file__).resolve().parent / "golden" / "config4_oracle.json").read_text())
    worlds = scenes.config4()
    assert len(worlds) == len(golden["worlds"])
    batch = solver.Batch(worlds)
    infos = np.concatenate([batch.step(3), batch.step(golden["steps"] - 3)])
    bubbles, offsets = batch.read_all_bubbles()
    mismatched = []
    for w, (n, executed, status, substeps, objects, sha) in enumerate(golden["worlds"]):
        col = infos[:, w]
        ran = col[col["status"] != abi.BP_ERR_UNAVAILABLE]
        last = ran[-1]
        ok = ran["status"] == abi.BP_OK
        got_objects = []
        if last["status"] == abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED:
            got_objects = [int(last["limiter_object"])]
        elif last["status"] == abi.BP_ERR_CONTACT_VIOLATION:
            got_objects = sorted(int(x) for x in last["detail"]["object_ids"] if x)
        got = [int(offsets[w + 1] - offsets[w]), len(ran), int(last["status"]), int(ran["substep_count"][ok].sum()),
               got_objects, digest(bubbles[offsets[w]:offsets[w + 1]])]
        if got != [n, executed, status, substeps, objects, sha]:
            mismatched.append((w, got, [n, executed, status, substeps, objects, sha]))
    assert mismatched == [], mismatched[:5]
    failed = sum(1 for r in golden["worlds"] if r[2] != abi.BP_OK)
    print(f"[config4] {len(worlds)} worlds equal the oracle; {failed} fail in both", flush=True)
