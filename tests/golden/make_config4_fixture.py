"""Writes config4_oracle.json: the oracle's run of every world of the batched workload
(scenes.config4(): 4,096 worlds, bubbles log-uniform in [16, 4096]) for 23 steps, the 3 warm-up
and 20 timed steps of the benchmark. Per world: bubble count, steps executed (a failing step ends
the run and counts), status of the last executed step, substeps summed over the successful steps,
the limiter object of a substep-limit failure or the violation objects of a contact violation, and
the first 16 hex digits of a SHA-256 of the state bits after the last successful step. The GPU test
compares the HIP batch against this file world by world.

Run from the repository root: python tests/golden/make_config4_fixture.py [threads]
"""
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from bubblephysics_amd import abi, scenes  # noqa: E402
from oracle import bp_oracle  # noqa: E402

STEPS = 23
STATE_FIELDS = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def state_digest(bubbles):
    h = hashlib.sha256()
    for f in STATE_FIELDS:
        h.update(bubbles[f].tobytes())
    return h.hexdigest()[:16]


def run_world(scene):
    cur = scene.copy()
    substeps, executed, status, detail = 0, 0, abi.BP_OK, []
    for _ in range(STEPS):
        b, bodies, frames, _ = bp_oracle.run(cur, 1)
        f = frames[0]
        executed += 1
        status = int(f["status"])
        if status != abi.BP_OK:
            if status == abi.BP_ERR_SUBSTEP_LIMIT_EXCEEDED:
                detail = [int(f["limiter_object"])]
            elif status == abi.BP_ERR_CONTACT_VIOLATION:
                detail = sorted(int(x) for x in f["violation_objects"] if x)
            break
        substeps += int(f["substeps"])
        cur.bubbles, cur.bodies = b, bodies
    return [len(scene.bubbles), executed, status, substeps, detail, state_digest(cur.bubbles)]


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    t = time.time()
    worlds = scenes.config4()
    print(f"built {len(worlds)} worlds in {time.time() - t:.1f}s", flush=True)
    bp_oracle.load()
    t = time.time()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        rows = list(pool.map(run_world, worlds))
    failed = [k for k, r in enumerate(rows) if r[2] != abi.BP_OK]
    print(f"ran in {time.time() - t:.1f}s on {threads} threads; {len(failed)} worlds fail", flush=True)
    out = {"workload": "config4: 4096 worlds, bubbles log-uniform in [16, 4096], 0-16 kinematic bodies, walls, dt 1/60",
           "steps": STEPS,
           "columns": ["bubbles", "executed", "status", "substeps", "failure objects", "state sha256[:16]"],
           "worlds": rows}
    path = Path(__file__).resolve().parent / "config4_oracle.json"
    path.write_text(json.dumps(out, separators=(",", ":")) + "\n")
    print(f"wrote {path} ({path.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
