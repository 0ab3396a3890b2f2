"""Writes config3_oracle.json: the oracle's first steps of the 262,144-bubble world of
scenes.config3() (counters and a SHA-256 of the state bits after each step). The oracle takes about
a minute per step at this size, so the GPU test compares against this file instead of running it.

Run from the repository root: python tests/golden/make_config3_fixture.py [steps]
"""
import hashlib
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from bubblephysics_amd import scenes  # noqa: E402
from oracle import bp_oracle  # noqa: E402

STATE_FIELDS = ("center", "previous_center", "velocity", "rotation", "angular_velocity")


def state_digest(bubbles):
    h = hashlib.sha256()
    for f in STATE_FIELDS:
        h.update(bubbles[f].tobytes())
    return h.hexdigest()


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 2
    scene = scenes.config3()
    cur = scene.copy()
    out = {"workload": "config3: 262144 bubbles, 4096 kinematic bodies + 4 walls, 16384^2, speed N(0,50)", "steps": []}
    for k in range(steps):
        t = time.time()
        b, bodies, frames, _ = bp_oracle.run(cur, 1)
        f = frames[0]
        out["steps"].append({"status": int(f["status"]), "substeps": int(f["substeps"]), "rebuilds": int(f["rebuilds"]),
                             "candidate_pairs": int(f["candidate_pairs"]), "max_candidates": int(f["max_candidates"]),
                             "state_sha256": state_digest(b)})
        print(f"step {k}: {out['steps'][-1]} ({time.time() - t:.1f}s)", flush=True)
        cur.bubbles, cur.bodies = b, bodies
    (Path(__file__).resolve().parent / "config3_oracle.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
